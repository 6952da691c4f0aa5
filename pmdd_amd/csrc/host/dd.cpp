// Decision-diagram constructors and flattening.
#include "dd.hpp"

#include <algorithm>
#include <atomic>
#include <thread>
#include <climits>
#include <cstdlib>
#include <cstring>

#include "word_map.hpp"

namespace pmdd {

pmdd_dd_view FlatDD::view() const {
    pmdd_dd_view v;
    std::memset(&v, 0, sizeof(v));
    v.d = d;
    v.num_layers = num_layers;
    v.kind = kind;
    v.root = root;
    v.terminal = terminal;
    v.layer_off = layer_off.data();
    v.in_off = in_off.data();
    v.arc_src = arc_src.data();
    v.arc_val = arc_val.data();
    v.arc_w = arc_w.data();
    v.min_weight = min_weight.empty() ? nullptr : min_weight.data();
    v.sd_partner = sd_partner.empty() ? nullptr : sd_partner.data();
    return v;
}

int64_t BinaryDD::width() const {
    int64_t w = 0;
    for (int l = 0; l < num_layers; ++l) {
        w = std::max(w, layer_size(l));
    }
    return w;
}

int64_t BinaryDD::num_nodes() const {
    int64_t n = 0;
    for (int l = 0; l < num_layers; ++l) {
        n += layer_size(l);
    }
    return n;
}

// ---------------------------------------------------------------------------
// Knapsack
// ---------------------------------------------------------------------------

BinaryDD build_knapsack_exact(const KnapsackInstance &inst, int d) {
    const int n = inst.n_vars;
    const int nc = inst.n_cons;
    BinaryDD bdd;
    bdd.d = d;
    bdd.num_layers = n + 1;
    bdd.child.resize(n);
    bdd.one_w.assign(n, std::vector<int32_t>(d, 0));
    for (int l = 0; l < n; ++l) {
        for (int o = 0; o < d; ++o) {
            bdd.one_w[l][o] = inst.obj_coeffs[l][o];
        }
    }

    // States of the current layer, one word per constraint (int32 stored in a
    // uint64 so that the map can hash it).
    std::vector<uint64_t> cur(nc, 0); // root: nothing used
    WordKeyMap next_map(nc);
    std::vector<uint64_t> key(nc);
    std::vector<int64_t> used(nc);

    for (int l = 0; l < n; ++l) {
        const int64_t width = static_cast<int64_t>(cur.size()) / nc;
        bdd.child[l].assign(width, {-1, -1});
        next_map.clear();
        for (int64_t i = 0; i < width; ++i) {
            const uint64_t *s = cur.data() + i * nc;
            // one-arc feasibility: every constraint stays within its rhs
            bool feasible = true;
            for (int c = 0; c < nc; ++c) {
                used[c] = static_cast<int64_t>(static_cast<int32_t>(s[c])) + inst.coeffs[c][l];
                feasible = feasible && used[c] <= inst.rhs[c];
            }
            if (l < n - 1) {
                bool ins = false;
                bdd.child[l][i][0] = static_cast<int32_t>(next_map.insert(s, &ins));
                if (feasible) {
                    for (int c = 0; c < nc; ++c) {
                        key[c] = static_cast<uint64_t>(static_cast<uint32_t>(used[c]));
                    }
                    bdd.child[l][i][1] = static_cast<int32_t>(next_map.insert(key.data(), &ins));
                }
            } else {
                bdd.child[l][i][0] = 0; // terminal
                if (feasible) {
                    bdd.child[l][i][1] = 0;
                }
            }
        }
        if (l < n - 1) {
            const int64_t nn = next_map.size();
            cur.assign(next_map.key(0), next_map.key(0) + nn * nc);
        }
    }
    return bdd;
}

void reduce_bdd(BinaryDD *bdd) {
    const int L = bdd->num_layers;
    // map from old index to new index for the layer below the one processed
    std::vector<int32_t> below_map(1, 0); // terminal layer keeps its single node
    WordKeyMap pairs(1);
    for (int l = L - 2; l >= 0; --l) {
        ChildLayer &layer = bdd->child[l];
        std::vector<int32_t> this_map(layer.size(), -1);
        ChildLayer kept;
        kept.reserve(layer.size());
        pairs.clear();
        for (size_t i = 0; i < layer.size(); ++i) {
            int32_t c0 = layer[i][0] >= 0 ? below_map[layer[i][0]] : -1;
            int32_t c1 = layer[i][1] >= 0 ? below_map[layer[i][1]] : -1;
            if (c0 < 0 && c1 < 0) {
                continue; // no way down: removed, its parents lose the arc
            }
            const uint64_t k = (static_cast<uint64_t>(static_cast<uint32_t>(c0)) << 32) |
                               static_cast<uint32_t>(c1);
            bool ins = false;
            const int64_t id = pairs.insert(&k, &ins);
            if (ins) {
                kept.push_back({c0, c1});
            }
            this_map[i] = static_cast<int32_t>(id);
        }
        layer.swap(kept);
        below_map.swap(this_map);
    }
}

std::vector<std::vector<int32_t>> knapsack_min_weights(const BinaryDD &bdd,
                                                        const KnapsackInstance &inst) {
    const int L = bdd.num_layers;
    std::vector<std::vector<int32_t>> mw(L);
    mw[0].assign(bdd.layer_size(0), 0);
    for (int l = 1; l < L; ++l) {
        mw[l].assign(bdd.layer_size(l), INT_MAX);
        const int32_t a = inst.coeffs[0][l - 1];
        for (size_t i = 0; i < bdd.child[l - 1].size(); ++i) {
            const int32_t p = mw[l - 1][i];
            const int32_t c0 = bdd.child[l - 1][i][0];
            const int32_t c1 = bdd.child[l - 1][i][1];
            if (c0 >= 0) {
                mw[l][c0] = std::min(mw[l][c0], p);
            }
            if (c1 >= 0) {
                mw[l][c1] = std::min(mw[l][c1], p + a);
            }
        }
    }
    return mw;
}

// ---------------------------------------------------------------------------
// Set packing as an independent-set BDD
// ---------------------------------------------------------------------------

// Host threads a constructor may use: the machine's hardware threads, at
// most 8, and at most PMDD_HOST_THREADS when that is set to a positive
// number (read on every build).  The diagrams do not depend on it.
static int host_threads() {
    int t = static_cast<int>(std::min<unsigned>(8, std::max(1u, std::thread::hardware_concurrency())));
    if (const char *e = std::getenv("PMDD_HOST_THREADS")) {
        const int v = std::atoi(e);
        if (v > 0) {
            t = std::min(t, v);
        }
    }
    return t;
}

// Runs f(begin, end, t) over T contiguous ranges of [0, n) on T host threads
// (T = 1: on the calling thread).
template <class F> static void parallel_ranges(int64_t n, int T, F f) {
    if (T <= 1) {
        f(int64_t(0), n, 0);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t) {
        pool.emplace_back([&, t]() { f(n * t / T, n * (t + 1) / T, t); });
    }
    for (std::thread &th : pool) {
        th.join();
    }
}

// Set-packing state dominance (filter_dominance_setpacking_cpu,
// dominance.cpp:88-173): the nodes of a layer are bucketed by the size of
// their state (eligible vertices); a node in a bucket of more than one node
// may have its labels dropped by the labels of the first node of the next
// non-empty bucket that has the same smallest vertex and a superset state,
// skipping pairs of single-parent nodes fed by the same parent through
// different arcs.  Such a partner has every completion of the node (its
// eligible set is larger), so dropping the node's labels that a partner
// label weakly dominates keeps the frontier.  The scan over the candidates
// of a node stops after kPartnerScan of them (then no partner: less pruning,
// same frontier).
static constexpr int kPartnerScan = 64;

static std::vector<int32_t> setpacking_partners(const std::vector<uint64_t> &states, int W,
                                                int n_vars,
                                                const ChildLayer &parents) {
    const int64_t N = static_cast<int64_t>(states.size()) / W;
    std::vector<int32_t> partner(N, -1);
    if (N < 2) {
        return partner;
    }
    const int T = N >= 65536 ? host_threads() : 1;
    // single in-arc: its parent and arc type (sp, sa are read only where the
    // in-degree is 1, so only one writer matters, and they need no fill)
    std::vector<int32_t> indeg(N, 0);
    RawVec<int32_t> sp(N), sa(N);
    const int64_t P = static_cast<int64_t>(parents.size());
    parallel_ranges(P, T, [&](int64_t p0, int64_t p1, int) {
        for (int64_t p = p0; p < p1; ++p) {
            for (int t = 0; t < 2; ++t) {
                const int32_t c = parents[p][t];
                if (c >= 0) {
                    __atomic_fetch_add(&indeg[c], 1, __ATOMIC_RELAXED);
                    __atomic_store_n(&sp[c], static_cast<int32_t>(p), __ATOMIC_RELAXED);
                    __atomic_store_n(&sa[c], t, __ATOMIC_RELAXED);
                }
            }
        }
    });
    // bucket of a node: its state size; the nodes sorted stably by (size,
    // smallest vertex), so a bucket's nodes with one smallest vertex are a
    // run in node order
    const int64_t nf = n_vars + 1; // smallest vertex + 1 (0: empty state)
    const int64_t nkeys = static_cast<int64_t>(n_vars + 1) * nf;
    RawVec<int32_t> count(N), first(N);
    std::vector<int64_t> hist(static_cast<size_t>(nkeys) * T, 0);
    parallel_ranges(N, T, [&](int64_t i0, int64_t i1, int t) {
        int64_t *ht = hist.data() + static_cast<size_t>(nkeys) * t;
        for (int64_t i = i0; i < i1; ++i) {
            const uint64_t *s = states.data() + i * W;
            int c = 0, f = -1;
            for (int w = 0; w < W; ++w) {
                c += __builtin_popcountll(s[w]);
                if (f < 0 && s[w] != 0) {
                    f = w * 64 + __builtin_ctzll(s[w]);
                }
            }
            count[i] = c;
            first[i] = f;
            ++ht[c * nf + (f + 1)];
        }
    });
    // key_off[k]: start of key k in the sorted order; hist[t][k] becomes
    // where thread t's range writes its nodes of key k
    std::vector<int64_t> key_off(static_cast<size_t>(nkeys) + 1, 0);
    int64_t run = 0;
    for (int64_t k = 0; k < nkeys; ++k) {
        key_off[k] = run;
        for (int t = 0; t < T; ++t) {
            int64_t &h = hist[static_cast<size_t>(nkeys) * t + k];
            const int64_t c = h;
            h = run;
            run += c;
        }
    }
    key_off[nkeys] = run;
    RawVec<int32_t> sorted(N);
    parallel_ranges(N, T, [&](int64_t i0, int64_t i1, int t) {
        int64_t *at = hist.data() + static_cast<size_t>(nkeys) * t;
        for (int64_t i = i0; i < i1; ++i) {
            sorted[at[count[i] * nf + (first[i] + 1)]++] = static_cast<int32_t>(i);
        }
    });
    // next non-empty bucket after each bucket (n_vars + 1: none)
    std::vector<int64_t> bucket_size(n_vars + 1);
    for (int c = 0; c <= n_vars; ++c) {
        bucket_size[c] = key_off[(c + 1) * nf] - key_off[c * nf];
    }
    std::vector<int> next_bucket(n_vars + 1, n_vars + 1);
    for (int c = n_vars - 1, nb = n_vars + 1; c >= 0; --c) {
        if (bucket_size[c + 1] > 0) {
            nb = c + 1;
        }
        next_bucket[c] = nb;
    }
    // every node's search is its own: a node of a bucket with more than one
    // node takes the first of (at most kPartnerScan) nodes of the next
    // non-empty bucket with its smallest vertex, in node order, whose state
    // contains its own.  Nodes are taken in the sorted order, so the nodes of
    // one key scan the same run of candidates one after another, and states
    // are read from a copy in that order.
    RawVec<uint64_t> ss(static_cast<size_t>(N) * W);
    parallel_ranges(N, T, [&](int64_t q0, int64_t q1, int) {
        for (int64_t q = q0; q < q1; ++q) {
            const uint64_t *s = states.data() + static_cast<int64_t>(sorted[q]) * W;
            std::copy(s, s + W, ss.data() + q * W);
        }
    });
    parallel_ranges(N, T, [&](int64_t q0, int64_t q1, int) {
        for (int64_t q = q0; q < q1; ++q) {
            const int32_t i = sorted[q];
            const int c = count[i];
            if (bucket_size[c] <= 1 || next_bucket[c] > n_vars) {
                continue;
            }
            const int64_t k = static_cast<int64_t>(next_bucket[c]) * nf + (first[i] + 1);
            const uint64_t *si = ss.data() + q * W;
            const int64_t r1 = std::min<int64_t>(key_off[k + 1], key_off[k] + kPartnerScan);
            for (int64_t r = key_off[k]; r < r1; ++r) {
                const uint64_t *sj = ss.data() + r * W;
                bool subset = true;
                for (int w = 0; w < W && subset; ++w) {
                    subset = (si[w] & ~sj[w]) == 0;
                }
                if (!subset) {
                    continue;
                }
                // siblings of one parent through different arcs are skipped
                const int32_t j = sorted[r];
                if (indeg[i] == 1 && indeg[j] == 1 && sp[i] == sp[j] && sa[i] != sa[j]) {
                    continue;
                }
                partner[i] = j;
                break;
            }
        }
    });
    return partner;
}

// Hash of a key of W words (SplitMix64 finaliser folded over the words).
static inline uint64_t word_hash(const uint64_t *k, int W) {
    uint64_t x = 0x9e3779b97f4a7c15ULL * static_cast<uint64_t>(W);
    for (int i = 0; i < W; ++i) {
        uint64_t z = k[i] + 0x9e3779b97f4a7c15ULL + (x << 6) + (x >> 2);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        x ^= z ^ (z >> 31);
    }
    return x;
}

// Working buffers of dedup_in_order, kept between the layers of one build so
// that a layer neither allocates nor clears them (they only grow).
struct DedupScratch {
    std::vector<uint64_t> h;
    std::vector<int64_t> rep, order;
    std::vector<std::vector<uint64_t>> table; // one per partition

    void reserve(int64_t n, int T) {
        if (h.size() < static_cast<size_t>(n)) {
            h.resize(static_cast<size_t>(n));
            rep.resize(static_cast<size_t>(n));
            order.resize(static_cast<size_t>(n));
        }
        if (table.size() < static_cast<size_t>(T)) {
            table.resize(static_cast<size_t>(T));
        }
    }
};

// Distinct keys of keys[0 .. n) (W words each) in order of first occurrence:
// ids[c] is the rank of key c's first occurrence among the distinct keys, the
// same ids a map that hands out ids in insertion order gives, and uniq holds
// the distinct keys in id order.  On T threads: the children are bucketed by
// hash partition (ascending within a bucket), and each thread keeps the keys
// of one partition in its own table, taking its bucket in order.  Which
// partition a key falls in does not change the ids: a key's first occurrence
// is the first child of its bucket that carries it.
static void dedup_in_order(const uint64_t *keys, int64_t n, int W, int T,
                           std::vector<int64_t> *ids, std::vector<uint64_t> *uniq,
                           DedupScratch *scratch) {
    // A table slot holds a child index in the low kIdBits bits and the top
    // bits of its hash above them, so most probes skip the key compare.
    constexpr int kIdBits = 40;
    constexpr uint64_t kIdMask = (uint64_t(1) << kIdBits) - 1;
    constexpr uint64_t kEmpty = ~uint64_t(0);
    constexpr size_t kAhead = 16;
    scratch->reserve(n, T);
    uint64_t *h = scratch->h.data();
    int64_t *rep = scratch->rep.data();
    int64_t *order = scratch->order.data();
    // partition of a child: its hash's top bits scaled to [0, T)
    auto part = [&](uint64_t x) {
        return static_cast<int>(((x >> kIdBits) * static_cast<uint64_t>(T)) >> (64 - kIdBits));
    };
    // cnt[t * T + p]: children of thread t's range in partition p
    std::vector<int64_t> cnt(static_cast<size_t>(T) * T, 0);
    parallel_ranges(n, T, [&](int64_t c0, int64_t c1, int t) {
        int64_t *ct = cnt.data() + static_cast<size_t>(t) * T;
        for (int64_t c = c0; c < c1; ++c) {
            h[c] = word_hash(keys + c * W, W);
            ++ct[part(h[c])];
        }
    });
    // bucket p holds its children in ascending order: the ranges of threads
    // 0 .. T-1 one after another
    std::vector<int64_t> at(static_cast<size_t>(T) * T), bucket(static_cast<size_t>(T) + 1, 0);
    int64_t run = 0;
    for (int p = 0; p < T; ++p) {
        bucket[p] = run;
        for (int t = 0; t < T; ++t) {
            at[static_cast<size_t>(t) * T + p] = run;
            run += cnt[static_cast<size_t>(t) * T + p];
        }
    }
    bucket[T] = run;
    parallel_ranges(n, T, [&](int64_t c0, int64_t c1, int t) {
        int64_t *a = at.data() + static_cast<size_t>(t) * T;
        for (int64_t c = c0; c < c1; ++c) {
            order[a[part(h[c])]++] = c;
        }
    });
    // partition p: one table of the first child index of each distinct key
    parallel_ranges(T, T, [&](int64_t, int64_t, int p) {
        const int64_t *mine = order + bucket[p];
        const size_t count = static_cast<size_t>(bucket[p + 1] - bucket[p]);
        size_t cap = 16;
        while (cap < 2 * count) {
            cap *= 2;
        }
        std::vector<uint64_t> &table = scratch->table[p];
        if (table.size() < cap) {
            table.resize(cap);
        }
        std::fill(table.begin(), table.begin() + static_cast<std::ptrdiff_t>(cap), kEmpty);
        const size_t mask = cap - 1;
        for (size_t i = 0; i < std::min(count, kAhead); ++i) {
            __builtin_prefetch(&table[static_cast<size_t>(h[mine[i]]) & mask]);
        }
        for (size_t i = 0; i < count; ++i) {
            if (i + kAhead < count) {
                __builtin_prefetch(&table[static_cast<size_t>(h[mine[i + kAhead]]) & mask]);
            }
            const int64_t c = mine[i];
            const uint64_t fp = h[c] >> kIdBits;
            size_t pos = static_cast<size_t>(h[c]) & mask;
            while (true) {
                const uint64_t e = table[pos];
                if (e == kEmpty) {
                    table[pos] = (fp << kIdBits) | static_cast<uint64_t>(c);
                    rep[c] = c;
                    break;
                }
                const int64_t ec = static_cast<int64_t>(e & kIdMask);
                if ((e >> kIdBits) == fp &&
                    std::memcmp(keys + ec * W, keys + c * W, sizeof(uint64_t) * W) == 0) {
                    rep[c] = ec;
                    break;
                }
                pos = (pos + 1) & mask;
            }
        }
    });
    // ids of first occurrences by a prefix count, then every child's id
    std::vector<int64_t> base(static_cast<size_t>(T) + 1, 0);
    parallel_ranges(n, T, [&](int64_t c0, int64_t c1, int t) {
        int64_t k = 0;
        for (int64_t c = c0; c < c1; ++c) {
            k += rep[c] == c ? 1 : 0;
        }
        base[t + 1] = k;
    });
    for (int t = 0; t < T; ++t) {
        base[t + 1] += base[t];
    }
    ids->resize(static_cast<size_t>(n));
    uniq->resize(static_cast<size_t>(base[T]) * W);
    parallel_ranges(n, T, [&](int64_t c0, int64_t c1, int t) {
        int64_t next = base[t];
        for (int64_t c = c0; c < c1; ++c) {
            if (rep[c] == c) {
                (*ids)[c] = next;
                std::copy(keys + c * W, keys + (c + 1) * W, uniq->data() + next * W);
                ++next;
            }
        }
    });
    parallel_ranges(n, T, [&](int64_t c0, int64_t c1, int) {
        for (int64_t c = c0; c < c1; ++c) {
            if (rep[c] != c) {
                (*ids)[c] = (*ids)[rep[c]];
            }
        }
    });
}

BinaryDD build_setpacking_exact(const SetPackingInstance &inst, int d,
                                std::vector<int32_t> *var_layer,
                                std::vector<std::vector<int32_t>> *partners) {
    const int n = inst.n_vars;
    const int W = (n + 63) / 64;
    // Conflict graph: variables sharing a constraint are adjacent.  compl[v]
    // keeps the vertices still eligible after v is chosen (not v, not N(v)).
    std::vector<uint64_t> compl_mask(static_cast<size_t>(n) * W, 0);
    for (int v = 0; v < n; ++v) {
        for (int w = 0; w < n; ++w) {
            compl_mask[static_cast<size_t>(v) * W + w / 64] |= (1ULL << (w % 64));
        }
    }
    auto clear_bit = [&](int v, int w) {
        compl_mask[static_cast<size_t>(v) * W + w / 64] &= ~(1ULL << (w % 64));
    };
    for (int v = 0; v < n; ++v) {
        clear_bit(v, v);
    }
    for (const std::vector<int32_t> &cons : inst.vars_cons) {
        for (size_t i = 0; i < cons.size(); ++i) {
            for (size_t j = i + 1; j < cons.size(); ++j) {
                if (cons[i] != cons[j]) {
                    clear_bit(cons[i], cons[j]);
                    clear_bit(cons[j], cons[i]);
                }
            }
        }
    }

    BinaryDD bdd;
    bdd.d = d;
    bdd.num_layers = n + 1;
    bdd.child.resize(n);
    bdd.one_w.assign(n, std::vector<int32_t>(d, 0));
    if (var_layer != nullptr) {
        var_layer->assign(n, -1);
    }

    std::vector<uint64_t> cur(W, 0);
    for (int v = 0; v < n; ++v) {
        cur[v / 64] |= (1ULL << (v % 64));
    }
    if (partners != nullptr) {
        partners->assign(n + 1, std::vector<int32_t>());
        partners->at(0).assign(1, -1);
    }
    std::vector<int32_t> active(n);
    for (int v = 0; v < n; ++v) {
        active[v] = v;
    }
    std::vector<int64_t> counter(n, 0);
    const size_t hist_len = static_cast<size_t>(W) * 8 * 256;
    // wide layers are split over host threads (the node order is unchanged)
    const int max_threads = host_threads();
    std::vector<int64_t> byte_hist(hist_len * max_threads, 0);
    WordKeyMap next_map(W);
    std::vector<uint64_t> child_keys;
    std::vector<int64_t> child_slot, child_ids, child_base(max_threads + 1);
    DedupScratch dedup_scratch;

    for (int l = 1; l <= n; ++l) {
        const int64_t width = static_cast<int64_t>(cur.size()) / W;
        const int T = width >= 65536 ? max_threads : 1;
        // min-in-state: the active vertex contained in the fewest states,
        // first in active-list order on ties
        // per byte position, how many states hold each byte value (per
        // thread, then summed); a vertex's count is the sum over the values
        // with its bit set
        std::fill(byte_hist.begin(), byte_hist.begin() + hist_len * T, 0);
        parallel_ranges(width, T, [&](int64_t i0, int64_t i1, int t) {
            int64_t *ht = byte_hist.data() + hist_len * t;
            for (int64_t i = i0; i < i1; ++i) {
                const uint64_t *s = cur.data() + i * W;
                for (int w = 0; w < W; ++w) {
                    uint64_t bits = s[w];
                    int64_t *h = ht + static_cast<size_t>(w) * 8 * 256;
                    for (int b = 0; b < 8; ++b) {
                        h[b * 256 + (bits & 0xFF)]++;
                        bits >>= 8;
                    }
                }
            }
        });
        for (int t = 1; t < T; ++t) {
            const int64_t *ht = byte_hist.data() + hist_len * t;
            for (size_t k = 0; k < hist_len; ++k) {
                byte_hist[k] += ht[k];
            }
        }
        for (int32_t v : active) {
            const int64_t *h = byte_hist.data() + static_cast<size_t>(v / 8) * 256;
            const int bit = v % 8;
            int64_t c = 0;
            for (int val = 0; val < 256; ++val) {
                c += ((val >> bit) & 1) ? h[val] : 0;
            }
            counter[v] = c;
        }
        size_t sel = 0;
        for (size_t i = 1; i < active.size(); ++i) {
            if (counter[active[i]] < counter[active[sel]]) {
                sel = i;
            }
        }
        const int vertex = active[sel];
        active[sel] = active.back();
        active.pop_back();
        if (var_layer != nullptr) {
            (*var_layer)[l - 1] = vertex;
        }
        for (int o = 0; o < d; ++o) {
            bdd.one_w[l - 1][o] = inst.objs[o][vertex];
        }

        next_map.clear();
        const int vw = vertex / 64;
        const uint64_t vb = 1ULL << (vertex % 64);
        const uint64_t *cm = compl_mask.data() + static_cast<size_t>(vertex) * W;
        // child states in state order (0-child, then the 1-child when the
        // vertex is eligible), inserted in that order; each thread writes
        // the children of its states at the offset its predecessors leave
        std::fill(child_base.begin(), child_base.end(), 0);
        parallel_ranges(width, T, [&](int64_t i0, int64_t i1, int t) {
            int64_t c = 0;
            for (int64_t i = i0; i < i1; ++i) {
                c += 1 + ((cur[static_cast<size_t>(i) * W + vw] & vb) != 0 ? 1 : 0);
            }
            child_base[t + 1] = c;
        });
        for (int t = 0; t < T; ++t) {
            child_base[t + 1] += child_base[t];
        }
        const int64_t n_children = child_base[T];
        child_keys.resize(static_cast<size_t>(n_children) * W);
        child_slot.resize(static_cast<size_t>(n_children));
        parallel_ranges(width, T, [&](int64_t i0, int64_t i1, int t) {
            int64_t at = child_base[t];
            for (int64_t i = i0; i < i1; ++i) {
                const uint64_t *s = cur.data() + i * W;
                const bool was_set = (s[vw] & vb) != 0;
                uint64_t *k0 = child_keys.data() + at * W;
                std::copy(s, s + W, k0);
                k0[vw] &= ~vb;
                child_slot[at++] = 2 * i;
                if (was_set) {
                    uint64_t *k1 = child_keys.data() + at * W;
                    for (int w = 0; w < W; ++w) {
                        k1[w] = k0[w] & cm[w];
                    }
                    child_slot[at++] = 2 * i + 1;
                }
            }
        });
        if (T > 1) {
            dedup_in_order(child_keys.data(), n_children, W, T, &child_ids, &cur, &dedup_scratch);
        } else {
            child_ids.resize(child_slot.size());
            next_map.insert_many(child_keys.data(), n_children, child_ids.data());
            cur.assign(next_map.key(0), next_map.key(0) + next_map.size() * W);
        }
        // each parent's children: -1 where absent; the thread that generated
        // a range of parents' children fills those parents
        ChildLayer &layer = bdd.child[l - 1];
        layer.resize(static_cast<size_t>(width));
        parallel_ranges(width, T, [&](int64_t i0, int64_t i1, int t) {
            for (int64_t i = i0; i < i1; ++i) {
                layer[i] = {-1, -1};
            }
            for (int64_t c = child_base[t]; c < child_base[t + 1]; ++c) {
                layer[child_slot[c] / 2][child_slot[c] % 2] = static_cast<int32_t>(child_ids[c]);
            }
        });
        if (partners != nullptr) {
            (*partners)[l] = setpacking_partners(cur, W, n, bdd.child[l - 1]);
        }
    }
    return bdd;
}

// ---------------------------------------------------------------------------
// Flattening
// ---------------------------------------------------------------------------

void flatten_bdd(const BinaryDD &bdd, const std::vector<std::vector<int32_t>> *min_weight,
                 FlatDD *out) {
    const int L = bdd.num_layers;
    const int d = bdd.d;
    out->d = d;
    out->kind = PMDD_DD_BDD;
    out->num_layers = L;
    out->root = 0;
    out->terminal = 0;
    out->layer_off.assign(L + 1, 0);
    for (int l = 0; l < L; ++l) {
        out->layer_off[l + 1] = out->layer_off[l] + bdd.layer_size(l);
    }
    const int64_t total = out->layer_off[L];
    // layers 1..L-1 (the root has no in-arcs), each on one of the host
    // threads, taken in order
    const int threads = static_cast<int>(std::min<int64_t>(
        {host_threads(), std::max(L - 1, 1), std::max<int64_t>(1, total / 100000)}));
    auto for_layers = [&](auto &&f) {
        std::atomic<int> next_layer{1};
        auto worker = [&]() {
            for (int l = next_layer++; l < L; l = next_layer++) {
                f(l);
            }
        };
        if (threads <= 1) {
            worker();
            return;
        }
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) {
            pool.emplace_back(worker);
        }
        for (std::thread &th : pool) {
            th.join();
        }
    };
    // arcs into each layer, so every layer knows where its arcs start and the
    // arc arrays are sized once
    std::vector<int64_t> layer_arc0(L + 1, 0);
    for_layers([&](int l) {
        int64_t a = 0;
        for (const std::array<int32_t, 2> &c : bdd.child[l - 1]) {
            a += (c[0] >= 0 ? 1 : 0) + (c[1] >= 0 ? 1 : 0);
        }
        layer_arc0[l + 1] = a;
    });
    for (int l = 1; l < L; ++l) {
        layer_arc0[l + 1] += layer_arc0[l];
    }
    const int64_t total_arcs = layer_arc0[L];
    // sized without a fill: fill_layer writes every element (its layer's
    // offsets and arcs, zero weights on 0-arcs), the root layer's offsets and
    // the closing offset are set below
    out->in_off.resize(total + 1);
    out->arc_src.resize(total_arcs);
    out->arc_val.resize(total_arcs);
    out->arc_w.resize(total_arcs * d);
    std::fill(out->in_off.begin(), out->in_off.begin() + out->layer_off[1], int64_t(0));

    // One layer's in-arcs: each head's 1-arcs then 0-arcs, parents in
    // ascending index order within each group.  Writes in_off of the layer's
    // own nodes and its own arc range only.
    auto fill_layer = [&](int l) {
        const int64_t n_here = bdd.layer_size(l);
        const ChildLayer &parents = bdd.child[l - 1];
        std::vector<int64_t> cnt1(n_here, 0), cnt0(n_here, 0);
        for (size_t p = 0; p < parents.size(); ++p) {
            if (parents[p][1] >= 0) {
                cnt1[parents[p][1]]++;
            }
            if (parents[p][0] >= 0) {
                cnt0[parents[p][0]]++;
            }
        }
        const int64_t base_node = out->layer_off[l];
        std::vector<int64_t> &pos1 = cnt1, &pos0 = cnt0; // counts become cursors
        int64_t cursor = layer_arc0[l];
        for (int64_t v = 0; v < n_here; ++v) {
            out->in_off[base_node + v] = cursor;
            const int64_t c1 = cnt1[v], c0 = cnt0[v];
            pos1[v] = cursor;
            pos0[v] = cursor + c1;
            cursor += c1 + c0;
        }
        for (size_t p = 0; p < parents.size(); ++p) {
            for (int t = 1; t >= 0; --t) {
                const int32_t h = parents[p][t];
                if (h < 0) {
                    continue;
                }
                const int64_t a = (t == 1) ? pos1[h]++ : pos0[h]++;
                out->arc_src[a] = static_cast<int32_t>(p);
                out->arc_val[a] = t;
                if (t == 1) {
                    std::copy(bdd.one_w[l - 1].begin(), bdd.one_w[l - 1].end(),
                              out->arc_w.begin() + a * d);
                } else {
                    std::fill_n(out->arc_w.begin() + a * d, d, 0);
                }
            }
        }
    };
    for_layers(fill_layer);
    out->in_off[total] = total_arcs;
    if (min_weight != nullptr) {
        out->min_weight.assign(total, 0);
        for (int l = 0; l < L; ++l) {
            std::copy((*min_weight)[l].begin(), (*min_weight)[l].end(),
                      out->min_weight.begin() + out->layer_off[l]);
        }
    } else {
        out->min_weight.clear();
    }
}

static void fill_arc_structure(FlatDD *out) {
    int64_t max_in = 0;
    for (int l = 0; l < out->num_layers; ++l) {
        const int64_t a0 = out->in_off[out->layer_off[l]];
        const int64_t a1 = out->in_off[out->layer_off[l + 1]];
        max_in = std::max(max_in, a1 - a0);
    }
    out->structure.max_num_in_arcs = max_in;
    out->structure.total_num_in_arcs = out->num_arcs();
    out->structure.num_layers = out->num_layers;
    out->structure.d = out->d;
}

// ---------------------------------------------------------------------------
// TSP
// ---------------------------------------------------------------------------

void build_tsp_exact(const TspInstance &inst, int d, FlatDD *out) {
    const int n = inst.n_cities;
    const int L = n + 2;
    out->d = d;
    out->kind = PMDD_DD_MDD;
    out->num_layers = L;
    out->root = 0;
    out->terminal = 0;

    struct ArcRec {
        int32_t tail, head, val;
    };
    // layer sizes and arcs (arcs of index l go from layer l to layer l+1)
    std::vector<int64_t> sizes(L, 0);
    std::vector<std::vector<ArcRec>> arcs(L - 1);
    std::vector<std::vector<int32_t>> weights(L - 1);

    sizes[0] = 1;
    sizes[1] = 1;
    arcs[0].push_back({0, 0, 0});
    weights[0].assign(d, 0);

    // states of layer 1: visited {0}, last city 0
    std::vector<uint64_t> cur = {1ULL, 0ULL}; // (mask, last)
    WordKeyMap next_map(2);
    uint64_t key[2];
    for (int l = 1; l < n; ++l) {
        next_map.clear();
        const int64_t width = static_cast<int64_t>(cur.size()) / 2;
        for (int64_t i = 0; i < width; ++i) {
            const uint64_t mask = cur[2 * i];
            const int last = static_cast<int>(cur[2 * i + 1]);
            for (int v = 0; v < n; ++v) {
                if (mask & (1ULL << v)) {
                    continue;
                }
                key[0] = mask | (1ULL << v);
                key[1] = static_cast<uint64_t>(v);
                bool ins = false;
                const int64_t h = next_map.insert(key, &ins);
                arcs[l].push_back({static_cast<int32_t>(i), static_cast<int32_t>(h), v});
                for (int o = 0; o < d; ++o) {
                    weights[l].push_back(-inst.at(o, last, v));
                }
            }
        }
        sizes[l + 1] = next_map.size();
        cur.assign(next_map.key(0), next_map.key(0) + next_map.size() * 2);
    }
    // closing arcs into the terminal
    sizes[n + 1] = 1;
    {
        const int64_t width = static_cast<int64_t>(cur.size()) / 2;
        for (int64_t i = 0; i < width; ++i) {
            const int last = static_cast<int>(cur[2 * i + 1]);
            arcs[n].push_back({static_cast<int32_t>(i), 0, 0});
            for (int o = 0; o < d; ++o) {
                weights[n].push_back(-inst.at(o, last, 0));
            }
        }
    }

    out->layer_off.assign(L + 1, 0);
    for (int l = 0; l < L; ++l) {
        out->layer_off[l + 1] = out->layer_off[l] + sizes[l];
    }
    const int64_t total = out->layer_off[L];
    out->in_off.assign(total + 1, 0);
    int64_t num_arcs = 0;
    for (int l = 0; l < L - 1; ++l) {
        num_arcs += static_cast<int64_t>(arcs[l].size());
    }
    out->arc_src.assign(num_arcs, 0);
    out->arc_val.assign(num_arcs, 0);
    out->arc_w.assign(num_arcs * d, 0);
    int64_t cursor = 0;
    for (int l = 1; l < L; ++l) {
        const std::vector<ArcRec> &la = arcs[l - 1];
        const int64_t base_node = out->layer_off[l];
        std::vector<int64_t> cnt(sizes[l], 0);
        for (const ArcRec &a : la) {
            cnt[a.head]++;
        }
        std::vector<int64_t> pos(sizes[l]);
        for (int64_t v = 0; v < sizes[l]; ++v) {
            out->in_off[base_node + v] = cursor;
            pos[v] = cursor;
            cursor += cnt[v];
        }
        // stable: in-arcs of a head keep insertion order (in_arcs_list order)
        for (size_t k = 0; k < la.size(); ++k) {
            const int64_t a = pos[la[k].head]++;
            out->arc_src[a] = la[k].tail;
            out->arc_val[a] = la[k].val;
            for (int o = 0; o < d; ++o) {
                out->arc_w[a * d + o] = weights[l - 1][k * d + o];
            }
        }
    }
    out->in_off[total] = cursor;
    out->min_weight.clear();
}

// ---------------------------------------------------------------------------
// main()'s construction sequence
// ---------------------------------------------------------------------------

static bool check_d(int file_d, int expect_d, std::string *error) {
    if (file_d < 1) {
        *error = "Error - instance declares no objectives";
        return false;
    }
    if (expect_d > 0 && file_d != expect_d) {
        *error = "Error - instance has " + std::to_string(file_d) +
                 " objectives but this run expects " + std::to_string(expect_d);
        return false;
    }
    return true;
}

bool build_from_file(const std::string &path, int problem_type, int expect_d, bool state_dominance,
                     FlatDD *out, std::string *error) {
    if (problem_type == 1) {
        KnapsackInstance inst;
        if (!read_knapsack(path, &inst, error) || !check_d(inst.n_objs, expect_d, error)) {
            return false;
        }
        const int d = inst.n_objs;
        BinaryDD bdd = build_knapsack_exact(inst, d);
        const int64_t original_width = bdd.width();
        const int64_t original_nodes = bdd.num_nodes();
        reduce_bdd(&bdd);
        std::vector<std::vector<int32_t>> mw = knapsack_min_weights(bdd, inst);
        flatten_bdd(bdd, &mw, out);
        out->structure.original_width = original_width;
        out->structure.original_num_nodes = original_nodes;
        out->structure.reduced_width = bdd.width();
        out->structure.reduced_num_nodes = bdd.num_nodes();
        fill_arc_structure(out);
        return true;
    }
    if (problem_type == 2) {
        SetPackingInstance inst;
        if (!read_setpacking(path, &inst, error) || !check_d(inst.n_objs, expect_d, error)) {
            return false;
        }
        const int d = inst.n_objs;
        std::vector<std::vector<int32_t>> partners;
        BinaryDD bdd =
            build_setpacking_exact(inst, d, nullptr, state_dominance ? &partners : nullptr);
        flatten_bdd(bdd, nullptr, out);
        out->sd_partner.clear();
        if (state_dominance) {
            out->sd_partner.reserve(static_cast<size_t>(out->layer_off[out->num_layers]));
            for (const std::vector<int32_t> &layer : partners) {
                out->sd_partner.insert(out->sd_partner.end(), layer.begin(), layer.end());
            }
        }
        out->structure.original_width = bdd.width();
        out->structure.original_num_nodes = bdd.num_nodes();
        out->structure.reduced_width = bdd.width();
        out->structure.reduced_num_nodes = bdd.num_nodes();
        fill_arc_structure(out);
        return true;
    }
    if (problem_type == 3) {
        TspInstance inst;
        if (!read_tsp(path, &inst, error) || !check_d(inst.n_objs, expect_d, error)) {
            return false;
        }
        build_tsp_exact(inst, inst.n_objs, out);
        out->structure.original_width = -1;
        out->structure.original_num_nodes = -1;
        out->structure.reduced_width = -1;
        out->structure.reduced_num_nodes = -1;
        fill_arc_structure(out);
        return true;
    }
    *error = "Error - problem type not recognized";
    return false;
}

} // namespace pmdd
