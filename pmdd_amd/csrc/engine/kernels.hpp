// Device kernels of the frontier engine (gfx950, wave64).
//
// Points are int32 vectors of D objectives stored structure-of-arrays in HBM
// (objective o of point i at pts[o * stride + i]).  Larger is better.  The key
// of a point is the sum of its objectives; every node frontier is kept in
// descending key order.
//
// The segmented Pareto filter works on "groups" (the candidates of one node
// during a layer expansion; the products of one coupling batch) cut into
// chunks of at most CH candidates:
//
//   k_filter_chunks   one workgroup per chunk (a node of at most CH
//                     candidates).  Loads its candidates straight from the
//                     parent frontiers (parent point + arc weight) or a dense
//                     key-ordered source into LDS, sorts them by key,
//                     descending, with a bitonic network (skipped when the
//                     source is in key order), then runs a sort-filter
//                     skyline: a candidate survives iff no earlier
//                     (higher-key) candidate is >= it in every objective.  Ties
//                     (equal points) keep the first copy.  Survivors go to HBM
//                     in key order.
//   k_gather_runs,    big nodes: the parent runs of each node are copied out
//   k_merge_pass      and merged pairwise into one key-ordered run.
//   k_stage_chunks    cuts a key-ordered run into chunks, all points alive.
//   k_cross_chunks    the chunks of a group in waves of W: a point dies when
//                     a point of the group's front (survivors of the earlier
//                     waves), an alive point of an earlier chunk of its wave,
//                     or an earlier point of its own chunk is >= it.
//   k_append_front    appends a wave's survivors to the group's front.
//   k_compact_chunks  writes the alive survivors of each chunk densely.
//
// Coupling (method 3): k_join_products writes a batch's products; the refill
// check k_refill_vs_block drops products that a block of the running frontier
// R dominates and, in flag mode, the points of R that a block of the batch
// frontier dominates.  State dominance: k_state_dominance drops the labels of
// a node that a label of its comparison nodes dominates; k_count_kept and
// k_compact_nodes rebuild the layer.
//
// Why it is exact: q >= p everywhere and q != p implies key(q) > key(p), so
// after sorting by key descending only earlier points can dominate a point;
// and if an earlier point dominates p, so does some earlier survivor
// (transitivity), hence checking survivors only is enough, and checking
// against a point that has itself died is harmless.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace pmdd_k {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kBlock / kWave;

template <int D> struct Pad {
    static constexpr int DP = D <= 4 ? 4 : 8; // survivor rows padded to 16 or 32 bytes
};

// Chunk capacity per objective count: LDS = CH*(4D + 8 + 4*DP) bytes <= 80 KB,
// so two workgroups fit one CU's 160 KB.
template <int D> struct Cap {
    static constexpr int CH = D <= 4 ? 2048 : 1024;
};

struct ChunkDesc {
    int64_t start;       // first candidate, as an index into the group's virtual list
    int64_t out_base;    // slot of the chunk's first survivor in the pass buffer
    int32_t group;       // node of the layer (expansion); unused for coupling
    int32_t len;         // candidates in the chunk, 1..CH
    int32_t group_first; // first chunk of the group
    int32_t group_count; // chunks in the group
};

// Candidate source of a layer expansion: node v's candidates are the parent
// frontiers of its arcs, concatenated in arc order, each shifted by its arc
// weight.  run_off[a] = node-local index of arc a's first candidate.
struct ExpandSrc {
    const int64_t *prev_off; // [N_prev + 1] frontier offsets of the parent layer
    const int32_t *prev_pts; // SoA, stride prev_stride
    int64_t prev_stride;
    const int64_t *arc_off; // [N + 1] arcs of each node (layer-local arc ids)
    const int32_t *arc_src; // [E] parent node of each arc
    const int32_t *arc_w;   // SoA [D][E]
    int64_t w_stride;       // = E
    const int64_t *run_off; // [E]
};

// Candidate source of a coupling batch: piece k (a node's TD rows
// r0 .. r0 + rows times its BU columns c0 .. c0 + col_size[k]) contributes the
// products td[row_base[k] + a] + bu[col_base[k] + b], numbered
// a * col_size[k] + b after prod_off[k]; row_base and col_base are the first
// TD and BU points of the piece (a node's points are contiguous).
struct JoinSrc {
    const int64_t *prod_off; // [K + 1]
    const int64_t *row_base; // [K]
    const int64_t *col_base; // [K]
    const int64_t *col_size; // [K]
    int32_t K;
    const int32_t *td_pts;
    int64_t td_stride;
    const int32_t *bu_pts;
    int64_t bu_stride;
};

template <int D>
__device__ __forceinline__ void load_candidate(const ExpandSrc &s, const ChunkDesc &c, int64_t t,
                                               int32_t (&p)[D]) {
    const int64_t q = c.start + t;
    int64_t lo = s.arc_off[c.group];
    int64_t hi = s.arc_off[c.group + 1] - 1;
    // last arc whose run starts at or before q (empty runs share a start)
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (s.run_off[mid] <= q) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    const int64_t e = lo;
    const int64_t pos = s.prev_off[s.arc_src[e]] + (q - s.run_off[e]);
#pragma unroll
    for (int o = 0; o < D; ++o) {
        p[o] = s.prev_pts[o * s.prev_stride + pos] + s.arc_w[o * s.w_stride + e];
    }
}

template <int D>
__device__ __forceinline__ void load_candidate(const JoinSrc &s, const ChunkDesc &c, int64_t t,
                                               int32_t (&p)[D]) {
    const int64_t q = c.start + t;
    int lo = 0;
    int hi = s.K - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s.prod_off[mid] <= q) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    const int64_t local = q - s.prod_off[lo];
    const int64_t bs = s.col_size[lo];
    const int64_t a = local / bs;
    const int64_t b = local - a * bs;
    const int64_t ti = s.row_base[lo] + a;
    const int64_t bi = s.col_base[lo] + b;
#pragma unroll
    for (int o = 0; o < D; ++o) {
        p[o] = s.td_pts[o * s.td_stride + ti] + s.bu_pts[o * s.bu_stride + bi];
    }
}

// Candidates already materialised in key order (SoA); c.start is the
// position of the chunk's first candidate in that buffer.
struct SortedSrc {
    const int32_t *pts;
    int64_t stride;
};

template <int D>
__device__ __forceinline__ void load_candidate(const SortedSrc &s, const ChunkDesc &c, int64_t t,
                                               int32_t (&p)[D]) {
#pragma unroll
    for (int o = 0; o < D; ++o) {
        p[o] = s.pts[o * s.stride + c.start + t];
    }
}

// Sort word: key (sum of objectives, |key| < 2^35) biased to unsigned in the
// high 48 bits, (0xFFFF - slot) in the low 16, so that descending order of
// the word is key descending, then slot ascending.
__device__ __forceinline__ uint64_t pack_key(int64_t key, int slot) {
    return (static_cast<uint64_t>(key + (int64_t(1) << 47)) << 16) |
           static_cast<uint64_t>(0xFFFF - slot);
}
__device__ __forceinline__ int unpack_slot(uint64_t w) { return 0xFFFF - static_cast<int>(w & 0xFFFF); }
__device__ __forceinline__ int64_t unpack_key(uint64_t w) {
    return static_cast<int64_t>(w >> 16) - (int64_t(1) << 47);
}

// Block-wide exclusive prefix of a 0/1 flag; returns the rank and writes the
// block total to *total.  Uses wave ballots and one LDS slot per wave.  It
// opens with a barrier, so no wave can overwrite its wave_tot slot while a
// slower wave still reads the slots of an earlier block_rank or block_sum:
// the call is safe after any other use of wave_tot.  Like every barrier, it
// must be reached by all threads of the workgroup (every caller calls it in
// workgroup-uniform control flow).  The barrier also orders the caller's
// earlier LDS reads before whatever it writes with the returned rank.
template <int NW = kWaves>
__device__ __forceinline__ int block_rank(bool flag, int *wave_tot, int *total) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const uint64_t mask = __ballot(flag);
    const int in_wave = __popcll(mask & ((lane == 0) ? 0ULL : (~0ULL >> (kWave - lane))));
    __syncthreads();
    if (lane == 0) {
        wave_tot[wave] = __popcll(mask);
    }
    __syncthreads();
    int before = 0;
    int sum = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int t = wave_tot[w];
        before += (w < wave) ? t : 0;
        sum += t;
    }
    *total = sum;
    return before + in_wave;
}

// Key (sum of objectives) of point i of a SoA buffer.
template <int D>
__device__ __forceinline__ int64_t key_at(const int32_t *pts, int64_t stride, int64_t i) {
    int64_t k = 0;
#pragma unroll
    for (int o = 0; o < D; ++o) {
        k += pts[o * stride + i];
    }
    return k;
}

// A row of DP ints (16-byte aligned) read into registers.
template <int D, int DP>
__device__ __forceinline__ void load_row(const int32_t *row, int32_t (&q)[D]) {
    const int4 a = *reinterpret_cast<const int4 *>(row);
    q[0] = a.x;
    if (D > 1) q[D > 1 ? 1 : 0] = a.y;
    if (D > 2) q[D > 2 ? 2 : 0] = a.z;
    if (D > 3) q[D > 3 ? 3 : 0] = a.w;
    if (D > 4) {
        const int4 b = *reinterpret_cast<const int4 *>(row + 4);
        q[D > 4 ? 4 : 0] = b.x;
        if (D > 5) q[D > 5 ? 5 : 0] = b.y;
        if (D > 6) q[D > 6 ? 6 : 0] = b.z;
        if (D > 7) q[D > 7 ? 7 : 0] = b.w;
    }
}

// Dominance margin of row q over point m: the least over objectives of
// q[o] - m[o], each a saturating int32 subtraction (so its sign is exact).
// q >= m in every objective iff the margin is >= 0.  A scan keeps the running
// maximum of the margins of the rows it meets and tests the sign once: per
// row that is D subtractions with clamp, a v_min3 chain and a v_max on the
// VALU, with no lane-mask arithmetic on the scalar unit (a chain of d
// v_cmp / s_and pairs per row left both units half idle).
template <int D>
__device__ __forceinline__ int32_t margin_vals(const int32_t (&q)[D], const int32_t (&m)[D]) {
    int32_t r = __builtin_elementwise_sub_sat(q[0], m[0]);
#pragma unroll
    for (int o = 1; o < D; ++o) {
        r = min(r, __builtin_elementwise_sub_sat(q[o], m[o]));
    }
    return r;
}

// The margin of a staged row (16-byte aligned, padded to DP ints).
template <int D, int DP>
__device__ __forceinline__ int32_t margin_row(const int32_t *row, const int32_t (&m)[D]) {
    int32_t q[D];
    load_row<D, DP>(row, q);
    return margin_vals<D>(q, m);
}

// Rows [0, n) of a staged table (DP ints per row, every lane reading the
// same row) against the lane's point m: returns dead, or'ed with whether
// some row the lane meets is >= m.  With LIM a lane meets only the rows
// below its own bound lim <= n (else every row).  Rows go eight at a time so
// that their LDS reads overlap, and the wave stops once none of its lanes is
// alive; n is the same in every lane.  tests counts the rows a lane met
// while alive.
template <int D, int DP, bool LIM>
__device__ __forceinline__ bool scan_until_dead(const int32_t *S, int n, int lim,
                                                const int32_t (&m)[D], bool dead,
                                                unsigned long long &tests) {
    constexpr int kB = 8;
    const int bound = LIM ? lim : n;
    int s = 0;
    for (; s + kB <= n; s += kB) {
        if (!__any(!dead && s < bound)) {
            return dead;
        }
        if (!dead) {
            tests += static_cast<unsigned long long>(max(0, min(kB, bound - s)));
        }
        int32_t acc = INT32_MIN;
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            int32_t mg = margin_row<D, DP>(&S[(s + k) * DP], m);
            if (LIM) {
                // a row at or past the bound gets a negative margin; below
                // it the bound term is >= 0 and leaves the sign as it was
                mg = min(mg, bound - 1 - s - k);
            }
            acc = max(acc, mg);
        }
        dead |= acc >= 0;
    }
    for (; s < n && s < bound && !dead; ++s) {
        ++tests;
        dead = margin_row<D, DP>(&S[s * DP], m) >= 0;
    }
    return dead;
}

// ---------------------------------------------------------------------------
// Chunk filter: per-chunk load, presort, sort-filter skyline.  The body is
// shared by k_filter_chunks (one workgroup per listed chunk) and
// k_filter_queue (persistent workgroups taking chunks from a queue).
// ---------------------------------------------------------------------------
// Nodes of at most this many arcs are put in key order by merging their
// parent runs (below); others by the bitonic network.
constexpr int kMergeRuns = 16;

template <int D, int CH, int NT> struct FilterLds {
    int32_t cs[D][CH];
    uint64_t sk[CH];
    alignas(16) int32_t S[CH * Pad<D>::DP];
    int wave_tot[NT / kWave];
    int rs[kMergeRuns + 1]; // run starts of a merged node
};

// Filters chunk ci; adds the pair tests it ran to `tests`.  All threads of
// the workgroup call it (it has barriers); it ends with one, so the caller
// may reuse the LDS at once.
template <int D, int CH, class Src, bool PRESORTED, int NT>
__device__ __forceinline__ void filter_chunk(int ci, const ChunkDesc *__restrict__ chunks,
                                             const Src &src, int32_t *__restrict__ out_pts,
                                             int64_t out_stride, uint8_t *__restrict__ alive,
                                             int32_t *__restrict__ chunk_nsurv,
                                             int32_t *__restrict__ chunk_nalive,
                                             FilterLds<D, CH, NT> &L, unsigned long long &tests) {
    constexpr int DP = Pad<D>::DP;
    auto &cs = L.cs;
    auto &sk = L.sk;
    auto &S = L.S;
    auto &wave_tot = L.wave_tot;
    const ChunkDesc c = chunks[ci];
    const int n = c.len;
    const int tid = threadIdx.x;

    int npad = 1;
    while (npad < n) {
        npad <<= 1;
    }
    for (int t = tid; t < npad; t += NT) {
        if (t < n) {
            int32_t p[D];
            load_candidate<D>(src, c, t, p);
            int64_t key = 0;
#pragma unroll
            for (int o = 0; o < D; ++o) {
                cs[o][t] = p[o];
                key += p[o];
            }
            sk[t] = pack_key(key, t);
        } else {
            sk[t] = 0; // below every real word
        }
    }
    __syncthreads();

    // Key order.  A node's candidates are its arcs' parent runs, each already
    // in descending key order (a parent frontier is, and an arc adds the same
    // weight to all of its points), so within a run the sort words are
    // strictly descending (equal keys: slots ascending).  With few runs, a
    // candidate's place is its index in its own run plus, for every other
    // run, the number of that run's words above its own (a binary search):
    // the order the bitonic network gives, at a fraction of its passes.
    bool merged = false;
    if constexpr (std::is_same<Src, ExpandSrc>::value && !PRESORTED) {
        // an ExpandSrc chunk is a whole node (c.start is 0, c.len its count)
        const int64_t e0 = src.arc_off[c.group];
        const int nr = static_cast<int>(src.arc_off[c.group + 1] - e0);
        if (nr <= kMergeRuns) {
            merged = true;
            auto &rs = L.rs;
            if (tid <= nr) {
                rs[tid] = tid < nr ? static_cast<int>(src.run_off[e0 + tid]) : n;
            }
            __syncthreads();
            uint64_t *out = reinterpret_cast<uint64_t *>(&S[0]); // free until the skyline
            for (int t = tid; t < n; t += NT) {
                const uint64_t w = sk[t];
                int a = 0;
                while (a + 1 < nr && rs[a + 1] <= t) {
                    ++a; // last run starting at or before t (runs may be empty)
                }
                int rank = t - rs[a];
                for (int b = 0; b < nr; ++b) {
                    if (b == a) {
                        continue;
                    }
                    int lo = rs[b], hi = rs[b + 1];
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (sk[mid] > w) {
                            lo = mid + 1;
                        } else {
                            hi = mid;
                        }
                    }
                    rank += lo - rs[b];
                }
                out[rank] = w;
            }
            __syncthreads();
            for (int t = tid; t < n; t += NT) {
                sk[t] = out[t];
            }
            __syncthreads();
        }
    }
    // bitonic sort, descending (skipped when the source is in key order)
    for (int k = 2; !PRESORTED && !merged && k <= npad; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < (npad >> 1); t += NT) {
                const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                const int l = i + j;
                const uint64_t a = sk[i];
                const uint64_t b = sk[l];
                const bool desc = (i & k) == 0;
                if (desc ? (a < b) : (a > b)) {
                    sk[i] = b;
                    sk[l] = a;
                }
            }
            __syncthreads();
        }
    }

    // sort-filter skyline, NT candidates per step
    int nS = 0;
    for (int base = 0; base < n; base += NT) {
        const int i = base + tid;
        const bool valid = i < n;
        const int tcount = min(NT, n - base);
        int32_t m[D];
        if (valid) {
            const int slot = unpack_slot(sk[i]);
#pragma unroll
            for (int o = 0; o < D; ++o) {
                m[o] = cs[o][slot];
                S[(nS + tid) * DP + o] = m[o];
            }
        } else {
#pragma unroll
            for (int o = 0; o < D; ++o) {
                m[o] = 0;
            }
        }
        __syncthreads();
        bool dead = !valid;
        // against survivors of earlier steps
        dead = scan_until_dead<D, DP, false>(S, nS, nS, m, dead, tests);
        // against earlier candidates of this step (slots below the lane's own)
        const int wave_end = __builtin_amdgcn_readfirstlane(
            min(tcount, static_cast<int>((tid | (kWave - 1)))));
        dead = scan_until_dead<D, DP, true>(S + nS * DP, wave_end, min(tid, tcount), m, dead, tests);
        // block_rank's opening barrier ends every lane's reads of S before
        // the survivors are written after it
        int total = 0;
        const int rank = block_rank<NT / kWave>(!dead, wave_tot, &total);
        if (!dead) {
#pragma unroll
            for (int o = 0; o < D; ++o) {
                S[(nS + rank) * DP + o] = m[o];
            }
        }
        nS += total;
        __syncthreads();
    }

    for (int s = tid; s < nS; s += NT) {
        const int64_t pos = c.out_base + s;
#pragma unroll
        for (int o = 0; o < D; ++o) {
            out_pts[o * out_stride + pos] = S[s * DP + o];
        }
        alive[pos] = 1;
    }
    if (tid == 0) {
        chunk_nsurv[ci] = nS;
        chunk_nalive[ci] = nS;
    }
    __syncthreads(); // every read of L is done
}

// One workgroup per listed chunk (every chunk when `which` is null).
template <int D, int CH, class Src, bool PRESORTED, int NT = kBlock>
__global__ __launch_bounds__(NT) void k_filter_chunks(
    const int32_t *__restrict__ which, const ChunkDesc *__restrict__ chunks, Src src,
    int32_t *__restrict__ out_pts,
    int64_t out_stride, uint8_t *__restrict__ alive, int32_t *__restrict__ chunk_nsurv,
    int32_t *__restrict__ chunk_nalive, unsigned long long *__restrict__ pair_tests) {
    __shared__ FilterLds<D, CH, NT> L;
    const int ci = which != nullptr ? which[blockIdx.x] : static_cast<int>(blockIdx.x);
    unsigned long long tests = 0;
    filter_chunk<D, CH, Src, PRESORTED, NT>(ci, chunks, src, out_pts, out_stride, alive,
                                            chunk_nsurv, chunk_nalive, L, tests);
    if (pair_tests != nullptr) {
        atomicAdd(pair_tests, tests);
    }
}

// One tier of a queued expansion pass, read by k_filter_queue from device
// memory, so that the launch's arguments stay the same from one pass to the
// next (a captured graph of the pass's launches can be replayed).
struct QueueTier {
    ExpandSrc src;
    const ChunkDesc *chunks;
    const int32_t *which; // chunks of this tier
    int32_t count;        // entries of which
    int32_t pad;
    int32_t *out_pts;
    int64_t out_stride;
    uint8_t *alive;
    int32_t *nsurv;
    int32_t *nalive;
};

// Persistent chunk filter: a fixed grid of workgroups takes the chunks of a
// tier one at a time from a queue (an atomic counter, *head, zero before the
// launch) until every listed chunk is filtered.  Chunks of many nodes, and
// of many instances when a layer packs several, share one launch.
template <int D, int CH, int NT>
__global__ __launch_bounds__(NT) void k_filter_queue(const QueueTier *__restrict__ tier,
                                                     int32_t *__restrict__ head,
                                                     unsigned long long *__restrict__ pair_tests) {
    __shared__ FilterLds<D, CH, NT> L;
    __shared__ int item;
    const QueueTier t = *tier;
    unsigned long long tests = 0;
    while (true) {
        if (threadIdx.x == 0) {
            item = atomicAdd(head, 1);
        }
        __syncthreads();
        const int k = item;
        __syncthreads(); // every thread has read item before thread 0 writes it again
        if (k >= t.count) {
            break;
        }
        filter_chunk<D, CH, ExpandSrc, false, NT>(t.which[k], t.chunks, t.src, t.out_pts,
                                                  t.out_stride, t.alive, t.nsurv, t.nalive, L,
                                                  tests);
    }
    if (pair_tests != nullptr && tests != 0) {
        atomicAdd(pair_tests, tests);
    }
}

// Block-wide sum of one int per thread (one LDS slot per wave).  Like
// block_rank it opens with a barrier before writing its slot; it also ends
// with one after every wave has read the slots, which orders the caller's
// earlier LDS reads (a tile scan) before whatever it writes next.
__device__ __forceinline__ int block_sum(int v, int *wave_tot) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
    }
    __syncthreads();
    if (lane == 0) {
        wave_tot[wave] = v;
    }
    __syncthreads();
    int sum = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        sum += wave_tot[w];
    }
    __syncthreads();
    return sum;
}

// The first NR of a thread's R rows against tn staged rows (DP ints each);
// with OWN, staged row j (slot TS[j]) only kills rows in later slots.
template <int D, int DP, int NR, int R, bool OWN>
__device__ __forceinline__ void scan_rows(const int32_t *U, const int16_t *TS, int tn,
                                          const int32_t (&m)[R][D], const int (&slot)[R],
                                          bool (&dead)[R]) {
    int32_t acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        acc[r] = INT32_MIN;
    }
#pragma unroll 2
    for (int j = 0; j < tn; ++j) {
        int32_t q[D];
        load_row<D, DP>(&U[j * DP], q);
        const int sj = OWN ? TS[j] : 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            int32_t mg = margin_vals<D>(q, m[r]);
            if (OWN) {
                mg = sj < slot[r] ? mg : INT32_MIN;
            }
            acc[r] = max(acc[r], mg);
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        dead[r] |= acc[r] >= 0;
    }
}

// scan_rows for a row count known only at run time (the same in every lane).
template <int D, int DP, int R, bool OWN, int NR = R>
__device__ __forceinline__ void scan_rows_upto(int rows, const int32_t *U, const int16_t *TS,
                                               int tn, const int32_t (&m)[R][D],
                                               const int (&slot)[R], bool (&dead)[R]) {
    if constexpr (NR >= 1) {
        if (rows == NR) {
            scan_rows<D, DP, NR, R, OWN>(U, TS, tn, m, slot, dead);
        } else {
            scan_rows_upto<D, DP, R, OWN, NR - 1>(rows, U, TS, tn, m, slot, dead);
        }
    }
}

// ---------------------------------------------------------------------------
// k_cross_chunks: the chunks of a group are consecutive blocks of descending
// keys, taken in waves of W chunks; every alive point of a chunk dies when it
// is weakly dominated by
//   (a) a point of the group's front: the final survivors of the earlier
//       waves, kept densely in key order at fv[fbase ..] (count
//       fcount[group_first], filled by k_append_front), or
//   (b) an alive point of an earlier chunk of the same wave, or
//   (c) an alive point in an earlier slot of the chunk itself.
// G workgroups per chunk listed in `which` (workgroup g of chunk which[i] is
// blockIdx.x = i * G + g); workgroup g takes the 256-point tiles g, g + G, ...
// of that work, so one chunk's scan is spread over G compute units.  A point
// that dies is flagged at once, and after every tile each workgroup drops the
// points the others have flagged.  Flags of chunks in the same wave may change
// while they are read: either value is correct (a point killed by cross
// checks is itself dominated by an earlier point that also dominates whatever
// it dominates).  The live points are repacked into fewer rows per thread
// whenever they fit, so the few that scan everything do so without the dead.
// Alive counts are left to k_append_front.
// ---------------------------------------------------------------------------
template <int D, int CH>
__global__ __launch_bounds__(kBlock) void k_cross_chunks(
    const int32_t *__restrict__ which, const ChunkDesc *__restrict__ chunks,
    const int32_t *__restrict__ pts, int64_t stride, uint8_t *alive,
    const int32_t *__restrict__ chunk_nsurv, const int32_t *__restrict__ fv, int64_t fv_stride,
    const int64_t *__restrict__ fcount, int W, int G, int only_wave,
    unsigned long long *__restrict__ pair_tests) {
    constexpr int DP = Pad<D>::DP;
    constexpr int R = CH / kBlock; // rows of points per thread
    constexpr int kU = (kBlock * DP > CH * D) ? kBlock * DP : CH * D;
    __shared__ __align__(16) int32_t U[kU]; // tile of earlier points, or repack buffer
    __shared__ int16_t PS[CH];              // repacked slots
    __shared__ int16_t TS[kBlock];          // slots of a tile of the chunk itself
    __shared__ int wave_tot[kWaves];

    const int me = which != nullptr ? which[blockIdx.x / G] : static_cast<int>(blockIdx.x / G);
    const int g = static_cast<int>(blockIdx.x % G);
    const ChunkDesc c = chunks[me];
    const int k_in = me - c.group_first;
    if (k_in >= c.group_count || (only_wave >= 0 && k_in / W != only_wave)) {
        // a chunk past its group's count (a table sized by a bound), or of
        // another wave (the list holds every wave's chunks)
        return;
    }
    const int nS = chunk_nsurv[me];
    const int tid = threadIdx.x;
    volatile uint8_t *valive = alive;
    const int wave_first = c.group_first + (k_in / W) * W;
    const int64_t fbase = c.start - static_cast<int64_t>(k_in) * CH;
    const int64_t fn = fcount[c.group_first];

    int32_t m[R][D];
    int slot[R]; // -1: no live point
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int s = r * kBlock + tid;
        const bool v = s < nS && valive[c.out_base + s] != 0;
        slot[r] = v ? s : -1;
#pragma unroll
        for (int o = 0; o < D; ++o) {
            m[r][o] = v ? pts[o * stride + c.out_base + s] : INT32_MAX;
        }
    }
    int rows = R;
    unsigned long long tests = 0;
    int live = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        live += slot[r] >= 0 ? 1 : 0;
    }
    int total = block_sum(live, wave_tot);
    bool go = total > 0;

    // tn staged rows in U against the live points (rows of the chunk itself
    // kill only points in later slots); then deaths are flagged, other
    // workgroups' kills taken, and the live points counted and repacked
    auto test_tile = [&](int tn, bool own) {
        bool dead[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            dead[r] = false;
        }
        // rows is the same in every lane: one branch-free scan per row count
        const int urows = __builtin_amdgcn_readfirstlane(rows);
        if (own) {
            scan_rows_upto<D, DP, R, true>(urows, U, TS, tn, m, slot, dead);
        } else {
            scan_rows_upto<D, DP, R, false>(urows, U, TS, tn, m, slot, dead);
        }
        live = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r < rows && slot[r] >= 0) {
                if (pair_tests != nullptr) {
                    tests += static_cast<unsigned long long>(tn);
                }
                if (dead[r] || valive[c.out_base + slot[r]] == 0) {
                    valive[c.out_base + slot[r]] = 0;
                    slot[r] = -1;
                } else {
                    ++live;
                }
            }
        }
        total = block_sum(live, wave_tot); // its barriers also end the tile's reads
        go = total > 0;
        const int need = (total + kBlock - 1) / kBlock;
        if (go && need < rows) {
            int base = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r < rows) {
                    const bool v = slot[r] >= 0;
                    int cnt = 0;
                    const int rk = block_rank(v, wave_tot, &cnt);
                    if (v) {
#pragma unroll
                        for (int o = 0; o < D; ++o) {
                            U[(base + rk) * D + o] = m[r][o];
                        }
                        PS[base + rk] = static_cast<int16_t>(slot[r]);
                    }
                    base += cnt;
                    __syncthreads(); // wave_tot is reused by the next rank
                }
            }
            rows = need;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int s = r * kBlock + tid;
                const bool v = r < rows && s < total;
                slot[r] = v ? PS[s] : -1;
#pragma unroll
                for (int o = 0; o < D; ++o) {
                    m[r][o] = v ? U[s * D + o] : INT32_MAX;
                }
            }
            __syncthreads();
        }
    };

    // (a) the front, dense
    for (int64_t tb = static_cast<int64_t>(g) * kBlock; go && tb < fn;
         tb += static_cast<int64_t>(G) * kBlock) {
        const int tn = static_cast<int>(min(static_cast<int64_t>(kBlock), fn - tb));
        if (tid < tn) {
#pragma unroll
            for (int o = 0; o < D; ++o) {
                U[tid * DP + o] = fv[o * fv_stride + fbase + tb + tid];
            }
        }
        __syncthreads();
        test_tile(tn, false);
    }
    // (b) earlier chunks of this wave, alive points only
    int idx = 0;
    for (int other = wave_first; go && other < me; ++other) {
        const ChunkDesc oc = chunks[other];
        const int on = chunk_nsurv[other];
        for (int tb = 0; go && tb < on; tb += kBlock, ++idx) {
            if (idx % G != g) {
                continue;
            }
            const int i = tb + tid;
            const bool a = i < on && valive[oc.out_base + i] != 0;
            int32_t q[D];
#pragma unroll
            for (int o = 0; o < D; ++o) {
                q[o] = a ? pts[o * stride + oc.out_base + i] : 0;
            }
            int tn = 0;
            const int pos = block_rank(a, wave_tot, &tn);
            if (a) {
#pragma unroll
                for (int o = 0; o < D; ++o) {
                    U[pos * DP + o] = q[o];
                }
            }
            __syncthreads();
            test_tile(tn, false);
        }
    }
    // (c) earlier points of the chunk itself, alive ones only
    for (int tb = 0; go && tb < nS; tb += kBlock, ++idx) {
        if (idx % G != g) {
            continue;
        }
        const int i = tb + tid;
        const bool a = i < nS && valive[c.out_base + i] != 0;
        int32_t q[D];
#pragma unroll
        for (int o = 0; o < D; ++o) {
            q[o] = a ? pts[o * stride + c.out_base + i] : 0;
        }
        int tn = 0;
        const int pos = block_rank(a, wave_tot, &tn);
        if (a) {
#pragma unroll
            for (int o = 0; o < D; ++o) {
                U[pos * DP + o] = q[o];
            }
            TS[pos] = static_cast<int16_t>(i);
        }
        __syncthreads();
        test_tile(tn, true);
    }
    if (pair_tests != nullptr) {
        atomicAdd(pair_tests, tests);
    }
}

// ---------------------------------------------------------------------------
// k_stage_chunks: the candidates of the listed chunks copied from a key-ordered
// source to the pass buffer, all alive (no local filter: k_cross_chunks checks
// a chunk against its own earlier points as well).
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_stage_chunks(const int32_t *__restrict__ which,
                                                         const ChunkDesc *__restrict__ chunks,
                                                         SortedSrc src, int32_t *__restrict__ out,
                                                         int64_t stride, uint8_t *__restrict__ alive,
                                                         int32_t *__restrict__ chunk_nsurv,
                                                         int32_t *__restrict__ chunk_nalive) {
    const int me = which != nullptr ? which[blockIdx.x] : static_cast<int>(blockIdx.x);
    const ChunkDesc c = chunks[me];
    for (int s = threadIdx.x; s < c.len; s += kBlock) {
#pragma unroll
        for (int o = 0; o < D; ++o) {
            out[o * stride + c.out_base + s] = src.pts[o * src.stride + c.start + s];
        }
        alive[c.out_base + s] = 1;
    }
    if (threadIdx.x == 0) {
        chunk_nsurv[me] = c.len;
        chunk_nalive[me] = c.len;
    }
}

// ---------------------------------------------------------------------------
// k_append_front: after the cross check of a wave, every chunk of the wave
// (listed in `which`) counts its alive points into chunk_nalive and appends
// them, in slot order, to its group's front fv[fbase + at ..], where at =
// fcount_in[g] plus the alive points of the wave's earlier chunks of the
// group, so the front stays in key order; the wave's last chunk of each group
// writes the new count to fcount_out[g] (g = group_first).
// ---------------------------------------------------------------------------
template <int D, int CH>
__global__ __launch_bounds__(kBlock) void k_append_front(
    const int32_t *__restrict__ which, const ChunkDesc *__restrict__ chunks,
    const int32_t *__restrict__ pts, int64_t stride, const uint8_t *__restrict__ alive,
    const int32_t *__restrict__ chunk_nsurv, int32_t *__restrict__ chunk_nalive,
    int32_t *__restrict__ fv, int64_t fv_stride, const int64_t *__restrict__ fcount_in,
    int64_t *__restrict__ fcount_out, int W, int only_wave, int d,
    unsigned long long *__restrict__ bytes) {
    __shared__ int wave_tot[kWaves];
    const int me = which != nullptr ? which[blockIdx.x] : static_cast<int>(blockIdx.x);
    const ChunkDesc c = chunks[me];
    const int k_in = me - c.group_first;
    if (k_in >= c.group_count || (only_wave >= 0 && k_in / W != only_wave)) {
        return;
    }
    const int wave_first = c.group_first + (k_in / W) * W;
    const int wave_last = min(wave_first + W, c.group_first + c.group_count) - 1;
    const int64_t fbase = c.start - static_cast<int64_t>(k_in) * CH;
    int64_t at = fcount_in[c.group_first];
    for (int other = wave_first; other < me; ++other) {
        const ChunkDesc oc = chunks[other];
        const int on = chunk_nsurv[other];
        for (int tb = 0; tb < on; tb += kBlock) {
            const int i = tb + threadIdx.x;
            int cnt = 0;
            block_rank(i < on && alive[oc.out_base + i] != 0, wave_tot, &cnt);
            at += cnt;
            __syncthreads();
        }
    }
    const int nS = chunk_nsurv[me];
    int written = 0;
    for (int tb = 0; tb < nS; tb += kBlock) {
        const int s = tb + threadIdx.x;
        const bool a = s < nS && alive[c.out_base + s] != 0;
        int cnt = 0;
        const int rk = block_rank(a, wave_tot, &cnt);
        if (a) {
#pragma unroll
            for (int o = 0; o < D; ++o) {
                fv[o * fv_stride + fbase + at + written + rk] = pts[o * stride + c.out_base + s];
            }
        }
        written += cnt;
        __syncthreads(); // wave_tot is reused by the next rank
    }
    if (threadIdx.x == 0) {
        chunk_nalive[me] = written;
        if (me == wave_last) {
            fcount_out[c.group_first] = at + written;
        }
        if (bytes != nullptr) {
            // compulsory bytes of the wave: each point read once, its flag
            // once, each survivor appended once, and (once per group and
            // wave, by its last chunk) the front read once
            unsigned long long b = 4ull * d * (static_cast<unsigned long long>(c.len) + written) +
                                   static_cast<unsigned long long>(c.len);
            if (me == wave_last) {
                b += 4ull * d * static_cast<unsigned long long>(fcount_in[c.group_first]);
            }
            atomicAdd(bytes, b);
        }
    }
}

// ---------------------------------------------------------------------------
// k_compact_chunks: alive survivors of each chunk to dense output at
// chunk_out[c] (order kept).
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_compact_chunks(
    const ChunkDesc *__restrict__ chunks, const int32_t *__restrict__ chunk_nsurv,
    const int64_t *__restrict__ chunk_out, const int32_t *__restrict__ src, int64_t src_stride,
    const uint8_t *__restrict__ alive, int32_t *__restrict__ dst, int64_t dst_stride) {
    __shared__ int wave_tot[kWaves];
    const ChunkDesc c = chunks[blockIdx.x];
    const int nS = chunk_nsurv[blockIdx.x];
    int64_t out = chunk_out[blockIdx.x];
    for (int base = 0; base < nS; base += kBlock) {
        const int s = base + threadIdx.x;
        const bool a = s < nS && alive[c.out_base + s] != 0;
        int total = 0;
        const int rank = block_rank(a, wave_tot, &total);
        if (a) {
#pragma unroll
            for (int o = 0; o < D; ++o) {
                dst[o * dst_stride + out + rank] = src[o * src_stride + c.out_base + s];
            }
        }
        out += total;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Knapsack state dominance (method 1): a label of a target node is dropped
// when some label of one of its (up to two) comparison nodes is >= it in every
// objective.  Reads the unfiltered layer; writes keep flags per point.
// ---------------------------------------------------------------------------
// One target per workgroup: every workgroup's (count null), or the first
// *count (the grid may be larger; the rest exit at once).
template <int D>
__global__ __launch_bounds__(kBlock) void k_state_dominance(
    const int32_t *__restrict__ targets, // [n][4]: node, cmpA, cmpB (-1 = none), first label
    const int64_t *__restrict__ count,
    const int64_t *__restrict__ off, const int32_t *__restrict__ pts, int64_t stride,
    uint8_t *__restrict__ keep, unsigned long long *__restrict__ pair_tests) {
    constexpr int DP = Pad<D>::DP;
    __shared__ __align__(16) int32_t T[kBlock * DP];
    unsigned long long tests = 0;
    const int64_t ti = blockIdx.x;
    if (count != nullptr && ti >= *count) {
        return;
    }
    const int32_t *tg = targets + 4 * ti;
    const int node = tg[0];
    const int64_t b0 = off[node];
    const int64_t n = off[node + 1] - b0;
    const int64_t base = tg[3];
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n;
    int32_t m[D];
#pragma unroll
    for (int o = 0; o < D; ++o) {
        m[o] = valid ? pts[o * stride + b0 + i] : 0;
    }
    bool dead = !valid;
    volatile const uint8_t *vkeep = keep;
    // frontiers are in descending key order: the last label of this block has
    // the smallest key, and a comparison label can only be >= a label whose
    // key does not exceed its own
    const int64_t last = min(base + kBlock, n) - 1;
    const int64_t block_min = key_at<D>(pts, stride, b0 + last);
    for (int cidx = 1; cidx <= 2; ++cidx) {
        const int cmp = tg[cidx];
        if (cmp < 0) {
            continue;
        }
        const int64_t c0 = off[cmp];
        const int64_t cn = off[cmp + 1] - c0;
        // the gridDim.y workgroups of a target take every gridDim.y-th tile
        // of the comparison labels, and each drops the labels the others
        // have flagged (flags only go from 1 to 0)
        const int64_t first = static_cast<int64_t>(blockIdx.y) * kBlock;
        const int64_t step = static_cast<int64_t>(gridDim.y) * kBlock;
        for (int64_t tb = first; tb < cn; tb += step) {
            if (key_at<D>(pts, stride, c0 + tb) < block_min) {
                break;
            }
            if (gridDim.y > 1 && valid && !dead) {
                dead = vkeep[b0 + i] == 0;
            }
            if (!__syncthreads_or(!dead)) {
                break;
            }
            const int tn = static_cast<int>(min(static_cast<int64_t>(kBlock), cn - tb));
            if (static_cast<int>(threadIdx.x) < tn) {
#pragma unroll
                for (int o = 0; o < D; ++o) {
                    T[threadIdx.x * DP + o] = pts[o * stride + c0 + tb + threadIdx.x];
                }
            }
            __syncthreads();
            dead = scan_until_dead<D, DP, false>(T, tn, tn, m, dead, tests);
        }
    }
    if (valid && dead) {
        keep[b0 + i] = 0;
    }
    if (pair_tests != nullptr && tests != 0) {
        atomicAdd(pair_tests, tests);
    }
}

// Per-node copy of kept points into a dense frontier with new offsets.
template <int D>
__global__ __launch_bounds__(kBlock) void k_compact_nodes(
    const int64_t *__restrict__ off_in, const int32_t *__restrict__ src, int64_t src_stride,
    const uint8_t *__restrict__ keep, const int64_t *__restrict__ off_out,
    int32_t *__restrict__ dst, int64_t dst_stride) {
    __shared__ int wave_tot[kWaves];
    const int node = blockIdx.x;
    const int64_t b0 = off_in[node];
    const int64_t n = off_in[node + 1] - b0;
    int64_t out = off_out[node];
    for (int64_t base = 0; base < n; base += kBlock) {
        const int64_t i = base + threadIdx.x;
        const bool a = i < n && keep[b0 + i] != 0;
        int total = 0;
        const int rank = block_rank(a, wave_tot, &total);
        if (a) {
#pragma unroll
            for (int o = 0; o < D; ++o) {
                dst[o * dst_stride + out + rank] = src[o * src_stride + b0 + i];
            }
        }
        out += total;
        __syncthreads();
    }
}

// Per-node count of kept points.
__global__ __launch_bounds__(kBlock) void k_count_kept(const int64_t *__restrict__ off,
                                                       const uint8_t *__restrict__ keep,
                                                       int64_t *__restrict__ counts,
                                                       int64_t n_nodes) {
    __shared__ int wave_tot[kWaves];
    const int node = blockIdx.x;
    const int64_t b0 = off[node];
    const int64_t n = off[node + 1] - b0;
    int64_t kept = 0;
    for (int64_t base = 0; base < n; base += kBlock) {
        const int64_t i = base + threadIdx.x;
        const bool a = i < n && keep[b0 + i] != 0;
        int total = 0;
        block_rank(a, wave_tot, &total);
        kept += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counts[node] = kept;
    }
}

// Y gets the points X[q], q < *n_in, whose flag is still 1 (order not
// kept), and *n_out counts them; one atomic per wave (grid-stride).
template <int D>
__global__ __launch_bounds__(kBlock) void k_flag_append(
    const int32_t *__restrict__ X, int64_t xstride, const unsigned long long *__restrict__ n_in,
    const uint8_t *__restrict__ flag, int32_t *__restrict__ Y, int64_t ystride,
    unsigned long long *__restrict__ n_out) {
    const int64_t n = static_cast<int64_t>(*n_in);
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (kWave - lane));
    const int64_t step = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t b0 = static_cast<int64_t>(blockIdx.x) * kBlock; b0 < n; b0 += step) {
        const int64_t q = b0 + threadIdx.x;
        const bool keep = q < n && flag[q] != 0;
        const uint64_t mask = __ballot(keep);
        if (mask == 0) {
            continue;
        }
        unsigned long long base = 0;
        if (lane == 0) {
            base = atomicAdd(n_out, static_cast<unsigned long long>(__popcll(mask)));
        }
        base = __shfl(base, 0);
        if (keep) {
            const int64_t at = static_cast<int64_t>(base) + __popcll(mask & below);
#pragma unroll
            for (int o = 0; o < D; ++o) {
                Y[o * ystride + at] = X[o * xstride + q];
            }
        }
    }
}

// Fill a byte array (used instead of memset so the engine never records a
// memset node when captured into a graph).
__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = v;
    }
}

// Per-chunk count of alive survivors.
__global__ __launch_bounds__(kBlock) void k_count_chunks(const ChunkDesc *__restrict__ chunks,
                                                         const int32_t *__restrict__ chunk_nsurv,
                                                         const uint8_t *__restrict__ alive,
                                                         int32_t *__restrict__ chunk_nalive) {
    __shared__ int wave_tot[kWaves];
    const ChunkDesc c = chunks[blockIdx.x];
    const int nS = chunk_nsurv[blockIdx.x];
    int kept = 0;
    for (int base = 0; base < nS; base += kBlock) {
        const int s = base + threadIdx.x;
        const bool a = s < nS && alive[c.out_base + s] != 0;
        int total = 0;
        block_rank(a, wave_tot, &total);
        kept += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        chunk_nalive[blockIdx.x] = kept;
    }
}

// ---------------------------------------------------------------------------
// Big groups: runs in key order merged pairwise, one pass per level.
//
// Elements of several groups lie in one buffer; group g owns runs
// [grp_run[g], grp_run[g+1]) and run r occupies [run_start[r], run_start[r+1]).
// In pass p the runs of a group are taken in segments of 2^p runs; segment s is
// merged with segment s^1 (stable: on equal keys the lower segment first).
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_gather_runs(
    const int64_t *__restrict__ big_base, // [G + 1] element offsets of the groups
    const int32_t *__restrict__ big_node, // [G] node of each group
    int32_t G, ExpandSrc src, int32_t *__restrict__ out, int64_t out_stride) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= big_base[G]) {
        return;
    }
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (big_base[mid] <= q) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    ChunkDesc c;
    c.start = 0;
    c.group = big_node[lo];
    int32_t p[D];
    load_candidate<D>(src, c, q - big_base[lo], p);
#pragma unroll
    for (int o = 0; o < D; ++o) {
        out[o * out_stride + q] = p[o];
    }
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_merge_pass(
    const int32_t *__restrict__ in, int32_t *__restrict__ out, int64_t stride,
    const int64_t *__restrict__ run_start, const int32_t *__restrict__ grp_run, int32_t G,
    int pass, int64_t total) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    // elements end at the table's sentinel run_start[grp_run[G]]; `total`
    // bounds the grid (it may exceed the elements when they are counted on
    // the device)
    if (q >= total || q >= run_start[grp_run[G]]) {
        return;
    }
    // group of q: last g with run_start[grp_run[g]] <= q
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (run_start[grp_run[mid]] <= q) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    const int64_t gr0 = grp_run[lo];
    const int64_t nr = grp_run[lo + 1] - gr0;
    // run of q: last r in [gr0, gr0 + nr) with run_start[r] <= q
    int64_t rl = gr0, rh = gr0 + nr - 1;
    while (rl < rh) {
        const int64_t mid = (rl + rh + 1) >> 1;
        if (run_start[mid] <= q) {
            rl = mid;
        } else {
            rh = mid - 1;
        }
    }
    const int64_t seg = (rl - gr0) >> pass;
    const int64_t partner = seg ^ 1;
    auto run_pos = [&](int64_t r) { return run_start[gr0 + (r < nr ? r : nr)]; };
    const int64_t a0 = run_pos(seg << pass);
    const int64_t b0 = run_pos(partner << pass);
    const int64_t b1 = run_pos((partner + 1) << pass);
    int32_t v[D];
    int64_t key = 0;
#pragma unroll
    for (int o = 0; o < D; ++o) {
        v[o] = in[o * stride + q];
        key += v[o];
    }
    int64_t rank = q - a0;
    if (b1 > b0) {
        // partner elements placed before q: keys > key (q in the lower
        // segment) or keys >= key (q in the upper segment)
        const bool lower = (seg & 1) == 0;
        int64_t l = b0, h = b1;
        while (l < h) {
            const int64_t mid = (l + h) >> 1;
            const int64_t km = key_at<D>(in, stride, mid);
            const bool before = lower ? (km > key) : (km >= key);
            if (before) {
                l = mid + 1;
            } else {
                h = mid;
            }
        }
        rank += l - b0;
    }
    const int64_t dst = (a0 < b0 ? a0 : b0) + rank;
#pragma unroll
    for (int o = 0; o < D; ++o) {
        out[o * stride + dst] = v[o];
    }
}


// ---------------------------------------------------------------------------
// k_join_products: the n products of a coupling batch, written densely (SoA,
// stride xstride) in product order.
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_join_products(JoinSrc src, int64_t n,
                                                          int32_t *__restrict__ X,
                                                          int64_t xstride) {
    ChunkDesc c{};
    const int64_t step = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; q < n; q += step) {
        int32_t p[D];
        load_candidate<D>(src, c, q, p);
#pragma unroll
        for (int o = 0; o < D; ++o) {
            X[o * xstride + q] = p[o];
        }
    }
}

// ---------------------------------------------------------------------------
// k_refill_vs_block: points X[0 .. *n_in) die when some point of the block
// R[lo .. hi) (descending key order) is >= them in every objective (equal
// points die).  Survivors are appended to Y (order not kept) and counted in
// *n_out.
//
// One wave per workgroup.  Each lane holds one point; the wave walks the
// block round-robin in sub-tiles of kSub rows (64-row tiles staged in LDS, so
// every lane reads the same row, a broadcast), and after every sub-tile a lane
// whose point has died, or has met the whole block since it was taken, takes
// the next point of the wave's stream.  One hard point therefore never holds
// 63 finished lanes, which is what a check in lockstep over fixed groups pays:
// most points meet a dominator within a few rows, a few need thousands.
// kSub = 64 (a whole tile between refills) measured faster than 16 or 32:
// the refill's bookkeeping costs more than the tests a dead lane idles over.
// Wave w's stream is the 64-point chunks w, w + G, w + 2G, ... (G waves), so
// taking points needs no atomics; survivors are staged in LDS and written
// out kStage at a time with one atomic each.
// A point whose key is above the block's largest key meets no test: no point
// of the block can be >= it.
// ---------------------------------------------------------------------------
//
// FLAGS mode: X keeps its order; a point that dies gets flag[q] = 0 (points
// whose flag is 0 already are skipped), and nothing is appended.  In FLAGS
// mode an empty block leaves every point as it is; otherwise a launch over an
// empty block appends nothing, so its caller must not rely on the survivors
// (the engine never launches one there).
template <int D, bool FLAGS>
__global__ __launch_bounds__(kWave) void k_refill_vs_block(
    const int32_t *__restrict__ X, int64_t xstride, const unsigned long long *__restrict__ n_in,
    const int32_t *__restrict__ R, int64_t rstride, int64_t lo_arg, int64_t hi_arg,
    const int64_t *__restrict__ hi_dev,
    int32_t *__restrict__ Y, int64_t ystride, unsigned long long *__restrict__ n_out,
    uint8_t *__restrict__ flag, unsigned long long *__restrict__ pair_tests,
    int64_t blen = 0) {
    constexpr int DP = Pad<D>::DP;
    constexpr int kSub = 64;                  // rows between two refills
    constexpr int kSubPerTile = kWave / kSub; // sub-tiles per staged tile
    constexpr int kStage = 128;               // survivors staged per flush
    __shared__ __align__(16) int32_t T[kWave * DP];
    __shared__ int32_t SB[kStage * D];
    const int lane = threadIdx.x;
    const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (kWave - lane));
    const int64_t n = static_cast<int64_t>(*n_in);
    const int64_t G = gridDim.x;
    const int64_t w = blockIdx.x;
    // the block ends at hi_arg, or at *hi_dev when that is set (a block whose
    // length is known on the device only).  With blen > 0 (FLAGS only), row
    // y of the grid takes the y-th block of blen rows from lo_arg instead.
    int64_t lo = lo_arg;
    int64_t hi = hi_dev != nullptr ? *hi_dev : hi_arg;
    if (FLAGS && blen > 0) {
        lo = lo_arg + static_cast<int64_t>(blockIdx.y) * blen;
        hi = min(hi, lo + blen);
    }
    if (w * kWave >= n || hi <= lo) {
        return; // an empty block kills nothing (flags stay, nothing appended)
    }
    const int64_t len = hi - lo;
    const int nsub = static_cast<int>((len + kSub - 1) / kSub);
    const int ntile = static_cast<int>((len + kWave - 1) / kWave);
    const int64_t top_key = key_at<D>(R, rstride, lo);
    int sub = static_cast<int>(w % ntile) * kSubPerTile;
    if (sub >= nsub) {
        sub = 0;
    }
    int loaded = -1;       // tile staged in T
    int64_t t_next = 0;    // next position in the wave's stream
    bool exhausted = false;
    bool holds = false;    // lane holds a point
    int64_t held = 0;      // its index in X
    int left = 0;          // sub-tiles the lane's point has still to meet
    int32_t m[D];
#pragma unroll
    for (int o = 0; o < D; ++o) {
        m[o] = 0;
    }
    int sb = 0;
    unsigned long long tests = 0;

    auto flush = [&]() {
        __syncthreads();
        unsigned long long base = 0;
        if (lane == 0) {
            base = atomicAdd(n_out, static_cast<unsigned long long>(sb));
        }
        base = __shfl(base, 0);
        for (int i = lane; i < sb; i += kWave) {
#pragma unroll
            for (int o = 0; o < D; ++o) {
                Y[o * ystride + static_cast<int64_t>(base) + i] = SB[i * D + o];
            }
        }
        __syncthreads();
        sb = 0;
    };
    auto stage = [&](bool keep) {
        if (FLAGS) {
            return;
        }
        const uint64_t mask = __ballot(keep);
        if (mask == 0) {
            return;
        }
        const int cnt = __popcll(mask);
        if (sb + cnt > kStage) {
            flush();
        }
        if (keep) {
            const int pos = sb + __popcll(mask & below);
#pragma unroll
            for (int o = 0; o < D; ++o) {
                SB[pos * D + o] = m[o];
            }
        }
        sb += cnt;
    };

    while (true) {
        // refill lanes without a point; points above the block's top key pass
        while (!exhausted) {
            const bool need = !holds;
            const uint64_t mask = __ballot(need);
            if (mask == 0) {
                break;
            }
            bool pass = false;
            bool past = false;
            if (need) {
                const int64_t t = t_next + __popcll(mask & below);
                const int64_t q = (w + (t >> 6) * G) * kWave + (t & (kWave - 1));
                if (q < n && FLAGS && flag[q] == 0) {
                    // already dead: nothing to do, the lane stays empty
                } else if (q < n) {
                    int64_t k = 0;
#pragma unroll
                    for (int o = 0; o < D; ++o) {
                        m[o] = X[o * xstride + q];
                        k += m[o];
                    }
                    if (k > top_key) {
                        pass = true;
                    } else {
                        holds = true;
                        held = q;
                        left = nsub;
                    }
                } else {
                    past = true;
                }
            }
            t_next += __popcll(mask);
            exhausted = __ballot(past) != 0;
            stage(pass);
        }
        if (__ballot(holds) == 0) {
            break;
        }
        const int tile = sub / kSubPerTile;
        if (tile != loaded) {
            __syncthreads();
            const int64_t tb = lo + static_cast<int64_t>(tile) * kWave;
            if (tb + lane < hi) {
#pragma unroll
                for (int o = 0; o < D; ++o) {
                    T[lane * DP + o] = R[o * rstride + tb + lane];
                }
            }
            __syncthreads();
            loaded = tile;
        }
        const int j0 = (sub - tile * kSubPerTile) * kSub;
        const int je = static_cast<int>(min(static_cast<int64_t>(j0 + kSub),
                                            hi - lo - static_cast<int64_t>(tile) * kWave));
        bool dead = false;
        if (holds) {
            int32_t acc = INT32_MIN;
            if (je - j0 == kSub) {
#pragma unroll
                for (int j = 0; j < kSub; ++j) {
                    acc = max(acc, margin_row<D, DP>(&T[(j0 + j) * DP], m));
                }
            } else {
                for (int j = j0; j < je; ++j) {
                    acc = max(acc, margin_row<D, DP>(&T[j * DP], m));
                }
            }
            dead = acc >= 0;
            tests += static_cast<unsigned long long>(je - j0);
        }
        bool survive = false;
        if (holds) {
            --left;
            if (dead) {
                holds = false;
                if (FLAGS) {
                    flag[held] = 0;
                }
            } else if (left == 0) {
                survive = true;
                holds = false;
            }
        }
        stage(survive);
        sub = (sub + 1 == nsub) ? 0 : sub + 1;
    }
    if (!FLAGS && sb > 0) {
        flush();
    }
    if (pair_tests != nullptr) {
        for (int off = kWave / 2; off > 0; off >>= 1) {
            tests += __shfl_xor(tests, off);
        }
        if (lane == 0) {
            atomicAdd(pair_tests, tests);
        }
    }
}

// ---------------------------------------------------------------------------
// Packed layers (serving): the layers of several in-flight instances taken as
// one.  Part i contributes its layer's arcs at [arc_base, arc_base + arcs) of
// the packed layer, their parents moved by node_add (the first node of the
// instance's current layer in the packed frontier).
// ---------------------------------------------------------------------------
struct PackPart {
    const int32_t *src; // [arcs] parent of each arc, layer-local
    const int32_t *w;   // SoA [d][arcs]
    int64_t arcs;
    int64_t arc_base;
    int64_t node_add;
};

__global__ __launch_bounds__(kBlock) void k_pack_arcs(const PackPart *__restrict__ parts,
                                                      int32_t nparts, int64_t total, int d,
                                                      int32_t *__restrict__ src_out,
                                                      int32_t *__restrict__ w_out) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < total;
         e += step) {
        int lo = 0, hi = nparts - 1; // last part whose arcs start at or before e
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (parts[mid].arc_base <= e) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        const PackPart p = parts[lo];
        const int64_t k = e - p.arc_base;
        src_out[e] = static_cast<int32_t>(p.src[k] + p.node_add);
        for (int o = 0; o < d; ++o) {
            w_out[o * total + e] = p.w[o * p.arcs + k];
        }
    }
}

} // namespace pmdd_k
