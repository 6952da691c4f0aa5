// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference's OpenMP enumerators, used by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
// checker and as the timed CPU baseline.  Nothing in the product links,
// imports or calls this file.
//
// What it restates (files under src/pmdd/ of the reference):
//   Frontier::merge_shift  ParetoFrontier::merge(frontier, shift)
//                          enum/pareto_frontier.hpp:237-303
//   Frontier::merge        ParetoFrontier::merge(frontier)  pareto_frontier.hpp:169-232
//   Frontier::remove_empty pareto_frontier.hpp:317-345
//   Frontier::convolute    pareto_frontier.hpp:353-365
//   topdown                topdown_cpu_enumerate / topdown_mdd_cpu_enumerate
//                          enum/cpu/enum.cpp:14-121, expand_layer_topdown_cpu
//                          enum/cpu/topdown.cpp:12-33
//   bottomup               bottomup_cpu_enumerate enum/cpu/enum.cpp:123-229,
//                          expand_layer_bottomup(_mdd) enum/cpu/bottomup.cpp:12-83
//   coupled                coupled_cpu_enumerate / coupled_mdd_cpu_enumerate
//                          enum/cpu/enum.cpp:231-554; layer values
//                          enum/cpu/couple.cpp:15-48; CompareNode
//                          enum/cpu/cpu_helpers.hpp:99-104; tree reduction
//                          cpu_helpers.hpp:49-88
//   knapsack_state_dominance  filter_dominance_knapsack_cpu
//                          enum/cpu/dominance.cpp:24-86
// The DOMINATED sentinel (-9999999, pareto_frontier.hpp:8) is replaced by an
// explicit per-point flag; nothing else changes.
//
// Brute force (independent of the diagrams): all feasible knapsack subsets,
// all independent sets of the set-packing conflict graph, all tours; the
// nondominated set of their objective vectors is the closed form every method
// must reproduce.  The parsers here are written separately from the product
// parsers so that a parser mistake in either shows up as a mismatch.
//
// Parity pinning of each function (what checks it) is listed in
// tests/test_oracle.py.

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include <omp.h>

#include "../include/pmdd_engine.h"

namespace {

// ---------------------------------------------------------------------------
// Frontier: flat point list with the reference's incremental merge rules.
// ---------------------------------------------------------------------------
struct Frontier {
    int d = 0;
    std::vector<int32_t> sols;
    std::vector<uint8_t> dead; // one flag per point, replaces the sentinel

    explicit Frontier(int d_) : d(d_) {}
    int64_t size() const { return static_cast<int64_t>(dead.size()); }

    void push(const int32_t *p) {
        sols.insert(sols.end(), p, p + d);
        dead.push_back(0);
    }

    // pareto_frontier.hpp:317-345: move the last live point into each dead slot.
    void remove_empty() {
        if (dead.empty()) {
            return;
        }
        int64_t last = size() - 1;
        while (last >= 0 && dead[last]) {
            --last;
        }
        if (last < 0) {
            sols.clear();
            dead.clear();
            return;
        }
        for (int64_t i = 0; i < last; ++i) {
            if (dead[i]) {
                std::copy(sols.begin() + last * d, sols.begin() + last * d + d,
                          sols.begin() + i * d);
                dead[i] = 0;
                dead[last] = 1;
                --last;
                while (dead[last]) {
                    --last;
                }
            }
        }
        sols.resize((last + 1) * d);
        dead.resize(last + 1);
    }

    // pareto_frontier.hpp:237-303.  Each shifted point is compared with the
    // points present before this call: dropped if weakly dominated (ties
    // drop); when it dominates a point it overwrites the first such point and
    // marks the later ones dead; appended if it was neither.
    void merge_shift(const Frontier &f, const int32_t *shift) {
        const int64_t end = size();
        std::vector<int32_t> aux(d);
        for (int64_t j = 0; j < f.size(); ++j) {
            for (int o = 0; o < d; ++o) {
                aux[o] = f.sols[j * d + o] + shift[o];
            }
            bool must_add = true;
            for (int64_t i = 0; i < end; ++i) {
                if (dead[i]) {
                    continue;
                }
                bool dominates = true, dominated = true;
                for (int o = 0; o < d && (dominates || dominated); ++o) {
                    dominates &= (aux[o] >= sols[i * d + o]);
                    dominated &= (aux[o] <= sols[i * d + o]);
                }
                if (dominated) {
                    must_add = false;
                    break;
                } else if (dominates) {
                    if (must_add) {
                        std::copy(aux.begin(), aux.end(), sols.begin() + i * d);
                        must_add = false;
                    } else {
                        dead[i] = 1;
                    }
                }
            }
            if (must_add) {
                push(aux.data());
            }
        }
        remove_empty();
    }

    // pareto_frontier.hpp:169-232 (no shift)
    void merge(const Frontier &f) {
        const int64_t end = size();
        bool modified = false;
        for (int64_t j = 0; j < f.size(); ++j) {
            const int32_t *q = &f.sols[j * d];
            bool must_add = true;
            for (int64_t i = 0; i < end; ++i) {
                if (dead[i]) {
                    continue;
                }
                bool dominates = true, dominated = true;
                for (int o = 0; o < d && (dominates || dominated); ++o) {
                    dominates &= (q[o] >= sols[i * d + o]);
                    dominated &= (q[o] <= sols[i * d + o]);
                }
                if (dominated) {
                    must_add = false;
                    break;
                } else if (dominates) {
                    if (must_add) {
                        std::copy(q, q + d, sols.begin() + i * d);
                        must_add = false;
                    } else {
                        dead[i] = 1;
                        modified = true;
                    }
                }
            }
            if (must_add) {
                push(q);
            }
        }
        if (modified) {
            remove_empty();
        }
    }

    // pareto_frontier.hpp:353-365: shift the larger frontier by each point of
    // the smaller one.
    void convolute(const Frontier &a, const Frontier &b) {
        if (a.sols.size() < b.sols.size()) {
            for (int64_t j = 0; j < a.size(); ++j) {
                merge_shift(b, &a.sols[j * d]);
            }
        } else {
            for (int64_t j = 0; j < b.size(); ++j) {
                merge_shift(a, &b.sols[j * d]);
            }
        }
    }

    int64_t sum() const {
        int64_t s = 0;
        for (int32_t x : sols) {
            s += x;
        }
        return s;
    }
};

// Per-layer adjacency derived from the view.
struct Graph {
    const pmdd_dd_view *v;
    int L, d;
    std::vector<int64_t> layer_size;
    // bottom-up: out-arcs per global node, as arc ids
    std::vector<int64_t> out_off;
    std::vector<int64_t> out_arc;

    explicit Graph(const pmdd_dd_view *view) : v(view), L(view->num_layers), d(view->d) {
        layer_size.resize(L);
        for (int l = 0; l < L; ++l) {
            layer_size[l] = v->layer_off[l + 1] - v->layer_off[l];
        }
        const int64_t total = v->layer_off[L];
        const int64_t A = v->in_off[total];
        out_off.assign(total + 1, 0);
        // tail global id of arc a
        std::vector<int64_t> tail(A);
        for (int l = 1; l < L; ++l) {
            for (int64_t g = v->layer_off[l]; g < v->layer_off[l + 1]; ++g) {
                for (int64_t a = v->in_off[g]; a < v->in_off[g + 1]; ++a) {
                    tail[a] = v->layer_off[l - 1] + v->arc_src[a];
                    out_off[tail[a] + 1]++;
                }
            }
        }
        for (int64_t g = 0; g < total; ++g) {
            out_off[g + 1] += out_off[g];
        }
        out_arc.assign(A, 0);
        std::vector<int64_t> pos(out_off.begin(), out_off.end() - 1);
        for (int64_t a = 0; a < A; ++a) {
            out_arc[pos[tail[a]]++] = a;
        }
        // out-arcs of a node in the reference's bottom-up merge order: BDD
        // 1-arc then 0-arc (bottomup.cpp:23-31); MDD by arc value, the order
        // the TSP constructor adds them (tsp_mdd.cpp:140-170)
        for (int64_t g = 0; g < total; ++g) {
            std::stable_sort(out_arc.begin() + out_off[g], out_arc.begin() + out_off[g + 1],
                             [&](int64_t x, int64_t y) {
                                 if (v->kind == PMDD_DD_BDD) {
                                     return v->arc_val[x] > v->arc_val[y];
                                 }
                                 return v->arc_val[x] < v->arc_val[y];
                             });
        }
    }

    int64_t gid(int l, int64_t i) const { return v->layer_off[l] + i; }
    int64_t in_deg(int l, int64_t i) const {
        const int64_t g = gid(l, i);
        return v->in_off[g + 1] - v->in_off[g];
    }
    int64_t out_deg(int l, int64_t i) const {
        const int64_t g = gid(l, i);
        return out_off[g + 1] - out_off[g];
    }
};

struct OracleRun {
    Graph g;
    int problem_type;
    int sd;
    int threads;
    int score_rule; // 0: per-node truncation (CPU), 1: truncate the layer sum (GPU)
    std::vector<std::vector<Frontier *>> td, bu;
    // stats
    int64_t cand_total = 0, cand_peak = 0, surv_total = 0, surv_peak = 0;
    int64_t join_products = 0, filtered = 0;
    int64_t max_node_frontier = 0, max_node_candidates = 0;
    int32_t layer_coupling = 0;
    std::vector<double> std_cand, std_surv;

    OracleRun(const pmdd_dd_view *v, int pt, int sd_, int thr, int rule)
        : g(v), problem_type(pt), sd(sd_), threads(thr), score_rule(rule) {
        td.resize(g.L);
        bu.resize(g.L);
        std_cand.assign(g.L, 0.0);
        std_surv.assign(g.L, 0.0);
    }
    ~OracleRun() {
        for (auto &layer : td) {
            for (Frontier *f : layer) {
                delete f;
            }
        }
        for (auto &layer : bu) {
            for (Frontier *f : layer) {
                delete f;
            }
        }
    }

    void release_td(int l) {
        for (Frontier *f : td[l]) {
            delete f;
        }
        td[l].clear();
    }
    void release_bu(int l) {
        for (Frontier *f : bu[l]) {
            delete f;
        }
        bu[l].clear();
    }

    static double pop_std(const std::vector<int64_t> &x) {
        if (x.empty()) {
            return 0.0;
        }
        double m = 0.0;
        for (int64_t v : x) {
            m += static_cast<double>(v);
        }
        m /= static_cast<double>(x.size());
        double s = 0.0;
        for (int64_t v : x) {
            s += (static_cast<double>(v) - m) * (static_cast<double>(v) - m);
        }
        return std::sqrt(s / static_cast<double>(x.size()));
    }

    void count_layer(const std::vector<int64_t> &cand, const std::vector<Frontier *> &fr, int l,
                     bool record_std) {
        int64_t c = 0, s = 0;
        std::vector<int64_t> sizes(fr.size());
        for (size_t i = 0; i < fr.size(); ++i) {
            c += cand[i];
            sizes[i] = fr[i]->size();
            s += sizes[i];
            max_node_frontier = std::max(max_node_frontier, sizes[i]);
            max_node_candidates = std::max(max_node_candidates, cand[i]);
        }
        cand_total += c;
        cand_peak = std::max(cand_peak, c);
        surv_total += s;
        surv_peak = std::max(surv_peak, s);
        if (record_std) {
            std_cand[l] = pop_std(cand);
            std_surv[l] = pop_std(sizes);
        }
    }

    // expand_layer_topdown_cpu (topdown.cpp:12-33): merge each in-arc's
    // parent frontier shifted by the arc weight, in CSR order.
    void expand_td(int l, bool record_std) {
        const pmdd_dd_view *v = g.v;
        const int64_t n = g.layer_size[l];
        td[l].assign(n, nullptr);
        std::vector<int64_t> cand(n, 0);
#pragma omp parallel for schedule(dynamic) num_threads(threads) if (threads > 1)
        for (int64_t i = 0; i < n; ++i) {
            Frontier *f = new Frontier(g.d);
            const int64_t gid = g.gid(l, i);
            for (int64_t a = v->in_off[gid]; a < v->in_off[gid + 1]; ++a) {
                const Frontier *p = td[l - 1][v->arc_src[a]];
                cand[i] += p->size();
                f->merge_shift(*p, &v->arc_w[a * g.d]);
            }
            td[l][i] = f;
        }
        if (sd > 0 && problem_type == 1 && v->min_weight != nullptr) {
            knapsack_state_dominance(l);
        }
        count_layer(cand, td[l], l, record_std);
    }

    // expand_layer_bottomup (bottomup.cpp:12-83)
    void expand_bu(int l, bool record_std) {
        const pmdd_dd_view *v = g.v;
        const int64_t n = g.layer_size[l];
        bu[l].assign(n, nullptr);
        std::vector<int64_t> cand(n, 0);
#pragma omp parallel for schedule(dynamic) num_threads(threads) if (threads > 1)
        for (int64_t i = 0; i < n; ++i) {
            Frontier *f = new Frontier(g.d);
            const int64_t gid = g.gid(l, i);
            for (int64_t k = g.out_off[gid]; k < g.out_off[gid + 1]; ++k) {
                const int64_t a = g.out_arc[k];
                const int64_t head = head_of(a, l + 1);
                const Frontier *c = bu[l + 1][head];
                cand[i] += c->size();
                f->merge_shift(*c, &v->arc_w[a * g.d]);
            }
            bu[l][i] = f;
        }
        count_layer(cand, bu[l], l, record_std);
    }

    // Local index of the head of arc a (head lies in layer hl).
    int64_t head_of(int64_t a, int hl) const {
        const pmdd_dd_view *v = g.v;
        const int64_t *lo = v->in_off + v->layer_off[hl];
        const int64_t *hi = v->in_off + v->layer_off[hl + 1];
        // last node whose in_off <= a
        const int64_t *it = std::upper_bound(lo, hi + 1, a);
        return (it - lo) - 1;
    }

    // filter_dominance_knapsack_cpu (dominance.cpp:24-86): nonempty nodes in
    // decreasing min_weight; a label of node k is dropped when some label of
    // the node at position k+1 or k+2 is >= it in every objective, skipping
    // pairs of single-in-arc nodes fed by the same parent through different
    // arc types.  Comparison nodes are read before their own filtering.
    void knapsack_state_dominance(int l) {
        const pmdd_dd_view *v = g.v;
        const int64_t n = g.layer_size[l];
        if (n <= 1) {
            return;
        }
        std::vector<std::pair<int64_t, int32_t>> order;
        for (int64_t i = 0; i < n; ++i) {
            if (td[l][i]->size() > 0) {
                order.push_back({i, v->min_weight[g.gid(l, i)]});
            }
        }
        std::stable_sort(order.begin(), order.end(),
                         [](const std::pair<int64_t, int32_t> &a,
                            const std::pair<int64_t, int32_t> &b) { return a.second > b.second; });
        if (order.empty()) {
            return;
        }
        for (size_t i = 0; i + 1 < order.size(); ++i) {
            const int64_t n1 = order[i].first;
            Frontier *f1 = td[l][n1];
            int64_t num_dominated = 0;
            for (size_t j = i + 1; j < std::min(order.size(), i + 3); ++j) {
                const int64_t n2 = order[j].first;
                if (same_parent_other_arc(l, n1, n2)) {
                    continue;
                }
                const Frontier *f2 = td[l][n2];
                for (int64_t s1 = 0; s1 < f1->size(); ++s1) {
                    if (f1->dead[s1]) {
                        continue;
                    }
                    for (int64_t s2 = 0; s2 < f2->size(); ++s2) {
                        bool dom = true;
                        for (int o = 0; o < g.d && dom; ++o) {
                            dom = f2->sols[s2 * g.d + o] >= f1->sols[s1 * g.d + o];
                        }
                        if (dom) {
                            f1->dead[s1] = 1;
                            ++num_dominated;
                            break;
                        }
                    }
                }
            }
            if (num_dominated > 0) {
                f1->remove_empty();
                filtered += num_dominated;
            }
        }
    }

    bool same_parent_other_arc(int l, int64_t n1, int64_t n2) const {
        const pmdd_dd_view *v = g.v;
        const int64_t g1 = g.gid(l, n1), g2 = g.gid(l, n2);
        if (v->in_off[g1 + 1] - v->in_off[g1] != 1 || v->in_off[g2 + 1] - v->in_off[g2] != 1) {
            return false;
        }
        const int64_t a1 = v->in_off[g1], a2 = v->in_off[g2];
        return v->arc_src[a1] == v->arc_src[a2] && v->arc_val[a1] != v->arc_val[a2];
    }

    Frontier *run_topdown() {
        const pmdd_dd_view *v = g.v;
        td[0].assign(g.layer_size[0], nullptr);
        for (int64_t i = 0; i < g.layer_size[0]; ++i) {
            td[0][i] = new Frontier(g.d);
        }
        std::vector<int32_t> zero(g.d, 0);
        td[0][v->root]->push(zero.data());
        for (int l = 1; l < g.L; ++l) {
            expand_td(l, true);
            release_td(l - 1);
        }
        Frontier *res = td[g.L - 1][v->terminal];
        td[g.L - 1][v->terminal] = nullptr;
        return res;
    }

    Frontier *run_bottomup() {
        const pmdd_dd_view *v = g.v;
        bu[g.L - 1].assign(g.layer_size[g.L - 1], nullptr);
        for (int64_t i = 0; i < g.layer_size[g.L - 1]; ++i) {
            bu[g.L - 1][i] = new Frontier(g.d);
        }
        std::vector<int32_t> zero(g.d, 0);
        bu[g.L - 1][v->terminal]->push(zero.data());
        for (int l = g.L - 2; l >= 0; --l) {
            expand_bu(l, true);
            release_bu(l + 1);
        }
        Frontier *res = bu[0][v->root];
        bu[0][v->root] = nullptr;
        return res;
    }

    // couple.cpp:15-48
    int64_t td_value(int l) const {
        int64_t total = 0;
        for (int64_t i = 0; i < g.layer_size[l]; ++i) {
            total += td[l][i]->size() * g.out_deg(l, i);
        }
        return total;
    }
    int64_t bu_value(int l) const {
        int64_t total = 0;
        if (score_rule == 0) {
            for (int64_t i = 0; i < g.layer_size[l]; ++i) {
                total += static_cast<int64_t>(1.5 * static_cast<double>(bu[l][i]->size() *
                                                                        g.in_deg(l, i)));
            }
            return total;
        }
        for (int64_t i = 0; i < g.layer_size[l]; ++i) {
            total += bu[l][i]->size() * g.in_deg(l, i);
        }
        return static_cast<int64_t>(1.5 * static_cast<double>(total));
    }

    Frontier *run_coupled() {
        const pmdd_dd_view *v = g.v;
        std::vector<int32_t> zero(g.d, 0);
        td[0].assign(g.layer_size[0], nullptr);
        for (int64_t i = 0; i < g.layer_size[0]; ++i) {
            td[0][i] = new Frontier(g.d);
        }
        td[0][v->root]->push(zero.data());
        bu[g.L - 1].assign(g.layer_size[g.L - 1], nullptr);
        for (int64_t i = 0; i < g.layer_size[g.L - 1]; ++i) {
            bu[g.L - 1][i] = new Frontier(g.d);
        }
        bu[g.L - 1][v->terminal]->push(zero.data());

        int lt = 0, lb = g.L - 1;
        int64_t val_td = 0, val_bu = 0;
        while (lt != lb) {
            if (val_td <= val_bu) {
                ++lt;
                expand_td(lt, false);
                release_td(lt - 1);
                val_td = td_value(lt);
            } else {
                --lb;
                expand_bu(lb, false);
                release_bu(lb + 1);
                val_bu = bu_value(lb);
            }
        }
        layer_coupling = lt;
        std::vector<int64_t> cut(g.layer_size[lt]);
        for (int64_t i = 0; i < g.layer_size[lt]; ++i) {
            cut[i] = i;
            join_products += td[lt][i]->size() * bu[lt][i]->size();
        }
        // CompareNode: larger frontier sums first
        std::vector<int64_t> key(cut.size());
        for (size_t i = 0; i < cut.size(); ++i) {
            key[i] = td[lt][i]->sum() + bu[lt][i]->sum();
        }
        std::stable_sort(cut.begin(), cut.end(),
                         [&](int64_t a, int64_t b) { return key[a] > key[b]; });
        Frontier *res = nullptr;
        if (threads > 1 && cut.size() > 1) {
            std::vector<Frontier *> partial(threads, nullptr);
#pragma omp parallel num_threads(threads)
            {
                const int tid = omp_get_thread_num();
                Frontier *local = new Frontier(g.d);
                partial[tid] = local;
#pragma omp for schedule(dynamic)
                for (int64_t k = 0; k < static_cast<int64_t>(cut.size()); ++k) {
                    const int64_t i = cut[k];
                    local->convolute(*td[lt][i], *bu[lt][i]);
                }
            }
            // pairwise tree reduction (cpu_helpers.hpp:49-88)
            size_t active = partial.size();
            while (active > 1) {
                const size_t pairs = active / 2;
#pragma omp parallel for schedule(dynamic) num_threads(threads)
                for (int64_t p = 0; p < static_cast<int64_t>(pairs); ++p) {
                    Frontier *lhs = partial[2 * p];
                    Frontier *rhs = partial[2 * p + 1];
                    lhs->merge(*rhs);
                    delete rhs;
                    partial[p] = lhs;
                }
                if (active % 2 != 0) {
                    partial[pairs] = partial[active - 1];
                    active = pairs + 1;
                } else {
                    active = pairs;
                }
            }
            res = partial[0];
        } else {
            res = new Frontier(g.d);
            for (int64_t i : cut) {
                res->convolute(*td[lt][i], *bu[lt][i]);
            }
        }
        surv_total += res->size();
        surv_peak = std::max(surv_peak, res->size());
        return res;
    }
};

// ---------------------------------------------------------------------------
// Brute force
// ---------------------------------------------------------------------------

bool read_ints(const char *path, std::vector<long long> *out) {
    std::ifstream in(path);
    if (!in.is_open()) {
        return false;
    }
    long long x;
    while (in >> x) {
        out->push_back(x);
    }
    return true;
}

// Nondominated subset with duplicates merged, by plain pairwise comparison.
void nondominated(std::vector<int32_t> *pts, int d) {
    const int64_t n = static_cast<int64_t>(pts->size()) / d;
    // sort lexicographically descending and drop exact duplicates first
    std::vector<int64_t> idx(n);
    for (int64_t i = 0; i < n; ++i) {
        idx[i] = i;
    }
    std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        for (int o = 0; o < d; ++o) {
            if ((*pts)[a * d + o] != (*pts)[b * d + o]) {
                return (*pts)[a * d + o] > (*pts)[b * d + o];
            }
        }
        return a < b;
    });
    std::vector<int32_t> uniq;
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = idx[k];
        if (k > 0 && std::equal(pts->begin() + i * d, pts->begin() + i * d + d,
                                uniq.end() - d)) {
            continue;
        }
        uniq.insert(uniq.end(), pts->begin() + i * d, pts->begin() + i * d + d);
    }
    const int64_t m = static_cast<int64_t>(uniq.size()) / d;
    std::vector<int32_t> res;
    for (int64_t i = 0; i < m; ++i) {
        bool dominated = false;
        for (int64_t j = 0; j < m && !dominated; ++j) {
            if (j == i) {
                continue;
            }
            bool ge = true;
            for (int o = 0; o < d && ge; ++o) {
                ge = uniq[j * d + o] >= uniq[i * d + o];
            }
            dominated = ge; // distinct points, so >= everywhere means strictly better
        }
        if (!dominated) {
            res.insert(res.end(), uniq.begin() + i * d, uniq.begin() + i * d + d);
        }
    }
    pts->swap(res);
}

int brute_knapsack(const std::vector<long long> &t, std::vector<int32_t> *pts, int *d_out) {
    if (t.size() < 3) {
        return PMDD_E_INVALID;
    }
    const int n = static_cast<int>(t[0]), m = static_cast<int>(t[1]), d = static_cast<int>(t[2]);
    if (n > 26 || t.size() < 3 + static_cast<size_t>(d) * n + static_cast<size_t>(m) * (n + 1)) {
        return PMDD_E_INVALID;
    }
    auto prof = [&](int o, int i) { return t[3 + static_cast<size_t>(o) * n + i]; };
    const size_t cbase = 3 + static_cast<size_t>(d) * n;
    auto coef = [&](int c, int i) { return t[cbase + static_cast<size_t>(c) * (n + 1) + i]; };
    auto rhs = [&](int c) { return t[cbase + static_cast<size_t>(c) * (n + 1) + n]; };
    std::vector<int32_t> p(d);
    for (uint64_t mask = 0; mask < (1ULL << n); ++mask) {
        bool ok = true;
        for (int c = 0; c < m && ok; ++c) {
            long long used = 0;
            for (int i = 0; i < n; ++i) {
                if (mask >> i & 1ULL) {
                    used += coef(c, i);
                }
            }
            ok = used <= rhs(c);
        }
        if (!ok) {
            continue;
        }
        for (int o = 0; o < d; ++o) {
            long long s = 0;
            for (int i = 0; i < n; ++i) {
                if (mask >> i & 1ULL) {
                    s += prof(o, i);
                }
            }
            p[o] = static_cast<int32_t>(s);
        }
        pts->insert(pts->end(), p.begin(), p.end());
    }
    *d_out = d;
    return PMDD_OK;
}

void enum_indep(int v, int n, const std::vector<uint64_t> &adj, uint64_t chosen,
                const std::vector<std::vector<long long>> &obj, int d, std::vector<int32_t> *pts) {
    if (v == n) {
        for (int o = 0; o < d; ++o) {
            long long s = 0;
            for (int i = 0; i < n; ++i) {
                if (chosen >> i & 1ULL) {
                    s += obj[o][i];
                }
            }
            pts->push_back(static_cast<int32_t>(s));
        }
        return;
    }
    enum_indep(v + 1, n, adj, chosen, obj, d, pts);
    if ((adj[v] & chosen) == 0) {
        enum_indep(v + 1, n, adj, chosen | (1ULL << v), obj, d, pts);
    }
}

int brute_setpacking(const std::vector<long long> &t, std::vector<int32_t> *pts, int *d_out) {
    if (t.size() < 3) {
        return PMDD_E_INVALID;
    }
    const int n = static_cast<int>(t[0]), m = static_cast<int>(t[1]), d = static_cast<int>(t[2]);
    if (n > 40) {
        return PMDD_E_INVALID;
    }
    size_t k = 3;
    std::vector<std::vector<long long>> obj(d, std::vector<long long>(n));
    for (int o = 0; o < d; ++o) {
        for (int i = 0; i < n; ++i) {
            obj[o][i] = t.at(k++);
        }
    }
    std::vector<uint64_t> adj(n, 0);
    for (int c = 0; c < m; ++c) {
        const int cnt = static_cast<int>(t.at(k++));
        std::vector<int> vs;
        for (int i = 0; i < cnt; ++i) {
            vs.push_back(static_cast<int>(t.at(k++)) - 1);
        }
        for (int a : vs) {
            for (int b : vs) {
                if (a != b) {
                    adj[a] |= 1ULL << b;
                }
            }
        }
    }
    enum_indep(0, n, adj, 0, obj, d, pts);
    *d_out = d;
    return PMDD_OK;
}

int brute_tsp(const std::vector<long long> &t, std::vector<int32_t> *pts, int *d_out) {
    if (t.size() < 2) {
        return PMDD_E_INVALID;
    }
    const int d = static_cast<int>(t[0]), n = static_cast<int>(t[1]);
    if (n > 10 || t.size() < 2 + static_cast<size_t>(d) * n * n) {
        return PMDD_E_INVALID;
    }
    auto dist = [&](int o, int i, int j) {
        return t[2 + (static_cast<size_t>(o) * n + i) * n + j];
    };
    std::vector<int> perm;
    for (int c = 1; c < n; ++c) {
        perm.push_back(c);
    }
    do {
        for (int o = 0; o < d; ++o) {
            long long s = 0;
            int prev = 0;
            for (int c : perm) {
                s += dist(o, prev, c);
                prev = c;
            }
            s += dist(o, prev, 0);
            pts->push_back(static_cast<int32_t>(-s));
        }
    } while (std::next_permutation(perm.begin(), perm.end()));
    *d_out = d;
    return PMDD_OK;
}

int32_t *export_points(const std::vector<int32_t> &v) {
    int32_t *p = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * (v.size() + 1)));
    if (p != nullptr && !v.empty()) {
        std::memcpy(p, v.data(), sizeof(int32_t) * v.size());
    }
    return p;
}

} // namespace

extern "C" {

typedef struct oracle_stats {
    int64_t work_candidates_total;
    int64_t work_candidates_peak;
    int64_t work_frontier_survivors_total;
    int64_t work_frontier_peak_points;
    int64_t work_join_products_total;
    int64_t dominance_filtered_total;
    int32_t layer_coupling;
    int32_t reserved;
    int64_t max_node_frontier;
    int64_t max_node_candidates;
} oracle_stats;

// method: 1 top-down, 2 bottom-up, 3 coupled.  score_rule picks the method-3
// bottom-up layer value: 0 truncates 1.5*|F|*in_arcs per node (couple.cpp:36),
// 1 truncates 1.5 times the layer sum (the GPU rule, enum.cu:735).
// std_cand / std_surv (nullable) receive num_layers entries for methods 1/2.
int oracle_enumerate(const pmdd_dd_view *v, int method, int problem_type, int state_dominance,
                     int threads, int score_rule, int32_t **out_points, int64_t *out_count,
                     oracle_stats *st, double *std_cand, double *std_surv) {
    if (v == nullptr || out_points == nullptr || out_count == nullptr) {
        return PMDD_E_INVALID;
    }
    if (threads < 1) {
        threads = 1;
    }
    OracleRun run(v, problem_type, state_dominance, threads, score_rule);
    Frontier *f = nullptr;
    if (method == 1) {
        f = run.run_topdown();
    } else if (method == 2) {
        f = run.run_bottomup();
    } else if (method == 3) {
        f = run.run_coupled();
    } else {
        return PMDD_E_INVALID;
    }
    *out_points = export_points(f->sols);
    *out_count = f->size();
    delete f;
    if (st != nullptr) {
        st->work_candidates_total = run.cand_total;
        st->work_candidates_peak = run.cand_peak;
        st->work_frontier_survivors_total = run.surv_total;
        st->work_frontier_peak_points = run.surv_peak;
        st->work_join_products_total = run.join_products;
        st->dominance_filtered_total = run.filtered;
        st->layer_coupling = run.layer_coupling;
        st->max_node_frontier = run.max_node_frontier;
        st->max_node_candidates = run.max_node_candidates;
    }
    if (std_cand != nullptr) {
        std::copy(run.std_cand.begin(), run.std_cand.end(), std_cand);
    }
    if (std_surv != nullptr) {
        std::copy(run.std_surv.begin(), run.std_surv.end(), std_surv);
    }
    return PMDD_OK;
}

// Brute-force nondominated set straight from an instance file.
int oracle_bruteforce_file(const char *path, int problem_type, int32_t **out_points,
                           int64_t *out_count, int *out_d) {
    std::vector<long long> t;
    if (!read_ints(path, &t)) {
        return PMDD_E_INVALID;
    }
    std::vector<int32_t> pts;
    int d = 0;
    int rc = PMDD_E_INVALID;
    if (problem_type == 1) {
        rc = brute_knapsack(t, &pts, &d);
    } else if (problem_type == 2) {
        rc = brute_setpacking(t, &pts, &d);
    } else if (problem_type == 3) {
        rc = brute_tsp(t, &pts, &d);
    }
    if (rc != PMDD_OK) {
        return rc;
    }
    nondominated(&pts, d);
    *out_points = export_points(pts);
    *out_count = static_cast<int64_t>(pts.size()) / d;
    *out_d = d;
    return PMDD_OK;
}

// Nondominated subset of an arbitrary point set (duplicates merged), for
// tests of the primitive itself.
int oracle_nondominated(const int32_t *pts, int64_t n, int d, int32_t **out_points,
                        int64_t *out_count) {
    std::vector<int32_t> v(pts, pts + n * d);
    nondominated(&v, d);
    *out_points = export_points(v);
    *out_count = static_cast<int64_t>(v.size()) / d;
    return PMDD_OK;
}

void oracle_free(void *p) { std::free(p); }

} // extern "C"
