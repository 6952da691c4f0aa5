// Layered decision diagrams built on the host, flattened into the engine's
// top-down CSR view (pmdd_dd_view in pmdd_engine.h).
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/pmdd_engine.h"
#include "../../../include/pmdd_host.h"
#include "instances.hpp"

namespace pmdd {

// Allocator that leaves elements uninitialised when a vector grows without a
// value (resize(n)), so that arrays of hundreds of MB sized on one thread and
// filled on host threads are not first zeroed on that one thread.
template <class T> struct NoInitAlloc : std::allocator<T> {
    template <class U> struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() noexcept = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <class U> void construct(U *p) noexcept(noexcept(::new (static_cast<void *>(p)) U)) {
        ::new (static_cast<void *>(p)) U;
    }
    template <class U, class... A> void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
template <class T> using RawVec = std::vector<T, NoInitAlloc<T>>;

// Flattened diagram: owns the arrays a pmdd_dd_view points into.
struct FlatDD {
    int32_t d = 0;
    int32_t kind = PMDD_DD_BDD;
    int32_t num_layers = 0;
    int32_t root = 0;
    int32_t terminal = 0;
    std::vector<int64_t> layer_off;  // [L+1]
    RawVec<int64_t> in_off;          // [total_nodes+1]
    RawVec<int32_t> arc_src;         // [A]
    RawVec<int32_t> arc_val;         // [A]
    RawVec<int32_t> arc_w;           // [A*d]
    std::vector<int32_t> min_weight; // [total_nodes] or empty
    std::vector<int32_t> sd_partner; // [total_nodes] or empty: set-packing dominance partner
    pmdd_dd_structure structure{};

    int64_t layer_size(int l) const { return layer_off[l + 1] - layer_off[l]; }
    int64_t num_arcs() const { return static_cast<int64_t>(arc_src.size()); }
    pmdd_dd_view view() const;
};

// A binary diagram as the constructors produce it: per layer, each node's
// 0-child and 1-child in the next layer (-1 when absent).  0-arcs weigh zero
// and every 1-arc leaving layer l carries the same weight vector one_w[l], as
// in the knapsack (knapsack_bdd.cpp:36-41) and independent-set
// (indepset_bdd.cpp:98-102) constructors.
// One layer's children, [node] -> {0-child, 1-child}.
using ChildLayer = RawVec<std::array<int32_t, 2>>;

struct BinaryDD {
    int d = 0;
    int num_layers = 0;
    std::vector<ChildLayer> child;           // [L-1][node]
    std::vector<std::vector<int32_t>> one_w; // [L-1][d]

    int64_t layer_size(int l) const {
        return l < num_layers - 1 ? static_cast<int64_t>(child[l].size()) : 1;
    }
    int64_t width() const;
    int64_t num_nodes() const;
};

// Exact knapsack BDD: state = used capacity per constraint
// (KnapsackBDDConstructor::generate_exact, knapsack_bdd.cpp:10-118).
BinaryDD build_knapsack_exact(const KnapsackInstance &inst, int d);

// Merge nodes with equal (0-child, 1-child) bottom-up and drop nodes without
// children (BDDAlg::reduce, bdd_alg.hpp:24-93).
void reduce_bdd(BinaryDD *bdd);

// Minimum used capacity of constraint 0 over root paths
// (KnapsackBDDConstructor::update_node_weights, knapsack_bdd.cpp:123-150).
std::vector<std::vector<int32_t>> knapsack_min_weights(const BinaryDD &bdd,
                                                        const KnapsackInstance &inst);

// Exact independent-set BDD with the min-in-state variable order
// (IndepSetBDDConstructor::generate_exact, indepset_bdd.cpp:50-165, and
// choose_next_vertex_min_size_next_layer, indepset_bdd.hpp:89-114) over the
// conflict graph of a set packing instance (create_indepset_instance,
// setpacking_instance.hpp:92-126).  var_layer receives the chosen order.
// With `partners` non-null, also the set-packing state-dominance partner of
// every node (layer-local index, -1 for none; see setpacking_partners).
BinaryDD build_setpacking_exact(const SetPackingInstance &inst, int d,
                                std::vector<int32_t> *var_layer,
                                std::vector<std::vector<int32_t>> *partners = nullptr);

// Flatten a binary diagram (1-arcs before 0-arcs per head, pack_bdd_layers
// order enum.cu:444-445).
void flatten_bdd(const BinaryDD &bdd, const std::vector<std::vector<int32_t>> *min_weight,
                 FlatDD *out);

// Exact TSP MDD: state = (visited set, last city), city 0 first, arc weights
// are negated distances, closing arcs to the terminal
// (MDDTSPConstructor::generate_exact, tsp_mdd.cpp:21-196).
void build_tsp_exact(const TspInstance &inst, int d, FlatDD *out);

// main()'s construction sequence for a problem type (main.cpp:113-199).
// state_dominance: also compute what the state-dominance rules need that only
// the construction knows (the set-packing partners).
bool build_from_file(const std::string &path, int problem_type, int expect_d, bool state_dominance,
                     FlatDD *out,
                     std::string *error);

} // namespace pmdd
