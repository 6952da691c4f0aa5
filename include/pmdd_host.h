/*
 * pmdd_host.h -- C ABI of the host surface: instance parsers and decision
 * diagram constructors.
 *
 * These entry points wrap what the reference does in main() before it calls
 * the enumeration backend:
 *   - KnapsackInstance::read            src/pmdd/instances/knapsack_instance.cpp:17-47
 *   - SetPackingInstance(...)           src/pmdd/instances/setpacking_instance.hpp:49-89
 *   - create_indepset_instance()        src/pmdd/instances/setpacking_instance.hpp:92-126
 *   - TSPInstance::read                 src/pmdd/instances/tsp_instance.cpp:11-41
 *   - KnapsackBDDConstructor::generate_exact + BDDAlg::reduce +
 *     update_node_weights (twice)       src/pmdd/main.cpp:113-158
 *   - IndepSetBDDConstructor::generate_exact (no reduction)
 *                                       src/pmdd/main.cpp:161-181
 *   - MDDTSPConstructor::generate_exact src/pmdd/main.cpp:184-199
 * and the packing step that flattens the diagram into per-layer CSR
 * (pack_bdd_layers src/pmdd/enum/gpu/enum.cu:438-593,
 *  pack_mdd_layers src/pmdd/enum/gpu/enum.cu:85-158).
 *
 * No exception crosses this boundary: every call returns PMDD_OK or a
 * negative PMDD_E_* code and fills err (when non-NULL) with a message.
 */
#ifndef PMDD_HOST_H_
#define PMDD_HOST_H_

#include <stddef.h>
#include <stdint.h>

#include "pmdd_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque decision diagram built on the host. */
typedef struct pmdd_dd pmdd_dd;

/* Order-independent structure figures reported in the stats JSONL
 * "structure" object (src/pmdd/util/output_utils.cpp:254-268). */
typedef struct pmdd_dd_structure {
    int64_t original_width;      /* -1 for TSP, as in main.cpp:283-286 */
    int64_t reduced_width;
    int64_t original_num_nodes;
    int64_t reduced_num_nodes;
    int64_t max_num_in_arcs;     /* BDD::get_max_num_in_arcs  bdd.hpp:262-276 */
    int64_t total_num_in_arcs;   /* BDD::get_total_num_in_arcs bdd.hpp:248-257 */
    int32_t num_layers;
    int32_t d;
} pmdd_dd_structure;

/* Read an instance file in the reference .dat format and build the decision
 * diagram that main() builds for it.  problem_type: 1 knapsack, 2 set packing,
 * 3 TSP.  expect_d > 0 rejects files whose objective count differs (the
 * reference binary is compiled for one NOBJS, src/pmdd/makefile:54-58);
 * expect_d == 0 takes d from the file.  flags: PMDD_BUILD_STATE_DOMINANCE
 * also records what state dominance needs that only the construction knows
 * (for set packing, each node's dominance partner, view.sd_partner).
 * Wide set-packing layers and the flattening run on up to 8 host threads;
 * the environment variable PMDD_HOST_THREADS, when set to a positive number,
 * caps them (the diagram is the same on any number of threads).
 * Returns PMDD_OK or PMDD_E_*. */
#define PMDD_BUILD_STATE_DOMINANCE 1
int pmdd_dd_build_from_file(const char *path, int problem_type, int expect_d, int flags,
                            pmdd_dd **out, char *err, size_t errlen);

/* Borrowed view of the flattened diagram; valid until pmdd_dd_free. */
const pmdd_dd_view *pmdd_dd_get_view(const pmdd_dd *dd);

/* Structure figures of the diagram. */
void pmdd_dd_get_structure(const pmdd_dd *dd, pmdd_dd_structure *out);

/* Node count of layer l (0 <= l < num_layers); -1 when out of range. */
int64_t pmdd_dd_layer_size(const pmdd_dd *dd, int layer);

void pmdd_dd_free(pmdd_dd *dd);

/* Sort a frontier's points lexicographically ascending, in place: the order
 * main() writes (ParetoFrontier::sort_lexicographic_ascending,
 * src/pmdd/enum/pareto_frontier.hpp:418-433, called at src/pmdd/main.cpp:403).
 * The engine returns frontiers unsorted, as the reference's GPU entry points
 * do.  Returns PMDD_OK, or PMDD_E_NOMEM when the host has no memory for the
 * sort (the points are then left as they were). */
int pmdd_frontier_sort(pmdd_frontier *fr);

#ifdef __cplusplus
}
#endif

#endif /* PMDD_HOST_H_ */
