/*
 * pmdd_engine.h -- C ABI of the MI355X frontier-enumeration engine.
 *
 * This is the seam that replaces the reference's GPU backend
 * (src/pmdd/enum/gpu/).  The reference binds four C++ functions there,
 * declared in src/pmdd/enum/gpu/cuda_wrappers.hpp:11-21:
 *
 *   topdown_cuda_enumerate(BDD*, maximization, problem_type, state_dominance,
 *                          EnumerationStats*, std::string* reason, gpu_batch_size)
 *   topdown_mdd_cuda_enumerate(MDD*, EnumerationStats*, std::string*, gpu_batch_size)
 *   coupled_cuda_enumerate(MDD*, EnumerationStats*, std::string*, gpu_batch_size, gpu_max_prod)
 *   coupled_bdd_cuda_enumerate(BDD*, maximization, problem_type, state_dominance,
 *                              EnumerationStats*, std::string*, gpu_batch_size, gpu_max_prod)
 *
 * Here the diagram crosses the boundary as plain arrays (pmdd_dd_view), the
 * result as a malloc'd point array (pmdd_frontier), errors as a negative code
 * plus a message in a caller buffer (the reference returns NULL and sets
 * *reason, src/pmdd/enum/gpu/enum_types.cuh:76-97), and the statistics as a
 * POD struct (the GPU-filled subset of EnumerationStats,
 * src/pmdd/util/stats.hpp:18-85).  No C++ type and no exception crosses it.
 *
 * Semantics (all integer, exact): maximisation everywhere; the frontier of a
 * node is the set of nondominated root->node path-weight vectors with
 * duplicates merged; the returned frontier is the nondominated set of all
 * root->terminal path sums, unsorted (the caller sorts, as main.cpp:403 does).
 */
#ifndef PMDD_ENGINE_H_
#define PMDD_ENGINE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define PMDD_OK 0
#define PMDD_E_INVALID (-1)     /* malformed view / options */
#define PMDD_E_NODEVICE (-2)    /* no HIP device ("No CUDA device found", enum.cu:37-39) */
#define PMDD_E_HIP (-3)         /* a HIP runtime call failed */
#define PMDD_E_NOMEM (-4)       /* device or host allocation failed */
#define PMDD_E_OVERFLOW (-5)    /* an int32 path sum could overflow */
#define PMDD_E_UNSUPPORTED (-6) /* objective count or method outside what the engine builds */
#define PMDD_E_REJECTED (-7)    /* serving admission refused the request */

/* ---- diagram kinds ---- */
#define PMDD_DD_BDD 0
#define PMDD_DD_MDD 1

/* Flattened layered decision diagram.
 *
 * Layers 0..num_layers-1; layer 0 holds the root, the last layer the terminal.
 * Nodes are numbered globally: layer l owns ids [layer_off[l], layer_off[l+1]).
 * Arcs are stored once, grouped by head node (top-down CSR): the in-arcs of
 * global node g are arcs [in_off[g], in_off[g+1]).  For a BDD the in-arcs of
 * a node list the 1-arcs first, then the 0-arcs (the maximisation order of
 * pack_bdd_layers, src/pmdd/enum/gpu/enum.cu:444-445,460-494).
 *
 * arc_src[a]  local index (within the previous layer) of the arc's tail
 * arc_val[a]  BDD: arc type 0/1; MDD: arc value (the city for TSP)
 * arc_w[a*d+o] objective o weight of arc a (int32, AoS like td_edge_weights)
 * min_weight  per global node, knapsack used capacity (constraint 0) as set
 *             by update_node_weights (knapsack_bdd.cpp:123-150); NULL when
 *             the diagram is not a knapsack BDD.
 */
typedef struct pmdd_dd_view {
    int32_t d;
    int32_t num_layers;
    int32_t kind; /* PMDD_DD_BDD or PMDD_DD_MDD */
    int32_t root;     /* local index of the root in layer 0 */
    int32_t terminal; /* local index of the terminal in the last layer */
    int32_t reserved;
    const int64_t *layer_off; /* [num_layers + 1] */
    const int64_t *in_off;    /* [layer_off[num_layers] + 1] */
    const int32_t *arc_src;   /* [num_arcs] */
    const int32_t *arc_val;   /* [num_arcs] */
    const int32_t *arc_w;     /* [num_arcs * d] */
    const int32_t *min_weight; /* [layer_off[num_layers]] or NULL */
    /* Set-packing state dominance (no GPU counterpart in the reference; its
     * CPU rule, dominance.cpp:88-173): per node, the layer-local index of the
     * node whose labels may drop its labels, or -1; NULL when not built. */
    const int32_t *sd_partner;
} pmdd_dd_view;

/* Options of one enumeration, mirroring the reference CLI
 * (src/pmdd/util/cli_parser.hpp:10-30). */
typedef struct pmdd_opts {
    int32_t method;          /* 1 = top-down, 2 = bottom-up (the reference runs it on the
                              * CPU only), 3 = dynamic layer cutset (coupled) */
    int32_t problem_type;    /* 1 knapsack, 2 set packing, 3 TSP */
    int32_t state_dominance; /* 0 or 1 */
    int32_t count_pair_tests;/* 1 = count executed dominance pair tests (slower) */
    int64_t max_cand;        /* --max-cand, candidates per expansion pass (default 20M) */
    int64_t max_prod;        /* --max-prod: > 0 caps the products of every coupling batch
                              * (couple.cu:393-403; a node with more products than the cap
                              * is split over several batches, so no batch exceeds it);
                              * PMDD_MAX_PROD_ADAPTIVE (0): batches start at 32K products
                              * and double up to 32M */
    int32_t phase_timing;    /* 1 = time the phases of every layer with HIP events
                              * (kernel_*_s below); 0 = record no per-layer event (the
                              * events cost about 6% of a run's wall time).  The named
                              * entry points set it when given stats, as the reference
                              * times its kernels whenever it is given stats. */
    int32_t reserved;
} pmdd_opts;

#define PMDD_DEFAULT_MAX_CAND 20000000LL
#define PMDD_DEFAULT_MAX_PROD 625000LL
#define PMDD_MAX_PROD_ADAPTIVE 0LL

/* Result: count points of d int32 objectives, row-major (point i at
 * points[i*d .. i*d+d-1]).  Allocated with malloc by the engine; release with
 * pmdd_frontier_free. */
typedef struct pmdd_frontier {
    int32_t d;
    int32_t reserved;
    int64_t count;
    int32_t *points;
} pmdd_frontier;

/* GPU-side statistics (subset of EnumerationStats that the GPU path fills,
 * src/pmdd/util/stats.hpp:18-85).  The engine overwrites every scalar field.
 * The per-layer arrays are caller-owned and optional: when non-NULL the engine
 * writes min(std_capacity, num_layers) entries (methods 1 and 2; method 3
 * leaves them at 0, as the reference clears them, enum.cu:610-613). */
typedef struct pmdd_stats {
    /* Every kernel_*_s field is 0 unless opts.phase_timing is set. */
    double wall_pack_transfer_s;
    double wall_join_s;
    double kernel_expand_td_s;
    double kernel_dominance_s;
    double kernel_total_s;
    int64_t gpu_mem_peak_used_bytes;
    int64_t gpu_mem_peak_reserved_bytes;
    int64_t work_candidates_total;
    int64_t work_candidates_peak;
    int64_t work_frontier_survivors_total;
    int64_t work_frontier_peak_points;
    int64_t work_join_products_total;
    int64_t dominance_filtered_total;
    int64_t dominance_pair_tests; /* executed pair tests; 0 unless opts.count_pair_tests */
    /* Engine-specific figures for roofline accounting (no reference counterpart):
     * the fused filter kernel (k_filter_chunks) alone -- its total device time,
     * launch count and candidates loaded -- and the compulsory HBM bytes of the
     * expanded layers, 4*(d*(S_prev+S_next) + E*(d+1) + 2*(N+1) + (N_prev+1)). */
    double kernel_filter_s;
    int64_t filter_launches;
    int64_t filter_candidates;
    int64_t algorithmic_bytes;
    /* The coupling kernel k_refill_vs_block, both of its uses: products
     * checked against the running frontier block by block (per launch: its
     * input points read once, its survivors written once, the block of the
     * frontier read once), and the running frontier's points flagged when a
     * block of the batch frontier dominates them (per launch: the frontier and
     * the block read once, the flags read and written).  Device time,
     * launches, compulsory bytes.  Part of kernel_dominance_s. */
    double kernel_frontier_check_s;
    int64_t frontier_check_launches;
    int64_t frontier_check_bytes;
    /* The block cross filter of big nodes and coupling batches
     * (k_cross_chunks with k_append_front): device time, launches, compulsory
     * bytes (per wave: the group's front and the wave's points read once,
     * the survivors appended once, one flag per point). */
    double kernel_cross_s;
    int64_t cross_launches;
    int64_t cross_bytes;
    /* Executed pair tests per kernel family (0 unless opts.count_pair_tests):
     * the chunk filter, the frontier check (both modes), the cross filter,
     * state dominance.  Their sum is dominance_pair_tests. */
    int64_t pair_tests_filter;
    int64_t pair_tests_frontier_check;
    int64_t pair_tests_cross;
    int64_t pair_tests_state_dominance;
    /* State dominance (k_state_dominance) device time and launches; part of
     * kernel_dominance_s. */
    double kernel_state_dominance_s;
    int64_t state_dominance_launches;
    /* Compulsory HBM bytes of the coupling: every TD and BU point of the
     * cutset read once per batch it takes part in, the final frontier
     * written once, 4*d*(sum over batches of (#TD rows + #BU columns) + #R). */
    int64_t join_bytes;
    /* Replays of the captured graph of a queued expansion pass (packed
     * serving layers), one per pass. */
    int64_t graph_launches;
    /* Coupling batches (method 3) and the most products any one held. */
    int64_t join_batches;
    int64_t join_batch_peak_products;
    /* Device bytes this call held at its peak: the context's buffers (kept
     * from earlier calls included) plus the diagram, counted exactly from the
     * engine's own allocations (gpu_mem_peak_used_bytes is device-wide). */
    int64_t ctx_mem_peak_bytes;
    /* The same, counting only the bytes this call asked for (each buffer at
     * the largest size the call needed of it, the diagram included): what
     * the call needs on a context that kept nothing from earlier calls. */
    int64_t mem_need_peak_bytes;
    /* Serving only: the admission estimate of the request (bytes), the
     * times it was retried alone after running out of device memory next to
     * other requests, and seconds from submit to admission. */
    int64_t mem_estimate_bytes;
    int32_t nomem_retries;
    int32_t reserved2;
    double queue_wait_s;
    int32_t layer_coupling;
    int32_t std_capacity;
    double *std_candidates_per_layer;
    double *std_frontier_survivors_per_layer;
} pmdd_stats;

/* Per-device engine context: two HIP streams (default priority; the second
 * runs the small-node filters of a layer with big nodes beside them, joined
 * before the layer's placement, and every call ends with both idle) and the
 * device buffers its enumerations reuse.  Those buffers (pass scratch and frontier
 * buffers) stay at the largest size any call needed until pmdd_ctx_destroy,
 * so repeated enumerations allocate nothing.  Calls on one context must not
 * overlap; use one context per thread (or the server below). */
typedef struct pmdd_ctx pmdd_ctx;

/* A diagram uploaded to device memory, reusable across enumerations. */
typedef struct pmdd_resident pmdd_resident;

/* Create a context on HIP device `device` (the reference hard-codes device 0,
 * enum.cu:42-47).  Fails with PMDD_E_NODEVICE when no device is visible. */
int pmdd_ctx_create(int device, pmdd_ctx **out, char *err, size_t errlen);
void pmdd_ctx_destroy(pmdd_ctx *ctx);

/* Launch shapes of a context, per objective count d (no reference
 * counterpart).  A context starts with the shapes of the tuning table
 * (pmdd_amd/tune/gfx950.json next to the library, or the file named by the
 * environment variable PMDD_TUNE_FILE) and built-in defaults where the table
 * has none; the tuner (python -m pmdd_amd.tune) writes that table.  Shapes
 * never change a result, only the time it takes. */
typedef struct pmdd_tuning {
    int32_t filter_threads[3]; /* threads of the chunk filter for nodes of at most 256
                                * candidates, at most 1024 (d <= 4), at most a whole
                                * chunk: each 128, 256 or 512 (defaults 128, 512, 512) */
    int32_t cross_workgroups;  /* workgroups a wave of the block cross filter is spread
                                * over, 256..16384 (default 2048) */
    int32_t cross_split_max;   /* most workgroups per block in that wave, 1..16 (16) */
    int32_t refill_waves;      /* one-wave workgroups of the frontier check,
                                * 256..32768 (8192) */
    int32_t source;            /* out: 0 built-in defaults, 1 the table, 2 set by the caller */
} pmdd_tuning;

/* The shapes a context uses for objective count d (1..8); set_tuning
 * validates them (PMDD_E_INVALID when out of range). */
int pmdd_ctx_get_tuning(pmdd_ctx *ctx, int d, pmdd_tuning *out);
int pmdd_ctx_set_tuning(pmdd_ctx *ctx, int d, const pmdd_tuning *t);

/* Number of visible HIP devices (0 when none); never fails. */
int pmdd_device_count(void);

/* One-shot enumeration: upload, enumerate, copy the result back, free device
 * memory.  Replaces the four cuda_wrappers.hpp entry points; method and
 * problem type come from opts, the diagram kind from the view. */
int pmdd_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *view, const pmdd_opts *opts,
                   pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen);

/* Named forms of the four reference entry points (cuda_wrappers.hpp:11-21).
 * They fill a pmdd_opts and call pmdd_enumerate; the view kind must match. */
int pmdd_topdown_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *bdd, int problem_type,
                           int state_dominance, int64_t max_cand, pmdd_frontier *out,
                           pmdd_stats *stats, char *err, size_t errlen);
int pmdd_topdown_mdd_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *mdd, int64_t max_cand,
                               pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen);
int pmdd_coupled_mdd_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *mdd, int64_t max_cand,
                               int64_t max_prod, pmdd_frontier *out, pmdd_stats *stats, char *err,
                               size_t errlen);
int pmdd_coupled_bdd_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *bdd, int problem_type,
                               int state_dominance, int64_t max_cand, int64_t max_prod,
                               pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen);

/* Upload a diagram once (pack_bdd_layers' H2D step, enum.cu:438-593) and keep
 * it resident in HBM; the view may be freed afterwards. */
int pmdd_upload(pmdd_ctx *ctx, const pmdd_dd_view *view, pmdd_resident **out, char *err,
                size_t errlen);
/* Enumerate a resident diagram (inputs already in HBM). */
int pmdd_enumerate_resident(pmdd_ctx *ctx, pmdd_resident *dd, const pmdd_opts *opts,
                            pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen);
void pmdd_resident_free(pmdd_resident *dd);

/* Device bytes of the scratch array of the no-allocation entry point below:
 * each GPU the library runs on holds one such array for the life of the
 * process. */
#define PMDD_NOALLOC_SCRATCH_BYTES (128LL << 20)

/* No-allocation variant of pmdd_enumerate_resident.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The call makes no hipMalloc, hipFree or pinned host
 * allocation: every device buffer the enumeration needs beyond its input
 * (the resident diagram) is a block of the device's static scratch array
 * (PMDD_NOALLOC_SCRATCH_BYTES), buffers the context already holds are used
 * when large enough, and blocks the context stops using are freed by its
 * next allocating call.  The frontier (unsorted, row-major, d int32 per
 * point) goes to the caller's host buffer points[capacity * d]; *count
 * receives its size.
 *   PMDD_E_INVALID      capacity < *count (*count is set; call again with more)
 *   PMDD_E_UNSUPPORTED  the diagram's own tables exceed the scratch array
 *   PMDD_E_NOMEM        the enumeration outgrew the scratch array mid-run
 * Every call on one GPU shares that GPU's array, which keeps the last call's
 * contents: two no-allocation calls that overlap on one GPU (from different
 * contexts or threads) race; make them one at a time per device.  Adaptive
 * coupling batches (max_prod 0) stay at 512K products in this call. */
int pmdd_enumerate_resident_noalloc(pmdd_ctx *ctx, pmdd_resident *dd, const pmdd_opts *opts,
                                    int32_t *points, int64_t capacity, int64_t *count,
                                    pmdd_stats *stats, char *err, size_t errlen);

void pmdd_frontier_free(pmdd_frontier *f);

/* Engine build identity, e.g. "pmdd-engine gfx950 d=1..8". */
const char *pmdd_engine_version(void);

/* ---- serving: a stream of instances with continuous batching ----
 *
 * No counterpart in the reference (which enumerates one instance per
 * process).  A server owns `slots` worker contexts on one device, each with
 * its own stream and resident scratch (with two slots or more, a slot's
 * layers stay on its one stream).  Submitted requests queue in arrival
 * order; a request is admitted when a slot is free and its memory estimate
 * fits the budget next to the requests already running, so requests that
 * arrive while others are on the GPU overlap with them.  A request whose
 * estimate exceeds the whole budget completes with PMDD_E_REJECTED; any
 * failure stays with its own request. */
typedef struct pmdd_server_config {
    int32_t device;
    int32_t slots;         /* concurrent requests (>= 1); 1 = one at a time */
    int64_t memory_budget; /* bytes of HBM for admitted requests; 0 = 80% of free memory */
    /* Packed serving (1) or not (0).  When on, requests that run method 1
     * without state dominance on diagrams of at most 8,192 nodes per layer
     * and 1M nodes in all go to a packer of their objective count instead of
     * a slot: one host thread and one stream that expands the current layers
     * of up to pack_inflight in-flight requests as one packed layer per
     * step, so the nodes of many instances (at different layers) share every
     * launch.  Requests join a packer at any step as room and memory free up
     * and leave it at their terminal layer.  The other requests use the
     * slots. */
    int32_t pack;
    int32_t pack_inflight; /* requests per packer (0 = 32) */
} pmdd_server_config;

typedef struct pmdd_server pmdd_server;

#define PMDD_PENDING 1 /* pmdd_server_poll: request not finished yet */

int pmdd_server_create(const pmdd_server_config *cfg, pmdd_server **out, char *err,
                       size_t errlen);
/* Queue a request; the view's arrays are copied, so the caller may free
 * them on return.  *ticket identifies the request. */
int pmdd_server_submit(pmdd_server *srv, const pmdd_dd_view *view, const pmdd_opts *opts,
                       int64_t *ticket, char *err, size_t errlen);
/* PMDD_PENDING while running; otherwise the request's own return code, with
 * the frontier and statistics moved to the caller (the ticket is then
 * retired). */
int pmdd_server_poll(pmdd_server *srv, int64_t ticket, pmdd_frontier *out, pmdd_stats *stats,
                     char *err, size_t errlen);
/* As pmdd_server_poll, but blocks until the request has finished. */
int pmdd_server_wait(pmdd_server *srv, int64_t ticket, pmdd_frontier *out, pmdd_stats *stats,
                     char *err, size_t errlen);
/* Blocks until some request has finished and collects the one that finished
 * first among those not collected yet: *ticket names it, the return code is
 * its own.  PMDD_E_INVALID when no request is outstanding.  Lets one thread
 * collect results in completion order while others submit. */
int pmdd_server_wait_any(pmdd_server *srv, int64_t *ticket, pmdd_frontier *out,
                         pmdd_stats *stats, char *err, size_t errlen);
/* Finishes the running requests, drops queued ones, joins the workers and
 * frees the frontiers nobody collected.  No other call on srv may be in
 * progress (a thread blocked in wait or wait_any must have returned) or
 * follow. */
void pmdd_server_destroy(pmdd_server *srv);

/* Admission record of a server (no reference counterpart).  A slot reserves
 * its request's estimate while it runs and the device buffers its context
 * keeps while it is idle; the head of the queue is admitted when its
 * estimate fits the budget next to every other slot's reservation (or
 * nothing runs).  When it does not fit, idle slots release their kept
 * buffers.  A request that runs out of device memory while others run is
 * queued again at the head and runs alone. */
typedef struct pmdd_server_info {
    int64_t budget;         /* bytes */
    int64_t reserved_peak;  /* most bytes reserved at once */
    int64_t admitted;       /* requests started (a retry counts again) */
    int64_t rejected;       /* requests refused at submit (estimate above the budget) */
    int64_t memory_waits;   /* times the head of the queue waited for memory with a slot free */
    int64_t trims;          /* idle slots that released their kept buffers for the head */
    int64_t nomem_retries;  /* requests queued again to run alone after PMDD_E_NOMEM */
    int64_t packed;         /* requests served by a packer */
    int64_t pack_steps;     /* packed layers expanded */
    int64_t pack_requests_per_step_x100; /* 100 x the mean requests in a packed step */
    int32_t running_peak;   /* most requests running at once */
    int32_t slots;
} pmdd_server_info;

int pmdd_server_get_info(pmdd_server *srv, pmdd_server_info *out);

/* Testing hook (no reference counterpart): the nth device allocation the
 * engine makes from now on, in any context, fails as out of memory before
 * any kernel uses the buffer; once.  nth <= 0 disarms it. */
void pmdd_debug_fail_alloc(int64_t nth);

#ifdef __cplusplus
}
#endif

#endif /* PMDD_ENGINE_H_ */
