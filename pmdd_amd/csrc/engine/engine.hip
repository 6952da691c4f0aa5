// Frontier enumeration engine for MI355X (gfx950): host orchestration and the
// C ABI declared in include/pmdd_engine.h.
//
// Replaces the reference GPU backend src/pmdd/enum/gpu/:
//   enumerate_bdd_topdown      topdown.cu:552-1043      -> Run::topdown
//   enumerate_mdd_topdown      enum.cu:164-289          -> Run::topdown
//   coupled_bdd_cuda_enumerate enum.cu:595-756          -> Run::coupled
//   enumerate_mdd_coupled      enum.cu:295-436          -> Run::coupled
//   expand_layer_frontiers     bottomup.cu:242-533      -> Run::expand
//   apply_knapsack_state_dominance topdown.cu:363-521   -> Run::state_dominance
//   couple_cutsets_cuda        couple.cu:340-498        -> Run::couple
// with kernels of its own design (kernels.hpp).  Also method 2 on the GPU
// (bottomup_cpu_enumerate enum/cpu/enum.cpp:123-229, CPU only in the
// reference) -> Run::bottomup, and a serving loop with no reference
// counterpart (pmdd_server_*: worker slots with their own streams, memory
// admission).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/pmdd_engine.h"
#include "kernels.hpp"
#include "plan.hpp"

using pmdd_k::ChunkDesc;
using pmdd_k::kBlock;

// Scratch array of the no-allocation entry point
// (pmdd_enumerate_resident_noalloc): every GPU the library runs on holds one
// copy, at its own address, for the life of the process.  Host code never
// passes this name to a kernel or a copy: it is a host-side stand-in, and the
// device address comes from hipGetSymbolAddress (scratch_address below).
constexpr size_t kScratchBytes = static_cast<size_t>(PMDD_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct EngineError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg) { throw EngineError{code, msg}; }

void hip_check(hipError_t e, const char *where) {
    if (e == hipSuccess) {
        return;
    }
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PMDD_E_NOMEM
                                                                                  : PMDD_E_HIP;
    fail(code, std::string(where) + ": " + hipGetErrorString(e));
}

// Device allocations made so far (device memory use changes only with them).
static std::atomic<uint64_t> g_device_allocs{0};
// Testing hook: allocations left before one fails as out of memory (0: off).
static std::atomic<int64_t> g_fail_alloc_in{0};

// Device bytes held by one context's buffers, and their peak since the last
// reset.  A buffer is charged to the account that is current on the thread
// when it is allocated (the context running a call on that thread).
struct MemAcct {
    std::atomic<int64_t> cur{0};
    std::atomic<int64_t> peak{0};
    // What the current call asked for: every buffer it grew or reused counts
    // at the largest size the call asked of it (buffers kept from earlier
    // calls and not touched do not count).  Only the thread running the call
    // updates these.
    int64_t epoch = 0;
    int64_t need_cur = 0;
    int64_t need_peak = 0;
    void begin_call() {
        ++epoch;
        need_cur = 0;
        need_peak = 0;
    }
    void add(int64_t b) {
        const int64_t now = cur.fetch_add(b, std::memory_order_relaxed) + b;
        int64_t p = peak.load(std::memory_order_relaxed);
        while (now > p && !peak.compare_exchange_weak(p, now, std::memory_order_relaxed)) {
        }
    }
    void sub(int64_t b) { cur.fetch_sub(b, std::memory_order_relaxed); }
    void reset_peak() { peak.store(cur.load(std::memory_order_relaxed), std::memory_order_relaxed); }
};
static thread_local MemAcct *t_acct = nullptr;

// Makes `a` the current account of this thread for a scope.
struct AcctScope {
    MemAcct *prev;
    explicit AcctScope(MemAcct *a) : prev(t_acct) { t_acct = a; }
    ~AcctScope() { t_acct = prev; }
    AcctScope(const AcctScope &) = delete;
    AcctScope &operator=(const AcctScope &) = delete;
};

// No-allocation calls (pmdd_enumerate_resident_noalloc): every device buffer
// the call grows is carved from the device's scratch array instead of
// hipMalloc'd; allocations a buffer held before are not freed during the
// call but handed to the context's deferred list.
struct ScratchArena {
    char *base = nullptr; // device address of this device's scratch array
    size_t cap = 0;
    size_t used = 0;
    std::vector<void *> *deferred = nullptr; // hipMalloc'd blocks to free later
};
static thread_local ScratchArena *t_arena = nullptr;

// Frees a hipMalloc'd block, or defers it during a no-allocation call.
inline void free_block(void *p) {
    if (t_arena != nullptr && t_arena->deferred != nullptr) {
        t_arena->deferred->push_back(p);
    } else {
        (void)hipFree(p);
    }
}

// Device buffer; grows on demand, contents not kept unless asked.

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool in_scratch = false; // carved from the scratch array: never freed
    MemAcct *acct = nullptr;
    MemAcct *need_acct = nullptr; // account of the call that last asked for this buffer
    int64_t epoch = -1;           // that call
    int64_t call_need = 0;        // the most that call asked for
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept { take(o); }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            take(o);
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void take(DevBuf &o) noexcept {
        p = o.p;
        bytes = o.bytes;
        in_scratch = o.in_scratch;
        acct = o.acct;
        need_acct = o.need_acct;
        epoch = o.epoch;
        call_need = o.call_need;
        o.p = nullptr;
        o.bytes = 0;
        o.in_scratch = false;
        o.acct = nullptr;
        o.need_acct = nullptr;
        o.epoch = -1;
        o.call_need = 0;
    }
    void release() {
        if (p != nullptr && !in_scratch) {
            free_block(p);
            if (acct != nullptr) {
                acct->sub(static_cast<int64_t>(bytes));
            }
        }
        if (need_acct != nullptr && need_acct == t_acct && epoch == t_acct->epoch) {
            t_acct->need_cur -= call_need;
        }
        p = nullptr;
        bytes = 0;
        in_scratch = false;
        acct = nullptr;
        need_acct = nullptr;
        epoch = -1;
        call_need = 0;
    }
    // Forgets a scratch-array block (nothing to free).
    void drop_scratch() {
        if (in_scratch) {
            p = nullptr;
            bytes = 0;
            in_scratch = false;
            need_acct = nullptr;
            epoch = -1;
            call_need = 0;
        }
    }
    // Exchanges the allocations of two buffers (their accounts go with them).
    void swap(DevBuf &o) noexcept {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(in_scratch, o.in_scratch);
        std::swap(acct, o.acct);
        std::swap(need_acct, o.need_acct);
        std::swap(epoch, o.epoch);
        std::swap(call_need, o.call_need);
    }
    // Records that the current call asks for `need` bytes of this buffer.
    void note_need(size_t need) {
        MemAcct *a = t_acct;
        if (a == nullptr) {
            return;
        }
        if (need_acct != a || epoch != a->epoch) {
            need_acct = a;
            epoch = a->epoch;
            call_need = 0;
        }
        const int64_t n = static_cast<int64_t>(need);
        if (n > call_need) {
            a->need_cur += n - call_need;
            call_need = n;
            a->need_peak = std::max(a->need_peak, a->need_cur);
        }
    }
    // Grows by half again (with at most 64 MiB of slack), so a buffer sized
    // by a widening layer is not reallocated on every layer.
    void ensure(size_t need, const char *what) {
        if (need <= bytes && p != nullptr) {
            note_need(need);
            return;
        }
        if (t_arena != nullptr) {
            // no-allocation call: a block of the scratch array (256-aligned)
            const size_t nb = (std::max<size_t>(need, 256) + 255) & ~size_t(255);
            if (t_arena->used + nb > t_arena->cap) {
                fail(PMDD_E_NOMEM, std::string(what) + ": the no-allocation scratch array (" +
                                       std::to_string(t_arena->cap >> 20) +
                                       " MiB) is exhausted; use pmdd_enumerate_resident");
            }
            release();
            p = t_arena->base + t_arena->used;
            t_arena->used += nb;
            bytes = nb;
            in_scratch = true;
            note_need(need);
            return;
        }
        size_t nb = std::max<size_t>(need, 256);
        if (bytes > 0) {
            nb = std::max(nb, std::min(bytes + bytes / 2, need + (size_t(64) << 20)));
        }
        release();
        if (g_fail_alloc_in.load(std::memory_order_relaxed) > 0 &&
            g_fail_alloc_in.fetch_sub(1, std::memory_order_relaxed) == 1) {
            fail(PMDD_E_NOMEM, std::string(what) + ": injected allocation failure");
        }
        hip_check(hipMalloc(&p, nb), what);
        bytes = nb;
        acct = t_acct;
        if (acct != nullptr) {
            acct->add(static_cast<int64_t>(nb));
        }
        g_device_allocs.fetch_add(1, std::memory_order_relaxed);
        note_need(need); // release() took back what the call asked of the old one
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Frontier of one layer on the device: offsets per node + SoA points.
// ---------------------------------------------------------------------------
struct DevFrontier {
    int64_t n_nodes = 0;
    std::vector<int64_t> h_off; // [n_nodes + 1], when on_host
    DevBuf off;
    DevBuf pts;
    int64_t cap = 0; // SoA stride
    // A layer placed on the device has its offsets there only (h_off empty)
    // until the host asks for them; `step` is the expansion that built it.
    bool on_host = true;
    int step = -1;
    int64_t total() const { return h_off.empty() ? 0 : h_off.back(); }
    int64_t size(int64_t v) const { return h_off[v + 1] - h_off[v]; }
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

} // namespace

// ---------------------------------------------------------------------------
// Context and resident diagram
// ---------------------------------------------------------------------------
// Device tables of one chunk list: descriptors, per-chunk survivor counts
// (local and after cross checks), key range of the survivors, output slots.
struct ChunkBufs {
    DevBuf chunks, nsurv, nalive, out;
};

// Pinned host memory for the small host<->device copies of an enumeration
// (chunk tables, offsets, counts): copies from it are asynchronous, so the
// host queues a layer's work without waiting.  A bump allocator: regions are
// handed out until the next stream synchronization, which frees them all.
struct Staging {
    char *base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    Staging() = default;
    Staging(const Staging &) = delete;
    Staging &operator=(const Staging &) = delete;
    ~Staging() {
        if (base != nullptr) {
            (void)hipHostFree(base);
        }
    }
    // n bytes (256-aligned), or nullptr when the arena has no room left.
    void *take(size_t n) {
        const size_t at = (used + 255) & ~static_cast<size_t>(255);
        if (base == nullptr || at + n > cap) {
            return nullptr;
        }
        used = at + n;
        return base + at;
    }
    // Only when no copy from the arena is in flight.
    void reset() { used = 0; }
    void grow(size_t need) {
        if (base != nullptr) {
            (void)hipHostFree(base);
            base = nullptr;
        }
        cap = std::max<size_t>({need, 2 * cap, size_t(1) << 20});
        if (hipHostMalloc(reinterpret_cast<void **>(&base), cap, hipHostMallocDefault) !=
            hipSuccess) {
            base = nullptr;
            cap = 0;
        }
        used = 0;
    }
};

struct pmdd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for the small-node filters of a
    // device-planned layer, which run beside its big nodes' path; the main
    // stream waits for it (ev_join) before the layer's placement
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr; // untimed
    // whether layers fork onto the side stream: not in the contexts of a
    // server with several slots, whose streams already keep the chip busy
    // and share its few hardware queues (a fork's event wait in a shared
    // queue holds back the other slots' launches behind it)
    bool use_side = true;
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_c = nullptr, ev_d = nullptr;
    hipEvent_t ev_e = nullptr, ev_f = nullptr, ev_g = nullptr;
    // scratch reused across layers and calls
    ChunkBufs cb_main, cb_aux;
    DevBuf run_off, which, which_a, which_b, which_c, which_s, which_m;
    DevBuf surv_pts, alive, counter, join_meta, keep_a, keep_b;
    DevBuf sort_a, sort_b, big_base, big_node, run_start, grp_run;
    DevBuf join_x, join_y, join_ctr, front_count, merge_a, merge_b;
    DevBuf join_flag; // a batch's product flags for the check against R's later blocks
    // device-planned layers: per sweep direction the plan of its next layer
    // (arc run offsets, candidate counts, workgroup prefixes); the block the
    // host reads back (headers and scores) and its pinned mirror; the first
    // chunk of each node, the prefix scratch of the survivor placement, the
    // per-expansion figures and run counters
    DevBuf plan_run_off[2], plan_gamma[2], plan_part[2];
    DevBuf plan_rb, node_chunk, scan_part, step_stats, run_ctr;
    DevBuf sd_flag, sd_pos, sd_active, sd_cmp; // device-planned state dominance
    DevBuf batch_ctr;                          // a coupling batch's device counts
    // the readback block in host memory the device writes (mapped,
    // coherent), its device address, and the last sequence number published
    pmdd_k::PlanReadback *h_rb = nullptr;
    pmdd_k::PlanReadback *d_rb_host = nullptr;
    unsigned long long rb_seq = 0;
    std::vector<hipEvent_t> ev_pool; // start/end pairs around single launches
    std::vector<hipEvent_t> ev_late; // start/end pairs read after a later wait
    Staging staging;
    // frontier buffers handed from one enumeration to the next (FrontierLease)
    std::vector<DevFrontier> frontier_pool;
    // this device's scratch array (no-allocation calls) and the hipMalloc'd
    // blocks such a call stopped using, freed by the next allocating call
    char *scratch = nullptr;
    std::vector<void *> deferred;
    // launch shapes per objective count (index d)
    pmdd_tuning tuning[9];
    // device bytes held by this context's buffers
    MemAcct acct;
    bool in_call = false; // inside a one-shot pmdd_enumerate (its peak includes the upload)
    // queued passes (packed layers): the tier descriptors, the queue heads,
    // and the captured graph of a pass's launches with what it was captured
    // for (objective count, the two buffers and the pair-test counter)
    DevBuf qtier, qhead;
    hipGraphExec_t pass_graph = nullptr;
    const void *graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t graph_replays = 0;
    // largest adaptive coupling batch (products); 0: kJoinBatchCap.  A server
    // slot sets it so that a request's join stays within its admission
    // estimate.
    int64_t join_cap = 0;

    // Every device buffer the context keeps between calls.
    template <class F> void for_each_buffer(F &&f) {
        for (ChunkBufs *cb : {&cb_main, &cb_aux}) {
            for (DevBuf *b : {&cb->chunks, &cb->nsurv, &cb->nalive, &cb->out}) {
                f(*b);
            }
        }
        for (DevBuf *b : {&run_off, &which, &which_a, &which_b, &which_c, &which_s, &which_m,
                          &surv_pts, &alive, &counter, &join_meta, &keep_a, &keep_b, &sort_a,
                          &sort_b, &big_base, &big_node, &run_start, &grp_run, &join_x,
                          &join_y, &join_ctr, &front_count, &merge_a, &merge_b, &qtier, &qhead,
                          &plan_run_off[0], &plan_run_off[1], &plan_gamma[0], &plan_gamma[1],
                          &plan_part[0], &plan_part[1], &plan_rb, &node_chunk, &scan_part,
                          &step_stats, &run_ctr, &sd_flag, &sd_pos, &sd_active, &sd_cmp,
                          &batch_ctr, &join_flag}) {
            f(*b);
        }
        for (DevFrontier &fr : frontier_pool) {
            f(fr.off);
            f(fr.pts);
        }
    }
    // Frees the blocks a no-allocation call stopped using (stream idle).
    void free_deferred() {
        for (void *p : deferred) {
            (void)hipFree(p);
        }
        deferred.clear();
    }
    // Releases every device buffer the context keeps between calls (the
    // stream must be idle).
    void drop_graph() {
        if (pass_graph != nullptr) {
            (void)hipGraphExecDestroy(pass_graph);
            pass_graph = nullptr;
        }
        for (const void *&k : graph_key) {
            k = nullptr;
        }
    }
    void trim() {
        drop_graph();
        for_each_buffer([](DevBuf &b) { b.release(); });
        frontier_pool.clear();
        free_deferred();
    }
};

// A frontier borrowed from the context's pool for one enumeration and handed
// back at its end, so a context's frontier buffers grow once, not per call.
// The SoA stride is recomputed for this enumeration's objective count.
struct FrontierLease {
    pmdd_ctx *ctx;
    DevFrontier f;
    FrontierLease(pmdd_ctx *c, int d) : ctx(c) {
        if (!c->frontier_pool.empty()) {
            f = std::move(c->frontier_pool.back());
            c->frontier_pool.pop_back();
        }
        f.n_nodes = 0;
        f.h_off.clear();
        f.on_host = true;
        f.step = -1;
        f.cap = f.pts.p != nullptr ? static_cast<int64_t>(f.pts.bytes / (4 * static_cast<size_t>(d)))
                                   : 0;
    }
    ~FrontierLease() { ctx->frontier_pool.push_back(std::move(f)); }
    FrontierLease(const FrontierLease &) = delete;
    FrontierLease &operator=(const FrontierLease &) = delete;
};

struct LayerCSR {
    // host copies (planning) and device arrays
    std::vector<int64_t> off; // [N + 1] layer-local arc ids
    std::vector<int32_t> src; // [E]
    int64_t n_arcs = 0;
    int64_t *d_off = nullptr;
    int32_t *d_src = nullptr;
    int32_t *d_w = nullptr; // SoA [d][E]
};

struct pmdd_resident {
    int device = 0;
    int d = 0;
    int L = 0;
    int kind = PMDD_DD_BDD;
    int root = 0;
    int terminal = 0;
    std::vector<int64_t> layer_size;
    std::vector<LayerCSR> td; // td[l]: arcs into layer l (l >= 1)
    std::vector<LayerCSR> bu; // bu[l]: arcs out of layer l (l <= L-2), "src" = child in l+1
    std::vector<std::vector<int32_t>> min_weight;    // per layer, knapsack only
    std::vector<std::vector<int32_t>> mw_order;      // per layer: nodes by descending
                                                     // min_weight (stable)
    std::vector<std::vector<int32_t>> single_parent; // per layer: parent index or -1
    std::vector<std::vector<int32_t>> single_arc;    // per layer: arc value of that parent
    std::vector<std::vector<int32_t>> sd_partner;    // per layer, set packing (or empty)
    DevBuf arena; // CSR arrays of both directions, then the state dominance tables
    // device copies of the state dominance tables, per layer (null: none)
    std::vector<const int32_t *> d_mw_order, d_single_parent, d_single_arc, d_partner;
};

namespace {

// ---------------------------------------------------------------------------
// Upload: pack the view into per-layer CSR in both directions
// (pack_bdd_layers enum.cu:438-593, pack_mdd_layers enum.cu:85-158).
// ---------------------------------------------------------------------------
// The checks of a view that are linear in its layers and nodes.
void validate_nodes(const pmdd_dd_view *v) {
    if (v == nullptr || v->layer_off == nullptr || v->in_off == nullptr) {
        fail(PMDD_E_INVALID, "decision diagram view is NULL or incomplete");
    }
    if (v->d < 1 || v->d > 8) {
        fail(PMDD_E_UNSUPPORTED,
             "objective count " + std::to_string(v->d) + " outside the engine's 1..8");
    }
    if (v->num_layers < 2) {
        fail(PMDD_E_INVALID, "decision diagram needs at least two layers");
    }
    if (v->kind != PMDD_DD_BDD && v->kind != PMDD_DD_MDD) {
        fail(PMDD_E_INVALID, "unknown decision diagram kind");
    }
    const int L = v->num_layers;
    if (v->layer_off[0] != 0) {
        fail(PMDD_E_INVALID, "layer_off[0] must be 0");
    }
    for (int l = 0; l < L; ++l) {
        if (v->layer_off[l + 1] <= v->layer_off[l]) {
            fail(PMDD_E_INVALID, "layer " + std::to_string(l) + " has no nodes");
        }
        if (v->layer_off[l + 1] - v->layer_off[l] > INT32_MAX) {
            fail(PMDD_E_UNSUPPORTED, "layer wider than 2^31 nodes");
        }
    }
    if (v->root < 0 || v->root >= v->layer_off[1] || v->terminal < 0 ||
        v->terminal >= v->layer_off[L] - v->layer_off[L - 1]) {
        fail(PMDD_E_INVALID, "root or terminal index out of range");
    }
    const int64_t total = v->layer_off[L];
    for (int64_t g = 0; g < v->layer_off[1]; ++g) {
        if (v->in_off[g + 1] != v->in_off[g]) {
            fail(PMDD_E_INVALID, "layer 0 nodes cannot have in-arcs");
        }
    }
    for (int64_t g = 0; g < total; ++g) {
        if (v->in_off[g + 1] < v->in_off[g]) {
            fail(PMDD_E_INVALID, "in_off must be non-decreasing");
        }
    }
    const int64_t A = v->in_off[total];
    if (A > 0 && (v->arc_src == nullptr || v->arc_w == nullptr || v->arc_val == nullptr)) {
        fail(PMDD_E_INVALID, "arc arrays missing");
    }
    // A set-packing state-dominance partner is -1 or another node of the same
    // layer (a node's own labels would all drop against themselves).
    if (v->sd_partner != nullptr) {
        for (int l = 0; l < L; ++l) {
            const int64_t n = v->layer_off[l + 1] - v->layer_off[l];
            const int32_t *pl = v->sd_partner + v->layer_off[l];
            for (int64_t i = 0; i < n; ++i) {
                if (pl[i] < -1 || pl[i] >= n || pl[i] == i) {
                    fail(PMDD_E_INVALID,
                         "state-dominance partner out of range in layer " + std::to_string(l));
                }
            }
        }
    }
}

// The arcs into layer l (1..L-1): sources in the previous layer, BDD arc
// values 0 or 1 (the bottom-up CSR is grouped so).  maxw[o] receives the
// largest |weight| of objective o over these arcs.
void validate_layer(const pmdd_dd_view *v, int l, int64_t *maxw) {
    const int64_t prev_n = v->layer_off[l] - v->layer_off[l - 1];
    const int64_t a0 = v->in_off[v->layer_off[l]], a1 = v->in_off[v->layer_off[l + 1]];
    const int d = v->d;
    const int32_t *src = v->arc_src;
    const int32_t *val = v->arc_val;
    const bool bdd = v->kind == PMDD_DD_BDD;
    bool bad_src = false, bad_val = false;
    int64_t mx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t a = a0; a < a1; ++a) {
        bad_src |= src[a] < 0 || src[a] >= prev_n;
        bad_val |= bdd && (val[a] & ~1) != 0;
        const int32_t *w = v->arc_w + a * d;
        for (int o = 0; o < d; ++o) {
            mx[o] = std::max<int64_t>(mx[o], std::llabs(static_cast<long long>(w[o])));
        }
    }
    if (bad_src) {
        fail(PMDD_E_INVALID, "arc source out of range in layer " + std::to_string(l));
    }
    if (bad_val) {
        fail(PMDD_E_INVALID, "BDD arc value other than 0 or 1 in layer " + std::to_string(l));
    }
    for (int o = 0; o < d; ++o) {
        maxw[o] = mx[o];
    }
}

// Every path sum stays inside int32: per objective, the sum over layers of
// the largest |arc weight| entering the layer bounds any partial or full path
// sum (the reference adds unchecked, bottomup.cu:124, couple.cu:93).  maxw
// holds validate_layer's figures, d per layer.
void check_path_bound(const pmdd_dd_view *v, const std::vector<int64_t> &maxw) {
    const int d = v->d;
    for (int o = 0; o < d; ++o) {
        int64_t bound = 0;
        for (int l = 1; l < v->num_layers; ++l) {
            bound += maxw[static_cast<size_t>(l) * d + o];
        }
        if (bound > INT32_MAX) {
            fail(PMDD_E_OVERFLOW, "objective " + std::to_string(o) +
                                      " path sums can exceed int32 (bound " +
                                      std::to_string(bound) + ")");
        }
    }
}

void validate_view(const pmdd_dd_view *v) {
    validate_nodes(v);
    std::vector<int64_t> maxw(static_cast<size_t>(v->num_layers) * v->d, 0);
    for (int l = 1; l < v->num_layers; ++l) {
        validate_layer(v, l, &maxw[static_cast<size_t>(l) * v->d]);
    }
    check_path_bound(v, maxw);
}

inline void *staging_take(pmdd_ctx *ctx, size_t n);

// Diagrams whose device arena is larger than this are packed into pageable
// host memory; smaller ones into the context's pinned staging, which keeps
// that size for later uploads.
constexpr size_t kPinnedUploadMax = size_t(256) << 20;

// The view packed into one device arena, laid out in host memory first and
// copied in one transfer: per layer l its top-down CSR (arcs into l, l >= 1:
// offsets, sources, weights SoA [d][E]) and its bottom-up CSR (arcs out of
// l, l <= L-2, grouped by tail, "src" = head in l+1), then the state
// dominance tables of every layer.  Each layer is checked and packed by one
// of a few host threads.
pmdd_resident *upload(pmdd_ctx *ctx, const pmdd_dd_view *v) {
    validate_nodes(v);
    std::unique_ptr<pmdd_resident> r(new pmdd_resident());
    const int L = v->num_layers;
    const int d = v->d;
    r->device = ctx->device;
    r->d = d;
    r->L = L;
    r->kind = v->kind;
    r->root = v->root;
    r->terminal = v->terminal;
    r->layer_size.resize(L);
    std::vector<int64_t> E(L + 1, 0); // E[l]: arcs into layer l
    for (int l = 0; l < L; ++l) {
        r->layer_size[l] = v->layer_off[l + 1] - v->layer_off[l];
        E[l] = v->in_off[v->layer_off[l + 1]] - v->in_off[v->layer_off[l]];
    }
    r->td.resize(L);
    r->bu.resize(L);
    const bool knap = v->min_weight != nullptr;
    const bool part = v->sd_partner != nullptr;

    // arena layout
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    struct Place {
        size_t off = 0, src = 0, w = 0;
    };
    std::vector<Place> td_at(L), bu_at(L);
    std::vector<size_t> sd_at(L, 0); // first table of layer l
    size_t bytes = 0;
    auto place = [&](Place &p, int64_t n, int64_t e) {
        p.off = bytes;
        bytes += align(static_cast<size_t>(n + 1) * 8);
        p.src = bytes;
        bytes += align(static_cast<size_t>(e) * 4);
        p.w = bytes;
        bytes += align(static_cast<size_t>(d) * e * 4);
    };
    for (int l = 0; l < L; ++l) {
        if (l >= 1) {
            place(td_at[l], r->layer_size[l], E[l]);
        }
        if (l <= L - 2) {
            place(bu_at[l], r->layer_size[l], E[l + 1]);
        }
    }
    const int tables = (knap ? 3 : 0) + (part ? 1 : 0);
    for (int l = 0; l < L; ++l) {
        sd_at[l] = bytes;
        bytes += static_cast<size_t>(tables) * align(static_cast<size_t>(r->layer_size[l]) * 4);
    }
    bytes = std::max<size_t>(bytes, 256);

    std::unique_ptr<char[]> pageable; // not cleared: every byte read is written
    char *host = nullptr;
    if (bytes <= kPinnedUploadMax) {
        host = static_cast<char *>(staging_take(ctx, bytes));
    }
    if (host == nullptr) {
        pageable.reset(new char[bytes]);
        host = pageable.get();
    }
    if (knap) {
        r->min_weight.resize(L);
        r->mw_order.resize(L);
        r->single_parent.resize(L);
        r->single_arc.resize(L);
    }
    if (part) {
        r->sd_partner.resize(L);
    }
    std::vector<int64_t> maxw(static_cast<size_t>(L) * d, 0);

    // layer l: its arcs checked, its top-down CSR, the bottom-up CSR of layer
    // l-1 (the same arcs grouped by tail), its state dominance tables
    auto pack_layer = [&](int l) {
        const int64_t n = r->layer_size[l];
        const int64_t g0 = v->layer_off[l];
        if (l >= 1) {
            validate_layer(v, l, &maxw[static_cast<size_t>(l) * d]);
            const int64_t a0 = v->in_off[g0];
            const int64_t e_n = E[l];
            LayerCSR &c = r->td[l];
            c.n_arcs = e_n;
            c.off.resize(n + 1);
            int64_t *hoff = reinterpret_cast<int64_t *>(host + td_at[l].off);
            for (int64_t i = 0; i <= n; ++i) {
                c.off[i] = v->in_off[g0 + i] - a0;
                hoff[i] = c.off[i];
            }
            c.src.assign(v->arc_src + a0, v->arc_src + a0 + e_n);
            if (e_n > 0) {
                std::memcpy(host + td_at[l].src, c.src.data(), static_cast<size_t>(e_n) * 4);
            }
            int32_t *tw = reinterpret_cast<int32_t *>(host + td_at[l].w);
            for (int64_t e = 0; e < e_n; ++e) {
                const int32_t *w = v->arc_w + (a0 + e) * d;
                for (int o = 0; o < d; ++o) {
                    tw[o * e_n + e] = w[o];
                }
            }
            // out-arcs grouped by tail in the reference merge order: BDD
            // 1-arcs then 0-arcs, MDD by arc value, each in in-arc order
            LayerCSR &b = r->bu[l - 1];
            const int64_t np = r->layer_size[l - 1];
            b.n_arcs = e_n;
            b.off.assign(np + 1, 0);
            for (int64_t e = 0; e < e_n; ++e) {
                b.off[c.src[e] + 1]++;
            }
            for (int64_t i = 0; i < np; ++i) {
                b.off[i + 1] += b.off[i];
            }
            std::vector<int64_t> order(e_n);
            std::vector<int64_t> pos(b.off.begin(), b.off.end() - 1);
            const int32_t *val = v->arc_val + a0;
            if (v->kind == PMDD_DD_BDD) {
                for (int t = 1; t >= 0; --t) {
                    for (int64_t e = 0; e < e_n; ++e) {
                        if (val[e] == t) {
                            order[pos[c.src[e]]++] = e;
                        }
                    }
                }
            } else {
                for (int64_t e = 0; e < e_n; ++e) {
                    order[pos[c.src[e]]++] = e;
                }
                for (int64_t i = 0; i < np; ++i) {
                    std::stable_sort(order.begin() + b.off[i], order.begin() + b.off[i + 1],
                                     [&](int64_t x, int64_t y) { return val[x] < val[y]; });
                }
            }
            std::vector<int32_t> head(e_n);
            for (int64_t i = 0; i < n; ++i) {
                for (int64_t e = c.off[i]; e < c.off[i + 1]; ++e) {
                    head[e] = static_cast<int32_t>(i);
                }
            }
            std::memcpy(host + bu_at[l - 1].off, b.off.data(), static_cast<size_t>(np + 1) * 8);
            b.src.resize(e_n);
            int32_t *bw = reinterpret_cast<int32_t *>(host + bu_at[l - 1].w);
            for (int64_t k = 0; k < e_n; ++k) {
                const int64_t e = order[k];
                b.src[k] = head[e];
                for (int o = 0; o < d; ++o) {
                    bw[o * e_n + k] = tw[o * e_n + e];
                }
            }
            if (e_n > 0) {
                std::memcpy(host + bu_at[l - 1].src, b.src.data(), static_cast<size_t>(e_n) * 4);
            }
        }
        // state dominance tables: knapsack nodes by descending min_weight
        // (stable) with their single parent and its arc value
        // (pack_bdd_layers enum.cu:496-540); set-packing partners
        size_t at = sd_at[l];
        auto put = [&](const std::vector<int32_t> &t) {
            std::memcpy(host + at, t.data(), t.size() * 4);
            at += align(t.size() * 4);
        };
        if (knap) {
            r->min_weight[l].assign(v->min_weight + g0, v->min_weight + g0 + n);
            const std::vector<int32_t> &mw = r->min_weight[l];
            // keys (descending min_weight, ascending index): a stable order
            std::vector<uint64_t> key(n);
            for (int64_t i = 0; i < n; ++i) {
                const uint32_t up = static_cast<uint32_t>(mw[i]) ^ 0x80000000u;
                key[i] = (static_cast<uint64_t>(~up) << 32) | static_cast<uint64_t>(i);
            }
            std::sort(key.begin(), key.end());
            std::vector<int32_t> &ord = r->mw_order[l];
            ord.resize(n);
            for (int64_t i = 0; i < n; ++i) {
                ord[i] = static_cast<int32_t>(key[i] & 0xffffffffu);
            }
            r->single_parent[l].assign(n, -1);
            r->single_arc[l].assign(n, -1);
            for (int64_t i = 0; i < n; ++i) {
                const int64_t a0 = v->in_off[g0 + i], a1 = v->in_off[g0 + i + 1];
                if (a1 - a0 == 1) {
                    r->single_parent[l][i] = v->arc_src[a0];
                    r->single_arc[l][i] = v->arc_val[a0];
                }
            }
            put(ord);
            put(r->single_parent[l]);
            put(r->single_arc[l]);
        }
        if (part) {
            r->sd_partner[l].assign(v->sd_partner + g0, v->sd_partner + g0 + n);
            put(r->sd_partner[l]);
        }
    };
    {
        // a thread per 16K arcs, at most 8; the first failure in layer order
        // is the one reported
        const int64_t arcs = v->in_off[v->layer_off[L]];
        const int workers = static_cast<int>(std::min<int64_t>(
            {8, std::max<int64_t>(1, std::thread::hardware_concurrency()),
             std::max<int64_t>(1, arcs >> 14), L}));
        std::vector<std::exception_ptr> failure(L);
        std::atomic<int> next_layer(0);
        auto work = [&] {
            for (int l = next_layer++; l < L; l = next_layer++) {
                try {
                    pack_layer(l);
                } catch (...) {
                    failure[l] = std::current_exception();
                }
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < workers; ++w) {
            pool.emplace_back(work);
        }
        work();
        for (std::thread &t : pool) {
            t.join();
        }
        for (const std::exception_ptr &e : failure) {
            if (e) {
                std::rethrow_exception(e);
            }
        }
    }
    check_path_bound(v, maxw);

    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    r->arena.ensure(bytes, "hipMalloc(diagram arena)");
    char *base = r->arena.as<char>();
    for (int l = 0; l < L; ++l) {
        if (l >= 1) {
            r->td[l].d_off = reinterpret_cast<int64_t *>(base + td_at[l].off);
            r->td[l].d_src = reinterpret_cast<int32_t *>(base + td_at[l].src);
            r->td[l].d_w = reinterpret_cast<int32_t *>(base + td_at[l].w);
        }
        if (l <= L - 2) {
            r->bu[l].d_off = reinterpret_cast<int64_t *>(base + bu_at[l].off);
            r->bu[l].d_src = reinterpret_cast<int32_t *>(base + bu_at[l].src);
            r->bu[l].d_w = reinterpret_cast<int32_t *>(base + bu_at[l].w);
        }
    }
    r->d_mw_order.assign(L, nullptr);
    r->d_single_parent.assign(L, nullptr);
    r->d_single_arc.assign(L, nullptr);
    r->d_partner.assign(L, nullptr);
    for (int l = 0; l < L; ++l) {
        const size_t t = align(static_cast<size_t>(r->layer_size[l]) * 4);
        size_t at = sd_at[l];
        if (knap) {
            r->d_mw_order[l] = reinterpret_cast<const int32_t *>(base + at);
            r->d_single_parent[l] = reinterpret_cast<const int32_t *>(base + at + t);
            r->d_single_arc[l] = reinterpret_cast<const int32_t *>(base + at + 2 * t);
            at += 3 * t;
        }
        if (part) {
            r->d_partner[l] = reinterpret_cast<const int32_t *>(base + at);
        }
    }
    hip_check(hipMemcpyAsync(base, host, bytes, hipMemcpyHostToDevice, ctx->stream),
              "hipMemcpy(diagram arena)");
    hip_check(hipStreamSynchronize(ctx->stream), "upload sync");
    return r.release();
}


// Grow a SoA point buffer to hold `need` points, keeping the first `keep`.
void grow_points(DevFrontier &f, int d, int64_t need, int64_t keep, hipStream_t s) {
    if (need <= f.cap && f.pts.p != nullptr) {
        f.pts.note_need(static_cast<size_t>(need) * d * 4);
        return;
    }
    const int64_t ncap = std::max<int64_t>(need, f.cap + f.cap / 2);
    DevBuf nb;
    nb.ensure(static_cast<size_t>(ncap) * d * 4, "hipMalloc(frontier points)");
    if (keep > 0) {
        for (int o = 0; o < d; ++o) {
            hip_check(hipMemcpyAsync(nb.as<int32_t>() + o * ncap, f.pts.as<int32_t>() + o * f.cap,
                                     static_cast<size_t>(keep) * 4, hipMemcpyDeviceToDevice, s),
                      "hipMemcpy(frontier grow)");
        }
    }
    f.pts.swap(nb);
    f.cap = ncap;
    hip_check(hipStreamSynchronize(s), "frontier grow sync");
}



// ---------------------------------------------------------------------------
// Kernel launchers, one instantiation per objective count.
// ---------------------------------------------------------------------------
using FilterFn = void (*)(dim3, hipStream_t, const int32_t *, const ChunkDesc *, const void *,
                          int32_t *, int64_t, uint8_t *, int32_t *, int32_t *,
                          unsigned long long *);
using CrossFn = void (*)(dim3, hipStream_t, const int32_t *, const ChunkDesc *, const int32_t *,
                         int64_t, uint8_t *, const int32_t *, const int32_t *, int64_t,
                         const int64_t *, int, int, int, unsigned long long *);
using AppendFn = void (*)(dim3, hipStream_t, const int32_t *, const ChunkDesc *, const int32_t *,
                          int64_t, const uint8_t *, const int32_t *, int32_t *, int32_t *, int64_t,
                          const int64_t *, int64_t *, int, int, unsigned long long *);

struct Launch {
    int chunk_cap;
    // ExpandSrc chunk filters by tier (nodes of at most kSmallNode
    // candidates, of at most mid_cap, of at most chunk_cap) and threads per
    // workgroup (128, 256, 512): the shapes the tuning picks from
    FilterFn filter_tier[3][3];
    int mid_cap;
    FilterFn filter_join;    // JoinSrc, presorted in LDS
    FilterFn filter_sorted;  // SortedSrc, source already in key order
    FilterFn filter_dense;   // SortedSrc in any order, presorted in LDS
    CrossFn cross;           // chunks of a group are descending key blocks
    AppendFn append_front;   // a wave's survivors onto their group's front
    void (*stage)(dim3, hipStream_t, const int32_t *, const ChunkDesc *, pmdd_k::SortedSrc,
                  int32_t *, int64_t, uint8_t *, int32_t *, int32_t *);
    void (*compact)(dim3, hipStream_t, const ChunkDesc *, const int32_t *, const int64_t *,
                    const int32_t *, int64_t, const uint8_t *, int32_t *, int64_t);
    void (*state_dom)(dim3, hipStream_t, const int32_t *, const int64_t *, const int64_t *,
                      const int32_t *, int64_t, uint8_t *, unsigned long long *);
    void (*compact_nodes)(dim3, hipStream_t, const int64_t *, const int32_t *, int64_t,
                          const uint8_t *, const int64_t *, int32_t *, int64_t);
    void (*gather_runs)(dim3, hipStream_t, const int64_t *, const int32_t *, int32_t,
                        pmdd_k::ExpandSrc, int32_t *, int64_t);
    void (*merge_pass)(dim3, hipStream_t, const int32_t *, int32_t *, int64_t, const int64_t *,
                       const int32_t *, int32_t, int, int64_t);
    void (*join_products)(dim3, hipStream_t, const pmdd_k::JoinSrc &, int64_t, int32_t *,
                          int64_t);
    void (*refill_vs)(dim3, hipStream_t, const int32_t *, int64_t, const unsigned long long *,
                      const int32_t *, int64_t, int64_t, int64_t, int32_t *, int64_t,
                      unsigned long long *, unsigned long long *);
    void (*refill_flags)(dim3, hipStream_t, const int32_t *, int64_t,
                         const unsigned long long *, const int32_t *, int64_t, int64_t, int64_t,
                         const int64_t *, uint8_t *, unsigned long long *);
    // the flag check over blocks of blen rows from lo, one grid row per block
    void (*refill_flag_rows)(dim3, hipStream_t, const int32_t *, int64_t,
                             const unsigned long long *, const int32_t *, int64_t, int64_t,
                             int64_t, const int64_t *, int64_t, uint8_t *, unsigned long long *);
    void (*flag_append)(dim3, hipStream_t, const int32_t *, int64_t, const unsigned long long *,
                        const uint8_t *, int32_t *, int64_t, unsigned long long *);
    // persistent chunk filters of a queued pass, one per tier (nodes of at
    // most kSmallNode candidates, of at most mid_cap, of at most CH), with
    // the grid that keeps every compute unit full for each
    void (*queue[3])(dim3, hipStream_t, const pmdd_k::QueueTier *, int32_t *,
                     unsigned long long *);
    int queue_grid[3];
    void (*queue_grids)(int *); // fills queue_grid (queries the current device)
};

// Nodes of at most kSmallNode candidates are filtered by an instantiation with
// an LDS footprint of kSmallNode points (about 9 KB instead of 40-72 KB), so
// that many of their workgroups fit one compute unit at once.
constexpr int kSmallNode = 256;
constexpr int kMidNode = 1024; // a middle tier where a full chunk holds 2048 points
// Layers of fewer nodes than this send every node above kSmallNode
// candidates to the big-node path (several workgroups per chunk).
constexpr int64_t kNarrowLayer = 256;
// Threads of the chunk filter: 128 for nodes of at most 256 candidates (most
// hold far fewer, so half the waves of a 256-thread workgroup would idle);
// for chunks of 1024 or more candidates: their
// LDS allows two workgroups per CU, so more threads means more waves to hide
// the LDS latency of the survivor scans (512 measured faster than 256 and 1024;
// a step of 1024 candidates tests too many peers that die anyway)
constexpr int kFilterThreadsBig = 512;

// Coupling: blocks of R after the first 2048 points hold kRefillBlock points,
// and so do the grid rows of the merged tail launch and of the check of R
// against a batch frontier (4096 measured fastest there, against 1024, 2048
// and 8192; when every block was its own launch, 2048 had beaten 8192);
// k_refill_vs_block runs at most kRefillWaves one-wave workgroups: 32 per CU,
// 8 per SIMD, which its register use allows (4096 measured 4% slower, 16384
// no faster).
constexpr int64_t kRefillBlock = 4096;
// A batch checks its products against R's first kSerialBlocks blocks one
// launch each, each launch keeping only the survivors; the products left
// then meet every later block in one launch (a grid row per block, deaths
// flagged), at most kTailWaves waves per block, and the unflagged ones are
// appended.  One block (R's first 512 points) first measured fastest: with
// none, products that R's first points kill within a few rows also walk
// every later block; with more, each later launch is a few waves walking
// one block alone.  2048 tail waves beat 512, 1024 and 4096.
constexpr int kSerialBlocks = 1;
// Workgroups per state-dominance target, each taking every kSdSplit-th tile
// of the comparison labels.
constexpr unsigned kSdSplit = 4;
constexpr int64_t kTailWaves = 2048;
constexpr int64_t kRefillWaves = 8192;
// Largest coupling batch (products) once the first batches have built R.
constexpr int64_t kJoinBatchCap = int64_t(1) << 25;
// The same for requests of a server (adaptive batches): their admission
// estimate covers a batch of at most this many products.
constexpr int64_t kServeJoinCap = int64_t(1) << 19;

template <int D> struct LaunchImpl {
    static constexpr int CH = pmdd_k::Cap<D>::CH;
    template <class Src, bool PRE, int C = CH,
              int NT = (C >= 1024 ? kFilterThreadsBig : (C <= 256 ? 128 : pmdd_k::kBlock))>
    static void filter(dim3 g, hipStream_t s, const int32_t *which, const ChunkDesc *c,
                       const void *src, int32_t *out, int64_t stride, uint8_t *alive, int32_t *ns,
                       int32_t *na, unsigned long long *pt) {
        hipLaunchKernelGGL((pmdd_k::k_filter_chunks<D, C, Src, PRE, NT>), g, dim3(NT), 0, s,
                           which, c, *static_cast<const Src *>(src), out, stride, alive, ns, na,
                           pt);
    }
    static void cross(dim3 g, hipStream_t s, const int32_t *which, const ChunkDesc *c,
                      const int32_t *pts, int64_t stride, uint8_t *alive, const int32_t *ns,
                      const int32_t *fv, int64_t fvs, const int64_t *fc, int W, int G,
                      int only_wave, unsigned long long *pt) {
        hipLaunchKernelGGL((pmdd_k::k_cross_chunks<D, CH>), g, dim3(kBlock), 0, s, which, c, pts,
                           stride, alive, ns, fv, fvs, fc, W, G, only_wave, pt);
    }
    static void append_front(dim3 g, hipStream_t s, const int32_t *which, const ChunkDesc *c,
                             const int32_t *pts, int64_t stride, const uint8_t *alive,
                             const int32_t *ns, int32_t *na, int32_t *fv, int64_t fvs,
                             const int64_t *fin, int64_t *fout, int W, int only_wave,
                             unsigned long long *bytes) {
        hipLaunchKernelGGL((pmdd_k::k_append_front<D, CH>), g, dim3(kBlock), 0, s, which, c, pts,
                           stride, alive, ns, na, fv, fvs, fin, fout, W, only_wave, D, bytes);
    }
    static void compact(dim3 g, hipStream_t s, const ChunkDesc *c, const int32_t *ns,
                        const int64_t *out, const int32_t *src, int64_t ss, const uint8_t *alive,
                        int32_t *dst, int64_t ds) {
        hipLaunchKernelGGL((pmdd_k::k_compact_chunks<D>), g, dim3(kBlock), 0, s, c, ns, out, src,
                           ss, alive, dst, ds);
    }
    static void state_dom(dim3 g, hipStream_t s, const int32_t *t, const int64_t *count,
                          const int64_t *off, const int32_t *pts, int64_t stride, uint8_t *keep,
                          unsigned long long *pt) {
        hipLaunchKernelGGL((pmdd_k::k_state_dominance<D>), dim3(g.x, kSdSplit), dim3(kBlock), 0,
                           s, t, count, off, pts, stride, keep, pt);
    }
    static void compact_nodes(dim3 g, hipStream_t s, const int64_t *oi, const int32_t *src,
                              int64_t ss, const uint8_t *keep, const int64_t *oo, int32_t *dst,
                              int64_t ds) {
        hipLaunchKernelGGL((pmdd_k::k_compact_nodes<D>), g, dim3(kBlock), 0, s, oi, src, ss, keep,
                           oo, dst, ds);
    }
    static void gather_runs(dim3 g, hipStream_t s, const int64_t *base, const int32_t *node,
                            int32_t G, pmdd_k::ExpandSrc src, int32_t *out, int64_t os) {
        hipLaunchKernelGGL((pmdd_k::k_gather_runs<D>), g, dim3(kBlock), 0, s, base, node, G, src,
                           out, os);
    }
    static void merge_pass(dim3 g, hipStream_t s, const int32_t *in, int32_t *out, int64_t stride,
                           const int64_t *run_start, const int32_t *grp_run, int32_t G, int pass,
                           int64_t total) {
        hipLaunchKernelGGL((pmdd_k::k_merge_pass<D>), g, dim3(kBlock), 0, s, in, out, stride,
                           run_start, grp_run, G, pass, total);
    }
    static void join_products(dim3 g, hipStream_t s, const pmdd_k::JoinSrc &src, int64_t n,
                              int32_t *X, int64_t xs) {
        hipLaunchKernelGGL((pmdd_k::k_join_products<D>), g, dim3(kBlock), 0, s, src, n, X, xs);
    }
    static void refill_vs(dim3 g, hipStream_t s, const int32_t *X, int64_t xs,
                          const unsigned long long *n_in, const int32_t *R, int64_t rs, int64_t lo,
                          int64_t hi, int32_t *Y, int64_t ys, unsigned long long *n_out,
                          unsigned long long *pt) {
        hipLaunchKernelGGL((pmdd_k::k_refill_vs_block<D, false>), g, dim3(pmdd_k::kWave), 0, s,
                           X, xs, n_in, R, rs, lo, hi, static_cast<const int64_t *>(nullptr), Y,
                           ys, n_out, nullptr, pt, int64_t(0));
    }
    static void refill_flags(dim3 g, hipStream_t s, const int32_t *X, int64_t xs,
                             const unsigned long long *n_in, const int32_t *R, int64_t rs,
                             int64_t lo, int64_t hi, const int64_t *hi_dev, uint8_t *flag,
                             unsigned long long *pt) {
        hipLaunchKernelGGL((pmdd_k::k_refill_vs_block<D, true>), g, dim3(pmdd_k::kWave), 0, s, X,
                           xs, n_in, R, rs, lo, hi, hi_dev, nullptr, 0, nullptr, flag, pt,
                           int64_t(0));
    }
    static void refill_flag_rows(dim3 g, hipStream_t s, const int32_t *X, int64_t xs,
                                 const unsigned long long *n_in, const int32_t *R, int64_t rs,
                                 int64_t lo, int64_t hi, const int64_t *hi_dev, int64_t blen,
                                 uint8_t *flag, unsigned long long *pt) {
        hipLaunchKernelGGL((pmdd_k::k_refill_vs_block<D, true>), g, dim3(pmdd_k::kWave), 0, s, X,
                           xs, n_in, R, rs, lo, hi, hi_dev, nullptr, 0, nullptr, flag, pt, blen);
    }
    static void flag_append(dim3 g, hipStream_t s, const int32_t *X, int64_t xs,
                            const unsigned long long *n_in, const uint8_t *flag, int32_t *Y,
                            int64_t ys, unsigned long long *n_out) {
        hipLaunchKernelGGL((pmdd_k::k_flag_append<D>), g, dim3(pmdd_k::kBlock), 0, s, X, xs, n_in,
                           flag, Y, ys, n_out);
    }
    static void stage(dim3 g, hipStream_t s, const int32_t *which, const ChunkDesc *c,
                      pmdd_k::SortedSrc src, int32_t *out, int64_t stride, uint8_t *alive,
                      int32_t *ns, int32_t *na) {
        hipLaunchKernelGGL((pmdd_k::k_stage_chunks<D>), g, dim3(kBlock), 0, s, which, c, src, out,
                           stride, alive, ns, na);
    }
    template <int C> static constexpr int threads_for() {
        return C >= 1024 ? kFilterThreadsBig : (C <= 256 ? 128 : pmdd_k::kBlock);
    }
    template <int C>
    static void queue_tier(dim3 g, hipStream_t s, const pmdd_k::QueueTier *t, int32_t *head,
                           unsigned long long *pt) {
        constexpr int NT = threads_for<C>();
        hipLaunchKernelGGL((pmdd_k::k_filter_queue<D, C, NT>), g, dim3(NT), 0, s, t, head, pt);
    }
    // Workgroups of k_filter_queue<D, C> resident on the whole device at
    // once (compute units times the kernel's occupancy per compute unit).
    template <int C> static int queue_grid_for() {
        static const int grid = [] {
            int dev = 0, cus = 0, per_cu = 0;
            hip_check(hipGetDevice(&dev), "hipGetDevice");
            hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                      "hipDeviceGetAttribute");
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &per_cu,
                          reinterpret_cast<const void *>(
                              &pmdd_k::k_filter_queue<D, C, threads_for<C>()>),
                          threads_for<C>(), 0),
                      "hipOccupancyMaxActiveBlocksPerMultiprocessor");
            return std::max(1, cus) * std::max(1, per_cu);
        }();
        return grid;
    }
    static Launch table() {
        Launch L;
        constexpr int MID = CH > kMidNode ? kMidNode : CH;
        L.filter_tier[0][0] = filter<pmdd_k::ExpandSrc, false, kSmallNode, 128>;
        L.filter_tier[0][1] = filter<pmdd_k::ExpandSrc, false, kSmallNode, 256>;
        L.filter_tier[0][2] = filter<pmdd_k::ExpandSrc, false, kSmallNode, 512>;
        L.filter_tier[1][0] = filter<pmdd_k::ExpandSrc, false, MID, 128>;
        L.filter_tier[1][1] = filter<pmdd_k::ExpandSrc, false, MID, 256>;
        L.filter_tier[1][2] = filter<pmdd_k::ExpandSrc, false, MID, 512>;
        L.filter_tier[2][0] = filter<pmdd_k::ExpandSrc, false, CH, 128>;
        L.filter_tier[2][1] = filter<pmdd_k::ExpandSrc, false, CH, 256>;
        L.filter_tier[2][2] = filter<pmdd_k::ExpandSrc, false, CH, 512>;
        L.queue[0] = queue_tier<kSmallNode>;
        L.queue[1] = queue_tier<MID>;
        L.queue[2] = queue_tier<CH>;
        L.queue_grid[0] = 0; // filled by queue_grids on first use (needs the device)
        L.queue_grid[1] = 0;
        L.queue_grid[2] = 0;
        L.queue_grids = [](int *g) {
            g[0] = queue_grid_for<kSmallNode>();
            g[1] = queue_grid_for<(CH > kMidNode ? kMidNode : CH)>();
            g[2] = queue_grid_for<CH>();
        };
        L.stage = stage;
        L.join_products = join_products;
        L.refill_vs = refill_vs;
        L.refill_flags = refill_flags;
        L.refill_flag_rows = refill_flag_rows;
        L.flag_append = flag_append;
        L.filter_dense = filter<pmdd_k::SortedSrc, false>;
        L.chunk_cap = CH;
        L.mid_cap = CH > kMidNode ? kMidNode : 0;
        L.filter_join = filter<pmdd_k::JoinSrc, false>;
        L.filter_sorted = filter<pmdd_k::SortedSrc, true>;
        L.cross = cross;
        L.append_front = append_front;
        L.compact = compact;
        L.state_dom = state_dom;
        L.compact_nodes = compact_nodes;
        L.gather_runs = gather_runs;
        L.merge_pass = merge_pass;
        return L;
    }
};

Launch launch_for(int d) {
    switch (d) {
    case 1: return LaunchImpl<1>::table();
    case 2: return LaunchImpl<2>::table();
    case 3: return LaunchImpl<3>::table();
    case 4: return LaunchImpl<4>::table();
    case 5: return LaunchImpl<5>::table();
    case 6: return LaunchImpl<6>::table();
    case 7: return LaunchImpl<7>::table();
    case 8: return LaunchImpl<8>::table();
    default: fail(PMDD_E_UNSUPPORTED, "objective count outside 1..8");
    }
}

// Index of a filter thread count in Launch::filter_tier (128, 256, 512).
inline int threads_index(int nt) { return nt <= 128 ? 0 : (nt <= 256 ? 1 : 2); }

// ---------------------------------------------------------------------------
// Tuning table: launch shapes per objective count, read once per process from
// a small JSON file (python -m pmdd_amd.tune writes it):
//   {"tuning": {"<d>": {"filter_threads": [t0, t1, t2], "cross_workgroups": n,
//                       "cross_split_max": n, "refill_waves": n}, ...}, ...}
// Other keys are ignored.  The file is PMDD_TUNE_FILE, or tune/gfx950.json
// beside the library's directory (pmdd_amd/lib/../tune/gfx950.json).
// ---------------------------------------------------------------------------
pmdd_tuning default_tuning() {
    pmdd_tuning t{};
    t.filter_threads[0] = 128;
    t.filter_threads[1] = kFilterThreadsBig;
    t.filter_threads[2] = kFilterThreadsBig;
    t.cross_workgroups = 2048;
    t.cross_split_max = 16;
    t.refill_waves = static_cast<int32_t>(kRefillWaves);
    t.source = 0;
    return t;
}

bool tuning_valid(const pmdd_tuning &t) {
    for (int i = 0; i < 3; ++i) {
        const int n = t.filter_threads[i];
        if (n != 128 && n != 256 && n != 512) {
            return false;
        }
    }
    return t.cross_workgroups >= 256 && t.cross_workgroups <= 16384 && t.cross_split_max >= 1 &&
           t.cross_split_max <= 16 && t.refill_waves >= 256 && t.refill_waves <= 32768;
}

// A JSON value, as much of it as the table needs.
struct JVal {
    enum Kind { NUL, NUM, STR, ARR, OBJ, BOOL } kind = NUL;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal *get(const std::string &k) const {
        for (const auto &kv : obj) {
            if (kv.first == k) {
                return &kv.second;
            }
        }
        return nullptr;
    }
};

struct JParser {
    const std::string &s;
    size_t i = 0;
    bool ok = true;
    explicit JParser(const std::string &x) : s(x) {}
    void ws() {
        while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) {
            ++i;
        }
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    std::string string_lit() {
        std::string out;
        if (!eat('"')) {
            ok = false;
            return out;
        }
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) {
                ++i;
            }
            out.push_back(s[i++]);
        }
        if (i >= s.size()) {
            ok = false;
        } else {
            ++i;
        }
        return out;
    }
    JVal value(int depth = 0) {
        JVal v;
        ws();
        if (!ok || i >= s.size() || depth > 32) {
            ok = false;
            return v;
        }
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = JVal::OBJ;
            if (eat('}')) {
                return v;
            }
            do {
                std::string k = string_lit();
                if (!eat(':')) {
                    ok = false;
                    return v;
                }
                v.obj.emplace_back(k, value(depth + 1));
            } while (ok && eat(','));
            if (!eat('}')) {
                ok = false;
            }
        } else if (c == '[') {
            ++i;
            v.kind = JVal::ARR;
            if (eat(']')) {
                return v;
            }
            do {
                v.arr.push_back(value(depth + 1));
            } while (ok && eat(','));
            if (!eat(']')) {
                ok = false;
            }
        } else if (c == '"') {
            v.kind = JVal::STR;
            v.str = string_lit();
        } else if (s.compare(i, 4, "true") == 0 || s.compare(i, 5, "false") == 0) {
            v.kind = JVal::BOOL;
            i += s[i] == 't' ? 4 : 5;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            char *end = nullptr;
            v.num = std::strtod(s.c_str() + i, &end);
            if (end == s.c_str() + i) {
                ok = false;
            } else {
                v.kind = JVal::NUM;
                i = static_cast<size_t>(end - s.c_str());
            }
        }
        return v;
    }
};

// Shapes per objective count from the table file (source 1), or the
// defaults (source 0) where the file, or its entry for d, is missing or bad.
struct TuneTable {
    pmdd_tuning t[9];
    std::string path;
};

std::string tune_file_path() {
    const char *env = std::getenv("PMDD_TUNE_FILE");
    if (env != nullptr && *env != '\0') {
        return env;
    }
    Dl_info info{};
    if (dladdr(reinterpret_cast<void *>(&default_tuning), &info) != 0 &&
        info.dli_fname != nullptr) {
        std::string lib(info.dli_fname);
        const size_t slash = lib.rfind('/');
        const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
        return dir + "/../tune/gfx950.json";
    }
    return "";
}

const TuneTable &tune_table() {
    static const TuneTable table = [] {
        TuneTable tt;
        for (pmdd_tuning &x : tt.t) {
            x = default_tuning();
        }
        tt.path = tune_file_path();
        std::FILE *f = tt.path.empty() ? nullptr : std::fopen(tt.path.c_str(), "rb");
        if (f == nullptr) {
            return tt;
        }
        std::string text;
        char buf[4096];
        size_t n = 0;
        while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) {
            text.append(buf, n);
        }
        std::fclose(f);
        JParser p(text);
        const JVal root = p.value();
        const JVal *tun = p.ok ? root.get("tuning") : nullptr;
        if (tun == nullptr || tun->kind != JVal::OBJ) {
            return tt;
        }
        for (int d = 1; d <= 8; ++d) {
            const JVal *e = tun->get(std::to_string(d));
            if (e == nullptr || e->kind != JVal::OBJ) {
                continue;
            }
            pmdd_tuning x = default_tuning();
            const JVal *ft = e->get("filter_threads");
            if (ft != nullptr && ft->kind == JVal::ARR && ft->arr.size() == 3) {
                for (int k = 0; k < 3; ++k) {
                    x.filter_threads[k] = static_cast<int32_t>(ft->arr[k].num);
                }
            }
            const JVal *v = nullptr;
            if ((v = e->get("cross_workgroups")) != nullptr && v->kind == JVal::NUM) {
                x.cross_workgroups = static_cast<int32_t>(v->num);
            }
            if ((v = e->get("cross_split_max")) != nullptr && v->kind == JVal::NUM) {
                x.cross_split_max = static_cast<int32_t>(v->num);
            }
            if ((v = e->get("refill_waves")) != nullptr && v->kind == JVal::NUM) {
                x.refill_waves = static_cast<int32_t>(v->num);
            }
            x.source = 1;
            if (tuning_valid(x)) {
                tt.t[d] = x;
            }
        }
        return tt;
    }();
    return table;
}

inline int ceil_log2(int64_t x) {
    int p = 0;
    while ((int64_t(1) << p) < x) {
        ++p;
    }
    return p;
}

// A pinned region of n bytes of the context's staging arena; when the arena
// is full, waits for the stream (every region is then free) and grows it.
// In a no-allocation call the arena is not grown: nullptr when it has no
// room, and the caller copies from or to its own (pageable) memory instead.
inline void *staging_take(pmdd_ctx *ctx, size_t n) {
    void *p = ctx->staging.take(n);
    if (p == nullptr) {
        hip_check(hipStreamSynchronize(ctx->stream), "staging sync");
        if (n + 256 > ctx->staging.cap) {
            if (t_arena != nullptr) {
                return nullptr;
            }
            ctx->staging.grow(n + 256);
        } else {
            ctx->staging.reset();
        }
        p = ctx->staging.take(n);
        if (p == nullptr) {
            fail(PMDD_E_NOMEM, "hipHostMalloc(staging) failed");
        }
    }
    return p;
}

template <class T>
void h2d(pmdd_ctx *ctx, DevBuf &b, const std::vector<T> &v, const char *what) {
    b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T), what);
    if (!v.empty()) {
        const size_t bytes = v.size() * sizeof(T);
        void *p = staging_take(ctx, bytes);
        if (p == nullptr) {
            // straight from the vector, which must outlive the copy
            hip_check(hipMemcpyAsync(b.p, v.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
                      what);
            hip_check(hipStreamSynchronize(ctx->stream), what);
            return;
        }
        std::memcpy(p, v.data(), bytes);
        hip_check(hipMemcpyAsync(b.p, p, bytes, hipMemcpyHostToDevice, ctx->stream), what);
    }
}

// Copies n elements back and waits for the stream, which frees the staging.
template <class T>
void d2h(pmdd_ctx *ctx, std::vector<T> &v, const DevBuf &b, size_t n, const char *what) {
    v.resize(n);
    void *p = nullptr;
    if (n > 0) {
        p = staging_take(ctx, n * sizeof(T));
        hip_check(hipMemcpyAsync(p != nullptr ? p : static_cast<void *>(v.data()), b.p,
                                 n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream),
                  what);
    }
    hip_check(hipStreamSynchronize(ctx->stream), what);
    if (n > 0 && p != nullptr) {
        std::memcpy(v.data(), p, n * sizeof(T));
    }
    ctx->staging.reset();
}

void upload_offsets(pmdd_ctx *ctx, DevFrontier &f) {
    h2d(ctx, f.off, f.h_off, "hipMemcpy(frontier offsets)");
    f.on_host = true;
    f.step = -1;
}

// Frontier holding {0^d} at one node of a layer, empty elsewhere.
void seed_frontier(pmdd_ctx *ctx, DevFrontier &f, int d, int64_t n_nodes, int64_t node) {
    hipStream_t s = ctx->stream;
    f.n_nodes = n_nodes;
    f.h_off.assign(n_nodes + 1, 0);
    for (int64_t v = node + 1; v <= n_nodes; ++v) {
        f.h_off[v] = 1;
    }
    grow_points(f, d, 1, 0, s); // keeps a leased buffer
    for (int o = 0; o < d; ++o) {
        hip_check(hipMemsetAsync(f.pts.as<int32_t>() + o * f.cap, 0, 4, s), "hipMemset(seed)");
    }
    upload_offsets(ctx, f);
}

// Launches forked onto the context's side stream: at construction the side
// stream waits for everything queued so far on the main stream; join() makes
// the main stream wait for everything queued since on the side stream.  The
// destructor joins too, so that when an error unwinds a layer, a later wait
// on the main stream also covers the side stream's launches.
struct SideFork {
    pmdd_ctx *ctx;
    hipStream_t main;
    bool open = false;
    SideFork(pmdd_ctx *c, hipStream_t m) : ctx(c), main(m) {
        hip_check(hipEventRecord(c->ev_fork, m), "hipEventRecord(fork)");
        hip_check(hipStreamWaitEvent(c->side, c->ev_fork, 0), "hipStreamWaitEvent(fork)");
        open = true;
    }
    SideFork(const SideFork &) = delete;
    SideFork &operator=(const SideFork &) = delete;
    void join() {
        if (open) {
            open = false;
            hip_check(hipEventRecord(ctx->ev_join, ctx->side), "hipEventRecord(join)");
            hip_check(hipStreamWaitEvent(main, ctx->ev_join, 0), "hipStreamWaitEvent(join)");
        }
    }
    ~SideFork() {
        if (open) {
            (void)hipEventRecord(ctx->ev_join, ctx->side);
            (void)hipStreamWaitEvent(main, ctx->ev_join, 0);
        }
    }
};

// ---------------------------------------------------------------------------
// One enumeration
// ---------------------------------------------------------------------------
struct Run {
    pmdd_ctx *ctx;
    pmdd_resident *dd;
    pmdd_opts opts;
    pmdd_stats *st;
    Launch K;
    int d;
    hipStream_t s;
    int64_t mem_base = 0;
    // executed pair-test counters (count_pair_tests): chunk filter, frontier
    // check, cross filter, state dominance; null when not counting
    unsigned long long *d_tests = nullptr, *t_check = nullptr, *t_cross = nullptr,
                       *t_sd = nullptr;
    FrontierLease sd_lease;
    DevFrontier &sd_spare; // state dominance output, swapped with the layer's frontier
    // per-phase device times (opts.phase_timing): HIP events bracket the
    // phases of every layer; off, no event is recorded
    const bool timing;
    // PMDD_DEBUG_JOIN=1 prints per-batch coupling counts to stderr
    const bool debug_join = std::getenv("PMDD_DEBUG_JOIN") != nullptr;

    Run(pmdd_ctx *c, pmdd_resident *r, const pmdd_opts &o, pmdd_stats *stats)
        : Run(c, r, r->d, o, stats) {}
    // r may be null for a run that only expands layers it is handed (the
    // serving packer's packed layers)
    Run(pmdd_ctx *c, pmdd_resident *r, int d_, const pmdd_opts &o, pmdd_stats *stats)
        : ctx(c), dd(r), opts(o), st(stats), K(launch_for(d_)), d(d_), s(c->stream),
          sd_lease(c, d_), sd_spare(sd_lease.f), timing(o.phase_timing != 0) {
        if (opts.max_cand <= 0) {
            opts.max_cand = PMDD_DEFAULT_MAX_CAND;
        }
        if (opts.max_prod < 0) {
            opts.max_prod = PMDD_MAX_PROD_ADAPTIVE;
        }
        size_t fr = 0, tot = 0;
        hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
        mem_base = static_cast<int64_t>(tot) - static_cast<int64_t>(fr);
        st->gpu_mem_peak_reserved_bytes = mem_base;
        if (opts.count_pair_tests) {
            ctx->counter.ensure(4 * 8, "hipMalloc(counter)");
            d_tests = ctx->counter.as<unsigned long long>();
            t_check = d_tests + 1;
            t_cross = d_tests + 2;
            t_sd = d_tests + 3;
            hipLaunchKernelGGL(pmdd_k::k_fill_u8, dim3(1), dim3(64), 0, s,
                               ctx->counter.as<uint8_t>(), int64_t(4 * 8), uint8_t(0));
            hip_check(hipGetLastError(), "k_fill_u8");
        }
    }

    // Peak device memory in use, sampled after allocations (it changes only
    // with them).
    uint64_t allocs_seen = ~uint64_t(0);
    void sample_memory() {
        const uint64_t a = g_device_allocs.load(std::memory_order_relaxed);
        if (a == allocs_seen) {
            return;
        }
        allocs_seen = a;
        size_t fr = 0, tot = 0;
        hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
        const int64_t used = static_cast<int64_t>(tot) - static_cast<int64_t>(fr);
        st->gpu_mem_peak_reserved_bytes = std::max(st->gpu_mem_peak_reserved_bytes, used);
        st->gpu_mem_peak_used_bytes =
            std::max(st->gpu_mem_peak_used_bytes, std::max<int64_t>(0, used - mem_base));
    }

    // Event i of the context's pool, created on first use.
    hipEvent_t pool_event(size_t i) {
        while (ctx->ev_pool.size() <= i) {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreate(&e), "hipEventCreate");
            ctx->ev_pool.push_back(e);
        }
        return ctx->ev_pool[i];
    }
    // Seconds between the pool's pairs (0,1), (2,3), ... for `pairs` pairs;
    // the stream must have passed them.
    double pool_seconds(int pairs) {
        double t = 0.0;
        if (!timing) {
            return t;
        }
        for (int i = 0; i < pairs; ++i) {
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, ctx->ev_pool[2 * i], ctx->ev_pool[2 * i + 1]),
                      "hipEventElapsedTime");
            t += ms * 1e-3;
        }
        return t;
    }

    // Host tables of a layer, kept between layers so that planning a layer
    // neither allocates nor touches fresh pages.
    struct HostTables {
        std::vector<int64_t> run_off, gamma, chunk_out, big_base, run_start, active;
        std::vector<ChunkDesc> chunks;
        std::vector<int32_t> which_small, which_tiny, which_mid, which_big, which_cross, nalive;
        std::vector<int32_t> big_node, grp_run, targets;
    } xh;

    // Timings read after a later synchronisation, so the host need not wait
    // for the work they bracket: mark() records event i of ctx->ev_late on
    // the stream, span(a, b, acc) adds the time between marks a and b to
    // *acc when flush_late runs.  Spans of one layer share its marks.
    size_t n_marks = 0;
    struct Span {
        size_t a, b;
        double *acc;
    };
    std::vector<Span> spans;
    void record(hipEvent_t e) {
        if (timing) {
            hip_check(hipEventRecord(e, s), "hipEventRecord");
        }
    }
    size_t mark() {
        if (!timing) {
            return 0;
        }
        const size_t i = n_marks++;
        while (ctx->ev_late.size() <= i) {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreate(&e), "hipEventCreate");
            ctx->ev_late.push_back(e);
        }
        record(ctx->ev_late[i]);
        return i;
    }
    void span(size_t a, size_t b, double *acc) {
        if (timing) {
            spans.push_back({a, b, acc});
        }
    }
    // Only once the stream has passed every mark.
    void flush_late() {
        for (const Span &x : spans) {
            add_time(ctx->ev_late[x.a], ctx->ev_late[x.b], x.acc);
        }
        spans.clear();
        n_marks = 0;
    }

    float elapsed_ms() {
        float ms = 0.f;
        if (!timing) {
            return ms;
        }
        hip_check(hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b), "hipEventElapsedTime");
        return ms;
    }

    // ---- chunk tables ----------------------------------------------------
    void upload_table(ChunkBufs &cb, const std::vector<ChunkDesc> &chunks) {
        const size_t n = std::max<size_t>(chunks.size(), 1);
        h2d(ctx, cb.chunks, chunks, "hipMemcpy(chunk table)");
        cb.nsurv.ensure(n * 4, "hipMalloc(chunk counts)");
        cb.nalive.ensure(n * 4, "hipMalloc(chunk counts)");
    }

    // A chunk list for run_filter, copied to the device ahead of the launch.
    void upload_list(DevBuf &wbuf, const std::vector<int32_t> &which) {
        if (!which.empty()) {
            h2d(ctx, wbuf, which, "hipMemcpy(chunk list)");
        }
    }

    // Launch a filter over the listed chunks (the list already uploaded to
    // wbuf), or over chunks 0 .. count-1 when `which` is empty; returns the
    // number of launches (0 or 1).
    int run_filter(FilterFn f, ChunkBufs &cb, const DevBuf &wbuf, const std::vector<int32_t> &which,
                   int64_t count, const void *src, int32_t *out, int64_t stride, uint8_t *alive) {
        const int32_t *w = nullptr;
        int64_t n = count;
        if (!which.empty()) {
            w = wbuf.as<int32_t>();
            n = static_cast<int64_t>(which.size());
        }
        if (n <= 0) {
            return 0;
        }
        f(dim3(static_cast<unsigned>(n)), s, w, cb.chunks.as<ChunkDesc>(), src, out, stride, alive,
          cb.nsurv.as<int32_t>(), cb.nalive.as<int32_t>(), d_tests);
        hip_check(hipGetLastError(), "k_filter_chunks");
        return 1;
    }

    // Packed layers (the serving packer): a pass's small-node chunks are
    // filtered by the persistent queue kernels, launched as one captured
    // graph that every pass replays.
    bool queued = false;

    // The three tier lists of a pass (already on the device) through
    // k_filter_queue: the tiers' descriptors go to ctx->qtier, then the
    // pass graph (queue heads zeroed by a fill kernel, then one persistent
    // launch per tier) is replayed, captured first when the buffers it was
    // captured for have moved.  Returns the number of tier launches.
    int run_queue_pass(ChunkBufs &cb, const pmdd_k::ExpandSrc &src,
                       const std::vector<int32_t> &tiny, const std::vector<int32_t> &mid,
                       const std::vector<int32_t> &full, int32_t *out, int64_t stride,
                       uint8_t *alive) {
        if (K.queue_grid[0] == 0) {
            K.queue_grids(K.queue_grid);
        }
        const std::vector<int32_t> *lists[3] = {&tiny, &mid, &full};
        const DevBuf *bufs[3] = {&ctx->which_s, &ctx->which_m, &ctx->which_a};
        std::vector<pmdd_k::QueueTier> tiers(3);
        for (int t = 0; t < 3; ++t) {
            pmdd_k::QueueTier &q = tiers[t];
            q.src = src;
            q.chunks = cb.chunks.as<ChunkDesc>();
            q.which = lists[t]->empty() ? nullptr : bufs[t]->as<int32_t>();
            q.count = static_cast<int32_t>(lists[t]->size());
            q.pad = 0;
            q.out_pts = out;
            q.out_stride = stride;
            q.alive = alive;
            q.nsurv = cb.nsurv.as<int32_t>();
            q.nalive = cb.nalive.as<int32_t>();
        }
        h2d(ctx, ctx->qtier, tiers, "hipMemcpy(queue tiers)");
        ctx->qhead.ensure(3 * sizeof(int32_t), "hipMalloc(queue heads)");
        const void *key[4] = {reinterpret_cast<const void *>(static_cast<intptr_t>(d)),
                              ctx->qtier.p, ctx->qhead.p, d_tests};
        if (ctx->pass_graph == nullptr || !std::equal(key, key + 4, ctx->graph_key)) {
            ctx->drop_graph();
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal),
                      "hipStreamBeginCapture");
            hipLaunchKernelGGL(pmdd_k::k_fill_u8, dim3(1), dim3(64), 0, s,
                               ctx->qhead.as<uint8_t>(), int64_t(3 * sizeof(int32_t)), uint8_t(0));
            const pmdd_k::QueueTier *qt = ctx->qtier.as<pmdd_k::QueueTier>();
            for (int t = 0; t < 3; ++t) {
                K.queue[t](dim3(static_cast<unsigned>(K.queue_grid[t])), s, qt + t,
                           ctx->qhead.as<int32_t>() + t, d_tests);
            }
            const hipError_t cap = hipGetLastError();
            hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
            hip_check(cap, "k_filter_queue (capture)");
            const hipError_t inst = hipGraphInstantiate(&ctx->pass_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(inst, "hipGraphInstantiate");
            std::copy(key, key + 4, ctx->graph_key);
        }
        hip_check(hipGraphLaunch(ctx->pass_graph, s), "hipGraphLaunch");
        ctx->graph_replays += 1;
        st->graph_launches += 1;
        return 3;
    }

    // Copy the listed chunks' candidates from a key-ordered source, all alive.
    void run_stage(ChunkBufs &cb, DevBuf &wbuf, const std::vector<int32_t> &which,
                   const pmdd_k::SortedSrc &src, int32_t *out, int64_t stride, uint8_t *alive) {
        if (which.empty()) {
            return;
        }
        h2d(ctx, wbuf, which, "hipMemcpy(chunk list)");
        K.stage(dim3(static_cast<unsigned>(which.size())), s, wbuf.as<int32_t>(),
                cb.chunks.as<ChunkDesc>(), src, out, stride, alive, cb.nsurv.as<int32_t>(),
                cb.nalive.as<int32_t>());
        hip_check(hipGetLastError(), "k_stage_chunks");
    }

    // One-sided block filter over every group with more than one chunk (the
    // chunks of a group are descending-key blocks), in waves of kCrossWave
    // chunks per group: a wave's chunks are checked against the group's front
    // (the survivors of its earlier waves, dense in key order, kept in fv at
    // the group's first candidate slot) and against the earlier chunks of the
    // same wave; the wave's survivors are then appended to the front.  `fv`
    // needs room for every candidate of the groups (stride fv_stride).
    // Returns the number of k_cross_chunks launches; launch i is bracketed by
    // pool events 2i and 2i+1.
    static constexpr int kCrossWave = 64;
    int run_cross_waves(ChunkBufs &cb, const std::vector<ChunkDesc> &chunks, const int32_t *pts,
                        int64_t stride, uint8_t *alive, int32_t *fv, int64_t fv_stride) {
        int32_t max_count = 1;
        for (const ChunkDesc &c : chunks) {
            max_count = std::max(max_count, c.group_count);
        }
        if (max_count <= 1) {
            return 0;
        }
        const int waves = static_cast<int>(ceil_div(max_count, kCrossWave));
        // per wave: every chunk of the groups with more than one chunk, in
        // chunk order (a counting pass, so the cost is chunks + waves)
        std::vector<size_t> start(waves + 1, 0);
        for (size_t i = 0; i < chunks.size(); ++i) {
            const ChunkDesc &c = chunks[i];
            if (c.group_count > 1) {
                start[(static_cast<int>(i) - c.group_first) / kCrossWave + 1] += 1;
            }
        }
        for (int w = 0; w < waves; ++w) {
            start[w + 1] += start[w];
        }
        std::vector<int32_t> lists(start[waves]);
        {
            std::vector<size_t> fill(start.begin(), start.end() - 1);
            for (size_t i = 0; i < chunks.size(); ++i) {
                const ChunkDesc &c = chunks[i];
                if (c.group_count > 1) {
                    lists[fill[(static_cast<int>(i) - c.group_first) / kCrossWave]++] =
                        static_cast<int32_t>(i);
                }
            }
        }
        h2d(ctx, ctx->which_c, lists, "hipMemcpy(wave lists)");
        // front counts, double-buffered by wave, indexed by a group's first chunk
        const size_t nc = chunks.size();
        ctx->front_count.ensure(2 * nc * sizeof(int64_t), "hipMalloc(front counts)");
        int64_t *fc = ctx->front_count.as<int64_t>();
        hipLaunchKernelGGL(pmdd_k::k_fill_u8,
                           dim3(static_cast<unsigned>(ceil_div(2 * nc * sizeof(int64_t), 256))),
                           dim3(256), 0, s, ctx->front_count.as<uint8_t>(),
                           static_cast<int64_t>(2 * nc * sizeof(int64_t)), uint8_t(0));
        hip_check(hipGetLastError(), "k_fill_u8");
        const int32_t *wl = ctx->which_c.as<int32_t>();
        int launches = 0;
        for (int w = 0; w < waves; ++w) {
            int64_t *fin = fc + (w % 2) * nc;
            int64_t *fout = fc + ((w + 1) % 2) * nc;
            const size_t n = start[w + 1] - start[w];
            if (n == 0) {
                continue;
            }
            // spread each chunk's scan over G workgroups, about
            // cross_workgroups in all (2048 by default; 1536 fit the chip at
            // this kernel's LDS use), at most cross_split_max per chunk
            const pmdd_tuning &T = ctx->tuning[d];
            const int G = static_cast<int>(std::min<int64_t>(
                T.cross_split_max, std::max<int64_t>(1, ceil_div(T.cross_workgroups, n))));
            record(pool_event(2 * launches));
            K.cross(dim3(static_cast<unsigned>(n * G)), s, wl + start[w], cb.chunks.as<ChunkDesc>(),
                    pts, stride, alive, cb.nsurv.as<int32_t>(), fv, fv_stride, fin, kCrossWave, G,
                    -1, t_cross);
            hip_check(hipGetLastError(), "k_cross_chunks");
            record(pool_event(2 * launches + 1));
            K.append_front(dim3(static_cast<unsigned>(n)), s, wl + start[w],
                           cb.chunks.as<ChunkDesc>(), pts, stride, alive, cb.nsurv.as<int32_t>(),
                           cb.nalive.as<int32_t>(), fv, fv_stride, fin, fout, kCrossWave, -1,
                           nullptr);
            hip_check(hipGetLastError(), "k_append_front");
            ++launches;
        }
        return launches;
    }

    // Compulsory bytes of the cross waves over `chunks` (groups of consecutive
    // chunks), once the final alive counts are known: per wave, the group's
    // front and the wave's points read once, the survivors appended once, one
    // flag per point.
    int64_t cross_bytes(const std::vector<ChunkDesc> &chunks,
                        const std::vector<int32_t> &nalive) const {
        int64_t b = 0;
        size_t i = 0;
        while (i < chunks.size()) {
            const ChunkDesc &c = chunks[i];
            if (c.group_count <= 1) {
                ++i;
                continue;
            }
            int64_t front = 0;
            for (int k = 0; k < c.group_count; k += kCrossWave) {
                int64_t pts = 0, kept = 0;
                for (int j = k; j < std::min(k + kCrossWave, c.group_count); ++j) {
                    pts += chunks[i + j].len;
                    kept += nalive[i + j];
                }
                b += 4 * static_cast<int64_t>(d) * (front + pts + kept) + pts;
                front += kept;
            }
            i += static_cast<size_t>(c.group_count);
        }
        return b;
    }

    void run_compact(ChunkBufs &cb, int64_t nch, const std::vector<int64_t> &outpos,
                     const int32_t *src, int64_t ss, const uint8_t *alive, int32_t *dst,
                     int64_t ds) {
        if (nch <= 0) {
            return;
        }
        h2d(ctx, cb.out, outpos, "hipMemcpy(chunk outputs)");
        K.compact(dim3(static_cast<unsigned>(nch)), s, cb.chunks.as<ChunkDesc>(),
                  cb.nsurv.as<int32_t>(), cb.out.as<int64_t>(), src, ss, alive, dst, ds);
        hip_check(hipGetLastError(), "k_compact_chunks");
    }

    // Merge the runs of each group (each run in descending key order) into one
    // run per group; returns the buffer (a or b) holding the result.
    int32_t *merge_runs(DevBuf &a, DevBuf &b, int64_t total, const std::vector<int64_t> &run_start,
                        const std::vector<int32_t> &grp_run, int64_t max_runs) {
        const int passes = ceil_log2(max_runs);
        if (passes == 0 || total == 0) {
            return a.as<int32_t>();
        }
        h2d(ctx, ctx->run_start, run_start, "hipMemcpy(run table)");
        h2d(ctx, ctx->grp_run, grp_run, "hipMemcpy(group table)");
        const int32_t G = static_cast<int32_t>(grp_run.size()) - 1;
        int32_t *cur = a.as<int32_t>();
        int32_t *oth = b.as<int32_t>();
        const dim3 grid(static_cast<unsigned>(ceil_div(total, kBlock)));
        for (int p = 0; p < passes; ++p) {
            K.merge_pass(grid, s, cur, oth, total, ctx->run_start.as<int64_t>(),
                         ctx->grp_run.as<int32_t>(), G, p, total);
            hip_check(hipGetLastError(), "k_merge_pass");
            std::swap(cur, oth);
        }
        return cur;
    }

    // One layer expansion (expand_layer_frontiers, bottomup.cu:242-533):
    // frontier of every node of the target layer from its arcs' parents.
    // Passes cover runs of whole nodes with at most max_cand candidates (a
    // single node may exceed it, as in the reference).  A node with at most CH
    // candidates is filtered by one workgroup; a larger node first has its
    // parent runs (each in key order) merged, then is cut into blocks of
    // descending keys, each block filtered locally and checked against the
    // survivors of the earlier blocks.  Every node's frontier leaves in
    // descending key order.
    void expand(const DevFrontier &prev, const LayerCSR &csr, int64_t n_nodes, DevFrontier &next,
                int64_t *cand_out, std::vector<int64_t> *gamma_out) {
        const int CH = K.chunk_cap;
        std::vector<int64_t> &run_off = xh.run_off;
        run_off.resize(std::max<int64_t>(csr.n_arcs, 1));
        std::vector<int64_t> &gamma = xh.gamma;
        gamma.assign(n_nodes, 0);
        int64_t C = 0;
        for (int64_t v = 0; v < n_nodes; ++v) {
            int64_t acc = 0;
            for (int64_t e = csr.off[v]; e < csr.off[v + 1]; ++e) {
                run_off[e] = acc;
                acc += prev.size(csr.src[e]);
            }
            gamma[v] = acc;
            C += acc;
        }
        *cand_out = C;
        h2d(ctx, ctx->run_off, run_off, "hipMemcpy(run offsets)");

        next.n_nodes = n_nodes;
        next.h_off.assign(n_nodes + 1, 0);
        int64_t next_total = 0;
        if (next.cap < 1) {
            grow_points(next, d, std::max<int64_t>(1024, std::min<int64_t>(C, 2 * prev.total())),
                        0, s);
        }

        pmdd_k::ExpandSrc src;
        src.prev_off = prev.off.as<int64_t>();
        src.prev_pts = prev.pts.as<int32_t>();
        src.prev_stride = prev.cap;
        src.arc_off = csr.d_off;
        src.arc_src = csr.d_src;
        src.arc_w = csr.d_w;
        src.w_stride = std::max<int64_t>(csr.n_arcs, 1);
        src.run_off = ctx->run_off.as<int64_t>();

        ChunkBufs &cb = ctx->cb_main;
        int64_t v = 0;
        std::vector<ChunkDesc> &chunks = xh.chunks;
        std::vector<int32_t> &which_small = xh.which_small, &which_tiny = xh.which_tiny,
                             &which_mid = xh.which_mid, &which_big = xh.which_big,
                             &which_cross = xh.which_cross, &nalive = xh.nalive;
        std::vector<int64_t> &chunk_out = xh.chunk_out, &big_base = xh.big_base,
                             &run_start = xh.run_start;
        std::vector<int32_t> &big_node = xh.big_node, &grp_run = xh.grp_run;
        while (v < n_nodes) {
            const int64_t v0 = v;
            int64_t pass_c = 0;
            while (v < n_nodes && (pass_c == 0 || pass_c + gamma[v] <= opts.max_cand)) {
                pass_c += gamma[v];
                ++v;
            }
            const int64_t v1 = v;
            chunks.clear();
            which_small.clear();
            which_tiny.clear();
            which_mid.clear();
            which_big.clear();
            which_cross.clear();
            big_base.clear();
            big_node.clear();
            run_start.clear();
            grp_run.clear();
            int64_t base = 0, bigpos = 0, max_runs = 1;
            for (int64_t u = v0; u < v1; ++u) {
                const int64_t g = gamma[u];
                if (g == 0) {
                    continue;
                }
                const int32_t first = static_cast<int32_t>(chunks.size());
                if (g <= CH) {
                    ChunkDesc c;
                    c.start = 0;
                    c.len = static_cast<int32_t>(g);
                    c.out_base = base;
                    c.group = static_cast<int32_t>(u);
                    c.group_first = first;
                    c.group_count = 1;
                    (g <= kSmallNode                      ? which_tiny
                     : (K.mid_cap > 0 && g <= K.mid_cap) ? which_mid
                                                         : which_small)
                        .push_back(first);
                    chunks.push_back(c);
                } else {
                    big_node.push_back(static_cast<int32_t>(u));
                    big_base.push_back(bigpos);
                    grp_run.push_back(static_cast<int32_t>(run_start.size()));
                    for (int64_t e = csr.off[u]; e < csr.off[u + 1]; ++e) {
                        run_start.push_back(bigpos + run_off[e]);
                    }
                    max_runs = std::max<int64_t>(max_runs, csr.off[u + 1] - csr.off[u]);
                    const int32_t cnt = static_cast<int32_t>(ceil_div(g, CH));
                    for (int32_t k = 0; k < cnt; ++k) {
                        ChunkDesc c;
                        c.start = bigpos + static_cast<int64_t>(k) * CH;
                        c.len = static_cast<int32_t>(std::min<int64_t>(CH, g - int64_t(k) * CH));
                        c.out_base = base + static_cast<int64_t>(k) * CH;
                        c.group = static_cast<int32_t>(u);
                        c.group_first = first;
                        c.group_count = cnt;
                        which_big.push_back(static_cast<int32_t>(chunks.size()));
                        if (k > 0) {
                            which_cross.push_back(static_cast<int32_t>(chunks.size()));
                        }
                        chunks.push_back(c);
                    }
                    bigpos += g;
                }
                base += g;
            }
            if (chunks.empty()) {
                for (int64_t u = v0; u < v1; ++u) {
                    next.h_off[u + 1] = next_total;
                }
                continue;
            }
            record(ctx->ev_a);
            int32_t *sorted = nullptr;
            if (!big_node.empty()) {
                grp_run.push_back(static_cast<int32_t>(run_start.size()));
                run_start.push_back(bigpos);
                big_base.push_back(bigpos);
                ctx->sort_a.ensure(static_cast<size_t>(bigpos) * d * 4, "hipMalloc(sort buffer)");
                ctx->sort_b.ensure(static_cast<size_t>(bigpos) * d * 4, "hipMalloc(sort buffer)");
                h2d(ctx, ctx->big_base, big_base, "hipMemcpy(big nodes)");
                h2d(ctx, ctx->big_node, big_node, "hipMemcpy(big nodes)");
                K.gather_runs(dim3(static_cast<unsigned>(ceil_div(bigpos, kBlock))), s,
                              ctx->big_base.as<int64_t>(), ctx->big_node.as<int32_t>(),
                              static_cast<int32_t>(big_node.size()), src,
                              ctx->sort_a.as<int32_t>(), bigpos);
                hip_check(hipGetLastError(), "k_gather_runs");
                sorted = merge_runs(ctx->sort_a, ctx->sort_b, bigpos, run_start, grp_run, max_runs);
            }
            upload_table(cb, chunks);
            ctx->surv_pts.ensure(static_cast<size_t>(pass_c) * d * 4, "hipMalloc(survivor buffer)");
            ctx->alive.ensure(static_cast<size_t>(pass_c), "hipMalloc(alive flags)");
            upload_list(ctx->which_s, which_tiny);
            upload_list(ctx->which_m, which_mid);
            upload_list(ctx->which_a, which_small);
            if (!which_big.empty()) {
                // big nodes: candidates staged as they are; the cross filter
                // does all of their checks
                pmdd_k::SortedSrc ss{sorted, bigpos};
                run_stage(cb, ctx->which_b, which_big, ss, ctx->surv_pts.as<int32_t>(), pass_c,
                          ctx->alive.as<uint8_t>());
            }
            // only k_filter_chunks launches between ev_d and ev_c
            record(ctx->ev_d);
            int filters = 0;
            if (queued) {
                filters = run_queue_pass(cb, src, which_tiny, which_mid, which_small,
                                         ctx->surv_pts.as<int32_t>(), pass_c,
                                         ctx->alive.as<uint8_t>());
            } else {
                const pmdd_tuning &T = ctx->tuning[d];
                filters = run_filter(K.filter_tier[0][threads_index(T.filter_threads[0])], cb,
                                     ctx->which_s, which_tiny, 0, &src,
                                     ctx->surv_pts.as<int32_t>(), pass_c, ctx->alive.as<uint8_t>());
                filters += run_filter(K.filter_tier[1][threads_index(T.filter_threads[1])], cb,
                                      ctx->which_m, which_mid, 0, &src,
                                      ctx->surv_pts.as<int32_t>(), pass_c,
                                      ctx->alive.as<uint8_t>());
                filters += run_filter(K.filter_tier[2][threads_index(T.filter_threads[2])], cb,
                                      ctx->which_a, which_small, 0, &src,
                                      ctx->surv_pts.as<int32_t>(), pass_c,
                                      ctx->alive.as<uint8_t>());
            }
            record(ctx->ev_c);
            int cross_waves = 0;
            if (!which_cross.empty()) {
                cross_waves = run_cross_waves(cb, chunks, ctx->surv_pts.as<int32_t>(), pass_c,
                                              ctx->alive.as<uint8_t>(), sorted, bigpos);
            }
            record(ctx->ev_b);
            d2h(ctx, nalive, cb.nalive, chunks.size(), "chunk survivor counts");
            flush_late();
            if (cross_waves > 0) {
                st->kernel_cross_s += pool_seconds(cross_waves);
                st->cross_launches += cross_waves;
                st->cross_bytes += cross_bytes(chunks, nalive);
            }
            add_time(ctx->ev_a, ctx->ev_d, &st->kernel_expand_td_s);
            add_time(ctx->ev_d, ctx->ev_c, &st->kernel_filter_s);
            add_time(ctx->ev_d, ctx->ev_b, &st->kernel_dominance_s);
            st->filter_launches += filters;
            st->filter_candidates += pass_c;

            // dense placement of the survivors, node by node
            chunk_out.resize(chunks.size());
            int64_t at = next_total;
            size_t ci = 0;
            for (int64_t u = v0; u < v1; ++u) {
                while (ci < chunks.size() && chunks[ci].group == u) {
                    chunk_out[ci] = at;
                    at += nalive[ci];
                    ++ci;
                }
                next.h_off[u + 1] = at;
            }
            grow_points(next, d, at, next_total, s);
            // no wait: the next pass or layer is planned while this runs
            const size_t tc = mark();
            run_compact(cb, static_cast<int64_t>(chunks.size()), chunk_out,
                        ctx->surv_pts.as<int32_t>(), pass_c, ctx->alive.as<uint8_t>(),
                        next.pts.as<int32_t>(), next.cap);
            span(tc, mark(), &st->kernel_expand_td_s);
            next_total = at;
            sample_memory();
        }
        upload_offsets(ctx, next);
        st->algorithmic_bytes +=
            4 * (static_cast<int64_t>(d) * (prev.total() + next_total) + csr.n_arcs * (d + 1) +
                 2 * (n_nodes + 1) + (prev.n_nodes + 1));
        if (gamma_out != nullptr) {
            gamma_out->swap(gamma);
        }
    }

    // Adds the seconds between two passed events to *acc; expansion and
    // dominance time also count in kernel_total_s (the filter, frontier-check
    // and cross figures are parts of the dominance time).
    void add_time(hipEvent_t a, hipEvent_t b, double *acc) {
        if (!timing) {
            return;
        }
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        *acc += ms * 1e-3;
        if (acc == &st->kernel_expand_td_s || acc == &st->kernel_dominance_s) {
            st->kernel_total_s += ms * 1e-3;
        }
    }

    // State dominance runs on top-down layers when asked for and its data came
    // with the diagram: knapsack min_weight, or set-packing partners.
    bool state_dominance_applies() const {
        if (opts.state_dominance <= 0 || dd->kind != PMDD_DD_BDD) {
            return false;
        }
        return (opts.problem_type == 1 && !dd->min_weight.empty()) ||
               (opts.problem_type == 2 && !dd->sd_partner.empty());
    }

    // Knapsack state dominance on a top-down layer (filter_dominance_knapsack_cuda,
    // topdown.cu:363-550; rule of filter_dominance_knapsack_cpu, dominance.cpp:24-86).
    void state_dominance(int l, DevFrontier &f) {
        const int64_t n = f.n_nodes;
        if (n <= 1) {
            return;
        }
        // nodes with labels: index order, or (knapsack) descending min_weight,
        // the stable order of the whole layer filtered
        std::vector<int64_t> &active = xh.active;
        active.clear();
        if (opts.problem_type == 2) {
            for (int64_t v = 0; v < n; ++v) {
                if (f.size(v) > 0) {
                    active.push_back(v);
                }
            }
        } else {
            for (int32_t v : dd->mw_order[l]) {
                if (f.size(v) > 0) {
                    active.push_back(v);
                }
            }
        }
        if (active.size() <= 1) {
            return;
        }
        std::vector<int32_t> &targets = xh.targets;
        targets.clear();
        if (opts.problem_type == 2) {
            // set packing: each node against its partner (pre-filter labels)
            const std::vector<int32_t> &pt = dd->sd_partner[l];
            for (int64_t n1 : active) {
                const int32_t n2 = pt[n1];
                if (n2 < 0 || f.size(n2) == 0) {
                    continue;
                }
                for (int64_t b = 0; b < f.size(n1); b += kBlock) {
                    targets.push_back(static_cast<int32_t>(n1));
                    targets.push_back(n2);
                    targets.push_back(-1);
                    targets.push_back(static_cast<int32_t>(b));
                }
            }
        } else {
            // knapsack: nodes in descending min_weight, each against the next
            // two (skipping same-parent, other-arc pairs)
            const std::vector<int32_t> &sp = dd->single_parent[l];
            const std::vector<int32_t> &sa = dd->single_arc[l];
            for (size_t i = 0; i + 1 < active.size(); ++i) {
                const int64_t n1 = active[i];
                int32_t cmp[2] = {-1, -1};
                int k = 0;
                for (size_t j = i + 1; j < std::min(active.size(), i + 3); ++j) {
                    const int64_t n2 = active[j];
                    if (sp[n1] >= 0 && sp[n2] >= 0 && sp[n1] == sp[n2] && sa[n1] != sa[n2]) {
                        continue;
                    }
                    cmp[k++] = static_cast<int32_t>(n2);
                }
                if (k == 0) {
                    continue;
                }
                // one workgroup per block of 256 labels of the target
                for (int64_t b = 0; b < f.size(n1); b += kBlock) {
                    targets.push_back(static_cast<int32_t>(n1));
                    targets.push_back(cmp[0]);
                    targets.push_back(cmp[1]);
                    targets.push_back(static_cast<int32_t>(b));
                }
            }
        }
        if (targets.empty()) {
            return;
        }
        const int64_t total = f.total();
        ctx->keep_a.ensure(std::max<int64_t>(total, 1), "hipMalloc(keep flags)");
        record(ctx->ev_a);
        hipLaunchKernelGGL(pmdd_k::k_fill_u8, dim3(static_cast<unsigned>(ceil_div(total, 256))),
                           dim3(256), 0, s, ctx->keep_a.as<uint8_t>(), total, uint8_t(1));
        h2d(ctx, ctx->which, targets, "hipMemcpy(dominance targets)");
        K.state_dom(dim3(static_cast<unsigned>(targets.size() / 4)), s, ctx->which.as<int32_t>(),
                    nullptr, f.off.as<int64_t>(), f.pts.as<int32_t>(), f.cap,
                    ctx->keep_a.as<uint8_t>(), t_sd);
        hip_check(hipGetLastError(), "k_state_dominance");
        // kept per node
        ctx->join_meta.ensure(static_cast<size_t>(n) * 8, "hipMalloc(kept counts)");
        hipLaunchKernelGGL(pmdd_k::k_count_kept, dim3(static_cast<unsigned>(n)), dim3(kBlock), 0,
                           s, f.off.as<int64_t>(), ctx->keep_a.as<uint8_t>(),
                           ctx->join_meta.as<int64_t>(), n);
        hip_check(hipGetLastError(), "k_count_kept");
        record(ctx->ev_b);
        std::vector<int64_t> kept;
        d2h(ctx, kept, ctx->join_meta, n, "kept counts");
        const double ms = elapsed_ms();
        st->kernel_dominance_s += ms * 1e-3;
        st->kernel_total_s += ms * 1e-3;
        st->kernel_state_dominance_s += ms * 1e-3;
        st->state_dominance_launches += 1;
        DevFrontier g;
        g.n_nodes = n;
        g.h_off.assign(n + 1, 0);
        for (int64_t v = 0; v < n; ++v) {
            g.h_off[v + 1] = g.h_off[v] + kept[v];
        }
        const int64_t dropped = total - g.total();
        if (dropped == 0) {
            return;
        }
        st->dominance_filtered_total += dropped;
        // the compacted labels go to the spare frontier, which then keeps the
        // old buffers for the next layer (no allocation per layer)
        sd_spare.n_nodes = n;
        std::swap(sd_spare.h_off, g.h_off);
        grow_points(sd_spare, d, std::max<int64_t>(sd_spare.total(), 1), 0, s);
        upload_offsets(ctx, sd_spare);
        K.compact_nodes(dim3(static_cast<unsigned>(n)), s, f.off.as<int64_t>(),
                        f.pts.as<int32_t>(), f.cap, ctx->keep_a.as<uint8_t>(),
                        sd_spare.off.as<int64_t>(), sd_spare.pts.as<int32_t>(), sd_spare.cap);
        hip_check(hipGetLastError(), "k_compact_nodes");
        std::swap(f.h_off, sd_spare.h_off);
        f.off.swap(sd_spare.off);
        f.pts.swap(sd_spare.pts);
        std::swap(f.cap, sd_spare.cap);
    }

    // State dominance on a layer whose offsets are on the device: the rule of
    // state_dominance with its targets planned by kernels (plan.hpp); f is
    // compacted into the spare frontier, which then becomes f, and the labels
    // removed are counted into expansion `step`.  cap bounds f's points (the
    // layer's candidates).
    void state_dominance_dev(int l, DevFrontier &f, int step, int64_t cap, bool size_std) {
        using namespace pmdd_k;
        const int64_t n = f.n_nodes;
        const bool knap = opts.problem_type != 2;
        const int32_t *order = knap ? dd->d_mw_order[l] : nullptr;
        const int32_t *sp = knap ? dd->d_single_parent[l] : nullptr;
        const int32_t *sa = knap ? dd->d_single_arc[l] : nullptr;
        const int32_t *partner = knap ? nullptr : dd->d_partner[l];
        const int64_t nb = ceil_div(n, kBlock);
        const int64_t max_targets = n + ceil_div(std::max<int64_t>(cap, 1), kBlock);
        ctx->sd_flag.ensure(static_cast<size_t>(std::max<int64_t>(n, 1)) * 4, "hipMalloc(sd flags)");
        ctx->sd_pos.ensure(static_cast<size_t>(n + 1) * 8, "hipMalloc(sd positions)");
        ctx->sd_active.ensure(static_cast<size_t>(std::max<int64_t>(n, 1)) * 4,
                              "hipMalloc(sd nodes)");
        ctx->sd_cmp.ensure(static_cast<size_t>(std::max<int64_t>(n, 1)) * 8, "hipMalloc(sd pairs)");
        ctx->scan_part.ensure(static_cast<size_t>(std::max<int64_t>(nb, 1)) * 8,
                              "hipMalloc(scan prefixes)");
        ctx->which.ensure(static_cast<size_t>(max_targets) * 16, "hipMalloc(dominance targets)");
        ctx->keep_a.ensure(static_cast<size_t>(std::max<int64_t>(cap, 1)), "hipMalloc(keep flags)");
        ctx->join_meta.ensure(static_cast<size_t>(std::max<int64_t>(n, 1)) * 8,
                              "hipMalloc(kept counts)");
        sd_spare.n_nodes = n;
        sd_spare.h_off.clear();
        sd_spare.off.ensure(static_cast<size_t>(n + 1) * 8, "hipMalloc(frontier offsets)");
        grow_points(sd_spare, d, std::max<int64_t>(cap, 1), 0, s);
        sample_memory();

        const size_t s0 = mark();
        int32_t *flag = ctx->sd_flag.as<int32_t>();
        int64_t *pos = ctx->sd_pos.as<int64_t>();
        int64_t *part = ctx->scan_part.as<int64_t>();
        const dim3 g(static_cast<unsigned>(nb));
        // exclusive prefix of cnt[0 .. n) into out[0 .. n]
        auto prefix = [&](auto *cnt, int64_t *out) {
            using C = std::remove_const_t<std::remove_pointer_t<decltype(cnt)>>;
            hipLaunchKernelGGL(k_block_sums<C>, g, dim3(kBlock), 0, s, n, cnt, part, arrivals(),
                               int64_t(0), out, static_cast<int64_t *>(nullptr));
            hipLaunchKernelGGL(k_scan_write<C>, g, dim3(kBlock), 0, s, n, cnt, part, int64_t(0),
                               out);
            hip_check(hipGetLastError(), "prefix scan");
        };
        // 1. the nodes holding labels, in the rule's order
        hipLaunchKernelGGL(k_sd_flags, g, dim3(kBlock), 0, s, n, order, f.off.as<int64_t>(), flag);
        hip_check(hipGetLastError(), "k_sd_flags");
        prefix(static_cast<const int32_t *>(flag), pos);
        hipLaunchKernelGGL(k_sd_active, g, dim3(kBlock), 0, s, n, order, flag, pos,
                           ctx->sd_active.as<int32_t>());
        hip_check(hipGetLastError(), "k_sd_active");
        // 2. comparison pairs and targets (the counts reuse the flags)
        hipLaunchKernelGGL(k_sd_count, g, dim3(kBlock), 0, s, n, pos + n,
                           ctx->sd_active.as<int32_t>(), f.off.as<int64_t>(), sp, sa, partner,
                           ctx->sd_cmp.as<int32_t>(), flag);
        hip_check(hipGetLastError(), "k_sd_count");
        prefix(static_cast<const int32_t *>(flag), pos);
        hipLaunchKernelGGL(k_sd_targets, g, dim3(kBlock), 0, s, n, flag, pos,
                           ctx->sd_active.as<int32_t>(), ctx->sd_cmp.as<int32_t>(),
                           ctx->which.as<int32_t>());
        hip_check(hipGetLastError(), "k_sd_targets");
        // 3. the rule, over the pre-filter labels
        hipLaunchKernelGGL(k_fill_u8, dim3(static_cast<unsigned>(ceil_div(std::max<int64_t>(cap, 1), 256))),
                           dim3(256), 0, s, ctx->keep_a.as<uint8_t>(), std::max<int64_t>(cap, 1),
                           uint8_t(1));
        // one workgroup per target, the grid sized by the bound on the
        // targets (the ones past the device count exit at once): hardware
        // dispatch of one target per workgroup measured faster than a
        // fixed grid taking targets from a queue (the rule: 42 ms against
        // 55-65 ms per bench step)
        const size_t s1 = mark();
        K.state_dom(dim3(static_cast<unsigned>(max_targets)), s,
                    ctx->which.as<int32_t>(), pos + n, f.off.as<int64_t>(), f.pts.as<int32_t>(),
                    f.cap, ctx->keep_a.as<uint8_t>(), t_sd);
        hip_check(hipGetLastError(), "k_state_dominance");
        span(s1, mark(), &st->kernel_state_dominance_s);
        // 4. kept labels per node, the new offsets, the compaction
        hipLaunchKernelGGL(k_count_kept, dim3(static_cast<unsigned>(n)), dim3(kBlock), 0, s,
                           f.off.as<int64_t>(), ctx->keep_a.as<uint8_t>(),
                           ctx->join_meta.as<int64_t>(), n);
        hip_check(hipGetLastError(), "k_count_kept");
        prefix(static_cast<const int64_t *>(ctx->join_meta.as<int64_t>()),
               sd_spare.off.as<int64_t>());
        hipLaunchKernelGGL(k_sd_dropped, dim3(1), dim3(64), 0, s, f.off.as<int64_t>() + n,
                           sd_spare.off.as<int64_t>() + n, &step_stat(step)->dropped);
        hip_check(hipGetLastError(), "k_sd_dropped");
        K.compact_nodes(dim3(static_cast<unsigned>(n)), s, f.off.as<int64_t>(),
                        f.pts.as<int32_t>(), f.cap, ctx->keep_a.as<uint8_t>(),
                        sd_spare.off.as<int64_t>(), sd_spare.pts.as<int32_t>(), sd_spare.cap);
        hip_check(hipGetLastError(), "k_compact_nodes");
        if (size_std) {
            hipLaunchKernelGGL(k_size_ssd, g, dim3(kBlock), 0, s, n, sd_spare.off.as<int64_t>(),
                               &step_stat(step)->ssd_size);
            hip_check(hipGetLastError(), "k_size_ssd");
        }
        span(s0, mark(), &st->kernel_dominance_s);
        st->state_dominance_launches += 1;
        // the compacted labels become the layer; the spare keeps the old
        // buffers for the next layer
        f.off.swap(sd_spare.off);
        f.pts.swap(sd_spare.pts);
        std::swap(f.cap, sd_spare.cap);
        f.h_off.clear();
        f.on_host = false;
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
        double acc = 0.0;
        for (int64_t v : x) {
            acc += (static_cast<double>(v) - m) * (static_cast<double>(v) - m);
        }
        return std::sqrt(acc / static_cast<double>(x.size()));
    }

    void count_layer(int64_t cand, int64_t surv) {
        st->work_candidates_total += cand;
        st->work_candidates_peak = std::max(st->work_candidates_peak, cand);
        st->work_frontier_survivors_total += surv;
        st->work_frontier_peak_points = std::max(st->work_frontier_peak_points, surv);
    }

    // Caller-owned output (no-allocation calls): the frontier is written
    // there instead of a malloc'd array.
    int32_t *user_pts = nullptr;
    int64_t user_cap = 0;
    int64_t *user_count = nullptr;

    void fetch_points(const DevFrontier &f, int64_t begin, int64_t count, pmdd_frontier *out) {
        out->d = d;
        out->count = count;
        if (user_count != nullptr) {
            *user_count = count;
            if (count > user_cap) {
                fail(PMDD_E_INVALID, "the output buffer holds " + std::to_string(user_cap) +
                                         " points; the frontier has " + std::to_string(count));
            }
        }
        std::vector<int32_t> soa(static_cast<size_t>(count) * d);
        for (int o = 0; o < d && count > 0; ++o) {
            hip_check(hipMemcpyAsync(soa.data() + o * count, f.pts.as<int32_t>() + o * f.cap + begin,
                                     static_cast<size_t>(count) * 4, hipMemcpyDeviceToHost, s),
                      "hipMemcpy(result)");
        }
        hip_check(hipStreamSynchronize(s), "result sync");
        int32_t *dst = user_pts;
        if (user_count == nullptr) {
            out->points = static_cast<int32_t *>(
                std::malloc(sizeof(int32_t) * std::max<int64_t>(count * d, 1)));
            if (out->points == nullptr) {
                fail(PMDD_E_NOMEM, "host allocation of the result failed");
            }
            dst = out->points;
        }
        for (int64_t i = 0; i < count; ++i) {
            for (int o = 0; o < d; ++o) {
                dst[i * d + o] = soa[o * count + i];
            }
        }
    }

    // ---- device-planned layers -------------------------------------------
    // A layer is planned on the device (plan.hpp) rather than by host loops
    // over its arcs and nodes: plan_layer queues the counting kernels as soon
    // as the parent frontier is queued, read_plans reads the small header
    // back (the one wait per layer), and expand_dev sizes the buffers from it
    // and queues the rest of the layer, the survivors' placement included.
    // The host planner (expand) still serves the packed serving layers, a
    // layer whose candidates exceed max_cand (cut into passes of whole
    // nodes), and every layer when PMDD_HOST_PLAN is set in the environment.
    const bool dev_plan = std::getenv("PMDD_HOST_PLAN") == nullptr;
    pmdd_k::PlanHeader hdr[2]{};
    int64_t score_rb[2] = {0, 0};

    // One expansion, folded into the statistics at the end of the run.
    struct StepRec {
        int layer = -1; // index into the per-layer std arrays (-1: none)
        int64_t n_nodes = 0, prev_nodes = 0, n_arcs = 0, cand = 0;
        int prev_step = -1;     // expansion whose survivors the parent holds, or -1
        int64_t prev_total = 0; // the parent's points when prev_step is -1
        bool dev = false;       // survivors and candidate std on the device
        bool dev_size_std = false;
        int64_t host_surv_pre = 0, host_dropped = 0;
        double host_std_c = 0.0, host_std_s = 0.0;
    };
    std::vector<StepRec> steps;
    bool want_std = false;

    // Tiers of a layer of n nodes.  A node of at most CH candidates is
    // normally filtered by one workgroup; in a layer of fewer nodes than the
    // chip has compute units those workgroups would run alone (a 2048-
    // candidate node takes about 110 us on one), so there every node above
    // the tiny tier takes the big-node path, whose cross filter spreads a
    // chunk over several workgroups.
    pmdd_k::PlanShape plan_shape(int64_t n) const {
        pmdd_k::PlanShape sh;
        sh.chunk_cap = K.chunk_cap;
        sh.small_cap = kSmallNode;
        sh.mid_cap = K.mid_cap;
        sh.big_min = n < kNarrowLayer ? kSmallNode : K.chunk_cap;
        return sh;
    }
    pmdd_k::PlanReadback *rb() const { return ctx->plan_rb.as<pmdd_k::PlanReadback>(); }
    // Arrival counter of the prefix kernels whose last workgroup finishes
    // the prefix (zero between launches; launches on the stream follow each
    // other).  Bytes 8..11 of the run counters.
    unsigned *arrivals() const { return reinterpret_cast<unsigned *>(ctx->run_ctr.as<char>() + 8); }
    // out[0 .. n] = exclusive prefix of cnt[0 .. n) (out[n] the sum), and
    // *total = the sum when total is set; two launches, no wait.
    template <class C> void device_prefix(const C *cnt, int64_t n, int64_t *out, int64_t *total) {
        const int64_t nb = ceil_div(n, kBlock);
        ctx->scan_part.ensure(static_cast<size_t>(std::max<int64_t>(nb, 1)) * 8,
                              "hipMalloc(scan prefixes)");
        hipLaunchKernelGGL(pmdd_k::k_block_sums<C>,
                           dim3(static_cast<unsigned>(std::max<int64_t>(nb, 1))), dim3(kBlock), 0, s,
                           n, cnt, ctx->scan_part.as<int64_t>(), arrivals(), int64_t(0), out,
                           total);
        hip_check(hipGetLastError(), "k_block_sums");
        if (nb > 0) {
            hipLaunchKernelGGL(pmdd_k::k_scan_write<C>, dim3(static_cast<unsigned>(nb)),
                               dim3(kBlock), 0, s, n, cnt, ctx->scan_part.as<int64_t>(),
                               int64_t(0), out);
            hip_check(hipGetLastError(), "k_scan_write");
        }
    }
    pmdd_k::StepStat *step_stat(int i) const {
        return ctx->step_stats.as<pmdd_k::StepStat>() + i;
    }
    void fill_bytes(void *p, int64_t n) {
        if (n <= 0) {
            return;
        }
        hipLaunchKernelGGL(pmdd_k::k_fill_u8, dim3(static_cast<unsigned>(ceil_div(n, 256))),
                           dim3(256), 0, s, static_cast<uint8_t *>(p), n, uint8_t(0));
        hip_check(hipGetLastError(), "k_fill_u8");
    }

    // Buffers whose size the diagram bounds, sized once per run (a plan is
    // queued while the other sweep's expansion may still run, so they must
    // not move during the run), and the run's counters zeroed.
    void prepare_plan(int sides) {
        int64_t max_e = 1, max_n = 1;
        for (int l = 0; l < dd->L; ++l) {
            max_e = std::max({max_e, dd->td[l].n_arcs, dd->bu[l].n_arcs});
            max_n = std::max(max_n, dd->layer_size[l]);
        }
        for (int k = 0; k < sides; ++k) {
            ctx->plan_run_off[k].ensure(static_cast<size_t>(max_e) * 8, "hipMalloc(plan runs)");
            ctx->plan_gamma[k].ensure(static_cast<size_t>(max_n) * 8, "hipMalloc(plan counts)");
            ctx->plan_part[k].ensure(static_cast<size_t>(ceil_div(max_n, kBlock)) *
                                         pmdd_k::kPlanPart * 8,
                                     "hipMalloc(plan prefixes)");
        }
        ctx->plan_rb.ensure(sizeof(pmdd_k::PlanReadback), "hipMalloc(plan header)");
        ctx->node_chunk.ensure(static_cast<size_t>(max_n + 1) * 4, "hipMalloc(node chunks)");
        const int64_t ss = static_cast<int64_t>(dd->L) * sizeof(pmdd_k::StepStat);
        ctx->step_stats.ensure(ss, "hipMalloc(expansion figures)");
        ctx->run_ctr.ensure(64, "hipMalloc(run counters)");
        fill_bytes(ctx->step_stats.p, ss);
        fill_bytes(ctx->run_ctr.p, 64);
        fill_bytes(ctx->plan_rb.p, sizeof(pmdd_k::PlanReadback));
        want_std = st->std_capacity > 0 && (st->std_candidates_per_layer != nullptr ||
                                             st->std_frontier_survivors_per_layer != nullptr);
        steps.reserve(dd->L);
    }

    // Queues the plan of expanding `prev` through `csr` into n nodes (its
    // header goes to the readback block of sweep direction `side`).
    void plan_layer(int side, const DevFrontier &prev, const LayerCSR &csr, int64_t n) {
        const int64_t nb = ceil_div(n, kBlock);
        const size_t p0 = mark();
        hipLaunchKernelGGL(pmdd_k::k_plan_count, dim3(static_cast<unsigned>(std::max<int64_t>(nb, 1))),
                           dim3(kBlock), 0, s, n, csr.d_off, csr.d_src, prev.off.as<int64_t>(),
                           ctx->plan_run_off[side].as<int64_t>(),
                           ctx->plan_gamma[side].as<int64_t>(), ctx->plan_part[side].as<int64_t>(),
                           plan_shape(n), arrivals(), &rb()->hdr[side]);
        hip_check(hipGetLastError(), "k_plan_count");
        span(p0, mark(), &st->kernel_expand_td_s);
    }

    // Reads the headers and scores back (waits for the stream).
    void read_plans() {
        // the device writes the block into host memory and then its
        // sequence number; the host waits for the number (polling the stream
        // now and then so that a failed stream ends the wait)
        const unsigned long long want = ++ctx->rb_seq;
        hipLaunchKernelGGL(pmdd_k::k_publish, dim3(1), dim3(64), 0, s,
                           static_cast<const pmdd_k::PlanReadback *>(rb()), ctx->d_rb_host, want);
        hip_check(hipGetLastError(), "k_publish");
        for (uint64_t spin = 1;; ++spin) {
            if (__atomic_load_n(&ctx->h_rb->seq, __ATOMIC_ACQUIRE) == want) {
                break;
            }
            if ((spin & 1023) == 0) {
                const hipError_t q = hipStreamQuery(s);
                if (q != hipSuccess && q != hipErrorNotReady) {
                    hip_check(q, "plan header wait");
                }
                if (q == hipSuccess && __atomic_load_n(&ctx->h_rb->seq, __ATOMIC_ACQUIRE) != want) {
                    fail(PMDD_E_HIP, "plan header wait: the stream is idle but the header "
                                     "was not published");
                }
            }
        }
        for (int k = 0; k < 2; ++k) {
            hdr[k] = ctx->h_rb->hdr[k];
            score_rb[k] = static_cast<int64_t>(ctx->h_rb->score[k]);
        }
    }

    // The host copy of a frontier's offsets (waits for the stream).
    void ensure_host(DevFrontier &f) {
        if (f.on_host) {
            return;
        }
        d2h(ctx, f.h_off, f.off, static_cast<size_t>(f.n_nodes + 1), "frontier offsets");
        flush_late();
        f.on_host = true;
    }

    // Queues the score of a just-built frontier (sum of |F(v)| times the
    // arcs of `deg` at v) into the readback block's slot `side`.
    void queue_score(int side, const DevFrontier &f, const LayerCSR &deg) {
        unsigned long long *slot = &rb()->score[side];
        fill_bytes(slot, sizeof(*slot));
        if (deg.d_off == nullptr || deg.off.empty() || f.n_nodes == 0) {
            return;
        }
        hipLaunchKernelGGL(pmdd_k::k_layer_score,
                           dim3(static_cast<unsigned>(ceil_div(f.n_nodes, kBlock))), dim3(kBlock),
                           0, s, f.n_nodes, f.off.as<int64_t>(), deg.d_off, slot);
        hip_check(hipGetLastError(), "k_layer_score");
    }

    // One layer of a sweep: the expansion (device-planned when its plan was
    // read and fits max_cand, else host-planned), then state dominance when
    // sd_layer >= 0, and its record.  `layer` indexes the std arrays (-1:
    // none).
    void expand_step(int side, DevFrontier &prev, const LayerCSR &csr, int64_t n_nodes,
                     DevFrontier &next, int layer, int sd_layer) {
        StepRec rec;
        rec.layer = want_std ? layer : -1;
        rec.n_nodes = n_nodes;
        rec.prev_nodes = prev.n_nodes;
        rec.n_arcs = csr.n_arcs;
        rec.prev_step = prev.on_host ? -1 : prev.step;
        rec.prev_total = prev.on_host ? prev.total() : 0;
        const bool dev = dev_plan && hdr[side].tot[pmdd_k::PV_CAND] <= opts.max_cand;
        const int step = static_cast<int>(steps.size());
        if (dev) {
            rec.dev = true;
            rec.cand = hdr[side].tot[pmdd_k::PV_CAND];
            rec.dev_size_std = rec.layer >= 0;
            expand_dev(side, prev, csr, n_nodes, next, step, rec.layer >= 0,
                       rec.layer >= 0 && sd_layer < 0);
            if (sd_layer >= 0) {
                state_dominance_dev(sd_layer, next, step, rec.cand, rec.layer >= 0);
            }
        } else {
            ensure_host(prev);
            std::vector<int64_t> gamma;
            expand(prev, csr, n_nodes, next, &rec.cand, rec.layer >= 0 ? &gamma : nullptr);
            rec.host_surv_pre = next.total();
            rec.host_std_c = pop_std(gamma);
        }
        if (sd_layer >= 0 && !dev) {
            ensure_host(next);
            const int64_t before = next.total();
            state_dominance(sd_layer, next);
            rec.host_dropped = before - next.total();
        }
        if (rec.layer >= 0 && !rec.dev_size_std) {
            ensure_host(next);
            std::vector<int64_t> sz(next.n_nodes);
            for (int64_t v = 0; v < next.n_nodes; ++v) {
                sz[v] = next.size(v);
            }
            rec.host_std_s = pop_std(sz);
        }
        steps.push_back(rec);
    }

    // Device-planned expansion of a layer whose header read_plans has read:
    // the stream is idle, so buffers are sized here from the header and the
    // whole layer is queued without waiting.
    void expand_dev(int side, const DevFrontier &prev, const LayerCSR &csr, int64_t n_nodes,
                    DevFrontier &next, int step, bool cand_std, bool size_std) {
        using namespace pmdd_k;
        const PlanHeader &h = hdr[side];
        const int64_t C = h.tot[PV_CAND];
        const int64_t nch = h.tot[PV_CHUNKS];
        const int64_t n_tiny = h.tot[PV_TINY], n_mid = h.tot[PV_MID], n_small = h.tot[PV_SMALL];
        const int64_t n_bign = h.tot[PV_BIGN], bigpos = h.tot[PV_BIGSZ], n_runs = h.tot[PV_RUNS];
        const int64_t n_bigch = nch - n_tiny - n_mid - n_small;
        ChunkBufs &cb = ctx->cb_main;
        next.n_nodes = n_nodes;
        next.h_off.clear();
        next.on_host = false;
        next.step = step;
        next.off.ensure(static_cast<size_t>(n_nodes + 1) * 8, "hipMalloc(frontier offsets)");
        grow_points(next, d, std::max<int64_t>(C, 1), 0, s);
        const size_t nc1 = static_cast<size_t>(std::max<int64_t>(nch, 1));
        cb.chunks.ensure(nc1 * sizeof(ChunkDesc), "hipMalloc(chunk table)");
        cb.nsurv.ensure(nc1 * 4, "hipMalloc(chunk counts)");
        cb.nalive.ensure(nc1 * 4, "hipMalloc(chunk counts)");
        cb.out.ensure((nc1 + 1) * 8, "hipMalloc(chunk outputs)");
        ctx->which_s.ensure(std::max<int64_t>(n_tiny, 1) * 4, "hipMalloc(chunk list)");
        ctx->which_m.ensure(std::max<int64_t>(n_mid, 1) * 4, "hipMalloc(chunk list)");
        ctx->which_a.ensure(std::max<int64_t>(n_small, 1) * 4, "hipMalloc(chunk list)");
        ctx->which_b.ensure(std::max<int64_t>(n_bigch, 1) * 4, "hipMalloc(chunk list)");
        ctx->big_node.ensure(std::max<int64_t>(n_bign, 1) * 4, "hipMalloc(big nodes)");
        ctx->big_base.ensure(static_cast<size_t>(n_bign + 1) * 8, "hipMalloc(big nodes)");
        ctx->grp_run.ensure(static_cast<size_t>(n_bign + 1) * 4, "hipMalloc(group table)");
        ctx->run_start.ensure(static_cast<size_t>(n_runs + 1) * 8, "hipMalloc(run table)");
        ctx->surv_pts.ensure(static_cast<size_t>(std::max<int64_t>(C, 1)) * d * 4,
                             "hipMalloc(survivor buffer)");
        ctx->alive.ensure(static_cast<size_t>(std::max<int64_t>(C, 1)), "hipMalloc(alive flags)");
        if (bigpos > 0) {
            ctx->sort_a.ensure(static_cast<size_t>(bigpos) * d * 4, "hipMalloc(sort buffer)");
            ctx->sort_b.ensure(static_cast<size_t>(bigpos) * d * 4, "hipMalloc(sort buffer)");
        }
        const int64_t nbc = ceil_div(nch, kBlock);
        ctx->scan_part.ensure(static_cast<size_t>(std::max<int64_t>(nbc, 1)) * 8,
                              "hipMalloc(scan prefixes)");
        sample_memory();

        const size_t m0 = mark();
        PlanTables t;
        t.chunks = cb.chunks.as<ChunkDesc>();
        t.node_chunk = ctx->node_chunk.as<int32_t>();
        t.tiny = ctx->which_s.as<int32_t>();
        t.mid = ctx->which_m.as<int32_t>();
        t.small = ctx->which_a.as<int32_t>();
        t.big_chunks = ctx->which_b.as<int32_t>();
        t.big_node = ctx->big_node.as<int32_t>();
        t.big_base = ctx->big_base.as<int64_t>();
        t.grp_run = ctx->grp_run.as<int32_t>();
        t.run_start = ctx->run_start.as<int64_t>();
        t.ssd_gamma = cand_std ? &step_stat(step)->ssd_gamma : nullptr;
        const int64_t nb = ceil_div(n_nodes, kBlock);
        if (nb > 0) {
            hipLaunchKernelGGL(k_plan_write, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s,
                               n_nodes, csr.d_off, ctx->plan_run_off[side].as<int64_t>(),
                               ctx->plan_gamma[side].as<int64_t>(),
                               ctx->plan_part[side].as<int64_t>(), &rb()->hdr[side], t,
                               plan_shape(n_nodes));
            hip_check(hipGetLastError(), "k_plan_write");
        }
        ExpandSrc src;
        src.prev_off = prev.off.as<int64_t>();
        src.prev_pts = prev.pts.as<int32_t>();
        src.prev_stride = prev.cap;
        src.arc_off = csr.d_off;
        src.arc_src = csr.d_src;
        src.arc_w = csr.d_w;
        src.w_stride = std::max<int64_t>(csr.n_arcs, 1);
        src.run_off = ctx->plan_run_off[side].as<int64_t>();
        int32_t *pts = ctx->surv_pts.as<int32_t>();
        uint8_t *alive = ctx->alive.as<uint8_t>();
        const pmdd_tuning &T = ctx->tuning[d];
        const int64_t counts[3] = {n_tiny, n_mid, n_small};
        const DevBuf *lists[3] = {&ctx->which_s, &ctx->which_m, &ctx->which_a};
        int filters = 0;
        // the chunk filters of the nodes of at most a chunk, per tier
        auto tiers = [&](hipStream_t q) {
            for (int k = 0; k < 3; ++k) {
                if (counts[k] <= 0) {
                    continue;
                }
                K.filter_tier[k][threads_index(T.filter_threads[k])](
                    dim3(static_cast<unsigned>(counts[k])), q, lists[k]->as<int32_t>(),
                    cb.chunks.as<ChunkDesc>(), &src, pts, C, alive, cb.nsurv.as<int32_t>(),
                    cb.nalive.as<int32_t>(), d_tests);
                hip_check(hipGetLastError(), "k_filter_chunks");
                ++filters;
            }
        };
        // With big nodes in the layer, the tiers run on the side stream
        // beside the big nodes' gather, merges and cross waves: both read
        // only the chunk tables and the previous frontier, and write only
        // their own chunks' slots, flags and counts.  (Without big nodes,
        // one tier beside the others measured no faster.)  Phase timing
        // keeps every launch on the main stream, where its events bracket
        // them.
        std::unique_ptr<SideFork> fork;
        if (!timing && ctx->use_side && n_bign > 0 && n_tiny + n_mid + n_small > 0) {
            fork.reset(new SideFork(ctx, s));
            tiers(ctx->side);
        }
        int32_t *sorted = nullptr;
        if (n_bign > 0) {
            // big nodes: parent runs gathered and merged into key order, then
            // staged as they are (the cross filter does all of their checks)
            K.gather_runs(dim3(static_cast<unsigned>(ceil_div(bigpos, kBlock))), s,
                          ctx->big_base.as<int64_t>(), ctx->big_node.as<int32_t>(),
                          static_cast<int32_t>(n_bign), src, ctx->sort_a.as<int32_t>(), bigpos);
            hip_check(hipGetLastError(), "k_gather_runs");
            sorted = ctx->sort_a.as<int32_t>();
            int32_t *oth = ctx->sort_b.as<int32_t>();
            const int passes = ceil_log2(h.max_runs);
            const dim3 grid(static_cast<unsigned>(ceil_div(bigpos, kBlock)));
            for (int p = 0; p < passes; ++p) {
                K.merge_pass(grid, s, sorted, oth, bigpos, ctx->run_start.as<int64_t>(),
                             ctx->grp_run.as<int32_t>(), static_cast<int32_t>(n_bign), p, bigpos);
                hip_check(hipGetLastError(), "k_merge_pass");
                std::swap(sorted, oth);
            }
            K.stage(dim3(static_cast<unsigned>(n_bigch)), s, ctx->which_b.as<int32_t>(),
                    cb.chunks.as<ChunkDesc>(), SortedSrc{sorted, bigpos}, pts, C, alive,
                    cb.nsurv.as<int32_t>(), cb.nalive.as<int32_t>());
            hip_check(hipGetLastError(), "k_stage_chunks");
        }
        const size_t m1 = mark();
        span(m0, m1, &st->kernel_expand_td_s);
        if (!fork) {
            tiers(s);
        }
        const size_t m2 = mark();
        span(m1, m2, &st->kernel_filter_s);
        if (n_bign > 0) {
            // every wave's chunks in one list; a launch per wave takes its own
            const int waves = static_cast<int>(ceil_div(h.max_chunks, kCrossWave));
            ctx->front_count.ensure(2 * nc1 * sizeof(int64_t), "hipMalloc(front counts)");
            fill_bytes(ctx->front_count.p, static_cast<int64_t>(2 * nc1 * sizeof(int64_t)));
            int64_t *fc = ctx->front_count.as<int64_t>();
            const int64_t per_wave = ceil_div(n_bigch, waves);
            const int G = static_cast<int>(std::min<int64_t>(
                T.cross_split_max, std::max<int64_t>(1, ceil_div(T.cross_workgroups, per_wave))));
            unsigned long long *cbytes = ctx->run_ctr.as<unsigned long long>();
            for (int w = 0; w < waves; ++w) {
                int64_t *fin = fc + (w % 2) * static_cast<int64_t>(nc1);
                int64_t *fout = fc + ((w + 1) % 2) * static_cast<int64_t>(nc1);
                const int only = waves > 1 ? w : -1;
                K.cross(dim3(static_cast<unsigned>(n_bigch * G)), s, ctx->which_b.as<int32_t>(),
                        cb.chunks.as<ChunkDesc>(), pts, C, alive, cb.nsurv.as<int32_t>(), sorted,
                        bigpos, fin, kCrossWave, G, only, t_cross);
                hip_check(hipGetLastError(), "k_cross_chunks");
                K.append_front(dim3(static_cast<unsigned>(n_bigch)), s, ctx->which_b.as<int32_t>(),
                               cb.chunks.as<ChunkDesc>(), pts, C, alive, cb.nsurv.as<int32_t>(),
                               cb.nalive.as<int32_t>(), sorted, bigpos, fin, fout, kCrossWave,
                               only, cbytes);
                hip_check(hipGetLastError(), "k_append_front");
            }
            st->cross_launches += waves;
        }
        if (fork) {
            fork->join(); // the placement reads every chunk's counts
        }
        const size_t m3 = mark();
        span(m1, m3, &st->kernel_dominance_s);
        if (n_bign > 0) {
            span(m2, m3, &st->kernel_cross_s); // the cross waves with their appends
        }
        st->filter_launches += filters;
        st->filter_candidates += C;

        // placement: chunk output slots, node offsets, then the survivors
        int64_t *slot = cb.out.as<int64_t>();
        hipLaunchKernelGGL(k_block_sums<int32_t>,
                           dim3(static_cast<unsigned>(std::max<int64_t>(nbc, 1))), dim3(kBlock), 0,
                           s, nch, static_cast<const int32_t *>(cb.nalive.as<int32_t>()),
                           ctx->scan_part.as<int64_t>(), arrivals(), int64_t(0), slot,
                           &step_stat(step)->surv);
        hip_check(hipGetLastError(), "k_block_sums");
        if (nbc > 0) {
            hipLaunchKernelGGL(k_scan_write, dim3(static_cast<unsigned>(nbc)), dim3(kBlock), 0, s,
                               nch, cb.nalive.as<int32_t>(), ctx->scan_part.as<int64_t>(),
                               int64_t(0), slot);
            hip_check(hipGetLastError(), "k_scan_write");
        }
        hipLaunchKernelGGL(k_node_offsets, dim3(static_cast<unsigned>(ceil_div(n_nodes + 1, kBlock))),
                           dim3(kBlock), 0, s, n_nodes, ctx->node_chunk.as<int32_t>(), slot,
                           next.off.as<int64_t>(),
                           size_std ? &step_stat(step)->ssd_size : static_cast<double *>(nullptr));
        hip_check(hipGetLastError(), "k_node_offsets");
        if (nch > 0) {
            K.compact(dim3(static_cast<unsigned>(nch)), s, cb.chunks.as<ChunkDesc>(),
                      cb.nsurv.as<int32_t>(), slot, pts, C, alive, next.pts.as<int32_t>(),
                      next.cap);
            hip_check(hipGetLastError(), "k_compact_chunks");
        }
        span(m3, mark(), &st->kernel_expand_td_s);
    }

    // Expansion records into the statistics (the stream is idle).
    void fold_steps() {
        if (steps.empty()) {
            return;
        }
        std::vector<pmdd_k::StepStat> ss;
        d2h(ctx, ss, ctx->step_stats, steps.size(), "expansion figures");
        std::vector<unsigned long long> ctr;
        d2h(ctx, ctr, ctx->run_ctr, 1, "run counters");
        st->cross_bytes += static_cast<int64_t>(ctr[0]);
        std::vector<int64_t> post(steps.size(), 0);
        for (size_t i = 0; i < steps.size(); ++i) {
            const StepRec &r = steps[i];
            const int64_t pre = r.dev ? ss[i].surv : r.host_surv_pre;
            const int64_t surv = pre - ss[i].dropped - r.host_dropped;
            post[i] = surv;
            count_layer(r.cand, surv);
            st->dominance_filtered_total += ss[i].dropped;
            if (r.dev) {
                const int64_t prev_total = r.prev_step >= 0 ? post[r.prev_step] : r.prev_total;
                st->algorithmic_bytes +=
                    4 * (static_cast<int64_t>(d) * (prev_total + pre) + r.n_arcs * (d + 1) +
                         2 * (r.n_nodes + 1) + (r.prev_nodes + 1));
            }
            if (r.layer >= 0 && r.layer < st->std_capacity) {
                const double n = static_cast<double>(std::max<int64_t>(r.n_nodes, 1));
                if (st->std_candidates_per_layer != nullptr) {
                    st->std_candidates_per_layer[r.layer] =
                        r.dev ? std::sqrt(ss[i].ssd_gamma / n) : r.host_std_c;
                }
                if (st->std_frontier_survivors_per_layer != nullptr) {
                    st->std_frontier_survivors_per_layer[r.layer] =
                        r.dev_size_std ? std::sqrt(ss[i].ssd_size / n) : r.host_std_s;
                }
            }
        }
        steps.clear();
    }

    // Method 1 (enumerate_bdd_topdown / enumerate_mdd_topdown).
    void topdown(pmdd_frontier *out) {
        const int L = dd->L;
        FrontierLease lc(ctx, d), ln(ctx, d);
        DevFrontier &cur = lc.f, &nxt = ln.f;
        seed_frontier(ctx, cur, d, dd->layer_size[0], dd->root);
        const bool sd = state_dominance_applies();
        prepare_plan(1);
        if (dev_plan) {
            plan_layer(0, cur, dd->td[1], dd->layer_size[1]);
        }
        for (int l = 1; l < L; ++l) {
            if (dev_plan) {
                read_plans();
            }
            expand_step(0, cur, dd->td[l], dd->layer_size[l], nxt, l, sd ? l : -1);
            if (dev_plan && l + 1 < L) {
                plan_layer(0, nxt, dd->td[l + 1], dd->layer_size[l + 1]);
            }
            std::swap(cur, nxt);
        }
        ensure_host(cur);
        fetch_points(cur, cur.h_off[dd->terminal], cur.size(dd->terminal), out);
    }

    // Method 2 on the GPU (the reference runs it on the CPU only,
    // bottomup_cpu_enumerate enum/cpu/enum.cpp:123-229, and rejects it for the
    // GPU, main.cpp:356-361): bottom-up expansion from the terminal with the
    // same kernels; the root's frontier is the result.  State dominance does
    // not apply (the CPU path ignores it too, enum.cpp:127).
    void bottomup(pmdd_frontier *out) {
        const int L = dd->L;
        FrontierLease lc(ctx, d), ln(ctx, d);
        DevFrontier &cur = lc.f, &nxt = ln.f;
        seed_frontier(ctx, cur, d, dd->layer_size[L - 1], dd->terminal);
        prepare_plan(1);
        if (dev_plan) {
            plan_layer(0, cur, dd->bu[L - 2], dd->layer_size[L - 2]);
        }
        for (int l = L - 2; l >= 0; --l) {
            if (dev_plan) {
                read_plans();
            }
            expand_step(0, cur, dd->bu[l], dd->layer_size[l], nxt, l, -1);
            if (dev_plan && l > 0) {
                plan_layer(0, nxt, dd->bu[l - 1], dd->layer_size[l - 1]);
            }
            std::swap(cur, nxt);
        }
        ensure_host(cur);
        fetch_points(cur, cur.h_off[dd->root], cur.size(dd->root), out);
    }

    // Expansion score of a frontier (sum of |F(v)| times v's arcs in `deg`),
    // on the host.
    static int64_t host_score(const DevFrontier &f, const LayerCSR &deg) {
        int64_t score = 0;
        for (int64_t v = 0; v < f.n_nodes; ++v) {
            const int64_t k = deg.off.empty() ? 0 : deg.off[v + 1] - deg.off[v];
            score += f.size(v) * k;
        }
        return score;
    }

    // Method 3 (coupled_bdd_cuda_enumerate enum.cu:595-756, enumerate_mdd_coupled
    // enum.cu:295-436): expand top-down while its score is not above the
    // bottom-up score, else bottom-up; the scores are sum |F(v)| * out-arcs(v)
    // and trunc(1.5 * sum |F(v)| * in-arcs(v)); stop where the sweeps meet.
    // Device-planned: after a sweep's expansion its score and the plan of its
    // next layer are queued together and read back in one wait; the other
    // sweep's plan, queued when it last moved, is still valid (its frontier
    // has not changed).
    void coupled(pmdd_frontier *out) {
        const int L = dd->L;
        FrontierLease lt_(ctx, d), lb_(ctx, d), lm_(ctx, d);
        DevFrontier &td = lt_.f, &bu = lb_.f, &tmp = lm_.f;
        seed_frontier(ctx, td, d, dd->layer_size[0], dd->root);
        seed_frontier(ctx, bu, d, dd->layer_size[L - 1], dd->terminal);
        // knapsack state dominance on the top-down sweep (the reference packs
        // its metadata for method 3 but does not apply it, enum.cu:618): the
        // rule is sound, so the frontier is unchanged and the join smaller
        const bool sd = state_dominance_applies();
        int lt = 0, lb = L - 1;
        int64_t td_score = 0, bu_score = 0;
        prepare_plan(2);
        if (dev_plan) {
            plan_layer(0, td, dd->td[1], dd->layer_size[1]);
            plan_layer(1, bu, dd->bu[L - 2], dd->layer_size[L - 2]);
            read_plans();
        }
        while (lt != lb) {
            if (td_score <= bu_score) {
                ++lt;
                expand_step(0, td, dd->td[lt], dd->layer_size[lt], tmp, -1, sd ? lt : -1);
                std::swap(td, tmp);
                if (!dev_plan) {
                    td_score = host_score(td, dd->bu[lt]);
                } else if (lt != lb) {
                    queue_score(0, td, dd->bu[lt]);
                    plan_layer(0, td, dd->td[lt + 1], dd->layer_size[lt + 1]);
                    read_plans();
                    td_score = score_rb[0];
                }
            } else {
                --lb;
                expand_step(1, bu, dd->bu[lb], dd->layer_size[lb], tmp, -1, -1);
                std::swap(bu, tmp);
                int64_t score = 0;
                if (!dev_plan) {
                    score = host_score(bu, dd->td[lb]);
                } else if (lt != lb) {
                    queue_score(1, bu, dd->td[lb]);
                    plan_layer(1, bu, dd->bu[lb - 1], dd->layer_size[lb - 1]);
                    read_plans();
                    score = score_rb[1];
                }
                bu_score = static_cast<int64_t>(1.5 * static_cast<double>(score));
            }
        }
        st->layer_coupling = lt;
        ensure_host(td);
        ensure_host(bu);
        const double t0 = now_s();
        couple(td, bu, out);
        st->wall_join_s += now_s() - t0;
    }

    // Cutset coupling (couple_cutsets_cuda, couple.cu:340-498): the result is
    // the nondominated set of the union over cutset nodes of TD(v) + BU(v).
    // Nodes with both frontiers non-empty, largest |TD|+|BU| first, go in
    // batches of at most max_prod products (a node with more is split into
    // pieces).  Per batch: the products are checked against the running
    // frontier R (kept in key order); the survivors are filtered chunk by
    // chunk and merged into key order, then in descending-key blocks filtered
    // locally and against earlier blocks, giving the batch frontier B; R
    // loses the points B strictly dominates and B is merged into R.  The host
    // waits twice per batch: for the products past R (they size the rest),
    // and at the end for the batch's counts; everything between is counted
    // and placed on the device.
    void couple(const DevFrontier &td, const DevFrontier &bu, pmdd_frontier *out) {
        const int CH = K.chunk_cap;
        const int64_t n = td.n_nodes;
        std::vector<int64_t> nodes;
        int64_t total_products = 0;
        for (int64_t v = 0; v < n; ++v) {
            const int64_t p = td.size(v) * bu.size(v);
            total_products += p;
            if (p > 0) {
                nodes.push_back(v);
            }
        }
        st->work_join_products_total += total_products;
        std::stable_sort(nodes.begin(), nodes.end(), [&](int64_t a, int64_t b) {
            return td.size(a) + bu.size(a) > td.size(b) + bu.size(b);
        });

        FrontierLease lr(ctx, d), lbt(ctx, d);
        DevFrontier &R = lr.f, &B = lbt.f;
        int64_t r_count = 0;
        size_t i = 0;
        std::vector<ChunkDesc> rchunks;
        ChunkBufs &cb = ctx->cb_main;
        ChunkBufs &cr = ctx->cb_aux;
        const std::vector<int32_t> none;
        // Batch sizes: with an explicit max_prod every batch holds at most
        // max_prod products.  As in the reference (couple.cu:393-403) a batch
        // takes whole nodes while they fit; a node with more products than
        // the cap is split into pieces (runs of its TD rows times all its BU
        // columns, or runs of columns of one row when one row exceeds the
        // cap), each its own batch, so unlike the reference no batch exceeds
        // it.  Adaptive (max_prod 0): batches start at 32K products and
        // double up to kJoinBatchCap, so the first batches build R cheaply and
        // the later, large ones are mostly filtered by it in few launches.
        // The result does not depend on the split.
        const bool adaptive = opts.max_prod <= 0;
        int64_t batch_cap = adaptive ? (int64_t(1) << 15) : opts.max_prod;
        const int64_t batch_limit =
            adaptive ? (ctx->join_cap > 0 ? ctx->join_cap : kJoinBatchCap) : opts.max_prod;
        // next piece: node nodes[i], its TD rows from cur_r, BU columns from cur_c
        int64_t cur_r = 0, cur_c = 0;
        struct Piece {
            int64_t v, r0, r1, c0, c1;
        };
        std::vector<Piece> pieces;
        while (i < nodes.size()) {
            int64_t bprod = 0;
            pieces.clear();
            while (i < nodes.size()) {
                const int64_t v = nodes[i];
                const int64_t T = td.size(v), Bn = bu.size(v);
                if (cur_r == 0 && cur_c == 0 && T * Bn <= batch_cap - bprod) {
                    pieces.push_back({v, 0, T, 0, Bn}); // the whole node
                    bprod += T * Bn;
                    ++i;
                    continue;
                }
                if (bprod > 0) {
                    break; // it does not fit next to the batch's nodes
                }
                if (cur_c == 0 && Bn <= batch_cap) {
                    const int64_t rows = std::min(T - cur_r, batch_cap / Bn);
                    pieces.push_back({v, cur_r, cur_r + rows, 0, Bn});
                    bprod += rows * Bn;
                    cur_r += rows;
                } else {
                    const int64_t cols = std::min(Bn - cur_c, batch_cap);
                    pieces.push_back({v, cur_r, cur_r + 1, cur_c, cur_c + cols});
                    bprod += cols;
                    cur_c += cols;
                    if (cur_c == Bn) {
                        cur_c = 0;
                        ++cur_r;
                    }
                }
                if (cur_r == T) {
                    cur_r = 0;
                    ++i;
                }
                break; // a piece fills its batch
            }
            batch_cap = std::min<int64_t>(batch_limit, batch_cap * 2);
            st->join_batches += 1;
            st->join_batch_peak_products = std::max(st->join_batch_peak_products, bprod);
            const int32_t Kn = static_cast<int32_t>(pieces.size());
            std::vector<int64_t> meta(4 * static_cast<size_t>(Kn) + 1);
            int64_t *prod_off = meta.data();
            int64_t *row_base = meta.data() + Kn + 1;
            int64_t *col_base = row_base + Kn;
            int64_t *col_size = col_base + Kn;
            prod_off[0] = 0;
            int64_t batch_inputs = 0;
            for (int32_t k = 0; k < Kn; ++k) {
                const Piece &pc = pieces[k];
                row_base[k] = td.h_off[pc.v] + pc.r0;
                col_base[k] = bu.h_off[pc.v] + pc.c0;
                col_size[k] = pc.c1 - pc.c0;
                prod_off[k + 1] = prod_off[k] + (pc.r1 - pc.r0) * (pc.c1 - pc.c0);
                batch_inputs += (pc.r1 - pc.r0) + (pc.c1 - pc.c0);
            }
            h2d(ctx, ctx->join_meta, meta, "hipMemcpy(batch table)");
            pmdd_k::JoinSrc src;
            src.prod_off = ctx->join_meta.as<int64_t>();
            src.row_base = src.prod_off + Kn + 1;
            src.col_base = src.row_base + Kn;
            src.col_size = src.col_base + Kn;
            src.K = Kn;
            src.td_pts = td.pts.as<int32_t>();
            src.td_stride = td.cap;
            src.bu_pts = bu.pts.as<int32_t>();
            src.bu_stride = bu.cap;

            // 1. the batch's products, densely, in product order; then the
            // check against the running frontier R (descending key order) in
            // blocks of growing length: most products meet a dominator among
            // R's first points, so the later blocks see few of them
            const size_t pbytes = static_cast<size_t>(bprod) * d * 4;
            ctx->join_x.ensure(pbytes, "hipMalloc(product buffer)");
            ctx->join_y.ensure(pbytes, "hipMalloc(product buffer)");
            std::vector<int64_t> bounds;
            if (r_count > 0) {
                bounds.assign(1, 0);
                for (int64_t b : {int64_t(512), int64_t(2048)}) {
                    if (b < r_count) {
                        bounds.push_back(b);
                    }
                }
                while (bounds.back() + kRefillBlock < r_count) {
                    bounds.push_back(bounds.back() + kRefillBlock);
                }
                bounds.push_back(r_count);
            }
            const int nbk = std::max<int>(0, static_cast<int>(bounds.size()) - 1);
            const int serial = nbk > kSerialBlocks + 1 ? kSerialBlocks : nbk;
            const int64_t tail_rows =
                serial < nbk ? ceil_div(r_count - bounds[serial], kRefillBlock) : 0;
            // counters: [0] products, [1 + b] survivors of R block b
            std::vector<unsigned long long> ctr(1 + static_cast<size_t>(nbk), 0);
            ctr[0] = static_cast<unsigned long long>(bprod);
            h2d(ctx, ctx->join_ctr, ctr, "hipMemcpy(join counters)");
            unsigned long long *dc = ctx->join_ctr.as<unsigned long long>();
            const size_t m0 = mark();
            K.join_products(dim3(static_cast<unsigned>(
                                std::min<int64_t>(ceil_div(bprod, pmdd_k::kBlock), 16384))),
                            s, src, bprod, ctx->join_x.as<int32_t>(), bprod);
            hip_check(hipGetLastError(), "k_join_products");
            const size_t m1 = mark();
            span(m0, m1, &st->kernel_expand_td_s);
            int32_t *F = ctx->join_x.as<int32_t>();
            if (nbk > 0) {
                const unsigned grid = static_cast<unsigned>(std::min<int64_t>(
                    ceil_div(bprod, pmdd_k::kWave), ctx->tuning[d].refill_waves));
                int32_t *in = ctx->join_x.as<int32_t>();
                int32_t *outb = ctx->join_y.as<int32_t>();
                for (int b = 0; b < serial; ++b) {
                    K.refill_vs(dim3(grid), s, in, bprod, dc + b, R.pts.as<int32_t>(), R.cap,
                                bounds[b], bounds[b + 1], outb, bprod, dc + 1 + b, t_check);
                    hip_check(hipGetLastError(), "k_refill_vs_block");
                    std::swap(in, outb);
                }
                if (serial < nbk) {
                    // R's rows from bounds[serial] on, in grid rows of
                    // kRefillBlock (the last one shorter, ending at r_count)
                    ctx->join_flag.ensure(static_cast<size_t>(bprod), "hipMalloc(product flags)");
                    uint8_t *pf = ctx->join_flag.as<uint8_t>();
                    hipLaunchKernelGGL(pmdd_k::k_fill_u8,
                                       dim3(static_cast<unsigned>(ceil_div(bprod, 256))), dim3(256),
                                       0, s, pf, bprod, uint8_t(1));
                    hip_check(hipGetLastError(), "k_fill_u8");
                    const unsigned gx = static_cast<unsigned>(
                        std::min<int64_t>(ceil_div(bprod, pmdd_k::kWave), kTailWaves));
                    K.refill_flag_rows(dim3(gx, static_cast<unsigned>(tail_rows)), s, in, bprod,
                                       dc + serial, R.pts.as<int32_t>(), R.cap, bounds[serial],
                                       r_count, static_cast<const int64_t *>(nullptr),
                                       kRefillBlock, pf, t_check);
                    hip_check(hipGetLastError(), "k_refill_vs_block");
                    K.flag_append(dim3(static_cast<unsigned>(
                                      std::min<int64_t>(ceil_div(bprod, kBlock), 4096))),
                                  s, in, bprod, dc + serial, pf, outb, bprod, dc + nbk);
                    hip_check(hipGetLastError(), "k_flag_append");
                    std::swap(in, outb);
                }
                F = in;
                const size_t m2 = mark();
                span(m1, m2, &st->kernel_frontier_check_s);
                span(m1, m2, &st->kernel_dominance_s);
                st->frontier_check_launches += serial + (serial < nbk ? 1 : 0);
            }
            // the first of the batch's two waits: the products past R size
            // the batch's own filter
            d2h(ctx, ctr, ctx->join_ctr, ctr.size(), "join counters");
            flush_late();
            const int64_t ytotal = static_cast<int64_t>(ctr[nbk]);
            for (int b = 0; b < serial; ++b) {
                const int64_t n_in = static_cast<int64_t>(ctr[b]);
                const int64_t n_out = static_cast<int64_t>(ctr[b + 1]);
                st->frontier_check_bytes +=
                    static_cast<int64_t>(d) * 4 * (n_in + n_out + bounds[b + 1] - bounds[b]);
            }
            if (serial < nbk) {
                // every later block reads the products and their flags and
                // the block; the append reads them once and writes the rest
                const int64_t n_in = static_cast<int64_t>(ctr[serial]);
                const int64_t rows = bounds[nbk] - bounds[serial];
                st->frontier_check_bytes +=
                    (static_cast<int64_t>(d) * 4 + 2) * n_in * tail_rows +
                    static_cast<int64_t>(d) * 4 * (rows + n_in + ytotal) + n_in;
            }
            if (debug_join) {
                std::fprintf(stderr, "join batch: nodes=%d products=%lld after_R=%lld", Kn,
                             static_cast<long long>(bprod), static_cast<long long>(ytotal));
            }
            if (ytotal == 0) {
                if (debug_join) {
                    std::fprintf(stderr, " R=%lld\n", static_cast<long long>(r_count));
                }
                continue;
            }
            // From here on every size is on the device (counted into bc:
            // [0] products past R, [1] their local survivors, [2] the batch
            // frontier B, [3] R's points B leaves, [4] the new R), every
            // table is sized by the products past R, and the batch ends with
            // its second wait.
            const int64_t cnt = ceil_div(ytotal, CH);
            ctx->batch_ctr.ensure(64, "hipMalloc(batch counts)");
            int64_t *bc = ctx->batch_ctr.as<int64_t>();
            cb.chunks.ensure(static_cast<size_t>(cnt) * sizeof(ChunkDesc), "hipMalloc(chunk table)");
            cb.nsurv.ensure(static_cast<size_t>(cnt) * 4, "hipMalloc(chunk counts)");
            cb.nalive.ensure(static_cast<size_t>(cnt) * 4, "hipMalloc(chunk counts)");
            cb.out.ensure(static_cast<size_t>(cnt + 1) * 8, "hipMalloc(chunk outputs)");
            ctx->surv_pts.ensure(static_cast<size_t>(ytotal) * d * 4, "hipMalloc(survivor buffer)");
            ctx->alive.ensure(static_cast<size_t>(ytotal), "hipMalloc(alive flags)");
            ctx->sort_a.ensure(static_cast<size_t>(ytotal) * d * 4, "hipMalloc(sort buffer)");
            ctx->sort_b.ensure(static_cast<size_t>(ytotal) * d * 4, "hipMalloc(sort buffer)");
            ctx->grp_run.ensure(2 * sizeof(int32_t), "hipMalloc(group table)");
            ctx->run_start.ensure(3 * sizeof(int64_t), "hipMalloc(run table)");
            grow_points(B, d, ytotal, 0, s);
            const dim3 cgrid(static_cast<unsigned>(ceil_div(cnt, kBlock)));
            int32_t *pts = ctx->surv_pts.as<int32_t>();
            uint8_t *alive = ctx->alive.as<uint8_t>();
            // 2. the survivors filtered chunk by chunk (presorted in LDS),
            // written in key order per chunk and merged into one key order
            hipLaunchKernelGGL(pmdd_k::k_count_to_i64, dim3(1), dim3(64), 0, s, dc + nbk, bc + 0);
            hipLaunchKernelGGL(pmdd_k::k_uniform_chunks, cgrid, dim3(kBlock), 0, s, bc + 0, cnt, CH,
                               false, cb.chunks.as<ChunkDesc>(), cb.nsurv.as<int32_t>(),
                               cb.nalive.as<int32_t>());
            hip_check(hipGetLastError(), "k_uniform_chunks");
            const size_t m3 = mark();
            pmdd_k::SortedSrc fsrc{F, bprod};
            const int dense = run_filter(K.filter_dense, cb, ctx->which_a, none, cnt, &fsrc, pts,
                                         ytotal, alive);
            const size_t m4 = mark();
            span(m3, m4, &st->kernel_filter_s);
            span(m3, m4, &st->kernel_dominance_s);
            int64_t *slot = cb.out.as<int64_t>();
            device_prefix(cb.nalive.as<int32_t>(), cnt, slot, bc + 1);
            K.compact(dim3(static_cast<unsigned>(cnt)), s, cb.chunks.as<ChunkDesc>(),
                      cb.nsurv.as<int32_t>(), slot, pts, ytotal, alive, ctx->sort_a.as<int32_t>(),
                      ytotal);
            hip_check(hipGetLastError(), "k_compact_chunks");
            // runs: the chunks' survivors (slot prefix, its end the count)
            h2d(ctx, ctx->grp_run, std::vector<int32_t>{0, static_cast<int32_t>(cnt)},
                "hipMemcpy(group table)");
            int32_t *ys = ctx->sort_a.as<int32_t>();
            {
                int32_t *oth = ctx->sort_b.as<int32_t>();
                const int passes = ceil_log2(cnt);
                const dim3 grid(static_cast<unsigned>(ceil_div(ytotal, kBlock)));
                for (int p = 0; p < passes; ++p) {
                    K.merge_pass(grid, s, ys, oth, ytotal, slot, ctx->grp_run.as<int32_t>(), 1, p,
                                 ytotal);
                    hip_check(hipGetLastError(), "k_merge_pass");
                    std::swap(ys, oth);
                }
            }
            const size_t m5 = mark();
            span(m4, m5, &st->kernel_expand_td_s);
            // 3. descending-key blocks of the merged survivors: a block's
            // local filter and the check against the earlier blocks (the
            // group's front), giving the batch frontier B
            hipLaunchKernelGGL(pmdd_k::k_uniform_chunks, cgrid, dim3(kBlock), 0, s, bc + 1, cnt, CH,
                               true, cb.chunks.as<ChunkDesc>(), cb.nsurv.as<int32_t>(),
                               cb.nalive.as<int32_t>());
            hip_check(hipGetLastError(), "k_uniform_chunks");
            pmdd_k::SortedSrc ss{ys, ytotal};
            int sorted_filters = 0, waves = 0;
            if (cnt == 1) {
                sorted_filters = run_filter(K.filter_sorted, cb, ctx->which_a, none, 1, &ss, pts,
                                            ytotal, alive);
            } else {
                K.stage(dim3(static_cast<unsigned>(cnt)), s, nullptr, cb.chunks.as<ChunkDesc>(), ss,
                        pts, ytotal, alive, cb.nsurv.as<int32_t>(), cb.nalive.as<int32_t>());
                hip_check(hipGetLastError(), "k_stage_chunks");
                waves = static_cast<int>(ceil_div(cnt, kCrossWave));
                ctx->front_count.ensure(2 * static_cast<size_t>(cnt) * sizeof(int64_t),
                                        "hipMalloc(front counts)");
                fill_bytes(ctx->front_count.p, 2 * cnt * static_cast<int64_t>(sizeof(int64_t)));
                int64_t *fc = ctx->front_count.as<int64_t>();
                const pmdd_tuning &T = ctx->tuning[d];
                const int64_t per_wave = ceil_div(cnt, waves);
                const int G = static_cast<int>(std::min<int64_t>(
                    T.cross_split_max, std::max<int64_t>(1, ceil_div(T.cross_workgroups, per_wave))));
                for (int w = 0; w < waves; ++w) {
                    int64_t *fin = fc + (w % 2) * cnt;
                    int64_t *fout = fc + ((w + 1) % 2) * cnt;
                    const int only = waves > 1 ? w : -1;
                    K.cross(dim3(static_cast<unsigned>(cnt * G)), s, nullptr,
                            cb.chunks.as<ChunkDesc>(), pts, ytotal, alive, cb.nsurv.as<int32_t>(),
                            ys, ytotal, fin, kCrossWave, G, only, t_cross);
                    hip_check(hipGetLastError(), "k_cross_chunks");
                    K.append_front(dim3(static_cast<unsigned>(cnt)), s, nullptr,
                                   cb.chunks.as<ChunkDesc>(), pts, ytotal, alive,
                                   cb.nsurv.as<int32_t>(), cb.nalive.as<int32_t>(), ys, ytotal, fin,
                                   fout, kCrossWave, only, ctx->run_ctr.as<unsigned long long>());
                    hip_check(hipGetLastError(), "k_append_front");
                }
                st->cross_launches += waves;
            }
            const size_t m6 = mark();
            span(m5, m6, &st->kernel_dominance_s);
            span(m5, m6, waves > 0 ? &st->kernel_cross_s : &st->kernel_filter_s);
            // 4. B, dense in key order
            device_prefix(cb.nalive.as<int32_t>(), cnt, slot, bc + 2);
            K.compact(dim3(static_cast<unsigned>(cnt)), s, cb.chunks.as<ChunkDesc>(),
                      cb.nsurv.as<int32_t>(), slot, pts, ytotal, alive, B.pts.as<int32_t>(), B.cap);
            hip_check(hipGetLastError(), "k_compact_chunks");
            const size_t m7 = mark();
            span(m6, m7, &st->kernel_expand_td_s);
            const int64_t r_prev = r_count;
            if (r_count == 0) {
                // the first batch frontier is R
                std::swap(R.pts, B.pts);
                std::swap(R.cap, B.cap);
            } else {
                // 5. R minus what B strictly dominates (B passed the check
                // against this R, so no point of B equals one of R, and
                // "strictly dominated" is "weakly dominated": the refill
                // check in flag mode, R in place against the whole of B),
                // then R and B merged by key
                rchunks.clear();
                const int32_t rn = static_cast<int32_t>(ceil_div(r_count, CH));
                std::vector<int32_t> rlen(rn);
                for (int32_t k = 0; k < rn; ++k) {
                    ChunkDesc c;
                    c.start = static_cast<int64_t>(k) * CH;
                    c.len = static_cast<int32_t>(std::min<int64_t>(CH, r_count - c.start));
                    c.out_base = c.start;
                    c.group = 0;
                    c.group_first = 0;
                    c.group_count = 1;
                    rchunks.push_back(c);
                    rlen[k] = c.len;
                }
                upload_table(cr, rchunks);
                cr.out.ensure(static_cast<size_t>(rn + 1) * 8, "hipMalloc(chunk outputs)");
                h2d(ctx, cr.nsurv, rlen, "hipMemcpy(frontier chunk sizes)");
                ctx->keep_b.ensure(static_cast<size_t>(r_count), "hipMalloc(frontier flags)");
                hipLaunchKernelGGL(pmdd_k::k_fill_u8,
                                   dim3(static_cast<unsigned>(ceil_div(r_count, 256))), dim3(256), 0,
                                   s, ctx->keep_b.as<uint8_t>(), r_count, uint8_t(1));
                h2d(ctx, ctx->join_ctr,
                    std::vector<unsigned long long>{static_cast<unsigned long long>(r_count)},
                    "hipMemcpy(frontier count)");
                const int64_t SM = r_count + ytotal; // merge stride: R's and B's bound
                DevBuf &M = ctx->merge_a;
                DevBuf &Rn = ctx->merge_b;
                M.ensure(static_cast<size_t>(SM) * d * 4, "hipMalloc(frontier merge)");
                Rn.ensure(static_cast<size_t>(SM) * d * 4, "hipMalloc(frontier merge)");
                const size_t e0 = mark();
                // B (its length a device count, at most ytotal) in grid rows
                // of kRefillBlock points (more when ytotal would ask for over
                // 64 rows), so that R's points, which mostly outlast B, walk
                // its blocks side by side; rows past B's end exit at once
                const int64_t blen = std::max<int64_t>(kRefillBlock, ceil_div(ytotal, 64));
                const int64_t brows = std::max<int64_t>(1, ceil_div(ytotal, blen));
                const unsigned rgrid = static_cast<unsigned>(std::min<int64_t>(
                    ceil_div(r_count, pmdd_k::kWave),
                    brows > 1 ? kTailWaves : ctx->tuning[d].refill_waves));
                K.refill_flag_rows(dim3(rgrid, static_cast<unsigned>(brows)), s,
                                   R.pts.as<int32_t>(), R.cap,
                                   ctx->join_ctr.as<unsigned long long>(), B.pts.as<int32_t>(),
                                   B.cap, 0, 0, bc + 2, blen, ctx->keep_b.as<uint8_t>(),
                                   t_check);
                hip_check(hipGetLastError(), "k_refill_vs_block");
                const size_t e1 = mark();
                span(e0, e1, &st->kernel_frontier_check_s);
                span(e0, e1, &st->kernel_dominance_s);
                st->frontier_check_launches += 1;
                hipLaunchKernelGGL(pmdd_k::k_count_chunks, dim3(static_cast<unsigned>(rn)),
                                   dim3(kBlock), 0, s, cr.chunks.as<ChunkDesc>(),
                                   cr.nsurv.as<int32_t>(), ctx->keep_b.as<uint8_t>(),
                                   cr.nalive.as<int32_t>());
                hip_check(hipGetLastError(), "k_count_chunks");
                device_prefix(cr.nalive.as<int32_t>(), rn, cr.out.as<int64_t>(), bc + 3);
                K.compact(dim3(static_cast<unsigned>(rn)), s, cr.chunks.as<ChunkDesc>(),
                          cr.nsurv.as<int32_t>(), cr.out.as<int64_t>(), R.pts.as<int32_t>(), R.cap,
                          ctx->keep_b.as<uint8_t>(), M.as<int32_t>(), SM);
                hip_check(hipGetLastError(), "k_compact_chunks");
                hipLaunchKernelGGL(pmdd_k::k_copy_at,
                                   dim3(static_cast<unsigned>(
                                       std::min<int64_t>(ceil_div(ytotal, kBlock), 4096))),
                                   dim3(kBlock), 0, s, B.pts.as<int32_t>(), B.cap, bc + 2, d,
                                   M.as<int32_t>(), SM, bc + 3);
                hip_check(hipGetLastError(), "k_copy_at");
                hipLaunchKernelGGL(pmdd_k::k_two_runs, dim3(1), dim3(64), 0, s, bc + 3, bc + 2,
                                   ctx->run_start.as<int64_t>(), ctx->grp_run.as<int32_t>(),
                                   bc + 4);
                hip_check(hipGetLastError(), "k_two_runs");
                K.merge_pass(dim3(static_cast<unsigned>(ceil_div(SM, kBlock))), s,
                             M.as<int32_t>(), Rn.as<int32_t>(), SM, ctx->run_start.as<int64_t>(),
                             ctx->grp_run.as<int32_t>(), 1, 0, SM);
                hip_check(hipGetLastError(), "k_merge_pass");
                span(e1, mark(), &st->kernel_expand_td_s);
                // R takes the merged buffer; its old buffer is kept for the next batch
                R.pts.swap(Rn);
                R.cap = SM;
            }
            // the batch's second wait: its counts
            std::vector<int64_t> h;
            d2h(ctx, h, ctx->batch_ctr, 5, "batch counts");
            flush_late();
            const int64_t xtotal = h[1], btotal = h[2];
            r_count = r_prev == 0 ? btotal : h[4];
            st->filter_launches += dense + sorted_filters;
            st->filter_candidates += ytotal + xtotal;
            st->algorithmic_bytes += 4 * static_cast<int64_t>(d) * (batch_inputs + btotal);
            st->join_bytes += 4 * static_cast<int64_t>(d) * batch_inputs;
            if (r_prev > 0) {
                // R and B read once, R's flags read and written
                st->frontier_check_bytes += static_cast<int64_t>(d) * 4 * (r_prev + btotal) +
                                            2 * r_prev;
            }
            if (debug_join) {
                std::fprintf(stderr, " local=%lld batch_frontier=%lld R=%lld\n",
                             static_cast<long long>(xtotal), static_cast<long long>(btotal),
                             static_cast<long long>(r_count));
            }
            sample_memory();
        }
        st->join_bytes += 4 * static_cast<int64_t>(d) * r_count;
        fetch_points(R, 0, r_count, out);
        count_layer(0, r_count);
    }

    void finish() {
        hip_check(hipStreamSynchronize(s), "finish sync");
        flush_late();
        fold_steps();
        if (d_tests != nullptr) {
            unsigned long long t[4] = {0, 0, 0, 0};
            hip_check(hipMemcpyAsync(t, d_tests, sizeof(t), hipMemcpyDeviceToHost, s),
                      "hipMemcpy(counter)");
            hip_check(hipStreamSynchronize(s), "counter sync");
            st->pair_tests_filter = static_cast<int64_t>(t[0]);
            st->pair_tests_frontier_check = static_cast<int64_t>(t[1]);
            st->pair_tests_cross = static_cast<int64_t>(t[2]);
            st->pair_tests_state_dominance = static_cast<int64_t>(t[3]);
            st->dominance_pair_tests = static_cast<int64_t>(t[0] + t[1] + t[2] + t[3]);
        }
        sample_memory();
    }
};

void write_err(char *err, size_t errlen, const std::string &msg) {
    if (err != nullptr && errlen > 0) {
        std::snprintf(err, errlen, "%s", msg.c_str());
    }
}

void reset_stats(pmdd_stats *st) {
    double *a = st->std_candidates_per_layer;
    double *b = st->std_frontier_survivors_per_layer;
    const int32_t cap = st->std_capacity;
    std::memset(st, 0, sizeof(*st));
    st->std_candidates_per_layer = a;
    st->std_frontier_survivors_per_layer = b;
    st->std_capacity = cap;
    if (cap > 0) {
        if (a != nullptr) {
            std::fill(a, a + cap, 0.0);
        }
        if (b != nullptr) {
            std::fill(b, b + cap, 0.0);
        }
    }
}

// Where a no-allocation call wants its frontier (host memory of the caller).
struct UserOut {
    int32_t *points;
    int64_t capacity;
    int64_t *count;
};

int run_enumeration(pmdd_ctx *ctx, pmdd_resident *dd, const pmdd_opts *opts, pmdd_frontier *out,
                    pmdd_stats *stats, const UserOut *user = nullptr) {
    if (opts == nullptr || out == nullptr) {
        fail(PMDD_E_INVALID, "NULL options or output");
    }
    if (opts->method < 1 || opts->method > 3) {
        fail(PMDD_E_UNSUPPORTED,
             "GPU backend is unsupported for method " + std::to_string(opts->method) + ".");
    }
    if (dd->device != ctx->device) {
        fail(PMDD_E_INVALID, "diagram was uploaded through a context of another device");
    }
    pmdd_stats local{};
    pmdd_stats *st = stats != nullptr ? stats : &local;
    reset_stats(st);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    out->d = dd->d;
    out->count = 0;
    out->points = nullptr;
    // buffers allocated from here on are charged to the context
    AcctScope scope(&ctx->acct);
    if (!ctx->in_call) {
        ctx->acct.reset_peak();
        ctx->acct.begin_call();
    }
    if (t_arena == nullptr && !ctx->deferred.empty()) {
        hip_check(hipStreamSynchronize(ctx->stream), "deferred free sync");
        ctx->free_deferred();
    }
    {
        Run run(ctx, dd, *opts, st);
        if (user != nullptr) {
            run.user_pts = user->points;
            run.user_cap = user->capacity;
            run.user_count = user->count;
        }
        if (opts->method == 1) {
            run.topdown(out);
        } else if (opts->method == 2) {
            run.bottomup(out);
        } else {
            run.coupled(out);
        }
        run.finish();
    }
    // the diagram counts too when it was uploaded outside this call
    const int64_t diagram = dd->arena.acct == &ctx->acct
                                ? 0
                                : static_cast<int64_t>(dd->arena.bytes);
    st->ctx_mem_peak_bytes = ctx->acct.peak.load(std::memory_order_relaxed) + diagram;
    st->mem_need_peak_bytes = ctx->acct.need_peak + diagram;
    return PMDD_OK;
}

template <class F> int guarded(char *err, size_t errlen, F &&f) {
    try {
        const int rc = f();
        if (rc == PMDD_OK) {
            write_err(err, errlen, "");
        }
        return rc;
    } catch (const EngineError &e) {
        write_err(err, errlen, e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        write_err(err, errlen, "host allocation failed");
        return PMDD_E_NOMEM;
    } catch (const std::exception &e) {
        write_err(err, errlen, e.what());
        return PMDD_E_INVALID;
    }
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int pmdd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

// Device address of this device's scratch array, resolved once per device
// with that device current.
static char *scratch_address(int device) {
    static std::mutex mu;
    static std::unordered_map<int, char *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        return it->second;
    }
    void *ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (ptr == nullptr) {
        fail(PMDD_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    }
    addr[device] = static_cast<char *>(ptr);
    return static_cast<char *>(ptr);
}

int pmdd_ctx_create(int device, pmdd_ctx **out, char *err, size_t errlen) {
    return guarded(err, errlen, [&]() -> int {
        if (out == nullptr) {
            fail(PMDD_E_INVALID, "pmdd_ctx_create: NULL output");
        }
        *out = nullptr;
        int n = 0;
        const hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(PMDD_E_NODEVICE, std::string("hipGetDeviceCount failed: ") + hipGetErrorString(e));
        }
        if (n <= 0) {
            fail(PMDD_E_NODEVICE, "No HIP device found");
        }
        if (device < 0 || device >= n) {
            fail(PMDD_E_NODEVICE, "HIP device " + std::to_string(device) + " not present");
        }
        std::unique_ptr<pmdd_ctx> c(new pmdd_ctx());
        c->device = device;
        hip_check(hipSetDevice(device), "hipSetDevice");
        c->scratch = scratch_address(device);
        const TuneTable &tt = tune_table();
        std::copy(tt.t, tt.t + 9, c->tuning);
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev_a), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev_b), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev_c), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev_d), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev_e), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev_f), "hipEventCreate");
        hip_check(hipEventCreate(&c->ev_g), "hipEventCreate");
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&c->h_rb), sizeof(pmdd_k::PlanReadback),
                                hipHostMallocMapped | hipHostMallocCoherent),
                  "hipHostMalloc(plan readback)");
        std::memset(c->h_rb, 0, sizeof(pmdd_k::PlanReadback));
        hip_check(hipHostGetDevicePointer(reinterpret_cast<void **>(&c->d_rb_host), c->h_rb, 0),
                  "hipHostGetDevicePointer(plan readback)");
        *out = c.release();
        return PMDD_OK;
    });
}

void pmdd_ctx_destroy(pmdd_ctx *ctx) {
    if (ctx == nullptr) {
        return;
    }
    (void)hipSetDevice(ctx->device);
    for (hipStream_t q : {ctx->side, ctx->stream}) {
        if (q != nullptr) {
            (void)hipStreamSynchronize(q);
        }
    }
    ctx->trim(); // release scratch before the stream
    if (ctx->ev_a != nullptr) {
        (void)hipEventDestroy(ctx->ev_a);
    }
    if (ctx->ev_b != nullptr) {
        (void)hipEventDestroy(ctx->ev_b);
    }
    if (ctx->ev_c != nullptr) {
        (void)hipEventDestroy(ctx->ev_c);
    }
    for (hipEvent_t e : {ctx->ev_d, ctx->ev_e, ctx->ev_f, ctx->ev_g, ctx->ev_fork, ctx->ev_join}) {
        if (e != nullptr) {
            (void)hipEventDestroy(e);
        }
    }
    for (hipEvent_t e : ctx->ev_pool) {
        (void)hipEventDestroy(e);
    }
    for (hipEvent_t e : ctx->ev_late) {
        (void)hipEventDestroy(e);
    }
    for (hipStream_t q : {ctx->side, ctx->stream}) {
        if (q != nullptr) {
            (void)hipStreamDestroy(q);
        }
    }
    if (ctx->h_rb != nullptr) {
        (void)hipHostFree(ctx->h_rb);
    }
    delete ctx;
}

int pmdd_ctx_get_tuning(pmdd_ctx *ctx, int d, pmdd_tuning *out) {
    if (ctx == nullptr || out == nullptr || d < 1 || d > 8) {
        return PMDD_E_INVALID;
    }
    *out = ctx->tuning[d];
    return PMDD_OK;
}

int pmdd_ctx_set_tuning(pmdd_ctx *ctx, int d, const pmdd_tuning *t) {
    if (ctx == nullptr || t == nullptr || d < 1 || d > 8 || !tuning_valid(*t)) {
        return PMDD_E_INVALID;
    }
    ctx->tuning[d] = *t;
    ctx->tuning[d].source = 2;
    return PMDD_OK;
}

int pmdd_upload(pmdd_ctx *ctx, const pmdd_dd_view *view, pmdd_resident **out, char *err,
                size_t errlen) {
    return guarded(err, errlen, [&]() -> int {
        if (ctx == nullptr || out == nullptr) {
            fail(PMDD_E_INVALID, "pmdd_upload: NULL argument");
        }
        *out = nullptr;
        *out = upload(ctx, view);
        return PMDD_OK;
    });
}

void pmdd_resident_free(pmdd_resident *dd) { delete dd; }

int pmdd_enumerate_resident(pmdd_ctx *ctx, pmdd_resident *dd, const pmdd_opts *opts,
                            pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen) {
    return guarded(err, errlen, [&]() -> int {
        if (ctx == nullptr || dd == nullptr) {
            fail(PMDD_E_INVALID, "pmdd_enumerate_resident: NULL argument");
        }
        return run_enumeration(ctx, dd, opts, out, stats);
    });
}

int pmdd_enumerate_resident_noalloc(pmdd_ctx *ctx, pmdd_resident *dd, const pmdd_opts *opts,
                                    int32_t *points, int64_t capacity, int64_t *count,
                                    pmdd_stats *stats, char *err, size_t errlen) {
    return guarded(err, errlen, [&]() -> int {
        if (ctx == nullptr || dd == nullptr || opts == nullptr || count == nullptr ||
            capacity < 0 || (capacity > 0 && points == nullptr)) {
            fail(PMDD_E_INVALID, "pmdd_enumerate_resident_noalloc: NULL or negative argument");
        }
        *count = 0;
        if (ctx->scratch == nullptr) {
            fail(PMDD_E_HIP, "no scratch array address for this device");
        }
        // the part of the need fixed by the diagram: arc run offsets and the
        // node offsets of the layer frontiers of the widest layer
        int64_t fixed = 0;
        for (int l = 0; l < dd->L; ++l) {
            const int64_t e = std::max(dd->td[l].n_arcs, dd->bu[l].n_arcs);
            fixed = std::max(fixed, 8 * e + 5 * 8 * (dd->layer_size[l] + 1));
        }
        if (fixed > static_cast<int64_t>(kScratchBytes)) {
            fail(PMDD_E_UNSUPPORTED,
                 "the diagram's tables alone need " + std::to_string(fixed) +
                     " bytes of scratch, more than the " + std::to_string(kScratchBytes >> 20) +
                     " MiB the no-allocation entry point has; use pmdd_enumerate_resident");
        }
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        ScratchArena arena;
        arena.base = ctx->scratch;
        arena.cap = kScratchBytes;
        arena.deferred = &ctx->deferred;
        // for the call: buffers come from the scratch array, adaptive
        // coupling batches stay at kServeJoinCap products; afterwards the
        // context forgets every scratch block, so no later call of this
        // context touches the array
        struct Scope {
            pmdd_ctx *c;
            ScratchArena *prev;
            int64_t join_cap;
            Scope(pmdd_ctx *x, ScratchArena *a) : c(x), prev(t_arena), join_cap(x->join_cap) {
                t_arena = a;
                c->join_cap = kServeJoinCap;
            }
            ~Scope() {
                (void)hipStreamSynchronize(c->stream);
                c->for_each_buffer([](DevBuf &b) { b.drop_scratch(); });
                c->join_cap = join_cap;
                t_arena = prev;
            }
        } scope(ctx, &arena);
        pmdd_frontier fr{};
        const UserOut user{points, capacity, count};
        return run_enumeration(ctx, dd, opts, &fr, stats, &user);
    });
}

int pmdd_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *view, const pmdd_opts *opts,
                   pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen) {
    return guarded(err, errlen, [&]() -> int {
        if (ctx == nullptr) {
            fail(PMDD_E_INVALID, "pmdd_enumerate: NULL context");
        }
        // the diagram's device arena is charged to the context for this call
        AcctScope scope(&ctx->acct);
        ctx->acct.reset_peak();
        ctx->acct.begin_call();
        struct InCall {
            pmdd_ctx *c;
            explicit InCall(pmdd_ctx *x) : c(x) { c->in_call = true; }
            ~InCall() { c->in_call = false; }
        } in_call(ctx);
        const double t0 = now_s();
        std::unique_ptr<pmdd_resident> r(upload(ctx, view));
        const double pack_s = now_s() - t0;
        const int rc = run_enumeration(ctx, r.get(), opts, out, stats);
        if (stats != nullptr) {
            stats->wall_pack_transfer_s = pack_s;
        }
        return rc;
    });
}

static int named_call(pmdd_ctx *ctx, const pmdd_dd_view *view, int kind, int method,
                      int problem_type, int sd, int64_t max_cand, int64_t max_prod,
                      pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen) {
    if (view != nullptr && view->kind != kind) {
        write_err(err, errlen, kind == PMDD_DD_BDD ? "expected a BDD view" : "expected an MDD view");
        return PMDD_E_INVALID;
    }
    pmdd_opts o{};
    o.method = method;
    o.problem_type = problem_type;
    o.state_dominance = sd;
    // the reference times its kernels whenever it is given stats
    // (topdown.cu:428-507 passes a timer only when stats != NULL)
    o.phase_timing = stats != nullptr ? 1 : 0;
    o.max_cand = max_cand;
    o.max_prod = max_prod;
    return pmdd_enumerate(ctx, view, &o, out, stats, err, errlen);
}

int pmdd_topdown_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *bdd, int problem_type,
                           int state_dominance, int64_t max_cand, pmdd_frontier *out,
                           pmdd_stats *stats, char *err, size_t errlen) {
    return named_call(ctx, bdd, PMDD_DD_BDD, 1, problem_type, state_dominance, max_cand,
                      PMDD_DEFAULT_MAX_PROD, out, stats, err, errlen);
}

int pmdd_topdown_mdd_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *mdd, int64_t max_cand,
                               pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen) {
    return named_call(ctx, mdd, PMDD_DD_MDD, 1, 3, 0, max_cand, PMDD_DEFAULT_MAX_PROD, out, stats,
                      err, errlen);
}

int pmdd_coupled_mdd_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *mdd, int64_t max_cand,
                               int64_t max_prod, pmdd_frontier *out, pmdd_stats *stats, char *err,
                               size_t errlen) {
    return named_call(ctx, mdd, PMDD_DD_MDD, 3, 3, 0, max_cand, max_prod, out, stats, err, errlen);
}

int pmdd_coupled_bdd_enumerate(pmdd_ctx *ctx, const pmdd_dd_view *bdd, int problem_type,
                               int state_dominance, int64_t max_cand, int64_t max_prod,
                               pmdd_frontier *out, pmdd_stats *stats, char *err, size_t errlen) {
    return named_call(ctx, bdd, PMDD_DD_BDD, 3, problem_type, state_dominance, max_cand, max_prod,
                      out, stats, err, errlen);
}

void pmdd_frontier_free(pmdd_frontier *f) {
    if (f != nullptr) {
        std::free(f->points);
        f->points = nullptr;
        f->count = 0;
    }
}

const char *pmdd_engine_version(void) { return "pmdd-engine gfx950 d=1..8"; }

} // extern "C"

// ---------------------------------------------------------------------------
// Serving: worker slots, each a context with its own stream; FIFO admission
// under a memory budget.
// ---------------------------------------------------------------------------
namespace {

struct ServerJob {
    int64_t ticket = 0;
    std::vector<int64_t> layer_off, in_off;
    std::vector<int32_t> arc_src, arc_val, arc_w, min_weight, sd_partner;
    pmdd_dd_view view{};
    pmdd_opts opts{};
    int64_t estimate = 0;
    bool exclusive = false; // runs alone (queued again after PMDD_E_NOMEM)
    int retries = 0;
    double submit_t = 0.0;
    double admit_t = 0.0;
    bool done = false;
    int rc = PMDD_OK;
    std::string err;
    pmdd_frontier fr{};
    pmdd_stats st{};
};

// Device bytes a request may claim, from its layer widths and budgets: its
// diagram (both CSR directions); the pass buffers of one expansion (survivor
// points with their flags, and the two merge buffers of big nodes) and three
// layer frontiers, each of at most min(max_cand, 256 * widest layer) points
// of d int32; for method 3, a coupling batch's two product buffers and its
// survivor, sort and merge buffers, at most five times the batch limit the
// engine uses (max_prod, or kServeJoinCap products when adaptive: a server
// slot caps the adaptive batches of its requests there); 16 MiB of tables.
// The frontier term is a bound on the usual case only: frontier sizes are
// not known before the run (the measured peak is ctx_mem_peak_bytes).
int64_t estimate_bytes(const pmdd_dd_view &v, const pmdd_opts &o) {
    const int L = v.num_layers;
    const int64_t N = v.layer_off[L];
    const int64_t A = v.in_off[N];
    const int64_t d = v.d;
    int64_t widest = 1;
    for (int l = 0; l < L; ++l) {
        widest = std::max<int64_t>(widest, v.layer_off[l + 1] - v.layer_off[l]);
    }
    const int64_t arena = 2 * (8 * (N + L) + 4 * A + 4 * d * A);
    const int64_t cap = o.max_cand > 0 ? o.max_cand : PMDD_DEFAULT_MAX_CAND;
    const int64_t pts = std::min<int64_t>(cap, 256 * widest);
    const int64_t pass = pts * (4 * d + 1) + 2 * pts * 4 * d;
    const int64_t fronts = 3 * pts * 4 * d;
    int64_t join = 0;
    if (o.method == 3) {
        const int64_t limit = o.max_prod > 0 ? o.max_prod : kServeJoinCap;
        join = 5 * limit * 4 * d;
    }
    return arena + pass + fronts + join + (int64_t(16) << 20);
}

// ---------------------------------------------------------------------------
// Packed serving: one packer per objective count expands the current layers
// of its in-flight requests as one layer per step (Run::expand on the packed
// layer, its small-node chunks filtered by the persistent queue kernels in a
// replayed graph).  Request i's current layer occupies nodes [base_i,
// base_i + N_i) of the packed frontier; its next layer's arcs are gathered
// into the packed layer with their parents moved by base_i.  Nodes of
// requests that have left stay in the packed frontier until the next step
// drops them (no arc refers to them).
// ---------------------------------------------------------------------------
// Packed requests: diagrams of at most kPackWidest nodes per layer and
// kPackNodes in all.  On the serving stream the wider ones (knapsacks with
// four or five objectives and n of about 40 or more, the large set-packing
// diagrams) hold almost all of the work (for the first 100 requests, the 42
// above 8192 nodes per layer take 99% of the CPU oracle's time): they run
// best next to each other in the slots, while packing pays for the many
// light requests whose layers are mostly launch and planning overhead.
constexpr int64_t kPackWidest = 8192;
constexpr int64_t kPackNodes = int64_t(1) << 20;
constexpr int kPackInflight = 32;

struct PackedReq {
    std::shared_ptr<ServerJob> job;
    pmdd_resident *dd = nullptr; // its diagram, uploaded when it joined
    int layer = 0;               // its layer now in the packed frontier
    int64_t base = 0;            // first node of that layer there
};

struct Packer {
    pmdd_ctx *ctx = nullptr;
    int d = 0;
    pmdd_stats st{}; // the packer's own totals
    std::unique_ptr<Run> run;
    std::unique_ptr<FrontierLease> la, lb;
    DevFrontier *cur = nullptr, *nxt = nullptr;
    std::vector<PackedReq> live;
    LayerCSR layer; // the packed layer (host tables for planning)
    DevBuf d_off, d_src, d_w, d_parts;
    std::vector<pmdd_k::PackPart> parts;
    // packed layers expanded, and requests they carried (read by the
    // server's other threads)
    std::atomic<int64_t> steps{0}, stepped{0};

    Packer(pmdd_ctx *c, int d_) : ctx(c), d(d_) { reset(); }
    ~Packer() { drop(); }

    void drop() {
        for (PackedReq &r : live) {
            delete r.dd;
        }
        live.clear();
        run.reset();
        la.reset();
        lb.reset();
    }
    // Empty packed frontier, no request.
    void reset() {
        drop();
        pmdd_opts o{};
        o.method = 1;
        o.max_cand = PMDD_DEFAULT_MAX_CAND;
        o.max_prod = PMDD_MAX_PROD_ADAPTIVE;
        la.reset(new FrontierLease(ctx, d));
        lb.reset(new FrontierLease(ctx, d));
        cur = &la->f;
        nxt = &lb->f;
        cur->n_nodes = 0;
        cur->h_off.assign(1, 0);
        run.reset(new Run(ctx, nullptr, d, o, &st));
        run->queued = true;
    }

    // Uploads a request's diagram and puts its root layer (frontier {0^d} at
    // the root) at the end of the packed frontier; upload_offsets follows.
    void add(const std::shared_ptr<ServerJob> &job) {
        PackedReq r;
        r.job = job;
        r.dd = upload(ctx, &job->view);
        r.base = cur->n_nodes;
        const int64_t n0 = r.dd->layer_size[0];
        const int64_t total = cur->total();
        for (int64_t v = 0; v < n0; ++v) {
            cur->h_off.push_back(total + (v >= r.dd->root ? 1 : 0));
        }
        cur->n_nodes += n0;
        grow_points(*cur, d, total + 1, total, ctx->stream);
        for (int o = 0; o < d; ++o) {
            hip_check(hipMemsetAsync(cur->pts.as<int32_t>() + o * cur->cap + total, 0, 4,
                                     ctx->stream),
                      "hipMemset(seed)");
        }
        live.push_back(r);
    }
    void upload_offsets_now() { upload_offsets(ctx, *cur); }

    // One packed layer: every live request's next layer.  Requests that
    // reach their terminal layer get their frontier and move to `done`.
    void step(std::vector<PackedReq> &done) {
        int64_t N = 0, E = 0;
        layer.off.assign(1, 0);
        layer.src.clear();
        parts.clear();
        for (const PackedReq &r : live) {
            const LayerCSR &c = r.dd->td[r.layer + 1];
            const int64_t n = r.dd->layer_size[r.layer + 1];
            for (int64_t v = 1; v <= n; ++v) {
                layer.off.push_back(E + c.off[v]);
            }
            for (int64_t e = 0; e < c.n_arcs; ++e) {
                layer.src.push_back(static_cast<int32_t>(c.src[e] + r.base));
            }
            parts.push_back({c.d_src, c.d_w, c.n_arcs, E, r.base});
            N += n;
            E += c.n_arcs;
        }
        layer.n_arcs = E;
        h2d(ctx, d_off, layer.off, "hipMemcpy(packed offsets)");
        h2d(ctx, d_parts, parts, "hipMemcpy(packed parts)");
        d_src.ensure(static_cast<size_t>(std::max<int64_t>(E, 1)) * 4, "hipMalloc(packed arcs)");
        d_w.ensure(static_cast<size_t>(std::max<int64_t>(E, 1)) * 4 * d,
                   "hipMalloc(packed weights)");
        if (E > 0) {
            hipLaunchKernelGGL(pmdd_k::k_pack_arcs,
                               dim3(static_cast<unsigned>(std::min<int64_t>(ceil_div(E, kBlock), 4096))),
                               dim3(kBlock), 0, ctx->stream, d_parts.as<pmdd_k::PackPart>(),
                               static_cast<int32_t>(parts.size()), E, d, d_src.as<int32_t>(),
                               d_w.as<int32_t>());
            hip_check(hipGetLastError(), "k_pack_arcs");
        }
        layer.d_off = d_off.as<int64_t>();
        layer.d_src = d_src.as<int32_t>();
        layer.d_w = d_w.as<int32_t>();
        int64_t cand = 0;
        run->expand(*cur, layer, N, *nxt, &cand, nullptr);
        steps += 1;
        stepped += static_cast<int64_t>(live.size());
        int64_t base = 0;
        std::vector<PackedReq> keep;
        for (PackedReq &r : live) {
            r.layer += 1;
            r.base = base;
            base += r.dd->layer_size[r.layer];
            if (r.layer == r.dd->L - 1) {
                const int64_t t = r.base + r.dd->terminal;
                run->fetch_points(*nxt, nxt->h_off[t], nxt->size(t), &r.job->fr);
                done.push_back(r);
            } else {
                keep.push_back(r);
            }
        }
        live.swap(keep);
        std::swap(cur, nxt);
    }
};

// Whether a request goes to a packer: method 1 without state dominance or
// per-request diagnostics, on a diagram small enough to share packed layers
// with others.
bool packable(const pmdd_dd_view &v, const pmdd_opts &o) {
    if (o.method != 1 || o.state_dominance != 0 || o.count_pair_tests != 0 ||
        o.phase_timing != 0) {
        return false;
    }
    int64_t widest = 0;
    for (int l = 0; l < v.num_layers; ++l) {
        widest = std::max<int64_t>(widest, v.layer_off[l + 1] - v.layer_off[l]);
    }
    return widest <= kPackWidest && v.layer_off[v.num_layers] <= kPackNodes;
}

} // namespace

struct pmdd_server {
    int device = 0;
    int slots = 1;
    int64_t budget = 0;
    int64_t next_ticket = 1;
    int running = 0;
    bool stopping = false;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<ServerJob>> queue;
    std::unordered_map<int64_t, std::shared_ptr<ServerJob>> jobs;
    std::deque<int64_t> finished; // done, in completion order (may hold collected tickets)
    std::vector<pmdd_ctx *> ctxs;
    std::vector<std::thread> workers;
    // per slot: bytes reserved (the running request's estimate, or the
    // buffers the idle context keeps), busy, asked to release its buffers
    std::vector<int64_t> reserve;
    std::vector<char> busy, trim_asked;
    int64_t blocked_ticket = 0; // head of the queue last seen waiting for memory
    pmdd_server_info info{};
    // packed serving: a packer (its own context, stream and thread) per
    // objective count, created on first use, with its queue of requests;
    // the estimates of the requests the packers hold are reserved
    bool pack = false;
    int pack_inflight = kPackInflight;
    std::unique_ptr<Packer> packers[9];
    pmdd_ctx *pack_ctx[9] = {};
    std::deque<std::shared_ptr<ServerJob>> pack_queue[9];
    std::vector<std::thread> pack_threads;
    int64_t pack_reserved = 0;

    int64_t reserved_except(int slot) const {
        int64_t r = pack_reserved;
        for (int s = 0; s < slots; ++s) {
            r += s == slot ? 0 : reserve[s];
        }
        return r;
    }

    // Starts the packer of objective count d (lock held).
    void start_packer(int d) {
        if (packers[d] != nullptr) {
            return;
        }
        pmdd_ctx *c = nullptr;
        char e[512] = {0};
        const int rc = pmdd_ctx_create(device, &c, e, sizeof(e));
        if (rc != PMDD_OK) {
            fail(rc, e);
        }
        c->use_side = false; // runs beside the slots
        pack_ctx[d] = c;
        {
            AcctScope scope(&c->acct);
            packers[d].reset(new Packer(c, d));
        }
        pack_threads.emplace_back([this, d] { packer_loop(d); });
    }

    void finish_job(const std::shared_ptr<ServerJob> &job) {
        job->st.mem_estimate_bytes = job->estimate;
        job->st.nomem_retries = job->retries;
        job->st.queue_wait_s = job->admit_t - job->submit_t;
        job->done = true;
        finished.push_back(job->ticket);
        cv_done.notify_all();
    }

    // Packer thread: admits queued requests of its objective count while
    // it has room and the budget allows, expands one packed layer, hands
    // back the requests that finished; repeats while any is in flight.
    void packer_loop(int d) {
        Packer &P = *packers[d];
        pmdd_ctx *ctx = pack_ctx[d];
        AcctScope scope(&ctx->acct);
        (void)hipSetDevice(ctx->device);
        std::unique_lock<std::mutex> lk(mu);
        std::deque<std::shared_ptr<ServerJob>> &q = pack_queue[d];
        while (true) {
            if (stopping) {
                return;
            }
            std::vector<std::shared_ptr<ServerJob>> admit;
            while (!q.empty() && static_cast<int>(P.live.size() + admit.size()) < pack_inflight) {
                const std::shared_ptr<ServerJob> &head = q.front();
                if (running > 0 && reserved_except(-1) + head->estimate > budget) {
                    if (blocked_ticket != head->ticket) {
                        blocked_ticket = head->ticket;
                        info.memory_waits += 1;
                    }
                    ask_trims(-1);
                    cv_work.notify_all();
                    break;
                }
                admit.push_back(head);
                q.pop_front();
                pack_reserved += head->estimate;
                running += 1;
                head->admit_t = now_s();
                info.admitted += 1;
                info.packed += 1;
                info.running_peak = std::max(info.running_peak, running);
                info.reserved_peak = std::max(info.reserved_peak, reserved_except(-1));
            }
            if (P.live.empty() && admit.empty()) {
                cv_work.wait(lk);
                continue;
            }
            lk.unlock();
            std::vector<PackedReq> done;
            std::vector<std::shared_ptr<ServerJob>> failed, requeue;
            int frc = PMDD_OK;
            std::string ferr;
            for (const std::shared_ptr<ServerJob> &job : admit) {
                try {
                    P.add(job);
                } catch (const EngineError &e) {
                    job->rc = e.code;
                    job->err = e.msg;
                    failed.push_back(job);
                } catch (const std::bad_alloc &) {
                    job->rc = PMDD_E_NOMEM;
                    job->err = "host allocation failed";
                    failed.push_back(job);
                }
            }
            try {
                P.upload_offsets_now();
                if (!P.live.empty()) {
                    P.step(done);
                }
            } catch (const EngineError &e) {
                frc = e.code;
                ferr = e.msg;
            } catch (const std::bad_alloc &) {
                frc = PMDD_E_NOMEM;
                ferr = "host allocation failed";
            }
            if (frc != PMDD_OK) {
                // the packed layer failed: its requests run again in the
                // slots, unpacked, and the packer starts empty
                (void)hipStreamSynchronize(ctx->stream);
                for (PackedReq &r : P.live) {
                    pmdd_frontier_free(&r.job->fr);
                    requeue.push_back(r.job);
                }
                for (PackedReq &r : done) {
                    pmdd_frontier_free(&r.job->fr);
                    requeue.push_back(r.job);
                    delete r.dd;
                }
                done.clear();
                try {
                    P.reset();
                } catch (const EngineError &) {
                }
            }
            for (PackedReq &r : done) {
                delete r.dd;
                r.dd = nullptr;
            }
            lk.lock();
            int64_t steps = 0, stepped = 0;
            for (const std::unique_ptr<Packer> &pk : packers) {
                if (pk != nullptr) {
                    steps += pk->steps.load(std::memory_order_relaxed);
                    stepped += pk->stepped.load(std::memory_order_relaxed);
                }
            }
            info.pack_steps = steps;
            info.pack_requests_per_step_x100 = steps > 0 ? 100 * stepped / steps : 0;
            for (const std::shared_ptr<ServerJob> &job : failed) {
                pack_reserved -= job->estimate;
                running -= 1;
                finish_job(job);
            }
            for (PackedReq &r : done) {
                pack_reserved -= r.job->estimate;
                running -= 1;
                r.job->rc = PMDD_OK;
                finish_job(r.job);
            }
            for (const std::shared_ptr<ServerJob> &job : requeue) {
                pack_reserved -= job->estimate;
                running -= 1;
                job->retries += 1;
                queue.push_front(job);
            }
            cv_work.notify_all();
        }
    }

    // Whether idle `slot` may start the head of the queue now: its estimate
    // (or the buffers this slot keeps, when larger) fits the budget next to
    // every other slot's reservation.  A request queued again after running
    // out of memory waits until no other slot runs or keeps buffers.
    bool admits(int slot, const ServerJob &head) const {
        const int64_t others = reserved_except(slot);
        if (head.exclusive) {
            return running == 0 && others == 0;
        }
        return others + std::max(head.estimate, reserve[slot]) <= budget;
    }

    // Asks every idle slot that keeps device buffers to release them (the
    // head of the queue is waiting for memory).
    void ask_trims(int except) {
        for (int s = 0; s < slots; ++s) {
            if (s != except && !busy[s] && reserve[s] > 0 && !trim_asked[s]) {
                trim_asked[s] = 1;
            }
        }
    }

    // Releases slot's kept buffers; called with the lock held, drops it
    // while the device calls run.
    void trim_slot(int slot, std::unique_lock<std::mutex> &lk) {
        pmdd_ctx *ctx = ctxs[slot];
        lk.unlock();
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->trim();
        lk.lock();
        reserve[slot] = ctx->acct.cur.load(std::memory_order_relaxed);
        info.trims += 1;
    }

    void worker(int slot) {
        pmdd_ctx *ctx = ctxs[slot];
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            if (stopping) {
                return;
            }
            if (trim_asked[slot]) {
                trim_asked[slot] = 0;
                trim_slot(slot, lk);
                cv_work.notify_all();
                continue;
            }
            if (queue.empty()) {
                cv_work.wait(lk);
                continue;
            }
            std::shared_ptr<ServerJob> job = queue.front();
            if (!admits(slot, *job)) {
                if (blocked_ticket != job->ticket) {
                    blocked_ticket = job->ticket;
                    info.memory_waits += 1;
                }
                if (reserve[slot] > 0 && (job->exclusive || reserve[slot] > job->estimate)) {
                    trim_slot(slot, lk); // this slot's own kept buffers are in the way
                }
                ask_trims(slot);
                cv_work.notify_all();
                if (!admits(slot, *job)) {
                    cv_work.wait(lk);
                }
                continue;
            }
            queue.pop_front();
            busy[slot] = 1;
            running += 1;
            reserve[slot] = std::max(job->estimate, reserve[slot]);
            info.admitted += 1;
            info.running_peak = std::max(info.running_peak, running);
            info.reserved_peak = std::max(info.reserved_peak, reserved_except(-1));
            job->admit_t = now_s();
            lk.unlock();
            cv_work.notify_all();
            char err[1024] = {0};
            job->st.std_capacity = 0;
            ctx->join_cap = kServeJoinCap;
            job->rc = pmdd_enumerate(ctx, &job->view, &job->opts, &job->fr, &job->st, err,
                                     sizeof(err));
            job->err = err;
            lk.lock();
            busy[slot] = 0;
            running -= 1;
            reserve[slot] = ctx->acct.cur.load(std::memory_order_relaxed);
            if (job->rc == PMDD_E_NOMEM && !job->exclusive && !stopping) {
                // out of device memory next to other requests: once more, alone
                pmdd_frontier_free(&job->fr);
                job->exclusive = true;
                job->retries += 1;
                info.nomem_retries += 1;
                queue.push_front(job);
                cv_work.notify_all();
                continue;
            }
            finish_job(job);
            cv_work.notify_all();
        }
    }
};

extern "C" {

int pmdd_server_create(const pmdd_server_config *cfg, pmdd_server **out, char *err,
                       size_t errlen) {
    return guarded(err, errlen, [&]() -> int {
        if (cfg == nullptr || out == nullptr || cfg->slots < 1) {
            fail(PMDD_E_INVALID, "pmdd_server_create: bad configuration");
        }
        *out = nullptr;
        std::unique_ptr<pmdd_server> srv(new pmdd_server());
        srv->device = cfg->device;
        srv->slots = cfg->slots;
        for (int i = 0; i < cfg->slots; ++i) {
            pmdd_ctx *c = nullptr;
            char e[512] = {0};
            const int rc = pmdd_ctx_create(cfg->device, &c, e, sizeof(e));
            if (rc != PMDD_OK) {
                for (pmdd_ctx *x : srv->ctxs) {
                    pmdd_ctx_destroy(x);
                }
                fail(rc, e);
            }
            c->use_side = cfg->slots == 1;
            srv->ctxs.push_back(c);
        }
        if (cfg->memory_budget > 0) {
            srv->budget = cfg->memory_budget;
        } else {
            size_t fr = 0, tot = 0;
            hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
            srv->budget = static_cast<int64_t>(fr) / 5 * 4;
        }
        srv->pack = cfg->pack != 0;
        srv->pack_inflight = cfg->pack_inflight > 0 ? cfg->pack_inflight : kPackInflight;
        srv->reserve.assign(cfg->slots, 0);
        srv->busy.assign(cfg->slots, 0);
        srv->trim_asked.assign(cfg->slots, 0);
        srv->info.budget = srv->budget;
        srv->info.slots = cfg->slots;
        pmdd_server *raw = srv.get();
        for (int i = 0; i < cfg->slots; ++i) {
            srv->workers.emplace_back([raw, i] { raw->worker(i); });
        }
        *out = srv.release();
        return PMDD_OK;
    });
}

int pmdd_server_submit(pmdd_server *srv, const pmdd_dd_view *view, const pmdd_opts *opts,
                       int64_t *ticket, char *err, size_t errlen) {
    return guarded(err, errlen, [&]() -> int {
        if (srv == nullptr || view == nullptr || opts == nullptr || ticket == nullptr) {
            fail(PMDD_E_INVALID, "pmdd_server_submit: NULL argument");
        }
        validate_view(view);
        auto job = std::make_shared<ServerJob>();
        job->submit_t = now_s();
        const int L = view->num_layers;
        const int64_t N = view->layer_off[L];
        const int64_t A = view->in_off[N];
        job->layer_off.assign(view->layer_off, view->layer_off + L + 1);
        job->in_off.assign(view->in_off, view->in_off + N + 1);
        job->arc_src.assign(view->arc_src, view->arc_src + A);
        job->arc_val.assign(view->arc_val, view->arc_val + A);
        job->arc_w.assign(view->arc_w, view->arc_w + A * view->d);
        if (view->min_weight != nullptr) {
            job->min_weight.assign(view->min_weight, view->min_weight + N);
        }
        if (view->sd_partner != nullptr) {
            job->sd_partner.assign(view->sd_partner, view->sd_partner + N);
        }
        job->view = *view;
        job->view.layer_off = job->layer_off.data();
        job->view.in_off = job->in_off.data();
        job->view.arc_src = job->arc_src.data();
        job->view.arc_val = job->arc_val.data();
        job->view.arc_w = job->arc_w.data();
        job->view.min_weight = view->min_weight != nullptr ? job->min_weight.data() : nullptr;
        job->view.sd_partner = view->sd_partner != nullptr ? job->sd_partner.data() : nullptr;
        job->opts = *opts;
        job->estimate = estimate_bytes(*view, *opts);
        std::lock_guard<std::mutex> lk(srv->mu);
        job->ticket = srv->next_ticket++;
        *ticket = job->ticket;
        srv->jobs[job->ticket] = job;
        if (job->estimate > srv->budget) {
            job->done = true;
            job->rc = PMDD_E_REJECTED;
            job->st.mem_estimate_bytes = job->estimate;
            job->err = "request needs about " + std::to_string(job->estimate) +
                       " bytes of device memory, more than the server budget of " +
                       std::to_string(srv->budget);
            srv->info.rejected += 1;
            srv->finished.push_back(job->ticket);
            srv->cv_done.notify_all();
            return PMDD_OK;
        }
        if (srv->pack && packable(*view, *opts)) {
            srv->start_packer(view->d);
            srv->pack_queue[view->d].push_back(job);
        } else {
            srv->queue.push_back(job);
        }
        srv->cv_work.notify_all();
        return PMDD_OK;
    });
}

// Moves a finished job's frontier and statistics to the caller.
static int server_take(const std::shared_ptr<ServerJob> &job, pmdd_frontier *out,
                       pmdd_stats *stats, char *err, size_t errlen) {
    if (out != nullptr) {
        *out = job->fr;
    } else {
        pmdd_frontier_free(&job->fr);
    }
    job->fr.points = nullptr;
    if (stats != nullptr) {
        double *a = stats->std_candidates_per_layer;
        double *b = stats->std_frontier_survivors_per_layer;
        const int32_t cap = stats->std_capacity;
        *stats = job->st;
        stats->std_candidates_per_layer = a;
        stats->std_frontier_survivors_per_layer = b;
        stats->std_capacity = cap;
    }
    write_err(err, errlen, job->err);
    return job->rc;
}

static int server_collect(pmdd_server *srv, int64_t ticket, bool block, pmdd_frontier *out,
                          pmdd_stats *stats, char *err, size_t errlen) {
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(ticket);
    if (it == srv->jobs.end()) {
        write_err(err, errlen, "unknown or already collected ticket");
        return PMDD_E_INVALID;
    }
    std::shared_ptr<ServerJob> job = it->second;
    if (block) {
        srv->cv_done.wait(lk, [&] { return job->done; });
    } else if (!job->done) {
        return PMDD_PENDING;
    }
    srv->jobs.erase(ticket);
    lk.unlock();
    srv->cv_done.notify_all(); // a wait_any may be waiting for the last job
    return server_take(job, out, stats, err, errlen);
}

int pmdd_server_wait_any(pmdd_server *srv, int64_t *ticket, pmdd_frontier *out,
                         pmdd_stats *stats, char *err, size_t errlen) {
    if (srv == nullptr || ticket == nullptr) {
        write_err(err, errlen, "pmdd_server_wait_any: NULL argument");
        return PMDD_E_INVALID;
    }
    std::unique_lock<std::mutex> lk(srv->mu);
    std::shared_ptr<ServerJob> job;
    while (job == nullptr) {
        srv->cv_done.wait(lk, [&] { return !srv->finished.empty() || srv->jobs.empty(); });
        if (srv->finished.empty()) {
            write_err(err, errlen, "no outstanding request");
            return PMDD_E_INVALID;
        }
        const int64_t t = srv->finished.front();
        srv->finished.pop_front();
        auto it = srv->jobs.find(t);
        if (it != srv->jobs.end()) { // not collected by a wait or poll already
            job = it->second;
            srv->jobs.erase(it);
        }
    }
    lk.unlock();
    *ticket = job->ticket;
    return server_take(job, out, stats, err, errlen);
}

int pmdd_server_poll(pmdd_server *srv, int64_t ticket, pmdd_frontier *out, pmdd_stats *stats,
                     char *err, size_t errlen) {
    if (srv == nullptr) {
        write_err(err, errlen, "pmdd_server_poll: NULL server");
        return PMDD_E_INVALID;
    }
    return server_collect(srv, ticket, false, out, stats, err, errlen);
}

int pmdd_server_wait(pmdd_server *srv, int64_t ticket, pmdd_frontier *out, pmdd_stats *stats,
                     char *err, size_t errlen) {
    if (srv == nullptr) {
        write_err(err, errlen, "pmdd_server_wait: NULL server");
        return PMDD_E_INVALID;
    }
    return server_collect(srv, ticket, true, out, stats, err, errlen);
}

int pmdd_server_get_info(pmdd_server *srv, pmdd_server_info *out) {
    if (srv == nullptr || out == nullptr) {
        return PMDD_E_INVALID;
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->info;
    return PMDD_OK;
}

void pmdd_server_destroy(pmdd_server *srv) {
    if (srv == nullptr) {
        return;
    }
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
        srv->queue.clear();
    }
    srv->cv_work.notify_all();
    for (std::thread &t : srv->workers) {
        t.join();
    }
    for (std::thread &t : srv->pack_threads) {
        t.join();
    }
    for (auto &kv : srv->jobs) {
        pmdd_frontier_free(&kv.second->fr);
    }
    for (int dd = 0; dd < 9; ++dd) {
        if (srv->packers[dd] != nullptr) {
            (void)hipSetDevice(srv->pack_ctx[dd]->device);
            (void)hipStreamSynchronize(srv->pack_ctx[dd]->stream);
            srv->packers[dd].reset();
            pmdd_ctx_destroy(srv->pack_ctx[dd]);
        }
    }
    for (pmdd_ctx *c : srv->ctxs) {
        pmdd_ctx_destroy(c);
    }
    delete srv;
}

void pmdd_debug_fail_alloc(int64_t nth) {
    g_fail_alloc_in.store(nth > 0 ? nth : 0, std::memory_order_relaxed);
}

} // extern "C"
