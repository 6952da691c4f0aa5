// Device planning of a layer expansion.
//
// The chunk filters of kernels.hpp work from tables: a chunk descriptor per
// chunk (a node of at most CH candidates is one chunk, a bigger node is cut
// into CH-candidate blocks of its merged, key-ordered parent runs), a list of
// chunks per filter tier, and for the big nodes their parent-run tables.  These
// kernels build those tables on the device from the parent frontier's device
// offsets and the layer's arc CSR, so planning a layer needs no host loop over
// its arcs and nodes:
//
//   k_plan_count  one thread per node: the node-local offset of each arc's
//                 candidate run (run_off), the node's candidate count (gamma),
//                 and per workgroup the sums of the node's plan values and the
//                 largest big node (chunks, arcs); the last workgroup to finish
//                 turns the sums into exclusive prefixes and writes the layer
//                 totals into the plan header (the one thing the host reads
//                 back: it sizes the buffers and picks the launches)
//   k_plan_write  one thread per node: its prefix of every plan value, then
//                 its chunk descriptors, tier-list entries and big-node tables
//
// After the filters, the survivors are placed without the host as well:
//   k_block_sums / k_scan_write  exclusive prefix of the per-chunk
//                 survivor counts (the chunks' output slots)
//   k_node_offsets  the new frontier's node offsets from the chunk slots
//
// The tables are the ones the host planner builds (the same chunk order,
// lists and run tables), so every filter launch is unchanged.
#pragma once

#include "kernels.hpp"

#include <cstddef>

namespace pmdd_k {

// Plan values of a node, summed over the layer as exclusive prefixes.
enum PlanVal {
    PV_CAND = 0, // candidates (gamma)
    PV_CHUNKS,   // chunk descriptors
    PV_TINY,     // 1 if one chunk of at most small_cap candidates
    PV_MID,      // 1 if one chunk of at most mid_cap (when mid_cap > 0)
    PV_SMALL,    // 1 if one chunk otherwise
    PV_BIGN,     // 1 if more than CH candidates (a big node)
    PV_BIGSZ,    // candidates of a big node
    PV_RUNS,     // arcs into a big node (its parent runs)
    kPlanVals
};
// Per workgroup of k_plan_count: the kPlanVals sums, then the two maxima.
constexpr int kPlanPart = kPlanVals + 2;

// What the host reads back after planning a layer.
struct PlanHeader {
    int64_t tot[kPlanVals]; // layer totals of the plan values
    int64_t max_chunks;     // most chunks of one big node (0: none)
    int64_t max_runs;       // most arcs into one big node (0: none)
};

// Per-expansion figures accumulated on the device and read once at the end
// of an enumeration.
struct StepStat {
    int64_t surv;      // survivors of the expansion (before state dominance)
    int64_t dropped;   // labels state dominance removed
    double ssd_gamma;  // sum over nodes of (candidates - mean)^2
    double ssd_size;   // sum over nodes of (|F(v)| - mean)^2, after state dominance
};

// The device block the host reads back in one copy: the plan header of each
// sweep direction's next layer and each direction's expansion score.
struct PlanReadback {
    PlanHeader hdr[2];
    unsigned long long score[2];
    unsigned long long seq; // the host copy's sequence number (k_publish)
};

// k_publish: one thread copies the readback block into host memory (mapped,
// coherent) and then, after a system-scope release, its sequence number;
// the host waits for the number instead of copying and synchronizing.
__global__ void k_publish(const PlanReadback *__restrict__ src, PlanReadback *dst,
                          unsigned long long seq) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const unsigned long long *a = reinterpret_cast<const unsigned long long *>(src);
        unsigned long long *b = reinterpret_cast<unsigned long long *>(dst);
        constexpr int kWords = static_cast<int>(offsetof(PlanReadback, seq) / 8);
        for (int i = 0; i < kWords; ++i) {
            b[i] = a[i];
        }
        __hip_atomic_store(&dst->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

struct PlanShape {
    int32_t chunk_cap; // CH
    int32_t small_cap; // tiny tier bound
    int32_t mid_cap;   // mid tier bound (0: no mid tier)
    int32_t big_min;   // nodes of more candidates take the big-node path (<= CH)
};

// Plan values of one node from its candidate count and in-degree.
__device__ __forceinline__ void plan_values(int64_t g, int64_t deg, const PlanShape &sh,
                                            int64_t (&pv)[kPlanVals]) {
#pragma unroll
    for (int k = 0; k < kPlanVals; ++k) {
        pv[k] = 0;
    }
    pv[PV_CAND] = g;
    if (g <= 0) {
        return;
    }
    if (g <= sh.big_min) {
        pv[PV_CHUNKS] = 1;
        if (g <= sh.small_cap) {
            pv[PV_TINY] = 1;
        } else if (sh.mid_cap > 0 && g <= sh.mid_cap) {
            pv[PV_MID] = 1;
        } else {
            pv[PV_SMALL] = 1;
        }
    } else {
        pv[PV_CHUNKS] = (g + sh.chunk_cap - 1) / sh.chunk_cap;
        pv[PV_BIGN] = 1;
        pv[PV_BIGSZ] = g;
        pv[PV_RUNS] = deg;
    }
}

// Wave-wide inclusive prefix sum (64 lanes).
__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t u = __shfl_up(v, off);
        if (lane >= off) {
            v += u;
        }
    }
    return v;
}

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
    }
    return v;
}

__device__ __forceinline__ int64_t wave_max(int64_t v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int64_t u = __shfl_xor(v, off);
        v = u > v ? u : v;
    }
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
    }
    return v;
}

// Exclusive prefix of one int64 per thread over a kBlock-thread workgroup;
// *total gets the workgroup sum.  Opens with a barrier (safe after any other
// use of lds) and ends with one (lds may be reused at once).  All threads
// call it.
__device__ __forceinline__ int64_t block_exscan64(int64_t v, int64_t *lds, int64_t *total) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int64_t incl = wave_incl_scan(v);
    __syncthreads();
    if (lane == kWave - 1) {
        lds[wave] = incl;
    }
    __syncthreads();
    int64_t before = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const int64_t t = lds[w];
        before += (w < wave) ? t : 0;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return before + incl - v;
}

// Workgroup sum of a double (one atomic add per workgroup into *acc).
__device__ __forceinline__ void block_add_f64(double v, double *lds, double *acc) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    v = wave_sum_f64(v);
    __syncthreads();
    if (lane == 0) {
        lds[wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kWaves; ++w) {
            s += lds[w];
        }
        atomicAdd(acc, s);
    }
}

// Called by every thread of every workgroup once the workgroup has written
// its partial results: true in the workgroup that arrives last, whose reads
// after the call see every workgroup's partials (the arrival counter is an
// agent-scope atomic with acquire-release order: each workgroup's writes are
// released to the device before its arrival, the last one acquires them).
// The last workgroup puts the counter back to zero for the next launch.  No
// workgroup waits for another.
__device__ __forceinline__ bool last_to_arrive(unsigned *counter) {
    __shared__ bool last;
    __syncthreads(); // the workgroup's partials are written before its arrival
    if (threadIdx.x == 0) {
        const unsigned prev =
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = prev == gridDim.x - 1;
        if (last) {
            __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    return last;
}

// Exclusive prefix, in place, of nb rows of kPlanPart values (the first
// kPlanVals summed, the last two maxima) into the plan header's totals and
// maxima; one workgroup.
__device__ __forceinline__ void plan_scan_rows(int64_t nb, int64_t *part, PlanHeader *hdr) {
    __shared__ int64_t lds[kWaves];
    __shared__ int64_t mx[2][kWaves];
    int64_t carry[kPlanVals];
#pragma unroll
    for (int k = 0; k < kPlanVals; ++k) {
        carry[k] = 0;
    }
    int64_t mx_chunks = 0, mx_runs = 0;
    for (int64_t base = 0; base < nb; base += kBlock) {
        const int64_t b = base + threadIdx.x;
        int64_t *row = part + b * kPlanPart;
        const bool ok = b < nb;
#pragma unroll
        for (int k = 0; k < kPlanVals; ++k) {
            const int64_t x = ok ? row[k] : 0;
            int64_t tot = 0;
            const int64_t ex = block_exscan64(x, lds, &tot);
            if (ok) {
                row[k] = carry[k] + ex;
            }
            carry[k] += tot;
        }
        if (ok) {
            mx_chunks = row[kPlanVals] > mx_chunks ? row[kPlanVals] : mx_chunks;
            mx_runs = row[kPlanVals + 1] > mx_runs ? row[kPlanVals + 1] : mx_runs;
        }
    }
    mx_chunks = wave_max(mx_chunks);
    mx_runs = wave_max(mx_runs);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    if (lane == 0) {
        mx[0][wave] = mx_chunks;
        mx[1][wave] = mx_runs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kPlanVals; ++k) {
            hdr->tot[k] = carry[k];
        }
        int64_t a = 0, r = 0;
        for (int w = 0; w < kWaves; ++w) {
            a = mx[0][w] > a ? mx[0][w] : a;
            r = mx[1][w] > r ? mx[1][w] : r;
        }
        hdr->max_chunks = a;
        hdr->max_runs = r;
    }
}

// ---------------------------------------------------------------------------
// k_plan_count: grid ceil(N / kBlock).  run_off[e] for every arc of the
// layer, gamma[v] for every node, part[b * kPlanPart ..] per workgroup.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_plan_count(
    int64_t n_nodes, const int64_t *__restrict__ arc_off, const int32_t *__restrict__ arc_src,
    const int64_t *__restrict__ prev_off, int64_t *__restrict__ run_off,
    int64_t *__restrict__ gamma, int64_t *__restrict__ part, PlanShape sh,
    unsigned *__restrict__ counter, PlanHeader *__restrict__ hdr) {
    __shared__ int64_t red[kWaves][kPlanPart];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    int64_t pv[kPlanVals];
    int64_t mx_chunks = 0, mx_runs = 0;
    if (v < n_nodes) {
        const int64_t e0 = arc_off[v];
        const int64_t e1 = arc_off[v + 1];
        int64_t acc = 0;
        for (int64_t e = e0; e < e1; ++e) {
            run_off[e] = acc;
            const int32_t p = arc_src[e];
            acc += prev_off[p + 1] - prev_off[p];
        }
        gamma[v] = acc;
        plan_values(acc, e1 - e0, sh, pv);
        if (pv[PV_BIGN] != 0) {
            mx_chunks = pv[PV_CHUNKS];
            mx_runs = pv[PV_RUNS];
        }
    } else {
        plan_values(0, 0, sh, pv);
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < kPlanVals; ++k) {
        const int64_t s = wave_sum(pv[k]);
        if (lane == 0) {
            red[wave][k] = s;
        }
    }
    mx_chunks = wave_max(mx_chunks);
    mx_runs = wave_max(mx_runs);
    if (lane == 0) {
        red[wave][kPlanVals] = mx_chunks;
        red[wave][kPlanVals + 1] = mx_runs;
    }
    __syncthreads();
    if (threadIdx.x < kPlanPart) {
        const int k = threadIdx.x;
        int64_t r = 0;
        for (int w = 0; w < kWaves; ++w) {
            const int64_t x = red[w][k];
            r = (k < kPlanVals) ? r + x : (x > r ? x : r);
        }
        part[static_cast<int64_t>(blockIdx.x) * kPlanPart + k] = r;
    }
    // the last workgroup turns the sums into prefixes and writes the header
    if (last_to_arrive(counter)) {
        plan_scan_rows(gridDim.x, part, hdr);
    }
}

// Tables k_plan_write fills (sized from the plan header).
struct PlanTables {
    ChunkDesc *chunks;   // [tot CHUNKS]
    int32_t *node_chunk; // [N + 1] first chunk of each node; [N] = tot CHUNKS
    int32_t *tiny;       // [tot TINY] chunk ids per tier
    int32_t *mid;        // [tot MID]
    int32_t *small;      // [tot SMALL]
    int32_t *big_chunks; // chunks of the big nodes, in chunk order
    int32_t *big_node;   // [tot BIGN]
    int64_t *big_base;   // [tot BIGN + 1] first slot of each big node in the run buffer
    int32_t *grp_run;    // [tot BIGN + 1] first run of each big node
    int64_t *run_start;  // [tot RUNS + 1] first slot of each run
    double *ssd_gamma;   // null, or += sum over nodes of (gamma - mean)^2
};

// ---------------------------------------------------------------------------
// k_plan_write: grid ceil(N / kBlock), after k_plan_count.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_plan_write(
    int64_t n_nodes, const int64_t *__restrict__ arc_off, const int64_t *__restrict__ run_off,
    const int64_t *__restrict__ gamma, const int64_t *__restrict__ part,
    const PlanHeader *__restrict__ hdr, PlanTables t, PlanShape sh) {
    __shared__ int64_t lds[kWaves];
    __shared__ double flds[kWaves];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool ok = v < n_nodes;
    const int64_t g = ok ? gamma[v] : 0;
    const int64_t e0 = ok ? arc_off[v] : 0;
    const int64_t deg = ok ? arc_off[v + 1] - e0 : 0;
    int64_t pv[kPlanVals];
    plan_values(g, deg, sh, pv);
    const int64_t *bp = part + static_cast<int64_t>(blockIdx.x) * kPlanPart;
    int64_t at[kPlanVals];
#pragma unroll
    for (int k = 0; k < kPlanVals; ++k) {
        int64_t tot = 0;
        at[k] = bp[k] + block_exscan64(pv[k], lds, &tot);
    }
    if (ok) {
        const int32_t first = static_cast<int32_t>(at[PV_CHUNKS]);
        t.node_chunk[v] = first;
        // a node in a tier (one chunk, filtered by one workgroup); a big node
        // may also have a single chunk (a narrow layer's)
        if (pv[PV_BIGN] == 0 && pv[PV_CHUNKS] == 1) {
            ChunkDesc c;
            c.start = 0;
            c.out_base = at[PV_CAND];
            c.group = static_cast<int32_t>(v);
            c.len = static_cast<int32_t>(g);
            c.group_first = first;
            c.group_count = 1;
            t.chunks[first] = c;
            if (pv[PV_TINY] != 0) {
                t.tiny[at[PV_TINY]] = first;
            } else if (pv[PV_MID] != 0) {
                t.mid[at[PV_MID]] = first;
            } else {
                t.small[at[PV_SMALL]] = first;
            }
        } else if (pv[PV_BIGN] != 0) {
            const int64_t bi = at[PV_BIGN];
            const int64_t bpos = at[PV_BIGSZ];
            t.big_node[bi] = static_cast<int32_t>(v);
            t.big_base[bi] = bpos;
            t.grp_run[bi] = static_cast<int32_t>(at[PV_RUNS]);
            for (int64_t j = 0; j < deg; ++j) {
                t.run_start[at[PV_RUNS] + j] = bpos + run_off[e0 + j];
            }
            const int32_t cnt = static_cast<int32_t>(pv[PV_CHUNKS]);
            // the big nodes' chunks follow each other in chunk order
            const int64_t list0 = at[PV_CHUNKS] - (at[PV_TINY] + at[PV_MID] + at[PV_SMALL]);
            for (int32_t k = 0; k < cnt; ++k) {
                ChunkDesc c;
                c.start = bpos + static_cast<int64_t>(k) * sh.chunk_cap;
                c.out_base = at[PV_CAND] + static_cast<int64_t>(k) * sh.chunk_cap;
                c.group = static_cast<int32_t>(v);
                const int64_t rest = g - static_cast<int64_t>(k) * sh.chunk_cap;
                c.len = static_cast<int32_t>(rest < sh.chunk_cap ? rest : sh.chunk_cap);
                c.group_first = first;
                c.group_count = cnt;
                t.chunks[first + k] = c;
                t.big_chunks[list0 + k] = first + k;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        t.node_chunk[n_nodes] = static_cast<int32_t>(hdr->tot[PV_CHUNKS]);
        t.big_base[hdr->tot[PV_BIGN]] = hdr->tot[PV_BIGSZ];
        t.grp_run[hdr->tot[PV_BIGN]] = static_cast<int32_t>(hdr->tot[PV_RUNS]);
        t.run_start[hdr->tot[PV_RUNS]] = hdr->tot[PV_BIGSZ];
    }
    if (t.ssd_gamma != nullptr) {
        const double mean = static_cast<double>(hdr->tot[PV_CAND]) / static_cast<double>(n_nodes);
        const double dv = ok ? static_cast<double>(g) - mean : 0.0;
        block_add_f64(dv * dv, flds, t.ssd_gamma);
    }
}

// ---------------------------------------------------------------------------
// Exclusive prefix of n counts (int32 or int64) into int64 slots: out[i] =
// base + sum of cnt[0 .. i), out[n] = base + sum of all.  k_block_sums (on
// ceil(n / kBlock) workgroups, at least one) writes the workgroup sums, and
// its last workgroup to finish turns them into exclusive prefixes, writes
// out[n] and, when total_out is set, *total_out = the sum; k_scan_write (on
// ceil(n / kBlock) workgroups, none when n is 0) writes out[0 .. n).
template <class C>
__global__ __launch_bounds__(kBlock) void k_block_sums(int64_t n, const C *__restrict__ cnt,
                                                       int64_t *__restrict__ part,
                                                       unsigned *__restrict__ counter,
                                                       int64_t base, int64_t *__restrict__ out,
                                                       int64_t *__restrict__ total_out) {
    __shared__ int64_t lds[kWaves];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    int64_t tot = 0;
    block_exscan64(i < n ? cnt[i] : 0, lds, &tot);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tot;
    }
    if (!last_to_arrive(counter)) {
        return;
    }
    const int64_t nb = gridDim.x;
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kBlock) {
        const int64_t b = b0 + threadIdx.x;
        int64_t t = 0;
        const int64_t ex = block_exscan64(b < nb ? part[b] : 0, lds, &t);
        if (b < nb) {
            part[b] = carry + ex;
        }
        carry += t;
    }
    if (threadIdx.x == 0) {
        out[n] = base + carry;
        if (total_out != nullptr) {
            *total_out = carry;
        }
    }
}

template <class C>
__global__ __launch_bounds__(kBlock) void k_scan_write(int64_t n, const C *__restrict__ cnt,
                                                       const int64_t *__restrict__ part,
                                                       int64_t base, int64_t *__restrict__ out) {
    __shared__ int64_t lds[kWaves];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    int64_t tot = 0;
    const int64_t ex = block_exscan64(i < n ? cnt[i] : 0, lds, &tot);
    if (i < n) {
        out[i] = base + part[blockIdx.x] + ex;
    }
}

// ---------------------------------------------------------------------------
// k_node_offsets: grid ceil((N + 1) / kBlock).  off[v] = slot[node_chunk[v]]
// for v in 0..N (node_chunk[N] = chunk count, slot[chunk count] = the end).
// With ssd set, adds the sum over nodes of (|F(v)| - mean)^2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_node_offsets(int64_t n_nodes,
                                                         const int32_t *__restrict__ node_chunk,
                                                         const int64_t *__restrict__ slot,
                                                         int64_t *__restrict__ off,
                                                         double *__restrict__ ssd) {
    __shared__ double flds[kWaves];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (v <= n_nodes) {
        off[v] = slot[node_chunk[v]];
    }
    if (ssd != nullptr) {
        const double total =
            static_cast<double>(slot[node_chunk[n_nodes]] - slot[node_chunk[0]]);
        const double mean = total / static_cast<double>(n_nodes);
        double dv = 0.0;
        if (v < n_nodes) {
            dv = static_cast<double>(slot[node_chunk[v + 1]] - slot[node_chunk[v]]) - mean;
        }
        block_add_f64(dv * dv, flds, ssd);
    }
}

// ---------------------------------------------------------------------------
// k_layer_score: *score += sum over nodes of |F(v)| * (deg_off[v+1] -
// deg_off[v]) (method 3's expansion score; *score zeroed before).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_layer_score(int64_t n_nodes,
                                                        const int64_t *__restrict__ off,
                                                        const int64_t *__restrict__ deg_off,
                                                        unsigned long long *__restrict__ score) {
    __shared__ int64_t lds[kWaves];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    int64_t x = 0;
    if (v < n_nodes) {
        x = (off[v + 1] - off[v]) * (deg_off[v + 1] - deg_off[v]);
    }
    int64_t tot = 0;
    block_exscan64(x, lds, &tot);
    if (threadIdx.x == 0 && tot != 0) {
        atomicAdd(score, static_cast<unsigned long long>(tot));
    }
}

// ---------------------------------------------------------------------------
// State dominance planned on the device: the comparison targets of a layer,
// one per 256 labels of a node that has comparison nodes, as the host rule
// builds them.  Knapsack: the nodes in descending min_weight order that hold
// labels, each against the next two of them, skipping a pair of
// single-parent nodes whose parent is the same and whose arcs differ.  Set
// packing: every node that holds labels against its partner, when the
// partner holds labels.  Nodes are taken through an order (null: index
// order); the ones holding labels are compacted first (k_sd_flags, a prefix
// of the flags, k_sd_active), then counted (k_sd_count), the counts prefixed
// and the targets written (k_sd_targets).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_sd_flags(int64_t n, const int32_t *__restrict__ order,
                                                     const int64_t *__restrict__ off,
                                                     int32_t *__restrict__ flag) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) {
        const int64_t v = order != nullptr ? order[i] : i;
        flag[i] = off[v + 1] > off[v] ? 1 : 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_sd_active(int64_t n, const int32_t *__restrict__ order,
                                                      const int32_t *__restrict__ flag,
                                                      const int64_t *__restrict__ apos,
                                                      int32_t *__restrict__ active) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n && flag[i] != 0) {
        active[apos[i]] = order != nullptr ? order[i] : static_cast<int32_t>(i);
    }
}

// Per active index a (a < n; active count apos_end[0]): the comparison nodes
// (cmp[2a], cmp[2a + 1], -1 for none) and the number of 256-label targets.
__global__ __launch_bounds__(kBlock) void k_sd_count(
    int64_t n, const int64_t *__restrict__ n_active, const int32_t *__restrict__ active,
    const int64_t *__restrict__ off, const int32_t *__restrict__ single_parent,
    const int32_t *__restrict__ single_arc, const int32_t *__restrict__ partner,
    int32_t *__restrict__ cmp, int32_t *__restrict__ cnt) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (a >= n) {
        return;
    }
    const int64_t na = *n_active;
    if (a >= na) {
        cnt[a] = 0;
        return;
    }
    const int32_t n1 = active[a];
    int32_t c[2] = {-1, -1};
    int k = 0;
    if (partner != nullptr) {
        const int32_t n2 = partner[n1];
        if (n2 >= 0 && off[n2 + 1] > off[n2]) {
            c[k++] = n2;
        }
    } else {
        const int64_t end = a + 3 < na ? a + 3 : na;
        for (int64_t j = a + 1; j < end; ++j) {
            const int32_t n2 = active[j];
            const int32_t p1 = single_parent[n1], p2 = single_parent[n2];
            if (p1 >= 0 && p2 >= 0 && p1 == p2 && single_arc[n1] != single_arc[n2]) {
                continue;
            }
            c[k++] = n2;
        }
    }
    const int64_t size = off[n1 + 1] - off[n1];
    cmp[2 * a] = c[0];
    cmp[2 * a + 1] = c[1];
    cnt[a] = k > 0 ? static_cast<int32_t>((size + kBlock - 1) / kBlock) : 0;
}

__global__ __launch_bounds__(kBlock) void k_sd_targets(int64_t n, const int32_t *__restrict__ cnt,
                                                       const int64_t *__restrict__ tpos,
                                                       const int32_t *__restrict__ active,
                                                       const int32_t *__restrict__ cmp,
                                                       int32_t *__restrict__ targets) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (a >= n) {
        return;
    }
    const int32_t k = cnt[a];
    for (int32_t b = 0; b < k; ++b) {
        int32_t *t = targets + 4 * (tpos[a] + b);
        t[0] = active[a];
        t[1] = cmp[2 * a];
        t[2] = cmp[2 * a + 1];
        t[3] = b * kBlock;
    }
}

// *acc += *old_end - *new_end (labels a compaction removed).
__global__ void k_sd_dropped(const int64_t *__restrict__ old_end,
                             const int64_t *__restrict__ new_end, int64_t *__restrict__ acc) {
    if (threadIdx.x == 0) {
        *acc += *old_end - *new_end;
    }
}

// *ssd += sum over nodes of (|F(v)| - mean)^2, from node offsets.
__global__ __launch_bounds__(kBlock) void k_size_ssd(int64_t n_nodes,
                                                     const int64_t *__restrict__ off,
                                                     double *__restrict__ ssd) {
    __shared__ double flds[kWaves];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const double mean = static_cast<double>(off[n_nodes] - off[0]) / static_cast<double>(n_nodes);
    double dv = 0.0;
    if (v < n_nodes) {
        dv = static_cast<double>(off[v + 1] - off[v]) - mean;
    }
    block_add_f64(dv * dv, flds, ssd);
}

// ---------------------------------------------------------------------------
// Coupling batches counted on the device.
// ---------------------------------------------------------------------------
// k_uniform_chunks: n_max chunk descriptors cutting *total points (dense,
// from 0) into CH-point blocks: chunk k < ceil(*total / CH) covers
// [k CH, min((k + 1) CH, *total)); the rest are empty (len 0 at the end, no
// survivors).  One group of all the non-empty chunks (group_count = their
// number) when `one_group`, else each chunk its own group.
__global__ __launch_bounds__(kBlock) void k_uniform_chunks(const int64_t *__restrict__ total,
                                                           int64_t n_max, int32_t CH,
                                                           bool one_group,
                                                           ChunkDesc *__restrict__ chunks,
                                                           int32_t *__restrict__ nsurv,
                                                           int32_t *__restrict__ nalive) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (k >= n_max) {
        return;
    }
    const int64_t t = *total;
    const int64_t cnt = (t + CH - 1) / CH;
    ChunkDesc c;
    c.start = k < cnt ? k * CH : t;
    c.out_base = c.start;
    c.group = 0;
    const int64_t rest = t - c.start;
    c.len = static_cast<int32_t>(k < cnt ? (rest < CH ? rest : CH) : 0);
    c.group_first = one_group ? 0 : static_cast<int32_t>(k);
    c.group_count = one_group ? static_cast<int32_t>(cnt) : 1;
    chunks[k] = c;
    nsurv[k] = 0;
    nalive[k] = 0;
}

// *dst = (int64) *src (a device count of another type).
__global__ void k_count_to_i64(const unsigned long long *__restrict__ src,
                               int64_t *__restrict__ dst) {
    if (threadIdx.x == 0) {
        *dst = static_cast<int64_t>(*src);
    }
}

// The run table of two dense runs, [0, *a) and [*a, *a + *b), one group:
// run_start = {0, a, a + b}, grp_run = {0, 2}; *sum = a + b.
__global__ void k_two_runs(const int64_t *__restrict__ a, const int64_t *__restrict__ b,
                           int64_t *__restrict__ run_start, int32_t *__restrict__ grp_run,
                           int64_t *__restrict__ sum) {
    if (threadIdx.x == 0) {
        run_start[0] = 0;
        run_start[1] = *a;
        run_start[2] = *a + *b;
        grp_run[0] = 0;
        grp_run[1] = 2;
        *sum = *a + *b;
    }
}

// dst[o * dst_stride + *at + i] = src[o * src_stride + i] for i < *n, every
// objective o < d (grid-stride).
__global__ __launch_bounds__(kBlock) void k_copy_at(const int32_t *__restrict__ src,
                                                    int64_t src_stride,
                                                    const int64_t *__restrict__ n, int d,
                                                    int32_t *__restrict__ dst, int64_t dst_stride,
                                                    const int64_t *__restrict__ at) {
    const int64_t cnt = *n;
    const int64_t base = *at;
    const int64_t step = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < cnt; i += step) {
        for (int o = 0; o < d; ++o) {
            dst[o * dst_stride + base + i] = src[o * src_stride + i];
        }
    }
}

} // namespace pmdd_k
