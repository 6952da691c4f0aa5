#!/usr/bin/env python3
"""Benchmark: exact Pareto-frontier enumeration on one MI355X.

Workload (one "step"): the single-large-knapsack configuration -- a
3-objective knapsack BDD with n=80 items and a 5-objective one with n=50
(weights and profits int32 uniform [1,1000], capacity = half the weight sum,
seeded SplitMix64), each enumerated with method 1 (top-down) and method 3
(dynamic layer cutset), state dominance 0 and 1: eight enumerations.  The
diagrams are built and uploaded to HBM before the timed region; every
enumeration in the timed region is complete (all layers, coupling, result
copied back to the host).

Prints one JSON line.  `value` = exact Pareto points per second over the
timed steps (sum of final frontier sizes / wall time).  The timed steps record
no per-layer event; as many steps again with phase timing (HIP events around
each kernel family's launches) give the per-family device times.  `roofline`
covers the kernel with the most device time; `cpu_baseline` times the CPU
restatement of the reference's OpenMP enumerators on a bounded sample on this
host.

`serving` replays a seeded prefix of the mixed serving stream (knapsack,
set packing and TSP requests of uneven size, d in {3,4,5}; --serving-requests,
2000 for the whole stream) through the plain one-at-a-time path (one slot)
and the batched path (--serving-slots worker slots with their own streams,
memory-based admission): first with every request queued at t=0 (capacity),
then with Poisson arrivals at 0.5x, 1x and 1.5x the plain path's capacity
(latency; --serving-rates), reporting
instances/s, Pareto points/s and p50/p99 latency from arrival to result
(host diagram construction included).  Every request's frontier must be the
same in the plain and the batched runs.

Profiling helpers (for rocprofv3 runs): --only NAME and --cases M:SD,... run a
subset of the step; --calib-copy-mib N runs one plain device copy of N MiB
(torch) and exits, to scale the FETCH_SIZE / WRITE_SIZE counters.

--paper runs the reference paper's benchmark instance families instead
(tests/golden/paper/, the instances of its experiments, 10 per family): each
instance one-shot (upload, enumeration, result copy) with method 3 and
method 1, printing one JSON line per instance and per family with the
paper's mean H100 times beside ours.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

T_START = time.perf_counter()
# coupling batches: adaptive (0), 32K products doubling to 32M
MAX_PROD = 0
HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, GB/s
VALU_PEAK_TOPS = 39.3           # 256 CU x 4 SIMD x 16 lanes x 2.4 GHz, int32 lane-ops/s (T)

WORKLOAD = [
    # (name, seed, n, d)
    ("kp_d3_n80", 80_3, 80, 3),
    ("kp_d5_n50", 50_5, 50, 5),
]
CASES = [(1, 0), (1, 1), (3, 0), (3, 1)]  # (method, state_dominance)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(tmpdir):
    from pmdd_amd import DecisionDiagram
    from pmdd_amd import instances as gen
    dds = []
    for name, seed, n, d in WORKLOAD:
        path = gen.write(gen.knapsack(seed, n, d), os.path.join(tmpdir, f"{name}.dat"))
        dds.append((name, path, DecisionDiagram.from_file(path, 1)))
    return dds


def read_traffic(kernel):
    """(HBM bytes per launch of `kernel`, source) from the committed rocprofv3
    PMC summary profiles/<kernel>_knapsack_traffic.txt, or (None, None).  The
    source names the file, the commit it was measured at, the traffic over the
    compulsory bytes and the counters' VALU issue fraction."""
    p = os.path.join(ROOT, "profiles", f"{kernel}_knapsack_traffic.txt")
    if not os.path.exists(p):
        return None, None
    vals = {}
    with open(p) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"]), {
            "file": f"profiles/{kernel}_knapsack_traffic.txt",
            "commit": vals.get("commit"),
            "traffic_over_compulsory": float(vals.get("traffic_over_compulsory", "nan")),
            # share of the VALU issue slots filled with any VALU instruction
            # (SQ_INSTS_VALU), beside the compare-only fraction of `frac`
            "valu_issue_frac_pmc": float(vals.get("valu_issue_frac", "nan"))}
    except (KeyError, ValueError):
        return None, None


def host_cores():
    """(physical cores, logical CPUs this process may run on) of this host."""
    logical = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (
        os.cpu_count() or 1)
    phys = None
    try:
        seen = set()
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":")[1].strip()
                elif not line.strip():
                    if cid is not None:
                        seen.add((pid, cid))
                    pid = cid = None
        phys = len(seen) or None
    except OSError:
        pass
    return phys, logical


# CPU sample: (instance, method) pairs of the step's workload, state dominance 0
CPU_SAMPLE = [("kp_d3_n80", 1), ("kp_d3_n80", 3), ("kp_d5_n50", 3), ("kp_d5_n50", 1)]


def cpu_baseline(dds, threads, budget_s):
    """Oracle (CPU restatement of the reference's OpenMP enumerators, methods
    1 and 3) on the step's two instances, state dominance 0, on `threads`
    host threads, in the order of CPU_SAMPLE, stopping before a pair once
    budget_s seconds are spent (the d=5 method-1 run is the longest)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    by_name = {name: dd for name, _, dd in dds}
    pts, spent, ran, skipped = 0, 0.0, [], []
    for name, method in CPU_SAMPLE:
        if name not in by_name:
            continue
        if ran and spent > budget_s:
            skipped.append(f"{name} m{method}")
            continue
        t0 = time.perf_counter()
        n = len(oracle.enumerate_dd(by_name[name], method, 0, threads=threads))
        dt = time.perf_counter() - t0
        pts += n
        spent += dt
        ran.append(f"{name} m{method}: {n} points in {dt:.2f} s")
        log(f"cpu baseline: {name} m{method} {dt:.2f} s")
    phys, logical = host_cores()
    return {
        "value": pts / spent,
        "unit": "Pareto pts/s",
        "cores": threads,
        "kind": "port",
        "sample": "oracle (OpenMP restatement of the reference's CPU enumerators), state "
                  "dominance 0: " + "; ".join(ran) +
                  (("; not run (time bound): " + ", ".join(skipped)) if skipped else ""),
        "host_physical_cores": phys,
        "host_logical_cpus": logical,
        "threads_note": "threads = the box's CPU share for one GPU (16) or fewer if the "
                        "process may use fewer CPUs",
    }


def serving_leg(tmpdir, requests, slots, method, rate_factors, deadline):
    """Plain (one slot) vs batched (`slots` slots) on the first `requests`
    requests of the serving stream: closed loop (all queued at t=0), then
    Poisson arrivals at each factor of the plain path's closed-loop capacity.
    The open-loop legs replay a prefix of the stream short enough to end
    before `deadline` (perf_counter seconds); the line states its length."""
    from pmdd_amd import serving
    specs = serving.mixed_stream(requests)
    paths = serving.write_stream(specs, os.path.join(tmpdir, "stream"))
    keys = ("requests", "inst_per_s", "pareto_pts_per_s", "p50_latency_s", "p99_latency_s",
            "mean_latency_s", "span_s", "failed", "mem_peak_over_estimate_max",
            "mem_peak_above_estimate", "admission")
    out = {"requests": requests, "method": method, "slots": slots,
           "mix": "60% knapsack n~U[20,70], 25% set packing n~U[60,150] m=n/5, "
                  "15% TSP n~U[8,11]; d~U{3,4,5}; seeded SplitMix64"}
    # the three paths: one request at a time; `slots` worker slots, each its
    # own context and stream; the same slots plus the packers (method-1
    # requests on small diagrams share packed layers)
    paths_cfg = [("plain", dict(slots=1)), ("batched", dict(slots=slots)),
                 ("packed", dict(slots=slots, pack=True))]
    t0 = time.perf_counter()
    closed = [0.0] * requests
    progress = 250 if requests > 250 else 0
    res, closed_out = {}, {}
    for name, kw in paths_cfg:
        res[name] = {}
        r = serving.run_stream(paths, specs, closed, method=method, results=res[name],
                               progress=progress, **kw)
        closed_out[name] = {k: r[k] for k in keys}
        log(f"serving: {name} closed loop {r['inst_per_s']:.2f} inst/s")
    same = all(i in res[p] and res[p][i].shape == res["plain"][i].shape and
               bool((res[p][i] == res["plain"][i]).all())
               for p in ("batched", "packed") for i in range(requests) if i in res["plain"])
    cap = closed_out["plain"]["inst_per_s"]
    out["closed_loop"] = closed_out
    sweep = []
    for f in rate_factors:
        lam = f * cap
        # the three paths replay n requests at rate lam: about 3 n / lam seconds
        left = deadline - time.perf_counter()
        n = min(requests, int(0.85 * left * lam / 3))
        if n < 50:
            log(f"serving: no time left for Poisson arrivals at {f}x")
            continue
        arrivals = serving.poisson_arrivals(n, lam)
        entry = {"rate_factor": f, "rate_per_s": lam, "requests": n}
        for name, kw in paths_cfg:
            r = serving.run_stream(paths[:n], specs[:n], arrivals, method=method,
                                   progress=progress, **kw)
            entry[name] = {k: r[k] for k in keys}
            log(f"serving: {name} open loop at {f}x capacity, {n} requests done")
        sweep.append(entry)
        if f == 1.0:
            out["poisson"] = {k: entry[k] for k in ("rate_per_s", "requests", "plain",
                                                    "batched", "packed")}
    out["poisson_sweep"] = sweep
    out["batched_over_plain_inst_per_s"] = (closed_out["batched"]["inst_per_s"] /
                                            max(cap, 1e-12))
    out["packed_over_plain_inst_per_s"] = closed_out["packed"]["inst_per_s"] / max(cap, 1e-12)
    out["bit_exact_batched_and_packed_vs_plain"] = same
    out["wall_s"] = time.perf_counter() - t0
    return out


# The reference paper's instance families (reports/aaai/aaai2026-article.tex,
# tables at lines 546-596 and 865-938): (family, problem type, state dominance,
# mean seconds on an H100 for PmddG (method 3) and PmddG-TD (method 1)).
PAPER_FAMILIES = [
    ("mokp_n40_k6", 1, 1, 2.0, 3.0),
    ("mokp_n40_k7", 1, 1, 4.0, 5.0),
    ("mokp_n50_k5", 1, 1, 4.0, 3.0),
    ("mokp_n70_k3", 1, 1, 2.0, 2.0),
    ("mospp_n150_k3", 2, 0, 5.0, 7.0),
    ("mospp_n150_k4", 2, 0, 6.0, 6.0),
    ("mospp_n150_k7", 2, 0, 17.0, 9.0),
    ("motsp_n15_k4", 3, 0, 3.0, 3.0),
    ("motsp_n15_k5", 3, 0, 48.0, 36.0),
    # the paper's hardest settings: method 1 solved 6/10 (4 out of memory)
    # and 0/10, method 3 10/10 and 6/10 of them
    ("mokp_n70_k6", 1, 1, 1010.0, 335.0),
    ("motsp_n15_k6", 3, 0, 369.0, None),
    ("motsp_n15_k7", 3, 0, 1828.0, None),
]


def paper_sets(guard_s, only, limit):
    """One-shot method 3 and method 1 on every instance of the paper's
    families; a family stops after an instance slower than guard_s."""
    from pmdd_amd import DecisionDiagram
    from pmdd_amd.engine import Engine
    eng = Engine(0)
    base = os.path.join(ROOT, "tests", "golden", "paper")
    warm = DecisionDiagram.from_file(os.path.join(base, "mokp_n70_k3",
                                                  sorted(os.listdir(os.path.join(
                                                      base, "mokp_n70_k3")))[0]), 1)
    eng.enumerate(warm, method=1)  # device and context set-up outside the timings
    for fam, pt, sd, paper_m3, paper_m1 in PAPER_FAMILIES:
        if only and fam not in only:
            continue
        files = sorted(os.listdir(os.path.join(base, fam)))
        if limit > 0:
            files = files[:limit]
        times = {1: [], 3: []}
        for f in files:
            t = time.perf_counter()
            dd = DecisionDiagram.from_file(os.path.join(base, fam, f), pt)
            build_s = time.perf_counter() - t
            row = {"family": fam, "instance": f, "build_s": build_s}
            pts = {}
            for method in (3, 1):
                t = time.perf_counter()
                out = eng.enumerate(dd, method=method, state_dominance=sd)
                dt = time.perf_counter() - t
                times[method].append(dt)
                pts[method] = out.points
                row[f"m{method}_s"] = dt
            row["points"] = len(pts[3])
            row["m1_points"] = len(pts[1])
            row["m1_equals_m3"] = bool(pts[1].shape == pts[3].shape and (pts[1] == pts[3]).all())
            print(json.dumps(row), flush=True)
            if max(row["m3_s"], row["m1_s"]) > guard_s:
                log(f"{fam}: {f} took more than {guard_s} s; rest of the family skipped")
                break
        mean = {m: sum(v) / len(v) for m, v in times.items() if v}
        print(json.dumps({"family": fam, "solved": len(times[3]), "of": len(files),
                          "mean_m3_s": mean.get(3), "mean_m1_s": mean.get(1),
                          "paper_h100_m3_s": paper_m3, "paper_h100_m1_s": paper_m1}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--skip-pair-count", action="store_true",
                    help="skip the extra run that counts executed pair tests")
    ap.add_argument("--only", default="", help="run only this workload instance")
    ap.add_argument("--cases", default="", help="method:state_dominance list, e.g. 3:0,1:0")
    ap.add_argument("--calib-copy-mib", type=int, default=0)
    ap.add_argument("--serving-requests", type=int, default=2000,
                    help="requests of the serving stream to replay (0: skip; 2000: all)")
    ap.add_argument("--time-budget-s", type=float, default=360.0,
                    help="the Poisson legs of the serving replay are cut to end by then")
    ap.add_argument("--cpu-budget-s", type=float, default=15.0,
                    help="the CPU baseline starts no new enumeration after this")
    ap.add_argument("--serving-slots", type=int, default=8)
    ap.add_argument("--serving-method", type=int, default=1)
    ap.add_argument("--serving-rates", default="1",
                    help="Poisson arrival rates as factors of the plain path's capacity")
    ap.add_argument("--paper", action="store_true",
                    help="run the reference paper's instance families instead of the step")
    ap.add_argument("--paper-guard-s", type=float, default=120.0)
    ap.add_argument("--paper-families", default="", help="comma list (default: all)")
    ap.add_argument("--paper-limit", type=int, default=0, help="first N instances per family")
    args = ap.parse_args()
    if args.paper:
        paper_sets(args.paper_guard_s, [f for f in args.paper_families.split(",") if f],
                   args.paper_limit)
        return
    if args.calib_copy_mib:
        import torch
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": n * 4}))
        return
    global WORKLOAD, CASES
    if args.only:
        WORKLOAD = [w for w in WORKLOAD if w[0] == args.only]
    if args.cases:
        CASES = [tuple(int(x) for x in c.split(":")) for c in args.cases.split(",")]
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                          "(instances shard across GPUs without communication)"}))
        sys.exit(2)

    from pmdd_amd.engine import Engine

    tmp = tempfile.mkdtemp(prefix="pmdd_bench_")
    dds = build_workload(tmp)
    eng = Engine(0)
    res = [(name, eng.upload(dd)) for name, _, dd in dds]
    log(f"engine {Engine.version()}; workload " + ", ".join(n for n, _ in res))

    def step(phase_timing=False):
        pts = 0
        agg = {"kernel_filter_s": 0.0, "filter_launches": 0, "algorithmic_bytes": 0,
               "kernel_frontier_check_s": 0.0, "frontier_check_launches": 0,
               "frontier_check_bytes": 0, "kernel_cross_s": 0.0, "cross_launches": 0,
               "cross_bytes": 0, "work_frontier_survivors_total": 0, "kernel_total_s": 0.0,
               "kernel_state_dominance_s": 0.0, "state_dominance_launches": 0, "join_bytes": 0}
        per = {}
        for name, r in res:
            ref = None
            for method, sd in CASES:
                out = r.enumerate(method=method, state_dominance=sd, max_prod=MAX_PROD,
                                  phase_timing=phase_timing)
                if ref is None:
                    ref = out.points
                elif out.points.shape != ref.shape or not (out.points == ref).all():
                    raise SystemExit(f"{name}: method {method} sd {sd} frontier differs")
                pts += len(out.points)
                for k in agg:
                    agg[k] += out.stats[k]
                per[f"{name}_m{method}_sd{sd}"] = len(out.points)
        return pts, agg, per

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")

    total_pts = 0
    tot = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        pts, agg, per = step()
        total_pts += pts
        if tot is None:
            tot = dict(agg)
        else:
            for k in agg:
                tot[k] += agg[k]
        log(f"step {i + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / max(args.steps, 1)

    # The timed steps record no per-layer event (phase timing costs about 6%
    # of the wall time).  The same steps again with phase timing give each
    # kernel family's device time (HIP events on the engine's stream around
    # the family's launches); the launches are the same in both.
    ttot = None
    t2 = time.perf_counter()
    for i in range(args.steps):
        _, agg, _ = step(phase_timing=True)
        if ttot is None:
            ttot = dict(agg)
        else:
            for k in agg:
                ttot[k] += agg[k]
    ms_per_step_phase_timed = 1000.0 * (time.perf_counter() - t2) / max(args.steps, 1)
    # (the bytes of the frontier check are not compared: they count the
    # survivors of each block of the running frontier, whose order varies
    # from run to run)
    for k in tot:
        if (k.endswith("_launches") or k == "work_frontier_survivors_total") and \
                tot[k] != ttot[k]:
            raise SystemExit(f"phase-timed steps differ from the timed steps in {k}")
    log("phase-timed steps done")
    enumerations = args.steps * len(res) * len(CASES)
    steps = max(args.steps, 1)

    # the same step with every diagram uploaded inside the timed call (the
    # reference's wall_enumeration_s wraps packing and the host-to-device copy,
    # main.cpp:327-394 -> enum.cu:438-593)
    t1 = time.perf_counter()
    for name, _, dd in dds:
        for method, sd in CASES:
            eng.enumerate(dd, method=method, state_dominance=sd, max_prod=MAX_PROD)
    ms_with_upload = 1000.0 * (time.perf_counter() - t1)
    log("upload-inclusive step done")

    # executed pair tests of one step per kernel family (deterministic), from
    # a counting run; each enumeration's tests count d compares
    fam_keys = {"k_filter_chunks": "pair_tests_filter",
                "k_refill_vs_block": "pair_tests_frontier_check",
                "k_cross_chunks": "pair_tests_cross",
                "k_state_dominance": "pair_tests_state_dominance"}
    tests = {k: 0 for k in fam_keys}
    lane_ops = {k: 0 for k in fam_keys}
    if not args.skip_pair_count:
        d_of = {name: d for name, _, _, d in WORKLOAD}
        for name, r in res:
            for method, sd in CASES:
                st = r.enumerate(method=method, state_dominance=sd, max_prod=MAX_PROD,
                                 count_pair_tests=True).stats
                for fam, key in fam_keys.items():
                    tests[fam] += st[key]
                    lane_ops[fam] += st[key] * d_of[name]
        log("pair-test count done")

    # per kernel family: device time (HIP events on the engine's stream over
    # the phase-timed steps), launches, compulsory HBM bytes, executed pair tests
    tot = ttot
    fam = {
        "k_filter_chunks": (tot["kernel_filter_s"], tot["filter_launches"],
                            tot["algorithmic_bytes"]),
        "k_refill_vs_block": (tot["kernel_frontier_check_s"], tot["frontier_check_launches"],
                              tot["join_bytes"]),
        "k_cross_chunks": (tot["kernel_cross_s"], tot["cross_launches"], tot["cross_bytes"]),
        "k_state_dominance": (tot["kernel_state_dominance_s"], tot["state_dominance_launches"],
                              None),
    }
    families = {}
    for k, (t, n, b) in fam.items():
        e = {"device_ms_per_step": t / steps * 1e3, "launches_per_step": n / steps,
             "avg_launch_ms": t / max(n, 1) * 1e3,
             "share_of_engine_kernel_time": t / max(tot["kernel_total_s"], 1e-12)}
        if not args.skip_pair_count and t > 0:
            tops = lane_ops[k] * steps / t / 1e12
            e.update({"pair_tests_per_step": tests[k],
                      "valu_achieved_tops": tops, "valu_frac": tops / VALU_PEAK_TOPS,
                      "tests_per_s": tests[k] * steps / t})
        families[k] = e
    kname = max(fam, key=lambda k: fam[k][0])
    k_time, k_launches, k_bytes = fam[kname]
    avg_launch_s = k_time / max(k_launches, 1)
    traffic, traffic_src = read_traffic(kname)
    if args.skip_pair_count:
        achieved = None
    else:
        achieved = lane_ops[kname] * steps / max(k_time, 1e-12) / 1e12
    roofline = {
        "kernel": kname,
        "bound": "valu",
        "achieved": achieved,
        "peak": VALU_PEAK_TOPS,
        "unit": "TOP/s (int32 compare lane-ops: d per executed pair test)",
        "frac": achieved / VALU_PEAK_TOPS if achieved is not None else None,
        "traffic": traffic,
        "traffic_source": traffic_src,
        "avg_launch_ms": avg_launch_s * 1e3,
        "launches_per_step": k_launches / steps,
        "share_of_engine_kernel_time": k_time / max(tot["kernel_total_s"], 1e-12),
        "families": families,
    }
    if k_bytes is not None and k_launches > 0:
        gbs = k_bytes / k_launches / avg_launch_s / 1e9
        roofline["hbm"] = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": gbs / HBM_PEAK_GBS,
                           "algorithmic_bytes_per_launch": k_bytes / k_launches,
                           "bytes": "coupling: 4*d*(TD and BU points of each batch + final "
                                    "frontier)" if kname == "k_refill_vs_block" else
                                    "compulsory bytes of the family"}

    cpu = None
    if not args.no_cpu_baseline:
        phys, logical = host_cores()
        threads = args.cpu_threads or min(logical, 16)
        cpu = cpu_baseline(dds, threads, args.cpu_budget_s)
        log(f"cpu baseline done: {cpu['value']:.1f} pts/s on {threads} threads")

    srv = None
    if args.serving_requests > 0:
        rates = tuple(float(x) for x in args.serving_rates.split(",") if x)
        deadline = T_START + args.time_budget_s
        srv = serving_leg(tmp, args.serving_requests, args.serving_slots, args.serving_method,
                          rates, deadline)
        if not srv["bit_exact_batched_and_packed_vs_plain"]:
            raise SystemExit("serving: batched and plain frontiers differ")

    line = {
        "metric": "exact Pareto pts/s, inst/s, p50/p99 latency: batched vs plain GPU",
        "value": total_pts / elapsed,
        "unit": "Pareto pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "ms_per_step_with_upload": ms_with_upload,
        "ms_per_step_phase_timed": ms_per_step_phase_timed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded SplitMix64 knapsack instances)",
        "config": {
            "workload": "single large knapsack BDDs: d=3 n=80 and d=5 n=50, methods 1 and 3, "
                        "state dominance 0 and 1 (8 enumerations per step), plain GPU path, "
                        "diagrams resident in HBM",
            "instances": [n for n, _, _ in dds],
            "max_cand": 20_000_000,
            "max_prod": "adaptive: coupling batches of 32K products doubling to 32M (the "
                        "frontier does not depend on the batch split)",
        },
        "inst_per_s": enumerations / elapsed,
        "frontier_points_processed_per_s": tot["work_frontier_survivors_total"] / elapsed,
        "comparisons_per_s": (sum(tests.values()) * args.steps / elapsed)
        if not args.skip_pair_count else None,
        "pareto_points_per_step": total_pts / steps,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "serving": srv,
        "wall_s": time.perf_counter() - T_START,
    }
    print(json.dumps(line), flush=True)
    for _, r in res:
        r.close()
    eng.close()


if __name__ == "__main__":
    main()
