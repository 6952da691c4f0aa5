"""Autotuner of the engine's launch shapes, per objective count.

    python -m pmdd_amd.tune [--out pmdd_amd/tune/gfx950.json] [--d 3,4,5,6,7,8]
                            [--reps 3]

For each objective count d it runs a few seeded workloads of that d
(methods 1 and 3, diagrams resident) under candidate shapes and keeps, per
kernel family, the shape with the least device time of that family
(coordinate descent, the minimum over --reps runs):
  - chunk filter (k_filter_chunks): threads per workgroup for nodes of at
    most 256 candidates, at most 1024 (d <= 4), at most a whole chunk, each
    128, 256 or 512 -- timed by kernel_filter_s;
  - block cross filter (k_cross_chunks): workgroups per wave and the most per
    block -- timed by kernel_cross_s;
  - frontier check (k_refill_vs_block): one-wave workgroups -- timed by
    kernel_frontier_check_s.
Every run's frontier must equal the first run's (shapes never change a
result).  The table holds, per d, the chosen shapes and the measured device
time of the workloads with the built-in shapes and with the chosen ones.
The engine reads the table when a context is created (pmdd_ctx_create).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import tempfile
import time

from . import DecisionDiagram
from . import instances as gen
from .engine import Engine

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune", "gfx950.json")
PAPER = os.path.join(_ROOT, "tests", "golden", "paper")

BUILTIN = {"filter_threads": [128, 512, 512], "cross_workgroups": 2048, "cross_split_max": 16,
           "refill_waves": 8192}


def workloads(d: int, tmp: str):
    """(name, diagram, methods) per d: the benchmark's knapsack configurations,
    the set-packing and TSP configurations, reference-paper instances."""
    out = []

    def gen_dd(name, inst):
        path = gen.write(inst, os.path.join(tmp, name + ".dat"))
        return DecisionDiagram.from_file(path, inst["kind"])

    if d == 3:
        out.append(("knapsack n=80", gen_dd("kp3", gen.knapsack(80_3, 80, 3)), (1, 3)))
        out.append(("tsp n=12", gen_dd("tsp3", gen.tsp(12_3, 12, 3)), (1, 3)))
    elif d == 4:
        out.append(("set packing n=150 m=30", gen_dd("sp4", gen.setpacking(150_4, 150, 4, m=30)),
                    (1, 3)))
        out.append(("knapsack n=60", gen_dd("kp4", gen.knapsack(60_4, 60, 4)), (1, 3)))
    elif d == 5:
        out.append(("knapsack n=50", gen_dd("kp5", gen.knapsack(50_5, 50, 5)), (1, 3)))
        out.append(("tsp-15 seed11093", DecisionDiagram.from_file(
            os.path.join(PAPER, "motsp_n15_k5", "tsp-nobj5-ncities15-seed11093.dat"), 3), (1, 3)))
    elif d == 6:
        out.append(("knapsack-40-6 seed253", DecisionDiagram.from_file(
            os.path.join(PAPER, "mokp_n40_k6", "knapsack-40-6-1000-253.dat"), 1), (1, 3)))
    elif d == 7:
        out.append(("knapsack-40-7 seed12144", DecisionDiagram.from_file(
            os.path.join(PAPER, "mokp_n40_k7", "knapsack-40-7-1000-12144.dat"), 1), (1, 3)))
        out.append(("bp-150-30-7 seed14816", DecisionDiagram.from_file(
            os.path.join(PAPER, "mospp_n150_k7", "bp-150-30-7-10-14816.dat"), 2), (1,)))
    else:
        out.append((f"knapsack n=30 d={d}", gen_dd(f"kp{d}", gen.knapsack(30_0 + d, 30, d)),
                    (1, 3)))
    return out


def digest(points) -> str:
    return hashlib.sha256(points.astype("<i4").tobytes()).hexdigest()


class Measure:
    """Runs the workloads of one d under a shape and checks the frontiers."""

    def __init__(self, eng: Engine, d: int, wl, reps: int):
        self.eng, self.d, self.reps = eng, d, reps
        self.res = [(name, eng.upload(dd), methods) for name, dd, methods in wl]
        self.ref = {}

    def __call__(self, cfg: dict) -> dict:
        self.eng.set_tuning(self.d, cfg)
        best = None
        for _ in range(self.reps):
            fam = {"filter": 0.0, "cross": 0.0, "check": 0.0, "kernel": 0.0}
            for name, r, methods in self.res:
                for m in methods:
                    out = r.enumerate(method=m, phase_timing=True)
                    key = (name, m)
                    dg = digest(out.points)
                    if self.ref.setdefault(key, dg) != dg:
                        raise SystemExit(f"tune: frontier of {name} method {m} changed under {cfg}")
                    s = out.stats
                    fam["filter"] += s["kernel_filter_s"]
                    fam["cross"] += s["kernel_cross_s"]
                    fam["check"] += s["kernel_frontier_check_s"]
                    fam["kernel"] += s["kernel_total_s"]
            best = fam if best is None else {k: min(best[k], fam[k]) for k in fam}
        return best

    def close(self):
        for _, r, _ in self.res:
            r.close()


def tune_d(eng: Engine, d: int, reps: int, tmp: str, log) -> dict:
    wl = workloads(d, tmp)
    meas = Measure(eng, d, wl, reps)
    try:
        cfg = json.loads(json.dumps(BUILTIN))
        base = meas(cfg)
        log(f"d={d} built-in: kernel {base['kernel'] * 1e3:.1f} ms (filter "
            f"{base['filter'] * 1e3:.1f}, cross {base['cross'] * 1e3:.1f}, check "
            f"{base['check'] * 1e3:.1f})")
        tried = []
        for tier in range(3):
            best_t, best_v = None, None
            for nt in (128, 256, 512):
                c = json.loads(json.dumps(cfg))
                c["filter_threads"][tier] = nt
                v = meas(c)["filter"]
                tried.append({"filter_tier": tier, "threads": nt, "filter_ms": v * 1e3})
                if best_v is None or v < best_v:
                    best_t, best_v = nt, v
            cfg["filter_threads"][tier] = best_t
        best_c, best_v = None, None
        for cw in (1024, 2048, 4096):
            for sm in (8, 16):
                c = dict(cfg, cross_workgroups=cw, cross_split_max=sm)
                v = meas(c)["cross"]
                tried.append({"cross_workgroups": cw, "cross_split_max": sm, "cross_ms": v * 1e3})
                if best_v is None or v < best_v:
                    best_c, best_v = (cw, sm), v
        cfg["cross_workgroups"], cfg["cross_split_max"] = best_c
        best_r, best_v = None, None
        for rw in (4096, 8192, 16384):
            c = dict(cfg, refill_waves=rw)
            v = meas(c)["check"]
            tried.append({"refill_waves": rw, "check_ms": v * 1e3})
            if best_v is None or v < best_v:
                best_r, best_v = rw, v
        cfg["refill_waves"] = best_r
        tuned = meas(cfg)
        log(f"d={d} tuned {cfg}: kernel {tuned['kernel'] * 1e3:.1f} ms")
        return {"tuning": cfg,
                "measured": {"workloads": [f"{n} methods {list(m)}" for n, _, m in wl],
                             "reps": reps,
                             "builtin_ms": {k: v * 1e3 for k, v in base.items()},
                             "tuned_ms": {k: v * 1e3 for k, v in tuned.items()},
                             "tried": tried}}
    finally:
        meas.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--d", default="3,4,5,6,7,8")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    eng = Engine(0)
    table = {"device": "gfx950", "engine": Engine.version(),
             "written": time.strftime("%Y-%m-%d %H:%M:%S"),
             "note": "launch shapes per objective count; device times are the sums of the "
                     "listed workloads' kernel families (minimum over reps)",
             "tuning": {}, "measured": {}}
    with tempfile.TemporaryDirectory() as tmp:
        for d in (int(x) for x in args.d.split(",") if x):
            r = tune_d(eng, d, args.reps, tmp, log)
            table["tuning"][str(d)] = r["tuning"]
            table["measured"][str(d)] = r["measured"]
    eng.close()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
