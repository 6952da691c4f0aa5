"""Writes tests/golden/bignode_digests.json: frontier size and SHA-256 of the
sorted int32 frontier of reference-paper instances with 6 and 7 objectives
whose nodes receive more candidates than one chunk of the engine holds
(1024 points for d >= 5), so that the engine's big-node path (parent runs
merged, the node cut into key-ordered blocks, blocks filtered against the
node's front and each other) runs on them.  Computed by the oracle (the CPU
restatement of the reference's OpenMP enumerators), method 3; method 1 too
where it finishes in a few minutes, and must agree.  max_node_candidates is
the oracle's largest per-node candidate count (method 1), recorded to show
that the big-node path is exercised.

Run from the repository root:
    python tests/golden/make_bignode_digests.py [THREADS]
It resumes: instances already in the file are skipped.
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from pmdd_amd import DecisionDiagram  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "bignode_digests.json")
# (family directory under tests/golden/paper, instance file, problem type)
INSTANCES = [
    ("mokp_n40_k6", "knapsack-40-6-1000-12237.dat", 1),
    ("mokp_n40_k6", "knapsack-40-6-1000-253.dat", 1),
    ("mokp_n40_k7", "knapsack-40-7-1000-12144.dat", 1),
    ("mospp_n150_k7", "bp-150-30-7-10-14816.dat", 2),
]
M1_BELOW_S = 300.0


def digest(points):
    return hashlib.sha256(points.astype("<i4").tobytes()).hexdigest()


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for fam, name, kind in INSTANCES:
        if name in data:
            continue
        dd = DecisionDiagram.from_file(os.path.join(ROOT, "tests", "golden", "paper", fam, name),
                                       kind)
        t0 = time.perf_counter()
        a = oracle.enumerate_dd(dd, 3, 0, threads=threads)
        dt = time.perf_counter() - t0
        entry = {"family": fam, "kind": kind, "d": int(dd.d), "count": int(len(a)),
                 "sha256_sorted": digest(a), "oracle_m3_s": round(dt, 2)}
        if dt < M1_BELOW_S:
            t0 = time.perf_counter()
            b, st, _, _ = oracle.enumerate_dd(dd, 1, 0, threads=threads, with_stats=True)
            assert b.shape == a.shape and (a == b).all(), f"{name}: oracle m1 != m3"
            entry["checked_m1"] = True
            entry["oracle_m1_s"] = round(time.perf_counter() - t0, 2)
            entry["max_node_candidates"] = int(st.max_node_candidates)
        data[name] = entry
        print(name, entry, flush=True)
        with open(OUT + ".tmp", "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(OUT + ".tmp", OUT)


if __name__ == "__main__":
    main()
