"""Writes tests/golden/serving_digests.json: for a prefix of the seeded mixed
serving stream (pmdd_amd.serving.mixed_stream(2000), the serving
configuration: 60% knapsack n~U[20,70], 25% set packing n~U[60,150] m=n/5,
15% TSP n~U[8,11], d~U{3,4,5}), the Pareto-frontier size and the SHA-256 of
the sorted int32 frontier of every request, computed by the oracle (the CPU
restatement of the reference's OpenMP enumerators) with method 3.  Where the
oracle's method 3 finishes in a few seconds, its method 1 runs too and must
agree (recorded as "checked_m1").

The GPU test (tests/test_gpu_serving_digests.py) serves the same prefix
through the batched server and compares every result with these digests.

Run from the repository root:
    python tests/golden/make_serving_digests.py [PREFIX] [THREADS]
It resumes: requests already in the file are skipped.  The heavy requests
(5-objective knapsacks with n >= 50) take minutes each on 8 cores.
"""
import hashlib
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from pmdd_amd import DecisionDiagram  # noqa: E402
from pmdd_amd import serving  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "serving_digests.json")
STREAM_LEN = 2000
M1_CHECK_BELOW_S = 5.0


def digest(points):
    return hashlib.sha256(points.astype("<i4").tobytes()).hexdigest()


def main():
    prefix = int(sys.argv[1]) if len(sys.argv) > 1 else 100
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else (os.cpu_count() or 1)
    data = {"stream": {"count": STREAM_LEN, "seed": 2000}, "requests": {}}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    specs = serving.mixed_stream(STREAM_LEN)
    with tempfile.TemporaryDirectory() as tmp:
        paths = serving.write_stream(specs[:prefix], tmp)
        for i in range(prefix):
            key = str(i)
            if key in data["requests"]:
                continue
            sp = specs[i]
            dd = DecisionDiagram.from_file(paths[i], sp["kind"])
            t0 = time.perf_counter()
            a = oracle.enumerate_dd(dd, 3, 0, threads=threads)
            dt = time.perf_counter() - t0
            checked = False
            if dt < M1_CHECK_BELOW_S:
                b = oracle.enumerate_dd(dd, 1, 0, threads=threads)
                assert b.shape == a.shape and (a == b).all(), f"request {i}: oracle m1 != m3"
                checked = True
            dd.close()
            data["requests"][key] = {"spec": sp, "count": int(len(a)),
                                     "sha256_sorted": digest(a), "oracle_m3_s": round(dt, 3),
                                     "checked_m1": checked}
            print(f"request {i} {sp} {len(a)} points, {dt:.2f} s", flush=True)
            with open(OUT + ".tmp", "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(OUT + ".tmp", OUT)


if __name__ == "__main__":
    main()
