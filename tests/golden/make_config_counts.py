"""Writes tests/golden/config_counts.json: Pareto-frontier sizes (and a
checksum of the sorted frontier) of the seeded full-size configuration
instances, computed by the oracle (the CPU restatement of the reference's
OpenMP enumerators) with methods 1 and 3, which must agree.

Run from the repository root:  python tests/golden/make_config_counts.py
(takes a few minutes on 8 cores).
"""
import hashlib
import json
import os
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from pmdd_amd import DecisionDiagram  # noqa: E402
from pmdd_amd import instances as gen  # noqa: E402

# (name, kind, seed, n, d, extra generator args)
CONFIGS = [
    ("kp_d3_n40", 1, 40_3, 40, 3, {}),
    ("kp_d3_n80", 1, 80_3, 80, 3, {}),
    ("kp_d5_n50", 1, 50_5, 50, 5, {}),
    ("sp_d4_n150", 2, 150_4, 150, 4, {"m": 30}),
    ("tsp_d3_n12", 3, 12_3, 12, 3, {}),
]


def digest(points):
    return hashlib.sha256(points.astype("<i4").tobytes()).hexdigest()


def main():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, kind, seed, n, d, kw in CONFIGS:
            path = gen.write(gen.generate(kind, seed, n, d, **kw), os.path.join(tmp, name + ".dat"))
            dd = DecisionDiagram.from_file(path, kind)
            a = oracle.enumerate_dd(dd, 1, 0, threads=os.cpu_count() or 1)
            b = oracle.enumerate_dd(dd, 3, 0, threads=os.cpu_count() or 1)
            assert a.shape == b.shape and (a == b).all(), name
            out[name] = {"kind": kind, "seed": seed, "n": n, "d": d, "extra": kw,
                         "count": int(len(a)), "sha256_sorted": digest(a)}
            print(name, len(a), flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "config_counts.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
