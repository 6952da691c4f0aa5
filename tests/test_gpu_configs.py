"""Full-size configurations on the GPU against golden frontiers.

tests/golden/config_counts.json holds, per seeded configuration instance,
the Pareto-frontier size and the SHA-256 of the sorted int32 frontier as the
oracle computed them (tests/golden/make_config_counts.py, methods 1 and 3
agreeing).  Every GPU method / state-dominance combination must reproduce the
digest bit for bit.
"""
import hashlib
import json
import os

import pytest

from pmdd_amd import DecisionDiagram
from pmdd_amd import instances as gen

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "config_counts.json")))


def digest(points):
    return hashlib.sha256(points.astype("<i4").tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_config_frontier_digest(engine, tmp_path, name):
    g = GOLDEN[name]
    path = gen.write(gen.generate(g["kind"], g["seed"], g["n"], g["d"], **g["extra"]),
                     str(tmp_path / f"{name}.dat"))
    dd = DecisionDiagram.from_file(path, g["kind"])
    sds = (0, 1) if g["kind"] == 1 else (0,)
    for method in (1, 3):
        for sd in sds:
            print(f"case {name} m{method} sd{sd}", flush=True)
            pts = engine.enumerate(dd, method=method, state_dominance=sd).points
            assert len(pts) == g["count"], (name, method, sd)
            assert digest(pts) == g["sha256_sorted"], (name, method, sd)
