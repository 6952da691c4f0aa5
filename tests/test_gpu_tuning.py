"""Launch shapes (pmdd_ctx_get_tuning / pmdd_ctx_set_tuning): a context
starts with the committed tuning table's shapes, refuses shapes out of range,
and gives the oracle's frontier under every shape the tuner may pick."""
import json
import os

import pytest

import oracle
from conftest import instance_path
from pmdd_amd import DecisionDiagram, abi
from pmdd_amd import instances as gen
from pmdd_amd.engine import Engine

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pmdd_amd",
                     "tune", "gfx950.json")


def _same(a, b):
    return a.shape == b.shape and bool((a == b).all())


def test_context_takes_the_table():
    table = json.load(open(TABLE))["tuning"]
    with Engine(0) as eng:
        for d in range(1, 9):
            t = eng.get_tuning(d)
            if str(d) in table:
                assert t["source"] == 1
                for k, v in table[str(d)].items():
                    assert t[k] == v, (d, k)
            else:
                assert t["source"] == 0


def test_out_of_range_shapes_are_refused(engine):
    good = engine.get_tuning(3)
    for bad in ({"filter_threads": [100, 512, 512]}, {"cross_workgroups": 10},
                {"cross_split_max": 0}, {"refill_waves": 1 << 20}):
        with pytest.raises(abi.PmddError) as e:
            engine.set_tuning(3, dict(good, **bad))
        assert e.value.code == abi.PMDD_E_INVALID
    assert engine.get_tuning(3)["source"] == good["source"]


@pytest.mark.parametrize("shape", [
    {"filter_threads": [128, 128, 128], "cross_workgroups": 256, "cross_split_max": 1,
     "refill_waves": 256},
    {"filter_threads": [512, 256, 256], "cross_workgroups": 16384, "cross_split_max": 16,
     "refill_waves": 32768},
])
def test_every_shape_gives_the_oracle_frontier(tmp_path, shape):
    cases = [(instance_path("KP_p-3_n-30_ins-2.dat"), 1),
             (gen.write(gen.knapsack(77, 34, 5), str(tmp_path / "k5.dat")), 1),
             (instance_path("tsp-nobj3-ncities10-seed12870.dat"), 3)]
    with Engine(0) as eng:
        for path, kind in cases:
            dd = DecisionDiagram.from_file(path, kind)
            ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
            eng.set_tuning(dd.d, shape)
            assert eng.get_tuning(dd.d)["source"] == 2
            for method in (1, 3):
                print(f"case tuning {os.path.basename(path)} m{method} {shape}", flush=True)
                assert _same(eng.enumerate(dd, method=method).points, ref)
