"""GPU: layers planned on the device (the default) against layers planned on
the host (PMDD_HOST_PLAN=1, the planner the packed serving layers and the
passes over --max-cand use).  Both must give the same frontier and the same
statistics: candidates and survivors per run and their peaks, the per-layer
std of candidates and survivors, the coupling layer, the labels state
dominance removes, the join's product count.  The oracle pins the frontiers
(test_gpu_parity.py); this pins that the two planners agree on everything
the statistics expose.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import instance_path, problem_type_of
from pmdd_amd import DecisionDiagram

pytestmark = pytest.mark.gpu

EXACT = ("work_candidates_total", "work_candidates_peak", "work_frontier_survivors_total",
         "work_frontier_peak_points", "work_join_products_total", "layer_coupling",
         "dominance_filtered_total")


def _run(engine, dd, host, **kw):
    old = os.environ.pop("PMDD_HOST_PLAN", None)
    try:
        if host:
            os.environ["PMDD_HOST_PLAN"] = "1"
        return engine.enumerate(dd, **kw)
    finally:
        os.environ.pop("PMDD_HOST_PLAN", None)
        if old is not None:
            os.environ["PMDD_HOST_PLAN"] = old


def _check(engine, dd, ref, **kw):
    dev = _run(engine, dd, False, **kw)
    host = _run(engine, dd, True, **kw)
    assert dev.points.shape == ref.shape and (dev.points == ref).all()
    assert host.points.shape == ref.shape and (host.points == ref).all()
    for k in EXACT:
        assert dev.stats[k] == host.stats[k], (k, dev.stats[k], host.stats[k])
    for k in ("std_candidates_per_layer", "std_frontier_survivors_per_layer"):
        np.testing.assert_allclose(dev.stats[k], host.stats[k], rtol=1e-12, atol=1e-9)


CASES = ["KP_p-3_n-30_ins-2.dat", "KP_p-5_n-10_ins-4.dat", "knapsack-20-6-1000-1065.dat",
         "bp-100-20-3-10-10192.dat", "tsp-nobj3-ncities10-seed12870.dat",
         "tsp-nobj7-ncities5-seed10689.dat"]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("method", [1, 2, 3])
def test_device_and_host_planning_agree(engine, name, method):
    print(f"case {name} method={method}", flush=True)
    pt = problem_type_of(name)
    dd = DecisionDiagram.from_file(instance_path(name), pt)
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    _check(engine, dd, ref, method=method)
    if pt != 3 and method != 2:
        # state dominance: planned on the device after a device-planned layer
        dds = DecisionDiagram.from_file(instance_path(name), pt, state_dominance=True)
        _check(engine, dds, ref, method=method, state_dominance=1)


@pytest.mark.parametrize("max_cand", [300, 3000])
def test_mixed_planning_under_a_candidate_budget(engine, max_cand):
    """A budget that some layers exceed: those layers go to the host planner
    (passes of whole nodes) between device-planned ones, in every method."""
    print(f"case max_cand={max_cand}", flush=True)
    name = "KP_p-3_n-30_ins-2.dat"
    dd = DecisionDiagram.from_file(instance_path(name), 1, state_dominance=True)
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    for method in (1, 2, 3):
        for sd in ((0, 1) if method != 2 else (0,)):
            _check(engine, dd, ref, method=method, state_dominance=sd, max_cand=max_cand)


def test_big_nodes_in_several_waves(engine):
    """MOTSP n=15, d=5 (117,171 points, kb/autoresearch/TSP_STATUS.md:50 of
    the reference): the last layers' nodes hold hundreds of thousands of
    candidates, more than 64 blocks of 1024, so the device-planned cross
    filter runs several waves over one list of the big nodes' chunks, each
    launch taking its own wave's.  Both planners give the reference's count
    and the same frontier and statistics."""
    name = "tsp-nobj5-ncities15-seed13095.dat"
    dd = DecisionDiagram.from_file(instance_path(name), 3)
    for method in (1, 3):
        print(f"case {name} method={method}", flush=True)
        dev = _run(engine, dd, False, method=method)
        host = _run(engine, dd, True, method=method)
        assert len(dev.points) == 117171
        assert dev.points.shape == host.points.shape and (dev.points == host.points).all()
        for k in EXACT:
            assert dev.stats[k] == host.stats[k], (k, dev.stats[k], host.stats[k])
        print(f"cross launches {dev.stats['cross_launches']} (device) "
              f"{host.stats['cross_launches']} (host)", flush=True)
        if method == 1:
            # one launch per wave: more launches than expansions means some
            # layer ran several waves
            assert dev.stats["cross_launches"] > dd.num_layers


def test_narrow_layers_take_the_big_node_path(engine):
    """KP_p-3_n-30 (d=3, 2048-candidate chunks): no node exceeds a chunk,
    but the narrow layers near the root and the terminal (fewer than 256
    nodes) send every node above 256 candidates to the big-node path, one
    chunk each, whose cross filter spreads it over several workgroups.  The
    cross filter therefore runs in the top-down sweep, and the frontier is
    the oracle's in every method."""
    name = "KP_p-3_n-30_ins-2.dat"
    dd = DecisionDiagram.from_file(instance_path(name), 1)
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    for method in (1, 2, 3):
        print(f"case {name} method={method}", flush=True)
        dev = _run(engine, dd, False, method=method)
        assert dev.points.shape == ref.shape and (dev.points == ref).all()
        if method == 1:
            # the top-down sweep meets such nodes (a probe counted 8 cross
            # launches; with the rule switched off, none)
            assert dev.stats["cross_launches"] > 0
        host = _run(engine, dd, True, method=method)
        for k in EXACT:
            assert dev.stats[k] == host.stats[k], (k, dev.stats[k], host.stats[k])


LAUNCHES = ("filter_launches", "cross_launches", "frontier_check_launches",
            "state_dominance_launches")


@pytest.mark.parametrize("name,kind,sd", [("KP_p-3_n-30_ins-2.dat", 1, 0),
                                          ("KP_p-3_n-30_ins-2.dat", 1, 1),
                                          ("bp-100-20-3-10-10192.dat", 2, 0)])
def test_two_stream_layers_match_one_stream(engine, name, kind, sd):
    """A layer with big nodes runs its small-node filters on the context's
    second stream, beside the big nodes' gather, merges and cross waves,
    unless phase timing is on (then every launch stays on one stream).  Both
    give the oracle's frontier, the same launches and the same statistics;
    method 1 of the knapsack meets big nodes (cross launches) and small ones
    (filter launches) in its sweep."""
    dd = DecisionDiagram.from_file(instance_path(name), kind, state_dominance=bool(sd))
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    for method in (1, 3):
        print(f"case two streams {name} sd={sd} method={method}", flush=True)
        two = engine.enumerate(dd, method=method, state_dominance=sd)
        one = engine.enumerate(dd, method=method, state_dominance=sd, phase_timing=True)
        assert two.points.shape == ref.shape and (two.points == ref).all()
        assert one.points.shape == ref.shape and (one.points == ref).all()
        if method == 1 and kind == 1:
            assert two.stats["cross_launches"] > 0 and two.stats["filter_launches"] > 0
        for k in EXACT + LAUNCHES:
            assert two.stats[k] == one.stats[k], (k, two.stats[k], one.stats[k])
