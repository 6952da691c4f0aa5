"""Serving: every request's frontier is the same alone and in the mix, and a
rejected request leaves the others unaffected."""
import gc

import numpy as np
import pytest

from pmdd_amd import DecisionDiagram, abi
from pmdd_amd import instances as gen
from pmdd_amd import serving

pytestmark = pytest.mark.gpu


def _same(a, b):
    return a.shape == b.shape and bool((a == b).all())


@pytest.mark.parametrize("method", [1, 3])
def test_mixed_stream_alone_vs_in_the_mix(engine, tmp_path, method):
    # the serving mix, capped in size so that the test stays short
    specs = [sp for sp in serving.mixed_stream(80, seed=11)
             if (sp["kind"] == 1 and sp["n"] <= 45) or (sp["kind"] == 2 and sp["n"] <= 100)
             or sp["kind"] == 3][:24]
    paths = serving.write_stream(specs, str(tmp_path))
    alone = []
    for sp, p in zip(specs, paths):
        alone.append(engine.enumerate(DecisionDiagram.from_file(p, sp["kind"]), method=method).points)
    results = {}
    arrivals = [0.0] * len(paths)   # everything arrives at once: maximal overlap
    m = serving.run_stream(paths, specs, arrivals, slots=6, method=method, results=results)
    assert m["failed"] == 0
    for i in range(len(paths)):
        assert _same(results[i], alone[i]), (i, specs[i])


def test_rejection_is_isolated(engine, tmp_path):
    small = gen.write(gen.tsp(3, 8, 3), str(tmp_path / "small.dat"))
    big = gen.write(gen.setpacking(5, 150, 4, m=30), str(tmp_path / "big.dat"))
    dd_s = DecisionDiagram.from_file(small, 3)
    dd_b = DecisionDiagram.from_file(big, 2)
    ref = engine.enumerate(dd_s, method=1).points
    with serving.Server(0, slots=2, memory_budget=64 << 20) as srv:
        t1 = srv.submit(dd_s, method=1)
        t2 = srv.submit(dd_b, method=1)
        t3 = srv.submit(dd_s, method=3)
        with pytest.raises(abi.PmddError) as e:
            srv.wait(t2)
        assert e.value.code == abi.PMDD_E_REJECTED
        assert _same(srv.wait(t1).points, ref)
        assert _same(srv.wait(t3).points, ref)


def test_poll_then_collect(engine, tmp_path):
    p = gen.write(gen.knapsack(9, 40, 4), str(tmp_path / "k.dat"))
    dd = DecisionDiagram.from_file(p, 1)
    ref = engine.enumerate(dd, method=3).points
    with serving.Server(0, slots=1) as srv:
        t = srv.submit(dd, method=3)
        res = None
        for _ in range(200000):
            res = srv.poll(t)
            if res is not None:
                break
        if res is None:
            res = srv.wait(t)
        assert _same(res.points, ref)
        with pytest.raises(abi.PmddError):
            srv.poll(t)  # collected tickets are retired


def test_wait_any_collects_every_ticket_once(engine, tmp_path):
    small = gen.write(gen.tsp(4, 8, 3), str(tmp_path / "t.dat"))
    big = gen.write(gen.setpacking(6, 150, 4, m=30), str(tmp_path / "b.dat"))
    dd_s = DecisionDiagram.from_file(small, 3)
    dd_b = DecisionDiagram.from_file(big, 2)
    ref = engine.enumerate(dd_s, method=1).points
    with serving.Server(0, slots=2, memory_budget=64 << 20) as srv:
        tickets = {srv.submit(dd_s, method=1): "ok", srv.submit(dd_b, method=1): "rejected",
                   srv.submit(dd_s, method=3): "ok"}
        seen = {}
        for _ in range(3):
            t, res = srv.wait_any()
            seen[t] = res
        assert set(seen) == set(tickets)
        for t, kind in tickets.items():
            if kind == "rejected":
                assert isinstance(seen[t], abi.PmddError)
                assert seen[t].code == abi.PMDD_E_REJECTED
            else:
                assert _same(seen[t].points, ref)
        with pytest.raises(abi.PmddError):
            srv.wait_any()  # nothing outstanding


def test_submit_copies_the_partner_array(engine, tmp_path):
    """The server copies every array of a submitted diagram, the set-packing
    state-dominance partners included, so the caller may free the diagram
    before its request runs."""
    paths = [gen.write(gen.setpacking(seed, 60, 3, m=12), str(tmp_path / f"s{seed}.dat"))
             for seed in (21, 22, 23)]
    refs = [engine.enumerate(DecisionDiagram.from_file(p, 2, state_dominance=True), method=1,
                             state_dominance=1).points for p in paths]
    with serving.Server(0, slots=1) as srv:
        tickets = []
        for p in paths:
            dd = DecisionDiagram.from_file(p, 2, state_dominance=True)
            tickets.append(srv.submit(dd, method=1, state_dominance=1))
            del dd  # frees the host arrays at once
        gc.collect()
        churn = [DecisionDiagram.from_file(p, 2, state_dominance=True) for p in paths]
        for t, ref in zip(tickets, refs):
            assert _same(srv.wait(t).points, ref)
        del churn

