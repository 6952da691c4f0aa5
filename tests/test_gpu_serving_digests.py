"""Served results of the serving configuration's stream against the oracle.

tests/golden/serving_digests.json holds, for a prefix of the seeded
2000-request mixed stream (60% knapsack n~U[20,70], 25% set packing
n~U[60,150] m=n/5, 15% TSP n~U[8,11], d~U{3,4,5}), each request's frontier
size and the SHA-256 of its sorted int32 frontier, as the oracle computed
them (tests/golden/make_serving_digests.py).  The prefix is served by the
batched server with every request arriving at once, so each runs next to
others, and every result must match its digest:
  - 8 slots, methods 1 and 3;
  - a memory budget so small that the head of the queue waits for memory;
  - an injected device allocation failure in the middle of a run, which the
    server retries with the request alone.
"""
import hashlib
import json
import os

import pytest

from pmdd_amd import abi
from pmdd_amd import serving

pytestmark = pytest.mark.gpu

GOLDEN_PATH = os.path.join(os.path.dirname(__file__), "golden", "serving_digests.json")
GOLDEN = json.load(open(GOLDEN_PATH))
REQ = GOLDEN["requests"]
PREFIX = 0
while str(PREFIX) in REQ:
    PREFIX += 1


def _digest(points):
    return hashlib.sha256(points.astype("<i4").tobytes()).hexdigest()


@pytest.fixture(scope="module")
def stream(tmp_path_factory):
    specs = serving.mixed_stream(GOLDEN["stream"]["count"], seed=GOLDEN["stream"]["seed"])
    specs = specs[:PREFIX]
    for i, sp in enumerate(specs):
        assert sp == REQ[str(i)]["spec"], i  # the generator still draws the same stream
    paths = serving.write_stream(specs, str(tmp_path_factory.mktemp("stream")))
    return specs, paths


def _check(results, indices):
    bad = []
    for i in indices:
        g = REQ[str(i)]
        pts = results.get(i)
        if pts is None or len(pts) != g["count"] or _digest(pts) != g["sha256_sorted"]:
            bad.append((i, g["spec"], None if pts is None else len(pts), g["count"]))
    assert not bad, bad


def _cheap(limit_s):
    return [i for i in range(PREFIX) if REQ[str(i)]["oracle_m3_s"] < limit_s]


@pytest.mark.parametrize("method", [1, 3])
def test_served_prefix_matches_oracle_digests(engine, stream, method):
    specs, paths = stream
    print(f"case serving digests prefix={PREFIX} method={method}", flush=True)
    results, stats = {}, {}
    m = serving.run_stream(paths, specs, [0.0] * PREFIX, slots=8, method=method,
                           results=results, stats=stats)
    assert m["failed"] == 0, m["errors"]
    _check(results, range(PREFIX))
    assert m["admission"]["admitted"] == PREFIX
    assert m["admission"]["running_peak"] > 1  # requests did run next to each other
    print(f"memory: peak/estimate max {m['mem_peak_over_estimate_max']:.3f}, "
          f"{m['mem_peak_above_estimate']} requests above their estimate", flush=True)


@pytest.mark.parametrize("inflight", [32, 3])
def test_packed_serving_matches_oracle_digests(engine, stream, inflight):
    """Packed serving: method-1 requests on small diagrams share packed
    layers (a packer per objective count expands the current layers of its
    in-flight requests as one layer, through a captured graph replayed every
    step); requests join at any step and leave at their terminal layer.
    Every served frontier matches its digest, requests were packed several
    to a step, and the rest of the stream ran in the slots."""
    specs, paths = stream
    print(f"case packed serving prefix={PREFIX} inflight={inflight}", flush=True)
    results = {}
    m = serving.run_stream(paths, specs, [0.0] * PREFIX, slots=8, method=1, results=results,
                           pack=True, pack_inflight=inflight)
    assert m["failed"] == 0, m["errors"]
    _check(results, range(PREFIX))
    a = m["admission"]
    assert a["admitted"] == PREFIX
    assert 0 < a["packed"] < PREFIX  # some packed, the large ones in the slots
    assert a["pack_steps"] > 0
    assert a["pack_requests_per_step_x100"] > 100  # more than one request per step
    print(f"admission: {a}", flush=True)


def test_small_budget_makes_the_head_wait(engine, stream):
    """A budget a little above the largest estimate: the head of the queue
    waits for memory while slots are free (it never does under the default
    budget), the slots' reservations never exceed the budget, and every
    result still matches."""
    specs, paths = stream
    idx = _cheap(1.0)[:24]
    sp = [specs[i] for i in idx]
    ps = [paths[i] for i in idx]
    print(f"case small budget, {len(idx)} requests", flush=True)
    first = {}
    m0 = serving.run_stream(ps, sp, [0.0] * len(idx), slots=8, method=1, stats=first)
    assert m0["admission"]["memory_waits"] == 0  # the default budget never binds here
    est = [first[k]["mem_estimate_bytes"] for k in range(len(idx))]
    budget = max(est) + min(est) // 2
    results = {}
    m = serving.run_stream(ps, sp, [0.0] * len(idx), slots=8, method=1, results=results,
                           memory_budget=budget)
    assert m["failed"] == 0, m["errors"]
    _check({idx[k]: v for k, v in results.items()}, idx)
    a = m["admission"]
    assert a["budget"] == budget
    assert a["memory_waits"] > 0
    assert a["reserved_peak"] <= budget
    print(f"admission: {a}", flush=True)


def test_injected_nomem_is_retried_alone(engine, stream):
    """The 30th device allocation from now fails as out of memory, in the
    middle of some request's run next to others; the server queues that
    request again to run alone, and every result matches its digest."""
    specs, paths = stream
    idx = _cheap(1.0)[:16]
    sp = [specs[i] for i in idx]
    ps = [paths[i] for i in idx]
    print(f"case injected NOMEM, {len(idx)} requests", flush=True)
    results, stats = {}, {}
    lib = abi.engine_lib()
    lib.pmdd_debug_fail_alloc(30)
    try:
        m = serving.run_stream(ps, sp, [0.0] * len(idx), slots=4, method=1, results=results,
                               stats=stats)
    finally:
        lib.pmdd_debug_fail_alloc(0)
    assert m["failed"] == 0, m["errors"]
    _check({idx[k]: v for k, v in results.items()}, idx)
    assert m["admission"]["nomem_retries"] == 1
    assert sum(s["nomem_retries"] for s in stats.values()) == 1
