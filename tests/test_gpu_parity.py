"""GPU parity: the HIP engine, called through the C ABI, against the oracle.

Every comparison is bit-exact on the frontier as a lexicographically sorted
set of int32 vectors, the way the reference writes it (main.cpp:403).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from conftest import instance_path, problem_type_of
from pmdd_amd import DecisionDiagram, abi
from pmdd_amd import instances as gen

pytestmark = pytest.mark.gpu

FIXTURES = sorted(os.listdir(os.path.join(os.path.dirname(__file__), "golden", "instances")))
# the 15-city TSP fixtures are covered by the known-count tests below
SMALL_FIXTURES = [f for f in FIXTURES if "ncities15" not in f]


def _dd(name, state_dominance=False):
    return DecisionDiagram.from_file(instance_path(name), problem_type_of(name),
                                     state_dominance=state_dominance)


def _same(a, b):
    return a.shape == b.shape and bool((a == b).all())


@pytest.mark.parametrize("name", SMALL_FIXTURES)
@pytest.mark.parametrize("method", [1, 2, 3])
def test_fixture_matches_oracle(engine, name, method):
    print(f"case {name} method={method}", flush=True)
    dd = _dd(name)
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    sds = [0, 1] if problem_type_of(name) != 3 else [0]
    for sd in sds:
        # state dominance data (set-packing partners) is built on request
        ddk = _dd(name, True) if sd else dd
        got = engine.enumerate(ddk, method=method, state_dominance=sd).points
        assert _same(got, ref), f"{name} m{method} sd{sd}: {len(got)} vs {len(ref)}"


@pytest.mark.parametrize("name", ["KP_p-3_n-30_ins-2.dat", "tsp-nobj3-ncities10-seed12870.dat",
                                  "bp-100-20-3-10-10192.dat"])
@pytest.mark.parametrize("max_cand", [1, 7, 1000])
def test_candidate_budget_invariance(engine, name, max_cand):
    print(f"case {name} max_cand={max_cand}", flush=True)
    dd = _dd(name)
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    for method in (1, 3):
        got = engine.enumerate(dd, method=method, max_cand=max_cand).points
        assert _same(got, ref)


@pytest.mark.parametrize("name", ["KP_p-3_n-30_ins-2.dat", "tsp-nobj3-ncities10-seed12870.dat"])
@pytest.mark.parametrize("max_prod", [1, 1000, 625000, 0])
def test_product_budget_invariance(engine, name, max_prod):
    """--max-prod caps every coupling batch (couple.cu:393-403); a node with
    more products is split over batches, so no batch exceeds the cap, and
    the frontier does not depend on the split.  0 is the adaptive mode."""
    print(f"case {name} max_prod={max_prod}", flush=True)
    dd = _dd(name)
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    res = engine.enumerate(dd, method=3, max_prod=max_prod)
    assert _same(res.points, ref)
    s = res.stats
    assert s["join_batches"] >= 1
    if max_prod > 0:
        assert s["join_batch_peak_products"] <= max_prod
    if max_prod == 1:
        # every product alone: every node was split down to single products
        assert s["join_batches"] == s["work_join_products_total"]


def test_tsp12_product_cap_holds(engine, tmp_path):
    """The 3-objective TSP n=12 configuration at --max-cand 20M --max-prod
    625K: no coupling batch holds more than 625K products, and methods 1 and 3
    give the oracle's frontier."""
    path = gen.write(gen.tsp(12_3, 12, 3), str(tmp_path / "tsp12.dat"))
    dd = DecisionDiagram.from_file(path, 3)
    ref = oracle.enumerate_dd(dd, 3, 0, threads=8)
    res = engine.enumerate(dd, method=3, max_cand=20_000_000, max_prod=625_000)
    assert _same(res.points, ref)
    assert 0 < res.stats["join_batch_peak_products"] <= 625_000
    assert _same(engine.enumerate(dd, method=1, max_cand=20_000_000).points, ref)


def _write_knapsack(path, profits, weights, capacity):
    """A knapsack file in the reference format (knapsack_instance.cpp:17-47):
    "n 1 d", d rows of profits, the weights, the capacity."""
    d, n = len(profits), len(weights)
    with open(path, "w") as f:
        f.write(f"{n} 1 {d}\n")
        for row in profits:
            f.write(" ".join(str(x) for x in row) + "\n")
        f.write(" ".join(str(x) for x in weights) + "\n")
        f.write(f"{capacity}\n")
    return str(path)


@pytest.mark.parametrize("capacity", [1, 2])
def test_path_sum_bound_exactly_int32_max(engine, tmp_path, capacity):
    """The upload check bounds every path sum by the sum over layers of the
    largest |arc weight|; at exactly INT32_MAX the enumeration runs and is
    exact (with capacity 2 the single Pareto point has objective 0 equal to
    INT32_MAX)."""
    big = 1 << 30
    path = _write_knapsack(tmp_path / "edge.dat", [[big, big - 1], [1, 2]], [1, 1], capacity)
    bf = oracle.bruteforce(path, 1)
    if capacity == 2:
        assert bf.tolist() == [[2**31 - 1, 3]]
    else:
        assert bf.tolist() == [[big - 1, 2], [big, 1]]
    dd = DecisionDiagram.from_file(path, 1)
    for method in (1, 2, 3):
        got = engine.enumerate(dd, method=method).points
        assert _same(got, bf), method


def test_path_sum_bound_above_int32_max_is_refused(engine, tmp_path):
    """One more than INT32_MAX: PMDD_E_OVERFLOW, before anything runs (the
    reference adds unchecked, bottomup.cu:124, couple.cu:93-94)."""
    big = 1 << 30
    path = _write_knapsack(tmp_path / "over.dat", [[big, big], [1, 2]], [1, 1], 2)
    dd = DecisionDiagram.from_file(path, 1)
    for method in (1, 3):
        with pytest.raises(abi.PmddError) as e:
            engine.enumerate(dd, method=method)
        assert e.value.code == abi.PMDD_E_OVERFLOW


def test_tsp15_seed108_known_count(engine):
    """831 solutions: reports/report.tex:226 of the reference."""
    dd = _dd("tsp-nobj3-ncities15-seed108.dat")
    ref = oracle.enumerate_dd(dd, 3, 0, threads=8)
    assert len(ref) == 831
    for method in (1, 3):
        got = engine.enumerate(dd, method=method).points
        assert _same(got, ref)


def test_tsp15_seed11417_known_count(engine):
    """5372 solutions: reports/report.tex:227 of the reference."""
    dd = _dd("tsp-nobj4-ncities15-seed11417.dat")
    a = engine.enumerate(dd, method=3).points
    assert len(a) == 5372
    b = engine.enumerate(dd, method=1).points
    assert _same(a, b)


@pytest.mark.parametrize("seed", range(6))
def test_generated_small_vs_bruteforce(engine, tmp_path, seed):
    cases = [
        gen.knapsack(seed, 14 + seed, 3),
        gen.knapsack(100 + seed, 12, 1 + seed % 8),
        gen.setpacking(200 + seed, 20 + 2 * seed, 2 + seed % 4, m=6),
        gen.tsp(300 + seed, 6 + seed % 3, 2 + seed % 6),
    ]
    for k, inst in enumerate(cases):
        path = gen.write(inst, str(tmp_path / f"g{seed}_{k}.dat"))
        print(f"case seed={seed} kind={inst['kind']} n={inst['n']} d={inst['d']}", flush=True)
        bf = oracle.bruteforce(path, inst["kind"])
        dds = [DecisionDiagram.from_file(path, inst["kind"]),
               DecisionDiagram.from_file(path, inst["kind"], state_dominance=True)]
        for method in (1, 2, 3):
            for sd in (0, 1):
                got = engine.enumerate(dds[sd], method=method, state_dominance=sd).points
                assert _same(got, bf), (seed, k, method, sd)


@pytest.mark.parametrize("d", [3, 5])
def test_generated_knapsack_stats_match_oracle(engine, tmp_path, d):
    """Order-independent work counters of method 1 without state dominance."""
    inst = gen.knapsack(7 + d, 30, d)
    path = gen.write(inst, str(tmp_path / "kp.dat"))
    dd = DecisionDiagram.from_file(path, 1)
    ref, st, std_c, std_s = oracle.enumerate_dd(dd, 1, 0, threads=4, with_stats=True)
    res = engine.enumerate(dd, method=1)
    assert _same(res.points, ref)
    s = res.stats
    assert s["work_candidates_total"] == st.work_candidates_total
    assert s["work_candidates_peak"] == st.work_candidates_peak
    assert s["work_frontier_survivors_total"] == st.work_frontier_survivors_total
    assert s["work_frontier_peak_points"] == st.work_frontier_peak_points
    np.testing.assert_allclose(s["std_candidates_per_layer"], std_c, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(s["std_frontier_survivors_per_layer"], std_s, rtol=1e-12,
                               atol=1e-9)


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_bottomup_stats_match_oracle(engine, tmp_path, kind):
    """Method 2: frontier and order-independent work counters and per-layer
    std against the oracle's bottom-up (bottomup_cpu_enumerate)."""
    inst = {1: gen.knapsack(21, 28, 4), 2: gen.setpacking(22, 40, 3, m=8), 3: gen.tsp(23, 8, 3)}[kind]
    path = gen.write(inst, str(tmp_path / "g.dat"))
    dd = DecisionDiagram.from_file(path, kind)
    print(f"case bottom-up kind={kind}", flush=True)
    ref, st, std_c, std_s = oracle.enumerate_dd(dd, 2, 0, threads=4, with_stats=True)
    assert _same(ref, oracle.enumerate_dd(dd, 1, 0, threads=4))
    res = engine.enumerate(dd, method=2)
    assert _same(res.points, ref)
    s = res.stats
    assert s["work_candidates_total"] == st.work_candidates_total
    assert s["work_candidates_peak"] == st.work_candidates_peak
    assert s["work_frontier_survivors_total"] == st.work_frontier_survivors_total
    assert s["work_frontier_peak_points"] == st.work_frontier_peak_points
    np.testing.assert_allclose(s["std_candidates_per_layer"], std_c, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(s["std_frontier_survivors_per_layer"], std_s, rtol=1e-12,
                               atol=1e-9)


@pytest.mark.parametrize("seed", range(4))
def test_setpacking_state_dominance_filters_and_keeps_frontier(engine, tmp_path, seed):
    """Set-packing state dominance (partners built with the diagram) drops
    labels on the top-down layers of methods 1 and 3 and leaves the frontier
    equal to the oracle's without it.  The number dropped depends on the
    partner choice, which differs from the reference's pairing, so only its
    sign is checked."""
    inst = gen.setpacking(500 + seed, 60 + 10 * seed, 3 + seed % 2, m=10)
    path = gen.write(inst, str(tmp_path / "sp.dat"))
    print(f"case sp-sd seed={seed} n={inst['n']} d={inst['d']}", flush=True)
    dd0 = DecisionDiagram.from_file(path, 2)
    dd1 = DecisionDiagram.from_file(path, 2, state_dominance=True)
    ref = oracle.enumerate_dd(dd0, 1, 0, threads=8)
    filtered = 0
    for method in (1, 3):
        r = engine.enumerate(dd1, method=method, state_dominance=1)
        assert _same(r.points, ref), (seed, method, len(r.points), len(ref))
        filtered += r.stats["dominance_filtered_total"]
        # the flag alone, without partners, filters nothing
        r0 = engine.enumerate(dd0, method=method, state_dominance=1)
        assert _same(r0.points, ref)
        assert r0.stats["dominance_filtered_total"] == 0
    assert filtered > 0


def test_coupled_layer_matches_oracle_gpu_rule(engine):
    dd = _dd("tsp-nobj3-ncities10-seed12870.dat")
    ref, st, _, _ = oracle.enumerate_dd(dd, 3, 0, threads=4, score_rule=1, with_stats=True)
    res = engine.enumerate(dd, method=3)
    assert _same(res.points, ref)
    assert res.stats["layer_coupling"] == st.layer_coupling
    assert res.stats["work_join_products_total"] == st.work_join_products_total


def test_resident_reuse(engine):
    dd = _dd("KP_p-3_n-20_ins-1.dat")
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    r = engine.upload(dd)
    for method in (1, 3, 1):
        assert _same(r.enumerate(method=method).points, ref)
    r.close()


def test_pair_test_counter_is_deterministic(engine):
    dd = _dd("KP_p-3_n-30_ins-2.dat")
    a = engine.enumerate(dd, method=1, count_pair_tests=True).stats["dominance_pair_tests"]
    b = engine.enumerate(dd, method=1, count_pair_tests=True).stats["dominance_pair_tests"]
    assert a == b and a > 0


def test_engine_reuse_across_objective_counts(engine, tmp_path):
    """One context runs instances of different d in turn: frontier buffers are
    reused between calls with the SoA stride recomputed for each d."""
    cases = [gen.knapsack(61, 24, 5), gen.knapsack(62, 30, 3), gen.setpacking(63, 40, 4, m=8),
             gen.knapsack(64, 22, 5)]
    for k, inst in enumerate(cases):
        path = gen.write(inst, str(tmp_path / f"r{k}.dat"))
        dd = DecisionDiagram.from_file(path, inst["kind"])
        ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
        for method in (1, 3, 2):
            print(f"case reuse k={k} d={inst['d']} method={method}", flush=True)
            assert _same(engine.enumerate(dd, method=method).points, ref), (k, method)


TIMES = ("kernel_expand_td_s", "kernel_dominance_s", "kernel_total_s", "kernel_filter_s",
         "kernel_frontier_check_s", "kernel_cross_s", "kernel_state_dominance_s")


def test_kernel_time_breakdown_is_consistent(engine):
    """With phase timing, kernel_total_s is expansion plus dominance time; the
    filter, frontier check and cross figures are parts of the dominance time.
    Without it every kernel time is 0 and the frontier and counts are the
    same."""
    dd = _dd("KP_p-3_n-30_ins-2.dat")
    for method in (1, 3):
        r = engine.enumerate(dd, method=method, state_dominance=1, phase_timing=True)
        s = r.stats
        assert s["kernel_total_s"] > 0
        assert s["kernel_total_s"] == pytest.approx(
            s["kernel_expand_td_s"] + s["kernel_dominance_s"], rel=1e-9, abs=1e-12)
        parts = s["kernel_filter_s"] + s["kernel_frontier_check_s"] + s["kernel_cross_s"]
        assert parts <= s["kernel_dominance_s"] * (1 + 1e-6) + 1e-9
        assert s["filter_launches"] > 0
        q = engine.enumerate(dd, method=method, state_dominance=1)
        assert all(q.stats[k] == 0.0 for k in TIMES), {k: q.stats[k] for k in TIMES}
        assert _same(q.points, r.points)
        for k in ("work_candidates_total", "work_frontier_survivors_total",
                  "dominance_filtered_total", "filter_launches", "cross_launches"):
            assert q.stats[k] == s[k], k


def test_named_entry_point_times_its_kernels(engine):
    """pmdd_topdown_enumerate and pmdd_coupled_bdd_enumerate, the entry points
    that replace the reference's topdown_cuda_enumerate and
    coupled_bdd_cuda_enumerate, fill the kernel times whenever they are given
    stats, as the reference does (topdown.cu:428-507), and give the oracle's
    frontier."""
    import ctypes as C
    from pmdd_amd.engine import _take_frontier
    dd = _dd("KP_p-3_n-30_ins-2.dat")
    ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
    lib = engine._lib
    for name in ("pmdd_topdown_enumerate", "pmdd_coupled_bdd_enumerate"):
        print(f"case {name}", flush=True)
        fr = abi.Frontier()
        st = abi.Stats()
        err = C.create_string_buffer(1024)
        args = [engine._h, dd.view_ptr, 1, 0, abi.DEFAULT_MAX_CAND]
        if name == "pmdd_coupled_bdd_enumerate":
            args.append(abi.DEFAULT_MAX_PROD)
        rc = getattr(lib, name)(*args, C.byref(fr), C.byref(st), err, len(err))
        assert rc == abi.PMDD_OK, err.value.decode()
        assert _same(_take_frontier(lib, fr, True), ref)
        assert st.kernel_total_s > 0 and st.kernel_dominance_s > 0
        assert st.kernel_total_s == pytest.approx(st.kernel_expand_td_s + st.kernel_dominance_s,
                                                  rel=1e-9, abs=1e-12)


def test_tsp15_5obj_seed13095_known_count(engine):
    """117171 solutions: the reference's CPU run of this instance,
    kb/autoresearch/TSP_STATUS.md:50 (the reference GPU path ran out of
    memory on it before chunking, kb/autoresearch/TSP_STATUS.md:28-37)."""
    dd = _dd("tsp-nobj5-ncities15-seed13095.dat")
    a = engine.enumerate(dd, method=3).points
    assert len(a) == 117171
    b = engine.enumerate(dd, method=1).points
    assert _same(a, b)


@pytest.mark.parametrize("family,name,kind,sd", [
    ("mokp_n40_k7", "knapsack-40-7-1000-12144.dat", 1, 1),
    ("mospp_n150_k7", "bp-150-30-7-10-14816.dat", 2, 0),
])
def test_paper_instances_methods_agree(engine, family, name, kind, sd):
    """Instances of the reference paper's benchmark (too large for the
    oracle): methods 1 and 3 (and 2, for the knapsack) give the same
    frontier."""
    path = os.path.join(os.path.dirname(__file__), "golden", "paper", family, name)
    print(f"case paper {family}/{name} sd={sd}", flush=True)
    dd = DecisionDiagram.from_file(path, kind)
    ref = engine.enumerate(dd, method=1, state_dominance=sd).points
    assert len(ref) > 1000
    methods = (2, 3) if kind == 1 else (3,)
    for method in methods:
        got = engine.enumerate(dd, method=method, state_dominance=sd).points
        assert _same(got, ref), (name, method, len(got), len(ref))


# Solution counts the reference reports for the 5-objective 15-city TSP
# family, kb/autoresearch/TSP_STATUS.md:28-37 (seed13095 is the known-count
# test above).
TSP5_COUNTS = {1081: 55198, 11093: 29505, 14792: 77303, 23138: 51306, 24645: 38883,
               30131: 24381, 32239: 78849, 6440: 32237, 7702: 53917}


@pytest.mark.parametrize("seed", sorted(TSP5_COUNTS))
def test_tsp15_5obj_reference_counts(engine, seed):
    path = os.path.join(os.path.dirname(__file__), "golden", "paper", "motsp_n15_k5",
                        f"tsp-nobj5-ncities15-seed{seed}.dat")
    print(f"case tsp15 5obj seed{seed}", flush=True)
    dd = DecisionDiagram.from_file(path, 3)
    a = engine.enumerate(dd, method=3).points
    assert len(a) == TSP5_COUNTS[seed]
    b = engine.enumerate(dd, method=1).points
    assert _same(a, b)


BIGNODE_PATH = os.path.join(os.path.dirname(__file__), "golden", "bignode_digests.json")
BIGNODE = json.load(open(BIGNODE_PATH)) if os.path.exists(BIGNODE_PATH) else {}


def _digest(points):
    return hashlib.sha256(points.astype("<i4").tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(BIGNODE))
def test_bignode_instances_match_oracle_digests(engine, name):
    """6- and 7-objective reference-paper instances whose nodes exceed one
    chunk (tests/golden/make_bignode_digests.py): the big-node path (merged
    parent runs, key-ordered blocks, cross filter against the node's front)
    against the oracle's frontier digest, methods 1 and 3 (knapsack also
    with state dominance)."""
    g = BIGNODE[name]
    path = os.path.join(os.path.dirname(__file__), "golden", "paper", g["family"], name)
    dd = DecisionDiagram.from_file(path, g["kind"])
    sds = (0, 1) if g["kind"] == 1 else (0,)
    for method in (1, 3):
        for sd in sds:
            print(f"case bignode {name} m{method} sd{sd}", flush=True)
            res = engine.enumerate(dd, method=method, state_dominance=sd)
            assert len(res.points) == g["count"], (name, method, sd)
            assert _digest(res.points) == g["sha256_sorted"], (name, method, sd)
            if method == 1 and sd == 0:
                assert res.stats["cross_launches"] > 0  # big nodes were met
