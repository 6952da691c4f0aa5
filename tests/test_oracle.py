"""CPU tests that pin the oracle before it is trusted as the checker.

What pins each oracle function (oracle/oracle.cpp):
  merge_shift / remove_empty / topdown   brute force on fixtures and generated
                                          instances (closed form: nondominated
                                          set of all feasible objective vectors)
  bottomup, coupled + convolute + tree   same brute force, and equality with
  reduction                              topdown on larger instances
  knapsack_state_dominance               sd=1 frontier == sd=0 frontier (the rule
                                          is sound, dominance.cpp:24-86) and
                                          brute force on small cases
  nondominated (brute-force filter)      hand-built point sets with known answers
  known answers                          831 solutions for TSP-15 seed108
                                          (reference reports/report.tex:226)
A planted mistake (one point dropped, one coordinate changed) must fail the
comparison used by every parity test.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import instance_path, problem_type_of
from pmdd_amd import DecisionDiagram
from pmdd_amd import instances as gen

SMALL = ["KP_p-3_n-10_ins-1.dat", "KP_p-3_n-20_ins-1.dat", "KP_p-4_n-20_ins-3.dat",
         "KP_p-5_n-10_ins-4.dat", "knapsack-20-6-1000-1065.dat",
         "tsp-nobj3-ncities5-seed11563.dat", "tsp-nobj7-ncities5-seed10689.dat"]


def same(a, b):
    return a.shape == b.shape and bool((a == b).all())


def test_nondominated_known_answers():
    pts = np.array([[1, 2, 3], [3, 2, 1], [1, 2, 3], [0, 0, 0], [1, 2, 2], [2, 2, 2],
                    [3, 2, 1]], dtype=np.int32)
    got = oracle.nondominated(pts)
    want = np.array([[1, 2, 3], [2, 2, 2], [3, 2, 1]], dtype=np.int32)
    assert same(got, want)
    # a chain: only the top survives
    chain = np.array([[i, i, i, i] for i in range(10)], dtype=np.int32)
    assert same(oracle.nondominated(chain), np.array([[9, 9, 9, 9]], dtype=np.int32))
    # an antichain: everything survives
    anti = np.array([[i, 9 - i] for i in range(10)], dtype=np.int32)
    assert len(oracle.nondominated(anti)) == 10


@pytest.mark.parametrize("name", SMALL)
def test_oracle_methods_match_bruteforce(name):
    pt = problem_type_of(name)
    path = instance_path(name)
    bf = oracle.bruteforce(path, pt)
    dd = DecisionDiagram.from_file(path, pt)
    for method in (1, 2, 3):
        for sd in ((0, 1) if pt == 1 else (0,)):
            for threads in (1, 4):
                got = oracle.enumerate_dd(dd, method, sd, threads=threads)
                assert same(got, bf), (name, method, sd, threads)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_generated_vs_bruteforce(tmp_path, seed):
    cases = [
        gen.knapsack(seed, 10 + seed, 1 + seed % 6),
        gen.setpacking(50 + seed, 18 + seed, 2 + seed % 5, m=5),
        gen.tsp(90 + seed, 5 + seed % 3, 2 + seed % 5),
    ]
    for k, inst in enumerate(cases):
        path = gen.write(inst, str(tmp_path / f"c{k}.dat"))
        bf = oracle.bruteforce(path, inst["kind"])
        dd = DecisionDiagram.from_file(path, inst["kind"])
        for method in (1, 2, 3):
            for sd in (0, 1):
                assert same(oracle.enumerate_dd(dd, method, sd, threads=2), bf), (seed, k, method)


def test_oracle_methods_agree_on_larger_instances():
    for name in ["KP_p-3_n-30_ins-2.dat", "bp-100-20-3-10-10192.dat",
                 "tsp-nobj3-ncities10-seed12870.dat"]:
        pt = problem_type_of(name)
        dd = DecisionDiagram.from_file(instance_path(name), pt)
        ref = oracle.enumerate_dd(dd, 1, 0, threads=4)
        for method, sd in [(2, 0), (3, 0), (3, 1), (1, 1)]:
            assert same(oracle.enumerate_dd(dd, method, sd, threads=4), ref), (name, method, sd)


def test_known_count_tsp15_seed108():
    dd = DecisionDiagram.from_file(instance_path("tsp-nobj3-ncities15-seed108.dat"), 3)
    assert len(oracle.enumerate_dd(dd, 3, 0, threads=8)) == 831


def test_state_dominance_filters_labels():
    """sd=1 must actually drop labels on a knapsack (otherwise the sd tests
    above pin nothing) and still return the same frontier."""
    path = instance_path("KP_p-3_n-30_ins-2.dat")
    dd = DecisionDiagram.from_file(path, 1)
    a, st0, _, _ = oracle.enumerate_dd(dd, 1, 0, threads=4, with_stats=True)
    b, st1, _, _ = oracle.enumerate_dd(dd, 1, 1, threads=4, with_stats=True)
    assert same(a, b)
    assert st0.dominance_filtered_total == 0
    assert st1.dominance_filtered_total > 0
    assert st1.work_frontier_survivors_total < st0.work_frontier_survivors_total


def test_planted_mistakes_are_caught():
    path = instance_path("KP_p-3_n-20_ins-1.dat")
    bf = oracle.bruteforce(path, 1)
    dropped = bf[1:]
    assert not same(dropped, bf)
    bumped = bf.copy()
    bumped[len(bumped) // 2, 1] += 1
    assert not same(bumped, bf)
    swapped = bf[:, ::-1].copy()
    assert not same(oracle.sort_lex(swapped), bf)


def test_work_counters_are_order_independent():
    """Candidates and survivors per layer do not depend on threads (they are
    set sizes), which is what lets the GPU counters be compared to them."""
    dd = DecisionDiagram.from_file(instance_path("bp-100-20-4-10-12796.dat"), 2)
    _, a, ca, sa = oracle.enumerate_dd(dd, 1, 0, threads=1, with_stats=True)
    _, b, cb, sb = oracle.enumerate_dd(dd, 1, 0, threads=8, with_stats=True)
    for f in ("work_candidates_total", "work_candidates_peak", "work_frontier_survivors_total",
              "work_frontier_peak_points"):
        assert getattr(a, f) == getattr(b, f)
    assert np.allclose(ca, cb) and np.allclose(sa, sb)


def test_splitmix64_published_vector():
    """SplitMix64 reference outputs for seed 1234567 (Steele, Lea, Flood 2014;
    the xoshiro authors' splitmix64.c)."""
    r = gen.SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
        16408922859458223821]


def test_uniform_is_in_range_and_covers():
    r = gen.SplitMix64(7)
    xs = [r.uniform(1, 6) for _ in range(6000)]
    assert min(xs) == 1 and max(xs) == 6
    counts = np.bincount(xs)[1:]
    assert counts.min() > 850  # each face near 1000
