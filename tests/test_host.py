"""Host surface on the CPU: parsers, diagram constructors, flattening, ABI
exports and the CLI's argument rules."""
import ctypes as C
import itertools
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import instance_path
from pmdd_amd import DecisionDiagram, abi
from pmdd_amd import instances as gen


def test_libraries_export_every_declared_symbol():
    host = C.CDLL(abi.HOST_LIB)
    for sym in abi.HOST_SYMBOLS:
        assert hasattr(host, sym), sym
    eng = C.CDLL(abi.ENGINE_LIB)
    for sym in abi.ENGINE_SYMBOLS:
        assert hasattr(eng, sym), sym


def test_header_declarations_match_symbol_lists():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    import re
    for header, syms in [("pmdd_engine.h", abi.ENGINE_SYMBOLS),
                         ("pmdd_host.h", abi.HOST_SYMBOLS)]:
        text = open(os.path.join(root, "include", header)).read()
        declared = set(re.findall(r"\b(pmdd_[a-z_]+)\s*\(", text))
        assert declared == set(syms), (header, declared ^ set(syms))


def test_engine_reports_missing_device_without_crashing():
    if abi.engine_lib().pmdd_device_count() > 0:
        pytest.skip("a device is visible here")
    from pmdd_amd.engine import Engine
    with pytest.raises(abi.PmddError) as e:
        Engine(0)
    assert e.value.code == abi.PMDD_E_NODEVICE


def test_tsp_layer_sizes_closed_form():
    """Layer l (2..n) of the TSP MDD holds C(n-1, l-1) * (l-1) states."""
    dd = DecisionDiagram.from_file(instance_path("tsp-nobj3-ncities10-seed12870.dat"), 3)
    n = 10
    sizes = dd.layer_sizes()
    assert len(sizes) == n + 2
    assert sizes[0] == 1 and sizes[1] == 1 and sizes[-1] == 1
    for l in range(2, n + 1):
        assert sizes[l] == math.comb(n - 1, l - 1) * (l - 1)
    s = dd.structure()
    assert s.original_width == -1 and s.reduced_num_nodes == -1
    arr = dd.arrays()
    assert s.total_num_in_arcs == len(arr["arc_src"])


def test_tsp_arc_weights_are_negated_distances(tmp_path):
    inst = gen.tsp(3, 5, 2)
    path = gen.write(inst, str(tmp_path / "t.dat"))
    dd = DecisionDiagram.from_file(path, 3)
    a = dd.arrays()
    lo = a["layer_off"]
    # closing arcs into the terminal: weight = -dist[last][0]; every weight <= 0
    assert (a["arc_w"] <= 0).all()
    term_arcs = slice(a["in_off"][lo[-2]], a["in_off"][lo[-1]])
    assert len(a["arc_src"][term_arcs]) == 4  # last city in {1..4}


def _reachable_paths(dd):
    """All root->terminal arc-label sequences of a small flattened BDD."""
    a = dd.arrays()
    lo, io = a["layer_off"], a["in_off"]
    L = dd.num_layers
    # out-arcs per node
    out = {}
    for l in range(1, L):
        for v in range(lo[l + 1] - lo[l]):
            g = lo[l] + v
            for e in range(io[g], io[g + 1]):
                out.setdefault((l - 1, int(a["arc_src"][e])), []).append((v, int(a["arc_val"][e])))
    paths = []

    def walk(l, v, acc):
        if l == L - 1:
            paths.append(tuple(acc))
            return
        for (h, val) in out.get((l, v), []):
            walk(l + 1, h, acc + [val])
    walk(0, 0, [])
    return paths


def test_knapsack_bdd_paths_are_exactly_the_feasible_subsets(tmp_path):
    inst = gen.knapsack(11, 9, 2)
    path = gen.write(inst, str(tmp_path / "k.dat"))
    dd = DecisionDiagram.from_file(path, 1)
    got = sorted(_reachable_paths(dd))
    w, cap = inst["weights"], inst["capacity"]
    want = sorted(x for x in itertools.product((0, 1), repeat=9)
                  if sum(wi * xi for wi, xi in zip(w, x)) <= cap)
    assert got == want
    s = dd.structure()
    assert s.reduced_num_nodes <= s.original_num_nodes
    assert s.reduced_width <= s.original_width


def test_knapsack_min_weight_is_min_used_capacity(tmp_path):
    inst = gen.knapsack(5, 8, 2)
    path = gen.write(inst, str(tmp_path / "k.dat"))
    dd = DecisionDiagram.from_file(path, 1)
    a = dd.arrays()
    lo, io = a["layer_off"], a["in_off"]
    w = inst["weights"]
    # min over in-arcs recomputed from the flattened arcs
    mw = {0: 0}
    for l in range(1, dd.num_layers):
        for v in range(lo[l + 1] - lo[l]):
            g = lo[l] + v
            best = None
            for e in range(io[g], io[g + 1]):
                p = mw[lo[l - 1] + int(a["arc_src"][e])] + (w[l - 1] if a["arc_val"][e] == 1 else 0)
                best = p if best is None else min(best, p)
            mw[g] = best
    assert all(int(a["min_weight"][g]) == mw[g] for g in mw)


def test_reduced_bdd_has_no_mergeable_nodes():
    dd = DecisionDiagram.from_file(instance_path("KP_p-3_n-30_ins-2.dat"), 1)
    a = dd.arrays()
    lo, io = a["layer_off"], a["in_off"]
    for l in range(0, dd.num_layers - 1):
        children = {}
        for v in range(lo[l + 2] - lo[l + 1]):
            g = lo[l + 1] + v
            for e in range(io[g], io[g + 1]):
                children.setdefault(int(a["arc_src"][e]), [-1, -1])[int(a["arc_val"][e])] = v
        pairs = [tuple(c) for c in children.values()]
        assert len(pairs) == len(set(pairs)), l


def test_bdd_in_arcs_list_one_arcs_first():
    dd = DecisionDiagram.from_file(instance_path("KP_p-3_n-20_ins-1.dat"), 1)
    a = dd.arrays()
    io = a["in_off"]
    for g in range(len(io) - 1):
        vals = a["arc_val"][io[g]:io[g + 1]]
        assert list(vals) == sorted(vals, reverse=True)


def _path_sums(dd):
    """Objective vectors of all root->terminal paths of a small flattened DD,
    with multiplicity."""
    a = dd.arrays()
    lo, io = a["layer_off"], a["in_off"]
    L, d = dd.num_layers, dd.d
    w = a["arc_w"].reshape(-1, d)
    sums = {0: [tuple([0] * d)]}
    for l in range(1, L):
        nxt = {}
        for v in range(lo[l + 1] - lo[l]):
            g = lo[l] + v
            acc = []
            for e in range(io[g], io[g + 1]):
                for p in sums.get(int(a["arc_src"][e]), []):
                    acc.append(tuple(int(x) + int(y) for x, y in zip(p, w[e])))
            nxt[v] = acc
        sums = nxt
    return sorted(sums[0])


def test_setpacking_bdd_paths_are_exactly_the_independent_sets(tmp_path):
    """Every independent set of the conflict graph is one path (and nothing
    else is), whatever vertex order the min-in-state rule picks."""
    inst = gen.setpacking(31, 13, 2, m=4)
    path = gen.write(inst, str(tmp_path / "s.dat"))
    dd = DecisionDiagram.from_file(path, 2)
    n = inst["n"]
    conflict = set()
    for c in inst["constraints"]:
        for u in c:
            for v in c:
                if u != v:
                    conflict.add((u, v))
    want = []
    for mask in range(1 << n):
        chosen = [v for v in range(n) if mask >> v & 1]
        if any((u, v) in conflict for u in chosen for v in chosen):
            continue
        want.append(tuple(sum(inst["objs"][o][v] for v in chosen) for o in range(2)))
    assert _path_sums(dd) == sorted(want)


def _completions(dd):
    """For every node (global index), the set of arc-label sequences from it
    to the terminal, for a small flattened diagram."""
    a = dd.arrays()
    lo, io = a["layer_off"], a["in_off"]
    L = dd.num_layers
    comp = {int(lo[L - 1]) + v: {()} for v in range(lo[L] - lo[L - 1])}
    for l in range(L - 1, 0, -1):
        parents = {}
        for v in range(lo[l + 1] - lo[l]):
            g = int(lo[l]) + v
            for e in range(io[g], io[g + 1]):
                p = int(lo[l - 1]) + int(a["arc_src"][e])
                parents.setdefault(p, set()).update(
                    (int(a["arc_val"][e]),) + c for c in comp[g])
        for p in range(int(lo[l - 1]), int(lo[l])):
            comp[p] = parents.get(p, set())
    return comp


def test_setpacking_partners_have_every_completion(tmp_path):
    """A set-packing state-dominance partner can complete every path its node
    can (its eligible set contains the node's), which is what makes dropping
    the node's dominated labels sound."""
    pairs = 0
    for seed in range(41, 47):
        inst = gen.setpacking(seed, 14, 2, m=4)
        path = gen.write(inst, str(tmp_path / f"s{seed}.dat"))
        dd = DecisionDiagram.from_file(path, 2, state_dominance=True)
        a = dd.arrays()
        lo = a["layer_off"]
        partner = a["sd_partner"]
        comp = _completions(dd)
        for l in range(dd.num_layers):
            for v in range(lo[l + 1] - lo[l]):
                q = int(partner[lo[l] + v])
                if q >= 0:
                    assert q != v
                    assert comp[int(lo[l]) + v] <= comp[int(lo[l]) + q], (seed, l, v, q)
                    pairs += 1
        # without the flag, no partners are recorded
        assert "sd_partner" not in DecisionDiagram.from_file(path, 2).arrays()
    assert pairs > 0


def _first_occurrence_ordered(a, num_layers):
    """Child ids of every layer follow the order in which a scan of the
    parents (ascending, 0-child then 1-child) first reaches them."""
    lo, io = a["layer_off"], a["in_off"]
    for l in range(1, num_layers):
        n_par = int(lo[l] - lo[l - 1])
        heads = np.repeat(np.arange(int(lo[l + 1] - lo[l])),
                          np.diff(io[int(lo[l]):int(lo[l + 1]) + 1]))
        e0, e1 = int(io[lo[l]]), int(io[lo[l + 1]])
        # position of each arc in the scan: parent-major, 0-arc before 1-arc
        pos = a["arc_src"][e0:e1].astype(np.int64) * 2 + a["arc_val"][e0:e1]
        assert pos.max() < 2 * n_par
        seq = heads[np.argsort(pos, kind="stable")]
        first = np.unique(seq, return_index=True)[1]
        if not np.array_equal(seq[np.sort(first)], np.arange(len(first))):
            return False
    return True


def test_setpacking_wide_layers_same_on_any_thread_count(tmp_path, monkeypatch):
    """Layers of 65536+ states are built on host threads (histogram, child
    generation, deduplication by hash partition, partner search) and the
    diagram is flattened on threads; the arrays must equal the one-thread
    build's, with node ids in first-occurrence order."""
    inst = gen.setpacking(5, 130, 2, m=26)
    path = gen.write(inst, str(tmp_path / "wide.dat"))
    for sd in (False, True):
        built = {}
        # "0", "-3" and "abc" are ignored: the machine's threads, at most 8
        for threads in ("1", "8") + (("0", "-3", "abc") if not sd else ()):
            monkeypatch.setenv("PMDD_HOST_THREADS", threads)
            dd = DecisionDiagram.from_file(path, 2, state_dominance=sd)
            if threads == "1":
                assert max(dd.layer_sizes()) >= 65536  # the threaded path is taken
            built[threads] = dd.arrays()
        one = built["1"]
        for threads, many in built.items():
            assert sorted(one) == sorted(many)
            for k in one:
                assert np.array_equal(one[k], many[k]), (sd, threads, k)
        many = built["8"]
        assert _first_occurrence_ordered(many, len(many["layer_off"]) - 1)


def test_setpacking_first_vertex_and_terminal(tmp_path):
    inst = gen.setpacking(3, 16, 3, m=4)
    path = gen.write(inst, str(tmp_path / "s.dat"))
    dd = DecisionDiagram.from_file(path, 2)
    sizes = dd.layer_sizes()
    assert sizes[0] == 1 and sizes[-1] == 1 and len(sizes) == 17
    a = dd.arrays()
    # min-in-state picks vertex 0 first: its 1-arc carries objs[.][0]
    first = [i for i in range(a["in_off"][a["layer_off"][2]]) if a["arc_val"][i] == 1]
    assert list(a["arc_w"][first[0]]) == [inst["objs"][o][0] for o in range(3)]


@pytest.mark.parametrize("content,problem,msg", [
    ("3 1\n", 1, "malformed"),
    ("2 1 1\n5 6\n", 1, "truncated"),
    ("3 1\n2\n1 2 3\n4 5 6\n2 1 9\n", 2, "out of range"),
    ("2 70\n", 3, "out of range"),
])
def test_parser_errors(tmp_path, content, problem, msg):
    p = tmp_path / "bad.dat"
    p.write_text(content)
    with pytest.raises(abi.PmddError) as e:
        DecisionDiagram.from_file(str(p), problem)
    assert msg in e.value.message


def test_missing_file_message():
    with pytest.raises(abi.PmddError) as e:
        DecisionDiagram.from_file("/nonexistent/x.dat", 1)
    assert "could not be found" in e.value.message


def test_expected_objective_count_is_enforced():
    with pytest.raises(abi.PmddError) as e:
        DecisionDiagram.from_file(instance_path("KP_p-3_n-10_ins-1.dat"), 1, expect_d=4)
    assert "expects 4" in e.value.message


def _cli(*args):
    return subprocess.run([abi.CLI_BIN, *args], capture_output=True, text=True)


@pytest.mark.parametrize("args,msg", [
    (["f", "4", "1", "0"], "Error - invalid problem type '4'"),
    (["f", "1", "1", "0", "--backend", "cuda"], "backend token 'cuda' is unsupported"),
    (["f", "1", "1", "0", "gpu", "3"], "gpu shorthand does not accept a kernel value"),
    (["f", "1", "1", "0", "--backend", "gpu", "--cpu-threads", "2"],
     "cpu thread options are not valid with backend=gpu"),
    (["f", "1", "1", "0", "--max-cand", "10"], "--max-cand is only valid with backend=gpu"),
    (["f", "1", "1", "0", "gpu", "--max-cand", "12X"], "invalid --max-cand value '12X'"),
    (["f", "1", "4", "0", "gpu"], "GPU backend is unsupported for method 4."),
    (["f", "1", "1", "0", "gpu", "--bogus"], "unrecognized optional argument '--bogus'"),
    (["f", "1", "1", "0", "gpu", "--backend", "gpu"],
     "cannot mix --backend with shorthand backend token"),
])
def test_cli_argument_errors(args, msg):
    r = _cli(*args)
    assert r.returncode == 1
    assert msg in r.stdout
    assert "Usage:" in r.stdout


def test_cli_cpu_backend_is_refused():
    r = _cli(instance_path("KP_p-3_n-10_ins-1.dat"), "1", "1", "0")
    assert r.returncode == 1
    assert "CPU backend is not part of this build" in r.stdout


def test_cli_without_device_fails_like_the_reference():
    if abi.engine_lib().pmdd_device_count() > 0:
        pytest.skip("a device is visible here")
    r = _cli(instance_path("KP_p-3_n-10_ins-1.dat"), "1", "1", "0", "gpu")
    assert r.returncode == 1
    assert r.stdout.startswith("Error - GPU backend requested but top-down enumeration failed: ")


def test_generator_is_deterministic_and_in_format(tmp_path):
    a = gen.to_text(gen.knapsack(42, 20, 3))
    b = gen.to_text(gen.knapsack(42, 20, 3))
    assert a == b
    inst = gen.knapsack(42, 20, 3)
    assert inst["capacity"] == sum(inst["weights"]) // 2
    assert all(1 <= x <= 1000 for x in inst["weights"])
    sp = gen.setpacking(1, 60, 4)
    covered = set(v for c in sp["constraints"] for v in c)
    assert covered == set(range(60)) and sp["m"] == 12
    assert all(5 <= len(c) <= 16 for c in sp["constraints"])
    t = gen.tsp(2, 7, 3)
    assert all(t["dist"][o][i][i] == 0 for o in range(3) for i in range(7))
    # round trip through the reader
    for k, inst in enumerate([inst, sp, t]):
        p = gen.write(inst, str(tmp_path / f"{k}.dat"))
        dd = DecisionDiagram.from_file(p, inst["kind"])
        assert dd.d == inst["d"]


def test_cli_program_name_fixes_objective_count(tmp_path):
    link = tmp_path / "multiobj_nobjs4"
    os.symlink(abi.CLI_BIN, link)
    r = subprocess.run([str(link), instance_path("KP_p-3_n-10_ins-1.dat"), "1", "1", "0", "gpu"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "expects 4" in r.stdout


@pytest.mark.parametrize("d", [1, 2, 3, 5, 8, 9])
def test_frontier_sort_matches_lexsort(d):
    """pmdd_frontier_sort orders rows as ParetoFrontier::sort_lexicographic_ascending
    (pareto_frontier.hpp:418-433): lexicographically ascending, duplicates kept."""
    import ctypes as C
    from pmdd_amd import abi
    rng = np.random.default_rng(100 + d)
    pts = rng.integers(-5, 6, size=(3000, d)).astype(np.int32)  # many ties
    want = pts[np.lexsort(pts.T[::-1])]
    buf = np.ascontiguousarray(pts.copy())
    fr = abi.Frontier()
    fr.d = d
    fr.count = len(buf)
    fr.points = buf.ctypes.data_as(C.POINTER(C.c_int32))
    assert abi.host_lib().pmdd_frontier_sort(C.byref(fr)) == abi.PMDD_OK
    assert (buf == want).all()


def test_paper_instance_families_build():
    """Every instance of the reference paper's families under
    tests/golden/paper/ (the data of `bench.py --paper`) parses and builds
    with its family's objective count."""
    import os
    base = os.path.join(os.path.dirname(__file__), "golden", "paper")
    kinds = {"mokp": 1, "mospp": 2, "motsp": 3}
    fams = sorted(os.listdir(base))
    assert len(fams) >= 9
    for fam in fams:
        kind = kinds[fam.split("_")[0]]
        k = int(fam.split("_k")[-1])
        files = sorted(os.listdir(os.path.join(base, fam)))
        assert len(files) == 10, fam
        for f in files:
            dd = DecisionDiagram.from_file(os.path.join(base, fam, f), kind, expect_d=k)
            assert dd.d == k and dd.num_layers > 1, (fam, f)
