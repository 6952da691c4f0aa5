"""ORACLE -- test infrastructure only.

ctypes binding of oracle/_build/liboracle.so (the CPU restatement of the
reference's OpenMP enumerators and the brute-force enumerators, oracle.cpp).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
as the checker and as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


class OracleStats(C.Structure):
    _fields_ = [
        ("work_candidates_total", C.c_int64),
        ("work_candidates_peak", C.c_int64),
        ("work_frontier_survivors_total", C.c_int64),
        ("work_frontier_peak_points", C.c_int64),
        ("work_join_products_total", C.c_int64),
        ("dominance_filtered_total", C.c_int64),
        ("layer_coupling", C.c_int32),
        ("reserved", C.c_int32),
        ("max_node_frontier", C.c_int64),
        ("max_node_candidates", C.c_int64),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing; run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.oracle_enumerate.argtypes = [
            C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
            C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.c_int64), C.POINTER(OracleStats),
            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_enumerate.restype = C.c_int
        L.oracle_bruteforce_file.argtypes = [
            C.c_char_p, C.c_int, C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.c_int64),
            C.POINTER(C.c_int)]
        L.oracle_bruteforce_file.restype = C.c_int
        L.oracle_nondominated.argtypes = [
            C.POINTER(C.c_int32), C.c_int64, C.c_int, C.POINTER(C.POINTER(C.c_int32)),
            C.POINTER(C.c_int64)]
        L.oracle_nondominated.restype = C.c_int
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_free.restype = None
        _lib = L
    return _lib


def _take(ptr, count: int, d: int) -> np.ndarray:
    if count:
        arr = np.ctypeslib.as_array(ptr, shape=(count * d,)).copy().reshape(count, d)
    else:
        arr = np.zeros((0, d), dtype=np.int32)
    lib().oracle_free(C.cast(ptr, C.c_void_p))
    return arr


def sort_lex(points: np.ndarray) -> np.ndarray:
    """Lexicographic ascending order, as main() writes the frontier
    (sort_lexicographic_ascending, pareto_frontier.hpp:418-433)."""
    if len(points) == 0:
        return points
    order = np.lexsort(points.T[::-1])
    return points[order]


def enumerate_dd(dd, method: int, state_dominance: int = 0, threads: int = 1,
                 score_rule: int = 1, with_stats: bool = False):
    """Run the CPU restatement on a pmdd_amd.dd.DecisionDiagram."""
    L = lib()
    pts = C.POINTER(C.c_int32)()
    cnt = C.c_int64()
    st = OracleStats()
    nl = dd.num_layers
    std_c = (C.c_double * nl)()
    std_s = (C.c_double * nl)()
    rc = L.oracle_enumerate(C.cast(dd.view_ptr, C.c_void_p), method, dd.problem_type,
                            state_dominance, threads, score_rule, C.byref(pts), C.byref(cnt),
                            C.byref(st), std_c, std_s)
    if rc != 0:
        raise RuntimeError(f"oracle_enumerate failed: {rc}")
    arr = sort_lex(_take(pts, cnt.value, dd.d))
    if with_stats:
        return arr, st, list(std_c), list(std_s)
    return arr


def bruteforce(path: str, problem_type: int) -> np.ndarray:
    L = lib()
    pts = C.POINTER(C.c_int32)()
    cnt = C.c_int64()
    d = C.c_int()
    rc = L.oracle_bruteforce_file(str(path).encode(), problem_type, C.byref(pts), C.byref(cnt),
                                  C.byref(d))
    if rc != 0:
        raise RuntimeError(f"oracle_bruteforce_file failed: {rc} ({path})")
    return sort_lex(_take(pts, cnt.value, d.value))


def nondominated(points: np.ndarray) -> np.ndarray:
    L = lib()
    p = np.ascontiguousarray(points, dtype=np.int32)
    n, d = p.shape
    out = C.POINTER(C.c_int32)()
    cnt = C.c_int64()
    rc = L.oracle_nondominated(p.ctypes.data_as(C.POINTER(C.c_int32)), n, d, C.byref(out),
                               C.byref(cnt))
    if rc != 0:
        raise RuntimeError("oracle_nondominated failed")
    return sort_lex(_take(out, cnt.value, d))
