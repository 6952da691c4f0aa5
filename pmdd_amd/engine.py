"""Python mirror of the engine's C ABI (libpmdd_engine.so).

``Engine.enumerate`` plays the role of the reference's
MultiobjEnum::pareto_frontier_{topdown,dynamic_layer_cutset}_cuda
(src/pmdd/enum/multiobj_enum.cpp:72-124): it takes a decision diagram, a
method (1 top-down, 2 bottom-up, 3 dynamic layer cutset), the problem type and state
dominance flag, plus the --max-cand / --max-prod budgets, and returns the
frontier.  max_prod defaults to adaptive coupling batches (0); a positive
value caps every batch, as the reference's --max-prod.  Errors raise PmddError carrying the engine's code and message
(the reference returns NULL with a reason string instead).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .abi import PmddError


@dataclass
class EnumerationResult:
    points: np.ndarray  # (count, d) int32, lexicographically ascending
    stats: dict = field(default_factory=dict)


def _stats_dict(st: abi.Stats, nl: int, std_c, std_s) -> dict:
    out = {name: getattr(st, name) for name, _ in abi.Stats._fields_
           if name not in ("std_candidates_per_layer", "std_frontier_survivors_per_layer",
                           "std_capacity")}
    if std_c is not None:
        out["std_candidates_per_layer"] = list(std_c)
        out["std_frontier_survivors_per_layer"] = list(std_s)
    return out


def _take_frontier(lib, fr: abi.Frontier, sort: bool = False) -> np.ndarray:
    """The points of a returned frontier as an (n, d) array (freeing it),
    lexicographically ascending when sort is set (pmdd_frontier_sort)."""
    n, d = int(fr.count), int(fr.d)
    if sort and n > 1:
        rc = abi.host_lib().pmdd_frontier_sort(C.byref(fr))
        if rc != abi.PMDD_OK:
            lib.pmdd_frontier_free(C.byref(fr))
            raise PmddError(rc, "host out of memory sorting the frontier")
    if n > 0:
        arr = np.ctypeslib.as_array(fr.points, shape=(n * d,)).copy().reshape(n, d)
    else:
        arr = np.zeros((0, max(d, 1)), dtype=np.int32)
    lib.pmdd_frontier_free(C.byref(fr))
    return arr


class Engine:
    """A per-device engine context (pmdd_ctx)."""

    def __init__(self, device: int = 0):
        lib = abi.engine_lib()
        self._lib = lib
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib.pmdd_ctx_create(int(device), C.byref(h), err, len(err))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, err.value.decode())
        self._h = h
        self.device = device

    @staticmethod
    def device_count() -> int:
        return int(abi.engine_lib().pmdd_device_count())

    @staticmethod
    def version() -> str:
        return abi.engine_lib().pmdd_engine_version().decode()

    def _opts(self, method, problem_type, state_dominance, max_cand, max_prod, count_pair_tests,
              phase_timing=False):
        o = abi.Opts()
        o.method = int(method)
        o.problem_type = int(problem_type)
        o.state_dominance = int(state_dominance)
        o.count_pair_tests = 1 if count_pair_tests else 0
        o.max_cand = int(max_cand)
        o.max_prod = int(max_prod)
        o.phase_timing = 1 if phase_timing else 0
        return o

    def enumerate(self, dd, method: int = 1, state_dominance: int = 0,
                  max_cand: int = abi.DEFAULT_MAX_CAND, max_prod: int = abi.MAX_PROD_ADAPTIVE,
                  count_pair_tests: bool = False, sort: bool = True,
                  phase_timing: bool = False) -> EnumerationResult:
        """phase_timing=True fills the kernel_*_s statistics (HIP events around
        the phases of every layer); they are 0 otherwise."""
        o = self._opts(method, dd.problem_type, state_dominance, max_cand, max_prod,
                       count_pair_tests, phase_timing)
        fr = abi.Frontier()
        st = abi.Stats()
        nl = dd.num_layers
        std_c = (C.c_double * nl)()
        std_s = (C.c_double * nl)()
        st.std_capacity = nl
        st.std_candidates_per_layer = std_c
        st.std_frontier_survivors_per_layer = std_s
        err = C.create_string_buffer(1024)
        rc = self._lib.pmdd_enumerate(self._h, dd.view_ptr, C.byref(o), C.byref(fr),
                                      C.byref(st), err, len(err))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, err.value.decode())
        pts = _take_frontier(self._lib, fr, sort)
        return EnumerationResult(pts, _stats_dict(st, nl, std_c, std_s))

    def get_tuning(self, d: int) -> dict:
        """Launch shapes this context uses for objective count d."""
        t = abi.Tuning()
        rc = self._lib.pmdd_ctx_get_tuning(self._h, int(d), C.byref(t))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, f"no tuning for d={d}")
        return {"filter_threads": list(t.filter_threads), "cross_workgroups": t.cross_workgroups,
                "cross_split_max": t.cross_split_max, "refill_waves": t.refill_waves,
                "source": t.source}

    def set_tuning(self, d: int, cfg: dict) -> None:
        t = abi.Tuning()
        for k in range(3):
            t.filter_threads[k] = int(cfg["filter_threads"][k])
        t.cross_workgroups = int(cfg["cross_workgroups"])
        t.cross_split_max = int(cfg["cross_split_max"])
        t.refill_waves = int(cfg["refill_waves"])
        rc = self._lib.pmdd_ctx_set_tuning(self._h, int(d), C.byref(t))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, f"invalid tuning {cfg}")

    def upload(self, dd) -> "ResidentDD":
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = self._lib.pmdd_upload(self._h, dd.view_ptr, C.byref(h), err, len(err))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, err.value.decode())
        return ResidentDD(self, h, dd.problem_type, dd.num_layers, dd.d)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            self._lib.pmdd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class ResidentDD:
    """A diagram resident in HBM (pmdd_resident); enumerations reuse it."""

    def __init__(self, engine: Engine, handle: C.c_void_p, problem_type: int, num_layers: int,
                 d: int):
        self.engine = engine
        self._h = handle
        self.problem_type = problem_type
        self.num_layers = num_layers
        self.d = d

    def enumerate(self, method: int = 1, state_dominance: int = 0,
                  max_cand: int = abi.DEFAULT_MAX_CAND, max_prod: int = abi.MAX_PROD_ADAPTIVE,
                  count_pair_tests: bool = False, sort: bool = True,
                  phase_timing: bool = False) -> EnumerationResult:
        eng = self.engine
        o = eng._opts(method, self.problem_type, state_dominance, max_cand, max_prod,
                      count_pair_tests, phase_timing)
        fr = abi.Frontier()
        st = abi.Stats()
        err = C.create_string_buffer(1024)
        rc = eng._lib.pmdd_enumerate_resident(eng._h, self._h, C.byref(o), C.byref(fr),
                                              C.byref(st), err, len(err))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, err.value.decode())
        pts = _take_frontier(eng._lib, fr, sort)
        return EnumerationResult(pts, _stats_dict(st, self.num_layers, None, None))

    def enumerate_noalloc(self, capacity: int, method: int = 1, state_dominance: int = 0,
                          max_cand: int = abi.DEFAULT_MAX_CAND,
                          max_prod: int = abi.MAX_PROD_ADAPTIVE):
        """pmdd_enumerate_resident_noalloc: the same enumeration with every
        device buffer taken from the engine's static scratch array.  Returns
        (points sorted lexicographically, stats); PmddError with code
        PMDD_E_INVALID when the frontier holds more than `capacity` points
        (its size is then the error's `count`)."""
        eng = self.engine
        o = eng._opts(method, self.problem_type, state_dominance, max_cand, max_prod, False)
        buf = np.zeros(max(capacity, 1) * self.d, dtype=np.int32)
        cnt = C.c_int64(0)
        st = abi.Stats()
        err = C.create_string_buffer(1024)
        rc = eng._lib.pmdd_enumerate_resident_noalloc(
            eng._h, self._h, C.byref(o), buf.ctypes.data_as(C.POINTER(C.c_int32)),
            int(capacity), C.byref(cnt), C.byref(st), err, len(err))
        if rc != abi.PMDD_OK:
            e = PmddError(rc, err.value.decode())
            e.count = int(cnt.value)
            raise e
        n = int(cnt.value)
        pts = buf[:n * self.d].reshape(n, self.d)
        if n > 1:
            pts = pts[np.lexsort(pts.T[::-1])]
        return pts, _stats_dict(st, self.num_layers, None, None)

    def close(self) -> None:
        if self._h is not None:
            self.engine._lib.pmdd_resident_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
