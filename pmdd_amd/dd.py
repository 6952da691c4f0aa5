"""Decision diagrams built by the host library (libpmdd_host.so).

``DecisionDiagram.from_file`` runs the reference main()'s construction
sequence for a problem type (src/pmdd/main.cpp:113-199): knapsack BDD built
exact, reduced and given min_weight; set packing as an unreduced
independent-set BDD; TSP as an exact MDD.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

PROBLEM_KNAPSACK = 1
PROBLEM_SETPACKING = 2
PROBLEM_TSP = 3


class DecisionDiagram:
    """Owns a pmdd_dd; exposes its flattened view as numpy arrays (copies)."""

    def __init__(self, handle: C.c_void_p, problem_type: int):
        self._h = handle
        self.problem_type = problem_type
        lib = abi.host_lib()
        self._view = lib.pmdd_dd_get_view(self._h)
        v = self._view.contents
        self.d = int(v.d)
        self.num_layers = int(v.num_layers)
        self.kind = int(v.kind)

    @classmethod
    def from_file(cls, path: str, problem_type: int, expect_d: int = 0,
                  state_dominance: bool = False) -> "DecisionDiagram":
        """state_dominance also records what the set-packing state-dominance
        rule needs (node partners), which only the construction knows."""
        lib = abi.host_lib()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        flags = abi.PMDD_BUILD_STATE_DOMINANCE if state_dominance else 0
        rc = lib.pmdd_dd_build_from_file(str(path).encode(), int(problem_type), int(expect_d),
                                         flags, C.byref(h), err, len(err))
        if rc != abi.PMDD_OK:
            raise abi.PmddError(rc, err.value.decode())
        return cls(h, problem_type)

    @property
    def view_ptr(self) -> "C.POINTER(abi.DDView)":
        if self._h is None:
            raise ValueError("decision diagram already freed")
        return self._view

    def structure(self) -> abi.DDStructure:
        s = abi.DDStructure()
        abi.host_lib().pmdd_dd_get_structure(self._h, C.byref(s))
        return s

    def layer_sizes(self) -> list[int]:
        lib = abi.host_lib()
        return [int(lib.pmdd_dd_layer_size(self._h, l)) for l in range(self.num_layers)]

    def arrays(self) -> dict:
        """Numpy copies of the view arrays (for host-side checks)."""
        v = self._view.contents
        L = self.num_layers
        layer_off = np.ctypeslib.as_array(v.layer_off, shape=(L + 1,)).copy()
        total = int(layer_off[-1])
        in_off = np.ctypeslib.as_array(v.in_off, shape=(total + 1,)).copy()
        A = int(in_off[-1])
        out = {
            "layer_off": layer_off,
            "in_off": in_off,
            "arc_src": np.ctypeslib.as_array(v.arc_src, shape=(A,)).copy() if A else
            np.zeros(0, np.int32),
            "arc_val": np.ctypeslib.as_array(v.arc_val, shape=(A,)).copy() if A else
            np.zeros(0, np.int32),
            "arc_w": (np.ctypeslib.as_array(v.arc_w, shape=(A * self.d,)).copy()
                      if A else np.zeros(0, np.int32)).reshape(A, self.d),
        }
        if bool(v.min_weight):
            out["min_weight"] = np.ctypeslib.as_array(v.min_weight, shape=(total,)).copy()
        if bool(v.sd_partner):
            out["sd_partner"] = np.ctypeslib.as_array(v.sd_partner, shape=(total,)).copy()
        return out

    def close(self) -> None:
        if self._h is not None:
            abi.host_lib().pmdd_dd_free(self._h)
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
