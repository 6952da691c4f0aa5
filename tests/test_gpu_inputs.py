"""Input checks of the engine's C ABI: a malformed diagram view is refused
with PMDD_E_INVALID when it is checked, before anything reaches the device."""
import ctypes as C

import numpy as np
import pytest

from pmdd_amd import DecisionDiagram, abi
from pmdd_amd import instances as gen

pytestmark = pytest.mark.gpu


def _same(a, b):
    return a.shape == b.shape and bool((a == b).all())


class _View:
    """A diagram's view with some int32 arrays replaced; keeps the diagram and
    the arrays alive."""

    def __init__(self, dd, **arrays):
        self._dd = dd
        self._view = abi.DDView.from_buffer_copy(bytes(dd.view_ptr.contents))
        self._arrays = {}
        for name, arr in arrays.items():
            arr = np.ascontiguousarray(arr, dtype=np.int32)
            self._arrays[name] = arr
            setattr(self._view, name, arr.ctypes.data_as(C.POINTER(C.c_int32)))
        self.view_ptr = C.pointer(self._view)
        self.problem_type = dd.problem_type
        self.num_layers = dd.num_layers


def test_partner_outside_its_layer_is_rejected(engine, tmp_path):
    """A partner index past its layer, or a node named as its own partner, is
    refused; the same partners unchanged give the diagram's own result."""
    p = gen.write(gen.setpacking(21, 40, 2, m=8), str(tmp_path / "s.dat"))
    dd = DecisionDiagram.from_file(p, 2, state_dominance=True)
    a = dd.arrays()
    lo, partner = a["layer_off"], a["sd_partner"]
    want = engine.enumerate(dd, method=1, state_dominance=1).points
    assert _same(engine.enumerate(_View(dd, sd_partner=partner), method=1,
                                  state_dominance=1).points, want)
    l = next(l for l in range(dd.num_layers) if lo[l + 1] - lo[l] >= 2)
    for bad in (0, int(lo[l + 1] - lo[l])):  # node 0 of layer l: itself, one past the layer
        q = partner.copy()
        q[lo[l]] = bad
        with pytest.raises(abi.PmddError) as e:
            engine.enumerate(_View(dd, sd_partner=q), method=1, state_dominance=1)
        assert e.value.code == abi.PMDD_E_INVALID and "partner" in e.value.message


def test_bdd_arc_value_other_than_0_or_1_is_rejected(engine, tmp_path):
    """A BDD arc is a 0-arc or a 1-arc; the same values unchanged give the
    diagram's own result."""
    p = gen.write(gen.knapsack(7, 12, 2), str(tmp_path / "k.dat"))
    dd = DecisionDiagram.from_file(p, 1)
    val = dd.arrays()["arc_val"]
    want = engine.enumerate(dd, method=1).points
    assert _same(engine.enumerate(_View(dd, arc_val=val), method=1).points, want)
    for bad in (2, -1):
        v = val.copy()
        v[len(v) // 2] = bad
        with pytest.raises(abi.PmddError) as e:
            engine.enumerate(_View(dd, arc_val=v), method=1)
        assert e.value.code == abi.PMDD_E_INVALID and "arc value" in e.value.message
