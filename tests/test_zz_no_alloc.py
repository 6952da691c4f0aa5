"""The no-allocation entry point (pmdd_enumerate_resident_noalloc) against
the allocating one (pmdd_enumerate_resident), through the C ABI.

With the engine's allocation hook armed to fail the very next device
allocation, a no-allocation call must still succeed: it allocates nothing.
Two calls in a row, the second on a smaller diagram (the scratch array keeps
the first call's contents), methods 1 and 3; then an allocating call on the
same context still works, and a too-small output buffer is refused with the
frontier's size."""
import numpy as np
import pytest

from conftest import instance_path
from pmdd_amd import DecisionDiagram, abi
from pmdd_amd import instances as gen

pytestmark = pytest.mark.gpu


def _same(a, b):
    return a.shape == b.shape and bool((a == b).all())


def _noalloc(res, **kw):
    """A no-allocation call with every device allocation set to fail."""
    lib = abi.engine_lib()
    lib.pmdd_debug_fail_alloc(1)
    try:
        return res.enumerate_noalloc(**kw)
    finally:
        lib.pmdd_debug_fail_alloc(0)


@pytest.mark.parametrize("method", [1, 3])
def test_noalloc_matches_allocating_entry_point(engine, tmp_path, method):
    big = gen.write(gen.knapsack(41, 40, 4), str(tmp_path / "big.dat"))
    small = instance_path("KP_p-3_n-20_ins-1.dat")
    dd_big = DecisionDiagram.from_file(big, 1)
    dd_small = DecisionDiagram.from_file(small, 1)
    r_big = engine.upload(dd_big)
    r_small = engine.upload(dd_small)
    try:
        ref_big = r_big.enumerate(method=method).points
        ref_small = r_small.enumerate(method=method).points
        print(f"case noalloc method={method} big={len(ref_big)} small={len(ref_small)}",
              flush=True)
        got_big, st = _noalloc(r_big, capacity=len(ref_big) + 10, method=method)
        assert _same(got_big, ref_big)
        assert st["mem_need_peak_bytes"] <= abi.NOALLOC_SCRATCH_BYTES
        got_small, _ = _noalloc(r_small, capacity=len(ref_small), method=method)
        assert _same(got_small, ref_small)
        # the context still serves allocating calls
        assert _same(r_big.enumerate(method=method).points, ref_big)
        # and the no-allocation call again after them
        got_big2, _ = _noalloc(r_big, capacity=len(ref_big), method=method)
        assert _same(got_big2, ref_big)
        with pytest.raises(abi.PmddError) as e:
            _noalloc(r_small, capacity=len(ref_small) - 1, method=method)
        assert e.value.code == abi.PMDD_E_INVALID
        assert e.value.count == len(ref_small)
    finally:
        r_big.close()
        r_small.close()


def test_noalloc_on_a_fresh_context(tmp_path):
    """A context that never allocated: every buffer comes from the scratch."""
    from pmdd_amd.engine import Engine
    path = gen.write(gen.tsp(77, 9, 3), str(tmp_path / "t.dat"))
    dd = DecisionDiagram.from_file(path, 3)
    with Engine(0) as ref_eng:
        ref = ref_eng.enumerate(dd, method=3).points
    with Engine(0) as eng:
        r = eng.upload(dd)
        for method in (3, 1):
            got, _ = _noalloc(r, capacity=len(ref), method=method)
            assert _same(got, ref), method
        r.close()
