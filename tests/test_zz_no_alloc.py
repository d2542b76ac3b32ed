"""The allocation-free setdt (mhd_compute_dt_noalloc) against mhd_compute_dt
through the C ABI: two calls in a row on different handles, the second on a
smaller grid, each after a step so the states differ; dt bit for bit, and
the handle's device dt slot left as mhd_compute_dt leaves it (the next
mhd_step/mhd_advance reads it)."""
import numpy as np
import pytest

from dasha_amd import harris

pytestmark = pytest.mark.gpu


def _pair(case):
    from dasha_amd.mhd import MHDStep
    a, b = MHDStep(case), MHDStep(case)
    for g in (a, b):
        g.init_residual()
    return a, b


def test_noalloc_dt_matches_compute_dt():
    big = harris.case_small(40, 16, 44, seed=31, di=0.2)
    small = harris.case_small(24, 8, 20, seed=32, di=0.0)
    print(f"[no-alloc] {big.name} then {small.name}", flush=True)
    for case in (big, small):
        ref, got = _pair(case)
        t = 0.0
        for s in range(2):
            d_ref = ref.compute_dt()
            d_got = got.compute_dt_noalloc()
            print(f"[no-alloc] {case.name} call {s + 1}: dt {d_ref!r} / {d_got!r}", flush=True)
            assert d_got == d_ref
            ref.step(t, d_ref)
            got.step(t, d_got)
            t += d_ref
        assert np.array_equal(ref.get_state(), got.get_state())
        ref.close()
        got.close()
