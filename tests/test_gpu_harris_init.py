"""The Harris-sheet initial state built on the device (mhd_init_harris)
against the host builder (dasha_amd.harris with arrays): the same
arithmetic, so every field agrees to the last bits of the device's and
numpy's tanh / cosh / exp (relative 1e-15 of each field's scale; the noise
factors are integer-exact); two steps from either start agree to 1e-13,
and the full 512x256x512 build takes a fraction of the host's time."""
import time

import numpy as np
import pytest

from dasha_amd import harris
from tests.util import field_errors

pytestmark = pytest.mark.gpu

CASES = {
    "periodic_stretched": lambda a: harris.case_small(40, 12, 36, arrays=a, seed=11),
    "nonperiodic_uniform_guide": lambda a: harris.case_small(48, 20, 40, stretched=False, periodic_y=False,
                                                             guide=0.1, x_offset=0.3, arrays=a, seed=5),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_state_matches_host(name):
    from dasha_amd.mhd import MHDStep
    host = CASES[name](True)
    dev = CASES[name](False)
    assert dev.x0 is None and dev.harris == host.harris
    print(f"[harris-init] {name}", flush=True)
    g = MHDStep(dev)
    x = g.get_state()
    for m in range(8):
        scale = max(np.abs(host.x0[m]).max(), 1e-300)
        assert np.abs(x[m] - host.x0[m]).max() <= 1e-15 * scale, m
    # the equilibrium enters the step: two steps from either start
    h = MHDStep(host)
    g.run(2)
    h.run(2)
    e = field_errors(g.get_state(), h.get_state())
    print("2-step errors", e)
    assert max(e) < 1e-13
    g.close()
    h.close()


def test_full_size_build_time():
    from dasha_amd.mhd import MHDStep
    t0 = time.perf_counter()
    c = harris.case_c3(arrays=False)
    g = MHDStep(c)
    dt = g.compute_dt()
    t_dev = time.perf_counter() - t0
    print(f"[harris-init] 512x256x512 on the device: {t_dev:.2f} s, dt {dt!r}", flush=True)
    x = g.get_state()
    # physical points (the padded ghost columns stay zero)
    assert np.isfinite(x).all() and x[0, 2:-2, :, 2:-2].min() > 0
    g.close()
    assert t_dev < 10.0
