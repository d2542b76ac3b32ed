"""Parity of the step's alternative kernel paths against the CPU oracle.

The library picks its kernels per handle at creation (environment knobs
read by mhd_create): the point-parallel one-pass sweeps with a separate RK
update and post-step filter (MHD_ZS=0, MHD_FUSE_RK=0), and the measured
defaults swapped for their alternatives (two-pass instead of three-pass last
avrgh launch, grid-order block tables, 64-plane chunks, the windowed right
part 2 on small volumes, an unpadded field stride, the side-stream launches
on small volumes).  Each must give the
oracle's step like the default path.
"""
import os

import pytest

from dasha_amd import harris
from tests.util import field_errors

pytestmark = pytest.mark.gpu

KNOBS = {
    "one_pass_sweeps": {"MHD_ZS": "0", "MHD_FUSE_RK": "0"},
    "alternatives": {"MHD_ZS3": "0", "MHD_ZORD": "0", "MHD_ZKC": "64", "MHD_RB2": "1", "MHD_PAD": "0"},
    # the side-stream launches (on by default from 2^22 points) on small volumes
    "side_streams": {"MHD_OVERLAP": "1", "MHD_RB2": "1"},
}
CASES = {
    "wide_nonperiodic_hall": lambda: harris.case_small(130, 30, 40, periodic_y=False, seed=17),
    "stretched_periodic_hall": lambda: harris.case_small(24, 8, 24),
    # several windows in x and in periodic y, chunks ending inside the volume
    "multiwindow_periodic_hall": lambda: harris.case_small(130, 36, 72, seed=5),
}


@pytest.mark.parametrize("knobs", sorted(KNOBS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_path_matches_oracle(knobs, name):
    from dasha_amd.mhd import MHDStep
    from oracle.oracle import Oracle
    case = CASES[name]()
    print(f"[paths] {knobs} {name}", flush=True)
    saved = {k: os.environ.get(k) for k in KNOBS[knobs]}
    os.environ.update(KNOBS[knobs])
    try:
        g = MHDStep(case)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    o = Oracle(case)
    g.run(3)
    o.run(3)
    e = field_errors(g.get_state(), o.get_state())
    print("3 step errs", e)
    assert max(e) < 1e-11
    g.close()
