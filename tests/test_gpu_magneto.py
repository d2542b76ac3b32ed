"""Magnetosphere mode (the reference's own default case) of the HIP step
against the oracle, through the C ABI: the frozen residual, one right() with
the driven inflow at both stage kinds, the current, setdt over r >= 1, three
steps from t = 0 and from the middle of the drive pulse (t = 20), the
diagnostics, and the device-dt loop against the host loop.  Three switch
sets: the reference's defaults with the Hall term on (mask, inflow, radial
Hall, totalE), the
totalE = .false. dipole terms without Hall, and a run without the inflow.
The third one also takes setdt's whistler term.  Tolerances as in
test_gpu_parity: 1e-12 relative L2 per field for one call, 1e-11 after
three steps."""
import numpy as np
import pytest

from dasha_amd import magneto
from tests.util import field_errors

pytestmark = pytest.mark.gpu

TIGHT = 1e-12

CASES = {
    "defaults": lambda: magneto.case_small(face=magneto.synthetic_face((13, 25)), di=0.2, hall=True),
    "partial_e": lambda: magneto.case_small(face=magneto.synthetic_face((13, 25)), hall=False,
                                            partial_e=True),
    # a uniform Hall coefficient needs the whistler speed in setdt to stay stable
    "no_inflow": lambda: magneto.case_small(29, 15, 21, inflow=False, radial_hall=False, di=0.15,
                                            hall=True, whistler_dt=True),
}


def _pair(case):
    from oracle.oracle import Oracle
    from dasha_amd.mhd import MHDStep
    print(f"[magneto] case={case.name} inflow={case.inflow} radial_hall={case.radial_hall} "
          f"partial_e={case.partial_e} hall={case.hall}", flush=True)
    return MHDStep(case), Oracle(case)


@pytest.mark.parametrize("name", sorted(CASES))
def test_residual_current_dt(name):
    g, o = _pair(CASES[name]())
    g.init_residual()
    o.init_residual()
    e = field_errors(g.get_residual(), o.get_residual())
    print("xresd", e)
    assert max(e) < TIGHT
    o.current()
    e = field_errors(g.get_current(), o.get_current())
    print("current", e)
    assert max(e) < TIGHT
    dg, do = g.compute_dt(), o.compute_dt()
    print("dt", dg, do)
    assert dg == do
    assert g.compute_dt_noalloc() == do
    # the state after the capture: its driven columns were overwritten too
    e = field_errors(g.get_state(), o.get_state())
    assert max(e) < TIGHT


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("km", [1, 2])
def test_rhs_with_drive(name, km):
    g, o = _pair(CASES[name]())
    for s in (g, o):
        s.init_residual()
        s.rhs_stage(20.0, 1e-3, km)
    e = field_errors(g.get_xdif(), o.get_xdif())
    print(f"xdif km={km}", e)
    assert max(e) < TIGHT


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("t0", [0.0, 20.0])
def test_three_steps(name, t0):
    g, o = _pair(CASES[name]())
    tg, dg = g.run(3, time=t0)
    to, do = o.run(3, time=t0)
    e = field_errors(g.get_state(), o.get_state())
    print(f"3 steps from t={t0}", e, dg, do)
    assert dg == do
    assert max(e) < 1e-11
    e = field_errors(g.get_current(), o.get_current())
    assert max(e) < 1e-11


def test_diagnostics_and_device_loop():
    from dasha_amd.mhd import MHDStep
    c = CASES["defaults"]()
    g, o = _pair(c)
    g.run(2, time=20.0)
    o.run(2, time=20.0)
    d = g.diagnostics()
    od, oe = o.diagn(), o.energy()
    got = [d["cxmax"], d["cxmin"], d["cymax"], d["cymin"], d["czmax"], d["czmin"]]
    assert np.allclose(got, od, rtol=1e-11, atol=1e-15)
    assert np.allclose([d["e_kin"], d["e_mag"], d["e_th"], d["e_mag_pert"]], oe, rtol=1e-11)
    dv = g.get_divb()
    ref = o.divb()
    assert np.linalg.norm(dv - ref) <= 1e-11 * np.linalg.norm(ref)
    # mhd_advance (setdt each step on the device) == the host loop
    a, b = MHDStep(c), MHDStep(c)
    a.init_residual()
    b.init_residual()
    ta, dla, _ = a.advance(4, 20.0)
    tb, dts = b.run(4, time=20.0, init=False)
    assert ta == tb and dla == dts[-1]
    assert np.array_equal(a.get_state(), b.get_state())


def test_magnetosphere_mode_refusals():
    """non-periodic y only; no ensemble or adjoint of this mode"""
    from dasha_amd.mhd import MHDEnsemble, MHDError, MHDStep
    c = CASES["defaults"]()
    g = MHDStep(c)
    with pytest.raises(MHDError):
        MHDEnsemble([g])
    with pytest.raises(MHDError):
        g.adjoint_run([1e-4], "magnetic_energy")
    bad = CASES["defaults"]()
    bad.periodic_y = True
    with pytest.raises(MHDError):
        MHDStep(bad)


@pytest.mark.parametrize("name", sorted(CASES))
def test_fill_forms_agree(name, monkeypatch):
    """the fills with the mask applied by the producers and the x, y and z
    rules in one launch (two for bndrytt) give the same state, bit for bit,
    as a mask launch per fill and separate x, y and z launches (the
    reference's routine order)"""
    from dasha_amd.mhd import MHDStep
    states = []
    for mask, split in (("1", "0"), ("0", "1")):
        monkeypatch.setenv("MHD_MAGMASK", mask)
        monkeypatch.setenv("MHD_FILLSPLIT", split)
        print(f"[magneto] {name} producer mask={mask} split fills={split}", flush=True)
        g = MHDStep(CASES[name]())
        g.run(3, time=20.0)
        states.append((g.get_state(), g.get_current()))
        g.close()
    assert np.array_equal(states[0][0], states[1][0])
    assert np.array_equal(states[0][1], states[1][1])
