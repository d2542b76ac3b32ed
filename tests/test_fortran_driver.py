"""The Fortran host program (dasha_amd/fortran/mhd_driver.f90) over the C ABI.

CPU tests cover the run-directory formats and the driver's input side: on a
machine without a GPU the driver reads mhd.nml, grid.bin, equil.bin and the
restart record, then stops at mhd_create with its "mhd_create failed"
message (any format mismatch stops it earlier with a runtime I/O error).
The GPU test runs the driver for 12 steps and requires its final state to
be bit-identical to the same setdt/stepon sequence driven from Python
through the same library, with the fort.20 and current.dat cadence of
t3dsdsc.f:171-181 and 1174-1182; a second one checks the m2d/m3d plot
files of an output step against the state and current at that step; a third
runs the driver's adjoint mode (mhd_adjoint_run called from Fortran) and
requires the Python call's value and gradients bit for bit; the last two
run the magnetosphere mode from the driver's own initial state, and from
the mesh the driver builds itself (gridgen).
"""
import os
import subprocess

import numpy as np
import pytest

from dasha_amd import harris, rundir

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "dasha_amd", "fortran", "mhd_driver")
LIB = os.path.join(ROOT, "dasha_amd", "libmhd_gpu.so")


def _driver():
    if not os.path.exists(DRIVER):
        if not os.path.exists(LIB):
            pytest.skip("libmhd_gpu.so not built")
        from dasha_amd import build
        if build.build_fortran() is None:
            pytest.skip("no Fortran compiler")
    return DRIVER


def test_restart_record_roundtrip(tmp_path):
    x = np.random.default_rng(0).standard_normal((8, 6, 5, 7))
    p = str(tmp_path / ("mhd" + rundir.cn(3) + rundir.cn(0)))
    rundir.write_restart(p, x, ncase=2, nstep=17, time=1.25, nst=3)
    assert os.path.basename(p) == "mhd003000"
    # record 1 is 20 bytes between 4-byte markers, record 2 the raw array
    assert os.path.getsize(p) == 4 + 20 + 4 + 4 + x.nbytes + 4
    hdr, y = rundir.read_restart(p, x.shape)
    assert hdr == dict(ncase=2, nstep=17, time=1.25, nst=3)
    assert np.array_equal(x, y)


def test_cn_range():
    assert rundir.cn(0) == "000" and rundir.cn(999) == "999"
    with pytest.raises(ValueError):
        rundir.cn(1000)


def test_driver_reads_run_directory(tmp_path):
    drv = _driver()
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present: covered by the GPU test")
    except ImportError:
        pass
    case = harris.case_small(12, 4, 12)
    rundir.write_case(str(tmp_path), case, nstop=3)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([drv], cwd=str(tmp_path), capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "mhd_create failed" in r.stderr, r.stdout + r.stderr


@pytest.mark.gpu
def test_driver_matches_python_loop(tmp_path):
    from dasha_amd.mhd import MHDStep
    drv = _driver()
    case = harris.case_small(24, 8, 24)
    nstop = 11                       # the loop exits once nstep > nstop: 12 steps
    rundir.write_case(str(tmp_path), case, nstop=nstop, dtime=1e9)
    print(f"[driver] case={case.name} steps={nstop + 1}", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "110", drv], cwd=str(tmp_path),
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    hdr, xf = rundir.read_restart(str(tmp_path / "mhdfinal"), case.x0.shape)
    assert hdr["nstep"] == nstop + 1

    g = MHDStep(case)
    t, dts = g.run(nstop + 1)
    xg = g.get_state()
    # x(mx+4, my, mz+4, 8) in Fortran order is the C array [8][mz+4][my][mx+4]
    assert np.array_equal(xf, xg)
    assert hdr["time"] == t

    lines = (tmp_path / "fort.20").read_text().split("\n")
    rows = [ln.split() for ln in lines if ln.strip()]
    assert [int(r_[2]) for r_ in rows] == list(range(1, nstop + 2))
    assert np.allclose([float(r_[0]) for r_ in rows], dts, atol=6e-6)
    cur = [ln.split() for ln in (tmp_path / "current.dat").read_text().split("\n") if ln.strip()]
    assert len(cur) == 2                              # nstep 0 and 10
    assert all(len(c) == 7 for c in cur)


@pytest.mark.gpu
def test_driver_plot_files(tmp_path):
    """recrd2 / recrd3 at an output step (dtime 0, dnstep 1: the cadence of
    t3dsdsc.f:185-199 fires after the first step, nst 1 -> 2)"""
    from dasha_amd.mhd import MHDStep
    drv = _driver()
    case = harris.case_small(24, 8, 24, seed=5)
    rundir.write_case(str(tmp_path), case, nstop=2, dtime=0.0, dnstep=1)
    print(f"[driver] plot files, case={case.name}", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "110", drv], cwd=str(tmp_path),
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    g = MHDStep(case)
    t, _ = g.run(1)
    x, cur = g.get_state(), g.get_current()
    mx, mz = case.mx, case.mz
    # physical points of the padded arrays [m][mz+4][my][mx+4]
    phys = np.concatenate([x[:, 2:mz + 2, :, 2:mx + 2], cur[:, 2:mz + 2, :, 2:mx + 2]])
    phys[1:4] /= phys[0]
    t2, c2, f2 = rundir.read_plot(str(tmp_path / "data" / "m2d002.dat"))
    assert t2 == pytest.approx(t, rel=1e-14)
    # e14.5 keeps 5 significant digits (relative rounding <= 5e-5), e10.4
    # keeps 4 (<= 5e-4)
    r2, r3 = 5.01e-5, 5.01e-4
    for got, want in ((c2[0], case.xx[2:mx + 2]), (c2[1], case.zz[2:mz + 2])):
        assert np.allclose(got, want, rtol=r2, atol=0), (got[:6], want[:6])
    nyh = case.my // 2                     # Fortran jy = my/2 + 1
    assert np.allclose(f2, phys[:, :, nyh, :], rtol=r2, atol=0)
    t3, c3, f3 = rundir.read_plot(str(tmp_path / "data" / "m3d002"))
    assert len(c3) == 3 and np.allclose(c3[1], case.yy, rtol=r3, atol=0)
    assert np.allclose(f3, phys, rtol=r3, atol=0)
    assert not (tmp_path / ("mhd" + rundir.cn(2) + rundir.cn(0))).exists()   # recrd only when mod(nst,3) == 1



@pytest.mark.gpu
def test_driver_adjoint_matches_python(tmp_path):
    """nadj = 4: the driver records the dt sequence, restores x0 and calls
    mhd_adjoint_run; same numbers as the Python binding of the same call"""
    import dataclasses
    from dasha_amd import adjoint
    from dasha_amd.mhd import MHDStep
    drv = _driver()
    ac = adjoint.small_case(24, 8, 24, steps=4, seed=16)
    case = dataclasses.replace(ac.case, x0=ac.state())
    rundir.write_case(str(tmp_path), case, nstop=1, nadj=4, adjobj=1)
    print(f"[driver] adjoint, case={case.name}", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "110", drv], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    val, geta, gx = rundir.read_adjoint(str(tmp_path / "adjoint.bin"), case.shape3)
    dts = adjoint.dt_sequence(ac)
    g = MHDStep(case)
    try:
        g.init_residual()
        v2, gx2, ge2, _ = g.adjoint_run(dts, "reconnected_flux")
    finally:
        g.close()
    assert val == v2 and geta == ge2
    assert np.array_equal(gx, gx2)


@pytest.mark.gpu
def test_driver_magnetosphere_initia(tmp_path):
    """magnetosphere mode: the driver builds initia's state from inp/bin*
    (mxdat) and reads the inflow face data from fcinp/bin* (mxdatfc); 4 steps
    equal the Python binding run from the driver's own initial state (mhdinit)
    with the face data read back in Python, bit for bit"""
    import dataclasses
    from dasha_amd import magneto
    from dasha_amd.mhd import MHDStep
    drv = _driver()
    face = magneto.synthetic_face((13, 25))
    case = magneto.case_small(face=face, di=0.2, hall=True)
    s = float(np.float32(-25.0) / np.float32(3.0))
    os.makedirs(tmp_path / "inp")
    os.makedirs(tmp_path / "fcinp")
    for m, c in enumerate("xyz"):
        rundir.write_inp_field(str(tmp_path / "inp" / f"bin{c}000"), case.x0[4 + m] / s, case.mx, case.my,
                               case.mz)
        rundir.write_fcinp_field(str(tmp_path / "fcinp" / f"bin{c}000"), face[m] / -50.0, case.my, case.mz)
    nstop = 3
    rundir.write_case(str(tmp_path), case, nstop=nstop, dtime=1e9, initia=True, mxdat=True, mxdatfc=True)
    print(f"[driver] magnetosphere initia, case={case.name}", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "110", drv], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    _, x0 = rundir.read_restart(str(tmp_path / "mhdinit"), case.x0.shape)
    np.testing.assert_allclose(x0[4:7], case.x0[4:7], rtol=1e-13, atol=1e-300)
    hdr, xf = rundir.read_restart(str(tmp_path / "mhdfinal"), case.x0.shape)
    assert hdr["nstep"] == nstop + 1
    fc = tuple(rundir.read_fcinp_field(str(tmp_path / "fcinp" / f"bin{c}000"), case.my, case.mz) for c in "xyz")
    pc = dataclasses.replace(case, x0=x0, rhoi=x0[0].copy(), bxi=x0[4].copy(), byi=x0[5].copy(),
                             bzi=x0[6].copy(), pri=x0[7].copy(), inflow_face=fc)
    g = MHDStep(pc)
    t, _ = g.run(nstop + 1)
    assert np.array_equal(xf, g.get_state())
    assert hdr["time"] == t


@pytest.mark.gpu
def test_driver_builds_mesh(tmp_path):
    """gridgen: the driver builds the reference's uniform mesh itself
    (gridpntnew2 with mhd03.for's switches, the extents of the small
    magnetosphere) instead of reading grid.bin; the run equals the Python
    binding run on the Python-built mesh bit for bit, and gridx.dat /
    grid.dat hold the global mesh in the reference's formats"""
    import dataclasses
    from dasha_amd import magneto
    from dasha_amd.mhd import MHDStep
    drv = _driver()
    case = magneto.case_small()
    lo = float(case.xx[2])
    mesh = dict(uniformx=True, uniformy=True, uniformz=True, halfx=False, halfy=False, halfz=False,
                xmin0=lo, xmin=lo, xmax=2.5, ymin=0.0, ymax=1.5, zmin=0.0, zmax=1.5, zmin0=-1.5)
    nstop = 2
    rundir.write_case(str(tmp_path), case, nstop=nstop, dtime=1e9, initia=True, mesh=mesh)
    # a grid.bin the driver must not read
    os.remove(tmp_path / "grid.bin")
    print(f"[driver] gridgen, case={case.name}", flush=True)
    r = subprocess.run(["timeout", "-k", "10", "110", drv], cwd=str(tmp_path), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    _, x0 = rundir.read_restart(str(tmp_path / "mhdinit"), case.x0.shape)
    np.testing.assert_allclose(x0, case.x0, rtol=1e-13, atol=1e-300)
    hdr, xf = rundir.read_restart(str(tmp_path / "mhdfinal"), case.x0.shape)
    # the Python run from the driver's initia state on the Python-built mesh
    g = MHDStep(dataclasses.replace(case, x0=x0, rhoi=x0[0].copy(), bxi=x0[4].copy(), byi=x0[5].copy(),
                                    bzi=x0[6].copy(), pri=x0[7].copy()))
    t, _ = g.run(nstop + 1)
    assert np.array_equal(xf, g.get_state()) and hdr["time"] == t
    # gridx.dat: (jx, x, dx) twice per line; grid.dat: x/dx, z/dz, y/dy pairs
    tok = open(tmp_path / "gridx.dat").read().split()
    assert len(tok) == 3 * case.mx and int(tok[-3]) == case.mx
    assert np.allclose([float(t) for t in tok[1::3]], case.xx[2:-2], atol=6e-6)
    vals = np.array(open(tmp_path / "grid.dat").read().split(), dtype=float)
    want = np.concatenate([np.stack([case.xx[2:-2], case.dx[2:-2]], 1).ravel(),
                           np.stack([case.zz[2:-2], case.dz[2:-2]], 1).ravel(),
                           np.stack([case.yy, case.dy], 1).ravel()])
    assert vals.shape == want.shape and np.allclose(vals, want, atol=6e-6)
