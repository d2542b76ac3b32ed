"""Run-directory files in the reference's formats, for the Fortran driver.

* restart files ``mhd<nst:3><rank:3>``: Fortran unformatted sequential,
  record 1 = ``ncase, nstep, time, nst`` (int32, int32, float64, int32),
  record 2 = ``x(mxpr, my, mzpr, 8)`` float64 (t3dsdsc.f:4840-4861 recrd,
  4814-4836 readin, 5416-5436 cn);
* ``grid.bin``: one record each of xx, yy, zz, dx, dy, dz (the arrays the
  reference builds in input/gridpntnew2, t3dsdsc.f:248-313);
* ``equil.bin``: one record each of rhoi, bxi, byi, bzi, pri (initia,
  t3dsdsc.f:806-814);
* ``mhd.nml``: namelist /run/ with the parameters the reference takes from
  mhd01.for / mhd03.for and the main-loop constants (t3dsdsc.f:88-92);
* ``inp/binx000`` .. ``binz000`` (initia's mxdat branch, t3dsdsc.f:674-726):
  list-directed values, z fastest, then x, then y, over the physical points
  (rank 000 of a single domain); ``fcinp/binx000`` .. ``binz000`` (mxdatfc,
  t3dsdsc.f:728-768): z fastest, then y, the inflow face data before the
  reference's -50 scaling.

Record markers are 4-byte little-endian lengths (the flang and gfortran
default).
"""
from __future__ import annotations

import os
import re
import struct

import numpy as np


def _write_record(f, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def _read_record(f) -> bytes:
    (n,) = struct.unpack("<i", f.read(4))
    data = f.read(n)
    (m,) = struct.unpack("<i", f.read(4))
    if m != n:
        raise ValueError("corrupt Fortran record")
    return data


def cn(n: int) -> str:
    """three-digit file number (t3dsdsc.f:5416-5436)"""
    if not 0 <= n <= 999:
        raise ValueError("file numbers run 0..999")
    return f"{n:03d}"


def write_restart(path: str, x: np.ndarray, *, ncase: int, nstep: int, time: float, nst: int):
    with open(path, "wb") as f:
        _write_record(f, struct.pack("<iidi", ncase, nstep, time, nst))
        _write_record(f, np.ascontiguousarray(x, dtype="<f8").tobytes())


def read_restart(path: str, shape):
    with open(path, "rb") as f:
        ncase, nstep, time, nst = struct.unpack("<iidi", _read_record(f))
        x = np.frombuffer(_read_record(f), dtype="<f8").reshape(shape).copy()
    return dict(ncase=ncase, nstep=nstep, time=time, nst=nst), x


def write_inp_field(path: str, a: np.ndarray, mx: int, my: int, mz: int):
    """one inp/bin* file from a padded (mz+4, my, mx+4) array: for each y row,
    for each x column, the z column's physical values"""
    with open(path, "w") as f:
        for jy in range(my):
            for jx in range(mx):
                col = a[2:mz + 2, jy, 2 + jx]
                f.write(" ".join("%.17g" % v for v in col) + "\n")


def read_inp_field(path: str, mx: int, my: int, mz: int) -> np.ndarray:
    """the padded (mz+4, my, mx+4) array of one inp/bin* file (zero ghosts)"""
    vals = np.array(open(path).read().split(), dtype=np.float64)[:mx * my * mz]
    out = np.zeros((mz + 4, my, mx + 4))
    out[2:mz + 2, :, 2:mx + 2] = vals.reshape(my, mx, mz).transpose(2, 0, 1)
    return out


def write_fcinp_field(path: str, f: np.ndarray, my: int, mz: int):
    """one fcinp/bin* file from unscaled face data (mz+4, my): for each y
    row, the physical z values"""
    with open(path, "w") as fh:
        for jy in range(my):
            fh.write(" ".join("%.17g" % v for v in f[2:mz + 2, jy]) + "\n")


def read_fcinp_field(path: str, my: int, mz: int) -> np.ndarray:
    """bxinfc(my, mz+4) as a C (mz+4, my) array: the file's values times -50
    (t3dsdsc.f:754-765), zero on the ghost planes"""
    vals = np.array(open(path).read().split(), dtype=np.float64)[:my * mz]
    out = np.zeros((mz + 4, my))
    out[2:mz + 2, :] = (vals.reshape(my, mz) * -50.0).T
    return out


def write_case(rundir: str, case, *, nstop: int, nst: int = 1, ncase: int = 1, dtime: float = 1.0,
               dnstep: int = 10, nadj: int = 0, adjobj: int = 0, initia: bool = False,
               mxdat: bool = False, mxdatfc: bool = False, mesh: dict | None = None):
    """Write grid.bin, equil.bin, mhd.nml and the initial restart file
    (nadj > 0: the driver differentiates nadj steps instead of running).
    For a magnetosphere case the namelist carries its switches; initia: the
    driver builds the initial state itself (mxdat: from inp/bin*), mxdatfc:
    it reads the inflow face data from fcinp/bin*.  mesh: the driver builds
    the mesh itself (gridgen) from these namelist entries (gridpntnew2's
    switches and extents: uniformx, halfx, xmin0, xmin, xmax, ymax, dymin,
    ...; booleans or floats) instead of reading grid.bin."""
    os.makedirs(rundir, exist_ok=True)
    with open(os.path.join(rundir, "grid.bin"), "wb") as f:
        for a in (case.xx, case.yy, case.zz, case.dx, case.dy, case.dz):
            _write_record(f, np.ascontiguousarray(a, dtype="<f8").tobytes())
    with open(os.path.join(rundir, "equil.bin"), "wb") as f:
        for a in (case.rhoi, case.bxi, case.byi, case.bzi, case.pri):
            _write_record(f, np.ascontiguousarray(a, dtype="<f8").tobytes())
    b = lambda v: ".true." if v else ".false."
    with open(os.path.join(rundir, "mhd.nml"), "w") as f:
        f.write("&run\n")
        f.write(f" mx={case.mx}, my={case.my}, mz={case.mz},\n")
        f.write(f" periody={b(case.periodic_y)}, hall={b(case.hall)}, whistler={b(case.whistler_dt)},\n")
        f.write(f" di={case.di!r}, eta0={case.eta0!r}, fmu0={case.fmu0!r}, frho0={case.frho0!r}, "
                f"fpr0={case.fpr0!r},\n")
        f.write(f" gamma={case.gamma!r}, cfl={case.cfl!r},\n")
        f.write(f" nstop={nstop}, nst={nst}, ncase={ncase}, dtime={dtime!r}, dnstep={dnstep},\n")
        f.write(f" nadj={nadj}, adjobj={adjobj}")
        if getattr(case, "magnetosphere", False):
            f.write(f",\n magneto=.true., initia={b(initia)}, mxdat={b(mxdat)}, mxdatfc={b(mxdatfc)},\n")
            f.write(f" inflow={b(case.inflow)}, radial_hall={b(case.radial_hall)}, "
                    f"partial_e={b(case.partial_e)}, rcr={case.rcr!r}")
        if mesh is not None:
            f.write(",\n gridgen=.true.")
            for k, v in mesh.items():
                f.write(f", {k}={b(v) if isinstance(v, bool) else repr(float(v))}")
        f.write("\n/\n")
    write_restart(os.path.join(rundir, "mhd" + cn(nst) + cn(0)), case.x0, ncase=ncase, nstep=0,
                  time=0.0, nst=nst)


def read_adjoint(path: str, shape3):
    """adjoint.bin of the driver: (value, d value/d eta0, d value/d x) with
    the gradient as a C array (8,) + shape3"""
    with open(path, "rb") as f:
        recs = [_read_record(f), _read_record(f)]
    val, geta = np.frombuffer(recs[0], dtype="<f8")
    gx = np.frombuffer(recs[1], dtype="<f8").reshape((8,) + tuple(shape3))
    return float(val), float(geta), gx


def read_plot(path: str):
    """A recrd2 (m2d) or recrd3 (m3d) plot file (t3dsdsc.f:5117-5157,
    5366-5401): returns (time, coordinate arrays, fields) with fields of shape
    (11, mz, mx) or (11, mz, my, mx), C order: rho, vx, vy, vz, bx, by, bz, p,
    jx, jy, jz."""
    with open(path) as f:
        tok = f.read().split()
    # Ew.d drops the E of a three-digit exponent (0.1234-100)
    fl = lambda t: float(re.sub(r"(?<=[0-9.])([+-])(?=[0-9]+$)", r"E\1", t))
    ndim = int(tok[0])
    dims = [int(t) for t in tok[1:1 + ndim]]
    pos = 1 + ndim
    time = fl(tok[pos])
    pos += 1
    coords = []
    for n in dims:
        coords.append(np.array([fl(t) for t in tok[pos:pos + n]]))
        pos += n
    npts = int(np.prod(dims))
    shape = tuple(reversed(dims))
    fields = np.empty((11,) + shape)
    for i in range(11):
        if int(tok[pos]) != i + 1:
            raise ValueError(f"{path}: field label {tok[pos]} where {i + 1} was expected")
        pos += 1
        fields[i] = np.array([fl(t) for t in tok[pos:pos + npts]]).reshape(shape)
        pos += npts
    return time, coords, fields

