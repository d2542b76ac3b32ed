"""The Fortran host program's I/O and set-up module (dasha_amd/fortran/mhd_io.f90)
on the CPU, through a small test program built with flang:

* inp/bin* reader (t3dsdsc.f:674-726): a file written y row by y row, x column
  by x column, z fastest, with values that encode their (jx, jy, jz) lands on
  the padded point (jx+2, jy, jz+2) of the field, ghosts zero;
* fcinp/bin* reader (t3dsdsc.f:728-768): y slowest, z fastest, times -50;
* initia (t3dsdsc.f:771-814): the program's dipole state equals the Python
  case builder's to rounding, and with mxdat the field is the inp field times
  the single-precision -25/3 exactly;
* recrd2 / recrd3 plot files (t3dsdsc.f:5116-5159, 5365-5408): the formatted
  lines byte for byte against lines written here from the Fortran rules for
  Ew.dEe and Ew.d (five E14.5E3 or eleven E10.4 values per line, each after
  one blank; E10.4 drops the optional leading zero of a negative value and
  the letter E of a three-digit exponent), and the list-directed header,
  time and field-label lines by their tokens.
"""
import os
import subprocess

import numpy as np
import pytest

from dasha_amd import build, magneto, rundir

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IO = os.path.join(ROOT, "dasha_amd", "fortran", "mhd_io.f90")

_PROG = """
program io_check
  use iso_c_binding
  use mhd_io
  implicit none
  integer, parameter :: mx = MX_, my = MY_, mz = MZ_
  real(c_double) :: xx(mx+4), yy(my), zz(mz+4), a(mx+4,my,mz+4), f(my,mz+4)
  real(c_double) :: bin(mx+4,my,mz+4,3), x(mx+4,my,mz+4,8)
  real(c_double), dimension(mx+4,my,mz+4) :: rhoi, bxi, byi, bzi, pri
  real(c_double) :: bl2(2,1,11), bl3(2,1,1,11)
  integer :: ios, u, i
  open(newunit=u, file='grid.raw', access='stream', form='unformatted', status='old')
  read(u) xx, yy, zz
  close(u)
  call read_inp_field('inp/binx000', mx, my, mz, a, ios)
  if (ios /= 0) stop 2
  call dump('inp.raw', a, size(a))
  call read_fcinp_field('fcinp/binx000', my, mz, f, ios)
  if (ios /= 0) stop 3
  call dump('fc.raw', f, size(f))
  bin = 0.d0
  call magnetosphere_initia(mx, my, mz, xx, yy, zz, .false., bin, x, rhoi, bxi, byi, bzi, pri)
  call dump('dipole.raw', x, size(x))
  bin(:,:,:,1) = a; bin(:,:,:,2) = 2.d0 * a; bin(:,:,:,3) = -a
  call magnetosphere_initia(mx, my, mz, xx, yy, zz, .true., bin, x, rhoi, bxi, byi, bzi, pri)
  call dump('mxdat.raw', x, size(x))
  do i = 1, 11
     bl2(1,1,i) = dble(i)
     bl2(2,1,i) = -0.25d0 * dble(i)
     bl3(1,1,1,i) = dble(i)
     bl3(2,1,1,i) = -0.25d0 * dble(i)
  end do
  bl2(1,1,1) = 0.d0; bl2(2,1,1) = 1.d-100
  bl3(1,1,1,1) = 1.d100; bl3(2,1,1,1) = 1.d-100
  call write_m2d('m2d.dat', 1.5d0, 2, 1, [1.d0, -0.5d0], [12345.678d0], bl2)
  call write_m3d('m3d.dat', 1.5d0, 2, 1, 1, [1.d0, -0.5d0], [0.d0], [12345.678d0], bl3)
contains
  subroutine dump(name, v, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(c_double), intent(in) :: v(n)
    integer :: w
    open(newunit=w, file=name, access='stream', form='unformatted', status='replace')
    write(w) v
    close(w)
  end subroutine dump
end program io_check
"""

MX, MY, MZ = 7, 5, 6


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    fc = build.FLANG if os.path.exists(build.FLANG) else None
    if fc is None:
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("fio")
    c = magneto.case_small(MX, MY, MZ, xmin=-2.05, xmax=1.3, ymin=-1.2, ymax=1.4, zmin=-1.1, zmax=1.5)
    with open(d / "grid.raw", "wb") as f:
        f.write(np.ascontiguousarray(c.xx).tobytes() + np.ascontiguousarray(c.yy).tobytes() +
                np.ascontiguousarray(c.zz).tobytes())
    os.makedirs(d / "inp")
    os.makedirs(d / "fcinp")
    # values that name their point: 1e4 jy + 100 jx + jz (1-based physical)
    with open(d / "inp" / "binx000", "w") as f:
        for jy in range(1, MY + 1):
            for jx in range(1, MX + 1):
                f.write(" ".join(str(10000 * jy + 100 * jx + jz) for jz in range(1, MZ + 1)) + "\n")
    with open(d / "fcinp" / "binx000", "w") as f:
        for jy in range(1, MY + 1):
            f.write(" ".join(str(0.5 + 100 * jy + jz) for jz in range(1, MZ + 1)) + "\n")
    src = d / "io_check.f90"
    src.write_text(_PROG.replace("MX_", str(MX)).replace("MY_", str(MY)).replace("MZ_", str(MZ)))
    exe = d / "io_check"
    subprocess.run([fc, "-O2", "-J", str(d), IO, str(src), "-o", str(exe)], check=True, cwd=str(d))
    subprocess.run([str(exe)], check=True, cwd=str(d))
    return d, c


def _raw(d, name, shape):
    return np.fromfile(d / name, dtype="<f8").reshape(shape)


def test_inp_reader_layout(run):
    d, _ = run
    a = _raw(d, "inp.raw", (MZ + 4, MY, MX + 4))      # Fortran a(mx+4, my, mz+4)
    for jz in range(1, MZ + 1):
        for jy in range(1, MY + 1):
            for jx in range(1, MX + 1):
                assert a[jz + 1, jy - 1, jx + 1] == 10000 * jy + 100 * jx + jz
    inner = np.zeros(a.shape, bool)
    inner[2:MZ + 2, :, 2:MX + 2] = True
    assert np.all(a[~inner] == 0.0)
    # the Python reader of the same file agrees
    assert np.array_equal(rundir.read_inp_field(str(d / "inp" / "binx000"), MX, MY, MZ), a)


def test_fcinp_reader_layout_and_scale(run):
    d, _ = run
    f = _raw(d, "fc.raw", (MZ + 4, MY))                # Fortran f(my, mz+4)
    for jz in range(1, MZ + 1):
        for jy in range(1, MY + 1):
            assert f[jz + 1, jy - 1] == (0.5 + 100 * jy + jz) * -50.0
    assert np.all(f[:2] == 0.0) and np.all(f[-2:] == 0.0)
    assert np.array_equal(rundir.read_fcinp_field(str(d / "fcinp" / "binx000"), MY, MZ), f)


def test_initia_dipole_and_mxdat(run):
    d, c = run
    x = _raw(d, "dipole.raw", (8, MZ + 4, MY, MX + 4))
    np.testing.assert_allclose(x, c.x0, rtol=1e-14, atol=1e-300)
    assert np.abs(x[4:7]).max() > 1.0
    a = _raw(d, "inp.raw", (MZ + 4, MY, MX + 4))
    xm = _raw(d, "mxdat.raw", (8, MZ + 4, MY, MX + 4))
    s = np.float64(np.float32(-25.0) / np.float32(3.0))     # -25./3. in single precision
    inner = np.s_[2:MZ + 2, :, 2:MX + 2]
    assert np.array_equal(xm[4][inner], a[inner] * s)
    assert np.array_equal(xm[5][inner], (2.0 * a[inner]) * s)
    assert np.array_equal(xm[6][inner], (-a[inner]) * s)
    np.testing.assert_allclose(xm[[0, 7]], c.x0[[0, 7]], rtol=1e-14)


E14 = {1.0: "  0.10000E+001", -0.5: " -0.50000E+000", 12345.678: "  0.12346E+005",
       0.0: "  0.00000E+000", 1e-100: "  0.10000E-099"}
E10 = {1.0: "0.1000E+01", -0.5: "-.5000E+00", 12345.678: "0.1235E+05", 0.0: "0.0000E+00",
       1e-100: "0.1000E-99", 1e100: "0.1000+101"}


def _lines(path):
    with open(path) as f:
        return f.read().split("\n")


def test_m2d_bytes(run):
    d, _ = run
    ln = _lines(d / "m2d.dat")
    assert ln[0].split() == ["2", "2", "1"]
    assert float(ln[1]) == 1.5
    assert ln[2] == "".join(" " + E14[v] for v in (1.0, -0.5, 12345.678))
    for i in range(1, 12):
        assert ln[1 + 2 * i].split() == [str(i)]
        vals = (0.0, 1e-100) if i == 1 else (float(i), -0.25 * i)
        want = "".join(" " + (E14[v] if v in E14 else "%14s" % _fmt_e14(v)) for v in vals)
        assert ln[2 + 2 * i] == want, i


def _fmt_e14(v):
    """Ew.dEe for the simple values i and -i/4 (i <= 11): 0.ddddd E+00e"""
    s = "-" if v < 0 else ""
    a = abs(v)
    e = 0
    while a >= 1.0:
        a /= 10.0
        e += 1
    digits = ("%.5f" % a)[1:]           # ".ddddd"
    return f"{s}0{digits}E+{e:03d}"


def test_m3d_bytes(run):
    d, _ = run
    ln = _lines(d / "m3d.dat")
    assert ln[0].split() == ["3", "2", "1", "1"]
    assert float(ln[1]) == 1.5
    assert ln[2] == "".join(" " + E10[v] for v in (1.0, -0.5, 0.0, 12345.678))
    assert ln[3].split() == ["1"]
    assert ln[4] == " " + E10[1e100] + " " + E10[1e-100]
    assert ln[5].split() == ["2"]
    assert ln[6] == " 0.2000E+01 -.5000E+00"
