"""The Fortran host program's mesh generator (mesh_x, mesh_yz in
dasha_amd/fortran/mhd_io.f90, gridpntnew2 of t3dsdsc.f:4380-4690) on the CPU,
through a small test program built with flang.  Pins:

* uniform axes: the closed form lo + (j-1) (hi - lo) / gaps with the
  reference's default-real extents (-2.1, 5.65 as single-precision values),
  y and z mirrored to [-ymax, ymax] and [zmin0, zmin0 + 2 zmax];
* graded y/z axes (both branches of the fine band): the mirror image about
  the origin bit for bit (the centre point the image -0.0 of the first),
  the end points at +-hi, the gaps summing to the points, the constant gaps
  of the band (dmax = dmin branch) and the small step down to the tanh rise
  outside it;
* the half axis with an even count: the reversed, negated axis with the
  last point the image of point n-2;
* the graded dayside x mesh (1441 points): -30 and 80 at the ends, the
  second segment symmetric about its middle bit for bit, the first
  segment's gaps falling towards x = 0, continuity at the segment joins,
  the last spacing 0 as the reference leaves it; other counts are refused.
"""
import os
import subprocess

import numpy as np
import pytest

from dasha_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IO = os.path.join(ROOT, "dasha_amd", "fortran", "mhd_io.f90")

F32 = lambda v: float(np.float32(v))   # a default-real Fortran literal as stored

_PROG = """
program mesh_check
  use iso_c_binding
  use mhd_io
  implicit none
  real(c_double) :: px(201), dxv(201), py(51), dyv(51), pz(105), dzv(105)
  real(c_double) :: gy(51), gdy(51), by(41), bdy(41), hy(50), hdy(50)
  real(c_double) :: gx(1441), gdx(1441), bad(1001), badd(1001)
  integer :: ios, ios2
  ! mhd03.for's defaults
  call mesh_x(201, dble(-2.1), dble(-2.1), dble(5.65), .true., .false., px, dxv, ios)
  call mesh_yz(51, 0.d0, 2.5d0, -2.5d0, .true., .false., dble(0.01), 1.d0, 0.d0, py, dyv)
  call mesh_yz(105, 0.d0, 2.5d0, -2.5d0, .true., .false., dble(0.01), 1.d0, 0.d0, pz, dzv)
  ! graded y: dmax /= dmin (fine band graded) and dmax = dmin with a band
  call mesh_yz(51, 0.d0, 2.5d0, -2.5d0, .false., .false., dble(0.01), 1.d0, 0.3d0, gy, gdy)
  call mesh_yz(41, 0.d0, 3.d0, -3.d0, .false., .false., 0.05d0, 0.05d0, 0.5d0, by, bdy)
  ! half axis, even count
  call mesh_yz(50, 0.d0, 2.d0, -2.d0, .true., .true., 0.01d0, 1.d0, 0.d0, hy, hdy)
  ! graded dayside x, and a count it does not take
  call mesh_x(1441, 0.d0, 0.d0, 0.d0, .false., .false., gx, gdx, ios)
  call mesh_x(1001, 0.d0, 0.d0, 0.d0, .false., .false., bad, badd, ios2)
  call dump('u.raw', [px, dxv, py, dyv, pz, dzv])
  call dump('g.raw', [gy, gdy, by, bdy, hy, hdy])
  call dump('x.raw', [gx, gdx, dble(ios), dble(ios2)])
contains
  subroutine dump(name, v)
    character(len=*), intent(in) :: name
    real(c_double), intent(in) :: v(:)
    integer :: w
    open(newunit=w, file=name, access='stream', form='unformatted', status='replace')
    write(w) v
    close(w)
  end subroutine dump
end program mesh_check
"""


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    fc = build.FLANG if os.path.exists(build.FLANG) else None
    if fc is None:
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("mesh")
    (d / "mesh_check.f90").write_text(_PROG)
    exe = d / "mesh_check"
    subprocess.run([fc, "-O2", "-J", str(d), IO, str(d / "mesh_check.f90"), "-o", str(exe)], check=True,
                   cwd=str(d))
    subprocess.run([str(exe)], check=True, cwd=str(d))
    raw = lambda n: np.fromfile(d / n, dtype="<f8")
    u, g, x = raw("u.raw"), raw("g.raw"), raw("x.raw")

    def split(v, sizes):
        out, o = [], 0
        for s in sizes:
            out.append(v[o:o + s])
            o += s
        return out

    return {"u": split(u, [201, 201, 51, 51, 105, 105]), "g": split(g, [51, 51, 41, 41, 50, 50]),
            "x": split(x, [1441, 1441, 1, 1])}


def test_uniform_closed_forms(run):
    px, dx, py, dy, pz, dz = run["u"]
    lo, hi = F32(-2.1), F32(5.65)
    step = (hi - lo) / 200.0
    assert np.array_equal(dx, np.full(201, step))
    assert np.array_equal(px, lo + np.arange(201) * step)
    sy = 2.5 / 25.0
    assert np.array_equal(py, -2.5 + np.arange(51) * sy) and np.array_equal(dy, np.full(51, sy))
    sz = 2.5 / 52.0
    assert np.array_equal(pz, -2.5 + np.arange(105) * sz) and np.array_equal(dz, np.full(105, sz))
    assert abs(py[-1] - 2.5) < 1e-15 and abs(pz[-1] - 2.5) < 1e-14


def _mirror_checks(p, d, hi):
    n = p.size
    c = (n - 1) // 2
    # mirror image about the origin, bit for bit; the centre is -0.0
    assert p[c] == 0.0 and np.signbit(p[c])
    assert np.array_equal(p[:c], -p[:c:-1])
    assert abs(p[-1] - hi) < 1e-13 * hi and abs(p[0] + hi) < 1e-13 * hi
    # the upper half: the gaps are the point differences (to rounding)
    up, gaps = p[c:], d[c:c + (n - 1) // 2]
    assert np.allclose(np.diff(up), gaps, rtol=1e-12, atol=1e-15)
    # the mirrored spacing array is the upper one reversed, shifted by one
    # index as the reference's mirror loop writes it
    assert np.array_equal(d[:c + 1], d[c:][::-1])
    return up, gaps


def test_graded_axis_fine_band_graded(run):
    gy, gdy = run["g"][0], run["g"][1]
    up, gaps = _mirror_checks(gy, gdy, 2.5)
    # the band (|y| < 0.3) is finer than the rest, the gaps rise beyond it
    inner = gaps[up[:-1] < 0.3 - 1e-12]
    outer = gaps[up[:-1] >= 0.3]
    assert inner.size >= 2 and outer.size >= 2 and inner.max() < outer.min()
    assert np.all(np.diff(outer) > 0)


def test_graded_axis_constant_band(run):
    by, bdy = run["g"][2], run["g"][3]
    up, gaps = _mirror_checks(by, bdy, 3.0)
    # dmax = dmin: nhr - 1 equal gaps from the origin (int(0.5/0.05 + 1) = 11),
    # then a strictly rising tanh profile
    nhr = int((0.5 - 0.0) / 0.05 + 1)
    assert np.all(gaps[:nhr - 1] == gaps[0])
    assert np.all(np.diff(gaps[nhr - 1:]) > 0)
    # the band's gaps sit at tanh(-width/2) of the rise's profile, the rise
    # starts at tanh(-2 width/3): the join steps down slightly
    assert gaps[nhr - 1] < gaps[0] < gaps[nhr - 1] * 1.001


def test_half_axis_even_count(run):
    hy, hdy = run["g"][4], run["g"][5]
    # the 49 computed points 0..2 reversed and negated, point 50 the image
    # of point 48 (the reference's yy(my1) = -yy(nny))
    step = 2.0 / 48.0
    base = 0.0 + np.arange(49) * step
    assert np.array_equal(hy[:49], -base[::-1])
    assert hy[49] == -hy[47]
    assert np.array_equal(hdy[:49], np.full(49, step)) and hdy[49] == hdy[47]


def test_dayside_x_mesh(run):
    gx, gdx, ios, ios2 = run["x"]
    assert ios[0] == 0 and ios2[0] == 1
    assert gx[0] == -30.0 and abs(gx[-1] - 80.0) < 1e-11
    assert gdx[-1] == 0.0 and np.all(gdx[:-1] > 0)
    assert np.allclose(np.diff(gx), gdx[:-1], rtol=1e-12, atol=1e-14)
    # segment 2 (gaps 401..1000) symmetric about its middle, bit for bit
    s2 = gdx[400:1000]
    assert np.array_equal(s2, s2[::-1])
    # segment 1 falls towards the nose; the joins are continuous to rounding
    assert np.all(np.diff(gdx[:400]) < 0)
    assert abs(gdx[400] - gdx[399]) < 1e-14 and abs(gdx[1000] - gdx[999]) < 1e-14
    # the finest gaps (0.02 + the common widening) sit at the joins
    assert gdx[:1440].min() == pytest.approx(gdx[399], rel=1e-12)
