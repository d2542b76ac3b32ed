"""Harris-sheet inputs for the Hall-MHD step, in the reference's layout.

The reference (t3dsdsc.f) builds its mesh in ``gridpntnew2`` (t3dsdsc.f:4380-4690)
and its initial state in ``initia`` (t3dsdsc.f:612-850).  Its own case is a
dipole magnetosphere that needs input files it does not ship, so this module
builds the Harris-sheet cases the step is measured on, in the reference's
array conventions:

* padded coordinates ``xx(mx+4)``, ``zz(mz+4)`` with the physical points at
  Fortran indices 3..mx+2 and zero ghost entries (a single domain never fills
  them, t3dsdsc.f:296-313);
* spacing arrays ``dx`` holding the forward spacing, last entry repeated
  (t3dsdsc.f:4448-4451, 4633-4635);
* y stored without ghosts when non-periodic, and with two ghost rows per
  side (``my = ny + 4``) when periodic;
* 3-D arrays as C-ordered numpy arrays of shape ``(mz+4, my, mx+4)``, which
  is the memory order of the Fortran array ``a(mx+4, my, mz+4)``; the state
  is ``(8, mz+4, my, mx+4)`` = Fortran ``x(mxpr, my, mzpr, 8)``.

Noise is counter based (splitmix64 finaliser on seed and global index) so any
implementation can regenerate it independently of array decomposition.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
KEYMUL = np.uint64(0xD1B54A32D192ED03)


def splitmix_uniform(seed: int, index: np.ndarray) -> np.ndarray:
    """U(-1, 1) from the splitmix64 finaliser of (seed, index), vectorised."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * GOLDEN + (index.astype(np.uint64) + np.uint64(1)) * KEYMUL
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0


def uniform_points(n: int, lo: float, hi: float):
    """n points lo..hi inclusive; spacing array = the constant step
    (the reference's uniform branch, t3dsdsc.f:4389-4396)."""
    d = (hi - lo) / (n - 1)
    pts = lo + np.arange(n, dtype=np.float64) * d
    return pts, np.full(n, d)


def stretched_points(n: int, dmin: float, ratio: float):
    """Symmetric tanh-stretched mesh about 0 for even n: gap i (between
    points i and i+1) is dmin*(1 + (ratio-1)*tanh((i+0.5-(n-1)/2)/w)^2) with
    w = n/6, so the sheet centre z=0 sits mid-gap; spacing array = forward
    gaps with the last repeated."""
    if n % 2:
        raise ValueError("stretched mesh needs an even point count")
    w = n / 6.0
    i = np.arange(n - 1, dtype=np.float64)
    gaps = dmin * (1.0 + (ratio - 1.0) * np.tanh((i + 0.5 - (n - 1) / 2.0) / w) ** 2)
    h = n // 2
    pos = np.empty(h)
    pos[0] = 0.5 * gaps[h - 1]
    for k in range(1, h):
        pos[k] = pos[k - 1] + gaps[h - 1 + k]
    pts = np.concatenate([-pos[::-1], pos])
    sp = np.empty(n)
    sp[:-1] = pts[1:] - pts[:-1]
    sp[-1] = sp[-2]
    return pts, sp


def pad2(a: np.ndarray) -> np.ndarray:
    out = np.zeros(a.size + 4)
    out[2:-2] = a
    return out


@dataclasses.dataclass
class MHDCase:
    """One configured run: grid, parameters, equilibrium and initial state."""
    name: str
    mx: int
    my: int               # full y extent (ny + 4 when periodic)
    mz: int
    periodic_y: bool
    hall: bool
    di: float
    eta0: float
    gamma: float
    cfl: float
    whistler_dt: bool
    steps: int
    xx: np.ndarray
    yy: np.ndarray
    zz: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray
    rhoi: np.ndarray
    bxi: np.ndarray
    byi: np.ndarray
    bzi: np.ndarray
    pri: np.ndarray
    x0: np.ndarray
    fmu0: float = 0.0
    frho0: float = 0.0
    fpr0: float = 0.0
    # the Harris-sheet parameters (mhd_init_harris); with the 3-D arrays left
    # out (None) the device builds the equilibrium and the state itself
    harris: dict | None = None

    @property
    def ny(self) -> int:
        return self.my - 4 if self.periodic_y else self.my

    @property
    def cells(self) -> int:
        """physical cells updated per step"""
        return self.mx * self.ny * self.mz

    @property
    def shape3(self):
        return (self.mz + 4, self.my, self.mx + 4)

    def config_dict(self) -> dict:
        return dict(mx=self.mx, my=self.my, mz=self.mz, periodic_y=int(self.periodic_y),
                    hall=int(self.hall), whistler_dt=int(self.whistler_dt), di=self.di,
                    eta0=self.eta0, fmu0=self.fmu0, frho0=self.frho0, fpr0=self.fpr0,
                    gamma=self.gamma, cfl=self.cfl)


def harris_case(name: str, mx: int, ny: int, mz: int, *, stretched: bool, lx: float = 20.0,
                lz: float = 20.0, ly: float = 10.0, dmin: float = 0.05, ratio: float = 3.0,
                lam: float = 0.5, beta_inf: float = 0.2, guide: float = 0.0, temp: float = 0.5,
                eta0: float = 0.005, di: float = 0.1, amp: float = 0.1, width: float = 1.0,
                noise: float = 1e-4, seed: int = 1, cfl: float = 0.4, steps: int = 100,
                periodic_y: bool = True, hall: bool | None = None, whistler_dt: bool = True,
                x_offset: float = 0.0, gamma: float = 5.0 / 3.0, arrays: bool = True) -> MHDCase:
    """Harris sheet Bx = tanh(z/lam), p = p_inf + 0.5 sech^2, rho = p/temp,
    uniform guide By, plus dAy = amp*exp(-(x^2+z^2)/width^2) and seeded
    noise rho*v_i = rho*noise*U(-1,1) on the unique physical points.
    arrays=False: the mesh and parameters only (the device builds the 3-D
    fields, mhd_init_harris, with the same arithmetic)."""
    if stretched:
        px, sx = stretched_points(mx, dmin, ratio)
        pz, sz = stretched_points(mz, dmin, ratio)
    else:
        px, sx = uniform_points(mx, -lx / 2, lx / 2)
        pz, sz = uniform_points(mz, -lz / 2, lz / 2)
    px = px + x_offset
    if periodic_y:
        dyv = ly / ny
        yy = (np.arange(ny + 4, dtype=np.float64) - 2.0) * dyv
        dy = np.full(ny + 4, dyv)
        my = ny + 4
    else:
        yy, dy = uniform_points(ny, -ly / 2, ly / 2)
        my = ny
    xx, zz, dx, dz = pad2(px), pad2(pz), pad2(sx), pad2(sz)
    if hall is None:
        hall = di != 0.0
    params = dict(lam=lam, beta_inf=beta_inf, guide=guide, temp=temp, amp=amp, width=width,
                  noise=noise, x_offset=x_offset, seed=seed)
    if not arrays:
        return MHDCase(name=name, mx=mx, my=my, mz=mz, periodic_y=periodic_y, hall=hall, di=di,
                       eta0=eta0, gamma=gamma, cfl=cfl, whistler_dt=whistler_dt, steps=steps,
                       xx=xx, yy=yy, zz=zz, dx=dx, dy=dy, dz=dz, rhoi=None, bxi=None, byi=None,
                       bzi=None, pri=None, x0=None, harris=params)
    shape = (mz + 4, my, mx + 4)
    Z = np.zeros(shape)
    Xc = np.zeros(shape)
    Z[2:-2] = pz[:, None, None]
    Xc[:, :, 2:-2] = px[None, None, :]
    sech2 = 1.0 / np.cosh(Z / lam) ** 2
    p_inf = 0.5 * beta_inf
    pri = p_inf + 0.5 * sech2
    rhoi = pri / temp
    bxi = np.tanh(Z / lam)
    byi = np.full(shape, guide)
    bzi = np.zeros(shape)
    # analytic perturbation: dBx = -d(dAy)/dz, dBz = d(dAy)/dx
    xr = Xc - x_offset
    g = amp * np.exp(-(xr ** 2 + Z ** 2) / width ** 2)
    dbx = 2.0 * Z / width ** 2 * g
    dbz = -2.0 * xr / width ** 2 * g
    x0 = np.zeros((8,) + shape)
    x0[0] = rhoi
    x0[4] = bxi + dbx
    x0[5] = byi
    x0[6] = bzi + dbz
    x0[7] = pri
    # seeded noise on rho*v, indexed by unique physical point
    nyu = ny
    k = np.arange(mz)[:, None, None]
    i = np.arange(mx)[None, None, :]
    if periodic_y:
        jrows = (np.arange(my) - 2) % ny          # ghost rows copy the periodic image
    else:
        jrows = np.arange(my)
    j = jrows[None, :, None]
    for c in range(3):
        idx = ((c * mz + k) * nyu + j) * mx + i
        u = splitmix_uniform(seed, idx.astype(np.int64))
        x0[1 + c, 2:-2, :, 2:-2] = rhoi[2:-2, :, 2:-2] * noise * u
    # ghost columns (never read by the step) stay zero, as in the reference
    for a in (rhoi, pri, bxi, byi, bzi):
        a[:2] = 0; a[-2:] = 0; a[:, :, :2] = 0; a[:, :, -2:] = 0
    x0[:, :2] = 0; x0[:, -2:] = 0; x0[:, :, :, :2] = 0; x0[:, :, :, -2:] = 0
    return MHDCase(name=name, mx=mx, my=my, mz=mz, periodic_y=periodic_y, hall=hall, di=di,
                   eta0=eta0, gamma=gamma, cfl=cfl, whistler_dt=whistler_dt, steps=steps,
                   xx=xx, yy=yy, zz=zz, dx=dx, dy=dy, dz=dz, rhoi=rhoi, bxi=bxi, byi=byi,
                   bzi=bzi, pri=pri, x0=x0, harris=params)


def case_c1(arrays: bool = True) -> MHDCase:
    """64x32x64 uniform, box 20x10x20, eta 0.005, d_i 0.1, 100 steps."""
    return harris_case("C1_64x32x64", 64, 32, 64, stretched=False, lam=0.5, beta_inf=0.2,
                       eta0=0.005, di=0.1, amp=0.1, seed=1, steps=100, arrays=arrays)


def case_c2(arrays: bool = True) -> MHDCase:
    """256x128x256 stretched (dmin 0.05, ratio 3), eta 0.003, d_i 0.2, 1000 steps."""
    return harris_case("C2_256x128x256", 256, 128, 256, stretched=True, lam=0.5, beta_inf=0.2,
                       eta0=0.003, di=0.2, amp=0.05, seed=2, steps=1000, arrays=arrays)


def case_c3(arrays: bool = True) -> MHDCase:
    """512x256x512 stretched, beta 0.1, guide By 0.1, eta 0.002, d_i 0.25, 200 steps."""
    return harris_case("C3_512x256x512", 512, 256, 512, stretched=True, lam=0.5, beta_inf=0.1,
                       guide=0.1, eta0=0.002, di=0.25, amp=0.05, seed=3, steps=200, arrays=arrays)


def case_small(mx: int = 24, ny: int = 8, mz: int = 24, *, stretched: bool = True,
               periodic_y: bool = True, di: float = 0.15, eta0: float = 0.005, seed: int = 7,
               **kw) -> MHDCase:
    """Tiny Harris case for parity tests."""
    lx = kw.pop("lx", 8.0)
    lz = kw.pop("lz", 8.0)
    ly = kw.pop("ly", 4.0)
    dmin = kw.pop("dmin", 0.2)
    return harris_case(f"small_{mx}x{ny}x{mz}", mx, ny, mz, stretched=stretched, lx=lx, lz=lz,
                       ly=ly, dmin=dmin, periodic_y=periodic_y, di=di, eta0=eta0, seed=seed,
                       **kw)


def case_c4(seed: int = 4, nruns: int = 48) -> list:
    """The batched ensemble: nruns Harris-sheet runs, uniform meshes on the
    20 x 10 x 20 box of C1, size scale s in {1, 1.25, ..., 3} drawn from the
    seed -> (64s, 32s, 64s) rounded to multiples of 16; eta ~ U[1e-3, 1e-2],
    d_i ~ U[0, 0.5], perturbation amplitude ~ U[0.02, 0.2], sheet thickness
    ~ U[0.3, 1.0]; 500 steps each."""
    rng = np.random.default_rng(seed)

    def r16(v: float) -> int:
        return int(16 * np.floor(v / 16 + 0.5))

    out = []
    for r in range(nruns):
        s = 1.0 + 0.25 * int(rng.integers(0, 9))
        mx, ny = r16(64 * s), r16(32 * s)
        eta, di = rng.uniform(1e-3, 1e-2), rng.uniform(0.0, 0.5)
        amp, lam = rng.uniform(0.02, 0.2), rng.uniform(0.3, 1.0)
        out.append(harris_case(f"C4_{r:02d}_{mx}x{ny}x{mx}", mx, ny, mx, stretched=False, lam=lam,
                               beta_inf=0.2, eta0=eta, di=di, amp=amp, seed=100 + r, steps=500))
    return out


CASES = {"C1": case_c1, "C2": case_c2, "C3": case_c3}
