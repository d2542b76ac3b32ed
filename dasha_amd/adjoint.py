"""Gradients of the step through its discrete adjoint (``mhd_adjoint_run``).

The reference has no derivatives; this module differentiates its RK4 step
(stepon/right, t3dsdsc.f:1063-1673) in reverse mode on the GPU:

* controls: the initial perturbation potential dAy(x, z) on the mx x mz
  physical points (y independent) and the resistivity eta;
* state map: x0 = base + (dBx, dBz) = base + (-d dAy/dz, d dAy/dx) on every y
  row, with the derivative of the physical mesh defined below;
* objectives at time T (after a frozen dt sequence): the magnetic energy
  1/2 sum |B|^2 dV, or the reconnected flux 1/2 sum_x |Bz~| dx averaged over
  y, Bz~ the mean of the two z planes nearest z = 0.

``AdjointObjective`` is a ``torch.autograd.Function`` over the same C ABI:
its forward runs the forward pass and the reverse sweep in one library call
and its backward returns the saved gradients times the incoming one.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import harris
from .mhd import MHDStep


# ---------------------------------------------------------------- control map
def mesh_derivative(f: np.ndarray, c: np.ndarray, axis: int) -> np.ndarray:
    """d f / d c along `axis` on a non-uniform mesh: central
    (f[i+1] - f[i-1]) / (c[i+1] - c[i-1]) inside, one-sided at both ends"""
    f = np.moveaxis(np.asarray(f, dtype=np.float64), axis, 0)
    out = np.empty_like(f)
    sh = (-1,) + (1,) * (f.ndim - 1)
    out[1:-1] = (f[2:] - f[:-2]) / (c[2:] - c[:-2]).reshape(sh)
    out[0] = (f[1] - f[0]) / (c[1] - c[0])
    out[-1] = (f[-1] - f[-2]) / (c[-1] - c[-2])
    return np.moveaxis(out, 0, axis)


def mesh_derivative_t(g: np.ndarray, c: np.ndarray, axis: int) -> np.ndarray:
    """transpose of mesh_derivative along `axis`"""
    g = np.moveaxis(np.asarray(g, dtype=np.float64), axis, 0)
    out = np.zeros_like(g)
    sh = (-1,) + (1,) * (g.ndim - 1)
    w = g[1:-1] / (c[2:] - c[:-2]).reshape(sh)
    out[2:] += w
    out[:-2] -= w
    w0 = g[0] / (c[1] - c[0])
    out[1] += w0
    out[0] -= w0
    w1 = g[-1] / (c[-1] - c[-2])
    out[-1] += w1
    out[-2] -= w1
    return np.moveaxis(out, 0, axis)


def physical_coords(case):
    return case.xx[2:case.mx + 2], case.zz[2:case.mz + 2]


def state_from_ay(case, base: np.ndarray, ay: np.ndarray) -> np.ndarray:
    """x0 = base + curl(ay y^): dBx = -d ay/dz, dBz = d ay/dx on the physical
    points of every y row (ay has shape (mz, mx))"""
    px, pz = physical_coords(case)
    ay = np.asarray(ay, dtype=np.float64)
    assert ay.shape == (case.mz, case.mx)
    x0 = np.array(base, dtype=np.float64, copy=True)
    dbx = -mesh_derivative(ay, pz, 0)
    dbz = mesh_derivative(ay, px, 1)
    x0[4, 2:-2, :, 2:-2] += dbx[:, None, :]
    x0[6, 2:-2, :, 2:-2] += dbz[:, None, :]
    return x0


def ay_gradient(case, grad_x: np.ndarray) -> np.ndarray:
    """d value / d ay from d value / d x0 (transpose of state_from_ay)"""
    px, pz = physical_coords(case)
    gbx = grad_x[4, 2:-2, :, 2:-2].sum(axis=1)
    gbz = grad_x[6, 2:-2, :, 2:-2].sum(axis=1)
    return -mesh_derivative_t(gbx, pz, 0) + mesh_derivative_t(gbz, px, 1)


# ---------------------------------------------------------------- the C5 case
@dataclasses.dataclass
class AdjointCase:
    case: object              # MHDCase with the base state (no dAy) as x0
    ay0: np.ndarray           # nominal control dAy on (mz, mx)
    steps: int

    def state(self, ay=None) -> np.ndarray:
        return state_from_ay(self.case, self.case.x0, self.ay0 if ay is None else ay)


def case_c5(steps: int = 500, mx: int = 128, ny: int = 64, mz: int = 128, *, amp: float = 0.1,
            width: float = 1.0, eta0: float = 0.005, di: float = 0.15, seed: int = 5,
            stretched: bool = True, **kw) -> AdjointCase:
    """128x64x128 tanh-stretched Harris sheet (dmin 0.05, ratio 3, Ly 10),
    eta 0.005, d_i 0.15, 500 steps; control dAy = amp exp(-(x^2+z^2)/w^2)"""
    base = harris.harris_case(f"C5_{mx}x{ny}x{mz}", mx, ny, mz, stretched=stretched, amp=0.0,
                              eta0=eta0, di=di, seed=seed, steps=steps, **kw)
    px, pz = physical_coords(base)
    ay0 = amp * np.exp(-(px[None, :] ** 2 + pz[:, None] ** 2) / width ** 2)
    return AdjointCase(case=base, ay0=ay0, steps=steps)


def small_case(mx: int = 24, ny: int = 8, mz: int = 24, steps: int = 3, *, fmu0: float = 0.0,
               frho0: float = 0.0, fpr0: float = 0.0, **kw) -> AdjointCase:
    """a small stretched case of the same form (tests); fmu0, frho0, fpr0
    switch on viscosity and diffusion"""
    base = dataclasses.replace(harris.case_small(mx, ny, mz, amp=0.0, **kw), fmu0=fmu0, frho0=frho0, fpr0=fpr0)
    px, pz = physical_coords(base)
    ay0 = 0.1 * np.exp(-(px[None, :] ** 2 + pz[:, None] ** 2))
    return AdjointCase(case=base, ay0=ay0, steps=steps)


def dt_sequence(ac: AdjointCase, ay=None) -> list:
    """the dt sequence of a plain run from the nominal state (setdt per step);
    the adjoint and every finite-difference run reuse it (frozen)"""
    g = MHDStep(ac.case)
    try:
        g.set_state(ac.state(ay))
        _, dts = g.run(ac.steps)
    finally:
        g.close()
    return dts


def with_eta(ac: AdjointCase, eta: float) -> AdjointCase:
    return dataclasses.replace(ac, case=dataclasses.replace(ac.case, eta0=float(eta)))


def objective(ac: AdjointCase, dts, ay=None, eta=None, obj="magnetic_energy", every: int = 0):
    """(value, d value/d ay, d value/d eta, (forward ms, reverse ms));
    every: checkpoint spacing (0: ceil(sqrt(N)), < 0: keep every step)"""
    c = ac if eta is None else with_eta(ac, eta)
    g = MHDStep(c.case)
    try:
        g.set_state(c.state(ay))
        g.init_residual()
        val, gx, ge, ms = g.adjoint_run(dts, obj, every=every)
    finally:
        g.close()
    return val, ay_gradient(c.case, gx), ge, ms


def forward_state(ac: AdjointCase, dts, ay=None, eta=None) -> np.ndarray:
    """the state after a plain forward run over the frozen dt sequence"""
    c = ac if eta is None else with_eta(ac, eta)
    g = MHDStep(c.case)
    try:
        g.set_state(c.state(ay))
        g.init_residual()
        t = 0.0
        for dt in dts:
            g.step(t, dt)
            t += dt
        return g.get_state()
    finally:
        g.close()


def forward_value(ac: AdjointCase, dts, ay=None, eta=None, obj="magnetic_energy") -> float:
    """the objective alone, from a plain forward run (finite differences)"""
    return objective_of_state(ac.case, forward_state(ac, dts, ay, eta), obj)


def fd_direction(ac: AdjointCase, obj: str, seed: int = 7) -> np.ndarray:
    """a direction in dAy for central differences, scaled to max |v| = 1.
    magnetic_energy: white noise; reconnected_flux: the control's shape with
    a +-50% random modulation, which keeps the sign of Bz~ at every point
    under a small step (|Bz~| has a kink where Bz~ = 0)"""
    rng = np.random.default_rng(seed)
    if obj == "magnetic_energy":
        v = rng.standard_normal(ac.ay0.shape)
    else:
        v = ac.ay0 * (1.0 + 0.5 * rng.uniform(-1.0, 1.0, ac.ay0.shape))
    return v / np.abs(v).max()


def central_difference(case, xp: np.ndarray, xm: np.ndarray, h: float, obj="magnetic_energy") -> float:
    """(J(xp) - J(xm)) / 2h, the difference taken inside one weighted sum
    (1/2 w (b+ - b-)(b+ + b-), or w (|bz+| - |bz-|)) instead of between two
    sums of ~1e6 terms, whose own rounding (~1e-12 of J) would otherwise
    dominate the quotient at h = 1e-6"""
    wx, rows, wy, wz, k0 = _weights(case)
    if obj == "magnetic_energy":
        bp, bm = xp[4:7, 2:-2, rows, 2:-2], xm[4:7, 2:-2, rows, 2:-2]
        w = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        return float(0.5 * np.sum(w[None] * (bp - bm) * (bp + bm))) / (2.0 * h)
    zp = 0.5 * (xp[6, 2 + k0, rows, 2:-2] + xp[6, 3 + k0, rows, 2:-2])
    zm = 0.5 * (xm[6, 2 + k0, rows, 2:-2] + xm[6, 3 + k0, rows, 2:-2])
    return float(np.sum(0.5 * wx[None, :] * (np.abs(zp) - np.abs(zm))) / case.ny) / (2.0 * h)


def _trapezoid(c: np.ndarray) -> np.ndarray:
    w = np.empty(c.size)
    w[1:-1] = 0.5 * (c[2:] - c[:-2])
    w[0] = 0.5 * (c[1] - c[0])
    w[-1] = 0.5 * (c[-1] - c[-2])
    return w


def _weights(case):
    """(wx, rows, wy, wz, k0): trapezoid weights over the unique physical
    points (periodic y: the rows 2..my-3, each dy wide), and the lower of the
    two z planes nearest z = 0"""
    px, pz = physical_coords(case)
    if case.periodic_y:
        rows = slice(2, case.my - 2)
        wy = np.full(case.ny, case.yy[3] - case.yy[2])
    else:
        rows = slice(0, case.my)
        wy = _trapezoid(case.yy)
    hits = np.nonzero((pz[:-1] < 0.0) & (pz[1:] >= 0.0))[0]
    k0 = int(hits[0]) if hits.size else -1
    return _trapezoid(px), rows, wy, _trapezoid(pz), k0


def objective_of_state(case, x: np.ndarray, obj="magnetic_energy") -> float:
    """host evaluation of the objectives from a state in the host layout"""
    wx, rows, wy, wz, k0 = _weights(case)
    if obj == "magnetic_energy":
        b = x[4:7, 2:-2, rows, 2:-2]
        w = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        return float(0.5 * np.sum(w[None] * b * b))
    bz = 0.5 * (x[6, 2 + k0, rows, 2:-2] + x[6, 3 + k0, rows, 2:-2])
    return float(np.sum(0.5 * wx[None, :] * np.abs(bz)) / case.ny)


def objective_seed(case, x: np.ndarray, obj="magnetic_energy") -> np.ndarray:
    """d objective_of_state / d x in the host layout (image rows 0)"""
    wx, rows, wy, wz, k0 = _weights(case)
    s = np.zeros_like(x)
    if obj == "magnetic_energy":
        w = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        s[4:7, 2:-2, rows, 2:-2] = w[None] * x[4:7, 2:-2, rows, 2:-2]
        return s
    bz = 0.5 * (x[6, 2 + k0, rows, 2:-2] + x[6, 3 + k0, rows, 2:-2])
    g = 0.5 * wx[None, :] * np.sign(bz) * 0.5 / case.ny
    s[6, 2 + k0, rows, 2:-2] = g
    s[6, 3 + k0, rows, 2:-2] = g
    return s


try:
    import torch

    class AdjointObjective(torch.autograd.Function):
        """value(ay, eta) with gradients from the discrete adjoint.

        forward(ay (mz, mx) float64, eta (0-d float64), ac, dts, obj) runs the
        library's forward pass and reverse sweep; backward scales the saved
        gradients by the incoming gradient."""

        @staticmethod
        def forward(ctx, ay, eta, ac, dts, obj="magnetic_energy"):
            val, gay, geta, _ = objective(ac, dts, ay.detach().cpu().numpy(), float(eta), obj)
            ctx.save_for_backward(torch.from_numpy(gay), torch.tensor(geta, dtype=torch.float64))
            return torch.tensor(val, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            gay, geta = ctx.saved_tensors
            return gout * gay, gout * geta, None, None, None
except ImportError:   # torch is optional for the library itself
    AdjointObjective = None


def sqrt_checkpoints(nsteps: int) -> int:
    """segment length of the default checkpointing (states kept every K)"""
    return int(math.ceil(math.sqrt(nsteps)))
