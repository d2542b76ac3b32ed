"""Pins of the CPU adjoint oracle (oracle/mhd_oracle_adj.c).

The reverse-mode restatement is checked against facts independent of its
code: the dot-product identity <A u, w> = <u, A^T w> of each linear primitive
against the forward primitive of mhd_oracle.c, and central differences of
the forward oracle for right(), one and three RK4 steps (with and without the
residual captured from the initial state), and the resistivity eta.
"""
import dataclasses

import numpy as np
import pytest

from dasha_amd import harris

orc = pytest.importorskip("oracle.oracle")


@pytest.fixture(scope="module")
def small():
    orc.build()
    c = harris.case_small(16, 8, 16, seed=3)
    o = orc.Oracle(c)
    yield c, o
    o.close()


def _direction(x0, rng, scale=1e-3):
    v = rng.standard_normal(x0.shape) * np.abs(x0).max(axis=(1, 2, 3))[:, None, None, None] * scale
    v[0] *= 0.1
    return v


@pytest.mark.parametrize("kind,nf,ck", [(0, 8, 0), (1, 8, 0), (2, 3, 0), (3, 3, 0), (3, 3, 1), (3, 3, 2)])
def test_primitive_transposes(small, kind, nf, ck):
    """bndry, bndrytt, bndryc/E/h and the three smoothing coefficients"""
    c, o = small
    rng = np.random.default_rng(10 + kind + ck)
    u = rng.standard_normal((nf,) + o.shape3)
    w = rng.standard_normal((nf,) + o.shape3)
    lhs = np.vdot(o.apply(kind, u, nf, ck), w)
    rhs = np.vdot(u, o.adj_apply(kind, w, nf, ck))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


@pytest.mark.parametrize("t,smooth", [(1.0, True), (0.0, False)])
def test_rhs_vjp_central_difference(small, t, smooth):
    c, o = small
    rng = np.random.default_rng(2)
    x0 = c.x0.copy()
    o.set_state(x0)
    o.init_residual()
    A = rng.standard_normal((8,) + o.shape3)
    V = _direction(x0, rng)
    o.set_state(x0)
    xb, rb, _ = o.adj_rhs(A, smooth=smooth)

    def J(x):
        o.set_state(x)
        o.rhs(t)
        return np.vdot(A, o.get_xdif())

    eps = 1e-3
    fd = (J(x0 + eps * V) - J(x0 - eps * V)) / (2 * eps)
    ad = np.vdot(V, xb)
    assert abs(fd - ad) <= 1e-8 * abs(ad)
    # the residual enters linearly: exact directional derivative
    W = rng.standard_normal((3,) + o.shape3) * 1e-3
    r0 = o.get_residual()
    o.set_state(x0)
    o.set_residual(r0 + W)
    jp = J(x0)
    o.set_residual(r0 - W)
    jm = J(x0)
    o.set_residual(r0)
    assert abs((jp - jm) / 2 - np.vdot(W, rb)) <= 1e-9 * abs(np.vdot(W, rb))


def _adjoint_run(o, x0, dts, A):
    o.set_state(x0)
    o.init_residual()
    states, ts, t = [], [], 0.0
    for dt in dts:
        states.append(o.get_state())
        ts.append(t)
        o.step(t, dt)
        t += dt
    xb = A.copy()
    rb = np.zeros((3,) + o.shape3)
    eb = 0.0
    for n in reversed(range(len(dts))):
        o.set_state(states[n])
        xb, eb = o.adj_step(ts[n], dts[n], xb, rb, eb)
    o.set_state(x0)
    o.adj_residual(rb, xb)
    return xb, eb


def _run(o, x0, dts, A):
    o.set_state(x0)
    o.init_residual()
    t = 0.0
    for dt in dts:
        o.step(t, dt)
        t += dt
    return np.vdot(A, o.get_state())


@pytest.mark.parametrize("nsteps", [1, 3])
def test_step_adjoint_central_difference(small, nsteps):
    """J = <A, x_N>, residual captured from x_0, frozen dt sequence"""
    c, o = small
    rng = np.random.default_rng(5 + nsteps)
    x0 = c.x0.copy()
    dts = [0.0056, 0.0055, 0.0057][:nsteps]
    A = rng.standard_normal((8,) + o.shape3)
    V = _direction(x0, rng)
    xb, eb = _adjoint_run(o, x0, dts, A)
    eps = 1e-3
    fd = (_run(o, x0 + eps * V, dts, A) - _run(o, x0 - eps * V, dts, A)) / (2 * eps)
    ad = np.vdot(V, xb)
    assert abs(fd - ad) <= 2e-8 * abs(ad)
    # eta
    e = 1e-5
    vals = []
    for de in (e, -e):
        o2 = orc.Oracle(dataclasses.replace(c, eta0=c.eta0 + de))
        vals.append(_run(o2, x0, dts, A))
        o2.close()
    fde = (vals[0] - vals[1]) / (2 * e)
    assert abs(fde - eb) <= 1e-6 * abs(eb)
