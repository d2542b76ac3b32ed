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


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("kind,nf,ck", [(0, 8, 0), (1, 8, 0), (2, 3, 0), (3, 3, 0), (3, 3, 1), (3, 3, 2)])
def test_primitive_transposes(small, kind, nf, ck, periodic):
    """bndry, bndrytt, bndryc/E/h and the three smoothing coefficients, with
    periodic and with non-periodic y"""
    c, o = small
    if not periodic:
        c = harris.case_small(16, 12, 16, periodic_y=False, seed=3)
        o = orc.Oracle(c)
    rng = np.random.default_rng(10 + kind + ck)
    u = rng.standard_normal((nf,) + o.shape3)
    w = rng.standard_normal((nf,) + o.shape3)
    lhs = np.vdot(o.apply(kind, u, nf, ck), w)
    rhs = np.vdot(u, o.adj_apply(kind, w, nf, ck))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    if not periodic:
        o.close()


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


@pytest.mark.parametrize("nsteps,variant", [(1, "stretched"), (3, "stretched"), (2, "no_hall"), (2, "uniform"),
                                            (2, "nonperiodic"), (2, "viscous")])
def test_step_adjoint_central_difference(small, nsteps, variant):
    """J = <A, x_N>, residual captured from x_0, frozen dt sequence"""
    c, o = small
    if variant != "stretched":
        c = {"no_hall": lambda: harris.case_small(16, 8, 16, seed=3, di=0.0),
             "uniform": lambda: harris.case_small(16, 8, 20, stretched=False, seed=3),
             "nonperiodic": lambda: harris.case_small(16, 12, 16, periodic_y=False, seed=3),
             "viscous": lambda: dataclasses.replace(harris.case_small(16, 8, 16, seed=3), fmu0=2e-3,
                                                    frho0=1e-3, fpr0=1.5e-3)}[variant]()
        o = orc.Oracle(c)
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
    if variant != "stretched":
        o.close()


# ------------------------------------------------ control map and objectives
def test_control_map_transposes():
    from dasha_amd import adjoint
    rng = np.random.default_rng(3)
    c = np.cumsum(rng.uniform(0.1, 0.5, 17))
    for axis in (0, 1):
        shp = (17, 5) if axis == 0 else (5, 17)
        f = rng.standard_normal(shp)
        g = rng.standard_normal(shp)
        lhs = np.vdot(adjoint.mesh_derivative(f, c, axis), g)
        rhs = np.vdot(f, adjoint.mesh_derivative_t(g, c, axis))
        assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # the central quotient of a quadratic is exact on a uniform mesh
    u = np.linspace(-1, 1, 11)
    d = adjoint.mesh_derivative((u ** 2)[:, None] * np.ones((1, 3)), u, 0)
    assert np.allclose(d[1:-1, 0], 2 * u[1:-1], atol=1e-13)
    ac = adjoint.small_case(16, 8, 16)
    ay = rng.standard_normal((16, 16))
    gx = rng.standard_normal((8,) + ac.case.shape3)
    dx0 = ac.state(ay) - ac.state(np.zeros_like(ay))
    assert abs(np.vdot(dx0, gx) - np.vdot(ay, adjoint.ay_gradient(ac.case, gx))) <= 1e-11 * abs(np.vdot(dx0, gx))


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("obj", ["magnetic_energy", "reconnected_flux"])
def test_objective_seed_central_difference(obj, periodic):
    from dasha_amd import adjoint
    ac = adjoint.small_case(16, 8 if periodic else 12, 16, periodic_y=periodic)
    x = ac.state()
    rng = np.random.default_rng(4)
    v = rng.standard_normal(x.shape) * 1e-3
    s = adjoint.objective_seed(ac.case, x, obj)
    eps = 1e-3
    fd = (adjoint.objective_of_state(ac.case, x + eps * v, obj) -
          adjoint.objective_of_state(ac.case, x - eps * v, obj)) / (2 * eps)
    assert abs(fd - np.vdot(s, v)) <= 1e-9 * abs(fd)


def oracle_adjoint(ac, dts, obj, ay=None):
    """(value, d value / d x0 folded onto unique rows, d value / d eta) of the
    CPU oracle: forward run, objective seed, reverse steps, residual vjp"""
    from dasha_amd import adjoint
    o = orc.Oracle(ac.case)
    x0 = ac.state(ay)
    o.set_state(x0)
    o.init_residual()
    states, ts, t = [], [], 0.0
    for dt in dts:
        states.append(o.get_state())
        ts.append(t)
        o.step(t, dt)
        t += dt
    xN = o.get_state()
    val = adjoint.objective_of_state(ac.case, xN, obj)
    xb = adjoint.objective_seed(ac.case, xN, obj)
    rb = np.zeros((3,) + o.shape3)
    eb = 0.0
    for n in reversed(range(len(dts))):
        o.set_state(states[n])
        xb, eb = o.adj_step(ts[n], dts[n], xb, rb, eb)
    o.set_state(x0)
    o.adj_residual(rb, xb)
    xb = o.fold_y(xb, 8)
    o.close()
    return val, xb, eb


@pytest.mark.parametrize("obj", ["magnetic_energy", "reconnected_flux"])
def test_oracle_ay_gradient_central_difference(obj):
    """the whole chain on the CPU: control map, steps, objective"""
    from dasha_amd import adjoint
    ac = adjoint.small_case(16, 8, 16)
    dts = [0.0056, 0.0055]
    val, xb, eb = oracle_adjoint(ac, dts, obj)
    gay = adjoint.ay_gradient(ac.case, xb)
    rng = np.random.default_rng(6)
    v = rng.standard_normal(ac.ay0.shape) * 0.01

    def J(ay):
        o = orc.Oracle(ac.case)
        o.set_state(ac.state(ay))
        o.init_residual()
        t = 0.0
        for dt in dts:
            o.step(t, dt)
            t += dt
        r = adjoint.objective_of_state(ac.case, o.get_state(), obj)
        o.close()
        return r

    eps = 1e-3
    fd = (J(ac.ay0 + eps * v) - J(ac.ay0 - eps * v)) / (2 * eps)
    assert abs(fd - np.vdot(gay, v)) <= 1e-7 * abs(fd)


@pytest.mark.parametrize("periodic", [True, False])
def test_central_difference_closed_forms(periodic):
    """adjoint.central_difference: the magnetic energy is quadratic, so its
    central difference along d is exactly sum w b.d for any h; the flux is
    |Bz~|, linear while no sign changes, so its quotient is seed.d"""
    from dasha_amd import adjoint
    ac = adjoint.small_case(16, 8 if periodic else 12, 16, periodic_y=periodic)
    x = ac.state()
    rng = np.random.default_rng(9)
    d = rng.standard_normal(x.shape)
    for h in (1e-6, 1e-2):
        for obj in ("magnetic_energy", "reconnected_flux"):
            q = adjoint.central_difference(ac.case, x + h * d, x - h * d, h, obj)
            seed = adjoint.objective_seed(ac.case, x, obj)
            exact = np.vdot(seed, d)
            if obj == "reconnected_flux" and h > 1e-4:
                continue            # large steps cross the zeros of Bz~
            # x +- h d rounds each term by ~1e-16 |x| / (h |d|)
            assert abs(q - exact) <= 1e-9 * np.vdot(np.abs(seed), np.abs(d)), (h, obj)
            assert adjoint.central_difference(ac.case, x, x, h, obj) == 0.0
    # the same quotient as the difference of two objective evaluations, to
    # the latter's rounding (two sums of O(1) that cancel to O(1e-3))
    xp, xm = x + 1e-3 * d, x - 1e-3 * d
    for obj in ("magnetic_energy", "reconnected_flux"):
        two = (adjoint.objective_of_state(ac.case, xp, obj) - adjoint.objective_of_state(ac.case, xm, obj)) / 2e-3
        assert adjoint.central_difference(ac.case, xp, xm, 1e-3, obj) == pytest.approx(two, rel=1e-8)


GOLDEN_FD = __import__("pathlib").Path(__file__).parent / "golden" / "c5_fd_oracle.json"


@pytest.mark.skipif(not GOLDEN_FD.exists(), reason="c5_fd_oracle.json not generated")
def test_c5_fd_fixture_is_consistent():
    """the oracle's central differences at 128x64x128 (make_c5_fd_oracle.py):
    the directions regenerate from their seed, the dt sequence is the full
    run's, and the differenced quotient agrees with the two-sum one to the
    latter's rounding"""
    import json
    from dasha_amd import adjoint
    fx = json.loads(GOLDEN_FD.read_text())
    ac = adjoint.case_c5()
    assert fx["h"] == 1e-6 and len(fx["dts"]) == ac.steps
    dts = np.array(fx["dts"])
    assert (dts > 0).all() and dts.max() / dts.min() < 2.0
    for obj, (s1, s2) in fx["direction_checksum"].items():
        v = adjoint.fd_direction(ac, obj, fx["seed"])
        assert np.sum(v) == pytest.approx(s1, rel=1e-12) and np.sum(v * v) == pytest.approx(s2, rel=1e-12)
    for key in ("fd_ay", "fd_eta"):
        for obj, q in fx[key].items():
            assert q == pytest.approx(fx[key + "_two_sums"][obj], rel=5e-6)
