"""ctypes front end of the CPU oracle (oracle/mhd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.
The method names mirror the product's step interface so a parity test can
drive both with the same calls.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmhd_oracle.so")


class OrcConfig(C.Structure):
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int), ("periodic_y", C.c_int),
                ("hall", C.c_int), ("whistler_dt", C.c_int), ("filters", C.c_int),
                ("bcs", C.c_int), ("di", C.c_double), ("eta0", C.c_double),
                ("fmu0", C.c_double), ("frho0", C.c_double), ("fpr0", C.c_double),
                ("gamma", C.c_double), ("cfl", C.c_double),
                ("magnetosphere", C.c_int), ("inflow", C.c_int), ("radial_hall", C.c_int),
                ("partial_e", C.c_int), ("rcr", C.c_double)]


def build(force: bool = False) -> str:
    srcs = [os.path.join(HERE, f) for f in ("mhd_oracle.c", "mhd_oracle_adj.c")]
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(map(os.path.getmtime, srcs)):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.POINTER(OrcConfig), D, D, D, D, D, D]
        L.orc_destroy.argtypes = [P]
        for n in ("orc_set_state", "orc_get_state", "orc_get_current", "orc_get_xdif",
                  "orc_get_efld", "orc_get_xdh", "orc_get_residual", "orc_set_residual",
                  "orc_diagn", "orc_energy", "orc_divb"):
            getattr(L, n).argtypes = [P, D]
        L.orc_set_equilibrium.argtypes = [P, D, D, D, D, D]
        L.orc_init_residual.argtypes = [P]
        L.orc_rhs.argtypes = [P, C.c_double]
        L.orc_rhs_stage.argtypes = [P, C.c_double, C.c_double, C.c_int]
        L.orc_set_inflow_face.argtypes = [P, D, D, D]
        L.orc_current.argtypes = [P]
        L.orc_current_raw.argtypes = [P]
        L.orc_step.argtypes = [P, C.c_double, C.c_double]
        L.orc_compute_dt.argtypes = [P]
        L.orc_compute_dt.restype = C.c_double
        L.orc_get_coeffs.argtypes = [P, C.c_int, D]
        L.orc_get_coeffs.restype = C.c_int
        L.orc_apply.argtypes = [P, C.c_int, D, C.c_int, C.c_int]
        L.orc_adj_apply.argtypes = [P, C.c_int, D, C.c_int, C.c_int]
        L.orc_adj_rhs.argtypes = [P, D, D, D, D, C.c_int, C.c_int]
        L.orc_adj_step.argtypes = [P, C.c_double, C.c_double, D, D, D]
        L.orc_adj_residual.argtypes = [P, D, D]
        L.orc_adj_fold_y.argtypes = [P, D, C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return lib().orc_max_threads()


class Oracle:
    """CPU restatement of stepon/setdt on one domain."""

    def __init__(self, case, *, filters: bool = True, bcs: bool = True):
        self.case = case
        cfg = OrcConfig(mx=case.mx, my=case.my, mz=case.mz, periodic_y=int(case.periodic_y),
                        hall=int(case.hall), whistler_dt=int(case.whistler_dt),
                        filters=int(filters), bcs=int(bcs), di=case.di, eta0=case.eta0,
                        fmu0=case.fmu0, frho0=case.frho0, fpr0=case.fpr0, gamma=case.gamma,
                        cfl=case.cfl, magnetosphere=int(getattr(case, "magnetosphere", False)),
                        inflow=int(getattr(case, "inflow", False)),
                        radial_hall=int(getattr(case, "radial_hall", False)),
                        partial_e=int(getattr(case, "partial_e", False)),
                        rcr=float(getattr(case, "rcr", 0.0)))
        self._keep = [np.ascontiguousarray(a, dtype=np.float64)
                      for a in (case.xx, case.yy, case.zz, case.dx, case.dy, case.dz)]
        self.h = lib().orc_create(C.byref(cfg), *[_p(a) for a in self._keep])
        self.shape3 = (case.mz + 4, case.my, case.mx + 4)
        eq = [np.ascontiguousarray(a, dtype=np.float64)
              for a in (case.rhoi, case.bxi, case.byi, case.bzi, case.pri)]
        lib().orc_set_equilibrium(self.h, *[_p(a) for a in eq])
        face = getattr(case, "inflow_face", None)
        if face is not None:
            self.set_inflow_face(*face)
        self.set_state(case.x0)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _get(self, fn, nf):
        out = np.empty((nf,) + self.shape3)
        getattr(lib(), fn)(self.h, _p(out))
        return out

    def set_state(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        assert a.shape == (8,) + self.shape3
        lib().orc_set_state(self.h, _p(a))

    def get_state(self):
        return self._get("orc_get_state", 8)

    def get_current(self):
        return self._get("orc_get_current", 3)

    def get_xdif(self):
        return self._get("orc_get_xdif", 8)

    def get_efld(self):
        return self._get("orc_get_efld", 3)

    def get_xdh(self):
        return self._get("orc_get_xdh", 3)

    def get_residual(self):
        return self._get("orc_get_residual", 3)

    def set_residual(self, r):
        a = np.ascontiguousarray(r, dtype=np.float64)
        lib().orc_set_residual(self.h, _p(a))

    def init_residual(self):
        lib().orc_init_residual(self.h)

    def compute_dt(self) -> float:
        return lib().orc_compute_dt(self.h)

    def step(self, time: float, dt: float):
        lib().orc_step(self.h, time, dt)

    def rhs(self, t: float):
        lib().orc_rhs(self.h, t)

    def rhs_stage(self, t: float, ddt: float, km: int):
        """right(t, ddt, km): the stage arguments the driven inflow reads"""
        lib().orc_rhs_stage(self.h, t, ddt, int(km))

    def set_inflow_face(self, bx, by, bz):
        """inflow face data bxinfc, byinfc, bzinfc: each (mz+4, my) C order"""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (bx, by, bz)]
        for a in arrs:
            assert a.shape == (self.case.mz + 4, self.case.my)
        lib().orc_set_inflow_face(self.h, *[_p(a) for a in arrs])

    def current(self):
        lib().orc_current(self.h)

    def current_raw(self):
        lib().orc_current_raw(self.h)

    def diagn(self):
        out = np.empty(6)
        lib().orc_diagn(self.h, _p(out))
        return out

    def energy(self):
        out = np.empty(4)
        lib().orc_energy(self.h, _p(out))
        return out

    def divb(self):
        out = np.empty(self.shape3)
        lib().orc_divb(self.h, _p(out))
        return out

    def coeffs(self, axis: int):
        n = (self.case.mx + 4, self.case.my, self.case.mz + 4)[axis]
        out = np.empty((14, n))
        lib().orc_get_coeffs(self.h, axis, _p(out))
        return out

    def apply(self, kind: int, a, nf: int = 3, coef_kind: int = 0):
        """0 bndry(xdif), 1 bndrytt(xm), 2 bndryc/E/h, 3 one smoothing sweep"""
        b = np.ascontiguousarray(a, dtype=np.float64).copy()
        lib().orc_apply(self.h, kind, _p(b), nf, coef_kind)
        return b

    # ---- reverse mode (oracle/mhd_oracle_adj.c) ----
    def adj_apply(self, kind: int, a, nf: int = 3, coef_kind: int = 0):
        """transpose of apply(kind, ...)"""
        b = np.ascontiguousarray(a, dtype=np.float64).copy()
        lib().orc_adj_apply(self.h, kind, _p(b), nf, coef_kind)
        return b

    def adj_rhs(self, abar, *, init: bool = False, smooth: bool = True):
        """(xbar, rbar, etabar): vjp of rhs(t) (smooth = t != 0) or of the
        residual capture (init) at the current state"""
        ab = np.ascontiguousarray(abar, dtype=np.float64)
        xb = np.zeros((8,) + self.shape3)
        rb = np.zeros((3,) + self.shape3)
        eb = np.zeros(1)
        lib().orc_adj_rhs(self.h, _p(ab), _p(xb), _p(rb), _p(eb), int(init), int(smooth))
        return xb, rb, float(eb[0])

    def adj_step(self, time: float, dt: float, xbar, rbar, etabar: float = 0.0):
        """advance one step (as step) and map xbar (adjoint of the new state)
        to the adjoint of the old one; rbar accumulates in place"""
        xb = np.ascontiguousarray(xbar, dtype=np.float64).copy()
        assert rbar.dtype == np.float64 and rbar.flags.c_contiguous
        eb = np.array([etabar])
        lib().orc_adj_step(self.h, time, dt, _p(xb), _p(rbar), _p(eb))
        return xb, float(eb[0])

    def adj_residual(self, rbar, xbar):
        """xbar += (d xresd / d x)^T rbar at the current state (in place)"""
        rb = np.ascontiguousarray(rbar, dtype=np.float64)
        assert xbar.dtype == np.float64 and xbar.flags.c_contiguous
        lib().orc_adj_residual(self.h, _p(rb), _p(xbar))

    def fold_y(self, a, nf: int):
        """adjoint of the periodic image rows folded onto the rows they copy"""
        b = np.ascontiguousarray(a, dtype=np.float64).copy()
        lib().orc_adj_fold_y(self.h, _p(b), nf)
        return b

    def run(self, steps: int, *, time: float = 0.0, init: bool = True):
        """driver loop t3dsdsc.f:136-202 without I/O: setdt, stepon, advance"""
        if init:
            self.init_residual()
        dts = []
        for _ in range(steps):
            dt = self.compute_dt()
            self.step(time, dt)
            time += dt
            dts.append(dt)
        return time, dts
