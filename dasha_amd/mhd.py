"""Host-side mirror of the reference's step interface over the C ABI.

The reference drives its step from ``program main`` (t3dsdsc.f:136-202):
``setdt`` -> dt-collapse check -> ``stepon`` -> ``nstep += 1; time += dt``.
``MHDStep`` exposes the same operations (``compute_dt``, ``step``,
``init_residual``, ``current``, ``diagnostics``) on the HIP implementation
in ``libmhd_gpu.so`` through ``include/mhd_gpu.h``.  There is no CPU path:
constructing an ``MHDStep`` without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmhd_gpu.so")

MHD_OK = 0
ERRORS = {-1: "bad argument", -2: "HIP error", -3: "out of device memory",
          -4: "dt collapse (dt < 0.1*dt0)", -5: "current blow-up (|Jy| >= 1e7)"}

# every symbol declared in include/mhd_gpu.h
EXPORTS = ("mhd_create", "mhd_set_equilibrium", "mhd_set_state", "mhd_get_state",
           "mhd_init_residual", "mhd_compute_dt", "mhd_step", "mhd_advance", "mhd_get_current",
           "mhd_diagnostics", "mhd_get_divb", "mhd_rhs", "mhd_get_xdif", "mhd_get_residual",
           "mhd_set_residual", "mhd_set_profiling", "mhd_kernel_times", "mhd_copy_rate",
           "mhd_last_error", "mhd_destroy", "mhd_ensemble_create", "mhd_ensemble_advance",
           "mhd_ensemble_status", "mhd_ensemble_error", "mhd_ensemble_destroy",
           "mhd_adjoint_run", "mhd_adjoint_rhs", "mhd_ensemble_tune", "mhd_compute_dt_noalloc",
           "mhd_set_inflow_face", "mhd_rhs_stage", "mhd_ensemble_tune_shapes", "mhd_ensemble_set_shape",
           "mhd_init_harris")

MHD_ADJ_RESIDUAL = 1
OBJECTIVES = {"magnetic_energy": 0, "reconnected_flux": 1}


class MHDConfig(C.Structure):
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int), ("periodic_y", C.c_int),
                ("hall", C.c_int), ("whistler_dt", C.c_int), ("di", C.c_double),
                ("eta0", C.c_double), ("fmu0", C.c_double), ("frho0", C.c_double),
                ("fpr0", C.c_double), ("gamma", C.c_double), ("cfl", C.c_double),
                ("magnetosphere", C.c_int), ("inflow", C.c_int), ("radial_hall", C.c_int),
                ("partial_e", C.c_int), ("rcr", C.c_double)]


class MHDHarris(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("lam", "beta_inf", "guide", "temp", "amp", "width", "noise",
                                           "x_offset")] + [("seed", C.c_ulonglong)]


class MHDDiag(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cxmax", "cxmin", "cymax", "cymin", "czmax", "czmin",
                                           "e_kin", "e_mag", "e_th", "e_mag_pert", "divb_max",
                                           "divb_rms")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# replay forms of an ensemble's step pairs (mhd_tile_shape.form)
FORM_RUNS, FORM_BATCH_GRAPH, FORM_KERNELS = 0, 1, 2


class MHDTileShape(C.Structure):
    _fields_ = [("kc", C.c_int), ("zkc", C.c_int), ("face_rows", C.c_int), ("form", C.c_int),
                ("rhs_rows", C.c_int), ("wave_cells", C.c_int), ("ms", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class MHDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIBPATH):
    """Load libmhd_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is missing: run __graft_entry__.build() first")
    L = C.CDLL(path)
    P, D, I = C.c_void_p, C.POINTER(C.c_double), C.c_int
    L.mhd_create.argtypes = [C.POINTER(MHDConfig), D, D, D, D, D, D, C.POINTER(P)]
    L.mhd_set_equilibrium.argtypes = [P, D, D, D, D, D]
    L.mhd_init_harris.argtypes = [P, C.POINTER(MHDHarris)]
    for n in ("mhd_set_state", "mhd_get_state", "mhd_get_current", "mhd_get_divb",
              "mhd_get_xdif", "mhd_get_residual", "mhd_set_residual"):
        getattr(L, n).argtypes = [P, D]
    L.mhd_init_residual.argtypes = [P]
    L.mhd_compute_dt.argtypes = [P, D]
    L.mhd_compute_dt_noalloc.argtypes = [P, D]
    L.mhd_step.argtypes = [P, C.c_double, C.c_double]
    L.mhd_advance.argtypes = [P, I, C.c_double, D, D, D]
    L.mhd_diagnostics.argtypes = [P, C.POINTER(MHDDiag)]
    L.mhd_rhs.argtypes = [P, C.c_double]
    L.mhd_rhs_stage.argtypes = [P, C.c_double, C.c_double, I]
    L.mhd_set_inflow_face.argtypes = [P, D, D, D]
    L.mhd_set_profiling.argtypes = [P, I]
    L.mhd_kernel_times.argtypes = [P, I, C.POINTER(C.c_char_p), D, C.POINTER(I)]
    L.mhd_copy_rate.argtypes = [C.c_long, I, D]
    L.mhd_last_error.argtypes = [P]
    L.mhd_last_error.restype = C.c_char_p
    L.mhd_destroy.argtypes = [P]
    L.mhd_ensemble_create.argtypes = [C.POINTER(P), I, C.POINTER(P)]
    L.mhd_ensemble_advance.argtypes = [P, I, D, D, D]
    L.mhd_ensemble_status.argtypes = [P, C.POINTER(I)]
    L.mhd_ensemble_error.argtypes = [P]
    L.mhd_ensemble_error.restype = C.c_char_p
    L.mhd_ensemble_destroy.argtypes = [P]
    L.mhd_adjoint_run.argtypes = [P, I, D, I, I, I, D, D, D, D]
    L.mhd_ensemble_tune.argtypes = [P, I, D, D, C.POINTER(I), D]
    L.mhd_adjoint_rhs.argtypes = [P, D, D, D, D, I, I]
    L.mhd_ensemble_tune_shapes.argtypes = [P, I, D, D, I, C.POINTER(MHDTileShape), C.POINTER(I),
                                           C.POINTER(MHDTileShape)]
    L.mhd_ensemble_set_shape.argtypes = [P, C.POINTER(MHDTileShape)]
    _lib = L
    return L


def copy_rate(nbytes: int = 2 << 30, iters: int = 10) -> float:
    """read+write bytes/s of a plain device copy (library kernel)"""
    L = load_library()
    v = C.c_double()
    rc = L.mhd_copy_rate(int(nbytes), int(iters), C.byref(v))
    if rc != MHD_OK:
        raise MHDError(rc, "copy probe failed")
    return v.value


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class MHDStep:
    """One configured domain on the GPU: stepon/setdt of t3dsdsc.f."""

    def __init__(self, case):
        self.L = load_library()
        self.case = case
        # open ensembles that use this handle; close() waits for them
        self._groups = 0
        self._close_pending = False
        cfg = MHDConfig(mx=case.mx, my=case.my, mz=case.mz, periodic_y=int(case.periodic_y),
                        hall=int(case.hall), whistler_dt=int(case.whistler_dt), di=case.di,
                        eta0=case.eta0, fmu0=case.fmu0, frho0=case.frho0, fpr0=case.fpr0,
                        gamma=case.gamma, cfl=case.cfl,
                        magnetosphere=int(getattr(case, "magnetosphere", False)),
                        inflow=int(getattr(case, "inflow", False)),
                        radial_hall=int(getattr(case, "radial_hall", False)),
                        partial_e=int(getattr(case, "partial_e", False)),
                        rcr=float(getattr(case, "rcr", 0.0)))
        geo = [np.ascontiguousarray(a, dtype=np.float64)
               for a in (case.xx, case.yy, case.zz, case.dx, case.dy, case.dz)]
        h = C.c_void_p()
        rc = self.L.mhd_create(C.byref(cfg), *[_p(a) for a in geo], C.byref(h))
        self.h = h
        self._check(rc)
        self.shape3 = (case.mz + 4, case.my, case.mx + 4)
        if case.x0 is None:
            # a Harris case without host arrays: the device builds them
            self.init_harris(case.harris)
            return
        eq = [np.ascontiguousarray(a, dtype=np.float64)
              for a in (case.rhoi, case.bxi, case.byi, case.bzi, case.pri)]
        self._check(self.L.mhd_set_equilibrium(self.h, *[_p(a) for a in eq]))
        face = getattr(case, "inflow_face", None)
        if face is not None:
            self.set_inflow_face(*face)
        self.set_state(case.x0)

    def _check(self, rc):
        if rc != MHD_OK:
            msg = self.L.mhd_last_error(self.h).decode() if self.h else ""
            raise MHDError(rc, msg)

    def close(self):
        """free the handle; while an open MHDEnsemble still uses it the
        handle is freed when the last such ensemble closes"""
        if getattr(self, "_groups", 0) > 0:
            self._close_pending = True
            return
        if getattr(self, "h", None):
            self.L.mhd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _get(self, fn, nf):
        out = np.empty((nf,) + self.shape3)
        self._check(getattr(self.L, fn)(self.h, _p(out)))
        return out

    def init_harris(self, params: dict):
        """the Harris equilibrium and initial state built on the device
        (mhd_init_harris) from the mesh and harris_case's parameters"""
        hp = MHDHarris(**{k: v for k, v in params.items() if k != "seed"}, seed=int(params["seed"]))
        self._check(self.L.mhd_init_harris(self.h, C.byref(hp)))

    def set_state(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        assert a.shape == (8,) + self.shape3
        self._check(self.L.mhd_set_state(self.h, _p(a)))

    def get_state(self):
        return self._get("mhd_get_state", 8)

    def get_current(self):
        return self._get("mhd_get_current", 3)

    def get_divb(self):
        return self._get("mhd_get_divb", 1)[0]

    def get_xdif(self):
        return self._get("mhd_get_xdif", 8)

    def get_residual(self):
        return self._get("mhd_get_residual", 3)

    def set_residual(self, r):
        a = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self.L.mhd_set_residual(self.h, _p(a)))

    def init_residual(self):
        self._check(self.L.mhd_init_residual(self.h))

    def compute_dt(self) -> float:
        v = C.c_double()
        self._check(self.L.mhd_compute_dt(self.h, C.byref(v)))
        return v.value

    def compute_dt_noalloc(self) -> float:
        """setdt with its partial minima in the library's fixed device array
        (the call allocates nothing)"""
        v = C.c_double()
        self._check(self.L.mhd_compute_dt_noalloc(self.h, C.byref(v)))
        return v.value

    def step(self, time: float, dt: float):
        self._check(self.L.mhd_step(self.h, time, dt))

    def rhs(self, t: float):
        self._check(self.L.mhd_rhs(self.h, t))

    def rhs_stage(self, t: float, ddt: float, km: int):
        """right(t, ddt, km): the stage arguments the driven inflow reads"""
        self._check(self.L.mhd_rhs_stage(self.h, t, ddt, int(km)))

    def set_inflow_face(self, bx, by, bz):
        """inflow face data bxinfc, byinfc, bzinfc: each (mz+4, my) C order"""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (bx, by, bz)]
        for a in arrs:
            assert a.shape == (self.case.mz + 4, self.case.my)
        self._check(self.L.mhd_set_inflow_face(self.h, *[_p(a) for a in arrs]))

    def advance(self, nsteps: int, time: float = 0.0, dt0: float = 0.0):
        """nsteps of setdt+stepon with dt on the device; returns
        (time, last dt, elapsed ms on the handle's stream)."""
        t = C.c_double(time)
        dl = C.c_double()
        ms = C.c_double()
        self._check(self.L.mhd_advance(self.h, int(nsteps), dt0, C.byref(t), C.byref(dl),
                                       C.byref(ms)))
        return t.value, dl.value, ms.value

    def diagnostics(self) -> dict:
        d = MHDDiag()
        rc = self.L.mhd_diagnostics(self.h, C.byref(d))
        if rc not in (MHD_OK, -5):
            self._check(rc)
        out = d.as_dict()
        out["status"] = rc
        return out

    def set_profiling(self, on: bool):
        self._check(self.L.mhd_set_profiling(self.h, int(on)))

    def kernel_times(self) -> dict:
        n = 32
        names = (C.c_char_p * n)()
        ms = (C.c_double * n)()
        calls = (C.c_int * n)()
        k = self.L.mhd_kernel_times(self.h, n, names, ms, calls)
        return {names[i].decode(): (ms[i], calls[i]) for i in range(k)}

    # ---- reverse mode ----
    def adjoint_rhs(self, xdif_bar, *, init: bool = False, smooth: bool = True):
        """(x_bar, xresd_bar, eta_bar): vector-Jacobian product of one
        right(t) at the current state (smooth = t != 0; init: the residual
        capture)"""
        ab = np.ascontiguousarray(xdif_bar, dtype=np.float64)
        assert ab.shape == (8,) + self.shape3
        xb = np.empty((8,) + self.shape3)
        rb = np.empty((3,) + self.shape3)
        eb = C.c_double()
        self._check(self.L.mhd_adjoint_rhs(self.h, _p(ab), _p(xb), _p(rb), C.byref(eb), int(init), int(smooth)))
        return xb, rb, eb.value

    def adjoint_run(self, dts, objective="magnetic_energy", *, every: int = 0, residual: bool = True):
        """Discrete adjoint of len(dts) steps from the current state with the
        frozen dt sequence: returns (value, d value/d x0, d value/d eta0,
        (forward ms, reverse ms)); the state ends at x_N."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        obj = OBJECTIVES[objective] if isinstance(objective, str) else int(objective)
        gx = np.empty((8,) + self.shape3)
        val, ge = C.c_double(), C.c_double()
        ms = (C.c_double * 2)()
        self._check(self.L.mhd_adjoint_run(self.h, int(d.size), _p(d), obj, int(every),
                                           MHD_ADJ_RESIDUAL if residual else 0, C.byref(val), _p(gx),
                                           C.byref(ge), ms))
        return val.value, gx, ge.value, (ms[0], ms[1])

    def run(self, steps: int, *, time: float = 0.0, init: bool = True):
        """host loop mirroring t3dsdsc.f:136-176 (setdt, stepon, time += dt)"""
        if init:
            self.init_residual()
        dts = []
        for _ in range(steps):
            dt = self.compute_dt()
            self.step(time, dt)
            time += dt
            dts.append(dt)
        return time, dts


class MHDEnsemble:
    """Independent runs (MHDStep objects, owned by the caller) advanced
    together: every run does setdt + stepon with its own device dt, all runs
    concurrently on their own streams."""

    def __init__(self, runs):
        self.L = load_library()
        self.runs = list(runs)
        self._held = []
        arr = (C.c_void_p * len(self.runs))(*[r.h for r in self.runs])
        e = C.c_void_p()
        rc = self.L.mhd_ensemble_create(arr, len(self.runs), C.byref(e))
        self.e = e
        if rc != MHD_OK:
            msg = self.L.mhd_ensemble_error(e).decode() if e else ""
            self.close()
            raise MHDError(rc, msg)
        # the runs' handles must outlive the ensemble (mhd_ensemble_destroy
        # restores their chunking and synchronises their streams)
        for r in self.runs:
            r._groups += 1
            self._held.append(r)

    def advance(self, nsteps: int, times, dt0s):
        """nsteps for every run from times[r] (dt0s[r] arms the collapse
        guard); returns (new times, elapsed ms of the batch)"""
        t = np.ascontiguousarray(times, dtype=np.float64).copy()
        d = np.ascontiguousarray(dt0s, dtype=np.float64)
        assert t.shape == d.shape == (len(self.runs),)
        ms = C.c_double()
        rc = self.L.mhd_ensemble_advance(self.e, int(nsteps), _p(d), _p(t), C.byref(ms))
        if rc != MHD_OK:
            raise MHDError(rc, self.L.mhd_ensemble_error(self.e).decode())
        return t, ms.value

    def tune(self, trial: int, times, dt0s):
        """time `trial` batch steps under each z-chunk policy and keep the
        fastest; returns (choice, trial ms per policy).  States and times
        are unchanged."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        d = np.ascontiguousarray(dt0s, dtype=np.float64)
        best = C.c_int()
        ms = np.zeros(4)
        rc = self.L.mhd_ensemble_tune(self.e, int(trial), _p(d), _p(t), C.byref(best), _p(ms))
        if rc != MHD_OK:
            raise MHDError(rc, self.L.mhd_ensemble_error(self.e).decode())
        return best.value, [float(m) for m in ms]

    def tune_shapes(self, trial: int, times, dt0s, max_trials: int = 32):
        """coordinate search over the runs' tile shapes (smoothing and tile
        z chunks, face-window rows, replay form); returns (best shape,
        every shape tried), each a dict with its trial time 'ms'.  The best
        stays in force; states and times are unchanged."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        d = np.ascontiguousarray(dt0s, dtype=np.float64)
        tr = (MHDTileShape * max_trials)()
        nt = C.c_int()
        best = MHDTileShape()
        rc = self.L.mhd_ensemble_tune_shapes(self.e, int(trial), _p(d), _p(t), max_trials, tr, C.byref(nt),
                                             C.byref(best))
        if rc != MHD_OK:
            raise MHDError(rc, self.L.mhd_ensemble_error(self.e).decode())
        return best.as_dict(), [tr[i].as_dict() for i in range(min(nt.value, max_trials))]

    def set_shape(self, kc: int, zkc: int, face_rows: int = 12, form: int = 0, rhs_rows: int = -1,
                  wave_cells: int = 0):
        """form: FORM_RUNS, FORM_BATCH_GRAPH or FORM_KERNELS; rhs_rows: right
        part 2 as the tile kernel (0), 16-row (1) or 8-row (2) windows, or
        each run's own (-1); wave_cells: most cells of one wave of batched
        launches (0: no limit)"""
        sh = MHDTileShape(kc=kc, zkc=zkc, face_rows=face_rows, form=int(form), rhs_rows=int(rhs_rows),
                          wave_cells=int(wave_cells), ms=0.0)
        rc = self.L.mhd_ensemble_set_shape(self.e, C.byref(sh))
        if rc != MHD_OK:
            raise MHDError(rc, self.L.mhd_ensemble_error(self.e).decode() if self.e else "")

    def status(self):
        st = (C.c_int * len(self.runs))()
        self.L.mhd_ensemble_status(self.e, st)
        return list(st)

    def close(self):
        if getattr(self, "e", None):
            self.L.mhd_ensemble_destroy(self.e)
            self.e = None
        held, self._held = getattr(self, "_held", []), []
        for r in held:
            r._groups -= 1
            if r._groups == 0 and r._close_pending:
                r._close_pending = False
                r.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
