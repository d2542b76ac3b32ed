"""Central differences of the CPU oracle on the adjoint workload, written to
c5_fd_oracle.json beside this script.

Workload: 128x64x128 tanh-stretched Harris sheet, eta 0.005, d_i 0.15, the
control dAy = 0.1 exp(-(x^2 + z^2)), 500 RK4 steps over the dt sequence of the
oracle's own setdt run from the nominal state (then frozen).  Step h = 1e-6:
dAy +- h v with max |v| = 1 (dasha_amd.adjoint.fd_direction, seed 7), and
eta +- h.

Legs (each a pair of 500-step oracle runs, ~150 s each on 16 host threads):
  nominal  the dt sequence and both objective values at T (one run)
  energy   d(magnetic energy)/d(dAy) along the white-noise direction
  flux     d(reconnected flux)/d(dAy) along the modulated control shape
  eta      d(objective)/d(eta) for both objectives from one pair of runs

    python tests/golden/make_c5_fd_oracle.py --legs nominal,energy --out a.json
    python tests/golden/make_c5_fd_oracle.py --legs flux,eta --fixture a.json --out b.json

--fixture supplies the legs already run (and the dt sequence); the output
holds them and the new legs.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dasha_amd import adjoint  # noqa: E402
from oracle import oracle as orc  # noqa: E402

H = 1e-6
SEED = 7
OBJECTIVES = ("magnetic_energy", "reconnected_flux")


def run(case, x0, dts=None, steps=0, label=""):
    """final state (and the dt sequence when dts is None) of an oracle run"""
    o = orc.Oracle(case)
    try:
        o.set_state(x0)
        o.init_residual()
        t, out = 0.0, []
        n = len(dts) if dts is not None else steps
        t0 = time.time()
        for i in range(n):
            dt = o.compute_dt() if dts is None else dts[i]
            o.step(t, dt)
            t += dt
            out.append(dt)
            if (i + 1) % 100 == 0:
                print(f"[fd] {label} step {i + 1}/{n} {time.time() - t0:.0f} s", flush=True)
        return o.get_state(), out
    finally:
        o.close()


def direction_checksum(v):
    return [float(np.sum(v)), float(np.sum(v * v))]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--legs", default="nominal,energy,flux,eta")
    ap.add_argument("--fixture", default=None, help="legs already run (JSON from an earlier call)")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "c5_fd_oracle.json"))
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    orc.build()
    if a.threads:
        orc.set_threads(a.threads)
    fx = {}
    if a.fixture:
        with open(a.fixture) as f:
            fx = json.load(f)
    ac = adjoint.case_c5()
    case = ac.case
    fx.setdefault("workload", f"{case.mx}x{case.ny}x{case.mz} stretched Harris sheet, eta {case.eta0}, "
                              f"d_i {case.di}, {ac.steps} RK4 steps, dAy = 0.1 exp(-(x^2+z^2))")
    fx.update(h=H, seed=SEED, threads=orc.max_threads())
    for leg in a.legs.split(","):
        t0 = time.time()
        if leg == "nominal":
            x, dts = run(case, ac.state(), steps=ac.steps, label="nominal")
            fx["dts"] = dts
            fx["value"] = {o: adjoint.objective_of_state(case, x, o) for o in OBJECTIVES}
        elif leg in ("energy", "flux"):
            obj = OBJECTIVES[0] if leg == "energy" else OBJECTIVES[1]
            v = adjoint.fd_direction(ac, obj, SEED)
            xp, _ = run(case, ac.state(ac.ay0 + H * v), fx["dts"], label=f"{leg} +h")
            xm, _ = run(case, ac.state(ac.ay0 - H * v), fx["dts"], label=f"{leg} -h")
            fx.setdefault("direction_checksum", {})[obj] = direction_checksum(v)
            fx.setdefault("fd_ay", {})[obj] = adjoint.central_difference(case, xp, xm, H, obj)
            fx.setdefault("fd_ay_two_sums", {})[obj] = (adjoint.objective_of_state(case, xp, obj) -
                                                        adjoint.objective_of_state(case, xm, obj)) / (2 * H)
        elif leg == "eta":
            e0 = case.eta0
            xp, _ = run(adjoint.with_eta(ac, e0 + H).case, ac.state(), fx["dts"], label="eta +h")
            xm, _ = run(adjoint.with_eta(ac, e0 - H).case, ac.state(), fx["dts"], label="eta -h")
            fx["fd_eta"] = {o: adjoint.central_difference(case, xp, xm, H, o) for o in OBJECTIVES}
            fx["fd_eta_two_sums"] = {o: (adjoint.objective_of_state(case, xp, o) -
                                         adjoint.objective_of_state(case, xm, o)) / (2 * H) for o in OBJECTIVES}
        else:
            raise SystemExit(f"unknown leg {leg}")
        fx.setdefault("seconds", {})[leg] = round(time.time() - t0, 1)
        with open(a.out, "w") as f:
            json.dump(fx, f, indent=1)
        print(f"[fd] leg {leg} done in {time.time() - t0:.0f} s -> {a.out}", flush=True)


if __name__ == "__main__":
    main()
