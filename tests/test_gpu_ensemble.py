"""The ensemble entry points against the same runs advanced one by one.

Every run of an ensemble executes the same kernels in the same order as
mhd_advance on that run alone, so states, times and dts must agree bit for
bit, through the captured step pairs, an odd trailing step and a second
call that continues from the other pointer state.
"""
import numpy as np
import pytest

from dasha_amd import harris

pytestmark = pytest.mark.gpu


def _cases():
    return [harris.case_small(24, 8, 24, seed=21),
            harris.case_small(20, 8, 28, stretched=False, seed=22, di=0.0),
            harris.case_small(24, 12, 20, periodic_y=False, seed=23, di=0.3)]


def test_ensemble_matches_single_runs():
    from dasha_amd.mhd import MHDEnsemble, MHDStep
    cases = _cases()
    print("[ensemble] runs", [c.name for c in cases], flush=True)
    singles = [MHDStep(c) for c in cases]
    members = [MHDStep(c) for c in cases]
    for g in singles + members:
        g.init_residual()
    dt0 = [g.compute_dt() for g in singles]
    ens = MHDEnsemble(members)
    t_ens = np.zeros(len(cases))
    for nsteps in (5, 4, 1):
        t_ref = []
        for g, d, t in zip(singles, dt0, t_ens):
            tn, _, _ = g.advance(nsteps, float(t), d)
            t_ref.append(tn)
        t_ens, ms = ens.advance(nsteps, t_ens, dt0)
        print(f"[ensemble] {nsteps} steps: {ms:.2f} ms", flush=True)
        assert list(t_ens) == t_ref
        for a, b in zip(singles, members):
            assert np.array_equal(a.get_state(), b.get_state())
    assert ens.status() == [0, 0, 0]
    ens.close()


def test_ensemble_rejects_bad_input():
    from dasha_amd.mhd import MHDEnsemble, MHDError, MHDStep
    g = MHDStep(_cases()[0])
    with pytest.raises(MHDError):
        MHDEnsemble([g, g])          # a run appears once
    with pytest.raises(MHDError):
        MHDEnsemble([])


def test_tuned_ensemble_matches_single_runs():
    """the tuner's trials leave states and times alone, and whatever chunk
    policy it keeps, results stay bit-identical to the runs alone"""
    from dasha_amd.mhd import MHDEnsemble, MHDStep
    cases = _cases()
    print("[ensemble] tune", [c.name for c in cases], flush=True)
    singles = [MHDStep(c) for c in cases]
    members = [MHDStep(c) for c in cases]
    for g in singles + members:
        g.init_residual()
    dt0 = [g.compute_dt() for g in singles]
    x_before = [g.get_state() for g in members]
    ens = MHDEnsemble(members)
    best, ms = ens.tune(3, np.zeros(len(cases)), dt0)
    print(f"[ensemble] tuned policy {best}, trial ms {ms}", flush=True)
    assert 0 <= best < 4 and all(m > 0 for m in ms)
    for g, x in zip(members, x_before):
        assert np.array_equal(g.get_state(), x)
    t_ens, _ = ens.advance(5, np.zeros(len(cases)), dt0)
    for g, d, t in zip(singles, dt0, t_ens):
        tn, _, _ = g.advance(5, 0.0, d)
        assert tn == t
    for a, b in zip(singles, members):
        assert np.array_equal(a.get_state(), b.get_state())
    ens.close()
