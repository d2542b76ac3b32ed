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


def test_c4_set_matches_single_runs_and_oracle():
    """The 48-run C4 set (64x32x64 .. 192x96x192, eta, d_i, amplitude and
    sheet thickness drawn from seed 4) through the ensemble for 3 steps (a
    captured step pair and an eager odd step): every run bit-identical to
    mhd_advance on that run alone (the run's own z chunking there, the
    ensemble's long chunks here), and the smallest run, the largest run and
    the run with the smallest d_i against the oracle."""
    import time
    from dasha_amd.mhd import MHDEnsemble, MHDStep
    from oracle.oracle import Oracle
    from tests.util import field_errors
    t0 = time.time()
    cases = harris.case_c4()
    print(f"[c4] {len(cases)} runs, {sum(c.cells for c in cases) / 1e6:.1f}M cells, "
          f"built in {time.time() - t0:.1f} s", flush=True)
    steps = 3
    alone = []
    for c in cases:
        g = MHDStep(c)
        g.init_residual()
        dt0 = g.compute_dt()
        tn, _, _ = g.advance(steps, 0.0, dt0)
        alone.append((dt0, tn, g.get_state()))
        g.close()
    print(f"[c4] runs alone done ({time.time() - t0:.1f} s)", flush=True)
    members = [MHDStep(c) for c in cases]
    for g in members:
        g.init_residual()
    ens = MHDEnsemble(members)
    t_ens, ms = ens.advance(steps, np.zeros(len(cases)), [a[0] for a in alone])
    print(f"[c4] ensemble {steps} steps {ms:.1f} ms ({time.time() - t0:.1f} s)", flush=True)
    assert ens.status() == [0] * len(cases)
    for r, (g, (_, tn, x)) in enumerate(zip(members, alone)):
        assert t_ens[r] == tn, r
        assert np.array_equal(g.get_state(), x), r
    pick = {min(range(len(cases)), key=lambda r: cases[r].cells),
            max(range(len(cases)), key=lambda r: cases[r].cells),
            min(range(len(cases)), key=lambda r: cases[r].di)}
    for r in sorted(pick):
        o = Oracle(cases[r])
        to, _ = o.run(steps)
        e = field_errors(alone[r][2], o.get_state())
        print(f"[c4] run {cases[r].name} d_i {cases[r].di:.4f} eta {cases[r].eta0:.4f}: "
              f"rel L2 vs oracle {max(e):.2e}", flush=True)
        assert abs(to - t_ens[r]) <= 1e-13 * to
        assert max(e) < 1e-12
        o.close()
    ens.close()
    for g in members:
        g.close()


def test_member_close_waits_for_the_ensemble():
    """closing a run that an open ensemble still uses defers the free until
    the ensemble closes (the ensemble's destroy touches every run's handle)"""
    from dasha_amd.mhd import MHDEnsemble, MHDStep
    cases = _cases()[:2]
    members = [MHDStep(c) for c in cases]
    for g in members:
        g.init_residual()
    ens = MHDEnsemble(members)
    members[0].close()
    assert members[0].h                      # still alive: the ensemble uses it
    t, _ = ens.advance(2, np.zeros(2), [0.0, 0.0])
    assert all(v > 0 for v in t)
    ens.close()
    assert not members[0].h                  # freed with the ensemble
    assert members[1].h
    members[1].close()
    assert not members[1].h


@pytest.mark.parametrize("shape", [dict(kc=8, zkc=16, face_rows=16, form=1),
                                   dict(kc=64, zkc=48, face_rows=12, form=0, rhs_rows=1),
                                   dict(kc=16, zkc=24, face_rows=12, form=1, rhs_rows=2),
                                   dict(kc=32, zkc=32, face_rows=12, form=2),
                                   dict(kc=8, zkc=64, face_rows=16, form=2, rhs_rows=1),
                                   dict(kc=16, zkc=16, face_rows=12, form=2, rhs_rows=2, more=True),
                                   dict(kc=16, zkc=32, face_rows=12, form=2, wave_cells=5000, more=True),
                                   dict(kc=32, zkc=24, face_rows=12, form=2, wave_cells=9000, more=True)])
def test_shapes_and_forms_match_single_runs(shape):
    """any tile shape and every replay form (per-run graphs, the
    whole-batch graph, the batched kernels) give the runs alone, bit for bit
    (pairs from both pointer states and an odd step).  The runs alone use
    the tile kernel for right part 2, so the windowed kernels (rhs_rows 1,
    2) must give its values exactly.  The three runs differ
    in their launch sequences (Hall off in one, non-periodic y with 16-row
    face windows in another), so the batched kernels run three groups; with
    ``more`` two runs of each kind share their groups, and a wave budget
    (wave_cells) splits them into waves of one run or keeps the pairs."""
    from dasha_amd.mhd import MHDEnsemble, MHDStep
    shape = dict(shape)
    cases = _cases()
    if shape.pop("more", False):
        cases += [harris.case_small(28, 8, 20, seed=24, di=0.2),
                  harris.case_small(16, 8, 24, stretched=False, seed=25, di=0.0),
                  harris.case_small(20, 12, 24, periodic_y=False, seed=26, di=0.1)]
    print(f"[ensemble] shape {shape}", flush=True)
    singles = [MHDStep(c) for c in cases]
    members = [MHDStep(c) for c in cases]
    for g in singles + members:
        g.init_residual()
    dt0 = [g.compute_dt() for g in singles]
    ens = MHDEnsemble(members)
    ens.set_shape(**shape)
    t_ens = np.zeros(len(cases))
    for nsteps in (4, 3, 2):
        t_ref = [g.advance(nsteps, float(t), d)[0] for g, d, t in zip(singles, dt0, t_ens)]
        t_ens, _ = ens.advance(nsteps, t_ens, dt0)
        assert list(t_ens) == t_ref
        for a, b in zip(singles, members):
            assert np.array_equal(a.get_state(), b.get_state())
    ens.close()


def test_shape_tuner_keeps_states():
    from dasha_amd.mhd import MHDEnsemble, MHDStep
    cases = _cases()
    members = [MHDStep(c) for c in cases]
    for g in members:
        g.init_residual()
    dt0 = [g.compute_dt() for g in members]
    x_before = [g.get_state() for g in members]
    ens = MHDEnsemble(members)
    best, trials = ens.tune_shapes(2, np.zeros(len(cases)), dt0)
    print(f"[ensemble] tuned shape {best}; {len(trials)} trials", flush=True)
    assert len(trials) == 18 and all(t["ms"] > 0 for t in trials)
    assert sorted({t["form"] for t in trials}) == [0, 1, 2]
    assert best["ms"] == min(t["ms"] for t in trials)
    for g, x in zip(members, x_before):
        assert np.array_equal(g.get_state(), x)
    ens.close()


def test_batched_kernels_refuse_viscosity():
    """a run with viscosity has no batched form: the kernels form refuses
    the ensemble and leaves the states alone; the per-run form still runs it"""
    from dasha_amd.mhd import FORM_KERNELS, FORM_RUNS, MHDEnsemble, MHDError, MHDStep
    cases = _cases()[:2]
    cases[1].fmu0 = 1e-3
    members = [MHDStep(c) for c in cases]
    for g in members:
        g.init_residual()
    dt0 = [g.compute_dt() for g in members]
    x_before = [g.get_state() for g in members]
    ens = MHDEnsemble(members)
    ens.set_shape(32, 32, 12, FORM_KERNELS)
    with pytest.raises(MHDError):
        ens.advance(2, np.zeros(2), dt0)
    for g, x in zip(members, x_before):
        assert np.array_equal(g.get_state(), x)
    ens.set_shape(32, 32, 12, FORM_RUNS)
    t, _ = ens.advance(2, np.zeros(2), dt0)
    assert all(v > 0 for v in t)
    ens.close()
