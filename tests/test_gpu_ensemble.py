"""Ensembles of independent 2D runs in one forest (one handle, one launch
per kernel for all runs): every run must compute exactly what it computes
alone.

The reference runs such a study as separate processes
(programs/parameter_study_2d/run_study.sh:1-5).  The single-run device
path equals the oracle bit for bit (test_gpu_parity.py); here each run of
the forest is compared with a single-run handle of the same case (and, for
the plain form, with the oracle directly), after the initial field solve
and after every time step: densities, potential and field bit for bit, and
the step results (dt, rejections, V-cycle counts) equal.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("CC_ELECTRON", "CC_POS_ION", "CC_NEG_ION", "CC_PHI", "CC_EFIELD")


def params_for(A, p, dt_safety=None, dt_max=None):
    q = A.ParamsT.from_buffer_copy(p)
    if dt_safety is not None:
        q.dt_safety_factor = dt_safety
    if dt_max is not None:
        q.dt_max = dt_max
    return q


def forest_and_singles(A, S, ens, precision, params=None, oracle=False):
    g = A.Path("gpu", 2, 8, A.CYL, precision)
    S.apply_ensemble(g, ens)
    if params is not None:
        g.set_params(params)
    singles = []
    for case in ens.cases:
        kind = "oracle" if oracle else "gpu"
        s = A.Path(kind, 2, 8, A.CYL, precision if kind == "gpu" else A.PREC_FP64)
        S.apply(s, case)
        if params is not None:
            s.set_params(params)
        singles.append(s)
    return g, singles


def compare_runs(A, ens, g, singles, what, exact=True, tol=0.0):
    for r, s in enumerate(singles):
        ids = ens.run_ids(r)
        for name in STATE:
            iv = getattr(A, name)
            a, b = g.get_cc(iv, ids), s.get_cc(iv)
            if exact:
                assert np.array_equal(a, b), "%s run %d %s differs (max %.3e)" % (
                    what, r, name, np.abs(a - b).max())
            else:
                scale = max(np.abs(b).max(), 1e-300)
                d = np.abs(a - b).max() / scale
                assert d <= tol, "%s run %d %s rel diff %.3e" % (what, r, name, d)


STEP_KEYS = ("it", "dt_used", "dt_next", "dt_lim", "n_rejected", "n_vcycles", "residual", "leaf_cells")


def run_and_compare(A, ens, g, singles, n_steps, what, exact=True, tol=0.0, dt_start=None):
    n_g, r_g = g.field_compute_runs(0, have_guess=False)
    for r, s in enumerate(singles):
        n_s, r_s = s.field_compute(0, have_guess=False)
        assert (int(n_g[r]), float(r_g[r])) == (n_s, r_s), (what, r, n_g[r], n_s, r_g[r], r_s)
    compare_runs(A, ens, g, singles, what + " field", exact, tol)
    for r, dt in (dt_start or {}).items():  # a restart with the given global dt
        g.set_run_time_state(r, 0, 0.0, dt)
        singles[r].set_time_state(0, 0.0, dt)
        assert g.get_run_time_state(r) == (0, 0.0, dt)
    rejected = 0
    for it in range(n_steps):
        infos = g.step_ensemble()
        for r, s in enumerate(singles):
            ref = s.step()
            got = {k: infos[r][k] for k in STEP_KEYS}
            exp = {k: ref[k] for k in STEP_KEYS}
            assert got == exp, (what, it, r, got, exp)
            rejected += ref["n_rejected"]
        print(what, "step", it, "dt", [round(i["dt_used"] * 1e12, 4) for i in infos],
              "vcycles", [i["n_vcycles"] for i in infos], flush=True)
        compare_runs(A, ens, g, singles, "%s step %d" % (what, it), exact, tol)
    return rejected


def test_forest_fp64_equals_oracle_per_run(A, S):
    ens = S.ensemble(level_list=[3, 4, 3, 2], num_vcycles=20, rel_residual=1e-10)
    print(ens.name, ens.voltages, flush=True)
    g, singles = forest_and_singles(A, S, ens, A.PREC_FP64, oracle=True)
    run_and_compare(A, ens, g, singles, 4, "fp64 vs oracle")


def test_forest_fp64_equals_single_runs(A, S):
    ens = S.ensemble(level_list=[5, 4, 6, 4, 5, 3], num_vcycles=20, rel_residual=1e-10)
    print(ens.name, flush=True)
    g, singles = forest_and_singles(A, S, ens, A.PREC_FP64)
    run_and_compare(A, ens, g, singles, 5, "fp64")


def test_forest_rejected_steps_redone_per_run(A, S):
    """Rejections are redone for the rejected runs only (the others are
    masked out): runs 1 and 3 restart with a dt far above their limits
    (dielectric relaxation below dt_max with a 1e21 seed), so their first
    step is rejected and redone; the safety factor 1 lets later steps
    reject whenever a limit shrinks."""
    ens = S.ensemble(level_list=[4, 5, 4, 5], num_vcycles=20, rel_residual=1e-10, seed_density=1e21)
    p = params_for(A, ens.cases[0].params, dt_safety=1.0)
    g, singles = forest_and_singles(A, S, ens, A.PREC_FP64, params=p)
    rejected = run_and_compare(A, ens, g, singles, 6, "fp64 rejections", dt_start={1: 5e-12, 3: 5e-12})
    print("rejected attempts", rejected, flush=True)
    assert rejected > 0, "the case was meant to reject some steps"


def test_forest_mixed_equals_single_runs(A, S):
    """Mixed precision: the runs share the batched exact solve of their
    level 5 (128^2 per run, one matrix-core transform for all runs) and
    stop on their own residual tests; each run equals its single-run form."""
    ens = S.ensemble(level_list=[5, 6, 5, 6], num_vcycles=20, rel_residual=1e-8)
    g, singles = forest_and_singles(A, S, ens, A.PREC_MIXED)
    run_and_compare(A, ens, g, singles, 4, "mixed")


@pytest.mark.parametrize("prec", ["PREC_FP64", "PREC_MIXED"])
def test_forest_step_graph_equals_single_runs(A, S, prec):
    """Each step's first attempt and field solve as one captured graph with
    the stopping tests on the device: the first steps hold fewer cycles than
    some runs need (they go on from the device's records), a run restarted
    with a too large dt is rejected and redone, and every run still equals
    its single-run form bit for bit.  (In mixed precision every run is as
    deep as the forest's exact-solve level, 128^2, as alone; a shallower run
    would give the whole forest a coarser exact-solve level, and so other
    correction iterates than that run alone: a valid but different solve.)"""
    levels = [5, 4, 6, 3] if prec == "PREC_FP64" else [5, 6, 5, 6]
    ens = S.ensemble(level_list=levels, num_vcycles=20, rel_residual=1e-10, seed_density=1e21)
    g, singles = forest_and_singles(A, S, ens, getattr(A, prec))
    g.set_option(A.OPT_STEP_GRAPH, 1)
    rejected = run_and_compare(A, ens, g, singles, 5, "step graph " + prec, dt_start={2: 5e-12})
    assert rejected > 0


@pytest.mark.parametrize("prec", ["PREC_FP64", "PREC_MIXED"])
def test_forest_photoionization_equals_single_runs(A, S, prec):
    """Photoionization in a forest (src/m_photoi_helmh.f90:162-191): each
    run solves the three Helmholtz modes to its own stopping test (its own
    max |rhs|) and is masked out of a mode's cycles once it has converged,
    so every run equals its single-run handle bit for bit: the photo
    source, the modes, the state and the step records, photoionization
    cycles included.  Run 2 restarts at another step count, so with
    photoionization every second step the runs are due at different
    steps."""
    ens = S.ensemble(level_list=[5, 6, 5, 5], num_vcycles=20, rel_residual=1e-8, photoi=True)
    p = params_for(A, ens.cases[0].params)
    p.photoi_per_steps = 2
    g, singles = forest_and_singles(A, S, ens, getattr(A, prec), params=p)
    n_g, r_g = g.field_compute_runs(0, have_guess=False)
    for r, s in enumerate(singles):
        assert (int(n_g[r]), float(r_g[r])) == s.field_compute(0, have_guess=False)
    g.set_run_time_state(2, 1, 0.0, 1e-13)
    singles[2].set_time_state(1, 0.0, 1e-13)
    g.photoi_set_src()
    for s in singles:
        s.photoi_set_src()
    photo = (A.CC_PHOTO, A.CC_HELMH, A.CC_HELMH + 1, A.CC_HELMH + 2)
    for r, s in enumerate(singles):
        for iv in photo:
            assert np.array_equal(g.get_cc(iv, ens.run_ids(r)), s.get_cc(iv)), ("set_src", r, iv)
    keys = STEP_KEYS + ("photoi_done", "photoi_cycles")
    due = set()
    for it in range(4):
        infos = g.step_ensemble()
        for r, s in enumerate(singles):
            ref = s.step()
            got = {k: infos[r][k] for k in keys}
            exp = {k: ref[k] for k in keys}
            assert got == exp, (prec, it, r, got, exp)
            if ref["photoi_done"]:
                due.add((it, r))
                assert ref["photoi_cycles"] > 0
        compare_runs(A, ens, g, singles, "%s photoi step %d" % (prec, it))
        for r, s in enumerate(singles):
            for iv in photo:
                assert np.array_equal(g.get_cc(iv, ens.run_ids(r)), s.get_cc(iv)), (it, r, iv)
    # both step parities had runs due (run 2 out of phase with the others)
    assert {it % 2 for it, r in due if r == 2} != {it % 2 for it, r in due if r == 0}


def test_forest_mixed_vs_fp64_per_run_error(A, S):
    """Per-run error of the mixed form against the FP64 form (the oracle's
    arithmetic) after a few steps of an uneven ensemble."""
    ens = S.ensemble(level_list=[5, 6, 7, 5], num_vcycles=20, rel_residual=1e-8)
    gm = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL, A.PREC_MIXED), ens)
    gd = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL, A.PREC_FP64), ens)
    for p in (gm, gd):
        p.field_compute_runs(0, have_guess=False)
    for it in range(5):
        im, idd = gm.step_ensemble(), gd.step_ensemble()
        for a, b in zip(im, idd):  # dt follows each form's own limits
            assert abs(a["dt_used"] - b["dt_used"]) <= 1e-6 * b["dt_used"]
    for r in range(len(ens.cases)):
        ids = ens.run_ids(r)
        for name, tol in (("CC_PHI", 1e-7), ("CC_EFIELD", 1e-5), ("CC_ELECTRON", 1e-7)):
            iv = getattr(A, name)
            a, b = gm.get_cc(iv, ids), gd.get_cc(iv, ids)
            d = np.abs(a - b).max() / np.abs(b).max()
            print("run %d %s max rel error %.3e" % (r, name, d), flush=True)
            assert d < tol, (r, name, d)


def test_forest_refuses_single_run_calls(A, S):
    ens = S.ensemble(level_list=[2, 3], num_vcycles=20)
    g = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL), ens)
    assert g.num_runs() == 2
    with pytest.raises(A.AfError):
        g.step()
    with pytest.raises(A.AfError):
        g.set_run_voltages([1.0, 2.0, 3.0])


def test_forest_rejected_steps_equal_oracle(A, S):
    """Rejected steps (the backup restored, copy_current_state /
    restore_previous_state, src/streamer.f90:254-294, 659-688) against the
    oracle: run 1 restarts with a dt far above its limits, so its first step
    is rejected and redone; every run's steps and state equal the oracle's
    bit for bit."""
    ens = S.ensemble(level_list=[4, 5, 4], num_vcycles=20, rel_residual=1e-10, seed_density=1e21)
    p = params_for(A, ens.cases[0].params, dt_safety=1.0)
    g, singles = forest_and_singles(A, S, ens, A.PREC_FP64, params=p, oracle=True)
    rejected = run_and_compare(A, ens, g, singles, 5, "fp64 rejections vs oracle", dt_start={1: 5e-12})
    print("rejected attempts", rejected, flush=True)
    assert rejected > 0
