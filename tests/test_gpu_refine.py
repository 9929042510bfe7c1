"""Refinement flags on the device equal the oracle's, box for box.

afgpu_refine_flags evaluates default_refinement (src/m_refine.f90:198-298)
on every cell of the leaves and of the boxes with a leaf child on the
device and finishes consistent_ref_flags (afivo/src/m_af_core.f90:929-1145)
on the host; the oracle does all of it on the CPU in the reference's order
(tests/test_oracle_refine.py pins it).  Both read the same FP64 state (the
plain form equals the oracle bit for bit), so the flags must be equal on
every box, for every setting below: buffer widths, alpha or effective
alpha, thresholds that refine and derefine, the initial-seed rule, fixed
regions and limits, and the spacing bounds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REFINE, DEREFINE = 2, -2


def cases(S):
    return {
        "cyl2d": S.small_amr(2, coord=2, uniform_lvl=3, max_lvl=5, frac=0.3, seed=3),
        "xyz3d_8": S.small_amr(3, nc=8, uniform_lvl=2, max_lvl=4, frac=0.3, seed=11),
        "xyz3d_16": S.small_amr(3, nc=16, uniform_lvl=2, max_lvl=3, frac=0.3, seed=7),
    }


def settings(A, case):
    """A mix of settings; the thresholds scale with the case so that some
    boxes are refined and some derefined."""
    t = case.extra["export"]
    D = case.D
    L = t["r_min"][:D] + t["dr"][:D] * case.nc  # the level-1 box
    mid = [0.5 * L[d] for d in range(D)]
    seed = ([0.0] * (D - 1) + [0.6 * L[-1]], [0.0] * (D - 1) + [0.8 * L[-1]], 0.2 * L[-1]) if case.coord == 2 else \
        ([0.5 * L[0]] * (D - 1) + [0.6 * L[-1]], [0.5 * L[0]] * (D - 1) + [0.8 * L[-1]], 0.1 * L[-1])
    fine = t["dr"][D * (t["n_boxes"] - 1)]
    out = {
        "default": A.refine_settings(global_time=1.0),
        "adx_low": A.refine_settings(refine_adx=1e-3, global_time=1.0),
        "deref": A.refine_settings(refine_adx=10.0, derefine_dx=1.0, global_time=1.0),
        "buffer0": A.refine_settings(refine_adx=1e-3, buffer_width=0, derefine_dx=1.0, global_time=1.0),
        "buffer_max": A.refine_settings(refine_adx=1e-2, buffer_width=case.nc, derefine_dx=1.0, global_time=1.0),
        "alpha_eff": A.refine_settings(refine_adx=1e-3, use_alpha_effective=1, refine_adx_fac=1.3,
                                       derefine_dx=1.0, global_time=1.0),
        "min_dens": A.refine_settings(refine_adx=1e-3, refine_min_dens=1e16, global_time=1.0),
        "seeds": A.refine_settings(refine_adx=1e3, global_time=0.0, refine_init_time=1e-9, seeds=[seed],
                                   derefine_dx=1.0),
        "regions_limits": A.refine_settings(refine_adx=1e-3, global_time=0.5,
                                            regions=[(0.5 * fine, 1.0, [0.0] * D, mid)],
                                            limits=[(4 * fine, mid, list(L))]),
        "max_dx": A.refine_settings(refine_max_dx=2 * fine, global_time=1.0),
        "min_dx": A.refine_settings(refine_adx=1e-3, refine_min_dx=0.6 * fine, global_time=1.0),
    }
    return out


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8", "xyz3d_16"])
def test_refine_flags_equal_oracle(A, S, name):
    case = cases(S)[name]
    case.params.mg_num_vcycles = 20
    case.params.mg_max_rel_residual = 1e-10
    g = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    o = S.apply(A.Path("oracle", case.D, case.nc, case.coord), case)
    n = case.extra["export"]["n_boxes"]
    for it in range(3):
        if it == 0:
            for p in (g, o):
                p.field_compute(0, have_guess=False)
        else:
            assert g.step() == o.step()
        for p in (g, o):
            p.restrict_tree(A.CC_ELECTRON)  # as before refinement (src/streamer.f90:399-400)
        for iv in (A.CC_EFIELD, A.CC_ELECTRON):
            assert np.array_equal(g.get_cc(iv), o.get_cc(iv))
        counts = {}
        for key, s in settings(A, case).items():
            fg, fo = g.refine_flags(s, n), o.refine_flags(s, n)
            assert np.array_equal(fg, fo), (name, it, key, np.nonzero(fg != fo)[0][:10])
            counts[key] = (int((fo == REFINE).sum()), int((fo == DEREFINE).sum()))
        print(name, "state", it, "refine/derefine per setting", counts, flush=True)
        if it == 0:
            assert sum(c[0] for c in counts.values()) > 0 and sum(c[1] for c in counts.values()) > 0


def test_refine_flags_forest_equals_runs(A, S):
    """In a forest of runs every run's boxes get the flags that run gets
    alone (runs are separate trees: no flag crosses between them)."""
    ens = S.ensemble(level_list=[3, 4, 3], num_vcycles=20, rel_residual=1e-10)
    g = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL), ens)
    g.field_compute_runs(0, have_guess=False)
    s = A.refine_settings(refine_adx=1e-3, derefine_dx=1.0, global_time=1.0)
    fg = g.refine_flags(s, ens.export["n_boxes"])
    for r, case in enumerate(ens.cases):
        single = S.apply(A.Path("gpu", 2, 8, A.CYL), case)
        single.field_compute(0, have_guess=False)
        fs = single.refine_flags(s, case.extra["export"]["n_boxes"])
        o = ens.offsets[r]
        assert np.array_equal(fg[o:o + len(fs)], fs), r


def inner_leaves(t, D):
    """Leaves whose face neighbours all exist (refining them keeps 2:1)."""
    n, nch, nnb = t["n_boxes"], 2 ** D, 2 * D
    return [b + 1 for b in range(n) if t["children"][b * nch] == 0
            and all(v != 0 for v in t["neighbors"][b * nnb:(b + 1) * nnb])]


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8", "xyz3d_16"])
def test_restricted_and_prolonged_boxes_equal_oracle(A, S, name):
    """auto_restrict into boxes about to lose their children, then new
    boxes from refining leaves (the tree synchronised with its new boxes,
    whose data come from auto_prolong on each side), then a time step on
    the new tree: every cell-centred variable equal to the oracle's."""
    case = cases(S)[name]
    case.params.mg_num_vcycles = 20
    case.params.mg_max_rel_residual = 1e-10
    case.params.photoi_enabled = 1
    case.params.photoi_per_steps = 1
    g = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    o = S.apply(A.Path("oracle", case.D, case.nc, case.coord), case)
    for p in (g, o):
        p.field_compute(0, have_guess=False)
    assert g.step() == o.step()
    for p in (g, o):
        p.restrict_tree(A.CC_ELECTRON)
        p.gc_tree(A.CC_ELECTRON)

    def same(what):
        for iv in range(A.N_CC):
            a, b = g.get_cc(iv), o.get_cc(iv)
            assert np.array_equal(a, b), (name, what, iv, np.abs(a - b).max())

    t0 = case.extra["export"]
    D, nch = case.D, 2 ** case.D
    ch = t0["children"].reshape(t0["n_boxes"], nch)
    derefine = [b + 1 for b in range(t0["n_boxes"]) if ch[b, 0] > 0 and (ch[ch[b] - 1, 0] == 0).all()]
    assert derefine
    for p in (g, o):
        p.restrict_boxes(derefine)
    same("restrict")
    rng = np.random.default_rng(5)
    leaves = inner_leaves(t0, D)
    pick = sorted(rng.choice(leaves, size=max(1, len(leaves) // 4), replace=False))
    case.tree.refine(np.asarray(pick, np.int32))
    t1 = case.tree.export()
    new = np.arange(t0["n_boxes"] + 1, t1["n_boxes"] + 1, dtype=np.int32)
    print(name, "derefined", len(derefine), "refined", len(pick), "new boxes", len(new), flush=True)
    for p in (g, o):
        p.sync_tree(t1)
        p.prolong_boxes(new)
    same("prolong")
    assert g.step() == o.step()
    same("step on the new tree")


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8"])
def test_refinement_cycles_equal_oracle(A, S, name):
    """The reference's loop with refinement (src/streamer.f90:397-427) for
    three cycles: two steps; restrict and fill the electron ghosts; the
    flags (afgpu_refine_flags); the boxes flagged for refinement refined
    (the tree change stays with the caller; derefinement is not applied
    here); restriction and prolongation of the changed boxes on each side;
    the field on the new mesh and the photoionization source.  Every
    cell-centred variable stays equal to the oracle's bit for bit."""
    case = cases(S)[name]
    case.params.mg_num_vcycles = 20
    case.params.mg_max_rel_residual = 1e-10
    case.params.photoi_enabled = 1
    case.params.photoi_per_steps = 1
    g = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    o = S.apply(A.Path("oracle", case.D, case.nc, case.coord), case)
    for p in (g, o):
        p.field_compute(0, have_guess=False)
    t = case.extra["export"]
    s = A.refine_settings(refine_adx=0.01, derefine_dx=1.0, global_time=1.0, max_lvl=int(t["lvl"].max()) + 1)
    added = 0
    for cycle in range(3):
        for _ in range(2):
            assert g.step() == o.step(), (name, cycle)
        for p in (g, o):
            p.restrict_tree(A.CC_ELECTRON)
            p.gc_tree(A.CC_ELECTRON)
        n = t["n_boxes"]
        fg, fo = g.refine_flags(s, n), o.refine_flags(s, n)
        assert np.array_equal(fg, fo)
        ref = np.nonzero(fo == REFINE)[0] + 1
        print(name, "cycle", cycle, "boxes", n, "refine", len(ref), flush=True)
        if len(ref) == 0:
            continue
        case.tree.refine(ref.astype(np.int32))
        t = case.tree.export()
        new = np.arange(n + 1, t["n_boxes"] + 1, dtype=np.int32)
        added += len(new)
        for p in (g, o):
            p.sync_tree(t)
            p.prolong_boxes(new)
            p.field_compute(0, have_guess=True)
            p.photoi_set_src()
        for iv in range(A.N_CC):
            assert np.array_equal(g.get_cc(iv), o.get_cc(iv)), (name, cycle, iv)
    assert added > 0
