"""The mixed-precision form (FP32 multigrid cycles inside FP64 defect
correction) against the plain FP64 form, which equals the CPU oracle bit for
bit (test_gpu_parity.py).

The two forms are different iterations with the same fixed point: the
composite leaf equations with the refined coarse cells equal to the
restriction of the finer solution.  Both are compared on leaf-box interiors
(the values the physics reads); refined coarse cells hold FAS auxiliary
values in the plain form.

* fixed point: both forms iterated to their round-off floor agree to
  ~1e-14 (asserted 2e-13);
* time steps at the production threshold (leaf residual 1e-8 of max |rhs|,
  or the absolute floor of m_field.f90:500-504 where that is larger): the
  potential, field and densities after several steps agree to the
  accuracy the threshold leaves (asserted at about ten times what is
  achieved: 2e-10 for the potential, 5e-9 for the field);
* photoionization to a 1e-8 relative residual: the photoionization source
  agrees to 1e-9.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


CASES = {
    "cyl2d_amr": lambda S: S.small_amr(2, coord=2, uniform_lvl=2, max_lvl=4, seed=3),
    "xyz3d_amr": lambda S: S.small_amr(3, nc=8, uniform_lvl=2, max_lvl=3, frac=0.3, seed=11),
    "xyz3d_16": lambda S: S.small_amr(3, nc=16, uniform_lvl=2, max_lvl=3, frac=0.3, seed=7),
}


def leaf_interior(case, p, iv):
    ex = case.extra["export"]
    D, nc = case.D, case.nc
    ch = ex["children"].reshape(ex["n_boxes"], 2 ** D)
    leaves = np.nonzero(ch[:, 0] == 0)[0] + 1
    a = p.get_cc(iv, ids=leaves).reshape((len(leaves),) + (nc + 2,) * D)
    sl = (slice(None),) + (slice(1, nc + 1),) * D
    return a[sl]


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def setup(case, photoi=False):
    case.params.mg_num_vcycles = 100
    case.params.mg_max_rel_residual = 1e-8
    if photoi:
        case.params.photoi_enabled = 1
        case.params.photoi_per_steps = 1
        case.params.photoi_max_rel_residual = 1e-8
    return case


def pair(A, S, case):
    f = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    m = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
    return f, m


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_poisson_fixed_point(A, S, name):
    case = setup(CASES[name](S))
    f, m = pair(A, S, case)
    nf, rf = f.field_compute(0, False)
    nm, rm = m.field_compute(0, False)
    print(name, "initial solve fp64", nf, rf, "mixed", nm, rm, flush=True)
    assert nm <= nf + 2
    for _ in range(10):
        f.mg_fas_vcycle(A.OP_POISSON, True, 0)
        m.field_compute(0, True)  # at least one more mixed cycle each call
    e = rel_max(leaf_interior(case, m, A.CC_PHI), leaf_interior(case, f, A.CC_PHI))
    print(name, "fixed point phi %.2e" % e, flush=True)
    # (achieved on MI355X: at most 1.5e-14)
    assert e < 2e-13


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_steps_match_fp64(A, S, name):
    case = setup(CASES[name](S))
    f, m = pair(A, S, case)
    f.field_compute(0, False)
    m.field_compute(0, False)
    for it in range(5):
        a, b = f.step(), m.step()
        print(name, it, "fp64", a["n_vcycles"], "%.2e" % a["rel_residual"],
              "mixed", b["n_vcycles"], "%.2e" % b["rel_residual"], flush=True)
        assert b["n_vcycles"] <= a["n_vcycles"] + 2
        # both stop on the same absolute threshold
        assert b["rel_residual"] < 10 * max(1e-8, a["rel_residual"])
    errs = {}
    for key, iv in (("phi", A.CC_PHI), ("field", A.CC_EFIELD), ("electron", A.CC_ELECTRON),
                    ("pos_ion", A.CC_POS_ION)):
        errs[key] = rel_max(leaf_interior(case, m, iv), leaf_interior(case, f, iv))
    print(name, "errors", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    # about ten times the largest achieved on MI355X over the three cases
    # (phi 1.6e-11, field 4.5e-10, electrons 5.8e-11, ions 5.0e-15): each
    # form stops on the same absolute threshold, so the difference is the
    # stopping tolerance's, not a precision floor
    assert errs["phi"] < 2e-10
    assert errs["field"] < 5e-9
    assert errs["electron"] < 6e-10
    assert errs["pos_ion"] < 1e-13


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_photoi_matches_fp64(A, S, name):
    case = setup(CASES[name](S), photoi=True)
    f, m = pair(A, S, case)
    f.field_compute(0, have_guess=False)
    m.put_cc(A.CC_PHI, f.get_cc(A.CC_PHI))
    m.field_from_potential()
    for rep in range(2):  # the second solve starts from the first one's modes
        for p in (f, m):
            p.photoi_set_src()
        e = rel_max(leaf_interior(case, m, A.CC_PHOTO), leaf_interior(case, f, A.CC_PHOTO))
        print(name, rep, "photo error %.2e" % e, flush=True)
        # (achieved on MI355X: at most 8.3e-11)
        assert e < 1e-9


@pytest.mark.parametrize("direct_n", [0, 16, 32, 64])
def test_mixed_inner_coarse_level(A, S, direct_n):
    """The inner cycle's coarsest level: the level-1 box (direct_n = 0) or
    the exact solve on the finest uniformly refined level with at most
    direct_n cells per axis.  All reach the FP64 fixed point."""
    case = setup(S.small_amr(2, coord=2, uniform_lvl=4, max_lvl=5, frac=0.3, seed=5))
    f, m = pair(A, S, case)
    m.set_option(A.OPT_DIRECT_N, direct_n)
    f.field_compute(0, False)
    n0, r0 = m.field_compute(0, False)
    for _ in range(8):
        f.mg_fas_vcycle(A.OP_POISSON, True, 0)
        m.field_compute(0, True)
    e = rel_max(leaf_interior(case, m, A.CC_PHI), leaf_interior(case, f, A.CC_PHI))
    print("direct_n", direct_n, "initial cycles", n0, "fixed point phi %.2e" % e, flush=True)
    assert e < 1e-11


def test_mixed_uniform_tree_is_one_direct_solve(A, S):
    """A uniformly refined tree is solved by the exact level solve alone:
    one correction cycle reaches the threshold."""
    case = setup(S.config0(uniform_lvl=4))  # 2D cylindrical, 64 x 64 cells
    f, m = pair(A, S, case)
    f.field_compute(0, False)
    m.field_compute(0, False)
    for it in range(3):
        a, b = f.step(), m.step()
        print(it, a["n_vcycles"], b["n_vcycles"], b["rel_residual"], flush=True)
        assert b["n_vcycles"] <= 2
    e = rel_max(leaf_interior(case, m, A.CC_ELECTRON), leaf_interior(case, f, A.CC_ELECTRON))
    assert e < 1e-9


BLOCK_CASES = {
    "cyl2d_amr": CASES["cyl2d_amr"],
    # level 4 has 512 boxes: one workgroup per box, so the block smoother runs there
    "xyz3d_l4": lambda S: S.small_amr(3, nc=8, uniform_lvl=4, max_lvl=5, frac=0.15, seed=21),
}


@pytest.mark.parametrize("name", sorted(BLOCK_CASES))
def test_mixed_block_smoother(A, S, name):
    """The inner cycle with the block smoother reaches the same fixed point,
    keeps the step results within the threshold's accuracy, and is
    deterministic (graph replays equal eager runs)."""
    case = setup(BLOCK_CASES[name](S))
    f, m = pair(A, S, case)
    m.set_option(A.OPT_BLOCK_SMOOTH, 1)
    m2 = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
    m2.set_option(A.OPT_BLOCK_SMOOTH, 1)
    m2.set_option(A.OPT_GRAPHS, 0)
    f.field_compute(0, False)
    assert m.field_compute(0, False) == m2.field_compute(0, False)
    for it in range(3):
        a, b, b2 = f.step(), m.step(), m2.step()
        print(name, it, "fp64", a["n_vcycles"], "block", b["n_vcycles"], "%.2e" % b["rel_residual"], flush=True)
        assert b == b2
        assert b["n_vcycles"] <= a["n_vcycles"] + 3
    assert np.array_equal(m.get_cc(A.CC_PHI), m2.get_cc(A.CC_PHI))
    e = rel_max(leaf_interior(case, m, A.CC_ELECTRON), leaf_interior(case, f, A.CC_ELECTRON))
    print(name, "electron error %.2e" % e, flush=True)
    assert e < 1e-6
    for _ in range(10):
        f.mg_fas_vcycle(A.OP_POISSON, True, 0)
        m.field_compute(0, True)
    e = rel_max(leaf_interior(case, m, A.CC_PHI), leaf_interior(case, f, A.CC_PHI))
    print(name, "fixed point phi %.2e" % e, flush=True)
    assert e < 1e-10


@pytest.mark.parametrize("name", ["cyl2d_amr", "xyz3d_16"])
def test_mixed_batched_modes_equal_sequential(A, S, name):
    """The photoionization modes solved side by side (parallel branches of
    one captured graph) give bit for bit the results of solving them one
    after the other, eagerly."""
    case = setup(CASES[name](S), photoi=True)
    case.params.photoi_max_rel_residual = 1e-4
    m1 = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
    m2 = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
    m2.set_option(A.OPT_BATCH_MODES, 0)
    m2.set_option(A.OPT_GRAPHS, 0)
    for p in (m1, m2):
        p.field_compute(0, have_guess=False)
    for rep in range(3):
        for p in (m1, m2):
            p.photoi_set_src()
        for iv in [A.CC_PHOTO] + [A.CC_HELMH + j for j in range(3)]:
            assert np.array_equal(m1.get_cc(iv), m2.get_cc(iv)), (rep, iv)


def test_mixed_direct_solve_128_matrix_cores(A, S):
    """3D, 16^3 boxes, levels 1-4 uniform (level 4 = 128^3 cells, solved
    exactly by the inner cycle): the FP32 and FP64 matrix-core transforms
    (the FP64 ones split by parity, the default, and unsplit) and the FP64
    vector transforms all reach the FP64 fixed point."""
    case = setup(S.small_amr(3, nc=16, uniform_lvl=4, max_lvl=5, frac=0.03, seed=4))
    f = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    f.field_compute(0, False)
    for _ in range(6):
        f.mg_fas_vcycle(A.OP_POISSON, True, 0)
    ref = leaf_interior(case, f, A.CC_PHI)
    for mfma, f32, fold in ((1, 1, 1), (1, 0, 1), (1, 0, 0), (0, 0, 0)):
        m = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
        m.set_option(A.OPT_DIRECT_MFMA, mfma)
        m.set_option(A.OPT_DIRECT_F32, f32)
        m.set_option(A.OPT_DIRECT_FOLD, fold)
        n0, r0 = m.field_compute(0, False)
        for _ in range(4):
            m.field_compute(0, True)
        e = rel_max(leaf_interior(case, m, A.CC_PHI), ref)
        print("mfma", mfma, "f32", f32, "fold", fold, "initial cycles", n0, "phi vs fp64 fixed point %.2e" % e,
              flush=True)
        assert e < 1e-11


def test_mixed_direct_solve_fused_passes_bit_for_bit(A, S):
    """The uniform-level solve with the gather and scatter folded into the
    x transforms and the last axis's forward transform, division and
    backward transform as one pass gives the same bits as the separate
    passes: 3D with photoionization (level 4 = 128^3), a few full steps.
    (The fused pass exists in FP64 only: both runs transform in FP64.)"""
    case = setup(S.small_amr(3, nc=16, uniform_lvl=4, max_lvl=5, frac=0.03, seed=4), photoi=True)
    runs = []
    for fuse in (1, 0):
        m = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
        m.set_option(A.OPT_DIRECT_F32, 0)
        m.set_option(A.OPT_DIRECT_FOLD, 0)  # (the fused pass transforms unsplit)
        m.set_option(A.OPT_FUSE_DIRECT, fuse)
        infos = [m.field_compute(0, False)] + [m.step() for _ in range(3)]
        runs.append((infos, [m.get_cc(iv) for iv in (A.CC_PHI, A.CC_ELECTRON, A.CC_PHOTO)]))
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)


def test_split_half_sweeps_bit_for_bit(A, S):
    """The inner cycle's half-sweeps staging only the other colour's block
    and the rhs (k_gsrb_split, the default for 3D 16^3 boxes) give the same
    bits as the half-sweeps staging the whole box in LDS (k_gsrb_lds).
    They run on levels of at least 256 boxes: the channel tree's level 5
    (1536 boxes, refinement boundaries and the domain's faces), and a
    uniform level 4 (512 boxes) smoothed above a 64^3 direct solve; Poisson
    and the photoionization modes (Helmholtz) over a few steps."""
    cases = ((setup(S.config1(photoi=True), photoi=True), {}),
             (setup(S.small_amr(3, nc=16, uniform_lvl=4, max_lvl=5, frac=0.05, seed=6), photoi=True),
              {A.OPT_DIRECT_N: 64}))
    for case, opts in cases:
        runs = []
        for box_lds in (0, 1):
            print(case.name, "box_lds", box_lds, flush=True)
            m = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
            for k, v in opts.items():
                m.set_option(k, v)
            m.set_option(A.OPT_BOX_LDS, box_lds)
            infos = [m.field_compute(0, False)] + [m.step() for _ in range(2)]
            runs.append((infos, [m.get_cc(iv) for iv in (A.CC_PHI, A.CC_EFIELD, A.CC_ELECTRON, A.CC_PHOTO)]))
            m.close()
        assert runs[0][0] == runs[1][0], case.name
        for a, b in zip(runs[0][1], runs[1][1]):
            assert np.array_equal(a, b), case.name


def test_split2d_half_sweeps_bit_for_bit(A, S):
    """The 2D half-sweep of 8^2 boxes that stages everything by DMA
    (k_gsrb_split2d, the default on levels of at least 256 boxes) gives the
    same bits as the generic box-staging half-sweep (k_gsrb_lds): a
    cylindrical and a Cartesian tree refined past a uniform level 5 (256
    boxes), so that refinement boundaries, the axis and the domain's faces
    all border levels of several hundred boxes; Poisson and the
    photoionization modes over a few steps; and a forest of runs of
    different depths (runs masked out as they converge)."""
    cases = [setup(S.small_amr(2, coord=A.CYL, uniform_lvl=5, max_lvl=7, frac=0.3, seed=17), photoi=True),
             setup(S.small_amr(2, coord=A.XYZ, uniform_lvl=5, max_lvl=7, frac=0.3, seed=19), photoi=True)]
    for case in cases:
        runs = []
        for box_lds in (0, 1):
            m = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
            m.set_option(A.OPT_BOX_LDS, box_lds)
            infos = [m.field_compute(0, False)] + [m.step() for _ in range(2)]
            runs.append((infos, [m.get_cc(iv) for iv in (A.CC_PHI, A.CC_EFIELD, A.CC_ELECTRON, A.CC_PHOTO)]))
            m.close()
        print(case.name, "cycles", [i if isinstance(i, tuple) else i["n_vcycles"] for i in runs[0][0]], flush=True)
        assert runs[0][0] == runs[1][0], case.name
        for a, b in zip(runs[0][1], runs[1][1]):
            assert np.array_equal(a, b), case.name
    ens = S.ensemble(level_list=[5, 6, 5, 6], num_vcycles=20, rel_residual=1e-8)
    out = []
    for box_lds in (0, 1):
        g = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL, A.PREC_MIXED), ens)
        g.set_option(A.OPT_BOX_LDS, box_lds)
        g.field_compute_runs(0, have_guess=False)
        infos = [g.step_ensemble() for _ in range(2)]
        out.append((infos, [g.get_cc(iv) for iv in (A.CC_PHI, A.CC_ELECTRON)]))
        g.close()
    assert out[0][0] == out[1][0]
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)


def test_mixed_direct_solve_parity_split(A, S):
    """The FP64 transforms split by parity (every axis of a 3D Cartesian
    level is symmetric under reversal) against the unsplit ones: 3D with
    photoionization (level 4 = 128^3), a few steps.  The split transform
    is the same solve in another summation order, so the runs agree to
    round-off: equal V-cycle counts (or one more) and fields within a few
    1e-13 of each other."""
    case = setup(S.small_amr(3, nc=16, uniform_lvl=4, max_lvl=5, frac=0.03, seed=4), photoi=True)
    runs = []
    for fold in (1, 0):
        m = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
        m.set_option(A.OPT_DIRECT_F32, 0)
        m.set_option(A.OPT_DIRECT_FOLD, fold)
        n0, _ = m.field_compute(0, False)
        infos = [m.step() for _ in range(3)]
        runs.append((m, [n0] + [i["n_vcycles"] for i in infos]))
    print("V-cycles split", runs[0][1], "unsplit", runs[1][1], flush=True)
    for a, b in zip(runs[0][1], runs[1][1]):
        assert abs(a - b) <= 1
    for iv, tol in ((A.CC_PHI, 1e-10), (A.CC_EFIELD, 1e-8), (A.CC_ELECTRON, 1e-9), (A.CC_PHOTO, 1e-6)):
        e = rel_max(leaf_interior(case, runs[0][0], iv), leaf_interior(case, runs[1][0], iv))
        print("var", iv, "split vs unsplit %.2e" % e, flush=True)
        assert e < tol


def test_mixed_direct_solve_f32_transforms(A, S):
    """The inner cycle's uniform-level solve in FP32 (the default) against
    the same solve in FP64: 3D with photoionization (level 4 = 128^3), a few
    steps.  The exact solve is one step of the FP32 correction cycle, so its
    rounding changes the iterates but not what the FP64 outer correction
    converges to: equal V-cycle counts (or one more), and potential, field
    and densities within the threshold's accuracy of each other."""
    case = setup(S.small_amr(3, nc=16, uniform_lvl=4, max_lvl=5, frac=0.03, seed=4), photoi=True)
    runs = []
    for f32 in (1, 0):
        m = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
        m.set_option(A.OPT_DIRECT_F32, f32)
        n0, _ = m.field_compute(0, False)
        infos = [m.step() for _ in range(4)]
        runs.append((m, [n0] + [i["n_vcycles"] for i in infos]))
    print("V-cycles f32", runs[0][1], "f64", runs[1][1], flush=True)
    for a, b in zip(runs[0][1], runs[1][1]):
        assert abs(a - b) <= 1
    for iv, tol in ((A.CC_PHI, 1e-7), (A.CC_EFIELD, 1e-5), (A.CC_ELECTRON, 1e-6), (A.CC_PHOTO, 1e-6)):
        e = rel_max(leaf_interior(case, runs[0][0], iv), leaf_interior(case, runs[1][0], iv))
        print("var", iv, "f32 vs f64 transforms %.2e" % e, flush=True)
        assert e < tol


def test_mixed_graph_replays_match_eager(A, S):
    case = setup(CASES["xyz3d_amr"](S), photoi=True)
    case.params.photoi_max_rel_residual = 1e-3
    g1 = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
    g2 = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED), case)
    g2.set_option(A.OPT_GRAPHS, 0)
    assert g1.field_compute(0, False) == g2.field_compute(0, False)
    assert np.array_equal(g1.get_cc(A.CC_PHI), g2.get_cc(A.CC_PHI))
    for it in range(5):
        a, b = g1.step(), g2.step()
        assert a == b
        for iv in (A.CC_ELECTRON, A.CC_PHI, A.CC_PHOTO):
            assert np.array_equal(g1.get_cc(iv), g2.get_cc(iv)), (it, iv)
