"""Work-decomposition shapes (boxes per workgroup, persistent grid) and the
autotuner that picks them: no shape changes any result, and tuning leaves
the simulation where it was."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("CC_ELECTRON", "CC_POS_ION", "CC_NEG_ION", "CC_PHI", "CC_EFIELD")


def snapshot(A, p, ids=None):
    return [p.get_cc(getattr(A, n), ids) for n in STATE]


@pytest.mark.parametrize("shape", [(1, 0), (2, 4), (4, 8), (8, 32), (1, 16)])
def test_every_shape_gives_the_same_steps(A, S, shape):
    """2D boxes of 8 x 8 (one wavefront each): every shape of every class,
    FP64 and mixed precision, against the default shapes, bit for bit."""
    for prec in (A.PREC_FP64, A.PREC_MIXED):
        ens = S.ensemble(level_list=[4, 5, 3], num_vcycles=20, rel_residual=1e-8)
        ref = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL, prec), ens)
        g = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL, prec), ens)
        g.set_shapes([shape] * A.N_KERNEL_CLASSES)
        assert g.get_shapes() == [shape] * A.N_KERNEL_CLASSES
        for p in (ref, g):
            p.field_compute_runs(0, have_guess=False)
        for it in range(3):
            assert ref.step_ensemble() == g.step_ensemble()
        for a, b in zip(snapshot(A, ref), snapshot(A, g)):
            assert np.array_equal(a, b)


def test_autotune_leaves_the_simulation_unchanged(A, S):
    ens = S.ensemble(level_list=[5, 4, 6], num_vcycles=20, rel_residual=1e-8)
    ref = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL, A.PREC_MIXED), ens)
    g = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL, A.PREC_MIXED), ens)
    for p in (ref, g):
        p.field_compute_runs(0, have_guess=False)
        p.step_ensemble()
    tuned = g.autotune()
    print(tuned, flush=True)
    assert all(bb in (1, 2, 4, 8) and 0 <= pc <= 64 for bb, pc, _ in tuned.values())
    assert [(bb, pc) for bb, pc, _ in tuned.values()] == g.get_shapes()
    for r in range(len(ens.cases)):
        assert g.get_run_time_state(r) == ref.get_run_time_state(r)
    for it in range(3):
        assert ref.step_ensemble() == g.step_ensemble()
    for a, b in zip(snapshot(A, ref), snapshot(A, g)):
        assert np.array_equal(a, b)


def test_autotune_single_3d_run(A, S):
    """Boxes of 16^3 (one workgroup each): only the persistent grid varies."""
    case = S.small_amr(3, nc=16, uniform_lvl=2, max_lvl=3, frac=0.3, seed=7, photoi=True)
    ref = S.apply(A.Path("gpu", 3, 16, A.XYZ, A.PREC_MIXED), case)
    g = S.apply(A.Path("gpu", 3, 16, A.XYZ, A.PREC_MIXED), case)
    for p in (ref, g):
        p.field_compute(0, have_guess=False)
    tuned = g.autotune()
    print(tuned, flush=True)
    assert all(bb == 1 for bb, _, _ in tuned.values())
    for it in range(2):
        assert ref.step() == g.step()
    for a, b in zip(snapshot(A, ref), snapshot(A, g)):
        assert np.array_equal(a, b)
    with pytest.raises(A.AfError):
        g.set_shapes([(4, 8)] * A.N_KERNEL_CLASSES)  # several boxes per workgroup: 8 x 8 boxes only


def test_non_default_shapes_equal_oracle(A, S):
    """A non-default work decomposition against the oracle itself: a forest
    with several 8 x 8 boxes per workgroup on a persistent grid, and a 3D
    16^3 run on a persistent grid, step by step bit for bit."""
    from test_gpu_ensemble import forest_and_singles, run_and_compare
    ens = S.ensemble(level_list=[4, 5, 3], num_vcycles=20, rel_residual=1e-10)
    g, singles = forest_and_singles(A, S, ens, A.PREC_FP64, oracle=True)
    g.set_shapes([(2, 16)] * A.N_KERNEL_CLASSES)
    run_and_compare(A, ens, g, singles, 3, "shape (2, 16) vs oracle")
    case = S.small_amr(3, nc=16, uniform_lvl=2, max_lvl=3, frac=0.3, seed=7)
    g = S.apply(A.Path("gpu", 3, 16, A.XYZ, A.PREC_FP64), case)
    c = S.apply(A.Path("oracle", 3, 16, A.XYZ), case)
    g.set_shapes([(1, 2)] * A.N_KERNEL_CLASSES)
    assert g.field_compute(0, False) == c.field_compute(0, False)
    for it in range(3):
        assert g.step() == c.step(), it
    for a, b in zip(snapshot(A, g), snapshot(A, c)):
        assert np.array_equal(a, b)
