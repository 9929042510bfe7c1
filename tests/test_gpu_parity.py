"""Parity of the HIP path with the CPU oracle (FP64): the same seeded
inputs, the same call sequence, compared after every stage.

The FP64 device path evaluates every cell value with the oracle's operation
order, so the arrays are expected to agree bit for bit; the asserted bar is
1e-12 relative (the required parity), and bit-exactness is asserted where
no reduction order is involved at all.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_diff(a, b):
    scale = max(np.abs(a).max(), np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def make_pair(A, S, case):
    g = A.Path("gpu", case.D, case.nc, case.coord)
    c = A.Path("oracle", case.D, case.nc, case.coord)
    S.apply(g, case)
    S.apply(c, case)
    return g, c


def compare(A, g, c, ivs, exact=True, tol=1e-12, what=""):
    for iv in ivs:
        a, b = g.get_cc(iv), c.get_cc(iv)
        d = rel_diff(a, b)
        print("%s cc %d: rel diff %.3e" % (what, iv, d), flush=True)
        if exact:
            assert np.array_equal(a, b), "%s: cc %d differs (rel %.3e)" % (what, iv, d)
        assert d <= tol, "%s: cc %d rel diff %.3e" % (what, iv, d)


CASES = {
    "cyl2d_uniform": lambda S: S.config0(uniform_lvl=4),
    "cyl2d_amr": lambda S: S.small_amr(2, coord=2, uniform_lvl=2, max_lvl=4, seed=3),
    "xyz2d_amr": lambda S: S.small_amr(2, coord=1, uniform_lvl=2, max_lvl=4, seed=5),
    "xyz3d_amr": lambda S: S.small_amr(3, nc=8, uniform_lvl=2, max_lvl=3, frac=0.3, seed=11),
    "xyz3d_16": lambda S: S.small_amr(3, nc=16, uniform_lvl=2, max_lvl=3, frac=0.3, seed=7),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_multigrid_parity(A, S, name):
    case = CASES[name](S)
    print("case", name, case.tree.info(), flush=True)
    g, c = make_pair(A, S, case)
    rng = np.random.default_rng(42)
    rhs = rng.standard_normal((case.extra["export"]["n_boxes"], g.ncell)) * 1e10
    for p in (g, c):
        p.put_cc(A.CC_RHS, rhs)
        p.mg_fas_fmg(A.OP_POISSON, True, False)
    compare(A, g, c, [A.CC_PHI, A.CC_TMP], what=name + " fmg")
    for p in (g, c):
        p.mg_fas_vcycle(A.OP_POISSON, True, 0)
    compare(A, g, c, [A.CC_PHI, A.CC_TMP, A.CC_RHS], what=name + " vcycle")
    assert g.tree_maxabs_cc(A.CC_TMP) == c.tree_maxabs_cc(A.CC_TMP)


@pytest.mark.parametrize("name", sorted(CASES))
def test_field_and_steps_parity(A, S, name):
    case = CASES[name](S)
    g, c = make_pair(A, S, case)
    ng, rg = g.field_compute(0, have_guess=False)
    nc_, rc = c.field_compute(0, have_guess=False)
    assert (ng, rg) == (nc_, rc)
    compare(A, g, c, [A.CC_PHI, A.CC_EFIELD], what=name + " field")
    for it in range(6):
        ig = g.step()
        ic = c.step()
        print(it, ig, flush=True)
        assert ig == ic, (ig, ic)
        compare(A, g, c, [A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION, A.CC_PHI, A.CC_EFIELD],
                what="%s step %d" % (name, it))
    fg, fc_ = g.get_fc(A.FC_FLUX), c.get_fc(A.FC_FLUX)
    assert np.array_equal(fg, fc_)


@pytest.mark.parametrize("name", ["cyl2d_amr", "xyz3d_amr", "xyz3d_16"])
def test_photoi_parity(A, S, name):
    case = CASES[name](S)
    g, c = make_pair(A, S, case)
    for p in (g, c):
        p.field_compute(0, have_guess=False)
        p.photoi_set_src()
    compare(A, g, c, [A.CC_PHOTO, A.CC_HELMH, A.CC_HELMH + 1, A.CC_HELMH + 2], what=name + " photoi")


def test_graph_replays_match_eager(A, S):
    """The HIP-graph path (captured V-cycle / FMG / field sequences, replayed
    many times with host synchronisation in between) gives exactly the eager
    results, step after step."""
    case = CASES["xyz3d_amr"](S)
    case.params.photoi_enabled = 1
    case.params.photoi_per_steps = 2
    g1 = S.apply(A.Path("gpu", case.D, case.nc, case.coord), case)
    g2 = S.apply(A.Path("gpu", case.D, case.nc, case.coord), case)
    g2.set_option(A.OPT_GRAPHS, 0)
    assert g1.field_compute(0, False) == g2.field_compute(0, False)
    for it in range(6):
        a, b = g1.step(), g2.step()
        assert a == b
        for iv in (A.CC_ELECTRON, A.CC_PHI, A.CC_PHOTO, A.CC_TMP):
            assert np.array_equal(g1.get_cc(iv), g2.get_cc(iv)), (it, iv)


def test_heun_steps_parity(A, S):
    """Heun's method (the reference's default integrator, m_dt.f90:80): two
    forward-Euler stages with the field recomputed in between."""
    case = S.small_amr(3, nc=8, uniform_lvl=2, max_lvl=3, frac=0.3, seed=11, integrator=A.HEUNS_METHOD)
    g, c = make_pair(A, S, case)
    assert g.field_compute(0, False) == c.field_compute(0, False)
    for it in range(4):
        ig, ic = g.step(), c.step()
        assert ig == ic, (ig, ic)
        compare(A, g, c, [A.CC_ELECTRON, A.CC_POS_ION, A.CC_PHI, A.CC_EFIELD], what="heun step %d" % it)


def test_config0_full_size_parity(A, S):
    """The 2D axisymmetric workload at its full size (uniform 256 x 256 grid in
    1024 boxes of 8 x 8, 20 kV, Poisson to 1e-10 of max |rhs|): several steps
    bit for bit."""
    case = S.config0(uniform_lvl=6, num_vcycles=100, rel_residual=1e-10)
    g, c = make_pair(A, S, case)
    assert g.field_compute(0, False) == c.field_compute(0, False)
    for it in range(8):
        ig, ic = g.step(), c.step()
        print(it, ig["n_vcycles"], ig["dt_used"], flush=True)
        assert ig == ic, (ig, ic)
    compare(A, g, c, [A.CC_ELECTRON, A.CC_POS_ION, A.CC_PHI, A.CC_EFIELD], what="config0")


def test_config1_full_size_parity(A, S):
    """The 3D AMR workload at its full size (2121 boxes of 16^3, 7,602,176 leaf
    cells, 15 kV, Poisson to 1e-8): the initial solve and two steps bit for
    bit against the CPU oracle."""
    case = S.config1(num_vcycles=100, rel_residual=1e-8)
    g, c = make_pair(A, S, case)
    assert g.field_compute(0, False) == c.field_compute(0, False)
    compare(A, g, c, [A.CC_PHI, A.CC_EFIELD], what="config1 field")
    for it in range(2):
        ig, ic = g.step(), c.step()
        print(it, ig["n_vcycles"], ig["dt_used"], flush=True)
        assert ig == ic, (ig, ic)
    compare(A, g, c, [A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION, A.CC_PHI, A.CC_EFIELD], what="config1 steps")


@pytest.mark.parametrize("name", ["cyl2d_amr", "xyz3d_16"])
def test_rejected_steps_parity(A, S, name):
    """A restart with a dt far above the limits: the first step is rejected,
    the backup restored (copy_current_state / restore_previous_state and the
    field recomputed from the restored potential, src/streamer.f90:254-294,
    659-688) and the step redone with a smaller dt; safety factor 1 lets
    later steps reject too.  Step results and states bit for bit."""
    case = CASES[name](S)
    case.params.dt_safety_factor = 1.0
    case.params.mg_num_vcycles = 20
    case.params.mg_max_rel_residual = 1e-10
    g, c = make_pair(A, S, case)
    assert g.field_compute(0, False) == c.field_compute(0, False)
    for p in (g, c):
        p.set_time_state(0, 0.0, 5e-11)
    rejected = 0
    for it in range(5):
        ig, ic = g.step(), c.step()
        assert ig == ic, (it, ig, ic)
        rejected += ic["n_rejected"]
        compare(A, g, c, [A.CC_ELECTRON, A.CC_POS_ION, A.CC_PHI, A.CC_EFIELD], what="%s reject %d" % (name, it))
    print(name, "rejected attempts", rejected, flush=True)
    assert rejected > 0
