"""Diagnostics from the device pool against the oracle.

- afgpu_regression_values: the regression log's statistics after time
  steps on 2D cylindrical and 3D AMR trees.  The maxima equal the oracle's
  exactly.  The sums add the same per-cell terms in another order (per box,
  then per run in a fixed order; the oracle adds them one by one as the
  reference does with one thread, and the reference's own order changes
  with its thread count), so they are compared to 1e-12 relative: every
  term is non-negative, and the rounding of any order is bounded by the
  number of terms times the unit roundoff times the sum.
- A forest of runs gives every run the values the run gets alone.
- afgpu_write_regression_log writes the reference's header and lines.
- afgpu_get_boxes / afgpu_put_boxes move whole boxes in af_write_tree's
  order, bit for bit.
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

SPECIES = ("e", "M_plus", "M_min")


def cases(S):
    return {
        "cyl2d": S.small_amr(2, coord=2, uniform_lvl=3, max_lvl=5, frac=0.3, seed=3),
        "xyz3d_8": S.small_amr(3, nc=8, uniform_lvl=2, max_lvl=4, frac=0.3, seed=11),
        "xyz3d_16": S.small_amr(3, nc=16, uniform_lvl=2, max_lvl=3, frac=0.3, seed=7),
    }


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_8", "xyz3d_16"])
def test_regression_values_equal_oracle(A, S, name):
    case = cases(S)[name]
    case.params.mg_num_vcycles = 20
    case.params.mg_max_rel_residual = 1e-10
    g = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    o = S.apply(A.Path("oracle", case.D, case.nc, case.coord), case)
    ivs = [A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION, A.CC_EFIELD]
    for p in (g, o):
        p.field_compute(0, have_guess=False)
    for it in range(3):
        vg, vo = g.regression_values(ivs), o.regression_values(ivs)
        assert vg.shape == vo.shape == (1, len(ivs), 3)
        assert np.array_equal(vg[..., 2], vo[..., 2]), (name, it)
        assert np.allclose(vg[..., :2], vo[..., :2], rtol=1e-12, atol=0), (name, it, vg - vo)
        # deterministic: a second call gives the same bits
        assert np.array_equal(g.regression_values(ivs), vg)
        assert g.step() == o.step()


def test_regression_values_forest_equal_runs(A, S):
    ens = S.ensemble(level_list=[3, 4, 3], num_vcycles=20, rel_residual=1e-10)
    g = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL), ens)
    g.field_compute_runs(0, have_guess=False)
    ivs = [A.CC_ELECTRON, A.CC_EFIELD]
    vf = g.regression_values(ivs)
    assert vf.shape == (len(ens.cases), len(ivs), 3)
    for r, case in enumerate(ens.cases):
        single = S.apply(A.Path("gpu", 2, 8, A.CYL), case)
        single.field_compute(0, have_guess=False)
        vs = single.regression_values(ivs)[0]
        assert np.array_equal(vf[r, :, 2], vs[:, 2]), r
        assert np.allclose(vf[r, :, :2], vs[:, :2], rtol=1e-13, atol=0), r


def test_write_regression_log(A, S, tmp_path):
    case = cases(S)["cyl2d"]
    g = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    g.field_compute(0, have_guess=False)
    ivs = [A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION]
    fn = str(tmp_path / "run_rtest.log")
    rows = []
    for cnt in range(3):
        g.write_regression_log(fn, cnt, ivs, SPECIES)
        it, t, dt = g.get_time_state()
        v = g.regression_values(ivs)[0]
        rows.append([t, dt] + list(v[:, 0]) + list(v[:, 1]) + list(v[:, 2]))
        g.step()
    lines = open(fn).read().splitlines()
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "regression_logs.json")))
    assert lines[0] == ref["programs/standard_2d/tests/test_cyl_rtest.log"]["header"]
    assert len(lines) == 4
    for cnt, (line, row) in enumerate(zip(lines[1:], rows)):
        assert line == A.format_regression_line(cnt, row)


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d_16"])
def test_box_io_equals_transfers(A, S, name):
    case = cases(S)[name]
    g = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    o = S.apply(A.Path("oracle", case.D, case.nc, case.coord), case)
    n = case.extra["export"]["n_boxes"]
    ivs = [A.CC_ELECTRON, A.CC_PHI, A.CC_POS_ION, A.CC_EFIELD]
    rng = np.random.default_rng(4)
    data = rng.normal(size=(n, len(ivs), g.ncell))
    ids = np.arange(1, n + 1, dtype=np.int32)
    rng.shuffle(ids)
    for p in (g, o):
        p.put_boxes(ivs, data[ids - 1], ids)
    for k, iv in enumerate(ivs):
        assert np.array_equal(g.get_cc(iv), data[:, k])
        assert np.array_equal(o.get_cc(iv), data[:, k])
    sub = ids[: max(1, n // 3)]
    assert np.array_equal(g.get_boxes(ivs, sub), o.get_boxes(ivs, sub))
    assert np.array_equal(g.get_boxes(ivs, sub), data[sub - 1])
