"""The reference's own regression logs replayed end to end on the device
(tests/rtest.py; the tolerances and their reasons are those of
tests/test_rtest_replay.py): the plain FP64 form against the reference's
logs and against the oracle's replay of the same run (times, time steps
and maxima equal, sums to 1e-12: they add the same terms in another
order; with a reaction list to 1e-9), and the mixed-precision form
against the reference's axisymmetric log."""
import numpy as np
import pytest

import rtest as R
from test_rtest_replay import CPU_CASES, check_rows, rel

pytestmark = pytest.mark.gpu

SUM_COLS = [3, 4, 5, 6, 7, 8]


@pytest.mark.parametrize("name", ["test_cyl", "test_2d", "test_3d", "test_2d_photoi"])
def test_device_replays_reference_log(A, S, name):
    print("replay", name, flush=True)
    g = R.Run(name, "gpu")
    rows = g.run()
    check_rows(name, rows)
    o = R.Run(name, "oracle").run()
    exact = [c for c in range(rows.shape[1]) if c not in SUM_COLS]
    assert np.array_equal(rows[:, exact], o[:, exact]), name
    assert rel(rows[:, SUM_COLS], o[:, SUM_COLS]).max() <= 1e-12, name
    assert g.steps > 50


@pytest.mark.parametrize("name", ["test_cyl_chem", "test_cyl_photoi_chem", "test_3d_chem", "test_3d_photoi_chem",
                                  "test_2d_photoi_chem"])
def test_device_replays_reference_chemistry_log(A, S, name):
    """The reaction-list rtests on the device: against the reference's log
    (the bounds of tests/test_rtest_replay.py) and, where the CPU suite runs
    the oracle, against the oracle's replay: the device's and the host's
    exp-type rates may differ in the last bit, so the rows agree to 1e-9."""
    print("replay", name, flush=True)
    g = R.Run(name, "gpu")
    rows = g.run()
    check_rows(name, rows)
    assert g.steps > 50
    if name in CPU_CASES:
        o = R.Run(name, "oracle").run()
        assert rel(rows, o).max() <= 1e-9, name


def test_mixed_form_replays_cylindrical_log(A, S):
    rows = R.Run("test_cyl", "gpu", A.PREC_MIXED).run()
    ref = R.reference_rows("test_cyl")
    err = rel(rows, ref)
    print("mixed test_cyl max relative error per column", np.array2string(err.max(axis=0), precision=2), flush=True)
    assert rows.shape == ref.shape
    assert np.isclose(rows, ref, rtol=1e-4, atol=1e-8).all()
