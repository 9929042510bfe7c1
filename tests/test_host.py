"""Host-side stand-in: tree invariants, initial conditions, lookup tables."""
import numpy as np


def test_tree_invariants(A, S):
    case = S.small_amr(3, nc=8, uniform_lvl=1, max_lvl=4, frac=0.4, seed=21)
    t = case.extra["export"]
    D, n = 3, t["n_boxes"]
    rev = [1, 0, 3, 2, 5, 4]
    for b in range(n):
        nb = t["neighbors"][b * 6:(b + 1) * 6]
        ch = t["children"][b * 8:(b + 1) * 8]
        if ch[0] > 0:
            assert all(c > 0 and t["parent"][c - 1] == b + 1 for c in ch)
        for q in range(6):
            if nb[q] > 0:
                assert t["neighbors"][(nb[q] - 1) * 6 + rev[q]] == b + 1
                assert t["lvl"][nb[q] - 1] == t["lvl"][b]
            elif nb[q] == 0:  # refinement boundary: parent's neighbour exists (2:1)
                par = t["parent"][b] - 1
                assert t["neighbors"][par * 6 + q] > 0
    assert case.tree.info()["highest_lvl"] == 4


import pytest


@pytest.mark.parametrize("D", [2, 3])
def test_neighbour_matrix_layout(A, S, D):
    """neighbor_mat(-1:1, -1:1[, -1:1]) is exported with 3^D entries per box
    in column-major order, (di+1) + 3(dj+1) [+ 9(dk+1)]: the centre is the
    box itself, every same-level box whose index differs by (di, dj[, dk])
    appears at that position, and the face entries agree with the
    neighbour list."""
    case = S.small_amr(D, nc=8, uniform_lvl=2, max_lvl=4, frac=0.4, seed=13)
    t = case.extra["export"]
    n, nm, nnb = t["n_boxes"], 3 ** D, 2 * D
    where = {(int(t["lvl"][b]), tuple(int(v) for v in t["ix"][b * D:(b + 1) * D])): b + 1 for b in range(n)}
    dirs = [(di, dj, dk) for dk in ((-1, 0, 1) if D == 3 else (0,)) for dj in (-1, 0, 1) for di in (-1, 0, 1)]
    n_diag = 0
    for b in range(n):
        mat = t["nbmat"][b * nm:(b + 1) * nm]
        lvl, ix = int(t["lvl"][b]), [int(v) for v in t["ix"][b * D:(b + 1) * D]]
        for d in dirs:
            q = (d[0] + 1) + 3 * (d[1] + 1) + (9 * (d[2] + 1) if D == 3 else 0)
            if d == (0, 0, 0):
                assert mat[q] == b + 1
                continue
            other = where.get((lvl, tuple(ix[k] + d[k] for k in range(D))))
            if other is not None:
                assert mat[q] == other, (b + 1, d, mat[q], other)
                n_diag += sum(1 for k in range(D) if d[k] != 0) > 1
            else:
                assert mat[q] <= 0, (b + 1, d)
            if sum(1 for k in range(D) if d[k] != 0) == 1:  # a face
                k = [k for k in range(D) if d[k] != 0][0]
                assert mat[q] == t["neighbors"][b * nnb + 2 * k + (d[k] > 0)], (b + 1, d)
    assert n_diag > 0


def test_initial_conditions_independent_of_box_order(A, S):
    case = S.small_amr(2, nc=8, uniform_lvl=2, max_lvl=3, frac=0.5, seed=2)
    ids = np.arange(1, case.extra["export"]["n_boxes"] + 1, dtype=np.int32)
    a = case.tree.init_cond(ids, 1e13, 0.1, 7, 1e20, 4e-4, [0.0, 3e-3])
    b = case.tree.init_cond(ids[::-1].copy(), 1e13, 0.1, 7, 1e20, 4e-4, [0.0, 3e-3])[::-1]
    assert np.array_equal(a, b)
    bg = a.min()
    assert 0.9e13 <= bg and a.max() <= 1.1e13 + 1e20


def test_tables_linear_interpolation_of_input(A, S):
    N = S.gas_number_density()
    E, mu, dif, alpha, eta = S.analytic_transport()
    tb = A.transport_tables(E, mu, dif, alpha, eta, N)
    n = tb["n"]
    x = tb["td_x_min"] + np.arange(n) / tb["td_inv_fac"]
    xx = E * 1e21 / N
    for c, y in enumerate((mu * N, dif * N, alpha / N, eta / N)):
        col = tb["td"][c * n:(c + 1) * n]
        ref = np.interp(x, xx, y)
        assert np.abs(col - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300)
    ion = tb["td"][2 * n:3 * n] * tb["td"][:n] * x * 1e-21 * N
    assert np.abs(tb["chem"][:n] - ion).max() <= 1e-12 * ion.max()


def test_forest_merge_offsets_every_reference(A, S):
    """An ensemble's forest is its runs' trees one after the other: box ids,
    and every positive tree reference, offset by the boxes before the run;
    boundary codes (0 refinement, < 0 physical) unchanged; the r-th level-1
    box is run r's root."""
    ens = S.ensemble(level_list=[2, 3, 1, 2], seed=5)
    t = ens.export
    D = 2
    assert t["n_boxes"] == sum(c.extra["export"]["n_boxes"] for c in ens.cases)
    roots = np.nonzero(t["lvl"] == 1)[0] + 1
    assert list(roots) == [o + 1 for o in ens.offsets]
    width = {"lvl": 1, "ix": D, "parent": 1, "children": 4, "neighbors": 4, "nbmat": 9,
             "r_min": D, "dr": D}
    for r, c in enumerate(ens.cases):
        e = c.extra["export"]
        o, n = ens.offsets[r], e["n_boxes"]
        for k, w in width.items():
            got = t[k][o * w:(o + n) * w]
            exp = np.array(e[k], copy=True)
            if k in ("parent", "children", "neighbors", "nbmat"):
                exp[exp > 0] += o
            assert np.array_equal(got, exp), (r, k)
        assert np.array_equal(ens.ne0[o:o + n], c.ne0)
    # distinct voltages and seeds per run, levels as requested
    assert len(set(ens.voltages)) == len(ens.cases)
    assert all(15e3 <= v <= 25e3 for v in ens.voltages)
    assert [c.tree.info()["highest_lvl"] for c in ens.cases] == [2, 3, 1, 2]
