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
