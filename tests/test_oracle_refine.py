"""Pins of the oracle's refinement flags (no GPU): af_adjust_refinement's
consistent_ref_flags (afivo/src/m_af_core.f90:929-1145) with the streamer's
default_refinement (src/m_refine.f90:198-298).

Each expectation is derived from the geometry and the closed forms, not
from a second copy of the algorithm:
* one flagged cell in a uniform tree refines its box and exactly those
  neighbours whose buffer slab (the buffer_width cells next to the shared
  side, edge or corner) holds the cell;
* the alpha*dx rule flips between refinement and keeping where max_dx times
  alpha(f E/N)/f N crosses refine_adx, with the table read as LT_get_col;
* the initial-condition rule refines the boxes with a cell centre within
  seed width + 2 dx of the seed segment (distances computed here with
  numpy's projection onto the segment);
* refine_max_dx refines everything, refine_min_dx forbids refinement;
* with no field everything may be derefined, and exactly the parents whose
  children are all leaves without refined same-level face neighbours are.
"""
import itertools

import numpy as np
import pytest

REFINE, DEREFINE = 2, -2


def uniform_tree(S, D, lvl, L=4e-3, nc=8):
    r0 = [L / 2] * (D - 1) + [0.7 * L]
    return S.make_case(D, nc, 1, [L] * D, 1e3, lvl, r0, seed_width=0.3e-3)


def oracle(A, S, case):
    p = A.Path("oracle", case.D, case.nc, case.coord)
    return S.apply(p, case)


def fields(p, case, E_fn=None, ne=1e15):
    """Put |E| (zero, or E_fn(box index, cell (i, j[, k]) arrays)) and n_e."""
    t = case.extra["export"]
    D, nc, n = case.D, case.nc, t["n_boxes"]
    E = np.zeros((n,) + (nc + 2,) * D)
    if E_fn is not None:
        for b in range(n):
            E[b] = E_fn(b)
    p.put_cc(p.A.CC_EFIELD, E.reshape(n, -1))
    p.put_cc(p.A.CC_ELECTRON, np.full((n, (nc + 2) ** D), ne))


def box_at(t, D, lvl, ix):
    for b in range(t["n_boxes"]):
        if t["lvl"][b] == lvl and list(t["ix"][b * D:(b + 1) * D]) == list(ix):
            return b
    return None


@pytest.mark.parametrize("D", [2, 3])
def test_one_flagged_cell_refines_its_box_and_the_buffer_neighbours(A, S, D):
    lvl = 3 if D == 2 else 2
    case = uniform_tree(S, D, lvl)
    p = oracle(A, S, case)
    p.A = A
    t = case.extra["export"]
    nc, n = case.nc, t["n_boxes"]
    nb_per_dim = 2 ** (lvl - 1)
    E_hot = 3e7  # alpha * dx ~ 30 on this level
    targets = [(2,) * D, (1,) * D, (nb_per_dim,) + (1,) * (D - 1)]
    cells = [(1,) * D, (nc,) * D, (2, nc - 1) + (nc // 2,) * (D - 2), (nc // 2,) * D, (nc - 3, 4) + (1,) * (D - 2)]
    for ix_box in targets:
        B = box_at(t, D, lvl, ix_box)
        for cell in cells:
            def E_fn(b, B=B, cell=cell):
                e = np.zeros((nc + 2,) * D)
                if b == B:
                    e[tuple(reversed(cell))] = E_hot
                return e
            fields(p, case, E_fn)
            for buf in (0, 1, 2, 4, nc):
                s = A.refine_settings(buffer_width=buf, derefine_dx=0.0, global_time=1.0)
                f = p.refine_flags(s, n)
                expect = np.zeros(n, np.int32)
                expect[B] = REFINE
                if buf > 0:
                    for d3 in itertools.product((-1, 0, 1), repeat=D):
                        if not any(d3):
                            continue
                        hit = all(dd == 0 or (dd < 0 and cell[k] <= buf) or (dd > 0 and cell[k] >= nc - buf + 1)
                                  for k, dd in enumerate(d3))
                        other = box_at(t, D, lvl, [ix_box[k] + d3[k] for k in range(D)])
                        if hit and other is not None:
                            expect[other] = REFINE
                assert np.array_equal(f, expect), (ix_box, cell, buf, np.nonzero(f != expect)[0])


def test_alpha_dx_threshold(A, S):
    """Uniform |E| on every cell: refinement where max_dx * alpha(f E)/f
    exceeds refine_adx, with alpha = N * LT_get_col(alpha/N, E * 1e21 / N)."""
    case = uniform_tree(S, 2, 3)
    p = oracle(A, S, case)
    p.A = A
    t = case.extra["export"]
    n = t["n_boxes"]
    td = case.tdata
    N = case.params.gas_number_density
    alpha_col = td["td"][2 * td["n"]:3 * td["n"]]
    x_nodes = td["td_x_min"] + np.arange(td["n"]) / td["td_inv_fac"]
    dx = t["dr"][2 * (n - 1)]  # a level-3 box (cubic: max_dx = dx)
    leaves = [b for b in range(n) if t["children"][b * 4] == 0]
    for E, fac in ((1.2e7, 1.0), (2.5e7, 1.0), (2.5e7, 1.5)):
        fields(p, case, lambda b: np.full((case.nc + 2,) * 2, E))
        alpha = np.interp(fac * E * 1e21 / N, x_nodes, alpha_col) * N / fac
        adx = dx * alpha
        for rel, expect in ((0.99, REFINE), (1.01, 0)):
            s = A.refine_settings(refine_adx=rel * adx, refine_adx_fac=fac, derefine_dx=0.0, global_time=1.0,
                                  buffer_width=0)
            f = p.refine_flags(s, n)
            assert all(f[b] == expect for b in leaves), (E, fac, rel, f[leaves])


def seg_dist(P, r0, r1):
    r0, r1 = np.asarray(r0, float), np.asarray(r1, float)
    d = r1 - r0
    tt = np.clip(((P - r0) @ d) / (d @ d), 0.0, 1.0)
    return np.linalg.norm(P - (r0 + tt[:, None] * d), axis=1)


def test_initial_seed_rule(A, S):
    case = uniform_tree(S, 2, 4)  # 8 x 8 leaf boxes, dx = 62.5 um
    p = oracle(A, S, case)
    p.A = A
    t = case.extra["export"]
    n, nc = t["n_boxes"], case.nc
    fields(p, case)
    r0, r1, w = (1.0e-3, 1.2e-3), (2.7e-3, 3.1e-3), 0.2e-3
    s = A.refine_settings(derefine_dx=0.0, global_time=0.0, refine_init_time=1e-9, buffer_width=0,
                          seeds=[(r0, r1, w)])
    f = p.refine_flags(s, n)
    n_ref = 0
    for b in range(n):
        if t["children"][b * 4] != 0:
            continue
        rmin, dr = t["r_min"][2 * b:2 * b + 2], t["dr"][2 * b:2 * b + 2]
        ii, jj = np.meshgrid(np.arange(1, nc + 1), np.arange(1, nc + 1), indexing="ij")
        P = np.stack([rmin[0] + (ii.ravel() - 0.5) * dr[0], rmin[1] + (jj.ravel() - 0.5) * dr[1]], axis=1)
        margin = seg_dist(P, r0, r1) - w - 2 * dr.max()
        assert np.abs(margin).min() > 1e-12  # no cell on the threshold
        expect = REFINE if (margin < 0).any() and dr.max() > 0.25 * w else 0
        assert f[b] == expect, b
        n_ref += expect == REFINE
    assert 0 < n_ref < 64


def test_grid_spacing_bounds(A, S):
    case = uniform_tree(S, 2, 3)
    p = oracle(A, S, case)
    p.A = A
    t = case.extra["export"]
    n = t["n_boxes"]
    leaves = [b for b in range(n) if t["children"][b * 4] == 0]
    dx = t["dr"][2 * leaves[0]]
    fields(p, case)  # no field: nothing to refine
    f = p.refine_flags(A.refine_settings(refine_max_dx=0.9 * dx, derefine_dx=0.0, global_time=1.0), n)
    assert all(f[b] == REFINE for b in leaves)
    fields(p, case, lambda b: np.full((case.nc + 2,) * 2, 3e7))  # everything above the threshold
    f = p.refine_flags(A.refine_settings(derefine_dx=0.0, global_time=1.0), n)
    assert all(f[b] == REFINE for b in leaves)
    f = p.refine_flags(A.refine_settings(refine_min_dx=0.6 * dx, derefine_dx=0.0, global_time=1.0), n)
    assert not (f == REFINE).any()


@pytest.mark.parametrize("D", [2, 3])
def test_derefinement_without_field(A, S, D):
    case = S.small_amr(D, nc=8, uniform_lvl=2, max_lvl=4, frac=0.4, seed=9)
    p = oracle(A, S, case)
    p.A = A
    t = case.extra["export"]
    n, nch, nnb = t["n_boxes"], 2 ** D, 2 * D
    fields(p, case)
    f = p.refine_flags(A.refine_settings(derefine_dx=1.0, global_time=1.0), n)
    children = t["children"].reshape(n, nch)
    nbs = t["neighbors"].reshape(n, nnb)
    has_ch = children[:, 0] > 0
    expect = np.zeros(n, np.int32)
    for b in range(n):
        if not has_ch[b]:
            continue
        ch = children[b] - 1
        if not all(not has_ch[c] for c in ch):
            continue
        stays = all(not any(nb > 0 and has_ch[nb - 1] for nb in nbs[c]) for c in ch)
        if stays:
            expect[b] = DEREFINE
            expect[ch] = -1
    assert np.array_equal(f, expect), np.nonzero(f != expect)[0]
    assert (expect == DEREFINE).sum() > 0
