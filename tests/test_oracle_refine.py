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


# ---- data of refined and derefined boxes (auto_prolong / auto_restrict) ------
def centres(nc, D, rmin, dr, ghosts=True):
    """Cell-centre coordinates of a box as arrays in (k, j, i) C order."""
    rng = np.arange(0, nc + 2) if ghosts else np.arange(1, nc + 1)
    axes = [rmin[d] + (rng - 0.5) * dr[d] for d in range(D)]
    return np.meshgrid(*axes[::-1], indexing="ij")[::-1]


def box_view(row, nc, D):
    return row.reshape((nc + 2,) * D)


def interior(v, D):
    return v[(slice(1, -1),) * D]


def refine_some(case, pick):
    """Refine the leaves chosen by pick(export) with the host tree; returns
    the new export and the ids of the new boxes (appended after the old)."""
    t0 = case.extra["export"]
    ids = pick(t0)
    case.tree.refine(np.asarray(ids, np.int32))
    t1 = case.tree.export()
    return t1, np.arange(t0["n_boxes"] + 1, t1["n_boxes"] + 1, dtype=np.int32)


def inner_leaves(t, D, lvl):
    """Leaves of a level whose face neighbours all exist (refining them keeps
    the 2:1 balance)."""
    n, nch, nnb = t["n_boxes"], 2 ** D, 2 * D
    return [b + 1 for b in range(n) if t["lvl"][b] == lvl and t["children"][b * nch] == 0
            and all(v != 0 for v in t["neighbors"][b * nnb:(b + 1) * nnb])]


AUTO = ("CC_PHOTO", "CC_EFIELD", "CC_PHI", "CC_ELECTRON", "CC_POS_ION", "CC_NEG_ION", "CC_RHS")


def centres_all(t, b, D, nc):
    rmin, dr = t["r_min"][b * D:(b + 1) * D], t["dr"][b * D:(b + 1) * D]
    return centres(nc, D, rmin, dr)


@pytest.mark.parametrize("D", [2, 3])
def test_prolongation_of_new_boxes_reproduces_linear_fields(A, S, D):
    case = uniform_tree(S, D, 3 if D == 2 else 2)
    case.params.photoi_enabled = 1
    p = oracle(A, S, case)
    t0 = case.extra["export"]
    n0, nc = t0["n_boxes"], case.nc
    g = [1.3e3, -0.7e3, 0.4e3][:D]
    for k, name in enumerate(AUTO):  # a different linear field per variable, ghosts included
        vals = np.stack([sum((k + 1) * g[d] * X[d] for d in range(D)) + 5.0 * (k + 1)
                         for X in (centres_all(t0, b, D, nc) for b in range(n0))])
        p.put_cc(getattr(A, name), vals.reshape(n0, -1))
    t1, new = refine_some(case, lambda t: inner_leaves(t, D, t["lvl"].max())[:3])
    p.sync_tree(t1)
    p.prolong_boxes(new)
    for k, name in enumerate(AUTO):
        u = p.get_cc(getattr(A, name))
        for b in new - 1:
            X = centres(nc, D, t1["r_min"][b * D:(b + 1) * D], t1["dr"][b * D:(b + 1) * D], ghosts=False)
            exact = sum((k + 1) * g[d] * X[d] for d in range(D)) + 5.0 * (k + 1)
            got = interior(box_view(u[b], nc, D), D)
            assert np.abs(got - exact).max() <= 1e-12 * np.abs(exact).max(), (name, b + 1)
    # restriction back into the parents gives the parents' (linear) values
    parents = sorted({int(t1["parent"][b]) for b in new - 1})
    before = {name: p.get_cc(getattr(A, name))[np.array(parents) - 1] for name in AUTO}
    p.restrict_boxes(parents)
    for name in AUTO:
        after = p.get_cc(getattr(A, name))[np.array(parents) - 1]
        assert np.abs(after - before[name]).max() <= 1e-12 * np.abs(before[name]).max(), name


def test_limited_prolongation_keeps_extrema_and_means(A, S):
    """Densities (af_prolong_limit): a parent cell at a local maximum gives
    its children its own value (the limited slopes vanish) where phi
    (af_prolong_linear) interpolates; every group of 2^D children keeps the
    parent cell's value as its mean."""
    D = 2
    case = uniform_tree(S, D, 3)
    p = oracle(A, S, case)
    t0 = case.extra["export"]
    n0, nc = t0["n_boxes"], case.nc
    rng = np.random.default_rng(4)
    base = rng.uniform(1.0, 2.0, (n0, nc + 2, nc + 2))
    leaf = inner_leaves(t0, D, 3)[0]
    peak = (5, 5)  # (j, i) of a parent cell inside the refined quadrant region
    base[leaf - 1][peak] = 10.0
    for name in ("CC_ELECTRON", "CC_PHI"):
        p.put_cc(getattr(A, name), base.reshape(n0, -1))
    t1, new = refine_some(case, lambda t: [leaf])
    p.sync_tree(t1)
    p.prolong_boxes(new)
    ne, phi = p.get_cc(A.CC_ELECTRON), p.get_cc(A.CC_PHI)
    H = nc // 2
    for b in new - 1:
        off = [(int(t1["ix"][b * D + d]) - 1) % 2 * H for d in range(D)]
        for j in range(1, H + 1):
            for i in range(1, H + 1):
                jc, ic = off[1] + j, off[0] + i
                kids = box_view(ne[b], nc, D)[2 * j - 1:2 * j + 1, 2 * i - 1:2 * i + 1]
                pv = base[leaf - 1][jc, ic]
                assert abs(kids.mean() - pv) <= 1e-14 * pv
                if (jc, ic) == peak:
                    assert np.all(kids == pv)
                    lin = box_view(phi[b], nc, D)[2 * j - 1:2 * j + 1, 2 * i - 1:2 * i + 1]
                    assert np.all(lin < pv)
