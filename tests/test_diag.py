"""Diagnostics (CPU): the regression-log line format against the reference's
own rtest logs, the oracle's regression statistics against closed forms,
and box I/O in af_write_tree's order.

- afgpu_format_regression_line is host code of the device library (no GPU
  use); every line of the 23 regression logs the reference's rtests compare
  against (tests/golden/regression_logs.json, from
  programs/*/tests/*_rtest.log) must come out character for character from
  its own numbers.
- The oracle's statistics (output_regression_log, src/m_output.f90:797-810)
  are pinned by integrals the midpoint rule gets exactly: constant and
  linear fields on AMR trees, a quadratic one on a uniform tree with the
  midpoint rule's known defect, the 2 pi r weight of cylindrical
  coordinates; the maximum reads leaf interiors only.
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "regression_logs.json")


def test_regression_lines_reproduce_reference_logs(A):
    logs = json.load(open(GOLDEN))
    assert len(logs) == 23
    n = 0
    for name, log in logs.items():
        for line in log["lines"]:
            tok = line.split()
            out = A.format_regression_line(int(tok[0]), [float(t) for t in tok[1:]])
            assert out == line, (name, line, out)
            n += 1
    assert n > 300


@pytest.mark.parametrize("x, text", [
    (0.0, "     0.00000000E+000"),
    (1.0, "     0.10000000E+001"),
    (-2.5e-7, "    -0.25000000E-006"),
    (0.999999999, "     0.10000000E+001"),    # the eighth digit carries into the exponent
    (1e-300, "     0.10000000E-299"),
    (6.02214076e23, "     0.60221408E+024"),
])
def test_fortran_e20_8e3(A, x, text):
    assert A.format_regression_line(7, [x]) == "7" + text


def test_format_parses_back(A):
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, 200) * 10.0 ** rng.integers(-250, 250, 200)
    line = A.format_regression_line(3, x)
    tok = line.split()
    assert tok[0] == "3" and len(line) == 1 + 20 * len(x)
    back = np.array([float(t) for t in tok[1:]])
    assert np.allclose(back, x, rtol=5.01e-8, atol=0)  # half a unit in the 8th digit
    for t in tok[1:]:
        m = t.lstrip("-")
        assert m.startswith("0.") and m[2] != "0" and len(m) == 15 and m[10] == "E"


def set_field(p, case, fn):
    """Every cell of every box (ghosts included) gets fn(coordinates)."""
    t = case.extra["export"]
    D, nc, n = case.D, case.nc, t["n_boxes"]
    vals = np.zeros((n, (nc + 2) ** D))
    for b in range(n):
        rmin, dr = t["r_min"][b * D:(b + 1) * D], t["dr"][b * D:(b + 1) * D]
        axes = [rmin[d] + (np.arange(nc + 2) - 0.5) * dr[d] for d in range(D)]
        X = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
        vals[b] = np.asarray(fn(*X), float).reshape(-1)
    return vals


def domain(case):
    t = case.extra["export"]
    b1 = int(np.nonzero(t["lvl"] == 1)[0][0])
    D = case.D
    return t["r_min"][b1 * D:(b1 + 1) * D], t["dr"][b1 * D:(b1 + 1) * D] * case.nc


def oracle_case(A, S, case):
    p = A.Path("oracle", case.D, case.nc, case.coord)
    return S.apply(p, case)


AMR_CASES = {
    "cyl2d": lambda S: S.small_amr(2, coord=2, uniform_lvl=3, max_lvl=5, frac=0.3, seed=3),
    "xyz2d": lambda S: S.small_amr(2, coord=1, uniform_lvl=2, max_lvl=4, frac=0.4, seed=5),
    "xyz3d": lambda S: S.small_amr(3, nc=8, uniform_lvl=2, max_lvl=3, frac=0.3, seed=11),
}


@pytest.mark.parametrize("name", list(AMR_CASES))
def test_regression_values_constant_and_linear_fields(A, S, name):
    """Constant c: (c, c^2, c).  Linear a + b z (z the last axis): mean
    a + b L/2 on any refinement, the midpoint rule being exact for it (and,
    in cylindrical coordinates, for the weight r)."""
    case = AMR_CASES[name](S)
    p = oracle_case(A, S, case)
    r0, L = domain(case)
    c = 3.7e14
    p.put_cc(A.CC_ELECTRON, set_field(p, case, lambda *X: np.full_like(X[0], c)))
    a, b = 2.0e14, 5.0e16
    p.put_cc(A.CC_POS_ION, set_field(p, case, lambda *X: a + b * (X[-1] - r0[-1])))
    v = p.regression_values([A.CC_ELECTRON, A.CC_POS_ION])[0]
    assert v[0, 0] == pytest.approx(c, rel=1e-14)
    assert v[0, 1] == pytest.approx(c * c, rel=1e-14)
    assert v[0, 2] == c
    assert v[1, 0] == pytest.approx(a + b * L[-1] / 2, rel=1e-14)


@pytest.mark.parametrize("coord, D", [(1, 2), (2, 2), (1, 3)])
def test_regression_values_quadratic_on_uniform_tree(A, S, coord, D):
    """n = a + b z on a uniform tree: the mean of n^2 is the integral mean
    minus the midpoint rule's defect b^2 h^2 / 12 (h the cell size along
    z); in cylindrical coordinates the mean of n = r (weighted by 2 pi r)
    follows from the same defect."""
    lvl = 3 if D == 2 else 2
    case = S.make_case(D, 8, coord, [4e-3] * D, 1e3, lvl, [2e-3] * (D - 1) + [2.8e-3])
    p = oracle_case(A, S, case)
    r0, L = domain(case)
    h = L / (8 * 2 ** (lvl - 1))
    a, b = 1.0, 250.0
    p.put_cc(A.CC_ELECTRON, set_field(p, case, lambda *X: a + b * (X[-1] - r0[-1])))
    v = p.regression_values([A.CC_ELECTRON])[0]
    Lz = L[-1]
    mean_sq = (a * a + a * b * Lz + b * b * Lz * Lz / 3) - b * b * h[-1] ** 2 / 12
    assert v[0, 1] == pytest.approx(mean_sq, rel=1e-13)
    if coord == 2:
        # n = r: sum over cells of r_c * (2 pi r_c h_r h_z) / (pi R^2 Lz); with
        # r_c = (m - 1/2) h_r the sum of r_c^2 h_r is R^3/3 - R h_r^2 / 12
        p.put_cc(A.CC_POS_ION, set_field(p, case, lambda *X: X[0]))
        v = p.regression_values([A.CC_POS_ION])[0]
        R, hr = L[0], h[0]
        assert v[0, 0] == pytest.approx(2 * (R ** 3 / 3 - R * hr * hr / 12) / R ** 2, rel=1e-13)


@pytest.mark.parametrize("name", ["cyl2d", "xyz3d"])
def test_regression_max_reads_leaf_interiors_only(A, S, name):
    """A huge value in every ghost cell and in the interiors of boxes with
    children is not seen; one planted leaf cell is the maximum."""
    case = AMR_CASES[name](S)
    p = oracle_case(A, S, case)
    t = case.extra["export"]
    D, nc, n = case.D, case.nc, t["n_boxes"]
    rng = np.random.default_rng(2)
    vals = np.full((n,) + (nc + 2,) * D, 1e30)
    inner = (slice(None),) + (slice(1, nc + 1),) * D
    vals[inner] = rng.uniform(0.0, 1.0, vals[inner].shape)
    nch = 2 ** D
    parents = [b for b in range(n) if t["children"][b * nch] > 0]
    leaves = [b for b in range(n) if t["children"][b * nch] == 0]
    vals[parents] = 1e30
    b = leaves[len(leaves) // 2]
    vals[(b,) + (3,) * D] = 7.5
    p.put_cc(A.CC_NEG_ION, vals.reshape(n, -1))
    assert p.regression_values([A.CC_NEG_ION])[0, 0, 2] == 7.5


def test_box_io_faces_and_species_oracle(A, S, CH):
    """Box data of cell variables, reaction-list species and face variables
    (AFGPU_VAR_FACE) in one list, af_write_tree's order
    (afivo/src/m_af_output.f90:90-114): per box the variables in turn, a
    cell variable or species (n_cell + 2)^D words, a face variable
    D (n_cell + 1)^D words as box%fc(DTIMES(:), :, ifc); each block equals
    the variable's own transfer, and put_boxes writes the same layout back.
    Species 3 and up live outside the standard pools."""
    import os
    from conftest import ROOT
    case = AMR_CASES["xyz2d"](S)
    p = oracle_case(A, S, case)
    N = case.params.gas_number_density
    E = S.analytic_transport()[0]
    net = CH.read_network(os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt"), {"N2": 0.8, "O2": 0.2}, N,
                          0.0, E[-1] * 1e21 / N, 1000)
    ns = len(net.species)
    chem = CH.Chem("oracle", ns, net.reactions, 0.0, E[-1] * 1e21 / N, net.fld_cols, None, 300.0)
    p.set_chemistry(chem, net.charges)
    n = case.extra["export"]["n_boxes"]
    rng = np.random.default_rng(21)
    for k in range(ns):
        p.put_species(k, rng.normal(size=(n, p.ncell)))
    for f in (A.FC_FLUX, A.FC_FIELD):
        p.put_fc(f, rng.normal(size=(n, p.nface)))
    p.put_cc(A.CC_PHI, rng.normal(size=(n, p.ncell)))
    ivs = [A.CC_PHI, A.VAR_SPECIES(1), A.VAR_SPECIES(5), A.VAR_FACE(A.FC_FLUX), A.VAR_FACE(A.FC_FIELD)]
    ids = np.array([4, 1, 7, 2], np.int32)
    got = p.get_boxes(ivs, ids)
    assert got.shape == (len(ids), 3 * p.ncell + 2 * p.nface)
    blocks = [p.get_cc(A.CC_PHI, ids), p.get_species(1, ids=ids), p.get_species(5, ids=ids),
              p.get_fc(A.FC_FLUX, ids), p.get_fc(A.FC_FIELD, ids)]
    assert np.array_equal(got, np.concatenate(blocks, axis=1))
    new = rng.normal(size=got.shape)
    p.put_boxes(ivs, new, ids)
    assert np.array_equal(p.get_boxes(ivs, ids), new)
    assert np.array_equal(p.get_species(5, ids=ids), new[:, 2 * p.ncell:3 * p.ncell])
    assert np.array_equal(p.get_fc(A.FC_FIELD, ids), new[:, 3 * p.ncell + p.nface:])


def test_box_io_order_oracle(A, S):
    """get_boxes returns box by box, variable by variable (af_write_tree's
    order); put_boxes writes the same layout back."""
    case = AMR_CASES["xyz2d"](S)
    p = oracle_case(A, S, case)
    n = case.extra["export"]["n_boxes"]
    ivs = [A.CC_ELECTRON, A.CC_PHI, A.CC_POS_ION]
    rng = np.random.default_rng(9)
    for iv in ivs:
        p.put_cc(iv, rng.normal(size=(n, p.ncell)))
    ids = np.array([5, 2, 9, 3], np.int32)
    got = p.get_boxes(ivs, ids)
    for k, iv in enumerate(ivs):
        assert np.array_equal(got[:, k], p.get_cc(iv, ids))
    new = rng.normal(size=got.shape)
    p.put_boxes(ivs, new, ids)
    assert np.array_equal(p.get_boxes(ivs, ids), new)
