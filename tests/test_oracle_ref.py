"""The oracle against the reference's own routines, bit for bit.

The vectors come from drivers (oracle/ref) that call afivo's and the
lookup table's routines compiled from the reference's sources
(oracle/ref/Makefile, amdflang, no fused multiply-adds) on seeded inputs;
the committed fixtures are tests/golden/ref/*.npz
(tests/golden/make_ref_fixtures.py).  Covered:
- af_limiter_koren, af_limiter_mc, af_limiter_gminmod43 on 4096 pairs that
  reach every branch (m_af_limiters.f90:72-151);
- the transport table of transport_data_initialize (table_set_column,
  LT_create, get_x) and the standard chemistry's field table, against the
  tables the host stand-in builds; LT_get_loc / LT_get_col at 3000 points
  inside, on and outside the table;
- on AMR trees (2D Cartesian and cylindrical, 8^2 boxes over four levels;
  3D, 4^3 boxes over three levels): af_gc_tree with limited and linear
  interpolation at refinement boundaries, Neumann and Dirichlet
  boundaries and corners; af_restrict_tree; af_gc2_box; stencil_gsrb_357
  and stencil_apply_357 with the multigrid's Laplacian and a Helmholtz
  term; stencil_prolong_248; af_prolong_limit and af_prolong_linear;
  af_consistent_fluxes; flux_upwind_tree with the fluid model's electron
  flux; flux_update_densities.
Where the reference build is present (oracle/_ref, built by build() from
the reference checkout), the drivers are run again: their output must equal
the committed fixtures, and a 3D tree of 8^3 boxes is checked as well."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import refvec

TREES = ("tree_xyz2d", "tree_cyl2d", "tree_xyz3d")
_P, _D, _I = C.c_void_p, C.c_double, C.c_int
_DP, _IP = C.POINTER(C.c_double), C.POINTER(C.c_int)
HOOKS = {
    "test_limiter": [_I, _I, _DP, _DP, _DP],
    "test_table": [_I, _D, _D, _DP, _I, _DP, _DP, _IP, _DP],
    "test_gsrb_boxes": [_P, _I, _IP, _I, _I, _I, _D],
    "test_apply_boxes": [_P, _I, _IP, _I, _I, _D],
    "test_prolong_boxes": [_P, _I, _IP, _I, _I, _I],
    "test_gc2_boxes": [_P, _I, _IP, _I, _DP],
    "test_consistent_fluxes": [_P],
    "test_flux_upwind_tree": [_P, _I, _DP],
    "test_flux_divergence": [_P, _D],
}


def dp(x):
    return x.ctypes.data_as(_DP)


def ip(x):
    return x.ctypes.data_as(_IP)


@pytest.fixture(scope="module")
def lib(A):
    L = A.load_oracle()
    for name, args in HOOKS.items():
        f = getattr(L, "afcpu_" + name)
        f.restype, f.argtypes = C.c_int, args
    return L


def same(a, b):
    """Equal bit for bit (NaN patterns included)."""
    a, b = np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def where_differs(a, b):
    d = np.nonzero(a.view(np.uint64) != b.view(np.uint64))[0]
    return "%d cells differ, first at %s: %r vs %r" % (len(d), d[:5], a[d[:5]], b[d[:5]])


# ---- point-wise routines -------------------------------------------------------
def test_limiters_equal_reference(lib):
    d = refvec.load("pointwise")
    a, b = d["lim.a"], d["lim.b"]
    # the pairs reach every branch of the Koren limiter and of gminmod
    ab, aa = a * b, a * a
    assert (ab <= 0).any() and (aa <= 0.25 * ab).any()
    assert ((aa > 0.25 * ab) & (aa <= 2.5 * ab)).any() and ((ab > 0) & (aa > 2.5 * ab)).any()
    for kind, name in enumerate(("lim.koren", "lim.mc", "lim.gminmod43")):
        out = np.zeros_like(a)
        assert lib.afcpu_test_limiter(kind, len(a), dp(a), dp(b), dp(out)) == 0
        assert same(out, d[name]), (name, where_differs(out, d[name]))


def test_transport_tables_equal_reference(A):
    """The host stand-in's tables (what the oracle and the device receive)
    equal transport_data_initialize's td_tbl and the standard chemistry's
    chemtbl_fld bit for bit: grid, x_min, inv_fac and every column."""
    d = refvec.load("pointwise")
    tb = A.transport_tables(d["in.E"], d["in.mu"], d["in.dif"], d["in.alpha"], d["in.eta"], d["in.N"][0])
    n = tb["n"]
    assert same(np.array([tb["td_x_min"], tb["td_inv_fac"]]), np.array([d["td.x_min"][0], d["td.inv_fac"][0]]))
    assert same(tb["td"], d["td.cols"]), where_differs(tb["td"], d["td.cols"])
    assert same(np.array([tb["chem_x_min"], tb["chem_inv_fac"]]),
                np.array([d["chem.x_min"][0], d["chem.inv_fac"][0]]))
    assert same(tb["chem"], d["chem.cols"]), where_differs(tb["chem"], d["chem.cols"])
    assert len(d["td.x"]) == n


def test_chemistry_field_table_equals_reference(CH):
    """chemistry.py's own table construction (the path of reaction-list
    files: lt_grid as LT_create's get_x, spaced_data as LT_lin_interp_list
    in LT_set_col) rebuilds the standard chemistry's field table that the
    reference's table_set_column made from the same rates on the same
    range (src/m_table_data.f90:82-118, m_lookup_table.f90:161-181, 240-255)
    bit for bit."""
    d = refvec.load("pointwise")
    x = d["td.x"]
    n_td = len(x)
    td = d["td.cols"].reshape(-1, n_td)  # rows_cols (n_points, n_cols), column by column
    n_chem = d["chem.cols"].size // 2
    chem = d["chem.cols"].reshape(2, n_chem)
    townsend_to_si, gas_n = 1.0e-21, d["in.N"][0]
    grid = CH.lt_grid(x[0], x[-1], n_chem)
    assert same(np.array([grid[0]]), d["chem.x_min"])
    for col, k in ((0, 2), (1, 3)):  # ionisation (alpha), attachment (eta)
        rate = td[k] * td[0] * x * townsend_to_si * gas_n
        got = CH.spaced_data(x, rate, grid)
        assert same(got, chem[col]), (col, where_differs(got, chem[col]))


def test_table_lookups_equal_reference(lib):
    d = refvec.load("pointwise")
    x = d["look.x"]
    n = len(d["td.x"])
    for tab, ncol in (("td", 4), ("chem", 2)):
        cols = d[tab + ".cols"].reshape(ncol, n)
        for c in range(ncol):
            y = np.zeros_like(x)
            ix = np.zeros(len(x), np.int32)
            fr = np.zeros_like(x)
            col = np.ascontiguousarray(cols[c])
            assert lib.afcpu_test_table(n, d[tab + ".x_min"][0], d[tab + ".inv_fac"][0], dp(col), len(x), dp(x),
                                        dp(y), ip(ix), dp(fr)) == 0
            assert same(y, d["look.%s%d" % (tab, c + 1)]), (tab, c, where_differs(y, d["look.%s%d" % (tab, c + 1)]))
            if tab == "td":
                assert np.array_equal(ix, d["look.low_ix"])
                assert same(fr, d["look.low_frac"])


# ---- tree routines ---------------------------------------------------------------
# the variables of the drivers' trees and the oracle variables with the same
# methods: a (Neumann zero, limited interpolation and prolongation) is the
# electron density, b (Neumann zero, linear interpolation) |E|, c (Dirichlet
# zero on the last dimension's faces, linear interpolation) photo
CC = {"a": "CC_ELECTRON", "b": "CC_EFIELD", "c": "CC_PHOTO", "phi": "CC_PHI", "rhs": "CC_RHS", "tmp": "CC_TMP"}
FC = {"flux": "FC_FLUX", "field": "FC_FIELD"}


def shape_of(d):
    t = refvec.tree_dict(d)
    D = 3 if len(t["nbmat"]) == 27 * t["n_boxes"] else 2
    nc = int(round((len(d["init.a"]) / t["n_boxes"]) ** (1.0 / D))) - 2
    return D, nc, int(d["tree.coord"][0])


def loaded(A, S, d, kind="oracle"):
    t = refvec.tree_dict(d)
    D, nc, coord = shape_of(d)
    p = A.Path(kind, D, nc, coord)
    p.sync_tree(t)
    prm = S.default_params(D, nc * t["dr"][D - 1], d["gas_n"][0])
    p.set_params(prm)
    n = len(d["td.cols"]) // 4
    inv = d["td.inv_fac"][0]
    p.set_tables({"n": n, "td_x_min": d["td.x_min"][0], "td_inv_fac": inv, "td": d["td.cols"],
                  "chem_x_min": d["td.x_min"][0], "chem_inv_fac": inv, "chem": np.zeros(2 * n)})
    for v, iv in CC.items():
        p.put_cc(getattr(A, iv), d["init." + v])
    for v, ifc in FC.items():
        p.put_fc(getattr(A, ifc), d["init." + v])
    return p, t


def check_cc(A, p, d, key):
    op, var = key.split(".")
    assert not same(d[key], d["init." + var]), key  # the operation changed something
    got = p.get_cc(getattr(A, CC[var])).ravel()
    assert same(got, d[key]), (key, where_differs(got, d[key]))


def check_fc(A, p, d, key):
    op, var = key.split(".")
    assert not same(d[key], d["init." + var]), key
    got = p.get_fc(getattr(A, FC[var])).ravel()
    assert same(got, d[key]), (key, where_differs(got, d[key]))


def all_ids(t):
    return np.arange(1, t["n_boxes"] + 1, dtype=np.int32)


def run_tree_ops(A, S, lib, d):
    """Every operation of the driver, each from the driver's inputs."""
    p, t = loaded(A, S, d)
    ids = all_ids(t)
    # ghost cells of every box, corners included (af_gc_tree)
    for v in ("a", "b", "c"):
        p.gc_tree(getattr(A, CC[v]))
        check_cc(A, p, d, "gc." + v)

    p, t = loaded(A, S, d)
    p.restrict_tree(A.CC_ELECTRON)
    check_cc(A, p, d, "restrict.a")

    p, t = loaded(A, S, d)
    leaves = np.ascontiguousarray(d["gc2.ids"], np.int32)
    halo = np.zeros_like(d["gc2.halo"])
    assert lib.afcpu_test_gc2_boxes(p.h, len(leaves), ip(leaves), A.CC_ELECTRON, dp(halo)) == 0
    assert same(halo, d["gc2.halo"]), where_differs(halo, d["gc2.halo"])
    check_cc(A, p, d, "gc2.a")

    lam = d["lambda2"][0]
    for tag, l2 in (("0", 0.0), ("L", lam)):
        p, t = loaded(A, S, d)
        for q in ids:  # box by box: red then black
            one = np.array([q], np.int32)
            for rb in (1, 2):
                assert lib.afcpu_test_gsrb_boxes(p.h, 1, ip(one), rb, A.CC_PHI, A.CC_RHS, l2) == 0
        check_cc(A, p, d, "gsrb%s.phi" % tag)
        p, t = loaded(A, S, d)
        assert lib.afcpu_test_apply_boxes(p.h, len(ids), ip(ids), A.CC_PHI, A.CC_TMP, l2) == 0
        check_cc(A, p, d, "apply%s.tmp" % tag)

    kids = np.ascontiguousarray(ids[t["parent"] > 0])
    for mode, key, iv, iv_to in ((0, "prolong.tmp", A.CC_PHI, A.CC_TMP), (1, "prolim.a", A.CC_ELECTRON, 0),
                                 (2, "prolin.b", A.CC_EFIELD, 0)):
        p, t = loaded(A, S, d)
        assert lib.afcpu_test_prolong_boxes(p.h, len(kids), ip(kids), mode, iv, iv_to) == 0
        check_cc(A, p, d, key)

    p, t = loaded(A, S, d)
    assert lib.afcpu_test_consistent_fluxes(p.h) == 0
    check_fc(A, p, d, "consistent.flux")

    p, t = loaded(A, S, d)
    lim = np.zeros(2)
    assert lib.afcpu_test_flux_upwind_tree(p.h, 0, dp(lim)) == 0
    check_fc(A, p, d, "flux.flux")
    check_cc(A, p, d, "flux.a")
    assert same(lim, d["flux.dt_lim"]), (lim, d["flux.dt_lim"])

    p, t = loaded(A, S, d)
    assert lib.afcpu_test_flux_divergence(p.h, d["dt_update"][0]) == 0
    check_cc(A, p, d, "update.a")


@pytest.mark.parametrize("name", TREES)
def test_tree_routines_equal_reference(A, S, lib, name):
    d = refvec.load(name)
    t = refvec.tree_dict(d)
    # the trees have refinement boundaries and boxes on every level
    assert (t["neighbors"] == 0).any() and (t["neighbors"] < 0).any()
    assert len(np.unique(t["lvl"])) >= 3
    run_tree_ops(A, S, lib, d)


# ---- the reference build itself, where it is present ------------------------------
REF_BUILT = os.path.exists(os.path.join(refvec.REPO, "oracle", "_ref", "2d", "drv_tree")) and \
    os.path.isdir("/root/reference")


@pytest.mark.skipif(not REF_BUILT, reason="the reference build (oracle/_ref) is not present")
def test_reference_build_reproduces_fixtures_and_3d_nc8(A, S, lib):
    subprocess.run(["make", "-s", "-C", os.path.join(refvec.REPO, "oracle", "ref"), "vectors"], check=True,
                   stdout=subprocess.DEVNULL)
    for name in ("pointwise",) + TREES:
        fresh = refvec.read_bin(os.path.join(refvec.REF_OUT, name + ".bin"))
        kept = refvec.load(name)
        assert sorted(fresh) == sorted(kept), name
        for k in fresh:
            assert np.array_equal(fresh[k].view(np.uint8), kept[k].view(np.uint8)), (name, k)
    run_tree_ops(A, S, lib, refvec.read_bin(os.path.join(refvec.REF_OUT, "tree_xyz3d_nc8.bin")))
