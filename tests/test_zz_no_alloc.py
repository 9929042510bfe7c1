"""afgpu_regression_values_noalloc, the entry point that allocates no device
memory (its scratch is a static device array of the library), against
afgpu_regression_values, which allocates its scratch: bit for bit, on two
calls in a row (the second with a smaller tree and fewer variables, so it
runs over what the first left in the shared array), on a plain tree and on
a forest of runs; requests beyond the array's size are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def noalloc_fn(A):
    lib = A.load_gpu()
    f = lib.afgpu_regression_values_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    return f


def noalloc_values(A, p, ivs, runs=1):
    ivs = np.ascontiguousarray(ivs, np.int32)
    out = np.full((runs, len(ivs), 3), np.nan)
    rc = noalloc_fn(A)(p.h, len(ivs), ivs.ctypes.data_as(C.POINTER(C.c_int)),
                       out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_noalloc_equals_allocating_entry_point(A, S):
    big = S.small_amr(3, nc=16, uniform_lvl=2, max_lvl=3, frac=0.3, seed=7)
    small = S.small_amr(2, coord=2, uniform_lvl=3, max_lvl=4, frac=0.3, seed=3)
    gb = S.apply(A.Path("gpu", big.D, big.nc, big.coord, A.PREC_FP64), big)
    gs = S.apply(A.Path("gpu", small.D, small.nc, small.coord, A.PREC_FP64), small)
    for p in (gb, gs):
        p.field_compute(0, have_guess=False)
        p.step()
    ivs_big = [A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION, A.CC_EFIELD, A.CC_PHI]
    ivs_small = [A.CC_EFIELD]
    print("noalloc: 3D 16^3 tree, 5 variables", flush=True)
    rc, v1 = noalloc_values(A, gb, ivs_big)
    assert rc == A.OK, A.last_error("gpu")
    assert np.array_equal(v1, gb.regression_values(ivs_big))
    print("noalloc: 2D cylindrical tree, 1 variable", flush=True)
    rc, v2 = noalloc_values(A, gs, ivs_small)
    assert rc == A.OK, A.last_error("gpu")
    assert np.array_equal(v2, gs.regression_values(ivs_small))
    # and back: the larger request after the smaller one
    rc, v3 = noalloc_values(A, gb, ivs_big)
    assert rc == A.OK and np.array_equal(v3, v1)


def test_noalloc_forest_of_runs(A, S):
    ens = S.ensemble(level_list=[4, 3, 5], num_vcycles=20, rel_residual=1e-10)
    g = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL), ens)
    g.field_compute_runs(0, have_guess=False)
    g.step_ensemble()
    ivs = [A.CC_ELECTRON, A.CC_EFIELD]
    print("noalloc: forest of 3 runs, 2 variables", flush=True)
    rc, v = noalloc_values(A, g, ivs, runs=g.num_runs())
    assert rc == A.OK, A.last_error("gpu")
    assert np.array_equal(v, g.regression_values(ivs))
    rc, v1 = noalloc_values(A, g, ivs[:1], runs=g.num_runs())
    assert rc == A.OK and np.array_equal(v1, g.regression_values(ivs[:1]))


def test_noalloc_refuses_requests_beyond_its_array(A, S):
    """A uniform 4096^2 tree (262144 leaf boxes) with 8 variables needs
    (262144 + 1) * 8 * 3 doubles, more than AFGPU_NOALLOC_SCRATCH_WORDS;
    with one variable it fits."""
    case = S.config0(uniform_lvl=10)
    g = S.apply(A.Path("gpu", case.D, case.nc, case.coord, A.PREC_FP64), case)
    print("noalloc: refusal on 262144 leaf boxes", flush=True)
    rc, _ = noalloc_values(A, g, [A.CC_ELECTRON] * 8)
    assert rc == 1  # AFGPU_ERR_ARG
    assert "AFGPU_NOALLOC_SCRATCH_WORDS" in A.last_error("gpu")
    rc, v = noalloc_values(A, g, [A.CC_ELECTRON])
    assert rc == A.OK, A.last_error("gpu")
    assert np.array_equal(v, g.regression_values([A.CC_ELECTRON]))
