"""The HIP path against the reference's own routines, bit for bit: the
tree-level ghost-cell fill and restriction of the device library
(afgpu_gc_tree, afgpu_restrict_tree) on the 2D trees of the committed
reference vectors (tests/golden/ref, written by afivo's af_gc_tree and
af_restrict_tree built from the reference's sources; see
tests/test_oracle_ref.py for the oracle's side of the same vectors)."""
import numpy as np
import pytest

import refvec
from test_oracle_ref import CC, loaded, same, shape_of, where_differs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["tree_xyz2d", "tree_cyl2d"])
def test_device_gc_and_restrict_equal_reference(A, S, name):
    d = refvec.load(name)
    assert shape_of(d)[1] == 8
    print("reference vectors", name, flush=True)
    for v in ("a", "b", "c"):
        g, _ = loaded(A, S, d, "gpu")
        g.gc_tree(getattr(A, CC[v]))
        got = g.get_cc(getattr(A, CC[v])).ravel()
        assert same(got, d["gc." + v]), (name, v, where_differs(got, d["gc." + v]))
    g, _ = loaded(A, S, d, "gpu")
    g.restrict_tree(A.CC_ELECTRON)
    got = g.get_cc(A.CC_ELECTRON).ravel()
    assert same(got, d["restrict.a"]), (name, where_differs(got, d["restrict.a"]))
