"""Golden vectors of the reference's own routines (written by the drivers
under oracle/ref, built from the reference's sources by oracle/ref/Makefile;
tests/golden/make_ref_fixtures.py turns them into the committed fixtures
tests/golden/ref/*.npz).

read_bin reads a driver's raw record stream; load reads a committed
fixture.  In a fixture, the arrays of an operation's result ("gc.a",
"prolim.a", ...) are stored as the XOR of their bit patterns with the
operation's input ("init.a"), so that the cells an operation leaves alone
compress to nothing; load undoes it."""
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "ref")
REPO = os.path.dirname(HERE)
REF_OUT = os.path.join(REPO, "oracle", "_ref", "out")


def read_bin(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            nm = f.read(32)
            if not nm:
                break
            code, = struct.unpack("<i", f.read(4))
            n, = struct.unpack("<q", f.read(8))
            dt = np.dtype("<f8") if code == 1 else np.dtype("<i4")
            out[nm.decode().strip()] = np.frombuffer(f.read(n * dt.itemsize), dt).copy()
    return out


def base_of(name, keys):
    """The input array an operation's result is stored against, or None."""
    op, _, var = name.partition(".")
    if op in ("init", "tree", "td", "in", "look", "lim", "chem") or not var:
        return None
    b = "init." + var
    return b if b in keys else None


def encode(d):
    out = {}
    for k, v in d.items():
        b = base_of(k, d)
        if b is not None and d[b].shape == v.shape and v.dtype == np.float64:
            out["xor:" + k] = v.view(np.uint64) ^ d[b].view(np.uint64)
        else:
            out[k] = v
    return out


def decode(z):
    d = {k: z[k] for k in z.files if not k.startswith("xor:")}
    for k in z.files:
        if k.startswith("xor:"):
            name = k[4:]
            d[name] = (z[k] ^ d[base_of(name, d)].view(np.uint64)).view(np.float64)
    return d


def load(name):
    with np.load(os.path.join(FIXTURES, name + ".npz")) as z:
        return decode(z)


def tree_dict(d):
    """The tree arrays in the form Path.sync_tree takes."""
    t = {k[5:]: d[k] for k in d if k.startswith("tree.")}
    t["n_boxes"] = len(t["lvl"])
    for k in ("lvl", "ix", "parent", "children", "neighbors", "nbmat"):
        t[k] = np.ascontiguousarray(t[k], np.int32)
    for k in ("r_min", "dr"):
        t[k] = np.ascontiguousarray(t[k], np.float64)
    return t
