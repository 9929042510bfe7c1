"""Writes tests/golden/ref/*.npz from the reference's own routines.

Builds the fixture drivers of oracle/ref against the reference's sources
(make -C oracle/ref vectors: needs the reference checkout and amdflang),
then stores the committed fixtures (pointwise, the 2D trees and the small
3D tree) in the compressed form tests/refvec.py reads."""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import refvec  # noqa: E402

COMMITTED = ("pointwise", "tree_xyz2d", "tree_cyl2d", "tree_xyz3d")


def main():
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle", "ref"), "vectors"], check=True)
    os.makedirs(refvec.FIXTURES, exist_ok=True)
    for name in COMMITTED:
        d = refvec.read_bin(os.path.join(refvec.REF_OUT, name + ".bin"))
        np.savez_compressed(os.path.join(refvec.FIXTURES, name + ".npz"), **refvec.encode(d))
        back = refvec.load(name)
        assert all(np.array_equal(back[k].view(np.uint8), d[k].view(np.uint8)) for k in d)
        print(name, os.path.getsize(os.path.join(refvec.FIXTURES, name + ".npz")), "bytes")


if __name__ == "__main__":
    main()
