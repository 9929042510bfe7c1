"""Index maps of the box pools (the colour-split layout of the FP32 pools
and afivo's layout of the FP64 ones), checked on the host: every cell has
its own position, the neighbour and child helpers agree with the plain
map, and the half-sweep's thread mapping enumerates each colour block in
order.  The check is compiled from tests/layout_check.hip with hipcc and
run on the CPU."""
import os
import subprocess

import pytest

from conftest import ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_pool_layout_maps(tmp_path):
    exe = str(tmp_path / "layout_check")
    csrc = os.path.join(ROOT, "afivo-streamer_amd", "csrc")
    subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950", "-I" + csrc,
                    "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "layout_check.hip"),
                    "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 7 and all(l.endswith("mismatches 0") for l in lines)
