"""The C-ABI driven from Fortran on the GPU: tests/fortran/drive_step.f90
(built by build() over include/m_afgpu.f90 and libafgpu.so) runs create,
sync_tree, set_params, set_tables, set_voltage, put_cc, the initial
field_compute and time steps on a small 3D AMR case and a 2D cylindrical
one; its step results and final electron density, potential and field
equal the CPU oracle's on the same inputs bit for bit."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "drive_step")


def write_input(path, A, case, n_steps):
    t = case.extra["export"]
    td = case.tdata
    with open(path, "wb") as f:
        f.write(struct.pack("<7i", case.D, case.nc, case.coord, A.PREC_FP64, t["n_boxes"], n_steps, td["n"]))
        for k in ("lvl", "ix", "parent", "children", "neighbors", "nbmat"):
            f.write(np.ascontiguousarray(t[k], "<i4").tobytes())
        for k in ("r_min", "dr"):
            f.write(np.ascontiguousarray(t[k], "<f8").tobytes())
        f.write(bytes(case.params))
        f.write(struct.pack("<4d", td["td_x_min"], td["td_inv_fac"], td["chem_x_min"], td["chem_inv_fac"]))
        f.write(np.ascontiguousarray(td["td"], "<f8").tobytes())
        f.write(np.ascontiguousarray(td["chem"], "<f8").tobytes())
        f.write(struct.pack("<d", case.voltage))
        f.write(np.ascontiguousarray(case.ne0, "<f8").tobytes())


def read_output(path, A, n_steps, n_values):
    raw = open(path, "rb").read()
    n_cyc, res, rec = struct.unpack("<idi", raw[:16])
    assert rec == A.C.sizeof(A.StepInfoT)
    infos = [A.StepInfoT.from_buffer_copy(raw, 16 + i * rec).as_dict() for i in range(n_steps)]
    off = 16 + n_steps * rec
    arrays = [np.frombuffer(raw, np.float64, n_values, off + q * 8 * n_values) for q in range(3)]
    return (n_cyc, res), infos, arrays


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="the Fortran driver is not built (no amdflang)")
@pytest.mark.parametrize("name", ["xyz3d_16", "cyl2d_amr"])
def test_fortran_driver_equals_oracle(A, S, tmp_path, name):
    case = {"xyz3d_16": lambda: S.small_amr(3, nc=16, uniform_lvl=2, max_lvl=3, frac=0.3, seed=7),
            "cyl2d_amr": lambda: S.small_amr(2, coord=2, uniform_lvl=2, max_lvl=4, seed=3)}[name]()
    n_steps = 3
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, A, case, n_steps)
    print("fortran driver", name, flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    n_boxes = case.extra["export"]["n_boxes"]
    field, infos, arrays = read_output(fout, A, n_steps, n_boxes * (case.nc + 2) ** case.D)
    c = S.apply(A.Path("oracle", case.D, case.nc, case.coord), case)
    assert field == c.field_compute(0, have_guess=False)
    for it in range(n_steps):
        assert infos[it] == c.step(), (name, it)
    for got, iv in zip(arrays, (A.CC_ELECTRON, A.CC_PHI, A.CC_EFIELD)):
        assert np.array_equal(got, c.get_cc(iv).ravel()), (name, iv)
