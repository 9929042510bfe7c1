"""The Fortran side of the boundary: include/m_afgpu.f90, the interface
module a Fortran driver uses (the reference's own precedent:
afivo/src/m_coarse_solver.f90:46-56, 187-189), compiled with amdflang.

- Every bind(c) type has C's size and every component C's offset
  (checked against offsetof / sizeof of include/afgpu.h, compiled with gcc).
- Host-only entry points called from a Fortran program
  (afgpu_coarse_factors, afgpu_format_regression_line, afgpu_create's
  argument check and afgpu_error_message) give what the C-ABI gives through
  ctypes, bit for bit.
The GPU-side driver (tests/fortran/drive_step.f90) is run by
tests/test_gpu_fortran.py."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FLANG = "/opt/rocm/bin/amdflang"
pytestmark = pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not present")

MODULE = os.path.join(ROOT, "include", "m_afgpu.f90")
PKG = os.path.join(ROOT, "afivo-streamer_amd")

# components of every bind(c) type, in C's order (array components by name)
TYPES = {
    "afgpu_params_t": ["gas_number_density", "time_integrator", "dt_max", "dt_min", "dt_safety_factor",
                       "dt_cfl_number", "dt_max_growth_factor", "dt_chemistry_limit_loss", "dt_chemistry_nmin",
                       "mg_num_vcycles", "mg_max_rel_residual", "mg_n_cycle_down", "mg_n_cycle_up",
                       "domain_len_last", "species_bc", "photoi_enabled", "photoi_per_steps", "photoi_species",
                       "photoi_eta_quench", "photoi_n_modes", "photoi_lambda", "photoi_coeff",
                       "photoi_max_rel_residual", "photoi_max_fmg_cycles"],
    "afgpu_step_info_t": ["it", "time", "dt_used", "dt_next", "dt_lim", "n_rejected", "n_vcycles", "residual",
                          "rel_residual", "photoi_done", "leaf_cells", "photoi_cycles"],
    "afgpu_refine_t": ["refine_adx", "refine_adx_fac", "derefine_dx", "refine_min_dx", "refine_max_dx",
                       "refine_min_dens", "refine_init_time", "refine_init_fac", "global_time",
                       "use_alpha_effective", "buffer_width", "max_lvl", "n_seeds", "seed_r0", "seed_r1",
                       "seed_width", "n_regions", "regions_dr", "regions_tstop", "regions_rmin", "regions_rmax",
                       "n_limits", "limits_dr", "limits_rmin", "limits_rmax"],
    "afgpu_shape_t": ["boxes_per_workgroup", "workgroups_per_cu", "ms"],
    "afgpu_reaction_t": ["rate_type", "lookup_col", "n_in", "n_out", "rate_factor", "rate_data", "ix_in", "ix_out",
                         "multiplicity_out"],
}


def run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, (cmd, r.stdout, r.stderr)
    return r.stdout


@pytest.fixture(scope="module")
def moddir(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("m_afgpu"))
    run([FLANG, "-c", MODULE, "-o", os.path.join(d, "m_afgpu.o"), "-module-dir", d], d)
    return d


def test_bind_c_types_match_c_layout(moddir):
    f = ["program layout", "  use iso_c_binding", "  use m_afgpu", "  implicit none",
         "  integer(c_intptr_t) :: b"]
    for t in TYPES:
        f.append("  type(%s), target :: v_%s" % (t, t))
    for t, fields in TYPES.items():
        f.append('  print "(A,1X,A,1X,I0)", "%s", "size", int(c_sizeof(v_%s))' % (t, t))
        f.append("  b = transfer(c_loc(v_%s), b)" % t)
        for fl in fields:
            f.append('  print "(A,1X,A,1X,I0)", "%s", "%s", transfer(c_loc(v_%s%%%s), b) - b' % (t, fl, t, fl))
    f.append("end program layout")
    c = ['#include <stdio.h>', '#include <stddef.h>', '#include "afgpu.h"', "int main(void) {"]
    for t, fields in TYPES.items():
        c.append('  printf("%s size %%zu\\n", sizeof(%s));' % (t, t))
        for fl in fields:
            c.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (t, fl, t, fl))
    c += ["  return 0;", "}"]
    with open(os.path.join(moddir, "layout.f90"), "w") as fh:
        fh.write("\n".join(f) + "\n")
    with open(os.path.join(moddir, "layout.c"), "w") as fh:
        fh.write("\n".join(c) + "\n")
    run([FLANG, "-I" + moddir, "layout.f90", "m_afgpu.o", "-L" + PKG, "-lafgpu", "-Wl,-rpath," + PKG,
         "-o", "layout_f"], moddir)
    run(["gcc", "-I" + os.path.join(ROOT, "include"), "layout.c", "-o", "layout_c"], moddir)
    got = run(["./layout_f"], moddir).split()
    exp = run(["./layout_c"], moddir).split()
    assert len(exp) == 3 * sum(len(v) + 1 for v in TYPES.values())
    assert got == exp, [(a, b) for a, b in zip(got, exp) if a != b][:6]


HOST_PROGRAM = """
program host_calls
  use iso_c_binding
  use m_afgpu
  implicit none
  real(c_double) :: r_min(3), dr(3), vinv(3*8*8), v(3*8*8), eig(3*8), vals(4)
  integer(c_int) :: bc(6), n, u
  character(kind=c_char) :: buf(256)
  type(c_ptr) :: h
  r_min = [0.0_c_double, 1.0e-3_c_double, 0.0_c_double]
  dr = [1.0e-3_c_double, 5.0e-4_c_double, 2.0e-3_c_double]
  bc = [0, 0, 0, 0, 1, 1]
  if (afgpu_coarse_factors(3, 8, afgpu_xyz, r_min, dr, bc, vinv, v, eig) /= afgpu_ok) &
       error stop "coarse factors"
  vals = [1.0_c_double, -2.5e-12_c_double, 0.0_c_double, 1.0e100_c_double]
  n = afgpu_format_regression_line(42, 4, vals, buf, 256)
  if (n < 0) error stop "format"
  ! an invalid dimension is refused with a message, before any GPU use
  if (afgpu_create(4, 8, afgpu_xyz, afgpu_prec_fp64, h) /= afgpu_err_arg) error stop "create check"
  open(newunit=u, file="host_out.bin", access="stream", form="unformatted", status="replace")
  write(u) vinv, v, eig, n, buf(1:n)
  close(u)
  print "(A)", afgpu_error_message()
end program host_calls
"""


def test_host_calls_from_fortran(A, moddir):
    src = os.path.join(moddir, "host_calls.f90")
    with open(src, "w") as fh:
        fh.write(HOST_PROGRAM)
    run([FLANG, "-I" + moddir, src, "m_afgpu.o", "-L" + PKG, "-lafgpu", "-Wl,-rpath," + PKG, "-o", "host_calls"],
        moddir)
    msg = run(["./host_calls"], moddir).strip()
    assert msg.startswith("afgpu_create: invalid arguments"), msg
    raw = open(os.path.join(moddir, "host_out.bin"), "rb").read()
    nn = 3 * 8 * 8
    vinv = np.frombuffer(raw, np.float64, nn, 0)
    v = np.frombuffer(raw, np.float64, nn, 8 * nn)
    eig = np.frombuffer(raw, np.float64, 24, 16 * nn)
    n, = struct.unpack("<i", raw[16 * nn + 192:16 * nn + 196])
    line = raw[16 * nn + 196:16 * nn + 196 + n].decode()
    ev, evv, ee = A.coarse_factors(3, 8, A.XYZ, [0.0, 1e-3, 0.0], [1e-3, 5e-4, 2e-3],
                                   np.array([0, 0, 0, 0, 1, 1], np.int32))
    assert np.array_equal(vinv, ev) and np.array_equal(v, evv) and np.array_equal(eig, ee)
    assert line == A.format_regression_line(42, [1.0, -2.5e-12, 0.0, 1e100])
