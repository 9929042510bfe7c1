"""The C-ABI library loads and exports every entry point declared in
include/afgpu.h (no GPU call is made)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "afgpu.h")).read()
    return sorted(set(re.findall(r"\b(afgpu_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("afgpu_create", "afgpu_sync_tree", "afgpu_mg_fas_vcycle", "afgpu_mg_fas_fmg",
              "afgpu_field_compute", "afgpu_forward_euler", "afgpu_photoi_set_src", "afgpu_step"):
        assert s in syms


def test_library_exports_all_declared(A):
    lib = ctypes.CDLL(A.GPU_LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_oracle_mirrors_entry_points(A):
    lib = ctypes.CDLL(A.ORACLE_LIB)
    skip = {"afgpu_coarse_factors", "afgpu_kernel_timing", "afgpu_kernel_stats", "afgpu_set_option",
            # ensembles are checked run by run against single-run oracle handles
            "afgpu_num_runs", "afgpu_set_run_voltages", "afgpu_field_compute_runs",
            "afgpu_step_ensemble", "afgpu_get_run_time_state", "afgpu_set_run_time_state",
            # work decomposition of the device kernels
            "afgpu_autotune", "afgpu_get_shapes", "afgpu_set_shapes",
            # the log writer is host code around afgpu_regression_values; its
            # format is checked against the reference's own logs
            "afgpu_format_regression_line", "afgpu_write_regression_log",
            # device memory for the chemistry arrays (the oracle's are host arrays)
            "afgpu_device_alloc", "afgpu_device_free", "afgpu_device_copy",
            # the variant that allocates nothing is checked against
            # afgpu_regression_values on the device
            "afgpu_regression_values_noalloc"}
    missing = [s for s in declared_symbols() if s not in skip and not hasattr(lib, s.replace("afgpu_", "afcpu_"))]
    assert not missing, missing


def test_coarse_factors_is_host_only(A):
    import numpy as np
    vinv, v, eig = A.coarse_factors(3, 8, A.XYZ, np.zeros(3), np.full(3, 1e-3),
                                    np.array([0, 0, 0, 0, 1, 1], np.int32))
    for d in range(3):
        Vi = vinv[d * 64:(d + 1) * 64].reshape(8, 8)
        V = v[d * 64:(d + 1) * 64].reshape(8, 8)
        assert np.abs(V @ Vi - np.eye(8)).max() < 1e-13
