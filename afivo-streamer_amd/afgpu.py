"""ctypes bindings for the afivo-streamer hot path on MI355X.

Three shared libraries are involved:

* ``libafgpu.so`` -- the product: HIP kernels for gfx950 behind the C-ABI in
  ``include/afgpu.h``.
* ``libafhost.so`` -- host-side stand-in for what stays on the CPU in the
  reference (AMR tree, initial conditions, lookup tables).
* ``oracle/liboracle_afcpu.so`` -- the CPU oracle (test infrastructure; only
  tests, the smoke check and the CPU-baseline leg of bench.py load it).

``Path`` wraps one handle of either the device library or the oracle with
the same method names as the reference's routines (``field_compute``,
``mg_fas_vcycle``, ``forward_euler`` ...), so that parity tests read like the
reference's own call sequence.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
GPU_LIB = os.path.join(PKG_DIR, "libafgpu.so")
HOST_LIB = os.path.join(PKG_DIR, "libafhost.so")
ORACLE_LIB = os.path.join(REPO_DIR, "oracle", "liboracle_afcpu.so")

# ---- constants mirrored from include/afgpu.h -------------------------------
OK = 0
XYZ, CYL = 1, 2
PREC_FP64, PREC_MIXED = 0, 1
FORWARD_EULER, HEUNS_METHOD = 1, 2
CC_ELECTRON, CC_POS_ION, CC_NEG_ION = 0, 3, 6
CC_PHI, CC_PHI_2, CC_EFIELD, CC_RHS, CC_TMP, CC_PHOTO, CC_HELMH = 9, 10, 11, 12, 13, 14, 15
N_CC = 18


def VAR_SPECIES(k):
    """Species k of a reaction list in variable lists (AFGPU_VAR_SPECIES)."""
    return 100 + k
FC_FLUX, FC_FIELD = 0, 1


def VAR_FACE(ifc):
    """Face variable ifc in the box I/O variable lists (AFGPU_VAR_FACE)."""
    return 200 + ifc
OP_POISSON, OP_HELMH = 0, 1
MAX_MODES = 3
OPT_GRAPHS = 1
OPT_DIRECT_N = 2
OPT_MIXED_SMOOTH = 3
OPT_BLOCK_SMOOTH = 4
OPT_BATCH_MODES = 5
OPT_DIRECT_MFMA = 6
OPT_PHOTOI_VCYCLE = 7
OPT_FLUX32 = 8
OPT_STEP_GRAPH = 9
OPT_FUSE_DIRECT = 10
OPT_DIRECT_F32 = 11
OPT_BOX_LDS = 12
OPT_DIRECT_FOLD = 13


class ParamsT(C.Structure):
    _fields_ = [
        ("gas_number_density", C.c_double),
        ("time_integrator", C.c_int),
        ("dt_max", C.c_double),
        ("dt_min", C.c_double),
        ("dt_safety_factor", C.c_double),
        ("dt_cfl_number", C.c_double),
        ("dt_max_growth_factor", C.c_double),
        ("dt_chemistry_limit_loss", C.c_int),
        ("dt_chemistry_nmin", C.c_double),
        ("mg_num_vcycles", C.c_int),
        ("mg_max_rel_residual", C.c_double),
        ("mg_n_cycle_down", C.c_int),
        ("mg_n_cycle_up", C.c_int),
        ("domain_len_last", C.c_double),
        ("species_bc", C.c_int),
        ("photoi_enabled", C.c_int),
        ("photoi_per_steps", C.c_int),
        ("photoi_species", C.c_int),
        ("photoi_eta_quench", C.c_double),
        ("photoi_n_modes", C.c_int),
        ("photoi_lambda", C.c_double * MAX_MODES),
        ("photoi_coeff", C.c_double * MAX_MODES),
        ("photoi_max_rel_residual", C.c_double),
        ("photoi_max_fmg_cycles", C.c_int),
    ]


REF_MAX_ITEMS = 8
REF_DEREFINE, REF_REFINE = -2, 2


class RefineT(C.Structure):
    """afgpu_refine_t: the refinement settings of src/m_refine.f90:8-83."""
    _fields_ = [
        ("refine_adx", C.c_double),
        ("refine_adx_fac", C.c_double),
        ("derefine_dx", C.c_double),
        ("refine_min_dx", C.c_double),
        ("refine_max_dx", C.c_double),
        ("refine_min_dens", C.c_double),
        ("refine_init_time", C.c_double),
        ("refine_init_fac", C.c_double),
        ("global_time", C.c_double),
        ("use_alpha_effective", C.c_int),
        ("buffer_width", C.c_int),
        ("max_lvl", C.c_int),
        ("n_seeds", C.c_int),
        ("seed_r0", (C.c_double * 3) * REF_MAX_ITEMS),
        ("seed_r1", (C.c_double * 3) * REF_MAX_ITEMS),
        ("seed_width", C.c_double * REF_MAX_ITEMS),
        ("n_regions", C.c_int),
        ("regions_dr", C.c_double * REF_MAX_ITEMS),
        ("regions_tstop", C.c_double * REF_MAX_ITEMS),
        ("regions_rmin", (C.c_double * 3) * REF_MAX_ITEMS),
        ("regions_rmax", (C.c_double * 3) * REF_MAX_ITEMS),
        ("n_limits", C.c_int),
        ("limits_dr", C.c_double * REF_MAX_ITEMS),
        ("limits_rmin", (C.c_double * 3) * REF_MAX_ITEMS),
        ("limits_rmax", (C.c_double * 3) * REF_MAX_ITEMS),
    ]


def refine_settings(**kw):
    """A RefineT with the reference's defaults (src/m_refine.f90:8-50:
    buffer 4, min dx 1e-7, max dx 1e-3, adx 1, adx factor 1, derefine dx
    1e-4, init time 10 ns, init factor 0.25, af_max_lvl 30) and no seeds,
    regions or limits; keyword arguments override fields, and seeds /
    regions / limits are given as lists of tuples (r0, r1, width),
    (dr, tstop, rmin, rmax) and (dr, rmin, rmax)."""
    r = RefineT()
    r.refine_adx, r.refine_adx_fac, r.derefine_dx = 1.0, 1.0, 1e-4
    r.refine_min_dx, r.refine_max_dx, r.refine_min_dens = 1e-7, 1e-3, -1e99
    r.refine_init_time, r.refine_init_fac, r.global_time = 10e-9, 0.25, 0.0
    r.use_alpha_effective, r.buffer_width, r.max_lvl = 0, 4, 30
    for k, v in kw.items():
        if k == "seeds":
            r.n_seeds = len(v)
            for n, (r0, r1, w) in enumerate(v):
                for d in range(len(r0)):
                    r.seed_r0[n][d], r.seed_r1[n][d] = r0[d], r1[d]
                r.seed_width[n] = w
        elif k == "regions":
            r.n_regions = len(v)
            for n, (dr, tstop, rmin, rmax) in enumerate(v):
                r.regions_dr[n], r.regions_tstop[n] = dr, tstop
                for d in range(len(rmin)):
                    r.regions_rmin[n][d], r.regions_rmax[n][d] = rmin[d], rmax[d]
        elif k == "limits":
            r.n_limits = len(v)
            for n, (dr, rmin, rmax) in enumerate(v):
                r.limits_dr[n] = dr
                for d in range(len(rmin)):
                    r.limits_rmin[n][d], r.limits_rmax[n][d] = rmin[d], rmax[d]
        else:
            setattr(r, k, v)
    return r


class StepInfoT(C.Structure):
    _fields_ = [
        ("it", C.c_int),
        ("time", C.c_double),
        ("dt_used", C.c_double),
        ("dt_next", C.c_double),
        ("dt_lim", C.c_double),
        ("n_rejected", C.c_int),
        ("n_vcycles", C.c_int),
        ("residual", C.c_double),
        ("rel_residual", C.c_double),
        ("photoi_done", C.c_int),
        ("leaf_cells", C.c_longlong),
        ("photoi_cycles", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ShapeT(C.Structure):
    _fields_ = [("boxes_per_workgroup", C.c_int), ("workgroups_per_cu", C.c_int), ("ms", C.c_double)]


N_KERNEL_CLASSES = 8
KERNEL_CLASSES = ("gsrb", "residual", "restrict", "prolong", "ghost", "boxop", "flux", "update")

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes) for the functions common to afgpu_/afcpu_
_COMMON = {
    "create": (_I, [_I, _I, _I, _I, C.POINTER(_P)]),
    "destroy": (_I, [_P]),
    "last_error": (C.c_char_p, []),
    "sync_tree": (_I, [_P, _I, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _DP]),
    "put_cc": (_I, [_P, _I, _I, _IP, _DP]),
    "get_cc": (_I, [_P, _I, _I, _IP, _DP]),
    "put_fc": (_I, [_P, _I, _I, _IP, _DP]),
    "get_fc": (_I, [_P, _I, _I, _IP, _DP]),
    "set_params": (_I, [_P, C.POINTER(ParamsT)]),
    "set_tables": (_I, [_P, _I, _D, _D, _DP, _D, _D, _DP]),
    "set_voltage": (_I, [_P, _D]),
    "mg_fas_vcycle": (_I, [_P, _I, _I, _I]),
    "mg_fas_fmg": (_I, [_P, _I, _I, _I]),
    "tree_maxabs_cc": (_I, [_P, _I, _DP]),
    "gc_tree": (_I, [_P, _I]),
    "restrict_tree": (_I, [_P, _I]),
    "field_compute": (_I, [_P, _I, _I, _IP, _DP]),
    "field_from_potential": (_I, [_P]),
    "forward_euler": (_I, [_P, _D, _I, _I, _IP, _DP, _I, _I, _I, _DP]),
    "advance": (_I, [_P, _D, _DP]),
    "photoi_set_src": (_I, [_P]),
    "copy_current_state": (_I, [_P]),
    "restore_previous_state": (_I, [_P]),
    "step": (_I, [_P, C.POINTER(StepInfoT)]),
    "num_leaf_cells": (C.c_longlong, [_P]),
    "get_time_state": (_I, [_P, _IP, _DP, _DP]),
    "set_time_state": (_I, [_P, _I, _D, _D]),
    "limit_step_dt": (_I, [_P, _D]),
    "regression_values": (_I, [_P, _I, _IP, _DP]),
    "set_chemistry": (_I, [_P, _P, _IP]),
    "get_species": (_I, [_P, _I, _I, _I, _IP, _DP]),
    "put_species": (_I, [_P, _I, _I, _I, _IP, _DP]),
    "get_boxes": (_I, [_P, _I, _IP, _I, _IP, _DP]),
    "put_boxes": (_I, [_P, _I, _IP, _I, _IP, _DP]),
}
_GPU_ONLY = {
    "format_regression_line": (_I, [_I, _I, _DP, C.c_char_p, _I]),
    "write_regression_log": (_I, [_P, C.c_char_p, _I, _I, _IP, C.POINTER(C.c_char_p)]),
    "coarse_factors": (_I, [_I, _I, _I, _DP, _DP, _IP, _DP, _DP, _DP]),
    "kernel_timing": (_I, [_P, _I]),
    "kernel_stats": (_I, [_P, _I, _I, _DP, C.POINTER(C.c_longlong), _DP]),
    "set_option": (_I, [_P, _I, _I]),
    "num_runs": (_I, [_P]),
    "set_run_voltages": (_I, [_P, _I, _DP]),
    "field_compute_runs": (_I, [_P, _I, _I, _I, _IP, _DP]),
    "step_ensemble": (_I, [_P, _I, C.POINTER(StepInfoT)]),
    "get_run_time_state": (_I, [_P, _I, _IP, _DP, _DP]),
    "autotune": (_I, [_P, _I, C.POINTER(ShapeT)]),
    "get_shapes": (_I, [_P, _I, C.POINTER(ShapeT)]),
    "set_shapes": (_I, [_P, _I, C.POINTER(ShapeT)]),
    "set_run_time_state": (_I, [_P, _I, _I, _D, _D]),
}
_ORACLE_ONLY = {
    "set_coarse_factors": (_I, [_P, _I, _DP, _DP, _DP]),
    "num_threads": (_I, []),
}

# general chemistry (afgpu_chem_*): rate types with the reference's codes
RATE_TABULATED_ENERGY, RATE_TABULATED_FIELD, RATE_CONSTANT, RATE_LINEAR = 0, 1, 2, 3
RATE_EXP_V1, RATE_EXP_V2, RATE_K1 = 4, 5, 6
(RATE_K3, RATE_K4, RATE_K5, RATE_K6, RATE_K7, RATE_K8, RATE_K9, RATE_K10, RATE_K11, RATE_K12, RATE_K13,
 RATE_K14, RATE_K15, RATE_K16) = range(8, 22)
CHEM_MAX_IN = CHEM_MAX_OUT = 8


class ReactionT(C.Structure):
    _fields_ = [("rate_type", C.c_int), ("lookup_col", C.c_int), ("n_in", C.c_int), ("n_out", C.c_int),
                ("rate_factor", C.c_double), ("rate_data", C.c_double * 9),
                ("ix_in", C.c_int * CHEM_MAX_IN), ("ix_out", C.c_int * CHEM_MAX_OUT),
                ("multiplicity_out", C.c_int * CHEM_MAX_OUT)]


_VP = C.c_void_p
_CHEM = {
    "chem_create": (_I, [_I, _I, C.POINTER(ReactionT), _I, _D, _D, _I, _DP, _DP, _D, C.POINTER(_VP)]),
    "chem_destroy": (_I, [_VP]),
    "chem_rates": (_I, [_VP, _I, _VP, _VP]),
    "chem_derivatives": (_I, [_VP, _I, _VP, _VP, _VP]),
    "chem_source": (_I, [_VP, _I, _VP, _VP, _VP]),
}
_GPU_MEM = {
    "device_alloc": (_I, [C.c_size_t, C.POINTER(_VP)]),
    "device_free": (_I, [_VP]),
    "device_copy": (_I, [_VP, _VP, C.c_size_t, _I]),
}

_libs = {}


class AfError(RuntimeError):
    pass


def _bind(lib, prefix, table):
    for name, (res, args) in table.items():
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args


def load_gpu():
    """Load the device library; raises if it is missing (no fallback)."""
    if "gpu" not in _libs:
        if not os.path.exists(GPU_LIB):
            raise AfError("libafgpu.so not built: run __graft_entry__.build()")
        lib = C.CDLL(GPU_LIB)
        _bind(lib, "afgpu_", _COMMON)
        _bind(lib, "afgpu_", _GPU_ONLY)
        _bind(lib, "afgpu_", _CHEM)
        _bind(lib, "afgpu_", _GPU_MEM)
        _libs["gpu"] = lib
    return _libs["gpu"]


def load_oracle():
    if "oracle" not in _libs:
        if not os.path.exists(ORACLE_LIB):
            raise AfError("oracle library not built: run __graft_entry__.build()")
        lib = C.CDLL(ORACLE_LIB)
        _bind(lib, "afcpu_", _COMMON)
        _bind(lib, "afcpu_", _ORACLE_ONLY)
        _bind(lib, "afcpu_", _CHEM)
        _libs["oracle"] = lib
    return _libs["oracle"]


def load_host():
    if "host" not in _libs:
        if not os.path.exists(HOST_LIB):
            raise AfError("libafhost.so not built: run __graft_entry__.build()")
        lib = C.CDLL(HOST_LIB)
        lib.afhost_tree_create.restype = _I
        lib.afhost_tree_create.argtypes = [_I, _I, _I, _DP, _DP, C.POINTER(_P)]
        lib.afhost_tree_destroy.argtypes = [_P]
        lib.afhost_refine.restype = _I
        lib.afhost_refine.argtypes = [_P, _I, _IP]
        lib.afhost_refine_uniform.restype = _I
        lib.afhost_refine_uniform.argtypes = [_P, _I]
        lib.afhost_derefine.restype = _I
        lib.afhost_derefine.argtypes = [_P, _I, _IP]
        lib.afhost_tree_info.restype = _I
        lib.afhost_tree_info.argtypes = [_P, _IP, _IP, C.POINTER(C.c_longlong), _IP]
        lib.afhost_tree_export.restype = _I
        lib.afhost_tree_export.argtypes = [_P, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _DP]
        lib.afhost_init_cond.restype = _I
        lib.afhost_init_cond.argtypes = [_P, _I, _IP, _D, _D, C.c_ulonglong, _D, _D, _DP, _DP]
        lib.afhost_transport_tables.restype = _I
        lib.afhost_transport_tables.argtypes = [_I, _DP, _DP, _DP, _DP, _DP, _D, _I, _D,
                                                _DP, _DP, _DP, _DP, _DP, _DP]
        _libs["host"] = lib
    return _libs["host"]


def _ip(a):
    return a.ctypes.data_as(_IP)


def _dp(a):
    return a.ctypes.data_as(_DP)


# ---- host tree --------------------------------------------------------------
class Tree:
    """The box tree (stand-in for afivo's af_t on the host)."""

    def __init__(self, ndim, nc, coord, r_min, domain_len):
        self.lib = load_host()
        self.D, self.nc, self.coord = ndim, nc, coord
        self.h = _P()
        r = np.ascontiguousarray(r_min, dtype=np.float64)
        L = np.ascontiguousarray(domain_len, dtype=np.float64)
        if self.lib.afhost_tree_create(ndim, nc, coord, _dp(r), _dp(L), C.byref(self.h)):
            raise AfError("afhost_tree_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.afhost_tree_destroy(self.h)
            self.h = None

    def refine_uniform(self, lvl):
        if self.lib.afhost_refine_uniform(self.h, lvl) < 0:
            raise AfError("refine_uniform failed")

    def refine(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        n = self.lib.afhost_refine(self.h, len(ids), _ip(ids))
        if n < 0:
            raise AfError("refine failed")
        return n

    def derefine(self, ids):
        """Remove the children (all leaves) of the given boxes; their slots
        become unused (level 0)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        n = self.lib.afhost_derefine(self.h, len(ids), _ip(ids))
        if n < 0:
            raise AfError("derefine failed")
        return n

    def info(self):
        nb, hl, nl = C.c_int(), C.c_int(), C.c_int()
        cells = C.c_longlong()
        self.lib.afhost_tree_info(self.h, C.byref(nb), C.byref(hl), C.byref(cells), C.byref(nl))
        return {"n_boxes": nb.value, "highest_lvl": hl.value, "leaf_cells": cells.value,
                "n_leaves": nl.value}

    def export(self):
        n = self.info()["n_boxes"]
        D = self.D
        a = {
            "lvl": np.zeros(n, np.int32),
            "ix": np.zeros(n * D, np.int32),
            "parent": np.zeros(n, np.int32),
            "children": np.zeros(n * 2 ** D, np.int32),
            "neighbors": np.zeros(n * 2 * D, np.int32),
            "nbmat": np.zeros(n * 3 ** D, np.int32),
            "r_min": np.zeros(n * D, np.float64),
            "dr": np.zeros(n * D, np.float64),
        }
        self.lib.afhost_tree_export(self.h, _ip(a["lvl"]), _ip(a["ix"]), _ip(a["parent"]),
                                    _ip(a["children"]), _ip(a["neighbors"]), _ip(a["nbmat"]),
                                    _dp(a["r_min"]), _dp(a["dr"]))
        a["n_boxes"] = n
        return a

    def init_cond(self, ids, background, noise, seed, seed_density, seed_width, seed_r0):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        ncell = (self.nc + 2) ** self.D
        out = np.zeros((len(ids), ncell), np.float64)
        r0 = np.ascontiguousarray(seed_r0, dtype=np.float64)
        if self.lib.afhost_init_cond(self.h, len(ids), _ip(ids), background, noise, seed,
                                     seed_density, seed_width, _dp(r0), _dp(out)):
            raise AfError("init_cond failed")
        return out


def transport_tables(E, mu, dif, alpha, eta, N, table_size=1000, table_max_td=-1.0):
    lib = load_host()
    E = np.ascontiguousarray(E, np.float64)
    args = [np.ascontiguousarray(x, np.float64) for x in (mu, dif, alpha, eta)]
    td = np.zeros(4 * table_size)
    ch = np.zeros(2 * table_size)
    tx0, tif, cx0, cif = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    if lib.afhost_transport_tables(len(E), _dp(E), *[_dp(x) for x in args], N, table_size,
                                   table_max_td, C.byref(tx0), C.byref(tif), _dp(td),
                                   C.byref(cx0), C.byref(cif), _dp(ch)):
        raise AfError("transport_tables failed")
    return {"n": table_size, "td_x_min": tx0.value, "td_inv_fac": tif.value, "td": td,
            "chem_x_min": cx0.value, "chem_inv_fac": cif.value, "chem": ch}


class DeviceArray:
    """A float64 array in device memory, allocated by the device library's
    own HIP runtime (afgpu_device_alloc); data_ptr() is its address."""

    def __init__(self, shape, host=None):
        self.lib = load_gpu()
        self.shape = tuple(np.atleast_1d(shape)) if host is None else np.shape(host)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.p = C.c_void_p()
        if self.lib.afgpu_device_alloc(self.nbytes, C.byref(self.p)) != OK:
            raise AfError("afgpu_device_alloc failed: " + last_error("gpu"))
        if host is not None:
            self.upload(host)

    def data_ptr(self):
        return self.p.value

    def upload(self, host):
        h = np.ascontiguousarray(host, np.float64)
        assert h.nbytes == self.nbytes
        if self.lib.afgpu_device_copy(self.p, h.ctypes.data_as(C.c_void_p), self.nbytes, 1) != OK:
            raise AfError("afgpu_device_copy failed: " + last_error("gpu"))

    def numpy(self):
        out = np.empty(self.shape)
        if self.lib.afgpu_device_copy(out.ctypes.data_as(C.c_void_p), self.p, self.nbytes, 0) != OK:
            raise AfError("afgpu_device_copy failed: " + last_error("gpu"))
        return out

    def __len__(self):
        return self.shape[0]

    def free(self):
        if getattr(self, "p", None) and self.p.value:
            self.lib.afgpu_device_free(self.p)
            self.p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def last_error(kind):
    lib = load_gpu() if kind == "gpu" else load_oracle()
    msg = (lib.afgpu_last_error if kind == "gpu" else lib.afcpu_last_error)()
    return msg.decode() if msg else ""


def format_regression_line(out_cnt, vals):
    """One regression-log line in the reference's "(I0, nE20.8e3)" format
    (host code of the device library; no GPU use)."""
    lib = load_gpu()
    vals = np.ascontiguousarray(vals, np.float64)
    buf = C.create_string_buffer(64 + 24 * len(vals))
    n = lib.afgpu_format_regression_line(out_cnt, len(vals), _dp(vals), buf, len(buf))
    if n < 0:
        raise AfError("afgpu_format_regression_line failed (%d)" % n)
    return buf.value.decode()


def coarse_factors(ndim, nc, coord, r_min, dr, bc_dirichlet):
    lib = load_gpu()
    r = np.ascontiguousarray(r_min, np.float64)
    d = np.ascontiguousarray(dr, np.float64)
    bc = np.ascontiguousarray(bc_dirichlet, np.int32)
    vinv = np.zeros(ndim * nc * nc)
    v = np.zeros(ndim * nc * nc)
    eig = np.zeros(ndim * nc)
    rc = lib.afgpu_coarse_factors(ndim, nc, coord, _dp(r), _dp(d), _ip(bc), _dp(vinv), _dp(v), _dp(eig))
    if rc:
        raise AfError("afgpu_coarse_factors failed (%d)" % rc)
    return vinv, v, eig


# ---- one handle of the device path or the oracle --------------------------
class Path:
    """A handle of the device library (kind='gpu') or of the oracle."""

    def __init__(self, kind, ndim, nc, coord, precision=PREC_FP64):
        self.kind = kind
        self.lib = load_gpu() if kind == "gpu" else load_oracle()
        self.pre = "afgpu_" if kind == "gpu" else "afcpu_"
        self.D, self.nc, self.coord = ndim, nc, coord
        self.ncell = (nc + 2) ** ndim
        self.nface = ndim * (nc + 1) ** ndim
        self.h = _P()
        self._call("create", ndim, nc, coord, precision, C.byref(self.h), with_handle=False)
        self.n_boxes = 0

    def _fn(self, name):
        return getattr(self.lib, self.pre + name)

    def _call(self, name, *args, with_handle=True):
        f = self._fn(name)
        rc = f(self.h, *args) if with_handle else f(*args)
        if rc != OK:
            msg = self._fn("last_error")()
            raise AfError("%s%s failed (%d): %s" % (self.pre, name, rc, msg.decode() if msg else ""))
        return rc

    def close(self):
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # tree / data -----------------------------------------------------------
    def sync_tree(self, t):
        self._call("sync_tree", t["n_boxes"], _ip(t["lvl"]), _ip(t["ix"]), _ip(t["parent"]),
                   _ip(t["children"]), _ip(t["neighbors"]), _ip(t["nbmat"]),
                   _dp(t["r_min"]), _dp(t["dr"]))
        self.n_boxes = t["n_boxes"]
        self.tree = t

    def _ids(self, ids):
        if ids is None:
            ids = np.arange(1, self.n_boxes + 1, dtype=np.int32)
        return np.ascontiguousarray(ids, dtype=np.int32)

    def put_cc(self, iv, data, ids=None):
        ids = self._ids(ids)
        data = np.ascontiguousarray(data, np.float64).reshape(len(ids), self.ncell)
        self._call("put_cc", iv, len(ids), _ip(ids), _dp(data))

    def get_cc(self, iv, ids=None):
        ids = self._ids(ids)
        out = np.zeros((len(ids), self.ncell))
        self._call("get_cc", iv, len(ids), _ip(ids), _dp(out))
        return out

    def put_fc(self, ifc, data, ids=None):
        ids = self._ids(ids)
        data = np.ascontiguousarray(data, np.float64).reshape(len(ids), self.nface)
        self._call("put_fc", ifc, len(ids), _ip(ids), _dp(data))

    def get_fc(self, ifc, ids=None):
        ids = self._ids(ids)
        out = np.zeros((len(ids), self.nface))
        self._call("get_fc", ifc, len(ids), _ip(ids), _dp(out))
        return out

    def set_params(self, p):
        self.params = p
        self._call("set_params", C.byref(p))

    def set_tables(self, tb):
        self._call("set_tables", tb["n"], tb["td_x_min"], tb["td_inv_fac"], _dp(tb["td"]),
                   tb["chem_x_min"], tb["chem_inv_fac"], _dp(tb["chem"]))

    def set_voltage(self, v):
        self._call("set_voltage", v)

    def set_coarse_factors(self, op, vinv, v, eig):
        """Oracle only: supply the exact level-1 solve factors."""
        self._call("set_coarse_factors", op, _dp(vinv), _dp(v), _dp(eig))

    # reference operator API -------------------------------------------------
    def mg_fas_vcycle(self, op, set_residual=True, max_lvl=0):
        self._call("mg_fas_vcycle", op, int(set_residual), max_lvl)

    def mg_fas_fmg(self, op, set_residual=True, have_guess=True):
        self._call("mg_fas_fmg", op, int(set_residual), int(have_guess))

    def tree_maxabs_cc(self, iv):
        out = C.c_double()
        self._call("tree_maxabs_cc", iv, C.byref(out))
        return out.value

    def gc_tree(self, iv):
        self._call("gc_tree", iv)

    def restrict_tree(self, iv):
        self._call("restrict_tree", iv)

    def field_compute(self, s_in=0, have_guess=True):
        n, r = C.c_int(), C.c_double()
        self._call("field_compute", s_in, int(have_guess), C.byref(n), C.byref(r))
        return n.value, r.value

    def field_from_potential(self):
        self._call("field_from_potential")

    def forward_euler(self, dt, s_deriv, s_prev, w_prev, s_out, i_step, n_steps):
        sp = np.ascontiguousarray(s_prev, np.int32)
        wp = np.ascontiguousarray(w_prev, np.float64)
        lim = C.c_double()
        self._call("forward_euler", dt, s_deriv, len(sp), _ip(sp), _dp(wp), s_out, i_step,
                   n_steps, C.byref(lim))
        return lim.value

    def advance(self, dt):
        lim = C.c_double()
        self._call("advance", dt, C.byref(lim))
        return lim.value

    def photoi_set_src(self):
        self._call("photoi_set_src")

    def copy_current_state(self):
        self._call("copy_current_state")

    def restore_previous_state(self):
        self._call("restore_previous_state")

    def step(self):
        info = StepInfoT()
        self._call("step", C.byref(info))
        return info.as_dict()

    def refine_flags(self, settings, n_boxes):
        """The reference's box refinement flags (af_adjust_refinement with
        default_refinement) for box ids 1..n_boxes."""
        out = np.zeros(n_boxes, np.int32)
        self._call("refine_flags", C.byref(settings), n_boxes, _ip(out))
        return out

    def restrict_boxes(self, ids):
        """auto_restrict: the children of the listed boxes into them."""
        ids = np.ascontiguousarray(ids, np.int32)
        self._call("restrict_boxes", len(ids), _ip(ids))

    def prolong_boxes(self, ids):
        """auto_prolong: the listed new boxes from their parents, then their
        ghost cells."""
        ids = np.ascontiguousarray(ids, np.int32)
        self._call("prolong_boxes", len(ids), _ip(ids))

    # diagnostics and box I/O ---------------------------------------------------
    def regression_values(self, ivs):
        """The statistics of output_regression_log for the variables ivs:
        array (runs, len(ivs), 3) of sum/volume, sum of squares/volume and
        maximum over the leaves (one run for a plain tree)."""
        ivs = np.ascontiguousarray(ivs, np.int32)
        runs = self.num_runs() if self.kind == "gpu" else 1
        out = np.zeros((runs, len(ivs), 3))
        self._call("regression_values", len(ivs), _ip(ivs), _dp(out))
        return out

    def write_regression_log(self, filename, out_cnt, ivs, names):
        """output_regression_log: the header at out_cnt 0, then one line."""
        ivs = np.ascontiguousarray(ivs, np.int32)
        arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
        self._call("write_regression_log", os.fsencode(filename), out_cnt, len(ivs), _ip(ivs), arr)

    def set_chemistry(self, chem, charges=None):
        """The step's source terms from a reaction list (a chemistry.Chem of
        the same kind; species 0 the electrons, charges per species), or
        the standard model (None).  Keep chem alive while it is set."""
        self._chem = chem
        q = None if charges is None else np.ascontiguousarray(charges, np.int32)
        self._call("set_chemistry", chem.h if chem is not None else None, _ip(q) if q is not None else None)

    def put_species(self, k, data, state=0, ids=None):
        ids = self._ids(ids)
        data = np.ascontiguousarray(data, np.float64).reshape(len(ids), self.ncell)
        self._call("put_species", k, state, len(ids), _ip(ids), _dp(data))

    def get_species(self, k, state=0, ids=None):
        ids = self._ids(ids)
        out = np.zeros((len(ids), self.ncell))
        self._call("get_species", k, state, len(ids), _ip(ids), _dp(out))
        return out

    def box_words(self, ivs):
        """Words of one box's data for the variable list ivs (a face
        variable AFGPU_VAR_FACE(ifc) takes nface, any other ncell)."""
        return sum(self.nface if 200 <= iv < 200 + 2 else self.ncell for iv in ivs)

    def get_boxes(self, ivs, ids=None):
        """Whole boxes of the variables ivs in af_write_tree's order: array
        (len(ids), len(ivs), ncell) for cell variables only, else
        (len(ids), box_words(ivs))."""
        ivs = np.ascontiguousarray(ivs, np.int32)
        ids = self._ids(ids)
        cells = self.box_words(ivs) == len(ivs) * self.ncell
        out = np.zeros((len(ids), len(ivs), self.ncell) if cells else (len(ids), self.box_words(ivs)))
        self._call("get_boxes", len(ivs), _ip(ivs), len(ids), _ip(ids), _dp(out))
        return out

    def put_boxes(self, ivs, data, ids=None):
        ivs = np.ascontiguousarray(ivs, np.int32)
        ids = self._ids(ids)
        data = np.ascontiguousarray(data, np.float64).reshape(len(ids), self.box_words(ivs))
        self._call("put_boxes", len(ivs), _ip(ivs), len(ids), _ip(ids), _dp(data))

    def get_time_state(self):
        it, t, dt = C.c_int(), C.c_double(), C.c_double()
        self._call("get_time_state", C.byref(it), C.byref(t), C.byref(dt))
        return it.value, t.value, dt.value

    def set_time_state(self, it, time, global_dt):
        self._call("set_time_state", it, time, global_dt)

    def limit_step_dt(self, dt_cap):
        """The next step's dt := min(dt, dt_cap); global_dt unchanged (the
        time loop's output-time clipping)."""
        self._call("limit_step_dt", dt_cap)

    def set_option(self, option, value):
        """Device library only (e.g. OPT_GRAPHS)."""
        self._call("set_option", option, int(value))

    def num_leaf_cells(self):
        return self._fn("num_leaf_cells")(self.h)

    # ensembles (device library only) ------------------------------------------
    def num_runs(self):
        return self._fn("num_runs")(self.h)

    def set_run_voltages(self, volts):
        v = np.ascontiguousarray(volts, np.float64)
        self._call("set_run_voltages", len(v), _dp(v))

    def field_compute_runs(self, s_in=0, have_guess=True):
        R = self.num_runs()
        n = np.zeros(R, np.int32)
        r = np.zeros(R)
        self._call("field_compute_runs", s_in, int(have_guess), R, _ip(n), _dp(r))
        return n, r

    def get_run_time_state(self, run):
        it, t, dt = C.c_int(), C.c_double(), C.c_double()
        self._call("get_run_time_state", run, C.byref(it), C.byref(t), C.byref(dt))
        return it.value, t.value, dt.value

    def set_run_time_state(self, run, it, time, global_dt):
        self._call("set_run_time_state", run, it, time, global_dt)

    def autotune(self):
        """Pick the work-decomposition shape of every kernel class by
        measurement; returns {class: (boxes per workgroup, workgroups per
        CU, ms)}."""
        arr = (ShapeT * N_KERNEL_CLASSES)()
        self._call("autotune", N_KERNEL_CLASSES, arr)
        return {KERNEL_CLASSES[k]: (arr[k].boxes_per_workgroup, arr[k].workgroups_per_cu, arr[k].ms)
                for k in range(N_KERNEL_CLASSES)}

    def get_shapes(self):
        arr = (ShapeT * N_KERNEL_CLASSES)()
        self._call("get_shapes", N_KERNEL_CLASSES, arr)
        return [(arr[k].boxes_per_workgroup, arr[k].workgroups_per_cu) for k in range(N_KERNEL_CLASSES)]

    def set_shapes(self, shapes):
        arr = (ShapeT * N_KERNEL_CLASSES)()
        for k, (bb, pc) in enumerate(shapes):
            arr[k].boxes_per_workgroup, arr[k].workgroups_per_cu = bb, pc
        self._call("set_shapes", len(shapes), arr)

    def step_ensemble(self):
        R = self.num_runs()
        infos = (StepInfoT * R)()
        self._call("step_ensemble", R, infos)
        return [infos[r].as_dict() for r in range(R)]
