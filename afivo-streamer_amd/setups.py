"""Configurations of the streamer hot path (trees, initial conditions,
tables, settings), built on the host the way the reference's driver builds
them before its time loop (src/streamer.f90:446-539).

``Case`` bundles everything a ``afgpu.Path`` needs; ``apply`` loads a case
into a handle of the device library or of the oracle.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

import afgpu as A

K_B = 1.380649e-23  # src/m_units_constants.f90:39


def gas_number_density(p_bar=1.0, T=300.0):
    """Ideal gas law, src/m_gas.f90:174-176."""
    return 1e5 * p_bar / (K_B * T)


def analytic_transport(n=1000, E_max=3e7):
    """Synthetic old-style transport data: constant mobility 0.03 m^2/Vs and
    diffusion 0.1 m^2/s, alpha(E) = 4.3e5 exp(-1.9e7/|E|) 1/m (0 at E = 0),
    no attachment, tabulated at n uniformly spaced |E| in [0, E_max]."""
    E = np.linspace(0.0, E_max, n)
    mu = np.full(n, 0.03)
    dif = np.full(n, 0.1)
    alpha = np.zeros(n)
    alpha[1:] = 4.3e5 * np.exp(-1.9e7 / E[1:])
    eta = np.zeros(n)
    return E, mu, dif, alpha, eta


def default_params(D, domain_len_last, N, photoi=False, integrator=A.FORWARD_EULER,
                   num_vcycles=2, rel_residual=1e-4):
    p = A.ParamsT()
    p.gas_number_density = N
    p.time_integrator = integrator
    p.dt_max = 1e-11
    p.dt_min = 1e-14
    p.dt_safety_factor = 0.9
    p.dt_cfl_number = 0.5
    p.dt_max_growth_factor = 2.0
    p.dt_chemistry_limit_loss = 1
    p.dt_chemistry_nmin = -1.0
    p.mg_num_vcycles = num_vcycles
    p.mg_max_rel_residual = rel_residual
    p.mg_n_cycle_down = 2
    p.mg_n_cycle_up = 2
    p.domain_len_last = domain_len_last
    p.species_bc = 0
    p.photoi_enabled = int(photoi)
    p.photoi_per_steps = 1
    p.photoi_species = 1  # M+
    p.photoi_eta_quench = 0.075
    # Helmholtz modes, custom author: lambda*p_O2 per (m Torr) with p_O2 = 150 Torr
    lam = np.array([0.0553, 0.1460, 0.89]) * 150.0 * 100.0  # 1/cm -> 1/m
    coeff = np.array([0.0021, 0.1775, 0.3324]) * (150.0 * 100.0) ** 2
    p.photoi_n_modes = 3
    for i in range(3):
        p.photoi_lambda[i] = lam[i]
        p.photoi_coeff[i] = coeff[i]
    p.photoi_max_rel_residual = 1e-2
    p.photoi_max_fmg_cycles = 10
    return p


@dataclass
class Case:
    D: int
    nc: int
    coord: int
    tree: A.Tree
    tdata: dict
    ne0: np.ndarray            # initial electron (= ion) density, all boxes
    params: A.ParamsT
    voltage: float
    name: str = ""
    extra: dict = field(default_factory=dict)


def make_case(D, nc, coord, domain_len, voltage, uniform_lvl, seed_r0, seed_width=0.4e-3,
              seed_density=1e20, background=1e13, noise=0.1, rng_seed=20260101,
              refine_fn=None, photoi=False, num_vcycles=2, rel_residual=1e-4,
              integrator=A.FORWARD_EULER, name=""):
    """Build a tree refined uniformly to ``uniform_lvl``, then refined further
    by ``refine_fn(tree) -> ids`` until it returns no ids; initial
    conditions as init_cond_set_box with seeded noise."""
    tree = A.Tree(D, nc, coord, np.zeros(D), np.asarray(domain_len, float))
    tree.refine_uniform(uniform_lvl)
    if refine_fn is not None:
        while True:
            ids = refine_fn(tree)
            if len(ids) == 0:
                break
            tree.refine(ids)
    t = tree.export()
    ids = np.arange(1, t["n_boxes"] + 1, dtype=np.int32)
    ne0 = tree.init_cond(ids, background, noise, rng_seed, seed_density, seed_width, seed_r0)
    N = gas_number_density()
    E, mu, dif, alpha, eta = analytic_transport()
    tdata = A.transport_tables(E, mu, dif, alpha, eta, N)
    params = default_params(D, domain_len[-1], N, photoi=photoi, integrator=integrator,
                            num_vcycles=num_vcycles, rel_residual=rel_residual)
    return Case(D, nc, coord, tree, tdata, ne0, params, voltage, name=name, extra={"export": t})


def apply(path, case):
    """Load a case into a Path (device library or oracle)."""
    t = case.extra["export"]
    path.sync_tree(t)
    path.set_params(case.params)
    path.set_tables(case.tdata)
    path.set_voltage(case.voltage)
    for iv in (A.CC_ELECTRON, A.CC_POS_ION):
        path.put_cc(iv, case.ne0)
    if path.kind != "gpu":
        # The oracle receives the exact level-1 solve factors; the device
        # library computes the same factors itself.
        lvl = t["lvl"]
        b1 = int(np.nonzero(lvl == 1)[0][0])
        D = case.D
        bc = np.zeros(2 * D, np.int32)
        bc[2 * (D - 1)] = 1
        bc[2 * (D - 1) + 1] = 1
        f = A.coarse_factors(D, case.nc, case.coord, t["r_min"][b1 * D:(b1 + 1) * D],
                             t["dr"][b1 * D:(b1 + 1) * D], bc)
        for op in range(1 + A.MAX_MODES):
            path.set_coarse_factors(op, *f)
    return path


# ---- named configurations --------------------------------------------------
def config0(uniform_lvl=6, voltage=2e4, **kw):
    """2D axisymmetric streamer, 10 mm x 10 mm, 8x8 boxes, uniform level
    ``uniform_lvl`` (6 -> 256^2 cells), 20 kV, seed on the axis at z = 8 mm."""
    return make_case(2, 8, A.CYL, [10e-3, 10e-3], voltage, uniform_lvl, [0.0, 8e-3],
                     name="cyl2d_uniform_lvl%d" % uniform_lvl, **kw)


def channel_refine(max_lvl, radius, axis_xy):
    """Refine leaves below ``max_lvl`` whose centre lies within ``radius`` of
    the vertical line through ``axis_xy`` (a streamer-channel region)."""
    def fn(tree):
        t = tree.export()
        D = tree.D
        out = []
        for b in range(t["n_boxes"]):
            if t["lvl"][b] >= max_lvl or t["children"][b * 2 ** D] > 0:
                continue
            rmin = t["r_min"][b * D:(b + 1) * D]
            dr = t["dr"][b * D:(b + 1) * D]
            ctr = rmin + 0.5 * dr * tree.nc
            if np.hypot(ctr[0] - axis_xy[0], ctr[1] - axis_xy[1]) <= radius:
                out.append(b + 1)
        return np.array(out, np.int32)
    return fn


def config1(uniform_lvl=4, max_lvl=5, radius=5.2e-3, nc=16, voltage=1.5e4, **kw):
    """3D Cartesian AMR, 16 mm cube, 16^3 boxes: uniform to ``uniform_lvl``,
    refined to ``max_lvl`` in a cylinder of ``radius`` around the vertical
    axis through the seed (1856 leaf boxes, 7.6M cells by default), 15 kV,
    Gaussian seed near the top."""
    L = 16e-3
    return make_case(3, nc, A.XYZ, [L, L, L], voltage, uniform_lvl, [L / 2, L / 2, 0.8 * L],
                     refine_fn=channel_refine(max_lvl, radius, (L / 2, L / 2)),
                     name="xyz3d_amr_l%d_r%.1fmm" % (max_lvl, radius * 1e3), **kw)


def random_refine(max_lvl, frac, seed):
    """Refine a seeded random fraction of leaves below ``max_lvl`` (once per
    level), to exercise every kind of refinement boundary."""
    rng = np.random.default_rng(seed)
    done = set()

    def fn(tree):
        t = tree.export()
        D = tree.D
        for lvl in range(1, max_lvl):
            if lvl in done:
                continue
            leaves = [b + 1 for b in range(t["n_boxes"])
                      if t["lvl"][b] == lvl and t["children"][b * 2 ** D] == 0]
            done.add(lvl)
            if not leaves:
                continue
            k = max(1, int(round(frac * len(leaves))))
            return np.array(sorted(rng.choice(leaves, size=k, replace=False)), np.int32)
        return np.array([], np.int32)
    return fn


def small_amr(D, nc=8, coord=None, uniform_lvl=2, max_lvl=4, frac=0.4, seed=7, voltage=1e4, **kw):
    """Small randomly refined tree for parity tests."""
    if coord is None:
        coord = A.XYZ
    L = [4e-3] * D
    r0 = [L[0] / 2] * (D - 1) + [0.7 * L[-1]]
    if coord == A.CYL:
        r0[0] = 0.0
    return make_case(D, nc, coord, L, voltage, uniform_lvl, r0, seed_width=0.3e-3,
                     refine_fn=random_refine(max_lvl, frac, seed),
                     name="amr%dd_%s_l%d_s%d" % (D, "cyl" if coord == A.CYL else "xyz", max_lvl, seed),
                     **kw)


# ---- ensembles ----------------------------------------------------------------
def merge_forest(cases):
    """One forest of independent runs from single-run cases of equal
    dimension, box size, geometry and level-1 box: the box ids of case r
    follow those of case r-1 (tree references offset accordingly), so the
    r-th level-1 box roots run r.  Returns (export, ne0, box offsets)."""
    D, nc, coord = cases[0].D, cases[0].nc, cases[0].coord
    keys = ("lvl", "ix", "parent", "children", "neighbors", "nbmat", "r_min", "dr")
    parts = {k: [] for k in keys}
    ne0, offs = [], []
    off = 0
    for c in cases:
        if (c.D, c.nc, c.coord) != (D, nc, coord):
            raise ValueError("runs of a forest must share dimension, box size and geometry")
        t = c.extra["export"]
        offs.append(off)
        for k in keys:
            a = np.array(t[k], copy=True)
            if k in ("parent", "children", "neighbors", "nbmat"):
                a[a > 0] += off  # 0: none / refinement boundary, < 0: physical boundary
            parts[k].append(a)
        ne0.append(c.ne0)
        off += t["n_boxes"]
    t = {k: np.ascontiguousarray(np.concatenate(parts[k])) for k in keys}
    t["n_boxes"] = off
    return t, np.concatenate(ne0), np.array(offs)


@dataclass
class Ensemble:
    cases: list               # the single-run cases, in run order
    export: dict              # the merged forest
    ne0: np.ndarray
    offsets: np.ndarray       # first box id - 1 of every run
    voltages: np.ndarray
    name: str = ""

    def run_ids(self, r):
        n = self.cases[r].extra["export"]["n_boxes"]
        return np.arange(self.offsets[r] + 1, self.offsets[r] + n + 1, dtype=np.int32)


def ensemble(n_runs=64, levels=(5, 6, 7, 8), v_range=(15e3, 25e3), z_range=(6e-3, 9e-3),
             seed=20260101, num_vcycles=100, rel_residual=1e-8, level_list=None, **kw):
    """Independent 2D axisymmetric runs of the config-0 physics (10 mm x
    10 mm, 8^2 boxes) with uneven sizes: each run's uniform level is drawn
    by seed from ``levels`` (5..8 -> 128^2 .. 1024^2 cells), its voltage
    uniformly from ``v_range`` and its seed height uniformly from
    ``z_range``; ``level_list`` fixes the levels instead.  Each run has its
    own noise seed."""
    rng = np.random.default_rng(seed)
    if level_list is None:
        level_list = rng.choice(np.asarray(levels), size=n_runs)
    else:
        level_list = np.asarray(level_list)
        n_runs = len(level_list)
    volts = rng.uniform(v_range[0], v_range[1], size=n_runs)
    zs = rng.uniform(z_range[0], z_range[1], size=n_runs)
    cases = []
    for r in range(n_runs):
        cases.append(make_case(2, 8, A.CYL, [10e-3, 10e-3], float(volts[r]), int(level_list[r]),
                               [0.0, float(zs[r])], rng_seed=seed + 1 + r, num_vcycles=num_vcycles,
                               rel_residual=rel_residual,
                               name="cyl2d_lvl%d_run%d" % (level_list[r], r), **kw))
    t, ne0, offs = merge_forest(cases)
    name = "ensemble_%druns_%dcells" % (n_runs, sum(c.tree.info()["leaf_cells"] for c in cases))
    return Ensemble(cases, t, ne0, offs, volts, name=name)


def apply_ensemble(path, ens):
    """Load an ensemble into a device-library handle (one forest)."""
    c0 = ens.cases[0]
    path.sync_tree(ens.export)
    path.set_params(c0.params)
    path.set_tables(c0.tdata)
    path.set_run_voltages(ens.voltages)
    for iv in (A.CC_ELECTRON, A.CC_POS_ION):
        path.put_cc(iv, ens.ne0)
    return path
