"""A host-side driver of the C-ABI that runs afivo-streamer's time loop the
way its regression tests run it, for the device library or the oracle
(the same calls through afgpu.Path): set_initial_conditions
(src/streamer.f90:477-520), then the loop of src/streamer.f90:181-432 with
output-time dt clipping, step rejection, the field solve, the regression
log at every output time and refinement every refine_per_steps steps
(restriction and ghost cells of the densities, the refinement flags on the
path, the tree changed on the host, the new boxes prolonged on the path,
the field and photoionization recomputed).

The settings are those of the reference's rtest configuration files
(programs/standard_2d/tests/test_cyl.cfg, test_2d.cfg, test_2d_photoi.cfg,
test_cyl_chem.cfg, test_cyl_photoi_chem.cfg, test_2d_photoi_chem.cfg,
programs/standard_3d/tests/test_3d.cfg, test_3d_chem.cfg,
test_3d_photoi_chem.cfg) with the reference's defaults for everything they
leave out; the input data are their td_air_siglo_swarm.txt (old-style
transport data), air_chemistry_v1.txt and air_chemistry_small_v1.txt
(reaction lists with their transport data; tests/golden)."""
import os

import numpy as np

import afgpu as A
import chemistry as CH
import setups as S

HERE = os.path.dirname(os.path.abspath(__file__))
TD_FILE = os.path.join(HERE, "golden", "td_air_siglo_swarm.txt")
UC_boltzmann_const = 1.380649e-23  # src/m_units_constants.f90


def seed(rel_r0, rel_r1, L, density, width, falloff, charge_type):
    return {"r0": np.asarray(rel_r0) * np.asarray(L), "r1": np.asarray(rel_r1) * np.asarray(L),
            "density": density, "width": width, "falloff": falloff, "charge_type": charge_type}


# the configuration files' settings (domain from the origin 0)
CASES = {
    "test_cyl": dict(   # programs/standard_2d/tests/test_cyl.cfg
        D=2, coord=A.CYL, nc=8, L=(16e-3, 16e-3), end_time=1.5e-9, output_dt=0.25e-9, dt_max=1e-10,
        field=-2e6, background=1e14, refine_adx=1.0, photoi=None,
        seeds=[seed((0.0, 0.52), (0.0, 0.48), (16e-3, 16e-3), 5e18, 2e-4, "smoothstep", 1)],
        log="programs/standard_2d/tests/test_cyl_rtest.log"),
    "test_2d": dict(   # programs/standard_2d/tests/test_2d.cfg
        D=2, coord=A.XYZ, nc=8, L=(32e-3, 32e-3), end_time=7e-9, output_dt=1e-9, dt_max=1e-10,
        field=2.5e6, background=1e14, refine_adx=1.0, photoi=None,
        seeds=[seed((0.5, 0.32), (0.5, 0.28), (32e-3, 32e-3), 1e18, 2e-4, "smoothstep", 1)],
        log="programs/standard_2d/tests/test_2d_rtest.log"),
    "test_cyl_chem": dict(   # programs/standard_2d/tests/test_cyl_chem.cfg
        D=2, coord=A.CYL, nc=8, L=(16e-3, 16e-3), end_time=1.5e-9, output_dt=0.25e-9, dt_max=1e-10,
        field=-2e6, background=1e14, refine_adx=1.0, photoi=None, chem="air_chemistry_v1.txt",
        seeds=[seed((0.0, 0.52), (0.0, 0.48), (16e-3, 16e-3), 5e18, 2e-4, "smoothstep", 1)],
        log="programs/standard_2d/tests/test_cyl_chem_rtest.log"),
    "test_cyl_photoi_chem": dict(   # programs/standard_2d/tests/test_cyl_photoi_chem.cfg
        D=2, coord=A.CYL, nc=8, L=(16e-3, 16e-3), end_time=1.5e-9, output_dt=0.25e-9, dt_max=1e-10,
        field=-2e6, background=1e14, refine_adx=1.0, photoi="Bourdon-3", chem="air_chemistry_v1.txt",
        seeds=[seed((0.0, 0.52), (0.0, 0.48), (16e-3, 16e-3), 5e18, 2e-4, "smoothstep", 1)],
        log="programs/standard_2d/tests/test_cyl_photoi_chem_rtest.log"),
    "test_2d_photoi_chem": dict(   # programs/standard_2d/tests/test_2d_photoi_chem.cfg
        D=2, coord=A.XYZ, nc=8, L=(32e-3, 32e-3), end_time=7e-9, output_dt=1e-9, dt_max=1e-10,
        field=2.5e6, background=1e11, refine_adx=1.0, photoi="Bourdon-3", chem="air_chemistry_v1.txt",
        seeds=[seed((0.5, 0.32), (0.5, 0.28), (32e-3, 32e-3), 1e18, 2e-4, "smoothstep", 1)],
        log="programs/standard_2d/tests/test_2d_photoi_chem_rtest.log"),
    "test_2d_photoi": dict(   # programs/standard_2d/tests/test_2d_photoi.cfg
        D=2, coord=A.XYZ, nc=8, L=(32e-3, 32e-3), end_time=7e-9, output_dt=1e-9, dt_max=1e-10,
        field=2.5e6, background=1e11, refine_adx=1.0, photoi="Bourdon-3",
        seeds=[seed((0.5, 0.32), (0.5, 0.28), (32e-3, 32e-3), 1e18, 2e-4, "smoothstep", 1)],
        log="programs/standard_2d/tests/test_2d_photoi_rtest.log"),
    "test_3d": dict(   # programs/standard_3d/tests/test_3d.cfg
        D=3, coord=A.XYZ, nc=8, L=(16e-3, 16e-3, 16e-3), end_time=3e-9, output_dt=0.2e-9, dt_max=5e-11,
        field=-2.5e6, background=1e14, refine_adx=1.5, photoi=None,
        seeds=[seed((0.5, 0.5, 0.52), (0.5, 0.5, 0.48), (16e-3, 16e-3, 16e-3), 5e18, 5e-4, "smoothstep", 0)],
        log="programs/standard_3d/tests/test_3d_rtest.log"),
    "test_3d_chem": dict(   # programs/standard_3d/tests/test_3d_chem.cfg
        D=3, coord=A.XYZ, nc=8, L=(16e-3, 16e-3, 16e-3), end_time=3e-9, output_dt=0.2e-9, dt_max=5e-11,
        field=-2.5e6, background=1e14, refine_adx=1.5, photoi=None, chem="air_chemistry_small_v1.txt",
        seeds=[seed((0.5, 0.5, 0.52), (0.5, 0.5, 0.48), (16e-3, 16e-3, 16e-3), 5e18, 5e-4, "smoothstep", 0)],
        log="programs/standard_3d/tests/test_3d_chem_rtest.log"),
    "test_3d_photoi_chem": dict(   # programs/standard_3d/tests/test_3d_photoi_chem.cfg
        D=3, coord=A.XYZ, nc=8, L=(16e-3, 16e-3, 16e-3), end_time=3e-9, output_dt=0.2e-9, dt_max=5e-11,
        field=-2.5e6, background=1e14, refine_adx=1.5, photoi="Bourdon-3", chem="air_chemistry_small_v1.txt",
        seeds=[seed((0.5, 0.5, 0.52), (0.5, 0.5, 0.48), (16e-3, 16e-3, 16e-3), 5e18, 5e-4, "smoothstep", 0)],
        log="programs/standard_3d/tests/test_3d_photoi_chem_rtest.log"),
}

REFINE_MAX_DX = 1e-3     # src/m_refine.f90:19
REFINE_PER_STEPS = 2     # src/m_refine.f90:13
PHOTOI_PER_STEPS = 5     # src/m_photoi.f90:53
GAS_PRESSURE = 1.0       # bar (the files' gas%pressure)
FRAC_O2 = 0.2            # gas%fractions (the default N2:O2 of 0.8:0.2)


def transport_tables(N):
    """td_tbl from the old-style file as transport_data_initialize forms it
    (src/m_transport_data.f90:85-126)."""
    lines = open(TD_FILE).read().splitlines()
    xs, ys = [], []
    for name in ("efield[V/m]_vs_mu[m2/Vs]", "efield[V/m]_vs_dif[m2/s]", "efield[V/m]_vs_alpha[1/m]",
                 "efield[V/m]_vs_eta[1/m]"):
        x, y = CH.table_from_file(lines, name)
        xs.append(x)
        ys.append(y)
    assert all(np.array_equal(x, xs[0]) for x in xs)
    return A.transport_tables(xs[0], *ys, N)


def transport_tables_new(path, N, table_size=1000):
    """td_tbl from a new-style input file (src/m_transport_data.f90:128-162):
    reduced field in Td, mobility and diffusion times N, alpha and eta over
    N, on LT_create's grid from 0 to the mobility data's last field.  The
    standard model's rate columns (unused with a reaction list) are formed
    as for old-style data.  Returns the tables and the grid's last field."""
    lines = open(path).read().splitlines()
    cols = [CH.table_from_file(lines, name) for name in
            ("Mobility *N (1/m/V/s)", "Diffusion coefficient *N (1/m/s)", "Townsend ioniz. coef. alpha/N (m2)",
             "Townsend attach. coef. eta/N (m2)")]
    x_max = cols[0][0][-1]
    grid = CH.lt_grid(0.0, x_max, table_size)
    td = [CH.spaced_data(np.asarray(x), np.asarray(y), grid) for x, y in cols]
    ion = td[2] * td[0] * grid * 1e-21 * N
    att = td[3] * td[0] * grid * 1e-21 * N
    tables = {"n": table_size, "td_x_min": 0.0, "td_inv_fac": (table_size - 1) / (x_max - 0.0),
              "td": np.concatenate(td), "chem_x_min": grid[0],
              "chem_inv_fac": (table_size - 1) / (grid[-1] - grid[0]), "chem": np.concatenate([ion, att])}
    return tables, x_max


def falloff_value(dist, width, kind):
    """GM_smoothstep / GM_gaussian (src/m_geometry.f90:84-123)."""
    if kind == "smoothstep":
        t = dist / width - 1
        return np.where(dist < width, 1.0, np.where(dist < 2 * width, 1 - (3 * t ** 2 - 2 * t ** 3), 0.0))
    if kind == "gaussian":
        return np.exp(-(dist / width) ** 2)
    raise ValueError(kind)


def density_line(r, s):
    """GM_density_line (src/m_geometry.f90:58-82): the fall-off of the
    distance to the segment r0-r1, times the density."""
    r0, r1 = s["r0"], s["r1"]
    line_len2 = np.sum((r1 - r0) ** 2)
    frac = np.sum((r - r0) * (r1 - r0), axis=-1)
    lo, hi = frac <= 0, frac >= line_len2
    f = np.where(lo, 0.0, np.where(hi, 1.0, frac / line_len2))
    proj = r0 + (frac / line_len2)[..., None] * (r1 - r0)
    dv = np.where(lo[..., None], r - r0, np.where(hi[..., None], r - r1, r - proj))
    dist = np.sqrt(np.sum(dv ** 2, axis=-1))
    val = falloff_value(dist, s["width"], s["falloff"])
    return val * (f * s["density"] + (1 - f) * s["density"])


def initial_densities(cfg, t, ids):
    """init_cond_set_box (src/m_init_cond.f90:217-291) on every cell (ghost
    cells included) of the listed boxes: electrons and positive ions at the
    background density, plus the seeds by their charge type."""
    D, nc = cfg["D"], cfg["nc"]
    ng = nc + 2
    q = np.arange(ng)
    e = np.zeros((len(ids), ng ** D))
    ion = np.zeros((len(ids), ng ** D))
    for n, bid in enumerate(ids):
        b = bid - 1
        rmin, dr = t["r_min"][b * D:(b + 1) * D], t["dr"][b * D:(b + 1) * D]
        grids = np.meshgrid(*[rmin[d] + (q - 0.5) * dr[d] for d in range(D)], indexing="ij")
        # box data is i fastest: transpose the (i, j, k) grid to (k, j, i)
        r = np.stack([g.transpose() for g in grids], axis=-1).reshape(-1, D)
        ev = np.full(len(r), cfg["background"])
        iv = np.full(len(r), cfg["background"])
        for s in cfg["seeds"]:
            dens = density_line(r, s)
            if s["charge_type"] in (-1, 0):
                ev = ev + dens
            if s["charge_type"] in (0, 1):
                iv = iv + dens
        e[n], ion[n] = ev, iv
    return e, ion


def params_for(cfg, N):
    p = S.default_params(cfg["D"], cfg["L"][-1], N, photoi=cfg["photoi"] is not None,
                         integrator=A.HEUNS_METHOD, num_vcycles=2, rel_residual=1e-4)
    p.dt_max = cfg["dt_max"]
    p.dt_min = 1e-14
    if cfg["photoi"] == "Bourdon-3":
        # src/m_photoi_helmh.f90:108-121 with O2 fraction and pressure
        lam = np.array([4147.85, 10950.93, 66755.67]) * FRAC_O2 * GAS_PRESSURE
        coeff = np.array([1117314.935, 28692377.5, 2748842283.0]) * (FRAC_O2 * GAS_PRESSURE) ** 2
        p.photoi_n_modes = 3
        for i in range(3):
            p.photoi_lambda[i], p.photoi_coeff[i] = lam[i], coeff[i]
        # photoi%eta and the quenching factor p_q / (p + p_q) (src/m_photoi.f90:38-47, 175-184)
        p.photoi_eta_quench = 0.05 * (40e-3 / (GAS_PRESSURE + 40e-3))
        p.photoi_per_steps = PHOTOI_PER_STEPS
        p.photoi_species = 1  # M_plus
        p.photoi_max_rel_residual = 1e-2
        p.photoi_max_fmg_cycles = 10
    return p


class Run:
    """One regression-test run on a path kind ("gpu" or "oracle")."""

    DENS = (A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION)

    def __init__(self, name, kind, precision=A.PREC_FP64):
        self.cfg = cfg = CASES[name]
        D, nc = cfg["D"], cfg["nc"]
        self.N = 1e5 * GAS_PRESSURE / (UC_boltzmann_const * 300.0)  # src/m_gas.f90:174-176
        self.tree = A.Tree(D, nc, cfg["coord"], np.zeros(D), np.asarray(cfg["L"]))
        # refine up to the maximal grid spacing (src/streamer.f90:484-488)
        lvl = 1
        while cfg["L"][0] / nc / 2 ** (lvl - 1) > REFINE_MAX_DX:
            lvl += 1
        self.tree.refine_uniform(lvl)
        self.t = self.tree.export()
        self.p = A.Path(kind, D, nc, cfg["coord"], precision if kind == "gpu" else A.PREC_FP64)
        self.p.sync_tree(self.t)
        self.p.set_params(params_for(cfg, self.N))
        self.net = None
        if cfg.get("chem"):
            # a reaction list with its transport data (input_data%old_style = f)
            path = os.path.join(HERE, "golden", cfg["chem"])
            tables, x_max = transport_tables_new(path, self.N)
            self.p.set_tables(tables)
            self.net = CH.read_network(path, {"N2": 0.8, "O2": 0.2}, self.N, 0.0, x_max, 1000)
            # td_tbl's mean energy column (eV, src/m_transport_data.f90:158-161),
            # read by the rate types that use the electron temperature
            lines = open(path).read().splitlines()
            xe, ye = CH.table_from_file(lines, "Mean energy (eV)")
            energy = CH.spaced_data(np.asarray(xe), np.asarray(ye), CH.lt_grid(0.0, x_max, 1000))
            self.chem = CH.Chem(kind, len(self.net.species), self.net.reactions, 0.0, x_max, self.net.fld_cols,
                                energy, 300.0)
            self.p.set_chemistry(self.chem, self.net.charges)
            if cfg["photoi"] is not None:
                # photoi%species, default O2_plus (src/m_photoi.f90:27-28)
                prm = params_for(cfg, self.N)
                prm.photoi_species = list(self.net.names).index("O2_plus")
                self.p.set_params(prm)
            self.dens = [A.VAR_SPECIES(k) for k in range(len(self.net.species))]
            # the first positive ion takes the ions of the initial conditions
            # (src/m_streamer.f90:228-238)
            self.i_pos = list(self.net.charges).index(1)
        else:
            self.p.set_tables(transport_tables(self.N))
            self.dens = list(self.DENS)
        # field_given_by field / field_amplitude: V = -L * E (src/m_field.f90:135-161)
        self.p.set_voltage(-cfg["L"][-1] * cfg["field"])
        if kind != "gpu":
            lvl1 = int(np.nonzero(self.t["lvl"] == 1)[0][0])
            bc = np.zeros(2 * D, np.int32)
            bc[2 * (D - 1)] = bc[2 * (D - 1) + 1] = 1
            f = A.coarse_factors(D, nc, cfg["coord"], self.t["r_min"][lvl1 * D:(lvl1 + 1) * D],
                                 self.t["dr"][lvl1 * D:(lvl1 + 1) * D], bc)
            for op in range(1 + A.MAX_MODES):
                self.p.set_coarse_factors(op, *f)
        ids = np.arange(1, self.t["n_boxes"] + 1, dtype=np.int32)
        self.set_initial(ids)
        self.rows = []
        self.steps = 0

    def set_initial(self, ids):
        e, ion = initial_densities(self.cfg, self.t, ids)
        if self.net is not None:
            for k in range(len(self.net.species)):
                self.p.put_species(k, e if k == 0 else ion if k == self.i_pos else np.zeros_like(e), ids=ids)
            return
        self.p.put_cc(A.CC_ELECTRON, e, ids)
        self.p.put_cc(A.CC_POS_ION, ion, ids)
        self.p.put_cc(A.CC_NEG_ION, np.zeros_like(e), ids)

    def refine_settings(self, global_time):
        seeds = [(s["r0"], s["r1"], s["width"]) for s in self.cfg["seeds"]]
        return A.refine_settings(global_time=global_time, refine_adx=self.cfg["refine_adx"], seeds=seeds)

    def adjust_refinement(self, global_time):
        """af_adjust_refinement with default_refinement: the flags on the
        path, auto_restrict of the boxes that lose their children, the tree
        changed on the host, auto_prolong of the new boxes on the path.
        Returns the new box ids and the number of removed boxes."""
        n = self.t["n_boxes"]
        flags = self.p.refine_flags(self.refine_settings(global_time), n)
        deref = np.nonzero(flags == A.REF_DEREFINE)[0] + 1
        ref = np.nonzero(flags == A.REF_REFINE)[0] + 1
        if len(deref):
            self.p.restrict_boxes(deref)
            self.tree.derefine(deref)
        if len(ref):
            self.tree.refine(ref)
        if len(deref) == 0 and len(ref) == 0:
            return np.zeros(0, np.int32), 0
        self.t = self.tree.export()
        self.p.sync_tree(self.t)
        new = np.arange(n + 1, self.t["n_boxes"] + 1, dtype=np.int32)
        if len(new):
            self.p.prolong_boxes(new)
        return new, len(deref) * 2 ** self.cfg["D"]

    def log_row(self, out_cnt):
        it, time, dt = self.p.get_time_state()
        # (species, (mean, mean of squares, max)), four variables per call
        v = np.concatenate([self.p.regression_values(self.dens[q:q + 4])[0] for q in range(0, len(self.dens), 4)])
        self.rows.append([out_cnt, time, dt] + list(v[:, 0]) + list(v[:, 1]) + list(v[:, 2]))

    def initial_conditions(self):
        """set_initial_conditions: up to 100 rounds of the initial field and
        refinement, the initial conditions set on every new box."""
        for _ in range(100):
            self.p.field_compute(0, have_guess=False)
            new, n_rm = self.adjust_refinement(0.0)
            if len(new):
                self.set_initial(new)
            if len(new) == 0:
                break
        self.log_row(0)

    def run(self, max_rows=None, verbose=False):
        cfg = self.cfg
        self.initial_conditions()
        it, time, time_last_output, out_cnt = 0, 0.0, 0.0, 0
        photoi = cfg["photoi"] is not None
        while True:
            it += 1
            _, time, global_dt = self.p.get_time_state()
            if time >= cfg["end_time"]:
                break
            # output-time clipping (src/streamer.f90:222-229)
            dt = global_dt  # (the step's dt is global_dt when not clipped)
            write_out = time + dt >= time_last_output + cfg["output_dt"]
            if write_out:
                self.p.limit_step_dt(max(0.0, time_last_output + cfg["output_dt"] - time))
            info = self.p.step()
            self.steps += 1
            if info["n_rejected"] > 0:
                write_out = False
            if write_out:
                out_cnt += 1
                _, time_last_output, _ = self.p.get_time_state()
                self.log_row(out_cnt)
                if verbose:
                    print("row", out_cnt, "it", it, "boxes", self.t["n_boxes"], self.rows[-1][:4], flush=True)
                if max_rows is not None and out_cnt >= max_rows:
                    break
            if it % REFINE_PER_STEPS == 0:
                for iv in self.dens:
                    self.p.restrict_tree(iv)
                    self.p.gc_tree(iv)
                _, gtime, _ = self.p.get_time_state()
                new, n_rm = self.adjust_refinement(gtime)
                if len(new) or n_rm:
                    self.p.field_compute(0, have_guess=True)
                    if photoi:
                        self.p.photoi_set_src()
        return np.array(self.rows)


def reference_rows(name):
    import json
    logs = json.load(open(os.path.join(HERE, "golden", "regression_logs.json")))
    lines = logs[CASES[name]["log"]]["lines"]
    return np.array([[float(x) for x in ln.split()] for ln in lines])
