"""Benchmark: cell-steps/s of the full 3D AMR streamer step (drift-diffusion
+ multigrid Poisson to 1e-8 relative residual + 3-term Helmholtz
photoionization) on one MI355X.

One "step" is one iteration of the reference's time loop
(src/streamer.f90:181-365 without output/refinement): photoionization
source and three Helmholtz FMG solves, backup copy, forward-Euler flux and
source update, then the Poisson solve to the residual threshold and the
field.  Cell-steps = leaf cells advanced per step (af_num_cells_used),
summed over the timed steps.

Prints one JSON line.  The default run measures the selected precision form
on the GPU, the plain FP64 GPU form on the same workload, and the CPU
oracle (the reference algorithm restated in C++/OpenMP) on a bounded sample
(one step of the same workload from the same state) on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "afivo-streamer_amd"))

import numpy as np  # noqa: E402

import afgpu as A  # noqa: E402
import setups as S  # noqa: E402

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dominant kernel (the single-precision
# half-sweep on the finest level of this workload), from the two rocprofv3
# counter passes in profiles/gsrb_f32_3d_amr_l5_counters.txt, each scaled by
# the same counter's reading for a plain copy of known size
# (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000).
MEASURED_TRAFFIC = {
    ("xyz3d_amr_l5_r5.2mm", "mixed", "gsrb"): (17678.0 * 2.000 + 12288.5 * 1.000) * 1024,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(args):
    case = S.config1(uniform_lvl=args.uniform_lvl, max_lvl=args.max_lvl, radius=args.radius * 1e-3,
                     photoi=True, num_vcycles=args.max_vcycles, rel_residual=args.rel_residual)
    case.params.photoi_per_steps = args.photoi_per_steps
    return case


FORCED_SHAPES = None   # class -> (boxes per workgroup, workgroups per CU)
SAVE_SHAPES = None
BOX_LDS = 0            # AFGPU_OPT_BOX_LDS of every handle


def tune(p, enabled):
    """Work-decomposition shapes for this workload (measured by the
    library's autotuner from the current state, which it leaves unchanged),
    or the shapes given on the command line."""
    if FORCED_SHAPES is not None:
        p.set_shapes([tuple(FORCED_SHAPES[k]) for k in A.KERNEL_CLASSES])
        return {k: list(FORCED_SHAPES[k]) for k in A.KERNEL_CLASSES}
    if not enabled:
        return None
    t0 = time.time()
    shapes = p.autotune()
    log("  autotune %.2f s: %s" % (time.time() - t0, {k: v[:2] for k, v in shapes.items()}))
    out = {k: list(v[:2]) for k, v in shapes.items()}
    global SAVE_SHAPES
    if SAVE_SHAPES:
        with open(SAVE_SHAPES, "w") as f:
            json.dump(out, f)
        SAVE_SHAPES = None  # (the measured run's shapes, not the reference run's)
    return out


def run_gpu(case, precision, warmup, steps, timing_steps, direct_n=None, smooth=None, block=None, batch=None,
            photoi_vcycle=None, autotune=True, fuse_direct=None, direct_f32=None):
    p = A.Path("gpu", case.D, case.nc, case.coord, precision)
    S.apply(p, case)
    p.set_option(A.OPT_BOX_LDS, BOX_LDS)
    if fuse_direct is not None:
        p.set_option(A.OPT_FUSE_DIRECT, fuse_direct)
    if direct_f32 is not None:
        p.set_option(A.OPT_DIRECT_F32, direct_f32)
    if direct_n is not None:
        p.set_option(A.OPT_DIRECT_N, direct_n)
    if smooth is not None:
        p.set_option(A.OPT_MIXED_SMOOTH, smooth)
    if block is not None:
        p.set_option(A.OPT_BLOCK_SMOOTH, block)
    if batch is not None:
        p.set_option(A.OPT_BATCH_MODES, batch)
    if photoi_vcycle is not None:
        p.set_option(A.OPT_PHOTOI_VCYCLE, photoi_vcycle)
    t0 = time.time()
    p.field_compute(0, have_guess=False)
    log("  initial field solve %.2f s" % (time.time() - t0))
    for i in range(warmup):
        info = p.step()
        log("  warmup", i, info["n_vcycles"], "%.3e" % info["rel_residual"])
    p.shapes = tune(p, autotune)
    # timed region: the production path (captured HIP graphs)
    infos = []
    t0 = time.perf_counter()
    for i in range(steps):
        infos.append(p.step())
    t1 = time.perf_counter()
    # per-kernel durations: HIP events recorded around every launch of the
    # GSRB / flux / update kernels on the library's stream, over further
    # steps of the same run executed eagerly (HIP does not time event nodes
    # inside graphs; the kernels and their inputs are the same)
    stats = {}
    if timing_steps > 0:
        # (modes in turn, so that no two timed kernels share the GPU)
        p.set_option(A.OPT_BATCH_MODES, 0)
        stats = kernel_stats(p, timing_steps, p.step)
        if batch is not None:
            p.set_option(A.OPT_BATCH_MODES, batch)
    return p, infos, t1 - t0, stats


def kernel_stats(p, n_steps, step):
    """Durations of the GSRB / flux / update launches over n_steps further
    calls of step(), executed eagerly with kernel timing on: {class:
    {total_ms, launches, algo_bytes}}."""
    import ctypes as C
    p.lib.afgpu_kernel_timing(p.h, 1)
    for k in range(3):
        p.lib.afgpu_kernel_stats(p.h, k, 1, None, None, None)
    for i in range(n_steps):
        step()
    stats = {}
    for k, name in enumerate(("gsrb", "flux", "update")):
        ms, n, b = C.c_double(), C.c_longlong(), C.c_double()
        p.lib.afgpu_kernel_stats(p.h, k, 1, C.byref(ms), C.byref(n), C.byref(b))
        stats[name] = {"total_ms": ms.value, "launches": n.value, "algo_bytes": b.value}
    p.lib.afgpu_kernel_timing(p.h, 0)
    return stats


def roofline_of(stats, timing_steps, traffic=None):
    """The roofline object of the dominant kernel class in stats."""
    dom = max(stats, key=lambda k: stats[k]["total_ms"]) if stats else None
    if not dom:
        return None
    st = stats[dom]
    avg_ms = st["total_ms"] / max(st["launches"], 1)
    achieved = st["algo_bytes"] / (st["total_ms"] * 1e-3) / 1e9 if st["total_ms"] > 0 else 0.0
    return {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_launch_ms": round(avg_ms, 5), "launches": st["launches"],
            "algo_bytes_per_launch": st["algo_bytes"] / max(st["launches"], 1),
            "timing": "start/stop events recorded by each launch's dispatch (hipExtLaunchKernel) on the "
                      "library's stream, %d eager steps after the timed region" % timing_steps}


def measure_refine(case, p, c=None, reps=10):
    """The refinement flags of the workload (afgpu_refine_flags, with the
    reference's default settings) on the device state, against the download
    of |E| and n_e that the flags replace and, given the oracle handle c,
    against the oracle computing the same flags on the host."""
    t = case.extra["export"]
    n = t["n_boxes"]
    D, nc = case.D, case.nc
    s = A.refine_settings(global_time=1.0)
    p.restrict_tree(A.CC_ELECTRON)
    f = p.refine_flags(s, n)
    t0 = time.perf_counter()
    for _ in range(reps):
        f = p.refine_flags(s, n)
    gpu_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(3):
        p.get_cc(A.CC_EFIELD)
        p.get_cc(A.CC_ELECTRON)
    dl_ms = (time.perf_counter() - t0) * 1e3 / 3
    ch = t["children"].reshape(n, 2 ** D)
    leaf = ch[:, 0] == 0
    evaluated = int(leaf.sum()) + int(sum(1 for b in range(n) if not leaf[b] and leaf[ch[b] - 1].any()))
    out = {"gpu_ms": round(gpu_ms, 3), "download_E_ne_ms": round(dl_ms, 3), "boxes_evaluated": evaluated,
           "cells_evaluated": evaluated * nc ** D, "algo_bytes": evaluated * nc ** D * 16,
           "refine": int((f == A.REF_REFINE).sum()), "derefine": int((f == A.REF_DEREFINE).sum()),
           "timing": "wall time of the entry point (list upload, device pass, flags back, host passes), "
                     "mean of %d calls" % reps}
    if c is not None:
        t0 = time.perf_counter()
        c.refine_flags(s, n)
        out["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        out["cpu_oracle_cores"] = 1
    log("  refinement flags", json.dumps(out))
    return out


def measure_diagnostics(case, p, c=None, reps=10):
    """The regression log's statistics of the three densities from the
    device pool (afgpu_regression_values) against downloading the densities
    for the host to reduce, and the oracle reducing them; whole boxes of
    seven variables in af_write_tree's order (afgpu_get_boxes) against seven
    per-variable downloads."""
    t = case.extra["export"]
    D, nc, n = case.D, case.nc, t["n_boxes"]
    ivs = [A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION]
    p.regression_values(ivs)
    t0 = time.perf_counter()
    for _ in range(reps):
        p.regression_values(ivs)
    gpu_ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for iv in ivs:
        p.get_cc(iv)
    dl_ms = (time.perf_counter() - t0) * 1e3
    leaves = int((t["children"].reshape(n, 2 ** D)[:, 0] == 0).sum())
    out = {"regression_values_ms": round(gpu_ms, 3), "download_densities_ms": round(dl_ms, 3),
           "leaf_cells": leaves * nc ** D, "algo_bytes": 3 * leaves * nc ** D * 8,
           "timing": "wall time of afgpu_regression_values (3 densities: sums, sums of squares, maxima; "
                     "result back), mean of %d calls" % reps}
    if c is not None:
        t0 = time.perf_counter()
        c.regression_values(ivs)
        out["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        out["cpu_oracle_cores"] = 1
    io_vars = [A.CC_PHOTO, A.CC_EFIELD, A.CC_PHI, A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION, A.CC_RHS]
    p.get_boxes(io_vars[:1], np.arange(1, 3, dtype=np.int32))
    t0 = time.perf_counter()
    boxes = p.get_boxes(io_vars)
    out["get_boxes_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
    out["get_boxes_bytes"] = int(boxes.nbytes)
    t0 = time.perf_counter()
    for iv in io_vars:
        p.get_cc(iv)
    out["get_cc_7_vars_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
    del boxes
    log("  diagnostics", json.dumps(out))
    return out


def measure_chemistry(n_cells, with_cpu, reps=10):
    """The source term of the reference's air network (8 species, 25
    reactions; tests/golden/air_chemistry_v1.txt) on n_cells cells resident
    in device memory (afgpu_chem_source), against the oracle on a sample."""
    import chemistry as CH
    path = os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt")
    net = CH.read_network(path, {"N2": 0.8, "O2": 0.2}, S.gas_number_density(), 0.0, 1000.0, 1000)
    ns = len(net.species)
    g = CH.Chem("gpu", ns, net.reactions, 0.0, 1000.0, net.fld_cols, None, 300.0)
    rng = np.random.default_rng(3)
    fields_h = rng.uniform(0.0, 300.0, n_cells)
    dens_h = 10.0 ** rng.uniform(12, 20, (ns, n_cells))
    fields = A.DeviceArray(None, fields_h)
    dens = A.DeviceArray(None, dens_h)
    derivs = A.DeviceArray((ns, n_cells))
    g.source(fields, dens, derivs)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.source(fields, dens, derivs)
    gpu_s = (time.perf_counter() - t0) / reps
    out = {"network": "air_chemistry_v1 (8 species, 25 reactions)", "cells": n_cells,
           "source_ms": round(gpu_s * 1e3, 3), "cells_per_s": n_cells / gpu_s,
           "algo_bytes": n_cells * (1 + 2 * ns) * 8,
           "timing": "wall time of afgpu_chem_source (one launch, synchronised), mean of %d calls" % reps}
    out["achieved_GBs"] = out["algo_bytes"] / gpu_s / 1e9
    if with_cpu:
        m = 200000
        o = CH.Chem("oracle", ns, net.reactions, 0.0, 1000.0, net.fld_cols, None, 300.0)
        f = fields_h[:m].copy()
        d = np.ascontiguousarray(dens_h[:, :m])
        dv = np.zeros((ns, m))
        t0 = time.perf_counter()
        o.source(f, d, dv)
        cs = time.perf_counter() - t0
        out["cpu_oracle_cells_per_s"] = m / cs
        out["cpu_oracle_sample"] = "%d cells, one host core, %.2f s" % (m, cs)
    for a in (fields, dens, derivs):
        a.free()
    g.close()
    log("  chemistry", json.dumps(out))
    return out


def measure_air_step(case, steps=5, warmup=2):
    """The time step with the reference's air network (8 species, 25
    reactions) in place of the standard model (afgpu_set_chemistry), mixed
    precision, on the same workload: cell-steps/s."""
    import chemistry as CH
    N = case.params.gas_number_density
    E = S.analytic_transport()[0]
    x_max = E[-1] * 1e21 / N
    net = CH.read_network(os.path.join(ROOT, "tests", "golden", "air_chemistry_v1.txt"), {"N2": 0.8, "O2": 0.2}, N,
                          0.0, x_max, 1000)
    p = A.Path("gpu", case.D, case.nc, case.coord, A.PREC_MIXED)
    S.apply(p, case)
    chem = CH.Chem("gpu", len(net.species), net.reactions, 0.0, x_max, net.fld_cols, None, 300.0)
    p.set_chemistry(chem, net.charges)
    ne0 = np.asarray(case.ne0)
    p.put_species(1, 0.8 * ne0)
    p.put_species(2, 0.2 * ne0)
    p.field_compute(0, have_guess=False)
    for _ in range(warmup):
        p.step()
    t0 = time.perf_counter()
    cells = sum(p.step()["leaf_cells"] for _ in range(steps))
    secs = time.perf_counter() - t0
    out = {"network": "air_chemistry_v1 (8 species, 25 reactions)", "precision": "mixed", "steps": steps,
           "value": cells / secs, "unit": "cell-steps/s", "ms_per_step": secs * 1e3 / steps}
    p.close()
    chem.close()
    log("  air network step", json.dumps(out))
    return out


def measure_new_boxes(case, p, c=None):
    """Data of new boxes on the device (afgpu_prolong_boxes) after refining
    a tenth of the finest leaves whose face neighbours all exist, against
    the host doing it: the oracle's prolongation and the upload of the new
    boxes' variables.  Changes the tree of p (and c): call last."""
    t0 = case.extra["export"]
    D, nc, n0 = case.D, case.nc, t0["n_boxes"]
    nch, nnb = 2 ** D, 2 * D
    top = t0["lvl"].max()
    inner = [b + 1 for b in range(n0) if t0["lvl"][b] == top and t0["children"][b * nch] == 0
             and all(v != 0 for v in t0["neighbors"][b * nnb:(b + 1) * nnb])]
    pick = inner[::10]
    case.tree.refine(np.asarray(pick, np.int32))
    t1 = case.tree.export()
    new = np.arange(n0 + 1, t1["n_boxes"] + 1, dtype=np.int32)
    p.sync_tree(t1)
    # the prolongation reads the parents only, so repeating it rewrites the
    # same values: one untimed call, then the mean of five
    p.prolong_boxes(new)
    t_0 = time.perf_counter()
    for _ in range(5):
        p.prolong_boxes(new)
    gpu_ms = (time.perf_counter() - t_0) * 1e3 / 5
    n_vars = 7 if case.params.photoi_enabled else 6
    out = {"new_boxes": int(len(new)), "refined": len(pick), "gpu_ms": round(gpu_ms, 3),
           "algo_bytes": int(len(new)) * n_vars * (nc ** D + nc ** D // nch) * 8,
           "timing": "wall time of afgpu_prolong_boxes (id upload, prolongation and ghost cells of %d "
                     "variables, synchronised), mean of 5 calls after one" % n_vars}
    data = [p.get_cc(iv, new) for iv in (A.CC_PHOTO, A.CC_EFIELD, A.CC_PHI, A.CC_ELECTRON, A.CC_POS_ION,
                                         A.CC_NEG_ION, A.CC_RHS)[7 - n_vars:]]
    t_0 = time.perf_counter()
    for iv, d in zip((A.CC_PHOTO, A.CC_EFIELD, A.CC_PHI, A.CC_ELECTRON, A.CC_POS_ION, A.CC_NEG_ION,
                      A.CC_RHS)[7 - n_vars:], data):
        p.put_cc(iv, d, new)
    out["upload_new_ms"] = round((time.perf_counter() - t_0) * 1e3, 3)
    if c is not None:
        c.sync_tree(t1)
        t_0 = time.perf_counter()
        c.prolong_boxes(new)
        out["cpu_oracle_ms"] = round((time.perf_counter() - t_0) * 1e3, 3)
        out["cpu_oracle_cores"] = 1
    log("  new boxes", json.dumps(out))
    return out


def rel_err(a, b):
    d = a - b
    return float(np.abs(d).max() / max(np.abs(b).max(), 1e-300)), \
        float(np.sqrt((d * d).sum() / max((b * b).sum(), 1e-300)))


def leaf_interior(case, p, iv):
    """Interior cells of the leaf boxes: the values the physics reads (the
    refined coarse boxes hold multigrid auxiliary values in the FP64 form)."""
    ex = case.extra["export"]
    ch = ex["children"].reshape(ex["n_boxes"], 2 ** case.D)
    leaves = np.nonzero(ch[:, 0] == 0)[0] + 1
    a = p.get_cc(iv, ids=leaves).reshape((len(leaves),) + (case.nc + 2,) * case.D)
    return a[(slice(None),) + (slice(1, case.nc + 1),) * case.D]


# Geometric factor of the coarse levels of a V-cycle (1 + 1/2^D + ...)
GEO = {2: 4.0 / 3.0, 3: 8.0 / 7.0}


def step_bytes(info, D, mixed, photoi):
    """Algorithmic bytes of one time step, counted as words per leaf cell
    touched once per kernel pass over interior cells (ghost layers, halos
    and the rejection backup excluded):
    - per step: flux 2 + 2D words, update and source 7 + D (+1 reading the
      photoionization source), rhs 4, gradient and norm 2 + D (27 in 3D);
    - per V-cycle (outer iteration of the mixed form): 22 words times the
      coarse levels' factor (4/3 in 2D, 8/7 in 3D): four GSRB sweeps of 3
      words, residual 3, restriction 2, correction 2, final residual 3;
    - photoionization: 4 words for the source and the mode sum, and
      (4/3) 22 words (times the same factor) per FMG cycle of a mode.
    FP64 words are 8 bytes.  In the mixed form the inner cycle's words
    (sweeps, restriction, correction: 16 of the 22) are FP32, 4 bytes; its
    FP64 residual and correction (6 of 22) and the per-step words stay 8."""
    g = GEO[D]
    cycle = (16 * 4 + 6 * 8) if mixed else 22 * 8
    fixed = (2 + 2 * D) + (7 + D + (1 if photoi else 0)) + 4 + (2 + D)
    b = fixed * 8 + info["n_vcycles"] * g * cycle
    if photoi and info["photoi_done"]:
        b += 4 * 8 + info["photoi_cycles"] * (4.0 / 3.0) * g * cycle
    return b * info["leaf_cells"]


def host_cpu():
    """The host's CPU model (as /proc/cpuinfo names it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_probe(mib):
    """A plain device-to-device copy of known size, for counter calibration:
    each launch reads and writes exactly `mib` MiB."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_launch": n * 4, "launches": 5}), flush=True)


def ensemble_leaf_interior(ens, r, p, iv):
    """Interior cells of run r's leaf boxes in an ensemble handle."""
    c = ens.cases[r]
    ex = c.extra["export"]
    ch = ex["children"].reshape(ex["n_boxes"], 4)
    leaves = np.nonzero(ch[:, 0] == 0)[0] + 1 + ens.offsets[r]
    a = p.get_cc(iv, ids=leaves).reshape((len(leaves), c.nc + 2, c.nc + 2))
    return a[:, 1:c.nc + 1, 1:c.nc + 1]


def run_ensemble(ens, precision, warmup, steps, autotune=True, step_graph=1, timing_steps=0):
    p = S.apply_ensemble(A.Path("gpu", 2, 8, A.CYL, precision), ens)
    p.set_option(A.OPT_STEP_GRAPH, step_graph)
    t0 = time.time()
    p.field_compute_runs(0, have_guess=False)
    log("  initial field solves %.2f s" % (time.time() - t0))
    for i in range(warmup):
        infos = p.step_ensemble()
        log("  warmup", i, "vcycles", [f["n_vcycles"] for f in infos][:16], "...")
    p.shapes = tune(p, autotune)
    all_infos = []
    t0 = time.perf_counter()
    for i in range(steps):
        all_infos.append(p.step_ensemble())
    t1 = time.perf_counter()
    p.kstats = kernel_stats(p, timing_steps, p.step_ensemble) if timing_steps > 0 else {}
    return p, all_infos, t1 - t0


def bench_ensemble(args):
    """The ensemble workload: independent 2D axisymmetric runs of uneven
    sizes advancing together in one forest (mixed precision at the given
    residual), with each run's error against the same runs in FP64."""
    t0 = time.time()
    ens = S.ensemble(n_runs=args.runs, num_vcycles=args.max_vcycles, rel_residual=args.rel_residual)
    sizes = [int(c.tree.info()["leaf_cells"]) for c in ens.cases]
    log("workload", ens.name, "built in %.1f s; sizes" % (time.time() - t0),
        {int(np.sqrt(k)): sizes.count(k) for k in sorted(set(sizes))})
    prec = A.PREC_MIXED if args.precision == "mixed" else A.PREC_FP64
    p, infos, secs = run_ensemble(ens, prec, args.warmup, args.steps, bool(args.autotune), args.step_graph,
                                  args.timing_steps)
    if p.kstats:
        log("  kernel stats", json.dumps(p.kstats))
    cell_steps = sum(f["leaf_cells"] for step in infos for f in step)
    value = cell_steps / secs
    nv = np.array([[f["n_vcycles"] for f in step] for step in infos])
    log("  %.3f s for %d steps: %.4e cell-steps/s; vcycles per run-step min %d mean %.2f max %d"
        % (secs, args.steps, value, nv.min(), nv.mean(), nv.max()))
    extra = {}
    if prec == A.PREC_MIXED and not args.no_fp64_ref:
        log("GPU fp64 ensemble")
        # (the same eager timing steps, so that both forms are compared at the
        # same step)
        ref, rinfos, rsecs = run_ensemble(ens, A.PREC_FP64, args.warmup, args.steps, bool(args.autotune),
                                          args.step_graph, args.timing_steps)
        fp64_value = sum(f["leaf_cells"] for step in rinfos for f in step) / rsecs
        extra["fp64_gpu_value"] = fp64_value
        extra["mixed_over_fp64"] = value / fp64_value
        worst = {}
        per_run = []
        for r in range(len(ens.cases)):
            e = {}
            for name, iv in (("phi", A.CC_PHI), ("efield", A.CC_EFIELD), ("electron", A.CC_ELECTRON)):
                mx, l2 = rel_err(ensemble_leaf_interior(ens, r, p, iv), ensemble_leaf_interior(ens, r, ref, iv))
                e[name] = mx
                worst[name] = max(worst.get(name, 0.0), mx)
            per_run.append(e)
        extra["error_vs_fp64"] = {"max_rel_over_runs": worst,
                                  "per_run_phi": [round(e["phi"], 14) for e in per_run],
                                  "cells": "leaf interiors of every run after %d steps, max relative "
                                           "to max |FP64| of that run"
                                           % (args.warmup + args.steps + args.timing_steps)}
        log("  fp64 %.4e cell-steps/s; mixed/fp64 %.3f; worst errors %s" % (fp64_value, value / fp64_value, worst))
    cpu = None
    if not args.no_cpu:
        if args.cpu_steps <= 0:
            args.cpu_steps = 8
        # every run of the ensemble, one after the other as the reference's
        # separate processes would, continued by the oracle for cpu_steps
        # steps from the GPU state (the FP64 form's when it ran)
        src = ref if "fp64_gpu_value" in extra else p
        threads = min(16, os.cpu_count() or 1)
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        os.environ.setdefault("OMP_PROC_BIND", "close")
        csecs, ccells, nthreads = 0.0, 0, 0
        for r, case in enumerate(ens.cases):
            c = A.Path("oracle", 2, 8, A.CYL)
            S.apply(c, case)
            ids = ens.run_ids(r)
            for iv in list(range(9)) + [A.CC_PHI, A.CC_EFIELD]:
                c.put_cc(iv, src.get_cc(iv, ids))
            c.put_fc(A.FC_FIELD, src.get_fc(A.FC_FIELD, ids))
            c.set_time_state(*src.get_run_time_state(r))
            c.lib.afcpu_num_threads.restype = A.C.c_int
            nthreads = c.lib.afcpu_num_threads()
            t0 = time.perf_counter()
            for _ in range(args.cpu_steps):
                ccells += c.step()["leaf_cells"]
            csecs += time.perf_counter() - t0
            c.close()
        cpu = {"value": ccells / csecs, "unit": "cell-steps/s", "cores": nthreads, "kind": "port",
               "sample": "%d step(s) of every run (%d cells) from the GPU FP64 state, %.1f s"
                         % (args.cpu_steps, int(sum(sizes)), csecs),
               "cpu_model": host_cpu(), "host_cpus": os.cpu_count(),
               "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
               "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
        log("  cpu %.4e cell-steps/s (%.1f s)" % (cpu["value"], csecs))
    out = {
        "metric": "cell-steps/s, ensemble of %d uneven 2D runs (mixed at %g)" % (len(ens.cases), args.rel_residual),
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if prec == A.PREC_MIXED else "f64", "data": "synthetic",
        "config": {"workload": ens.name, "runs": len(ens.cases), "cells": int(sum(sizes)),
                   "box_cells": "8^2", "precision": args.precision, "rel_residual": args.rel_residual,
                   "mean_vcycles": float(nv.mean()), "step_graph": args.step_graph, "shapes": p.shapes},
        "roofline": roofline_of(p.kstats, args.timing_steps), "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["amr3d", "ensemble"], default="amr3d",
                    help="amr3d: the 3D AMR step (the headline metric); ensemble: independent 2D runs")
    ap.add_argument("--runs", type=int, default=64, help="ensemble: number of runs")
    ap.add_argument("--step-graph", type=int, default=0,
                    help="ensemble: each step's first attempt and field solve as one captured graph")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--uniform-lvl", type=int, default=4)
    ap.add_argument("--max-lvl", type=int, default=5)
    ap.add_argument("--radius", type=float, default=5.2, help="refined channel radius (mm)")
    ap.add_argument("--rel-residual", type=float, default=1e-8)
    ap.add_argument("--max-vcycles", type=int, default=100)
    ap.add_argument("--photoi-per-steps", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="CPU oracle sample: steps of the workload (0: 25 for the 3D workload, about 10 s "
                         "on 16 cores; 8 of every run for the ensemble)")
    ap.add_argument("--timing-steps", type=int, default=2,
                    help="eager steps after the timed region with per-kernel HIP event timing")
    ap.add_argument("--direct-n", type=int, default=128,
                    help="mixed precision: largest uniform level (cells per axis) solved exactly "
                         "by the inner cycle; 0 = descend to level 1")
    ap.add_argument("--mixed-smooth", type=int, default=0,
                    help="mixed precision: GSRB sweeps per side in the inner cycle (0 = reference counts)")
    ap.add_argument("--block-smooth", type=int, default=0,
                    help="mixed precision: block smoother in the inner cycle (1) or red-black sweeps (0)")
    ap.add_argument("--batch-modes", type=int, default=1,
                    help="mixed precision: photoionization modes as parallel graph branches (1) or in turn (0)")
    ap.add_argument("--direct-f32", type=int, default=0,
                    help="mixed precision: uniform-level solve of the FP32 inner cycle in FP32 (1) or FP64 (0)")
    ap.add_argument("--fuse-direct", type=int, default=0,
                    help="mixed precision: uniform-level solve with the fused transform passes (1) or separate (0)")
    ap.add_argument("--photoi-vcycle", type=int, default=0,
                    help="mixed precision: photoionization inner cycle V-cycle (1) or FMG (0)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB a few times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes) and exit")
    ap.add_argument("--autotune", type=int, default=1,
                    help="pick the kernels' work-decomposition shapes by measurement before the timed region")
    ap.add_argument("--shapes", default="", metavar="BB,PER_CU|FILE",
                    help="work-decomposition shapes instead of the autotuner's: one shape for every "
                         "kernel class, or a JSON file of class -> [bb, per_cu] as --save-shapes writes")
    ap.add_argument("--save-shapes", default="", metavar="FILE",
                    help="write the autotuner's shapes (of the run measured first) to FILE")
    ap.add_argument("--box-lds", type=int, default=0,
                    help="mixed precision, 3D 16^3 boxes: stage whole boxes in LDS in the half-sweeps and the FP64 "
                         "residual (1) instead of only what they read (0); identical results")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-ref", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the mixed form's variants (FP32 coarsest solve, inner cycle to level 1)")
    ap.add_argument("--no-side", action="store_true",
                    help="skip the measurements beyond the time step (refinement flags, diagnostics, chemistry, "
                         "new boxes), e.g. to profile the time step alone")
    args = ap.parse_args()
    global FORCED_SHAPES, SAVE_SHAPES, BOX_LDS
    BOX_LDS = args.box_lds
    if args.shapes:
        if os.path.exists(args.shapes):
            with open(args.shapes) as f:
                FORCED_SHAPES = json.load(f)
        else:
            one = [int(x) for x in args.shapes.split(",")]
            FORCED_SHAPES = {k: one for k in A.KERNEL_CLASSES}
    SAVE_SHAPES = args.save_shapes or None
    if args.copy_probe > 0:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print("bench.py measures one GPU: the path runs on a single MI355X here "
              "(sharding across GPUs is a design, not built)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ensemble":
        return bench_ensemble(args)
    case = make_case(args)
    info = case.tree.info()
    log("workload", case.name, info)
    prec = A.PREC_MIXED if args.precision == "mixed" else A.PREC_FP64
    log("GPU %s run" % args.precision)
    mixed = prec == A.PREC_MIXED
    p, infos, secs, stats = run_gpu(case, prec, args.warmup, args.steps, args.timing_steps,
                                    args.direct_n if mixed else None, args.mixed_smooth if mixed else None,
                                    args.block_smooth if mixed else None, args.batch_modes if mixed else None,
                                    args.photoi_vcycle if mixed else None, bool(args.autotune),
                                    args.fuse_direct if mixed else None, args.direct_f32 if mixed else None)
    cell_steps = sum(i["leaf_cells"] for i in infos)
    value = cell_steps / secs
    nv = [i["n_vcycles"] for i in infos]
    log("  %.3f s for %d steps: %.4e cell-steps/s, vcycles/step %s" % (secs, args.steps, value, nv))

    # roofline of the dominant kernel class
    dom = max(stats, key=lambda k: stats[k]["total_ms"]) if stats else None
    roof = None
    if dom:
        traffic = None
        if not args.block_smooth and args.direct_n == 128 and args.mixed_smooth == 0:
            traffic = MEASURED_TRAFFIC.get((case.name, args.precision, dom))
        roof = roofline_of(stats, args.timing_steps, traffic)
        log("  kernel stats", json.dumps(stats))

    photoi = bool(case.params.photoi_enabled)
    sb = sum(step_bytes(i, case.D, mixed, photoi) for i in infos)
    step_roof = {"bound": "hbm", "algo_bytes_per_step": sb / len(infos), "achieved": round(sb / secs / 1e9, 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(sb / secs / 1e9 / HBM_PEAK_GBS, 4),
                 "mean_photoi_cycles": float(np.mean([i["photoi_cycles"] for i in infos])),
                 "rule": "words per leaf cell once per kernel pass: 27 per step (+1 with photoionization), "
                         "22 x 8/7 per V-cycle, 4 + (4/3) 22 x 8/7 per mode FMG cycle; FP64 words 8 B, "
                         "the mixed inner cycle's 16 of 22 words 4 B; timed steps' wall time"}
    log("  whole step: %.3e algorithmic bytes per step, %.1f GB/s (%.3f of peak)"
        % (step_roof["algo_bytes_per_step"], step_roof["achieved"], step_roof["frac"]))

    extra = {}
    ref = None
    if mixed and not args.no_variants:
        # variants of the mixed form beside the measured one: the inner
        # cycle's uniform-level solve in FP32, and the inner cycle descending
        # to level 1 as the plain V-cycle does (separates the precision from
        # the inner cycle's structure)
        for key, kw in (("mixed_f32_coarsest_value", {"direct_f32": 1 - args.direct_f32}),
                        ("mixed_vcycle_to_level1_value", {"direct_n": 0})):
            log("GPU mixed variant %s" % key)
            opts = dict(direct_n=args.direct_n, direct_f32=args.direct_f32)
            opts.update(kw)
            q, qinfos, qsecs, _ = run_gpu(case, A.PREC_MIXED, args.warmup, args.steps, 0, opts["direct_n"],
                                          args.mixed_smooth, args.block_smooth, args.batch_modes,
                                          args.photoi_vcycle, bool(args.autotune), args.fuse_direct,
                                          opts["direct_f32"])
            extra[key] = sum(i["leaf_cells"] for i in qinfos) / qsecs
            extra[key.replace("_value", "_config")] = opts
            log("  %s %.4e cell-steps/s" % (key, extra[key]))
            q.close()
    if prec == A.PREC_MIXED and not args.no_fp64_ref:
        log("GPU fp64 reference run")
        ref, rinfos, rsecs, _ = run_gpu(case, A.PREC_FP64, args.warmup, args.steps, args.timing_steps,
                                        autotune=bool(args.autotune))
        fp64_value = sum(i["leaf_cells"] for i in rinfos) / rsecs
        extra["fp64_gpu_value"] = fp64_value
        extra["mixed_over_fp64"] = value / fp64_value
        errs = {}
        for name, iv in (("phi", A.CC_PHI), ("efield", A.CC_EFIELD), ("electron", A.CC_ELECTRON),
                         ("photo", A.CC_PHOTO)):
            mx, l2 = rel_err(leaf_interior(case, p, iv), leaf_interior(case, ref, iv))
            errs[name] = {"max_rel": mx, "l2_rel": l2}
        errs["cells"] = "leaf interiors after %d steps, relative to max |FP64| (max) and the FP64 L2 norm (l2)" \
            % (args.warmup + args.steps + args.timing_steps)
        tm, tr = p.get_time_state(), ref.get_time_state()
        errs["time"] = {"mixed": tm[1], "fp64": tr[1]}
        extra["error_vs_fp64"] = errs
        log("  fp64 %.4e cell-steps/s; mixed/fp64 = %.3f; errors %s" % (fp64_value, value / fp64_value, errs))

    cpu = None
    if not args.no_cpu:
        if args.cpu_steps <= 0:
            args.cpu_steps = 25
        src = ref if ref is not None else p
        threads = min(16, os.cpu_count() or 1)
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        os.environ.setdefault("OMP_PROC_BIND", "close")
        c = A.Path("oracle", case.D, case.nc, case.coord)
        S.apply(c, case)
        for iv in list(range(9)) + [A.CC_PHI, A.CC_EFIELD, A.CC_PHOTO] + [A.CC_HELMH + k for k in range(3)]:
            c.put_cc(iv, src.get_cc(iv))
        c.put_fc(A.FC_FIELD, src.get_fc(A.FC_FIELD))
        # continue the time loop from the GPU state: same dt and iteration
        c.set_time_state(*src.get_time_state())
        c.lib.afcpu_num_threads.restype = A.C.c_int
        nthreads = c.lib.afcpu_num_threads()
        log("CPU oracle sample: %d step(s), %d threads" % (args.cpu_steps, nthreads))
        t0 = time.perf_counter()
        ccells = 0
        for _ in range(args.cpu_steps):
            ccells += c.step()["leaf_cells"]
        csecs = time.perf_counter() - t0
        cpu = {"value": ccells / csecs, "unit": "cell-steps/s", "cores": nthreads, "kind": "port",
               "sample": "%d step(s) of the same 3D AMR workload (%d leaf cells) from the GPU state "
                         "after warmup, %.1f s" % (args.cpu_steps, info["leaf_cells"], csecs),
               "cpu_model": host_cpu(), "host_cpus": os.cpu_count(),
               "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
               "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
        log("  cpu %.4e cell-steps/s (%.1f s)" % (cpu["value"], csecs))
    if not args.no_side:
        extra["refine_flags"] = measure_refine(case, p, None if args.no_cpu else c)
        extra["diagnostics"] = measure_diagnostics(case, p, None if args.no_cpu else c)
        extra["chemistry"] = measure_chemistry(int(info["leaf_cells"]), not args.no_cpu)
        extra["air_network_step"] = measure_air_step(case)
        extra["new_boxes"] = measure_new_boxes(case, p, None if args.no_cpu else c)

    out = {
        "metric": "cell-steps/s, 3D AMR step (DD+MG Poisson+photoi), FP64 vs mixed at 1e-8",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == A.PREC_MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": case.name, "leaf_boxes": info["n_leaves"], "leaf_cells": info["leaf_cells"],
                   "boxes": info["n_boxes"], "levels": info["highest_lvl"], "box_cells": "16^3",
                   "precision": args.precision, "rel_residual": args.rel_residual,
                   "photoi_per_steps": args.photoi_per_steps, "mean_vcycles": float(np.mean(nv)),
                   "direct_n": args.direct_n if prec == A.PREC_MIXED else None,
                   "mixed_smooth": args.mixed_smooth if prec == A.PREC_MIXED else None,
                   "block_smooth": args.block_smooth if prec == A.PREC_MIXED else None,
                   "batch_modes": args.batch_modes if prec == A.PREC_MIXED else None,
                   "photoi_vcycle": args.photoi_vcycle if prec == A.PREC_MIXED else None,
                   "box_lds": args.box_lds if prec == A.PREC_MIXED else None,
                   "fuse_direct": args.fuse_direct if prec == A.PREC_MIXED else None,
                   "direct_f32": args.direct_f32 if prec == A.PREC_MIXED else None,
                   "shapes": p.shapes},
        "roofline": roof,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
