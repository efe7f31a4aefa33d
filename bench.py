#!/usr/bin/env python
"""Headline benchmark: the isolated 3D FAS multigrid Poisson solve on a frozen
octree (boxes of 16^3 cells, uniform level 4 = 4096 boxes plus a sub-volume
refined to level 5; 64 seeded Gaussian charge blobs; phi = V / 0 on the
bottom / top faces, Neumann sides), solved to 1e-10 relative residual.

One step = one full solve from phi = 0 to max|rhs - A phi| / max|rhs| <=
1e-10 over the leaves. Two forms are timed on the same tree:
  fp64  - mg3_fas_fmg (first call) then mg3_fas_vcycle, the schedule of
          afivo/src/m_aX_mg.f90 with the drivers' n_cycle_base = 8; the
          number of V-cycles is fixed in warm-up, the residual re-checked
          after the timed steps;
  mixed - FP32 FAS multigrid inside an FP64 defect-correction loop (first
          correction: FMG + 3 V-cycles, later ones: 4 V-cycles); its
          potential is compared with the FP64 solution (max relative error
          over leaf cells).
The headline value is FP64's unless the mixed form meets the error gate of
1e-8 relative potential error and is faster by more than the run-to-run
spread of these timings (about 1%; the rule takes 2%); both forms, their
ratio and their rooflines are in the line.
Throughput = leaf cells x solves / s. Prints one JSON line. Inputs are
resident in HBM before timing starts.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_INNER = 4   # FP32 V-cycles per correction of the mixed solve
KIND_NAMES = ["gsrb_half_sweep", "ghost_cells", "residual_restrict", "coarse_rhs",
              "correction", "copy"]


def build_gpu(W, n_cc=3):
    import a5_streamer_amd as A
    T = A.Tree(3, W.n_cell, n_cc, 0, W.dr, coarsen_to=2, n_boxes=1000)
    W.build(T)
    kinds, vals = W.bc()
    i_phi, i_tmp, i_rhs = W.ivs
    T.mg_setup(i_phi, i_tmp, i_rhs, n_down=2, n_up=2, n_base=W.n_cycle_base,
               bc_kind=kinds, bc_val=vals, auto_ops=True)
    T.set_gaussians(i_rhs, W.blobs(), 0.0)
    return T


def solve(T, max_lvl, n_v):
    T.fas_fmg(False, True)
    for _ in range(n_v):
        T.fas_vcycle(max_lvl, False)


def host_cpu():
    """The host's processor model and the cores this process may run on."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except Exception:
        usable = os.cpu_count() or 1
    return model, usable


def _ref_octree(W, n_cc=3):
    """The octree W on the reference harness, with the rhs the GPU library
    evaluates on the bit-identical tree."""
    import refharness as R
    import a5_streamer_amd as A
    Tg = A.Tree(3, W.n_cell, n_cc, 0, W.dr, coarsen_to=2, n_boxes=1000)
    W.build(Tg)
    Tg.set_gaussians(3, W.blobs(), 0.0)
    rhs = Tg.get_cc_all(3)
    Tg.close()
    T = R.RefTree(3, W.n_cell, 3, 0, W.dr, coarsen_to=2, n_boxes=1000)
    W.build(T)
    b = T.boxes()
    for i in range(len(b["lvl"])):
        if b["in_use"][i]:
            nc = int(b["n_cell"][i])
            T.set_cc(i + 1, 3, rhs[i][: (nc + 2) ** 3])
    kinds, vals = W.bc()
    T.mg_setup(1, 2, 3, n_base=W.n_cycle_base, bc_kind=kinds, bc_val=vals, auto_ops=True)
    return T


def _ref_solves(T, n_v, reps):
    """Wall seconds of each of reps reference solves (FMG + n_v V-cycles)."""
    max_lvl = T.info()["max_lvl"]
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        T.fas_fmg(False, True)
        for _ in range(n_v):
            T.fas_vcycle(max_lvl, False)
        times.append(time.perf_counter() - t0)
    return times


def cpu_baseline(W, n_v, small_W=None, n_v_small=None):
    """The reference (oracle/_ref, built from the reference sources) timed on
    this host: the median of 3 solves of the benchmark octree with OpenMP on
    every usable core, and the median of 3 solves of a smaller octree of the
    same construction on one thread (a bounded sample: one full-size solve
    takes tens of seconds on one core). The rhs comes from the GPU library on
    the bit-identical tree."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refharness as R
    if not R.available():
        return None
    model, usable = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", usable))
    T = _ref_octree(W)
    cells = W.leaf_cells(T)
    R.set_threads(threads)
    times = _ref_solves(T, n_v, 3)
    med = float(np.median(times))
    out = {
        "value": cells / med / 1e6,
        "unit": "Mcell-steps/s",
        "cores": threads,
        "kind": "reference",
        "sample": "reference Fortran (flang -O2 -fopenmp) mg3_fas_fmg + %d mg3_fas_vcycle to "
                  "1e-10 on the benchmark octree (%d leaf cells), median of 3 solves "
                  "(%s s) on %d threads" % (n_v, cells, ", ".join("%.2f" % t for t in times),
                                            threads),
        "host": {"cpu": model, "usable_cores": usable},
    }
    if small_W is not None:
        Ts = _ref_octree(small_W)
        cs = small_W.leaf_cells(Ts)
        R.set_threads(1)
        ts = _ref_solves(Ts, n_v_small, 3)
        R.set_threads(threads)
        out["one_thread"] = {
            "value": cs / float(np.median(ts)) / 1e6, "unit": "Mcell-steps/s", "cores": 1,
            "sample": "the same solve on the octree with uniform level 3 and a 5x5x8-box "
                      "level-4 sub-volume (%d leaf cells), %d V-cycles, median of 3 "
                      "(%s s) on one thread" % (cs, n_v_small,
                                                ", ".join("%.2f" % t for t in ts))}
    out["drivers"] = driver_baselines(R, threads)
    return out


# The reference's own driver programs (built as modules into oracle/_ref)
# on the physics of the 2D workload and of the 3D production streamer, each
# on its own single base box and refinement; rates over the steps between
# two end times, so that initialisation cancels
DRIVER_BASELINES = {
    "2d_air_streamer": (dict(variant="2d", n_base=1, lvls_max=30), 20e-9, 40e-9),
    "3d_streamer_first_steps": (dict(variant="3d", n_base=1, lvls_max=30, domain_len=64e-3,
                                     applied_fld=2.0e6, seed_r0=[[32e-3, 32e-3, 56e-3]],
                                     seed_r1=[[32e-3, 32e-3, 56e-3]], n_boxes=20000),
                                5e-9, 12e-9),
}


def _ref_driver_pair(R, cfg, end1, end2):
    """(leaf cell-steps, seconds, steps) of the reference driver's steps
    between end1 and end2: two runs from t = 0, the difference of their
    step counters and wall times."""
    res = []
    for end in (end1, end2):
        D = R.RefDriver(cfg.variant)
        D.config(cfg.domain_len, cfg.box_size, cfg.applied_fld, dt_max=cfg.dt_max,
                 end_time=end, epsilon_diel=cfg.epsilon_diel, steps_amr=cfg.steps_amr)
        D.init_cond(cfg.bg_dens, cfg.seed_r0, cfg.seed_r1, cfg.seed_dens, cfg.seed_width,
                    cfg.seed_falloff)
        x, cols = cfg.transport_columns()
        D.table(cfg.table_size, cfg.table_max_fld, x, cols)
        D.counters(reset=True)
        t0 = time.perf_counter()
        D.run(cfg.box_size)
        w = time.perf_counter() - t0
        n, c = D.counters()
        res.append((c, w, n))
    return res[1][0] - res[0][0], res[1][1] - res[0][1], res[1][2] - res[0][2]


def driver_baselines(R, threads):
    """Mcell-steps/s of the reference driver programs on this host's cores:
    on every usable core the median of 3 step-range measurements, on one
    core one (a bounded sample)."""
    from a5_streamer_amd import streamer as ST
    out = {}
    for name, (kw, end1, end2) in DRIVER_BASELINES.items():
        cfg = ST.Config(**kw)
        row = {}
        for thr, reps in ((threads, 3), (1, 1)):
            R.set_threads(thr)
            runs = [_ref_driver_pair(R, cfg, end1, end2) for _ in range(reps)]
            rates = [c / w / 1e6 for c, w, _ in runs]
            k = int(np.argsort(rates)[len(rates) // 2])
            c, w, n = runs[k]
            row["all_cores" if thr == threads else "one_thread"] = {
                "value": rates[k], "unit": "Mcell-steps/s", "cores": thr,
                "sample": "steps %d after t = %g s, %.3g leaf cell-steps in %.2f s%s" % (
                    n, end1, c, w, ", median of %d" % reps if reps > 1 else "")}
        out[name] = row
    R.set_threads(threads)
    return out


def converge_count(T, max_lvl, tol=1e-10, cap=60):
    """FMG from phi = 0, then the number of V-cycles to reach tol."""
    T.fas_fmg(False, True)
    for n in range(cap + 1):
        r, b = T.residual_norms()
        if r <= tol * b:
            return n, r / b
        T.fas_vcycle(max_lvl, False)
    raise RuntimeError("multigrid did not reach %g in %d V-cycles" % (tol, cap))


def timed(T, steps, fn):
    """Wall time of `steps` calls of fn, the GPU idle before and drained after."""
    T.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    T.sync()
    return time.perf_counter() - t0


def profiled(T, fn):
    """Per-kernel-class launch count, duration and algorithmic bytes of one
    more call of fn, from HIP events recorded around each launch on the
    library's stream (kept out of the timed steps: the event pairs cost
    host time per launch)."""
    T.prof_enable(True)
    T.sync()
    fn()
    T.sync()
    prof = {}
    for k, name in enumerate(KIND_NAMES):
        prof[name] = T.prof_read(k, -999)
    prof["mixed_fp64_passes"] = T.prof_read(6, -999)
    T.prof_enable(False)
    return prof


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "octree16_gsrb_pmc_traffic.md")


def measured_traffic(kernel, precision):
    """HBM bytes per launch of the finest-level launch of `kernel`, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (None if
    absent)."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 6 and w[:2] == ["-", "traffic"] and w[2] == kernel and \
                        w[3] == "finest" and w[4] == precision:
                    return float(w[5])
    except OSError:
        pass
    return None


def copy_peak(T, reps=5):
    """Measured HBM rate of a plain copy kernel on this GPU: the whole-box copy
    of one variable (ghosts included) on the finest level, bytes read plus
    written over the mean HIP-event duration. Overwrites variable 4."""
    lvl = T.info()["max_lvl"]
    n_box = len(T.level(lvl)[0])
    nc = T.n_cell
    whole = n_box * (nc + 2) ** T.dim * 8.0 * 2
    T.tree_copy_cc(1, 4)
    T.prof_enable(True)
    T.sync()
    for _ in range(reps):
        T.tree_copy_cc(1, 4)
    T.sync()
    n, ms, _ = T.prof_read(5, lvl)
    T.prof_enable(False)
    return round(whole / (ms * 1e-3 / n) / 1e9, 1) if n else None


def roofline(T, precision, fn):
    """Roofline of the dominant kernel, the Gauss-Seidel half-sweep, on the
    finest level (its largest launches): algorithmic bytes per launch over
    the launch's mean duration from HIP events on the library's stream, for
    one more call of fn. The same figures over all levels ride along."""
    lvl = T.info()["max_lvl"]
    T.prof_enable(True)
    T.sync()
    fn()
    T.sync()
    n, ms, by = T.prof_read(0, lvl)
    na, msa, bya = T.prof_read(0, -999)
    T.prof_enable(False)
    achieved = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    tr = measured_traffic("gsrb_half_sweep", precision)
    return {
        "bound": "hbm", "kernel": "gsrb_half_sweep (finest level)", "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": tr, "launches": n, "avg_launch_us": round(ms * 1e3 / max(n, 1), 2),
        "alg_bytes_per_launch": round(by / max(n, 1)),
        "all_levels": {"launches": na, "avg_launch_us": round(msa * 1e3 / max(na, 1), 2),
                       "alg_bytes_per_launch": round(bya / max(na, 1)),
                       "achieved": round(bya / (msa * 1e-3) / 1e9, 1) if msa > 0 else 0.0},
    }


def solve_bytes(T, n_v):
    """Algorithmic HBM bytes of one FP64 solve (FMG from phi = 0, then n_v
    V-cycles) by the pass-structure count of the reference: per V-cycle
    16 + 10/2^D words per cell of every level it visits (two half-sweeps
    down and up of 1.5 words, residual + restriction 2 + 2/2^D, coarse rhs
    4 and correction 3 per parent cell, correction 2 + 1/2^D per child cell,
    counted per cell of the finer level); the FMG adds a V-cycle over levels
    <= T for every level T, its phi -> tmp copy of 2 words per cell of T,
    and the first call's restriction of the rhs, 1 + 2/2^D words per cell."""
    D = T.dim
    info = T.info()
    b = T.boxes()
    cells = {}
    for l in range(info["min_lvl"], info["max_lvl"] + 1):
        ids = T.level(l)[0]
        cells[l] = int(sum(int(b["n_cell"][i - 1]) ** D for i in ids))
    per = 16 + 10 / 2 ** D
    lv = sorted(cells)
    words = 0.0
    for t in lv:
        words += per * sum(cells[l] for l in lv if l <= t) + 2 * cells[t]
        if t > lv[0]:
            words += (1 + 2 / 2 ** D) * cells[t]
    words += n_v * per * sum(cells.values())
    return 8.0 * words


# kernel classes of the streamer step (the library's event-timing kinds)
STEP_KINDS = {0: "gsrb_half_sweep", 1: "ghost_cells", 2: "residual_restrict",
              3: "coarse_rhs", 4: "correction", 5: "copy", 7: "space_charge",
              8: "field_from_potential", 9: "koren_flux", 10: "flux_fix", 11: "update",
              12: "average", 13: "restrict", 14: "photoi_rate"}


TUNE_KERNELS = ("gsrb", "gc", "resid_rstr", "coarse_rhs", "corr")


def tune_summary(rec):
    """The autotuner's rows, per (kernel, box size, size class): the shape
    chosen and the microseconds per launch of shapes 0..3 (None: not on that
    kernel's menu) over the boxes timed."""
    out = {}
    for r in rec:
        k = "%s_nc%d_class%d" % (TUNE_KERNELS[r["kernel"]], r["nc"], r["size_class"])
        e = out.setdefault(k, {"boxes": r["boxes"], "us": [None] * 4, "chosen": 0})
        e["us"][r["shape"]] = round(r["us"], 2)
        if r["chosen"]:
            e["chosen"] = r["shape"]
    return out


def profile_block(T, step_block, steps_amr):
    """One more block with the library's per-kind event timers on: per kind
    the launches, ms per step, algorithmic GB/s and share of the kernel time;
    the total kernel ms and algorithmic bytes of the block and its wall time."""
    T.prof_enable(True)
    T.sync()
    t1 = time.perf_counter()
    step_block()
    T.sync()
    t_prof = time.perf_counter() - t1
    kinds = {}
    tot_ms = 0.0
    tot_bytes = 0.0
    for k, name in STEP_KINDS.items():
        n, ms, by = T.prof_read(k, -999)
        if n:
            kinds[name] = {"launches": n, "ms_per_step": round(ms / steps_amr, 3),
                           "alg_GBps": round(by / (ms * 1e-3) / 1e9, 1) if by else None}
            tot_ms += ms
            tot_bytes += by
    T.prof_enable(False)
    for v in kinds.values():
        v["share"] = round(v["ms_per_step"] * steps_amr / tot_ms, 4)
    return kinds, tot_ms, tot_bytes, t_prof


# the production 3D streamer: a 64 mm cube of one base box (the reference's
# init_tree), uniform level 6 (0.25 mm cells) and level 7 in the top quarter
# where the streamer starts, the seed refining further
PRODUCTION_L = 64e-3
PRODUCTION_REGIONS = [(6, [0.0, 0.0, 0.0], [PRODUCTION_L] * 3),
                      (7, [0.0, 0.0, 0.75 * PRODUCTION_L], [PRODUCTION_L] * 3)]


def streamer_step(args, steps=None, warmup=None):
    """The full 3D streamer step on one GPU: the 3D driver (a 64 mm cube of
    8^3-cell boxes on a base of step_base^3 boxes, 2 MV/m, the air seed and
    transport of the 2D benchmark moved to (32, 32, 56) mm, Helmholtz
    photoionization) after its initial refinement, stepped through the
    library. With one base box the mesh is refined at least to level 6
    everywhere and to level 7 in the top quarter (the production size, tens
    of millions of leaf cells) and the initial densities are evaluated on the
    device. Returns one JSON-able dict: leaf Mcell-steps/s, the leaf cells at
    the start and end, the whole step's algorithmic bytes over its time as a
    fraction of the HBM peak, and per kernel class the time share and rate."""
    import a5_streamer_amd as A
    from a5_streamer_amd import streamer as ST
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    production = args.step_base == 1
    cfg = ST.Config(variant="3d", domain_len=PRODUCTION_L, n_base=args.step_base,
                    lvls_max=args.step_lvls,
                    applied_fld=2.0e6, seed_r0=[[32e-3, 32e-3, 56e-3]],
                    seed_r1=[[32e-3, 32e-3, 56e-3]], photoi=not args.no_photoi,
                    photoi_warm=not args.photoi_cold, n_boxes=20000,
                    refine_regions=PRODUCTION_REGIONS if production else [],
                    init_on_device=production)
    run = ST.StreamerRun(A.Tree, cfg)
    t0 = time.perf_counter()
    run.initialise()
    t_init = time.perf_counter() - t0
    cells = run.leaf_cells()
    print("streamer step: initial refinement done in %.1f s, %d leaf cells, %d boxes" % (
        t_init, cells, run.T.info()["max_id"]), file=sys.stderr, flush=True)
    blocks = max(1, steps // cfg.steps_amr)
    tune = run.T.autotune(reps=10) if args.autotune else []
    for _ in range(warmup):
        run.step_block()
    T = run.T
    T.sync()
    c0 = run.leaf_cells()
    cyc0 = run.photoi_cycles
    t0 = time.perf_counter()
    for _ in range(blocks):
        run.step_block()
    T.sync()
    dt = time.perf_counter() - t0
    cyc1 = run.photoi_cycles
    c1 = run.leaf_cells()
    n_steps = blocks * cfg.steps_amr
    kinds, tot_ms, tot_bytes, t_prof = profile_block(T, run.step_block, cfg.steps_amr)
    ms_step = dt * 1e3 / n_steps
    # algorithmic bytes of one step (the profiled block's, per step) over the
    # time per step of the timed blocks
    step_bytes = tot_bytes / cfg.steps_amr
    step_ach = step_bytes / (ms_step * 1e-3) / 1e9
    out = {
        "metric": "3D streamer step, leaf Mcell-steps/s",
        "value": round(0.5 * (c0 + c1) * n_steps / dt / 1e6, 3),
        "unit": "Mcell-steps/s", "steps": n_steps, "warmup_blocks": warmup,
        "ms_per_step": round(ms_step, 3), "dtype": "f64",
        "config": {"workload": "3D streamer: 64 mm cube, 8^3-cell boxes, %d^3 base box(es), "
                               "2 MV/m, seed at (32, 32, 56) mm, lvls_max %d, %s%s" % (
                                   args.step_base, args.step_lvls,
                                   "uniform level 6, level 7 for z > 48 mm, "
                                   if production else "",
                                   "Helmholtz photoionization"
                                   if cfg.photoi else "no photoionization"),
                   "leaf_cells_start": c0, "leaf_cells_end": c1, "leaf_cells_init": cells,
                   "boxes": T.info()["max_id"], "init_s": round(t_init, 1)},
        "roofline": {"bound": "hbm", "step_bytes": round(step_bytes),
                     "step_achieved": round(step_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "step_frac": round(step_ach / HBM_PEAK_GBS, 4)},
        "photoi": {"start": "cold (FMG)" if args.photoi_cold else "warm (last step's)",
                   "v_cycles_per_step": round((cyc1 - cyc0) / n_steps, 2)}
        if cfg.photoi else None,
        "kernels": kinds,
        "profiled_block_s": round(t_prof, 3),
        "kernel_ms_per_step": round(tot_ms / cfg.steps_amr, 3),
        "autotune": tune_summary(tune),
    }
    T.close()
    return out


def batch_runs(args, n=None, blocks=None, sample=None):
    """Many cylindrical runs in one tree (BatchRun over a5g_set_instances)
    against the same runs stepped one by one, each on its own tree (the first
    `sample` of them, whose rate stands for all: the runs one by one take
    the sum of their times). Returns the line's dict: aggregate leaf
    Mcell-steps/s of both and their ratio."""
    import a5_streamer_amd as A
    from a5_streamer_amd import streamer as ST
    from a5_streamer_amd.workloads import BATCH_REGIONS, batch_configs
    n = n or args.batch
    regions = BATCH_REGIONS if args.batch_region else None
    cfgs = batch_configs(n, args.batch_lvls, regions, init_on_device=True)
    blocks = blocks or max(1, args.steps // cfgs[0].steps_amr)
    sample = min(n, sample or args.batch_sample or n)
    t0 = time.perf_counter()
    B = ST.BatchRun(A.Tree, cfgs)
    B.initialise()
    t_init = time.perf_counter() - t0
    print("batch of %d runs initialised in %.1f s" % (n, t_init), file=sys.stderr, flush=True)
    # the autotuner times shapes on the settled mesh; the warmup blocks then
    # capture the graphs with the chosen shapes
    tune = B.T.autotune(reps=10) if args.autotune else []
    for _ in range(args.warmup):
        B.step_block()
    B.T.sync()
    cells = []
    t0 = time.perf_counter()
    for _ in range(blocks):
        cells.append(B.total_leaf_cells())
        B.step_block()
    B.T.sync()
    dt_b = time.perf_counter() - t0
    steps = blocks * cfgs[0].steps_amr
    v_b = sum(cells) * cfgs[0].steps_amr / dt_b / 1e6
    kinds, tot_ms, tot_bytes, t_prof = profile_block(B.T, B.step_block, cfgs[0].steps_amr)
    step_bytes = tot_bytes / cfgs[0].steps_amr
    step_ach = step_bytes / (dt_b / steps) / 1e9
    nb = B._boxes()
    per_run = np.bincount(nb["run"][nb["in_use"] != 0], minlength=n)
    boxes = int((nb["in_use"] != 0).sum())
    B.T.close()
    print("batch timed: %.1f ms per step" % (dt_b * 1e3 / steps), file=sys.stderr, flush=True)
    # the same runs one by one (sample < 0: not timed)
    dt_s, cells_s = 0.0, 0
    for c in cfgs[:max(sample, 0)]:
        r = ST.StreamerRun(A.Tree, c)
        r.initialise()
        for _ in range(args.warmup):
            r.step_block()
        r.T.sync()
        t0 = time.perf_counter()
        for _ in range(blocks):
            cells_s += r.leaf_cells() * c.steps_amr
            r.step_block()
        r.T.sync()
        dt_s += time.perf_counter() - t0
        r.T.close()
    v_s = cells_s / dt_s / 1e6 if dt_s > 0 else None
    return {
        "metric": "batched cylindrical runs, aggregate leaf Mcell-steps/s",
        "value": round(v_b, 3), "unit": "Mcell-steps/s", "steps": steps,
        "warmup_blocks": args.warmup, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d cylindrical runs of 16 x 32 mm, 1 x 2 base boxes of 8 x 8 "
                               "cells, lvls_max %d, SplitMix64 seed 20261203 + k: field "
                               "1.5-2.5 MV/m, seed width 0.2-0.8 mm, background 1e14-1e16%s"
                               % (n, args.batch_lvls,
                                  ", refine regions %s" % regions if regions else ""),
                   "boxes_in_batch_tree": boxes,
                   "boxes_per_run": [int(per_run.min()), int(np.median(per_run)),
                                     int(per_run.max())],
                   "leaf_cells_start": cells[0], "init_s": round(t_init, 1)},
        "ms_per_step": round(dt_b * 1e3 / steps, 3),
        "one_by_one": {"value": round(v_s, 3), "runs_timed": sample,
                       "ms_per_step_all_runs": round(dt_s * n / sample * 1e3 / steps, 3)}
        if v_s else None,
        "speedup_vs_one_by_one": round(v_b / v_s, 2) if v_s else None,
        "autotune": tune_summary(tune),
        "roofline": {"bound": "hbm", "step_bytes": round(step_bytes),
                     "step_achieved": round(step_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "step_frac": round(step_ach / HBM_PEAK_GBS, 4)},
        "kernels": kinds,
        "profiled_block_s": round(t_prof, 3),
        "kernel_ms_per_step": round(tot_ms / cfgs[0].steps_amr, 3),
    }


# the adjoint workload: the 2D air streamer (4 x 4 base boxes of 8 x 8
# cells on 32 mm, 2.5 MV/m, seed at (16, 28) mm) on a frozen mesh of about
# 1.5M leaf cells: uniform level 6 and level 7 over a 10 x 14 mm block
# around the seed and the channel below it
ADJ_REGIONS = [(6, [0.0, 0.0], [32e-3, 32e-3]), (7, [11e-3, 18e-3], [21e-3, 32e-3])]


def adjoint_oracle(run, win, cfg, cells, k=2):
    """The oracle's FP64 reverse mode (the same FrozenWindow code over the
    oracle's tree) on the first k steps of the adjoint window: the same
    frozen mesh and the same time steps. Returns its forward + backward rate
    on one host core (the adjoint's CPU baseline, a bounded sample of the
    window) and the library's gradient of the same k-step window against
    it: the forward states agree bit for bit, the gradients to the order of
    the library's atomic sums."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from a5_streamer_amd import adjoint as AD
    from a5_streamer_amd import streamer as St
    win.reset()
    gw = AD.FrozenWindow(run, k, every=k, dts=win.dts[:k])
    t0 = time.perf_counter()
    ref = St.StreamerRun(O.OStreamTree, cfg)
    ref.initialise()
    t_init = time.perf_counter() - t0
    rw = AD.FrozenWindow(ref, k, every=k, dts=win.dts[:k])
    res = []
    for w in (gw, rw):
        geo = AD.LeafGeometry(w.T)
        t1 = time.perf_counter()
        w.forward()
        t_f = time.perf_counter() - t1
        ne = torch.from_numpy(w.T.get_cc_all(St.I_ELEC)).requires_grad_(True)
        E = torch.from_numpy(w.T.get_cc_all(St.I_FLD)).requires_grad_(True)
        AD.window_loss(ne, E, geo, cfg.domain_len, cfg.applied_fld)[0].backward()
        t1 = time.perf_counter()
        g = w.backward(ne.grad.numpy(), E.grad.numpy())
        res.append((g, ne.detach().numpy(), t_f, time.perf_counter() - t1))
    (gt, gv, gn, _), ne_g, _, _ = res[0]
    (rt, rv, rn, _), ne_r, tf, tb = res[1]
    del rw, ref

    def rel(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())

    return {
        "cpu_baseline": {
            "value": round(cells * k / (tf + tb) / 1e6, 4), "unit": "Mcell-steps/s", "cores": 1,
            "kind": "port",
            "sample": "the oracle's FP64 reverse mode (C, one thread) on the first %d steps of "
                      "the window, forward %.2f s + backward %.2f s (tree set-up %.1f s "
                      "not counted)" % (k, tf, tb, t_init)},
        "oracle_check": {"steps": k, "forward_bit_identical": bool(np.array_equal(ne_g, ne_r)),
                         "table_grad_rel": rel(gt, rt), "voltage_grad_rel": abs(gv - rv) / abs(rv),
                         "n_e_start_grad_rel": rel(gn, rn)},
    }


def adjoint_window(args, steps=100, every=10, fd_dirs=20):
    """Gradient of the adjoint workload's loss (electron count + softmax
    head position after `steps` steps on the frozen mesh) with respect to
    the transport table, the mobility and the applied voltage: forward pass
    with a checkpoint every `every` steps, backward pass by
    checkpoint-recompute, and central differences (h = 1e-6 relative) on
    fd_dirs seeded directions. Per direction: the difference's error
    relative to the directional derivative's absolute scale, and the gap
    between its one-sided differences, |f(h) + f(-h) - 2 f(0)| / (h scale),
    which is h f'' / scale on a smooth stretch and about the slope jump of
    a limiter or clamp kink that the pair straddles."""
    import torch
    import a5_streamer_amd as A
    from a5_streamer_amd import adjoint as AD
    from a5_streamer_amd import streamer as St
    cfg = St.Config("2d", adjoint=True, refine_regions=ADJ_REGIONS, n_boxes=40000)
    t0 = time.perf_counter()
    run = St.StreamerRun(A.Tree, cfg)
    run.initialise()
    T = run.T
    if args.autotune:
        T.autotune(reps=5)
    win = AD.FrozenWindow(run, steps, every=every)
    init_s = time.perf_counter() - t0
    cells = int(sum(len(T.level(l)[1]) for l in range(1, T.info()["max_lvl"] + 1))) * 64
    geo = AD.LeafGeometry(T)
    win.forward()                      # records the time steps (warm-up)
    T.sync()
    t1 = time.perf_counter()
    win.forward()
    T.sync()
    t_fwd = time.perf_counter() - t1
    ne = torch.from_numpy(T.get_cc_all(St.I_ELEC)).requires_grad_(True)
    E = torch.from_numpy(T.get_cc_all(St.I_FLD)).requires_grad_(True)
    loss, count, head = AD.window_loss(ne, E, geo, cfg.domain_len, cfg.applied_fld)
    loss0 = float(loss.detach())
    loss.backward()
    t1 = time.perf_counter()
    gt, gv, _, _ = win.backward(ne.grad.numpy(), E.grad.numpy())
    T.sync()
    t_bwd = time.perf_counter() - t1
    # the gradient with respect to the scalar mobility and the alpha column
    g_mob = float(np.sum(gt[St.T_MOB - 1]))
    rng = np.random.default_rng(20261204)
    tb0, v0 = win.table0, win.voltage0
    fd = []
    for k in range(fd_dirs):
        dtb = tb0 * rng.standard_normal(tb0.shape)
        dv = v0 * rng.standard_normal()
        vals = []
        for sgn in (1.0, -1.0):
            win.forward(tb0 + sgn * 1e-6 * dtb, v0 + sgn * 1e-6 * dv)
            ne_ = torch.from_numpy(T.get_cc_all(St.I_ELEC))
            E_ = torch.from_numpy(T.get_cc_all(St.I_FLD))
            vals.append(float(AD.window_loss(ne_, E_, geo, cfg.domain_len, cfg.applied_fld)[0]))
        f = (vals[0] - vals[1]) / 2e-6
        a = float(np.sum(gt * dtb)) + gv * dv
        sc = float(np.sum(np.abs(gt * dtb))) + abs(gv * dv)
        fd.append((abs(f - a) / sc, abs(vals[0] + vals[1] - 2 * loss0) / (1e-6 * sc)))
    errs = np.array([e for e, _ in fd])
    slots = win.store_slots()
    n_boxes = int(np.sum(T.boxes()["in_use"]))
    # the same window with every state kept in HBM (no recompute)
    win.reset()
    wk = AD.FrozenWindow(run, steps, keep_all=True, dts=win.dts)
    wk.forward()
    T.sync()
    t1 = time.perf_counter()
    wk.forward()
    T.sync()
    tk_f = time.perf_counter() - t1
    t1 = time.perf_counter()
    gk, gvk, _, _ = wk.backward(ne.grad.numpy(), E.grad.numpy())
    T.sync()
    tk_b = time.perf_counter() - t1
    keep_all = {
        "value": round(cells * steps / (tk_f + tk_b) / 1e6, 3),
        "forward_ms_per_step": round(tk_f * 1e3 / steps, 3),
        "backward_ms_per_step": round(tk_b * 1e3 / steps, 3),
        "store_slots": wk.store_slots(),
        "grad_rel_to_checkpointed": float(np.abs(gk - gt).max() / np.abs(gt).max()),
    }
    oracle_leg = {} if args.no_cpu_baseline else adjoint_oracle(run, wk, cfg, cells)
    T.close()
    return {
        "metric": "adjoint of the 2D streamer on a frozen mesh, leaf Mcell-steps/s of "
                  "forward + backward",
        "value": round(cells * steps / (t_fwd + t_bwd) / 1e6, 3),
        "unit": "Mcell-steps/s", "dtype": "f64",
        "config": {"workload": "2D air streamer (32 mm, 4 x 4 base boxes of 8 x 8 cells, "
                               "2.5 MV/m), frozen mesh: uniform level 6 + level 7 on "
                               "[11, 21] x [18, 32] mm, %d steps, checkpoint every %d"
                               % (steps, every),
                   "leaf_cells": cells, "boxes": n_boxes, "init_s": round(init_s, 1)},
        "forward_ms_per_step": round(t_fwd * 1e3 / steps, 3),
        "backward_ms_per_step": round(t_bwd * 1e3 / steps, 3),
        "backward_over_forward": round(t_bwd / t_fwd, 2),
        "store_slots": slots,
        "keep_all_states": keep_all,
        "loss": {"electron_count": float(count.detach()), "head_position_m": float(head.detach())},
        "grad": {"voltage": gv, "mobility": g_mob,
                 "alpha_table_norm": float(np.linalg.norm(gt[St.T_ALPHA - 1]))},
        "central_differences": {
            "directions": fd_dirs, "h": 1e-6,
            "rel_err_of_scale": {"max": float(errs.max()), "median": float(np.median(errs))},
            "per_direction": [{"rel_err": float("%.3g" % e), "one_sided_gap": float("%.3g" % g)}
                              for e, g in fd]} if fd_dirs else None,
        **oracle_leg,
    }


def level_breakdown(args):
    """Per level of the benchmark octree: launches and event-timed ms of each
    kernel class in one FP64 solve (FMG + the V-cycles to 1e-10) and in one
    mixed solve (FP64 defect correction around FP32 V-cycles), with the wall
    time of both (the mixed form's FP64 outer passes are kind 6)."""
    from a5_streamer_amd.workloads import OctreePoisson
    W = OctreePoisson()
    T = build_gpu(W, n_cc=4)
    max_lvl = T.info()["max_lvl"]
    min_lvl = T.info()["min_lvl"]
    tol = 1e-10
    n_v, _ = converge_count(T, max_lvl, tol)
    tune = T.autotune(reps=10) if args.autotune else []
    legs = {"fp64": lambda: solve(T, max_lvl, n_v),
            "mixed": lambda: T.solve_mixed(tol, 40, N_INNER, True)}
    out = {"metric": "per-level kernel time of one solve, FP64 vs mixed",
           "config": {"workload": "3D octree Poisson of the headline line", "tol": tol,
                      "fp64_v_cycles": n_v, "mixed_inner": N_INNER,
                      "boxes_per_level": {str(l): len(T.level(l)[0])
                                          for l in range(min_lvl, max_lvl + 1)
                                          if len(T.level(l)[0])}},
           "autotune": tune_summary(tune)}
    names = KIND_NAMES + ["fp64_outer"]
    for leg, fn in legs.items():
        for _ in range(2):
            fn()
        wall = timed(T, 3, fn) / 3
        T.prof_enable(True)
        T.sync()
        fn()
        T.sync()
        lv = {}
        for lvl in range(min_lvl, max_lvl + 1):
            row = {}
            for k, name in enumerate(names):
                n, ms, _ = T.prof_read(k, lvl)
                if n:
                    row[name] = [n, round(ms, 4)]
            if row:
                lv[str(lvl)] = row
        tot = {}
        for k, name in enumerate(names):
            n, ms, _ = T.prof_read(k, -999)
            if n:
                tot[name] = [n, round(ms, 4)]
        T.prof_enable(False)
        out[leg] = {"wall_ms": round(wall * 1e3, 3), "levels": lv, "total": tot}
    # outer corrections and wall time of the mixed solve per number of FP32
    # V-cycles in each correction
    sweep = {}
    for n_inner in (1, 2, 3, 4, 6):
        info = {}

        def one():
            info["it"], info["rel"] = T.solve_mixed(tol, 60, n_inner, True)
        one()
        sweep[str(n_inner)] = {"outer": info["it"], "rel_residual": info["rel"],
                               "wall_ms": round(timed(T, 3, one) / 3 * 1e3, 3)}
    out["mixed_inner_sweep"] = sweep
    T.close()
    return out


def main():
    # The JSON line goes to the process's stdout; everything native code
    # writes to file descriptor 1 (the reference library of the CPU baseline
    # reports box storage growth there, buffered until exit) goes to stderr.
    json_out = os.fdopen(os.dup(1), "w")
    sys.stdout.flush()
    os.dup2(2, 1)

    def emit(d):
        json_out.write(json.dumps(d) + "\n")
        json_out.flush()

    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--small", action="store_true", help="smaller octree (testing)")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--streamer-step", action="store_true",
                    help="measure the full 3D streamer step instead (own JSON line)")
    ap.add_argument("--step-lvls", type=int, default=9)
    ap.add_argument("--step-base", type=int, default=1,
                    help="base boxes per direction of the streamer step (1: the reference's "
                         "single base box, refined to the production size)")
    ap.add_argument("--no-streamer3d", action="store_true",
                    help="leave the production 3D streamer step out of the headline line")
    ap.add_argument("--no-photoi", action="store_true")
    ap.add_argument("--photoi-cold", action="store_true",
                    help="solve every photoionization term by FMG from zero each step")
    ap.add_argument("--level-breakdown", action="store_true",
                    help="per-level kernel times of one FP64 and one mixed solve of the "
                         "octree instead (own JSON line)")
    ap.add_argument("--no-batched", action="store_true",
                    help="leave the batched cylindrical runs out of the headline line")
    ap.add_argument("--batch", type=int, default=0,
                    help="measure this many batched cylindrical runs instead (own JSON line)")
    ap.add_argument("--batch-lvls", type=int, default=10)
    ap.add_argument("--batch-region", type=int, default=1,
                    help="1: refine every batched run to a developed streamer mesh (0: the "
                         "reference's criteria only)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: run the launch-shape autotuner on the tree before timing")
    ap.add_argument("--batch-sample", type=int, default=0,
                    help="time this many of the runs one by one (0: all, -1: none)")
    ap.add_argument("--adjoint", action="store_true",
                    help="measure the adjoint workload instead (own JSON line)")
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--no-adjoint", action="store_true",
                    help="leave the adjoint workload out of the headline line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.streamer_step:
        emit(streamer_step(args))
        return
    if args.batch:
        emit(batch_runs(args))
        return
    if args.level_breakdown:
        emit(level_breakdown(args))
        return
    if args.adjoint:
        emit(adjoint_window(args, steps=args.adjoint_steps))
        return

    from a5_streamer_amd.workloads import OctreePoisson
    W = OctreePoisson(uniform_lvl=3, fine_lvl=4, refine_frac=(5, 5, 8)) if args.small \
        else OctreePoisson()
    T = build_gpu(W, n_cc=4)
    info = T.info()
    max_lvl = info["max_lvl"]
    cells = W.leaf_cells(T)
    tol = 1e-10

    # --- FP64 ---------------------------------------------------------------
    n_v, _ = converge_count(T, max_lvl, tol)
    tune = T.autotune(reps=10) if args.autotune else []
    for _ in range(args.warmup):
        solve(T, max_lvl, n_v)
    dt64 = timed(T, args.steps, lambda: solve(T, max_lvl, n_v))
    res, rhs = T.residual_norms()
    prof64 = profiled(T, lambda: solve(T, max_lvl, n_v))
    rel64 = res / rhs
    if not rel64 <= tol:
        raise RuntimeError("timed FP64 solves did not converge: %g" % rel64)
    T.tree_copy_cc(1, 4)
    v64 = cells * args.steps / dt64 / 1e6
    sb = solve_bytes(T, n_v)

    out = {
        "metric": "Mcell-steps/s at \u22641e-8 rel. error, mixed vs FP64 on 1 GPU; "
                  "% HBM roofline",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "data": "synthetic (64 seeded Gaussian charge blobs)",
        "config": {
            "workload": "3D octree Poisson: 2x2x2 base, 16^3-cell boxes, uniform level %d "
                        "+ %dx%dx%d boxes refined to level %d" % (
                            W.uniform_lvl, *W.refine_boxes, W.fine_lvl),
            "leaf_cells": cells, "boxes": info["max_id"], "tol": tol,
        },
        "fp64": {"value": round(v64, 3), "ms_per_step": round(dt64 * 1e3 / args.steps, 3),
                 "v_cycles_per_solve": n_v, "rel_residual": rel64,
                 "roofline": roofline(T, "fp64", lambda: solve(T, max_lvl, n_v)),
                 "kernel_ms_per_step": {k: round(v[1], 3) for k, v in prof64.items()},
                 "autotune": tune_summary(tune)},
    }
    head = "fp64"
    if not args.fp64_only:
        info_mixed = {}
        for _ in range(args.warmup):
            T.solve_mixed(tol, 40, N_INNER, True)

        def one():
            info_mixed["it"], info_mixed["rel"] = T.solve_mixed(tol, 40, N_INNER, True)
        dtm = timed(T, args.steps, one)
        profm = profiled(T, one)
        if not info_mixed["rel"] <= tol:
            raise RuntimeError("mixed solves did not converge: %g" % info_mixed["rel"])
        d, b = T.tree_maxdiff(1, 4)
        err = d / b
        vm = cells * args.steps / dtm / 1e6
        out["mixed"] = {"value": round(vm, 3), "ms_per_step": round(dtm * 1e3 / args.steps, 3),
                        "outer_iterations": info_mixed["it"], "rel_residual": info_mixed["rel"],
                        "rel_potential_error_vs_fp64": err,
                        "speedup_vs_fp64": round(vm / v64, 3),
                        "roofline": roofline(T, "mixed", one),
                        "kernel_ms_per_step": {k: round(v[1], 3) for k, v in profm.items()}}
        if err <= 1e-8 and vm > 1.02 * v64:
            head = "mixed"
    H = out[head]
    out["value"] = H["value"]
    out["unit"] = "Mcell-steps/s"
    out["ms_per_step"] = H["ms_per_step"]
    out["dtype"] = "f64" if head == "fp64" else "f32+f64"
    out["config"]["precision"] = head
    out["roofline"] = H["roofline"]
    # the whole solve against the same peak: its algorithmic bytes (the
    # reference's passes, FP64 words) over the measured time per solve
    ach = sb / (out["fp64"]["ms_per_step"] * 1e-3) / 1e9
    out["fp64"]["roofline"]["step_bytes"] = round(sb)
    out["fp64"]["roofline"]["step_achieved"] = round(ach, 1)
    out["fp64"]["roofline"]["step_frac"] = round(ach / HBM_PEAK_GBS, 4)
    # the nominal peak above, and the rate a plain copy reaches here
    cp = copy_peak(T)
    out["roofline"]["copy_peak_measured"] = cp
    out["roofline"]["frac_of_copy_peak"] = round(out["roofline"]["achieved"] / cp, 4) if cp else None
    T.close()
    if not args.no_cpu_baseline:
        # the FP64 GPU path is bit-identical to the reference, so the
        # reference needs the same number of V-cycles
        if args.small:
            out["cpu_baseline"] = cpu_baseline(W, n_v)
        else:
            Ws = OctreePoisson(uniform_lvl=3, fine_lvl=4, refine_frac=(5, 5, 8))
            Tq = build_gpu(Ws)
            n_vs, _ = converge_count(Tq, Tq.info()["max_lvl"], tol)
            Tq.close()
            out["cpu_baseline"] = cpu_baseline(W, n_v, Ws, n_vs)
    if not args.no_streamer3d and not args.small:
        # the production 3D streamer step, a few timed steps after its
        # initial refinement
        out["streamer3d"] = streamer_step(args, steps=6, warmup=1)
    if not args.no_batched and not args.small:
        # the batched cylindrical runs: 256 runs of about 1,100 boxes each,
        # against a sample of them stepped one by one
        out["batched"] = batch_runs(args, n=256, blocks=2, sample=4)
    if not args.no_adjoint and not args.small:
        # the adjoint workload: 100 steps of the 2D streamer on a frozen mesh
        # of about 1.5M leaf cells, differentiated with checkpoints every 10
        out["adjoint"] = adjoint_window(args, steps=args.adjoint_steps)
    # keys first that the driver reads
    order = ["metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
             "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
             "roofline", "cpu_baseline"]
    line = {k: out[k] for k in order if k in out}
    line.update({k: v for k, v in out.items() if k not in line})
    emit(line)


if __name__ == "__main__":
    main()
