#!/usr/bin/env python3
"""Benchmark: batched acoustic RTM of 64 virtual shots of uneven size on one MI355X.

Workload (one "step" = migrating the whole batch once): 64 virtual shots on
aperture-cropped windows nz = 300, nx ~ U{600..1600} at dx = 5 m, nt ~
U{3000..6000} at dt = 0.5 ms, imaging every k = 2 steps, per-shot images
stacked in FP32.  Inputs are seeded synthetic data uploaded to HBM before the
timed region.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Gcell-updates/s + migrated shots/s (batched/served vs plain GPU), "
          "p50/p99 latency")
BYTES_PER_CELL_UPDATE = 26.0  # acoustic RTM at k=2: 20 B step + 24 B per imaging step / (2k)
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def physical_cores():
    """Physical cores among the CPUs this process may run on (one per
    (package, core id) pair, so SMT siblings count once) -> (cores, cpus)."""
    cpus = sorted(os.sched_getaffinity(0))
    seen = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "core_id") as f:
                core = f.read().strip()
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
        except OSError:
            core, pkg = str(c), "0"
        seen.add((pkg, core))
    return len(seen), len(cpus)


def cpu_worker(threads, budget_cell_updates, full):
    """Child-process side of the CPU baseline (no GPU use): the reference
    engine (wave2d.c compiled from source, oracle/_ref) driven by the oracle's
    C time loop on the first shots of the batched workload, with OpenMP
    placement fixed by the parent's environment; prints one JSON line."""
    from das_amd.workloads import BatchedWindows
    from oracle import oracle as ora

    cfg = BatchedWindows(n_shots=64).build(with_data=True)
    kind = "reference" if ora.ref_lib() is not None else "port"
    done, t, used = 0, 0.0, []
    for i, s in enumerate(cfg.shots):
        g = ora.Grid(s.nx, s.nz, cfg.dx, cfg.dx, cfg.dt, cfg.n_sponge)
        t0 = time.perf_counter()
        ora.migrate_shot(g, s.vel_fwd, s.vel_bwd, s.src, s.six, s.siz, s.data, s.rix, s.riz,
                         cfg.k, use_reference_kernels=(kind == "reference"))
        t += time.perf_counter() - t0
        done += 2 * s.nt * s.nx * s.nz
        used.append(f"shot {i} ({s.nz}x{s.nx}, nt={s.nt})")
        log(f"cpu baseline ({threads} threads): {used[-1]} {t:.1f}s")
        if done >= budget_cell_updates:
            break
    out = {"value": round(done / t / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
           "kind": kind,
           "sample": (f"{len(used)} of 64 shots: " + "; ".join(used) +
                      f"; {done / 1e9:.2f} G cell-updates in {t:.1f} s, per-step kernels "
                      "and snapshots in host RAM as the reference")}
    out.update(cpu_reference_points(kind, elastic=full))
    print(json.dumps(out), flush=True)


def cpu_baseline(budget_cell_updates, budget_one_core):
    """The CPU baseline on the box's host cores: once with one OpenMP thread
    per physical core (OMP_PROC_BIND=close, OMP_PLACES=cores) and once on one
    core, each in a child process so that the placement applies to the
    reference engine's OpenMP team from its start."""
    import subprocess

    cores, cpus = physical_cores()
    runs = {}
    for threads, budget, full in ((cores, budget_cell_updates, True),
                                  (1, budget_one_core, False)):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", str(threads),
               "--cpu-budget", str(budget / 1e9)] + ([] if full else ["--cpu-worker-lite"])
        r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"CPU baseline worker ({threads} threads) failed: {r.returncode}")
        runs[threads] = json.loads(r.stdout.strip().splitlines()[-1])
    out = runs[cores]
    out["placement"] = (f"OMP_NUM_THREADS={cores} (physical cores of the {cpus} CPUs this "
                        "process may use), OMP_PROC_BIND=close, OMP_PLACES=cores")
    one = runs[1]
    out["one_core"] = {"value": one["value"], "unit": one["unit"], "cores": 1,
                       "sample": one["sample"],
                       "acoustic_shot_200x400": one["acoustic_shot_200x400"]}
    out["speedup_vs_one_core"] = round(out["value"] / one["value"], 2)
    return out


def plain_shot_200x400():
    """The 200 x 400 acoustic shot (nt 4000, k 2), the GPU side of the CPU
    baseline's acoustic_shot_200x400, on two GPU paths, in ms (second of two
    runs each):
      * ``migrate_shot_ms``: TSWRTMEngine.migrate_shot from host arrays
        (tsw_migrate_shot, copies included), which runs the shot through the
        temporally blocked kernel as a one-shot batch;
      * ``per_step_ms``: the same shot resident, one launch of the per-step
        kernel per time step (no graph)."""
    import numpy as np
    import torch

    from das_amd.batch import ShotBatch
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.velocity_flood import FloodModels
    from das_amd.wave_solver import WaveSolver2D
    from das_amd.workloads import PlainAcousticShot, Shot

    c0 = PlainAcousticShot()
    solver = WaveSolver2D(c0.nx, c0.nz, c0.dx, c0.dx, c0.dt, n_sponge=c0.n_sponge)
    data0 = (np.random.default_rng(0).standard_normal((c0.nt, len(c0.rix))) * 1e-3
             ).astype(np.float32)
    fm = FloodModels(c0.mig_vel, c0.mig_vel, c0.mig_vel, 0, 0)
    eng = TSWRTMEngine(solver)
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.migrate_shot(fm, c0.src, c0.six, c0.siz, data0, c0.rix, c0.riz, c0.k)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    shot = Shot(nx=c0.nx, nz=c0.nz, nt=c0.nt, vel_fwd=c0.mig_vel, vel_bwd=c0.mig_vel,
                src=c0.src, six=c0.six, siz=c0.siz, data=data0, rix=c0.rix, riz=c0.riz)
    b = ShotBatch([shot], c0.dx, c0.dx, c0.dt, n_sponge=c0.n_sponge, k=c0.k, slots=1,
                  graph_steps=0, kernel=0)
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run("plain")
        torch.cuda.synchronize()
        ep = time.perf_counter() - t0
    b.close()
    return {"migrate_shot_ms": round(el * 1e3, 2), "per_step_ms": round(ep * 1e3, 2),
            "note": "migrate_shot: host arrays in and out, blocked kernel as a one-shot batch; "
                    "per_step: resident, one per-step kernel launch per time step"}


def check_against_per_step(cfg, batch, shots, log):
    """Shots of the timed batch against the same shots migrated alone with the
    per-step kernel (one launch per time step): bit-for-bit equality."""
    from das_amd.batch import ShotBatch

    out = {}
    for i in shots:
        s = cfg.shots[i]
        b = ShotBatch([s], cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge, k=cfg.k, slots=1,
                      graph_steps=100, kernel=0)
        b.run("batched")
        ref = b.shot_result(0)
        b.close()
        got = batch.shot_result(i)
        same = bool(np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]))
        diff = float(np.linalg.norm(got[0].astype(np.float64) - ref[0]) /
                     max(np.linalg.norm(ref[0].astype(np.float64)), 1e-300))
        out[f"shot_{i}"] = {"grid": f"{s.nz}x{s.nx}", "nt": s.nt, "bit_equal": same,
                            "image_rel_l2": diff}
        log(f"check: shot {i} ({s.nz}x{s.nx}, nt {s.nt}) bit-equal to per-step: {same}")
    return out


def cpu_reference_points(kind, elastic=True):
    """The same CPU engine on the single-shot workloads: the 200 x 400 acoustic
    shot (nt 4000, k 2) migrated in full, and (``elastic``) the forward of the
    400 x 1600 elastic shot over its first 800 of 8000 steps (the oracle's
    restatement: the reference elastic kernel reads a row above the grid)."""
    import numpy as np

    from das_amd.workloads import PlainAcousticShot, PlainElasticShot
    from oracle import oracle as ora

    c0 = PlainAcousticShot()
    g0 = ora.Grid(c0.nx, c0.nz, c0.dx, c0.dx, c0.dt, c0.n_sponge)
    data0 = (np.random.default_rng(0).standard_normal((c0.nt, len(c0.rix))) * 1e-3
             ).astype(np.float32)
    t0 = time.perf_counter()
    ora.migrate_shot(g0, c0.mig_vel, c0.mig_vel, c0.src, c0.six, c0.siz, data0, c0.rix, c0.riz,
                     c0.k, use_reference_kernels=(kind == "reference"))
    t_c0 = time.perf_counter() - t0
    pt0 = {"acoustic_shot_200x400": {"s": round(t_c0, 3), "kind": kind, "gcell_updates_per_s":
                                     round(2 * c0.nt * c0.nx * c0.nz / t_c0 / 1e9, 4)}}
    if not elastic:
        return pt0
    c1 = PlainElasticShot(nt=800)
    g1 = ora.Grid(c1.nx, c1.nz, c1.dx, c1.dx, c1.dt, c1.n_sponge)
    t0 = time.perf_counter()
    ora.elastic_forward(g1, c1.model_true, c1.src, c1.six, c1.siz, c1.rix, c1.riz)
    t_c1 = time.perf_counter() - t0
    log(f"cpu: acoustic 200x400 shot {t_c0:.2f}s, elastic forward 800 steps {t_c1:.2f}s")
    pt0["elastic_forward_400x1600"] = {"steps": c1.nt, "s": round(t_c1, 3), "kind": "port",
                                       "gcell_updates_per_s":
                                           round(c1.nt * c1.nx * c1.nz / t_c1 / 1e9, 4)}
    return pt0


ELASTIC_BYTES_PER_CELL_UPDATE = 60.0  # 5 fields read + written, 5 parameters read


def elastic_leg(args, log):
    """Plain 2D elastic P-SV RTM of one shot (400 x 1600, nt 8000, dt 0.4 ms,
    800 strain-rate receivers): data modelled in the true model, then the
    migration timed as a resident one-shot batch (hipGraph-captured steps)."""
    import torch

    from das_amd.batch import ShotBatch
    from das_amd.wave_solver import WaveSolver2D
    from das_amd.workloads import PlainElasticShot, ElasticShot

    cfg = PlainElasticShot()
    solver = WaveSolver2D(cfg.nx, cfg.nz, cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge,
                          solver_type="elastic")
    t0 = time.time()
    data, _ = solver.forward_propagate_elastic(cfg.model_true, cfg.src, cfg.six, cfg.siz,
                                               cfg.rix, cfg.riz)
    log(f"elastic data modelled in {time.time() - t0:.1f}s")
    shot = ElasticShot(nx=cfg.nx, nz=cfg.nz, nt=cfg.nt, model_fwd=cfg.model_mig,
                       model_bwd=cfg.model_mig, src=cfg.src, six=cfg.six, siz=cfg.siz,
                       data=data, rix=cfg.rix, riz=cfg.riz)
    b = ShotBatch([shot], cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge, k=cfg.k)
    b.run("batched")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run("batched")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms, nl = b.time_kernels("elastic")
    b.close()
    cu = cfg.cell_updates
    ach = ELASTIC_BYTES_PER_CELL_UPDATE * cu / (ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_ela_fused (velocity and stress passes, one launch per step)",
            "launches": nl, "avg_launch_us": round(ms * 1e3 / max(nl, 1), 3),
            "algorithmic_bytes_per_cell_update": ELASTIC_BYTES_PER_CELL_UPDATE}
    tp = os.path.join(ROOT, "profiles", "elastic_step_hbm_traffic.json")
    if os.path.exists(tp):
        # HBM bytes the step kernel moves per cell-update (rocprofv3 FETCH_SIZE +
        # WRITE_SIZE on a shortened run of this shot), at this run's kernel time
        with open(tp) as f:
            per = json.load(f)["step_total_B_per_cell_update"]
        gbs = per * cu / (ms * 1e-3) / 1e9
        roof["traffic"] = {"bytes_per_cell_update": per, "gbs": round(gbs, 1),
                           "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                           "source": "profiles/elastic_step_hbm_traffic.json"}
    return {"workload": "elastic P-SV RTM, 1 shot, 400x1600, nt=8000, dt=0.4 ms, 800 receivers",
            "value": round(cu / el / 1e9, 3), "unit": "Gcell-updates/s",
            "ms_per_shot": round(el * 1e3, 2), "roofline": roof}


def served_leg(args, acoustic_gcell, elastic_gcell, kernel, log):
    """Served RTM: a seeded Poisson stream of virtual-shot requests (nz in
    {256, 384, 512} x nx in {512, 1024, 2048}, nt ~ U{2000..8000}, 80% acoustic /
    20% elastic) against a 200 GB HBM budget.  The arrival rate is stated as a
    fraction of the capacity implied by the batched acoustic and elastic
    throughputs measured above."""
    from das_amd.serve import RTMServer, run_poisson
    from das_amd.workloads import ServedStream

    cfg = ServedStream(args.served_requests)
    t_req = np.mean([2 * s.nt * s.cells / 1e9 /
                     (elastic_gcell if s.kind == "elastic" else acoustic_gcell)
                     for s in cfg.specs])
    capacity = 1.0 / t_req
    rate = args.served_load * capacity
    log(f"served: {len(cfg.specs)} requests at {rate:.2f}/s "
        f"({args.served_load} x capacity {capacity:.2f}/s)")
    srv = RTMServer(cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge, k=cfg.k, budget_bytes=200e9,
                    max_active=args.served_slots, kernel=kernel, tick_steps=100)
    srv.set_models(cfg.vel, cfg.vel, cfg.ela, cfg.ela)  # one shared resident model
    r = run_poisson(srv, cfg.specs, rate, seed=11, make=lambda sp: cfg.make(sp, resident=True),
                    log=log)
    srv.close()
    n_el = sum(s.kind == "elastic" for s in cfg.specs)
    # algorithmic HBM traffic of the stream: 26 B per acoustic RTM cell-update
    # (k=2), 60 B per elastic cell-update, over the span of the run
    traffic = sum(2 * s.nt * s.cells * (ELASTIC_BYTES_PER_CELL_UPDATE if s.kind == "elastic"
                                        else BYTES_PER_CELL_UPDATE) for s in cfg.specs)
    return {"workload": f"{len(cfg.specs)} Poisson-arrival virtual-shot migrations, "
                        f"{len(cfg.specs) - n_el} acoustic / {n_el} elastic, nz {{256,384,512}} "
                        "x nx {512,1024,2048}, nt 2000-8000, dx 5 m, dt 0.4 ms, 200 GB budget, "
                        "windows of one shared resident model",
            "arrival_rate_per_s": round(rate, 3), "load_vs_capacity": args.served_load,
            "capacity_estimate_per_s": round(capacity, 3),
            "p50_latency_s": round(r["p50_latency_s"], 4),
            "p99_latency_s": round(r["p99_latency_s"], 4),
            "shots_per_s": round(r["shots_per_s"], 3),
            "gcell_updates_per_s": round(r["gcell_updates_per_s"], 2),
            "hbm_gbs_algorithmic": (round(traffic / r["span_s"] / 1e9, 1)
                                    if r["failed"] == 0 and r["span_s"] else None),
            "failed": r["failed"], "span_s": round(r["span_s"], 2),
            "gpu_busy_s": round(r["gpu_busy_s"], 2), "max_active": args.served_slots,
            "kernel": kernel}


FWI_BYTES_PER_SHOT_CELL_STEP = 76.0  # forward 20 B + recompute 24 B + adjoint 32 B (FP32)


def fwi_leg(args, log):
    """Misfit gradient (forward, checkpointed recompute and discrete adjoint)
    of 16 shots on 300 x 800, nt 5000, FP32, timed on resident inputs.  With
    --fwi-check, also the FP64 central-difference check along 20 seeded smooth
    directions and the correlation of the gradient with the stacked RTM image
    of the residuals."""
    import torch

    from das_amd.fwi import MisfitShot, WaveformMisfit
    from das_amd.workloads import GradientShots

    cfg = GradientShots()
    src = cfg.source()
    geo = cfg.geometry()

    def make_op(dtype, d_obs=None):
        npdt = np.float64 if dtype == torch.float64 else np.float32
        shots = [MisfitShot(src=src, six=six, siz=siz, rix=rix, riz=riz,
                            d_obs=(np.zeros((cfg.nt, len(rix)), npdt) if d_obs is None
                                   else d_obs[i]))
                 for i, (six, siz, rix, riz) in enumerate(geo)]
        return WaveformMisfit(cfg.nx, cfg.nz, cfg.dx, cfg.dx, cfg.dt, shots,
                              n_sponge=cfg.n_sponge, dtype=dtype)

    def observed(dtype):
        op = make_op(dtype)
        op.forward_loss(torch.tensor(cfg.true_vel, dtype=dtype, device="cuda"))
        d = [op.predicted(i) for i in range(len(geo))]
        op.close()
        return d

    t0 = time.time()
    d32 = observed(torch.float32)
    op = make_op(torch.float32, d32)
    if args.fwi_blocking >= 0:
        op.set_blocking(args.fwi_blocking)
    if args.fwi_adjoint_blocking >= 0:
        op.set_adjoint_blocking(args.fwi_adjoint_blocking)
    v = torch.tensor(cfg.start_vel, dtype=torch.float32, device="cuda")
    log(f"fwi: data modelled, operator resident in {time.time() - t0:.1f}s")
    op.forward_loss(v)
    op.gradient(v)
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        loss = op.forward_loss(v)
        g32 = op.gradient(v)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    steps = len(geo) * cfg.nt * cfg.cells()
    ach = FWI_BYTES_PER_SHOT_CELL_STEP * steps / el / 1e9
    out = {"workload": "misfit gradient, 16 shots, 300x800, nt=5000, dx=5 m, dt=0.5 ms, "
                       "400 receivers, FP32, sqrt(nt)=71-step checkpoints + recompute",
           "gradient_s": round(el, 4), "shots_per_s": round(len(geo) / el, 2),
           "gcell_updates_per_s": round(3 * steps / el / 1e9, 2),
           "effective_gbs": round(ach, 1), "frac_of_hbm_peak": round(ach / HBM_PEAK_GBS, 4),
           "algorithmic_bytes_per_shot_cell_step": FWI_BYTES_PER_SHOT_CELL_STEP,
           "blocking": args.fwi_blocking, "adjoint_blocking": args.fwi_adjoint_blocking,
           "loss": loss}
    tp = os.path.join(ROOT, "profiles", "gradient_kernels_hbm_traffic.json")
    if os.path.exists(tp) and args.fwi_blocking < 0 and args.fwi_adjoint_blocking < 0:
        # HBM bytes the default kernels move per shot-cell-step (rocprofv3 FETCH_SIZE +
        # WRITE_SIZE on a shortened run of this workload), at this run's time
        with open(tp) as f:
            per = json.load(f)["gradient_total_B_per_shot_cell_step"]
        out["traffic"] = {"bytes_per_shot_cell_step": per,
                          "gbs": round(per * steps / el / 1e9, 1),
                          "frac_of_hbm_peak": round(per * steps / el / 1e9 / HBM_PEAK_GBS, 4),
                          "source": "profiles/gradient_kernels_hbm_traffic.json"}
    log(f"fwi: gradient in {el:.3f}s")
    if args.fwi_check:
        out["check"] = fwi_check(cfg, geo, src, make_op, observed, op, g32, v, d32, log)
    op.close()
    return out


def fwi_check(cfg, geo, src, make_op, observed, op32, g32, v32, d32, log):
    import torch
    from scipy import ndimage

    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.velocity_flood import FloodModels
    from das_amd.wave_solver import WaveSolver2D

    t0 = time.time()
    d64 = observed(torch.float64)
    op = make_op(torch.float64, d64)
    v0 = cfg.start_vel.astype(np.float64)
    v = torch.tensor(v0, device="cuda")
    op.forward_loss(v)
    g = op.gradient(v).cpu().numpy()
    rng = np.random.default_rng(2024)
    errs = []
    for i in range(20):
        phi = ndimage.gaussian_filter(rng.standard_normal((cfg.nz, cfg.nx)), 100.0 / cfg.dx)
        phi /= np.abs(phi).max()
        row = []
        for eps in (1e-3, 1e-4):
            dv = eps * v0 * phi
            jp = op.forward_loss(torch.tensor(v0 + dv, device="cuda"))
            jm = op.forward_loss(torch.tensor(v0 - dv, device="cuda"))
            fd = (jp - jm) / 2.0
            row.append(abs(fd - float(np.sum(g * dv))) / abs(fd))
        errs.append(row)
        log(f"fwi check: direction {i}: rel err {row[0]:.2e} (eps 1e-3), {row[1]:.2e} (1e-4)")
    e = np.array(errs)
    g32n = g32.cpu().numpy().astype(np.float64)
    fp32_vs_fp64 = float(np.linalg.norm(g32n - g) / np.linalg.norm(g))
    op.close()
    # stacked RTM image of the FP32 residuals, migrated in the start model
    solver = WaveSolver2D(cfg.nx, cfg.nz, cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge)
    eng = TSWRTMEngine(solver)
    fm = FloodModels(cfg.start_vel, cfg.start_vel, cfg.start_vel, 0, 0)
    img = np.zeros((cfg.nz, cfg.nx))
    op32.forward_loss(v32)
    for i, (six, siz, rix, riz) in enumerate(geo):
        res = op32.predicted(i) - d32[i]
        im, _ = eng.migrate_shot(fm, src, six, siz, res, rix, riz, 2)
        img += im
    m = solver.sponge > 0.99
    m[:3, :] = False
    corr = float(np.corrcoef(g[m], img[m])[0, 1])
    log(f"fwi check done in {time.time() - t0:.1f}s")
    return {"fd_directions": 20, "fd_smoothing_m": 100.0,
            "fd_rel_err_eps1e-3": {"median": float(np.median(e[:, 0])),
                                   "max": float(e[:, 0].max())},
            "fd_rel_err_eps1e-4": {"median": float(np.median(e[:, 1])),
                                   "max": float(e[:, 1].max())},
            "fp32_vs_fp64_gradient_rel_l2": fp32_vs_fp64,
            "corr_gradient_vs_rtm_image_of_residual": corr}


FP64_PEAK_TFLOPS = 78.6


def gather_leg(args, log):
    """Virtual-shot gather construction before migration: correlation of 64
    virtual-source channels with 800 receiver channels over 10 min of 1-bit
    50 Hz data (30000 samples), lags up to 3.5 s (nlag 175), timed on resident
    data; then derivative, causal branch and resampling to 0.5 ms of one
    source's gather (host round trip, reported beside)."""
    import ctypes

    import torch

    from das_amd import _native as N
    from das_amd.interferometry import build_virtual_shot_gather, resample_gather_to_sim_dt

    nt, nch, nsrc, nrec, dt, max_lag = 30000, 1600, 64, 800, 0.02, 3.5
    nlag = int(round(max_lag / dt))
    g = torch.Generator(device="cpu").manual_seed(7)
    data = torch.sign(torch.randn(nt, nch, generator=g)).to(torch.float32).cuda()
    src = (np.linspace(20, nch // 4, nsrc)).astype(np.int32)
    rec = np.arange(nch // 3, nch // 3 + nrec).astype(np.int32)
    out = torch.empty(nsrc, 2 * nlag + 1, nrec, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        N.check(N.lib().tsw_xcorr_device(ctypes.c_void_p(data.data_ptr()), nt, nch, N.ptr(src),
                                         nsrc, N.ptr(rec), nrec, nlag, 0,
                                         ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))

    run()
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    macs = nsrc * nrec * (2 * nlag + 1) * nt
    tflops = 2.0 * macs / el / 1e12
    ccf = out[0].cpu().numpy()
    lags = np.arange(-nlag, nlag + 1) * dt
    t0 = time.perf_counter()
    vs = build_virtual_shot_gather(ccf, lags, 0, int(src[0]), np.arange(nch) * 4.0, rec)
    res = resample_gather_to_sim_dt(vs.shot_gather, vs.dt, 5e-4, 7000)
    post = time.perf_counter() - t0
    log(f"gathers: correlation {el * 1e3:.1f} ms ({tflops:.1f} TFLOP/s FP64)")
    pre = preprocessing_timing(log)
    return {"workload": f"virtual-shot gathers: {nsrc} source x {nrec} receiver channels, "
                        f"{nt} samples at 50 Hz (1-bit), lags -3.5..3.5 s ({2 * nlag + 1})",
            "correlation_ms": round(el * 1e3, 3), "pairs_per_s": round(nsrc * nrec / el, 1),
            "roofline": {"bound": "fp64", "achieved": round(tflops, 2),
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
                         "kernel": "k_xcorr (FP64 lag sums of exact FP32 products)",
                         "flops_per_launch": 2 * macs},
            "derivative_branch_resample_ms": round(post * 1e3, 3),
            "resampled_shape": list(res.shape),
            "preprocessing": pre}


def preprocessing_timing(log):
    """The chain before correlation on one 10-minute block of 800 channels
    sampled at 100 Hz (seeded noise): decimation to 50 Hz (order-8 zero-phase
    low-pass), 0.8-4.5 Hz band-pass (order 4, zero-phase), 1-bit, spectral
    whitening.  Host-pointer calls, so the times include the PCIe copies of
    each call; the zero-phase filters run one channel per thread because
    their FP64 recursion follows scipy's order sample by sample."""
    from das_amd.preprocessing import (bandpass_filter, decimate_time, spectral_whitening,
                                       temporal_normalization)

    nt, nch = 60000, 800
    x = np.random.default_rng(3).standard_normal((nt, nch)).astype(np.float32)

    def chain():
        times = {}
        t0 = time.perf_counter()
        y, dt, fs = decimate_time(x, 100.0, 50.0)
        times["decimate_ms"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        y = bandpass_filter(y, fs, 0.8, 4.5)
        times["bandpass_ms"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        y = temporal_normalization(y)
        times["one_bit_ms"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        spectral_whitening(y, fs, 0.8, 4.5)
        times["whitening_ms"] = time.perf_counter() - t0
        return times

    first = sum(chain().values())  # the first block also makes the FFT plans
    times = chain()
    total = sum(times.values())
    log(f"preprocessing: {total * 1e3:.1f} ms per block of {nch} channels x {nt} samples "
        f"(first block {first * 1e3:.1f} ms)")
    out = {"workload": f"{nch} channels x {nt} samples at 100 Hz (10 min) -> 50 Hz, band-pass "
                       "0.8-4.5 Hz, 1-bit, whitening; host-pointer calls (PCIe included); "
                       "a block after the first"}
    out.update({k: round(v * 1e3, 2) for k, v in times.items()})
    out["total_ms"] = round(total * 1e3, 2)
    out["first_block_ms"] = round(first * 1e3, 2)
    out["input_channel_samples_per_s"] = round(nt * nch / total, 1)
    # the same chain kept on the device (one copy of the block in), then the
    # block's correlations (64 sources x 800 receivers, lags -3.5..3.5 s)
    # added into a resident stack, as stacked_correlations runs them
    import torch

    from das_amd.interferometry import cross_correlations_device
    from das_amd.preprocessing import preprocess_block_device

    srcs = np.linspace(20, nch - 20, 64).astype(np.int32)
    w, dtp, _ = preprocess_block_device(x, 100.0, 50.0, 0.8, 4.5)
    stack, _ = cross_correlations_device(w, srcs, 3.5, dtp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w, dtp, _ = preprocess_block_device(x, 100.0, 50.0, 0.8, 4.5, out=w)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    cross_correlations_device(w, srcs, 3.5, dtp, out=stack)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out["device_chain_ms"] = round((t1 - t0) * 1e3, 2)
    out["device_chain_plus_correlation_ms"] = round((t2 - t0) * 1e3, 2)
    # several blocks per call, as stacked_correlations groups them (8 by default)
    from das_amd.preprocessing import preprocess_blocks_device

    group = 8
    ws, _, _ = preprocess_blocks_device([x] * group, 100.0, 50.0, 0.8, 4.5)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    preprocess_blocks_device([x] * group, 100.0, 50.0, 0.8, 4.5, outs=ws)
    torch.cuda.synchronize()
    out["device_chain_blocks_per_call"] = group
    out["device_chain_grouped_ms_per_block"] = round((time.perf_counter() - t3) * 1e3 / group, 2)
    log(f"preprocessing on the device: {(t1 - t0) * 1e3:.1f} ms per block, "
        f"{(t2 - t0) * 1e3:.1f} ms with its correlations into the stack, "
        f"{out['device_chain_grouped_ms_per_block']:.1f} ms per block {group} at a time")
    return out


def blocked_algorithmic_bytes(shots, T, k):
    """Compulsory HBM bytes of one run of the temporally blocked scheme (T
    steps per launch): per shot, phase and block of T steps each cell reads
    the two carried levels, the model coefficient and its image/illumination
    accumulator (16 B) and writes the two levels and the accumulator (12 B);
    the first block of a phase reads only the coefficient, the last writes
    only the accumulator; plus 4 B per cell per snapshot level, written by the
    forward phase and read by the back-propagation.  Cells that keep the level
    from three buffers back (the grid's outer two columns and bottom two rows)
    move one more level; they are left out (under 2% of the cells)."""
    total = 0
    for s in shots:
        n = s.nx * s.nz
        nb = -(-s.nt // T)
        per_phase = 8 if nb == 1 else 16 + 20 + 28 * (nb - 2)
        nsnap = (s.nt - 1) // k + 1
        total += 2 * n * per_phase + 2 * n * 4 * nsnap
    return total


def counted_traffic(name):
    """HBM bytes per cell-update of a kernel from a committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if present."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--slots", type=int, default=-1, help="-1: chosen by the autotuner")
    ap.add_argument("--graph-steps", type=int, default=20000,
                    help="time steps per hipGraph of the batched run (20000: the whole "
                         "workload in one graph, whose launch groups never join midway)")
    ap.add_argument("--kernel", type=int, default=-1,
                    help="-1: autotuned; 0: one launch per time step; c >= 1: temporally "
                         "blocked config c")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-elastic", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-fwi", action="store_true")
    ap.add_argument("--no-gathers", action="store_true")
    ap.add_argument("--fwi-blocking", type=int, default=-1,
                    help="misfit gradient forward passes: -1 library default, 0 one launch per "
                         "step, c >= 1 temporally blocked configuration c")
    ap.add_argument("--fwi-adjoint-blocking", type=int, default=-1,
                    help="misfit gradient adjoint: -1 library default, 0 one launch per step, "
                         "c >= 1 temporally blocked configuration c")
    ap.add_argument("--fwi-check", action="store_true",
                    help="FP64 finite-difference check of the gradient (about a minute)")
    ap.add_argument("--served-requests", type=int, default=400)
    ap.add_argument("--served-load", type=float, default=0.8)
    ap.add_argument("--served-slots", type=int, default=32)
    ap.add_argument("--cpu-budget", type=float, default=25.0,
                    help="G cell-updates of the CPU baseline sample (about 10 s)")
    ap.add_argument("--cpu-budget-one-core", type=float, default=3.0,
                    help="G cell-updates of the one-core CPU sample (the first shot)")
    ap.add_argument("--no-check", action="store_true",
                    help="skip comparing two shots of the timed batch with the per-step path")
    ap.add_argument("--cpu-worker", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-worker-lite", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args.cpu_worker, args.cpu_budget * 1e9, not args.cpu_worker_lite)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU design "
              "(shots sharded across GPUs, one image reduce) is not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from das_amd import _native as N
    from das_amd.batch import ShotBatch
    from das_amd.workloads import BatchedWindows

    N.check(N.lib().tsw_init(0))
    t0 = time.time()
    cfg = BatchedWindows(n_shots=args.shots).build(with_data=True)
    log(f"workload built in {time.time() - t0:.1f}s")
    tuning = None
    if args.kernel < 0 or args.slots < 0:
        from das_amd import autotune

        t0 = time.time()
        tuning = autotune.tune(cfg.shots, cfg.dx, cfg.dt, cfg.k, cfg.n_sponge, log=log,
                               graph_steps=args.graph_steps)
        log(f"autotune ({time.time() - t0:.1f}s): {tuning}")
        if args.kernel < 0:
            args.kernel = tuning["kernel"]
        if args.slots < 0:
            args.slots = tuning["slots"]
    t0 = time.time()
    batch = ShotBatch(cfg.shots, cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge, k=cfg.k,
                      slots=args.slots, graph_steps=args.graph_steps, kernel=args.kernel)
    info = batch.info()
    log(f"batch resident in HBM in {time.time() - t0:.1f}s: {info}")

    # one step: migrate every shot, then stack the per-shot images and
    # illuminations in FP32 into the extended grid at the window offsets
    gnz, gnx = cfg.nz, cfg.global_nx
    st_img = torch.empty((gnz, gnx), dtype=torch.float32, device="cuda")
    st_ill = torch.empty_like(st_img)

    def step():
        batch.run("batched")
        batch.stack_device(st_img, st_ill)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i} done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cu = info["cell_updates"]
    value = cu * args.steps / elapsed / 1e9
    shots_per_s = cfg.n_shots * args.steps / elapsed

    kernel_ms, launches = batch.time_kernels()
    from das_amd.batch import kernel_configs, kernel_name

    T = kernel_configs()[args.kernel][0] if args.kernel > 0 else 1
    kname = kernel_name(args.kernel) if args.kernel > 0 else "k_batch_step"
    family = kname.split("<")[0]
    if args.kernel > 0:
        alg = blocked_algorithmic_bytes(cfg.shots, T, cfg.k)
    else:
        alg = BYTES_PER_CELL_UPDATE * cu
    achieved = alg / (kernel_ms * 1e-3) / 1e9
    effective = BYTES_PER_CELL_UPDATE * cu / (kernel_ms * 1e-3) / 1e9
    counted = counted_traffic(f"{family}_hbm_traffic.json") if args.kernel > 0 else None
    if counted is not None and counted.get("kernel_config") != args.kernel:
        counted = None
    # HBM bytes per launch: the blocked scheme's algorithmic bytes times the
    # counted/algorithmic ratio of the committed rocprofv3 --pmc summary
    traffic = round(counted["counted_over_algorithmic"] * alg / launches) if counted else None
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched acoustic RTM: 64 virtual shots, nz=300, "
                               "nx~U{600..1600}, nt~U{3000..6000}, dx=5 m, dt=0.5 ms, k=2, "
                               "FP32 stack",
                   "global_batch": cfg.n_shots, "cell_updates_per_step": cu,
                   "slots": args.slots, "graph_steps": args.graph_steps,
                   "kernel": args.kernel,
                   "autotune": tuning,
                   "device_bytes": info["device_bytes"]},
        "shots_per_s": round(shots_per_s, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": (f"{kname}, configuration {args.kernel} (T={T})" if args.kernel
                                else "k_batch_step"),
                     "launches": launches,
                     "avg_launch_us": round(kernel_ms * 1e3 / max(launches, 1), 3),
                     "algorithmic_bytes_per_launch": round(alg / launches),
                     "algorithmic_bytes_per_cell_update": round(alg / cu, 3),
                     "algorithmic_model": (f"blocked scheme: 28 B per cell per {T}-step block "
                                           f"(16 read, 12 written) + 4 B per snapshot level "
                                           f"each way (k={cfg.k})" if args.kernel else
                                           "per-step scheme: 20 B per step + 24 B per imaging "
                                           "step"),
                     "effective_gbs": round(effective, 1),
                     "effective_note": ("26 B per cell-update (the per-step scheme's bytes) "
                                        "over the kernel time: the rate a per-step kernel "
                                        "would need, not bytes moved")},
    }
    if counted is not None:
        gbs = traffic * launches / (kernel_ms * 1e-3) / 1e9
        out["roofline"]["traffic_counted"] = {
            "hbm_bytes_per_launch": traffic,
            "bytes_per_cell_update": round(traffic * launches / cu, 3),
            "counted_over_algorithmic": counted["counted_over_algorithmic"],
            "gbs": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
            "source": f"profiles/{family}_hbm_traffic.json (rocprofv3 FETCH_SIZE, WRITE_SIZE, "
                      "copy-kernel calibrated)"}
    if args.kernel > 0 and not args.no_check:
        out["check"] = check_against_per_step(cfg, batch, (0, cfg.n_shots - 1), log)
    if not args.no_plain:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.run("plain")
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        out["plain_gpu"] = {"value": round(cu / tp / 1e9, 3), "unit": "Gcell-updates/s",
                            "shots_per_s": round(cfg.n_shots / tp, 3),
                            "mode": "one shot at a time, one launch per time step"}
        out["speedup_vs_plain"] = round(shots_per_s / (cfg.n_shots / tp), 2)
        out["plain_gpu"]["acoustic_shot_200x400"] = plain_shot_200x400()
    batch.close()
    if not args.no_elastic:
        out["elastic"] = elastic_leg(args, log)
    if not args.no_fwi:
        out["fwi"] = fwi_leg(args, log)
    if not args.no_gathers:
        out["gathers"] = gather_leg(args, log)
    if not args.no_served:
        el = out.get("elastic", {}).get("value", 50.0)
        out["served"] = served_leg(args, value, el, args.kernel, log)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget * 1e9, args.cpu_budget_one_core * 1e9)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
