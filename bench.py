#!/usr/bin/env python3
"""Heun cell-steps/s of the FDM LLG step on one MI355X.

Workload (the metric's configuration): thin film 512x512x8 cells of 4x4x2 nm,
Ms 8.6e5 A/m, A 1.3e-11 J/m, alpha 0.02, exchange + Newell demag (padded
1024x1024x16) + Zeeman with B_ext of seeded random direction and 0-50 mT
magnitude, m0 = seeded random unit vectors (planner generator), Heun dt 5e-14 s.

A step is one fullmag_fdm_backend_step() call through the C ABI, including the
per-step diagnostics it returns, exactly as the runner drives it.  The timed
region is bracketed by torch.cuda.synchronize(); inputs are resident on the
device before timing starts.  Prints one JSON line.

Also reported: the plain FP64 path's throughput on the same workload, the
mixed-vs-FP64 accuracy after --acc-steps steps from the same m0 (with the FP64
path's own divergence from a 1e-12 perturbation of m0 over the same steps:
this random-m0, alpha = 0.02 workload is chaotic, so that envelope is what a
step-count horizon means), the same comparison over --smooth-steps steps from
a smooth m0, the roofline position of the dominant kernel (per-kernel
HIP-event timing on the backend's stream), and the CPU oracle timed on the
host (bounded sample).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MU0 = 4.0 * np.pi * 1e-7
HBM_PEAK_GBS = 8000.0
METRIC = "Heun cell-steps/s (exch+demag+Zeeman) mixed-FP32 vs FP64 at max|Δm|≤1e-5"


def workload():
    import fullmag_amd as fa

    nx, ny, nz = 512, 512, 8
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(42, n)
    v = fa.generate_random_unit_vectors(7, 2)
    b_tesla = 0.050 * (v[1, 0] + 1.0) / 2.0  # magnitude in [0, 50] mT from the same stream
    h_ext = tuple((b_tesla / MU0) * v[0])
    base = dict(cells=(nx, ny, nz), cell_size=(4e-9, 4e-9, 2e-9), initial_magnetization=m0,
                saturation_magnetisation=8.6e5, exchange_stiffness=1.3e-11, damping=0.02,
                external_field=h_ext, integrator="heun", demag_kernel="device")
    return base, {"b_ext_mT": 1e3 * b_tesla, "b_dir": [float(x) for x in v[0]]}


def alg_bytes(nx, ny, nz, q, r):
    """Compulsory HBM bytes per launch of each kernel of one Heun step
    (each array read or written once per launch; the exchange stencil's
    neighbour re-reads are not counted).  q: state bytes, r: field bytes."""
    n = nx * ny * nz
    c = 2 * r
    h = nx + 1
    a1 = h * ny * nz
    a2 = h * 2 * ny * nz
    kern = 6 * (nx + 1) * (ny + 1) * (nz + 1)
    return {
        # one launch reads the state m, the other the stage point
        "demag_x_r2c": 3 * n * (q + r) / 2 + 3 * a1 * c,
        "demag_y_fwd": 3 * a1 * c + 3 * a2 * c,
        "demag_z_conv": 3 * a2 * c + kern * r + 3 * a2 * c,
        "demag_y_inv": 3 * a2 * c + 3 * a1 * c,
        "demag_x_c2r": 3 * a1 * c + 3 * n * r,
        # corrector: mt, Hd, m0, k1 -> m (the predictor runs in the stats
        # pass of the previous step: m, Hd -> stats, k1, mt)
        "llg_stage": 3 * n * (2 * q + 3 * r),
        "llg_stats": 3 * n * (q + 3 * r),
    }


def step_bytes(nx, ny, nz, q, r):
    """Algorithmic bytes of one Heun step: two demag evaluations plus the
    torque/predictor and corrector passes, stats fused."""
    n = nx * ny * nz
    c = 2 * r
    h = nx + 1
    a1, a2 = h * ny * nz, h * 2 * ny * nz
    kern = 6 * (nx + 1) * (ny + 1) * (nz + 1)
    d = (3 * n * q + 3 * a1 * c) + (3 * a1 * c + 3 * a2 * c) + (3 * a2 * c + kern * r + 3 * a2 * c) \
        + (3 * a2 * c + 3 * a1 * c) + (3 * a1 * c + 3 * n * r)
    return 2 * d + 3 * n * (2 * q + 2 * r) + 3 * n * (3 * q + 2 * r)


def smooth_m0(nx, ny, nz):
    """Uniform magnetisation with a gentle in-plane twist along x."""
    x = np.arange(nx)[None, None, :] * np.ones((nz, ny, 1))
    th = 0.3 * np.sin(2 * np.pi * x / nx)
    m = np.stack([np.cos(th), np.sin(th), 0.05 * np.ones_like(th)], -1).reshape(-1, 3)
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def accuracy(fa, base, dt, steps, perturb=None, chunk=0):
    """max|dm| and relative total-energy error of mixed vs FP64 after
    `steps` steps from the same m0; with `perturb`, also the FP64 path's
    divergence from m0 perturbed by that amplitude.  chunk > 0 advances in
    step_many chunks of that many steps (stats only at chunk ends)."""
    a = fa.NativeFdmBackend(fa.FdmPlan(precision="double", **base))
    b = fa.NativeFdmBackend(fa.FdmPlan(precision="mixed", **base))
    p = None
    if perturb:
        m0 = np.asarray(base["initial_magnetization"])
        rng = np.random.default_rng(1)
        mp = m0 + perturb * rng.standard_normal(m0.shape)
        mp /= np.linalg.norm(mp, axis=1, keepdims=True)
        p = fa.NativeFdmBackend(fa.FdmPlan(precision="double", **dict(base, initial_magnetization=mp)))
    if chunk > 0:
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            sa = a.step_many(dt, k)
            sb = b.step_many(dt, k)
            if p is not None:
                p.step_many(dt, k)
            done += k
    else:
        for _ in range(steps):
            sa = a.step(dt)
            sb = b.step(dt)
            if p is not None:
                p.step(dt)
    ma = a.copy_m()
    out = {"steps": steps, "max_abs_dm": float(np.abs(ma - b.copy_m()).max()),
           "rel_energy_err": abs(sb.e_total - sa.e_total) / abs(sa.e_total)}
    if p is not None:
        out["fp64_divergence_from_1e-12_perturbation"] = float(np.abs(ma - p.copy_m()).max())
        p.close()
    a.close()
    b.close()
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_heun_film512x512x8.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH + WRITE, calibrated against a
    copy kernel) from the committed rocprofv3 --pmc summary, or None."""
    try:
        text = open(TRAFFIC_FILE).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        k = json.loads(block)["kernels"].get(kernel)
        return None if k is None else k["fetch_bytes"] + k["write_bytes"]
    except (OSError, IndexError, KeyError, ValueError):
        return None


def time_steps(backend, steps, warmup, dt):
    import torch

    for _ in range(warmup):
        backend.step(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        backend.step(dt)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def kernel_profile(backend, dt, steps=10):
    backend.set_timing(True)
    for _ in range(steps):
        backend.step(dt)
    t = backend.timings()
    backend.set_timing(False)
    return t


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(base, dt):
    """The oracle (FP64 C++ restatement of fullmag-engine: OpenMP per-cell
    loops, serial FFT as crates/fullmag-engine/src/lib.rs:2768-2812) on the
    host: one Heun step of the 512x512x8 workload at nproc - 1 threads (the
    runner's default, crates/fullmag-runner/src/lib.rs:766-770) and at nproc
    threads, and the 64x64x1 parity workload (1000 steps) in full.  nproc is
    the CPU share the process may use (sched_getaffinity, capped by
    OMP_NUM_THREADS when that is set)."""
    from oracle import oracle as O

    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    if os.environ.get("OMP_NUM_THREADS"):
        nproc = min(nproc, int(os.environ["OMP_NUM_THREADS"]))
    nx, ny, nz = base["cells"]
    dx, dy, dz = base["cell_size"]
    cfg = O.OracleConfig(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, ms=base["saturation_magnetisation"],
                         a=base["exchange_stiffness"], alpha=base["damping"], h_ext=base["external_field"])
    p = O.OracleProblem(cfg, base["initial_magnetization"])
    per = {}
    for threads in sorted({max(1, nproc - 1), nproc}):
        O.set_num_threads(threads)
        t0 = time.perf_counter()
        p.step(dt)
        per[threads] = time.perf_counter() - t0
    best_t = min(per, key=per.get)
    # the 64x64x1 parity workload in full: 1000 steps of 1e-13 s
    import fullmag_amd as fa
    O.set_num_threads(nproc)
    c1 = O.OracleConfig(nx=64, ny=64, nz=1, dx=5e-9, dy=5e-9, dz=3e-9, ms=8e5, a=1.3e-11, alpha=0.5,
                        h_ext=(10e-3 / MU0, 0.0, 0.0))
    q = O.OracleProblem(c1, fa.generate_random_unit_vectors(42, 64 * 64))
    t0 = time.perf_counter()
    for _ in range(1000):
        q.step(1e-13)
    el1 = time.perf_counter() - t0
    n = nx * ny * nz
    return {"value": n / per[best_t], "unit": "cell-steps/s", "cores": best_t, "kind": "port",
            "sample": f"1 Heun step of the {nx}x{ny}x{nz} workload ({n} cells) per thread count, FP64 oracle "
                      "restating crates/fullmag-engine (setup and Newell build excluded)",
            "cpu_model": cpu_model(), "nproc": nproc,
            "by_threads": {str(t): {"cell_steps_per_s": n / v, "seconds": v} for t, v in sorted(per.items())},
            "config1_64x64x1_1000_steps": {"seconds": el1, "cell_steps_per_s": 64 * 64 * 1000 / el1,
                                           "threads": nproc}}


# ---------------------------------------------------------------------------
# Other workloads (opt-in with --workload; the default line is the film)
# ---------------------------------------------------------------------------
def line(metric, value, unit, steps, warmup, ms_per_step, dtype, config, higher=True, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms_per_step, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def sweep96(fa, torch, args):
    """96 independent instances: grids {32,64,128,256}^2 x {1,2,4,8} layers
    (6 of each), cells 5x5x3 nm, Ms ~ U[4e5,1.4e6], A ~ U[5e-12,2e-11],
    alpha ~ U[0.005,0.5], B = random unit direction x U[0,100] mT, seeded
    random m0, dt 1e-13 s, mixed-precision Heun advanced together through
    one hipGraph (20 steps per replay), schedules autotuned per grid."""
    rng = np.random.default_rng(5)
    shapes = [(s, s, l) for s in (32, 64, 128, 256) for l in (1, 2, 4, 8)]
    members, cells = [], 0
    t0 = time.perf_counter()
    for i in range(96):
        nx, ny, nz = shapes[i % 16]
        n = nx * ny * nz
        d = fa.generate_random_unit_vectors(1000 + i, 1)[0]
        b_t = 0.1 * rng.uniform()
        plan = fa.FdmPlan(cells=(nx, ny, nz), cell_size=(5e-9, 5e-9, 3e-9),
                          initial_magnetization=fa.generate_random_unit_vectors(2000 + i, n),
                          saturation_magnetisation=rng.uniform(4e5, 1.4e6), exchange_stiffness=rng.uniform(5e-12, 2e-11),
                          damping=rng.uniform(0.005, 0.5), external_field=tuple(b_t / MU0 * d), precision="mixed",
                          demag_kernel="device")
        members.append(fa.NativeFdmBackend(plan))
        cells += n
    setup_s = time.perf_counter() - t0
    schedules = {}
    for i, b in enumerate(members):
        key = shapes[i % 16]
        if key not in schedules:
            schedules[key] = b.autotune(reps=3)
        else:
            b.set_schedule(schedules[key])
    dt = 1e-13
    run = fa.BatchRun(members, [dt] * len(members), steps_per_graph=20)
    run.run(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = run.run(args.steps)
    el = time.perf_counter() - t0
    # the same members stepped one at a time without graphs, on a bounded sample
    ref_steps = max(1, args.steps // 10)
    t1 = time.perf_counter()
    for b in members:
        b.step_many(dt, ref_steps)
    el_eager = time.perf_counter() - t1
    run.close()
    finite = all(np.isfinite(s.e_total) for s in stats)
    for b in members:
        b.close()
    line("Heun cell-steps/s aggregate over a 96-instance batched sweep (mixed precision)", cells * args.steps / el,
         "cell-steps/s", args.steps, args.warmup, 1e3 * el / args.steps, "f32 fields / f64 state",
         {"workload": "batched sweep: 96 instances, grids {32,64,128,256}^2 x {1,2,4,8} layers (6 each), 5x5x3 nm, "
                      "Ms U[4e5,1.4e6], A U[5e-12,2e-11], alpha U[0.005,0.5], B random dir x U[0,100] mT, "
                      "seeded m0, Heun dt 1e-13 s, one hipGraph per 20 steps, autotuned schedules",
          "instances": 96, "cells_total": cells, "setup_s": setup_s, "all_finite": finite,
          "eager_step_many_cell_steps_per_s": cells * ref_steps / el_eager,
          "schedules": {f"{k[0]}x{k[1]}x{k[2]}": v for k, v in schedules.items()}})


def film_m0(nx, ny, nz):
    """In-plane state with gentle x and y modulations (no domain walls)."""
    x = np.arange(nx)[None, None, :] * np.ones((nz, ny, 1))
    y = np.arange(ny)[None, :, None] * np.ones((nz, 1, nx))
    th = 0.4 * np.sin(2 * np.pi * x / nx) + 0.2 * np.cos(2 * np.pi * y / ny)
    m = np.stack([np.cos(th), np.sin(th), 0.1 * np.ones_like(th)], -1).reshape(-1, 3)
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def vortex_m0(nx, ny, nz, c=0.1, noise=0.01, seed=3):
    y, x = np.meshgrid(np.arange(ny) - (ny - 1) / 2, np.arange(nx) - (nx - 1) / 2, indexing="ij")
    v = np.stack([-y, x, np.full_like(x, c)], -1)
    v = v / np.linalg.norm(v, axis=-1, keepdims=True)
    v = np.broadcast_to(v[None], (nz, ny, nx, 3)).reshape(-1, 3)
    v = v + noise * np.random.default_rng(seed).standard_normal(v.shape)
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def dp45_block(fa, torch, args):
    """256x256x64 cells of 2 nm, Ms 1.4e6, A 2.1e-11, alpha 0.01, no Zeeman,
    m0 = normalise((-y, x, 0.1) + 0.01 N(0,1)) (coordinates in cells from the
    grid centre, seed 3), DP45 with max error 1e-6 and dt feedback from
    suggested_next_dt, demag padded 512x512x128; --steps accepted steps, or
    with --sim-ns every accepted step up to that simulated time."""
    nx, ny, nz = 256, 256, 64
    n = nx * ny * nz
    base = dict(cells=(nx, ny, nz), cell_size=(2e-9, 2e-9, 2e-9), initial_magnetization=vortex_m0(nx, ny, nz),
                saturation_magnetisation=1.4e6, exchange_stiffness=2.1e-11, damping=0.01, integrator="dp45",
                demag_kernel="device", adaptive_max_error=1e-6)
    if args.sim_ns > 0:
        return dp45_block_horizon(fa, torch, base, n, args.sim_ns * 1e-9)
    res, final = {}, {}
    for prec in ("mixed", "double"):
        b = fa.NativeFdmBackend(fa.FdmPlan(precision=prec, **base))
        dt = 1e-13
        s = None
        for _ in range(args.warmup):
            s = b.step(dt)
            dt = s.dt_suggested or dt
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim0 = s.time if s is not None else 0.0
        dts = []
        for _ in range(args.steps):
            s = b.step(dt)
            dts.append(s.dt)
            dt = s.dt_suggested or dt
        el = time.perf_counter() - t0
        res[prec] = {"cell_steps_per_s": n * args.steps / el, "ms_per_step": 1e3 * el / args.steps,
                     "mean_dt": (s.time - sim0) / args.steps, "sim_time_s": s.time, "e_total": s.e_total}
        final[prec] = (b.copy_m(), s, np.array(dts))
        b.close()
    m = res["mixed"]
    (m_mx, s_mx, dt_mx), (m_64, s_64, dt_64) = final["mixed"], final["double"]
    acc = {"accepted_steps": args.warmup + args.steps, "max_abs_dm": float(np.abs(m_mx - m_64).max()),
           "rel_energy_err": abs(s_mx.e_total - s_64.e_total) / abs(s_64.e_total),
           "max_rel_dt_diff": float(np.abs(dt_mx - dt_64).max() / dt_64.max()),
           "rel_sim_time_diff": abs(s_mx.time - s_64.time) / s_64.time}
    line("DP45 accepted cell-steps/s, 256x256x64 block (mixed precision)", m["cell_steps_per_s"], "cell-steps/s",
         args.steps, args.warmup, m["ms_per_step"], "f32 fields / f64 state",
         {"workload": "3-D block 256x256x64 cells of 2 nm, Ms 1.4e6, A 2.1e-11, alpha 0.01, no Zeeman, vortex m0 "
                      "(-y, x, 0.1) + 1% noise, DP45 max error 1e-6, padded 512x512x128",
          "cells": n, "mixed": m, "double": res["double"], "mixed_vs_fp64_same_accepted_steps": acc,
          "speedup_vs_fp64": m["cell_steps_per_s"] / res["double"]["cell_steps_per_s"]})


def dp45_block_horizon(fa, torch, base, n, t_end):
    """The 3-D block DP45 run over its configured simulated time: each
    precision steps with dt feedback until t >= t_end.  No Zeeman term and
    alpha > 0, so the total energy may not rise; reported: accepted steps,
    wall time, the largest energy rise between accepted steps (relative),
    max ||m| - 1| at the end, and mixed against FP64 at the end (each at its
    own final time; the difference of those times is reported beside)."""
    res, final = {}, {}
    for prec in ("mixed", "double"):
        b = fa.NativeFdmBackend(fa.FdmPlan(precision=prec, **base))
        dt, steps, rise, e_prev = 1e-13, 0, 0.0, None
        e_first = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = None
        while s is None or s.time < t_end:
            s = b.step(min(dt, t_end - s.time) if s is not None else dt)
            steps += 1
            dt = s.dt_suggested or dt
            if e_prev is not None:
                rise = max(rise, (s.e_total - e_prev) / abs(e_prev))
            else:
                e_first = s.e_total
            e_prev = s.e_total
            if steps % 2000 == 0:
                print(f"{prec}: {steps} accepted steps, t = {s.time:.4e} s, E = {s.e_total:.6e} J", file=sys.stderr,
                      flush=True)
        el = time.perf_counter() - t0
        m = b.copy_m()
        res[prec] = {"accepted_steps": steps, "wall_s": el, "ms_per_step": 1e3 * el / steps,
                     "cell_steps_per_s": n * steps / el, "sim_time_s": s.time, "mean_dt": s.time / steps,
                     "e_first": e_first, "e_final": s.e_total, "max_rel_energy_rise": rise,
                     "max_norm_dev": float(np.abs(np.linalg.norm(m, axis=1) - 1.0).max())}
        final[prec] = (m, s)
        b.close()
    (m_mx, s_mx), (m_64, s_64) = final["mixed"], final["double"]
    mx = res["mixed"]
    line("DP45 accepted cell-steps/s, 256x256x64 block over its simulated time (mixed precision)",
         mx["cell_steps_per_s"], "cell-steps/s", mx["accepted_steps"], 0, mx["ms_per_step"], "f32 fields / f64 state",
         {"workload": "3-D block 256x256x64 cells of 2 nm, Ms 1.4e6, A 2.1e-11, alpha 0.01, no Zeeman, vortex m0 "
                      "(-y, x, 0.1) + 1% noise, DP45 max error 1e-6 with dt feedback, padded 512x512x128",
          "cells": n, "t_end_s": t_end, "mixed": mx, "double": res["double"],
          "mixed_vs_fp64_at_end": {"max_abs_dm": float(np.abs(m_mx - m_64).max()),
                                   "rel_energy_err": abs(s_mx.e_total - s_64.e_total) / abs(s_64.e_total),
                                   "final_time_diff_s": s_mx.time - s_64.time},
          "speedup_vs_fp64": res["double"]["wall_s"] / mx["wall_s"]})


def adjoint_film(fa, torch, args):
    """128x128x4 film of 5x5x3 nm, per-cell Ms ~ U[6e5, 9e5] (seed 4),
    A 1.3e-11, alpha 0.1, B = (20, 5, 0) mT, 2000 Heun steps of 1e-13 s from
    a smooth in-plane m0 (film_m0); loss = mean m_x over the centred 32x32
    window (x, y in [48, 80)) of all layers; checkpoints every 100 steps.
    Gradients w.r.t. per-cell Ms and B from the adjoint, checked against
    central differences (h = 1e-4 relative, FP64) at 32 seeded cells and the
    two in-plane B components.  (From a uniformly random m0 the relaxation is
    turbulent and the h = 1e-4 difference itself carries a 1e-3 truncation
    error; the difference converges to the adjoint as h shrinks.)"""
    nx, ny, nz = 128, 128, 4
    n = nx * ny * nz
    steps = 2000 if args.steps >= 300 else args.steps
    rng = np.random.default_rng(4)
    ms_np = rng.uniform(6e5, 9e5, n)
    m0 = film_m0(nx, ny, nz)
    b_mt = np.array([20.0, 5.0, 0.0])
    res = {}
    for prec in ("double", "mixed"):
        plan = fa.FdmPlan(cells=(nx, ny, nz), cell_size=(5e-9, 5e-9, 3e-9), initial_magnetization=m0,
                          saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.1, precision=prec,
                          demag_kernel="device")
        ms = torch.tensor(ms_np, dtype=torch.float64, requires_grad=True)
        bt = torch.tensor(b_mt, dtype=torch.float64, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m = fa.heun_rollout(m0, ms, bt * 1e-3 / MU0, plan, 1e-13, steps, 100)
        loss = m[:, 0].reshape(nz, ny, nx)[:, 48:80, 48:80].mean()
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        loss.backward()
        torch.cuda.synchronize()
        t_all = time.perf_counter() - t0
        res[prec] = {"loss": float(loss), "g_ms": ms.grad.numpy().copy(), "g_b": bt.grad.numpy().copy(),
                     "forward_s": t_fwd, "backward_s": t_all - t_fwd, "plan": plan}
    d = res["double"]

    def L(ms_v, b_v):
        with torch.no_grad():
            m = fa.heun_rollout(m0, torch.tensor(ms_v), torch.tensor(b_v * 1e-3 / MU0), d["plan"], 1e-13, steps,
                                steps)
            return float(m[:, 0].reshape(nz, ny, nx)[:, 48:80, 48:80].mean())

    cells = rng.choice(n, size=32, replace=False) if args.steps >= 300 else rng.choice(n, size=4, replace=False)
    errs, fds = [], []
    for i in cells:
        h = 1e-4 * ms_np[i]
        p, q = ms_np.copy(), ms_np.copy()
        p[i] += h
        q[i] -= h
        fd = (L(p, b_mt) - L(q, b_mt)) / (2 * h)
        fds.append(fd)
        errs.append(abs(fd - d["g_ms"][i]))
    scale = max(np.abs(fds).max(), 1e-300)
    rel = [e / max(abs(f), 1e-3 * scale) for e, f in zip(errs, fds)]
    b_errs = []
    for k in range(2):
        h = 1e-4 * np.abs(b_mt).max()
        p, q = b_mt.copy(), b_mt.copy()
        p[k] += h
        q[k] -= h
        fd = (L(ms_np, p) - L(ms_np, q)) / (2 * h)
        b_errs.append(abs(fd - d["g_b"][k]) / max(abs(fd), 1e-300))
    x = res["mixed"]
    mixed_vs_fp64 = float(np.abs(x["g_ms"] - d["g_ms"]).max() / np.abs(d["g_ms"]).max())
    line("adjoint cell-steps/s (Heun backward pass with checkpoint recompute), 128x128x4 film",
         n * steps / d["backward_s"], "cell-steps/s", steps, 0, 1e3 * d["backward_s"] / steps, "f64",
         {"workload": "differentiable film 128x128x4 of 5x5x3 nm, per-cell Ms U[6e5,9e5], A 1.3e-11, alpha 0.1, "
                      "B (20,5,0) mT, smooth in-plane m0, 2000 Heun steps dt 1e-13, loss mean m_x over centred 32x32 "
                      "window, checkpoint every 100",
          "fd_cells": len(cells), "max_rel_err_ms": float(max(rel)),
          "max_abs_err_over_max_fd_ms": float(max(errs) / scale),
          "max_rel_err_b": float(max(b_errs)), "loss": d["loss"],
          "fp64": {"forward_s": d["forward_s"], "backward_s": d["backward_s"]},
          "mixed": {"forward_s": x["forward_s"], "backward_s": x["backward_s"],
                    "max_rel_dev_g_ms_vs_fp64": mixed_vs_fp64}})


def stack_trilayer(fa, torch, args):
    """Three 512x512x1 films of 4x4x1 nm at z = 0, 2 and 4 nm (1 nm spacers),
    Ms 8e5 / 6e5 / 8e5 A/m, A 1.3e-11 J/m, alpha 0.02, seeded random m0,
    B = 10 mT along x, coupled by demag on the common 512x512x1 grid (self
    kernel and shifted kernels for +-2 and +-4 nm), mixed-precision Heun
    dt 5e-14 s; a step is one fullmag_fdm_ext_stack_step with its stats.
    Beside it: the same three films stepped as independent backends (no
    coupling), and the oracle's stack on the host at 3 x 128x128 (one step)."""
    nx, ny = 512, 512
    cell = (4e-9, 4e-9, 1e-9)
    h = (0.01 / MU0, 0.0, 0.0)
    specs = [(0.0, 8e5, 11), (2e-9, 6e5, 12), (4e-9, 8e5, 13)]

    def plan(n_x, n_y, prec):
        layers = [fa.FdmLayerPlan(magnet_name=f"L{i}", native_grid=(n_x, n_y, 1), native_cell_size=cell,
                                  native_origin=(0.0, 0.0, z), initial_magnetization=fa.generate_random_unit_vectors(
                                      seed, n_x * n_y), saturation_magnetisation=ms, exchange_stiffness=1.3e-11,
                                  damping=0.02) for i, (z, ms, seed) in enumerate(specs)]
        return fa.FdmMultilayerPlan(common_cells=(n_x, n_y, 1), convolution_cell_size=cell, layers=layers,
                                    external_field=h, precision=prec)

    dt = 5e-14
    cells = 3 * nx * ny
    stack = fa.MultilayerStack(plan(nx, ny, "mixed"))
    for _ in range(args.warmup):
        stack.step(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = stack.step(dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    stack.close()
    # the same films as three independent backends, each with its own demag
    films = [fa.NativeFdmBackend(fa.FdmPlan(cells=(nx, ny, 1), cell_size=cell, initial_magnetization=lp.initial_magnetization,
                                            saturation_magnetisation=lp.saturation_magnetisation,
                                            exchange_stiffness=1.3e-11, damping=0.02, external_field=h,
                                            precision="mixed", demag_kernel="device"))
             for lp in plan(nx, ny, "mixed").layers]
    el_f = sum(time_steps(b, args.steps, args.warmup, dt) for b in films)
    for b in films:
        b.close()
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        cn = 128
        probs = [O.OracleProblem(O.OracleConfig(nx=cn, ny=cn, nz=1, dx=cell[0], dy=cell[1], dz=cell[2], ms=ms,
                                                a=1.3e-11, alpha=0.02, demag=False, h_ext=h),
                                 O.random_unit_vectors(seed, cn * cn)) for _, ms, seed in specs]
        st = O.OracleStack(probs, [(0.0, 0.0, z) for z, _, _ in specs], (cn, cn, 1), cell)
        t1 = time.perf_counter()
        st.step(dt)
        el_c = time.perf_counter() - t1
        cpu = {"value": 3 * cn * cn / el_c, "unit": "cell-steps/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"1 Heun step of a 3 x {cn}x{cn}x1 stack (FP64 oracle, OpenMP)"}
    line("Heun cell-steps/s of a coupled three-layer stack (mixed precision)", cells * args.steps / el,
         "cell-steps/s", args.steps, args.warmup, 1e3 * el / args.steps, "f32 fields / f64 state",
         {"workload": "stack of three 512x512x1 films, 4x4x1 nm cells, z = 0 / 2 / 4 nm, Ms 8e5 / 6e5 / 8e5, "
                      "A 1.3e-11, alpha 0.02, B 10 mT x, seeded m0, Heun dt 5e-14 s, demag on the common grid",
          "cells": cells, "independent_films_ms_per_step": 1e3 * el_f / args.steps,
          "coupling_cost_ratio": el / el_f, "e_total": s.e_total},
         cpu_baseline=cpu)


def tiers_film(fa, torch, args):
    """The film workload's grid (512x512x8 cells of 4x4x2 nm) holding a
    cylinder of radius 860 nm sampled by the planner's rules
    (fullmag_amd.boundary): T0 face-link exchange ("volume"), volume-fraction
    demag source and the sparse correction rows, active mask phi > 0,
    Ms 8.6e5, A 1.3e-11, alpha 0.02, seeded m0, mixed-precision Heun
    dt 5e-14 s.  Beside it: the same masked cylinder without tiers."""
    from fullmag_amd import boundary

    nx, ny, nz = 512, 512, 8
    cell = (4e-9, 4e-9, 2e-9)
    t0 = time.perf_counter()
    g = boundary.compute_boundary_geometry(boundary.cylinder_sdf(860e-9, nx * cell[0] / 2, ny * cell[1] / 2), nx, ny,
                                           nz, *cell, False)
    geo_s = time.perf_counter() - t0
    mask = (g.volume_fraction > 0).astype(np.uint8)
    m0 = fa.generate_random_unit_vectors(21, nx * ny * nz)
    base = dict(cells=(nx, ny, nz), cell_size=cell, initial_magnetization=m0, saturation_magnetisation=8.6e5,
                exchange_stiffness=1.3e-11, damping=0.02, precision="mixed", demag_kernel="device", active_mask=mask)
    dt = 5e-14
    res = {}
    for key, kw in (("tiers", dict(boundary_correction="volume", boundary_geometry=g)), ("plain", {})):
        b = fa.NativeFdmBackend(fa.FdmPlan(**base, **kw))
        res[key] = time_steps(b, args.steps, args.warmup, dt)
        b.close()
    n = nx * ny * nz
    line("Heun cell-steps/s, 512x512x8 cylinder with the T0 boundary tier (mixed precision)",
         n * args.steps / res["tiers"], "cell-steps/s", args.steps, args.warmup, 1e3 * res["tiers"] / args.steps,
         "f32 fields / f64 state",
         {"workload": "512x512x8 cells of 4x4x2 nm holding a cylinder of radius 860 nm: T0 face links, phi-weighted "
                      "demag source, sparse correction rows from the planner's sampling, mask phi > 0, Ms 8.6e5, "
                      "A 1.3e-11, alpha 0.02, seeded m0, Heun dt 5e-14 s",
          "correction_rows": int(g.demag_corr_target_idx.size), "stencil": int(g.demag_corr_stencil_size),
          "geometry_s": geo_s, "plain_masked_ms_per_step": 1e3 * res["plain"] / args.steps,
          "tier_cost_ratio": res["tiers"] / res["plain"]})


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["film512", "sweep96", "dp45-block", "adjoint-film", "stack-trilayer",
                                           "tiers-film"], default="film512")
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=100)
    ap.add_argument("--acc-steps", type=int, default=50)
    ap.add_argument("--smooth-steps", type=int, default=1000)
    ap.add_argument("--horizon-steps", type=int, default=5000,
                    help="the workload's configured step count: mixed vs FP64 after this many steps")
    ap.add_argument("--sim-ns", type=float, default=0.0,
                    help="dp45-block: run each precision until this simulated time (ns) instead of --steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run two device-to-device copies, 256 MiB and 1 GiB (references for the HBM "
                         "counters; the second exceeds the 256 MB Infinity Cache)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this "
              "build", file=sys.stderr)
        sys.exit(2)

    import torch

    import fullmag_amd as fa

    torch.cuda.init()
    if args.copy_calibration:
        for mib in (256, 1024):
            src = torch.ones(mib * 2**18, dtype=torch.float32, device="cuda")
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            del src, dst
    if args.workload != "film512":
        {"sweep96": sweep96, "dp45-block": dp45_block, "adjoint-film": adjoint_film, "stack-trilayer": stack_trilayer,
         "tiers-film": tiers_film}[args.workload](fa, torch, args)
        return
    dt = 5e-14
    base, extra = workload()
    nx, ny, nz = base["cells"]
    n = nx * ny * nz

    # accuracy of the mixed path against the FP64 GPU path from the same m0
    # (these runs also bring the GPU to its working clocks before timing)
    acc = accuracy(fa, base, dt, args.acc_steps, perturb=1e-12) if args.acc_steps > 0 else None
    acc_smooth = None
    if args.smooth_steps > 0:
        acc_smooth = accuracy(fa, dict(base, initial_magnetization=smooth_m0(nx, ny, nz)), dt, args.smooth_steps,
                              chunk=500)
    acc_horizon = None
    if args.horizon_steps > 0:
        acc_horizon = accuracy(fa, base, dt, args.horizon_steps, perturb=1e-12, chunk=500)

    fp64 = fa.NativeFdmBackend(fa.FdmPlan(precision="double", **base))
    t64 = time_steps(fp64, args.fp64_steps, 5, dt)
    fp64_value = n * args.fp64_steps / t64
    prof64 = {} if args.no_profile else kernel_profile(fp64, dt)
    fp64.close()

    mixed = fa.NativeFdmBackend(fa.FdmPlan(precision="mixed", **base))
    t_mixed = time_steps(mixed, args.steps, args.warmup, dt)
    value = n * args.steps / t_mixed
    ms_per_step = 1e3 * t_mixed / args.steps
    prof = {} if args.no_profile else kernel_profile(mixed, dt)
    mixed.close()

    roofline = None
    if prof:
        ab = alg_bytes(nx, ny, nz, 8, 4)
        name, (tot_ms, cnt) = max(prof.items(), key=lambda kv: kv[1][0])
        avg_s = tot_ms / cnt / 1e3
        achieved = ab[name] / avg_s / 1e9 if name in ab else None
        roofline = {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": measured_traffic(name),
                    "avg_launch_us": 1e6 * avg_s, "alg_bytes_per_launch": ab.get(name),
                    "kernels_us": {k: round(1e3 * v[0] / v[1], 2) for k, v in sorted(prof.items())}}

    sb_bytes = step_bytes(nx, ny, nz, 8, 4)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields / f64 state",
        "data": "synthetic",
        "config": {"workload": "thin film 512x512x8 cells of 4x4x2 nm, Heun dt 5e-14 s, exchange + Newell demag "
                               "(padded 1024x1024x16) + Zeeman, Ms 8.6e5, A 1.3e-11, alpha 0.02, seeded random "
                               "m0 and B_ext", "cells": n, "b_ext_mT": extra["b_ext_mT"],
                   "precision": "mixed (FP32 fields/FFT/torque, FP64 state and Newell tensor)",
                   "fp64_cell_steps_per_s": fp64_value, "speedup_vs_fp64": value / fp64_value,
                   "fp64_ms_per_step": 1e3 * t64 / args.fp64_steps,
                   "accuracy_vs_fp64": acc, "accuracy_vs_fp64_smooth_m0": acc_smooth,
                   "accuracy_vs_fp64_at_horizon": acc_horizon,
                   "step_alg_bytes": sb_bytes, "step_achieved_GBs": sb_bytes / (ms_per_step / 1e3) / 1e9},
        "roofline": roofline,
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(base, dt),
    }
    if prof64:
        out["config"]["fp64_kernels_us"] = {k: round(1e3 * v[0] / v[1], 2) for k, v in sorted(prof64.items())}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
