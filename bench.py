"""Benchmark of the Natrix fluid step on one MI355X.

Scene: a 2048x2048 float32 grid (velocity, pressure, dye off) with a
circular obstacle (seeded centre in [0.3, 0.7]^2, radius 0.1*N, held in the
persistent obstacle map), inflow 1.0 on the left wall, vorticity
confinement 0.3, no viscosity, dt = 1/60, speed 500, developed by 10
fixed-iteration steps (80 Jacobi sweeps each).  Inputs are resident in HBM.

Default workload (`--workload learned`): steps whose pressure is solved to
FP64 relative residual <= 1e-4.  `value` is the steps/s of the learned
three-level cycle (fine grid, a middle grid 4x coarser, the U-Net's coarse
grid; Chebyshev smoothing on the two finer levels; `--mid 0` for the
two-level cycle); the same
criterion with plain Jacobi correction solves is timed on one step
(capped; its rate is then an upper bound).  The JSON line also carries the
roofline of the learned step's dominant kernel, the fine Chebyshev
smoother (`roofline`), and of the plain path's Jacobi kernel
(`roofline_plain_jacobi`, fixed-sweep steps), both from HIP events on the
simulator's stream, the plain fused step at fixed iterations, and the CPU
oracle timed on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--size N] [--tol 1e-4]
                    [--workload learned|single|batch|mixed] [--no-plain] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("sim steps/sec at 2048² @ rel. residual ≤1e-4 "
          "(learned+Jacobi vs plain); % HBM peak")
HBM_PEAK_GBS = 8000.0
JACOBI_BYTES_PER_CELL_SWEEP = 13   # read p 4 + b 4 + mask 1, write p 4
CHEB_BYTES_PER_CELL_ITER = 25      # read x 4 + r 4 + d 4 + mask 1, write x, r, d 12
CHEB_SMOOTH = 72                   # fine Chebyshev iterations per two-level cycle
CHEB_SMOOTH_3 = 18                 # the same for the three-level cycle (--mid)
CHEB_TILE = (8, 4, 8, 2)           # nx_smooth_chebyshev's default tile
STEP_BYTES_PER_CELL = 104          # the other stages, FP32, dye on, nu = 0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_header(name, **match):
    """The `key: value` header lines of the committed rocprofv3 summary
    profiles/<name>.md when every key named in `match` equals it (tile as a
    comma list), else None."""
    path = os.path.join(ROOT, "profiles", f"{name}.md")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        k, sep, v = line.partition(":")
        if sep and k.strip() and " " not in k.strip():
            kv.setdefault(k.strip(), v.strip())
    try:
        for k, want in match.items():
            got = kv[k]
            if isinstance(want, (list, tuple)):
                ok = [int(t) for t in got.split(",")] == [int(t) for t in want]
            else:
                ok = type(want)(got) == want
            if not ok:
                return None
    except (KeyError, ValueError):
        return None
    kv["_source"] = os.path.relpath(path, ROOT)
    return kv


def profile_traffic(name, **match):
    """HBM bytes per launch from the committed rocprofv3 summary
    profiles/<name>.md (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the copy-probe
    calibration) when its header matches, else (None, None)."""
    kv = profile_header(name, **match)
    try:
        return int(kv["traffic_bytes_per_launch"]), kv["_source"]
    except (TypeError, KeyError, ValueError):
        return None, None


def valu_entry(name, launch_us, **match):
    """VALU issue figure of a launch from the committed summary's
    SQ_INSTS_VALU count (wave instructions per launch): the time the chip
    needs just to issue them, each wave64 VALU instruction taking 4 cycles
    of one of the 1024 SIMDs (256 CUs x 4) at the 2.4 GHz peak engine
    clock, as a fraction of the measured launch time; None without a
    matching summary."""
    kv = profile_header(name, **match)
    try:
        n = float(kv["valu_instructions_per_launch"])
    except (TypeError, KeyError, ValueError):
        return None
    issue_us = n * 4 / (1024 * 2.4e9) * 1e6
    return {"instructions_per_launch": int(n), "issue_bound_us": round(issue_us, 3),
            "frac": round(issue_us / launch_us, 4),
            "peak": "1024 SIMDs x 2.4 GHz, 4 cycles per wave64 VALU instruction",
            "source": kv["_source"]}


def committed_traffic(N, iters, tile, kernel="k_jacobi_tb", suffix="fixed"):
    """HBM bytes per launch of `kernel` on an N x N grid at this iteration
    count and tile, from profiles/<kernel>_<N>_<suffix><iters>.md."""
    return profile_traffic(f"{kernel}_{N}_{suffix}{iters}", grid=N, iterations=iters,
                           tile=list(tile))


def roofline_entry(kernel, algo_bytes, launch_us, traffic, traffic_src, **extra):
    """The roofline block of one kernel.  `algo_bytes`: its compulsory bytes
    per launch (every array it reads or writes moved once per launch);
    `traffic`: HBM bytes per launch from the committed counter summary
    (FETCH_SIZE x 2 + WRITE_SIZE x 1 per dispatch, the copy calibration) or
    None.  `achieved` and `frac` are the measured traffic's rate when the
    summary matches this launch shape, else the compulsory bytes' rate;
    either way a fraction of the HBM peak, not an on-chip iteration rate."""
    t = launch_us * 1e-6
    algo_gbs = algo_bytes / t / 1e9
    out = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel": kernel,
           "avg_launch_us": round(launch_us, 3)}
    if traffic:
        ach = traffic / t / 1e9
        out.update(achieved=round(ach, 1), frac=round(ach / HBM_PEAK_GBS, 4),
                   basis="hbm counter bytes per launch / launch time", traffic=traffic,
                   traffic_source=traffic_src,
                   traffic_over_algorithmic=round(traffic / algo_bytes, 3))
    else:
        out.update(achieved=round(algo_gbs, 1), frac=round(algo_gbs / HBM_PEAK_GBS, 4),
                   basis="compulsory bytes per launch / launch time (no matching counter "
                         "summary)", traffic=None, traffic_source=None)
    out.update(algorithmic_bytes_per_launch=int(algo_bytes),
               algorithmic_GBps=round(algo_gbs, 1),
               algorithmic_frac=round(algo_gbs / HBM_PEAK_GBS, 4))
    out.update(extra)
    return out


def cheb_launch_plan(iterations, k):
    """(passes, compulsory bytes per cell) of each launch of one
    nx_smooth_chebyshev call with k passes per launch: the first launch
    spends a pass on the setup, the last one a pass on the FP64 residual
    (the library's launch loop)."""
    plan, done = [], 0
    while True:
        setup = 1 if done == 0 else 0
        last = setup + (iterations - done) + 1 <= k
        kk = setup + (iterations - done) + 1 if last else k
        its = kk - setup - (1 if last else 0)
        # reads: x, mask, and (r, d) or b; the last also b; writes x and,
        # unless last, r and d
        rd = 4 + 1 + (4 if setup else 8) + (4 if last and not setup else 0)
        wr = 4 if last else 12
        plan.append((kk, rd + wr))
        done += its
        if last:
            return plan


def baseline_formula(cells, K, ms_per_step):
    """The step-level figure of the path's bandwidth model:
    N^2 (104 + 13 K) bytes per step (each array once per stage, K Jacobi /
    smoother passes each 13 B) over the step time.  Temporal blocking keeps
    iterates on chip between passes, so this ratio to the HBM peak is not
    bounded by 1 and is not a roofline fraction."""
    by = cells * (STEP_BYTES_PER_CELL + JACOBI_BYTES_PER_CELL_SWEEP * K)
    tbs = by / (ms_per_step * 1e-3) / 1e12
    return {"bytes_per_step": int(by), "passes_per_step": round(float(K), 2),
            "TBps": round(tbs, 3), "ratio_to_peak": round(tbs / (HBM_PEAK_GBS / 1e3), 4)}


def make_scene(N, seed=2):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.3, 0.7, 2)
    X = np.arange(N, dtype=np.float32)[None, :]
    Y = np.arange(N, dtype=np.float32)[:, None]
    sx, sy = np.float32(c[0]) * np.float32(N), np.float32(c[1]) * np.float32(N)
    d = np.sqrt((sx - X) ** 2 + (sy - Y) ** 2)
    solid = d <= np.float32(0.1 * N)
    return solid, tuple(float(v) for v in c)


def configure(sim, args):
    sim.vorticity = 0.3
    sim.viscosity = 0.0
    sim.iterations = args.iters
    sim.bc_mode = "inflow"
    sim.inflow = 1.0
    if args.tile:
        sim.jacobi_tile = tuple(args.tile)
    if args.mode == "tol":
        sim.tol = args.tol
        sim.max_iters = args.max_iters
        sim.check_every = args.check_every


def threadpools():
    """Host thread pools (threadpoolctl): the oracle's NumPy ufuncs run on one
    thread whatever these report."""
    try:
        import threadpoolctl
        return [{"api": d.get("internal_api"), "threads": d.get("num_threads")}
                for d in threadpoolctl.threadpool_info()]
    except Exception as e:  # threadpoolctl missing or failing: report why
        return [{"error": type(e).__name__}]


def cpu_baseline(N, iters, budget_s):
    """The oracle (NumPy float32, one thread) on the same scene: whole steps
    until the time budget is spent (at least one)."""
    from oracle import fluid as F
    solid, _ = make_scene(N)
    prm = F.Params(vorticity=0.3, viscosity=0.0, iterations=iters,
                   bc_mode=F.BC_INFLOW)
    u = np.zeros((N, N, 2), np.float32)
    p = np.zeros((N, N), np.float32)
    obs = np.zeros((N, N, 2), np.float32)
    dt = np.float32(1 / 60)
    steps = 0
    t0 = time.perf_counter()
    while True:
        u, p, obs, _ = F.step(u, p, obs, dt, prm, persistent=solid)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return steps / el, steps, el


def config_256_splats(seed=1, n=3, N=256):
    """The CPU configuration's forcing: 3 Gaussian splats, centres U[0.1, 0.9]^2,
    radius 0.02 N, |F| ~ U(50, 200) times dt, random direction, dye 1.0."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        f = rng.uniform(50, 200) / 60.0
        a = rng.uniform(0, 2 * np.pi)
        out.append((rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), f * np.cos(a),
                    f * np.sin(a), 0.02 * N, 1.0))
    return out


def config_256(steps=200):
    """The CPU configuration in full: 256^2, free-slip box, vorticity 0.3, no
    viscosity, 40 sweeps from zero, dye on, 3 splats every step, 200 steps;
    seconds on the oracle (NumPy float32, one thread) and on the GPU."""
    import torch
    from natrix_amd import FluidSimulator
    from oracle import fluid as F
    N = 256
    splats = config_256_splats()
    s32 = [tuple(np.float32(v) for v in s) for s in splats]
    prm = F.Params(vorticity=0.3, viscosity=0.0, iterations=40, bc_mode=F.BC_FREE_SLIP)
    u = np.zeros((N, N, 2), np.float32)
    p = np.zeros((N, N), np.float32)
    q = np.zeros((N, N), np.float32)
    obs = np.zeros((N, N, 2), np.float32)
    dt = np.float32(1 / 60)
    t0 = time.perf_counter()
    for _ in range(steps):
        u, q = F.gaussian_splat(u, q, s32)
        u, p, obs, q = F.step(u, p, obs, dt, prm, dye=q)
    t_cpu = time.perf_counter() - t0
    sim = FluidSimulator(N, N, dye=True, bc_mode="free_slip")
    sim.vorticity = 0.3
    sim.viscosity = 0.0
    sim.iterations = 40
    for _ in range(3):
        sim.add_splats(splats)
        sim.update(dt)
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.add_splats(splats)
        sim.update(dt)
    sim.synchronize()
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    sim.destroy()
    return {"workload": "256x256, free-slip, vorticity 0.3, 40 sweeps, dye, 3 splats/step",
            "steps": steps, "cpu_seconds": round(t_cpu, 3), "cpu_cores": 1,
            "gpu_seconds": round(t_gpu, 4),
            "cpu_steps_per_s": round(steps / t_cpu, 2), "gpu_steps_per_s": round(steps / t_gpu, 1)}


def make_batch(n_scenes=64, iters=60):
    """The batched configuration's scenes (natrix_amd.batch.benchmark_scenes)."""
    from natrix_amd.batch import benchmark_scenes
    return benchmark_scenes(n_scenes, iters)


def bench_batch(args, copy_bandwidth):
    import torch
    iters = 60 if args.iters == 80 else args.iters
    b, sides, _, _ = make_batch(iters=iters)
    cells = b.cells
    log(f"bench batch: {len(sides)} scenes, {cells} cells, tile {b.jacobi_tile}")
    b.use_graph = args.graph
    for _ in range(args.warmup):
        b.update(1 / 60)
    b.synchronize()
    torch.cuda.synchronize()
    b.set_timing(not args.graph)
    b.collect_timing()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.update(1 / 60)
    b.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    timing = b.collect_timing()
    jac_ms, jac_n = timing["jacobi"]
    tb_sweeps = (iters - 1) * args.steps
    roof = None
    if jac_n and jac_ms > 0:
        # compulsory bytes per launch: every scene's p, b, mask read, p written
        traffic, tsrc = profile_traffic(f"kb_jacobi_batch{len(sides)}", scenes=len(sides),
                                        cells=cells, iterations=iters,
                                        tile=list(b.jacobi_tile))
        roof = roofline_entry(
            "kb_jacobi", JACOBI_BYTES_PER_CELL_SWEEP * cells, 1e3 * jac_ms / jac_n, traffic, tsrc,
            sweeps_per_launch=round(tb_sweeps / jac_n, 3),
            per_sweep_GBps=round(JACOBI_BYTES_PER_CELL_SWEEP * cells * tb_sweeps
                                 / (jac_ms * 1e-3) / 1e9, 1))
    ms = 1e3 * el / args.steps
    out = {"metric": "batched sim steps/sec (64 uneven scenes, one launch per stage)",
           "value": round(args.steps / el, 3), "unit": "steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{len(sides)} scenes {cells} cells, I={iters}, "
                                  f"free-slip walls, splats every step",
                      "jacobi_tile": list(b.jacobi_tile), "graph": bool(args.graph)},
           "cells_per_s": round(cells * args.steps / el, 1),
           "baseline_formula": baseline_formula(cells, iters, ms),
           "stage_ms_per_step": {k: round(v[0] / args.steps, 5)
                                 for k, v in timing.items() if v[1]},
           "roofline": roof, "cpu_baseline": None}
    b.destroy()
    print(json.dumps(out), flush=True)


def mixed_scene(N, storage, iters, dye=True):
    """4096^2-class scene: 8 seeded Gaussian splats (seed 4) every step,
    free-slip walls, vorticity 0.3, no viscosity, dye on."""
    from natrix_amd import FluidSimulator
    rng = np.random.default_rng(4)
    splats = []
    for _ in range(8):
        f = rng.uniform(50, 200) / 60.0
        ang = rng.uniform(0, 2 * np.pi)
        splats.append((rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9),
                       f * np.cos(ang), f * np.sin(ang), 0.02 * N, 1.0))
    sim = FluidSimulator(N, N, dye=dye, bc_mode="free_slip", storage=storage)
    sim.vorticity = 0.3
    sim.viscosity = 0.0
    sim.iterations = iters
    return sim, splats


def mixed_error_vs_f32(N, storage, iters, steps=500):
    """Error of each field of the reduced-storage rollout against the
    float32 GPU rollout of the same full-size scene (8 splats every step,
    `steps` steps from rest, the mixed configuration's length): relative
    max-norm and L2 error of velocity, pressure (the corrected float32
    pressure, mean-free over the grid) and dye."""
    fields = {}
    for st in (storage, "f32"):
        sim, splats = mixed_scene(N, st, iters, dye=True)
        for _ in range(steps):
            sim.add_splats(splats)
            sim.update(1 / 60)
        fields[st] = {"velocity": sim.download("velocity").astype(np.float64),
                      "pressure": sim.download("pressure32").astype(np.float64),
                      "dye": sim.download("dye").astype(np.float64)}
        sim.destroy()
    out = {"steps": steps, "reference": "float32 storage, same scene and steps"}
    for f in ("velocity", "pressure", "dye"):
        a, b = fields[storage][f], fields["f32"][f]
        if f == "pressure":
            a = a - a.mean()
            b = b - b.mean()
        d = a - b
        out[f] = {"max_rel": float(np.abs(d).max() / np.abs(b).max()),
                  "l2_rel": float(np.linalg.norm(d) / np.linalg.norm(b))}
    return out


def bench_mixed(args):
    import torch
    N = 4096 if args.size == 2048 else args.size
    iters = 120 if args.iters == 80 else args.iters
    sim, splats = mixed_scene(N, args.storage, iters)
    mixed_tile = tuple(sim.jacobi_tile)
    mixed_k = mixed_tile[0]
    log(f"bench mixed: {N}^2 {args.storage} I={iters} tile {sim.jacobi_tile}")
    if args.graph:
        sim.use_graph = True
    for _ in range(args.warmup):
        sim.add_splats(splats)
        sim.update(1 / 60)
    sim.synchronize()
    torch.cuda.synchronize()
    sim.set_timing(not args.graph)
    sim.collect_timing()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.add_splats(splats)
        sim.update(1 / 60)
    sim.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    timing = sim.collect_timing()
    sim.destroy()
    err32 = None if args.no_errors else mixed_error_vs_f32(N, args.storage, iters,
                                                           args.error_steps)
    ms = 1e3 * el / args.steps
    cells = N * N
    # bf16/f16 bytes per cell: 54 for the other stages + 7 per Jacobi sweep
    # (p 2 + b 2 + mask 1 in, p 2 out) + splats 12 + the correction launch
    # (p 2 + b 2 + mask 1 in, p32 4 out) + projection's float32 p (2 extra)
    step_bytes = cells * (54 + 7 * (iters - 1) + 12 + 9 + 2)
    step_tbs = step_bytes / (ms * 1e-3) / 1e12
    jac_ms, jac_n = timing["jacobi"]
    roof = None
    if jac_n and jac_ms > 0:
        # km_jacobi: per launch p 2 + b 2 + mask 1 in, p 2 out (7 B/cell);
        # the correction launch reads p 2 + b 2 + mask 1 and writes the
        # float32 pressure 4 (9 B/cell); launches per step from the tile
        k = mixed_k
        per_step_launches = -(-(iters - 1) // k) + 1
        algo = cells * (7 * (per_step_launches - 1) + 9) / per_step_launches
        traffic, tsrc = profile_traffic(f"km_jacobi_{args.storage}_{N}", grid=N,
                                        iterations=iters, tile=list(mixed_tile))
        roof = roofline_entry(
            "km_jacobi (+ the correction launch)", algo, 1e3 * jac_ms / jac_n, traffic, tsrc,
            per_sweep_GBps=round(cells * (7 * (iters - 1) + 9) * args.steps
                                 / (jac_ms * 1e-3) / 1e9, 1))
    out = {"metric": f"sim steps/sec at {N}^2, {args.storage} storage",
           "value": round(args.steps / el, 3), "unit": "steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": f"{args.storage} storage, f32 arithmetic", "data": "synthetic",
           "config": {"workload": f"{N}x{N}, 8 Gaussian splats/step, I={iters} + "
                                  f"8 float32 correction sweeps, free-slip walls, dye",
                      "jacobi_tile": None, "graph": bool(args.graph)},
           "baseline_formula": {"bytes_per_step": int(step_bytes), "TBps": round(step_tbs, 3),
                                "ratio_to_peak": round(step_tbs / (HBM_PEAK_GBS / 1e3), 4)},
           "stage_ms_per_step": {k: round(v[0] / args.steps, 5)
                                 for k, v in timing.items() if v[1]},
           "field_errors_vs_f32_gpu": err32,
           "roofline": roof, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def bench_learned(args, copy_bandwidth):
    """The headline workload: fluid steps at N^2 whose pressure is solved to
    FP64 relative residual <= tol, with the learned two-grid cycle + Jacobi
    (timed: `value`), against plain Jacobi correction solves under the same
    criterion (one step, capped at `plain_cap` sweeps: when the cap is hit
    its rate is an upper bound and the speedup a lower bound).  The roofline
    of the Jacobi kernel comes from a fixed-sweep phase of the plain fused
    step on the same grid (exact sweep count per launch); the CPU oracle is
    timed on fixed-sweep steps of the same scene and scaled by the plain
    path's sweeps per step."""
    import torch
    from natrix_amd import FluidSimulator
    from natrix_amd.learned.solver import LearnedGuess, solve_step
    from natrix_amd.learned.train import DEFAULT_WEIGHTS, load
    N = args.size
    net, meta = load(args.weights or DEFAULT_WEIGHTS)
    guess = LearnedGuess(net, coarse=meta.get("size", 128), mid=args.mid,
                         mid_iters=args.mid_iters)
    smooth_default = args.smooth or (CHEB_SMOOTH_3 if args.mid > 1 else CHEB_SMOOTH)
    solid, centre = make_scene(N)
    dt = 1.0 / 60.0
    cells = N * N

    def fresh(develop=None):
        sim = FluidSimulator(N, N)
        sim.vorticity = 0.3
        sim.viscosity = 0.0
        sim.iterations = args.iters
        sim.bc_mode = "inflow"
        sim.inflow = 1.0
        sim.set_obstacle_map(solid)
        if args.tile:
            sim.jacobi_tile = tuple(args.tile)
        if args.smoother_tile:
            sim.smoother_tile = tuple(args.smoother_tile)
        for _ in range(args.warmup if develop is None else develop):
            sim.update(dt)
        sim.synchronize()
        return sim

    def timed_learned(sim, steps, g=guess, smooth=None):
        smooth = smooth or smooth_default
        for _ in range(2):  # first-call costs (allocator, code objects) off the clock
            solve_step(sim, g, dt, tol=args.tol, smooth=smooth)
        sim.synchronize()
        torch.cuda.synchronize()
        # no per-launch events in this timed region (they cost ~10 % of a
        # learned step); the kernels' rooflines are taken in phases of their own
        sweeps, outer, conv = [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            st = solve_step(sim, g, dt, tol=args.tol, smooth=smooth)
            sweeps.append(st["sweeps"])
            outer.append(st["outer"])
            conv.append(st["converged"] and st["residual"] <= args.tol)
        sim.synchronize()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ms = 1e3 * el / steps
        K = float(np.mean(sweeps))
        return {"steps_per_s": round(steps / el, 3), "ms_per_step": round(ms, 3),
                "sweeps_per_step": K, "outer_per_step": float(np.mean(outer)),
                "all_converged": bool(all(conv)), "timed_steps": steps,
                "baseline_formula": baseline_formula(cells, K, ms)}

    log(f"bench learned: {N}x{N}, warmup {args.warmup} fixed steps (I={args.iters})")
    sim = fresh()
    learned = timed_learned(sim, args.steps)
    sim.destroy()
    learned["flow"] = f"steps {args.warmup + 3}..{args.warmup + 2 + args.steps} from rest"
    log(f"bench learned: {learned}")

    developed = None
    if args.develop > 0:
        # the same scene after `develop` fixed-iteration steps (the wake
        # behind the cylinder has formed), then learned steps
        sim = fresh(args.develop)
        developed = timed_learned(sim, args.developed_steps)
        sim.destroy()
        developed["flow"] = (f"steps {args.develop + 3}..{args.develop + 2 + args.developed_steps}"
                             f" ({args.develop} fixed-iteration steps of development)")
        log(f"bench learned, developed flow: {developed}")

    classical = None
    if args.classical_coarse > 0:
        # the same two-grid cycle with the network off: the coarse problem
        # is smoothed from zero with more Chebyshev iterations instead (the
        # method without the learned guess: 350 coarse iterations, the
        # learned cycle's shape and fine iterations otherwise)
        g0 = LearnedGuess(net, coarse=meta.get("size", 128),
                          coarse_iters=args.classical_coarse, use_network=False, mid=args.mid,
                          mid_iters=args.mid_iters)
        sim = fresh()
        classical = timed_learned(sim, args.steps, g0, smooth=args.classical_smooth)
        sim.destroy()
        classical["coarse_iterations"] = args.classical_coarse
        classical["fine_iterations_per_cycle"] = args.classical_smooth or smooth_default
        classical["flow"] = learned["flow"]
        log(f"bench two-grid without the network: {classical}")

    plain = None
    if not args.no_plain:
        sim = fresh()
        t0 = time.perf_counter()
        st = solve_step(sim, None, dt, tol=args.tol, max_inner_sweeps=args.plain_cap,
                        max_outer=args.plain_outer)
        sim.synchronize()
        t_p = time.perf_counter() - t0
        sim.destroy()
        done = st["converged"] and st["residual"] <= args.tol
        plain = {"steps_per_s": round(1.0 / t_p, 6), "ms_per_step": round(1e3 * t_p, 1),
                 "sweeps_per_step": st["sweeps"], "converged": bool(done),
                 "residual_reached": st["residual"], "sweep_cap_per_correction_solve": args.plain_cap,
                 "rate_is_upper_bound": not done,
                 "baseline_formula": baseline_formula(cells, st["sweeps"], 1e3 * t_p)}
        log(f"bench plain: {plain}")

    # the learned step's dominant kernel: the fine Chebyshev smoother, timed
    # per launch with HIP events on the simulator's stream over calls of the
    # step's own length (iterations per call, the simulator's smoother tile)
    import ctypes as C
    from natrix_amd import _lib
    sim = fresh()
    sim._push_params()
    _lib.call("nx_step_begin", sim._alive(), dt)
    lmin = guess.fine_lmin(N, N)
    n_iter, calls = smooth_default, 20
    tile = (C.c_int * 5)()
    _lib.call("nx_get_smoother_tile", sim._alive(), tile)
    tile = tuple(tile) if tile[4] > 0 else tuple(tile)[:4]
    _lib.call("nx_smooth_chebyshev", sim._alive(), n_iter, lmin, None, None)
    sim.set_timing(True)
    sim.collect_timing()
    for _ in range(calls):
        _lib.call("nx_smooth_chebyshev", sim._alive(), n_iter, lmin, None, None)
    tc = sim.collect_timing()
    sim.set_timing(False)
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()
    ch_ms, ch_n = tc["chebyshev"]
    plan = cheb_launch_plan(n_iter, tile[0])
    assert ch_n == calls * len(plan), (ch_n, len(plan))
    algo_call = cells * sum(b for _, b in plan)
    kname = "k_cheb_stream" if len(tile) > 4 else "k_cheb_tb"
    traffic_c, tsrc_c = committed_traffic(N, n_iter, tile, kname, "iters")
    roof_cheb = roofline_entry(
        kname, algo_call / len(plan), 1e3 * ch_ms / ch_n, traffic_c, tsrc_c,
        valu=valu_entry(f"{kname}_{N}_iters{n_iter}", 1e3 * ch_ms / ch_n, grid=N,
                        iterations=n_iter, tile=list(tile)),
        iterations_per_launch=round(n_iter / len(plan), 3),
        algorithmic_bytes_note=f"per call of {n_iter} iterations: the first launch reads x, b, mask "
                               "and writes x, r, d (21 B/cell), the middle ones read x, r, d, "
                               "mask and write x, r, d (25 B/cell), the last reads x, r, d, b, "
                               "mask and writes x (21 B/cell); averaged per launch",
        per_iteration_GBps=round(CHEB_BYTES_PER_CELL_ITER * cells * n_iter * calls
                                 / (ch_ms * 1e-3) / 1e9, 1),
        per_iteration_note="25 B per cell-iteration as if every iteration went to HBM "
                           "(temporal blocking keeps x, r, d on chip between passes)",
        lambda_min=lmin, tile=list(tile))
    log(f"bench chebyshev: {roof_cheb}")

    # fixed-sweep plain fused step: exact sweeps per Jacobi launch
    sim = fresh()
    sim.set_timing(True)
    sim.collect_timing()
    t0 = time.perf_counter()
    for _ in range(args.fixed_steps):
        sim.update(dt)
    sim.synchronize()
    el_f = time.perf_counter() - t0
    tf = sim.collect_timing()
    sim.set_timing(False)
    jtile = tuple(sim.jacobi_tile)
    sim.destroy()
    jac_ms, jac_n = tf["jacobi"]
    tb_sweeps = (args.iters - 1) * args.fixed_steps  # sweep 1 runs in the fused divergence kernel
    jname = "k_jacobi_stream" if len(jtile) > 4 and jtile[4] > 0 else "k_jacobi_tb"
    traffic, tsrc = committed_traffic(N, args.iters, jtile, jname)
    roof = roofline_entry(
        jname, JACOBI_BYTES_PER_CELL_SWEEP * cells, 1e3 * jac_ms / jac_n, traffic, tsrc,
        valu=valu_entry(f"{jname}_{N}_fixed{args.iters}", 1e3 * jac_ms / jac_n, grid=N,
                        iterations=args.iters, tile=list(jtile)),
        sweeps_per_launch=round(tb_sweeps / jac_n, 3),
        algorithmic_bytes_note="13 B per cell per launch: read p, b, mask, write p",
        per_sweep_GBps=round(JACOBI_BYTES_PER_CELL_SWEEP * cells * tb_sweeps
                             / (jac_ms * 1e-3) / 1e9, 1),
        tile=list(jtile))
    ms_f = 1e3 * el_f / args.fixed_steps
    fixed = {"steps_per_s": round(args.fixed_steps / el_f, 2), "ms_per_step": round(ms_f, 4),
             "iterations": args.iters, "jacobi_tile": list(jtile),
             "baseline_formula": baseline_formula(cells, args.iters, ms_f),
             "stage_ms_per_step": {k: round(v[0] / args.fixed_steps, 5) for k, v in tf.items() if v[1]}}
    log(f"bench fixed: {fixed} roofline {roof}")
    copy_gbs = copy_bandwidth(1 << 30, 20)
    attainable = copy_bandwidth(1 << 30, 20, -1)

    cpu = None
    if not args.no_cpu:
        log("bench: CPU oracle baseline")
        sps, nst, el_c = cpu_baseline(N, args.iters, args.cpu_budget)
        cpu_sweeps_per_s = sps * args.iters
        plain_sweeps = plain["sweeps_per_step"] if plain else None
        cpu = {"value": (float(f"{cpu_sweeps_per_s / plain_sweeps:.4g}") if plain_sweeps else None),
               "unit": "steps/s", "cores": 1, "kind": "port", "derived": True,
               "sample": f"oracle/fluid.py step() on the same {N}x{N} scene with {args.iters} "
                         f"Jacobi sweeps: {nst} steps in {el_c:.1f} s = {cpu_sweeps_per_s:.1f} "
                         f"sweeps/s (NumPy float32, one thread; host reports {os.cpu_count()} "
                         f"cpus); value is derived, not timed to the tolerance: that rate / "
                         f"the plain path's sweeps per step to the tolerance ({plain_sweeps}"
                         + (", a lower bound: capped" if plain and plain["rate_is_upper_bound"] else "")
                         + ") = the CPU's plain-Jacobi steps/s at the tolerance",
               "measured_sweeps_per_s": round(cpu_sweeps_per_s, 2),
               "fixed_iteration_steps_per_s": round(sps, 4),
               "threadpools": threadpools(),
               "config_256_200_steps": config_256()}

    speedup = (learned["steps_per_s"] / plain["steps_per_s"]) if plain else None
    out = {"metric": METRIC, "value": learned["steps_per_s"], "unit": "steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": learned["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": f"fluid step {N}x{N} f32, cylinder r=0.1N at "
                                  f"({centre[0]:.3f},{centre[1]:.3f}), left inflow 1.0, vorticity 0.3, "
                                  f"pressure to FP64 rel. residual <= {args.tol:g}",
                      "grid": N,
                      "solver": (f"learned three-level cycle ({N}² -> {N // args.mid}² -> "
                                 f"network grid; U-Net bf16 MFMA guess, face-weighted coarse "
                                 f"operators, Chebyshev smoothing: {args.mid_iters} middle and "
                                 f"{smooth_default} fine iterations per cycle), FP64 refinement"
                                 if args.mid > 1 else
                                 "learned two-grid cycle (U-Net bf16 MFMA guess, face-weighted "
                                 "coarse operator, Chebyshev coarse and fine smoothing), FP64 "
                                 "refinement"),
                      "weights": os.path.basename(args.weights or DEFAULT_WEIGHTS)},
           "learned": learned, "learned_developed_flow": developed,
           "two_grid_without_network": classical, "plain": plain,
           "speedup_learned_vs_plain": round(speedup, 1) if speedup else None,
           "speedup_is_lower_bound": bool(plain and plain["rate_is_upper_bound"]),
           "plain_fixed_iterations": fixed, "copy_probe_GBps": round(copy_gbs, 1),
           "attainable_copy_GBps": round(attainable, 1),
           "roofline": roof_cheb, "roofline_plain_jacobi": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 20 learned, 1000 single/batch/mixed)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default 10 learned, 20 otherwise)")
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--iters", type=int, default=80)
    ap.add_argument("--mode", choices=["fixed", "tol"], default="fixed")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--max-iters", type=int, default=200000)
    ap.add_argument("--check-every", type=int, default=1024)
    ap.add_argument("--tile", type=lambda s: [int(v) for v in s.split(",")],
                    default=None)
    ap.add_argument("--smoother-tile", type=lambda s: [int(v) for v in s.split(",")],
                    default=None,
                    help="fine Chebyshev smoother launch shape k,rows,waves,cols[,seg] "
                         "(learned workload; default: the autotuned shape)")
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["learned", "single", "batch", "mixed"],
                    default="learned",
                    help="learned: steps to rel. residual <= tol, learned two-grid cycle "
                         "vs plain Jacobi (the headline); single: the plain fused step at "
                         "fixed iterations; batch: 64 uneven scenes in one launch per "
                         "stage; mixed: 4096^2 with bf16/f16 storage")
    ap.add_argument("--storage", choices=["bf16", "f16"], default="bf16")
    ap.add_argument("--no-errors", action="store_true",
                    help="mixed workload: skip the error rollouts (for profiling)")
    ap.add_argument("--error-steps", type=int, default=500,
                    help="mixed workload: steps of the rollouts compared with float32")
    ap.add_argument("--weights", default=None, help="U-Net weights (learned workload)")
    ap.add_argument("--plain-cap", type=int, default=1_000_000,
                    help="sweep cap per correction solve of the plain path (learned workload)")
    ap.add_argument("--plain-outer", type=int, default=2,
                    help="refinement iterations of the capped plain step (learned workload)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--develop", type=int, default=500,
                    help="learned workload: also time learned steps after this many "
                         "fixed-iteration steps of development (0: skip)")
    ap.add_argument("--classical-coarse", type=int, default=350,
                    help="learned workload: also time the cycle without the network, "
                         "with this many coarse Chebyshev iterations (0: skip)")
    ap.add_argument("--classical-smooth", type=int, default=0,
                    help="fine Chebyshev iterations per cycle of the cycle without the network "
                         "(0: as the learned cycle)")
    ap.add_argument("--mid", type=int, default=4,
                    help="learned workload: middle grid factor of the three-level cycle "
                         "(0: the two-level cycle)")
    ap.add_argument("--mid-iters", type=int, default=12,
                    help="Chebyshev iterations on the middle grid per cycle")
    ap.add_argument("--smooth", type=int, default=0,
                    help=f"fine Chebyshev iterations per cycle (0: {CHEB_SMOOTH_3} with --mid, "
                         f"{CHEB_SMOOTH} without)")
    ap.add_argument("--developed-steps", type=int, default=10,
                    help="learned workload: timed steps of the developed-flow leg")
    ap.add_argument("--fixed-steps", type=int, default=100,
                    help="fixed-iteration steps timed for the Jacobi roofline (learned workload)")
    ap.add_argument("--probe-copy", type=int, default=0, metavar="BYTES",
                    help="only run the copy-bandwidth probe (counter calibration)")
    ap.add_argument("--probe-variant", type=int, default=0,
                    help="copy probe variant (0: float4 grid-stride, 5/6/7/8: 8/4/2/1-byte "
                         "accesses; -1: the best of all)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 20 if args.workload == "learned" else 1000
    if args.warmup is None:
        args.warmup = 10 if args.workload == "learned" else 20
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the fluid "
                          "step of one grid is not sharded (run with --gpus 1)"}))
        sys.exit(2)

    import ctypes

    import torch
    from natrix_amd import FluidSimulator, _lib

    def copy_bandwidth(nbytes, iters, variant=0):
        """GB/s of the counter-calibration copy (variant 0) or, variant -1,
        the best of the probe's copy variants (the attainable rate)."""
        g = ctypes.c_double()
        _lib.call("nx_copy_bandwidth_variant", nbytes, iters, variant, ctypes.byref(g))
        return g.value

    if args.probe_copy:
        gbs = copy_bandwidth(args.probe_copy, 10, args.probe_variant)
        print(json.dumps({"probe": "copy", "bytes": args.probe_copy,
                          "variant": args.probe_variant,
                          "launches": 11, "GBps": round(gbs, 1)}), flush=True)
        return

    if args.workload == "batch":
        return bench_batch(args, copy_bandwidth)
    if args.workload == "mixed":
        return bench_mixed(args)
    if args.workload == "learned":
        return bench_learned(args, copy_bandwidth)

    N = args.size
    solid, centre = make_scene(N)
    log(f"bench: {N}x{N} mode={args.mode} iters={args.iters} tile={args.tile}")
    sim = FluidSimulator(N, N)
    configure(sim, args)
    sim.set_obstacle_map(solid)
    if args.graph:
        sim.use_graph = True
    dt = 1.0 / 60.0
    for _ in range(args.warmup):
        sim.update(dt)
    sim.synchronize()
    torch.cuda.synchronize()
    log("bench: warmup done")

    sim.set_timing(not args.graph)
    sim.collect_timing()
    sweeps_total = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        sim.update(dt)
        if args.mode == "tol":
            sweeps_total += sim.stats()["sweeps"]
    sim.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timing = sim.collect_timing()
    sim.set_timing(False)
    if args.mode == "fixed":
        sweeps_total = args.iters * args.steps
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    rel, _ = sim.residual()
    st = sim.stats()
    log(f"bench: {args.steps} steps in {elapsed:.3f}s, residual {rel:.3e}")

    cells = N * N
    jac_ms, jac_launches = timing["jacobi"]
    # the first sweep of every step runs inside the fused divergence kernel
    tb_sweeps = sweeps_total - args.steps
    roof = None
    if jac_launches > 0 and jac_ms > 0:
        jtile = tuple(sim.jacobi_tile)
        jname = "k_jacobi_stream" if len(jtile) > 4 and jtile[4] > 0 else "k_jacobi_tb"
        traffic, tsrc = (committed_traffic(N, args.iters, jtile, jname)
                         if args.mode == "fixed" else (None, None))
        roof = roofline_entry(
            jname, JACOBI_BYTES_PER_CELL_SWEEP * cells, 1e3 * jac_ms / jac_launches, traffic,
            tsrc, sweeps_per_launch=round(tb_sweeps / jac_launches, 3),
            per_sweep_GBps=round(JACOBI_BYTES_PER_CELL_SWEEP * cells * tb_sweeps
                                 / (jac_ms * 1e-3) / 1e9, 1))

    copy_gbs = copy_bandwidth(1 << 30, 20)
    attainable = copy_bandwidth(1 << 30, 20, -1)
    log(f"bench: copy probe {copy_gbs:.0f} GB/s, best copy variant {attainable:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        log("bench: CPU oracle baseline")
        sps, nst, el = cpu_baseline(N, args.iters, args.cpu_budget)
        cpu = {"value": round(sps, 5), "unit": "steps/s", "cores": 1,
               "kind": "port",
               "sample": f"oracle/fluid.py step() on the same {N}x{N} scene, "
                         f"{args.iters} sweeps, {nst} steps in {el:.1f} s, "
                         f"NumPy float32 single thread "
                         f"(host reports {os.cpu_count()} cpus)"}

    out = {
        "metric": METRIC,
        "value": round(args.steps / elapsed, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"fluid step {N}x{N} f32, circular obstacle r=0.1N at "
                         f"({centre[0]:.3f},{centre[1]:.3f}), left inflow 1.0, "
                         f"vorticity 0.3, "
                         + (f"{args.iters} Jacobi sweeps/step" if args.mode == "fixed"
                            else f"Jacobi to rel. residual <= {args.tol:g}")),
            "grid": N, "solver": f"plain fused Jacobi ({args.mode})",
            "jacobi_tile": list(sim.jacobi_tile), "graph": bool(args.graph),
        },
        "sweeps_per_step": round(sweeps_total / args.steps, 2),
        "final_rel_residual": rel,
        "baseline_formula": baseline_formula(cells, sweeps_total / args.steps, ms_per_step),
        "stage_ms_per_step": {k: round(v[0] / args.steps, 5)
                              for k, v in timing.items() if v[1]},
        "copy_probe_GBps": round(copy_gbs, 1),
        "attainable_copy_GBps": round(attainable, 1),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if args.mode == "tol":
        out["converged_last_step"] = st["converged"]
    sim.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
