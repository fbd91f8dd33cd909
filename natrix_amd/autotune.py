"""Autotuner for the temporal-blocked Jacobi kernel's tile shape.

The search space is every (k, rows, waves, cols) the library instantiates
(k sweeps per launch on a (64*cols) x (rows*waves) register tile).  Each
candidate runs a fixed number of sweeps on a synthetic scene of the target
size and is timed with HIP events on the simulator's stream; the fastest
per sweep wins.  Results are cached per (architecture, dtype, grid) in
``tuned.json`` next to this file, which ``FluidSimulator`` consults when it
is created.  A batch of scenes is tuned as a whole (``tune_batch``): its
best tile depends on the mix of scene sizes, not on the total cell count,
and is cached under ``batch_key(sizes)`` for ``BatchedScenes``.

    python -m natrix_amd.autotune 2048 1024 512   # tunes and updates tuned.json
    python -m natrix_amd.autotune --batch 64      # the 64-scene benchmark batch

The fine Chebyshev smoother of the learned cycle is tuned per grid as well
(``tune_smoother``: tile and streaming launch shapes, timed on the
smoothing call the cycle makes) and cached under "cheb".
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuned.json")
ARCH = "gfx950"

# (rows, waves, cols) instantiated in csrc/nx_jacobi.hip (NX_TB_SHAPES)
SHAPES = [(4, 8, 1), (4, 16, 1), (8, 8, 1), (8, 16, 1), (16, 8, 1), (2, 8, 2), (2, 16, 2),
          (4, 4, 2), (4, 8, 2), (4, 16, 2), (8, 4, 2), (8, 8, 2)]
KS = (2, 4, 6, 8, 10, 12, 16)
DEFAULT = (6, 4, 8, 2)


def supported(k, rows, waves, cols):
    return (rows, waves, cols) in SHAPES and k >= 1 and \
        64 * cols - 2 * k >= 8 and rows * waves - 2 * k >= 8


def candidates():
    return [(k, r, w, c) for (r, w, c) in SHAPES for k in KS if supported(k, r, w, c)]


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(W, H, dtype="f32"):
    """Tuned tile for a W x H grid with `dtype` storage: the exact entry of
    that storage, else the exact float32 entry, else the float32 entry with
    the nearest cell count, else the built-in default."""
    data = _load().get(ARCH, {})
    key = f"{W}x{H}"
    if key in data.get(dtype, {}):
        return tuple(data[dtype][key]["tile"])
    table = data.get("f32", {})
    if key in table:
        return tuple(table[key]["tile"])
    best, dist = None, None
    for k, v in table.items():
        if k.startswith("batch"):
            continue
        w, h = (int(t) for t in k.split("x"))
        d = abs((w * h) - W * H)
        if dist is None or d < dist:
            best, dist = tuple(v["tile"]), d
    return best if best is not None else DEFAULT


SMOOTHER_DEFAULT = (8, 4, 8, 2)


def lookup_smoother(W, H):
    """Tuned launch shape of the fine Chebyshev smoother for a W x H grid
    (the entry of that grid, else of the nearest cell count, else the
    library's default tile)."""
    table = _load().get(ARCH, {}).get("cheb", {})
    key = f"{W}x{H}"
    if key in table:
        return tuple(table[key]["tile"])
    best, dist = None, None
    for k, v in table.items():
        w, h = (int(t) for t in k.split("x"))
        d = abs(w * h - W * H)
        if dist is None or d < dist:
            best, dist = tuple(v["tile"]), d
    return best if best is not None else SMOOTHER_DEFAULT


def tune(W, H, dtype="f32", sweeps_per_candidate=None, log=print):
    """Time every candidate on a W x H scene; returns (best tile, results)."""
    import numpy as np

    from .core.fluid_simulator import FluidSimulator
    sim = FluidSimulator(W, H)
    X = np.arange(W)[None, :]
    Y = np.arange(H)[:, None]
    solid = (X - 0.4 * W) ** 2 + (Y - 0.42 * H) ** 2 <= (0.1 * min(W, H)) ** 2
    sim.set_obstacle_map(solid)
    rng = np.random.default_rng(0)
    sim.upload("divergence", (0.01 * rng.standard_normal((H, W))).astype(np.float32))
    if sweeps_per_candidate is None:
        # about 1e9 cell-sweeps per candidate, at least 48 sweeps
        sweeps_per_candidate = max(48, int(1e9 // (W * H)))
    results = []
    try:
        for (k, r, w, c) in candidates():
            n = max(k, (sweeps_per_candidate // k) * k)
            sim.jacobi_blocked(2 * k, k, r, w, c)  # warm-up
            sim.synchronize()
            sim.set_timing(True)
            sim.collect_timing()
            sim.jacobi_blocked(n, k, r, w, c)
            ms, launches = sim.collect_timing()["jacobi"]
            sim.set_timing(False)
            us = 1e3 * ms / n
            results.append((us, (k, r, w, c)))
            log(f"  {W}x{H} k={k:2d} rows={r:2d} waves={w:2d} cols={c}: "
                f"{us:8.3f} us/sweep")
    finally:
        sim.destroy()
    results.sort()
    return results[0][1], results


def store(W, H, tile, us_per_sweep, dtype="f32"):
    data = _load()
    data.setdefault(ARCH, {}).setdefault(dtype, {})[f"{W}x{H}"] = {
        "tile": list(tile), "us_per_sweep": round(us_per_sweep, 4)}
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def batch_key(sizes):
    """Cache key of a batch: scene count and total cells."""
    return f"batch{len(sizes)}:{sum(w * h for w, h in sizes)}"


def lookup_batch(sizes, dtype="f32"):
    """Tuned tile of a batch of (W, H) scenes: its own entry, else the entry
    of the single grid nearest its total cell count."""
    table = _load().get(ARCH, {}).get(dtype, {})
    key = batch_key(sizes)
    if key in table:
        return tuple(table[key]["tile"])
    side = int(round(sum(w * h for w, h in sizes) ** 0.5))
    return lookup(side, side, dtype)


def tune_batch(batch, sweeps_per_candidate=240, log=print):
    """Time every candidate on the batch's current state (HIP events on its
    stream); returns (best tile, results).  The scenes' pressure fields are
    swept, nothing else changes."""
    results = []
    for (k, r, w, c) in candidates():
        n = max(k, (sweeps_per_candidate // k) * k)
        batch.jacobi_blocked(2 * k, k, r, w, c)  # warm-up
        batch.synchronize()
        batch.set_timing(True)
        batch.collect_timing()
        batch.jacobi_blocked(n, k, r, w, c)
        ms, launches = batch.collect_timing()["jacobi"]
        batch.set_timing(False)
        us = 1e3 * ms / n
        results.append((us, (k, r, w, c)))
        log(f"  batch k={k:2d} rows={r:2d} waves={w:2d} cols={c}: {us:8.3f} us/sweep")
    results.sort()
    return results[0][1], results


def store_batch(sizes, tile, us_per_sweep, dtype="f32"):
    data = _load()
    data.setdefault(ARCH, {}).setdefault(dtype, {})[batch_key(sizes)] = {
        "tile": list(tile), "us_per_sweep": round(us_per_sweep, 4)}
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def store_step_timed(W, H, tile, ms_per_step, dtype):
    """Entry for reduced storage, chosen by whole-step time (its Jacobi
    launches include the float32 correction launch)."""
    data = _load()
    data.setdefault(ARCH, {}).setdefault(dtype, {})[f"{W}x{H}"] = {
        "tile": list(tile), "ms_per_step": round(ms_per_step, 4)}
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


# the fine Chebyshev smoother's launch shapes: tile kernels (seg = 0;
# NX_CHEB_SHAPES in csrc/nx_cheb.hip) and streaming kernels (seg > 0;
# NX_STREAM_SHAPES in csrc/nx_stream.hip)
CHEB_SHAPES = [(2, 16, 2), (4, 8, 2), (4, 16, 2), (4, 16, 1), (8, 8, 1), (8, 16, 1)]
CHEB_KS = (4, 6, 8, 10, 12, 16)
STREAM_SHAPES = [(4, 8, 2), (2, 16, 2), (4, 16, 2), (8, 8, 2), (8, 4, 2)]
STREAM_KS = (8, 12, 16)
STREAM_SEGS = (48, 80, 112, 144, 208)


def smoother_candidates():
    out = []
    for (r, w, c) in CHEB_SHAPES:
        for k in CHEB_KS:
            if 64 * c - 2 * k >= 8 and r * w - 2 * k >= 8:
                out.append((k, r, w, c, 0))
    for (r, w, c) in STREAM_SHAPES:
        for k in STREAM_KS:
            for seg in STREAM_SEGS:
                out.append((k, r, w, c, seg))
    return out


def tune_smoother(W, H, iterations=72, lambda_min=0.006, calls=10, log=print):
    """Time the fine Chebyshev smoother (nx_smooth_chebyshev, `iterations`
    per call, as the learned cycle runs it) with every launch shape on a
    W x H scene inside a step; returns (best shape, results)."""
    import ctypes as C

    import numpy as np

    from . import _lib
    from .core.fluid_simulator import FluidSimulator
    sim = FluidSimulator(W, H, bc_mode="inflow")
    X = np.arange(W)[None, :]
    Y = np.arange(H)[:, None]
    sim.set_obstacle_map((X - 0.4 * W) ** 2 + (Y - 0.42 * H) ** 2 <= (0.1 * min(W, H)) ** 2)
    for _ in range(3):
        sim.update(1 / 60)
    results = []
    sim._push_params()
    _lib.call("nx_step_begin", sim._alive(), 1 / 60)
    try:
        for cfg in smoother_candidates():
            try:
                sim.smoother_tile = cfg
            except ValueError:
                continue
            _lib.call("nx_smooth_chebyshev", sim._alive(), iterations, lambda_min, None, None)
            sim.set_timing(True)
            sim.collect_timing()
            for _ in range(calls):
                _lib.call("nx_smooth_chebyshev", sim._alive(), iterations, lambda_min, None,
                          None)
            ms, launches = sim.collect_timing()["chebyshev"]
            sim.set_timing(False)
            us = 1e3 * ms / (calls * iterations)
            results.append((us, cfg))
            log(f"  {W}x{H} smoother {cfg}: {us:8.3f} us/iteration")
    finally:
        _lib.call("nx_step_finish", sim._alive())
        sim.destroy()
    results.sort()
    return results[0][1], results


def store_smoother(W, H, tile, us_per_iteration):
    data = _load()
    data.setdefault(ARCH, {}).setdefault("cheb", {})[f"{W}x{H}"] = {
        "tile": list(tile), "us_per_iteration": round(us_per_iteration, 4)}
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def main(argv):
    """python -m natrix_amd.autotune [--batch N] [--no-smoother] SIZE...:
    the Jacobi tile and the Chebyshev smoother's launch shape of every
    SIZE x SIZE grid, and with --batch N the Jacobi tile of the benchmark's
    batch of N scenes."""
    import argparse
    ap = argparse.ArgumentParser(prog="natrix_amd.autotune")
    ap.add_argument("sizes", type=int, nargs="*")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--no-smoother", action="store_true")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    for n in a.sizes or ([] if a.batch else [2048]):
        best, res = tune(n, n, log=log)
        store(n, n, best, res[0][0])
        log(f"{n}x{n} jacobi: {best} {res[0][0]:.4f} us/sweep")
        if not a.no_smoother:
            best, res = tune_smoother(n, n, log=log)
            store_smoother(n, n, best, res[0][0])
            log(f"{n}x{n} smoother: {best} {res[0][0]:.4f} us/iteration")
    if a.batch:
        from .batch import benchmark_scenes
        b, sides, _, _ = benchmark_scenes(a.batch)
        try:
            best, res = tune_batch(b, log=log)
            store_batch(list(zip(sides, sides)), best, res[0][0])
            log(f"batch of {a.batch}: {best} {res[0][0]:.4f} us/sweep")
        finally:
            b.destroy()
        print(f"best {n}x{n}: tile {best} {res[0][0]:.3f} us/sweep", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
