"""``BatchedScenes``: many independent fluid grids advanced together.

Each scene behaves exactly like its own ``FluidSimulator`` (same parameters,
same step), but every stage of every scene runs in one persistent HIP launch
(see csrc/nx_batch.hip).  Scenes may have uneven sizes and their own
parameter blocks, obstacle maps and per-step Gaussian splats.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FIELD_DIVERGENCE, FIELD_DYE, FIELD_MASK, FIELD_PRESSURE, FIELD_VELOCITY, Params

_FIELDS = {"velocity": FIELD_VELOCITY, "pressure": FIELD_PRESSURE,
           "divergence": FIELD_DIVERGENCE, "mask": FIELD_MASK, "dye": FIELD_DYE}
_BC = {"reference": 0, "free_slip": 1, "inflow": 2}
_KEYS = ("speed", "dissipation", "vorticity", "viscosity", "iterations",
         "has_borders", "bc_mode", "inflow", "dye_speed", "dye_dissipation")


def _params(base: Params, kw) -> Params:
    p = Params()
    C.memmove(C.byref(p), C.byref(base), C.sizeof(Params))
    for k, v in kw.items():
        if k == "jacobi_tile":
            p.jacobi_k, p.jacobi_rows, p.jacobi_waves, p.jacobi_cols = (int(t) for t in v)
            continue
        if k not in _KEYS:
            raise ValueError(f"unknown parameter {k!r}")
        if k == "bc_mode" and isinstance(v, str):
            v = _BC[v]
        if k == "has_borders":
            v = 1 if v else 0
        setattr(p, k, v)
    return p


class BatchedScenes:
    def __init__(self, widths, heights=None, dye=False):
        widths = [int(w) for w in widths]
        heights = widths if heights is None else [int(h) for h in heights]
        if len(widths) != len(heights) or not widths:
            raise ValueError("widths and heights must be non-empty and equal length")
        self.sizes = list(zip(widths, heights))
        self.dye = bool(dye)
        n = len(widths)
        h = C.c_void_p()
        _lib.call("nx_batch_create", n, (C.c_int * n)(*widths), (C.c_int * n)(*heights),
                  1 if dye else 0, C.byref(h))
        self._h = h
        self._base = Params()
        _lib.call("nx_default_params", C.byref(self._base))
        from .autotune import lookup_batch
        self.set_params(jacobi_tile=lookup_batch(self.sizes))

    def __len__(self):
        return len(self.sizes)

    def _alive(self):
        if self._h is None:
            raise _lib.NatrixError("batch has been destroyed")
        return self._h

    @property
    def cells(self):
        c = C.c_longlong()
        _lib.call("nx_batch_cells", self._alive(), C.byref(c))
        return c.value

    @property
    def jacobi_tile(self):
        b = self._base
        return (b.jacobi_k, b.jacobi_rows, b.jacobi_waves, b.jacobi_cols)

    def set_params(self, **kw):
        """Solver settings (iterations, jacobi_tile) and every scene's
        parameters."""
        p = _params(self._base, kw)
        _lib.call("nx_batch_set_params", self._alive(), C.byref(p))
        self._base = p

    def set_scene_params(self, scene, **kw):
        p = _params(self._base, kw)
        _lib.call("nx_batch_set_scene_params", self._alive(), int(scene), C.byref(p))

    def set_obstacle_map(self, scene, solid):
        w, h = self.sizes[scene]
        m = np.ascontiguousarray(np.asarray(solid) != 0, dtype=np.uint8)
        if m.shape != (h, w):
            raise ValueError("obstacle map must have shape (height, width)")
        _lib.call("nx_batch_set_obstacle_map", self._alive(), int(scene),
                  m.ctypes.data_as(C.POINTER(C.c_uint8)))

    def set_step_splats(self, splats):
        """Gaussian splats applied at the start of every update: iterable of
        (scene, px, py, ax, ay, radius, dye)."""
        a = np.ascontiguousarray(np.asarray(splats, np.float32).reshape(-1, 7))
        _lib.call("nx_batch_set_step_splats", self._alive(),
                  a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[0])

    def update(self, dt):
        _lib.call("nx_batch_update", self._alive(), float(dt))

    @property
    def use_graph(self):
        return getattr(self, "_graph", False)

    @use_graph.setter
    def use_graph(self, on):
        self._graph = bool(on)
        _lib.call("nx_batch_set_graph", self._alive(), int(self._graph))

    def add_circle_obstacle(self, scene, position, radius, static=False):
        """FluidSimulator.add_circle_obstacle for one scene (this step)."""
        _lib.call("nx_batch_add_circle_obstacle", self._alive(), int(scene),
                  float(position[0]), float(position[1]), float(radius), 1 if static else 0)

    def add_triangle_obstacle(self, scene, p1, p2, p3, static=False):
        """FluidSimulator.add_triangle_obstacle for one scene (this step)."""
        _lib.call("nx_batch_add_triangle_obstacle", self._alive(), int(scene),
                  *(float(v) for v in (*p1, *p2, *p3)), 1 if static else 0)

    def jacobi_blocked(self, sweeps, k, rows, waves, cols):
        _lib.call("nx_batch_jacobi_blocked", self._alive(), int(sweeps), int(k),
                  int(rows), int(waves), int(cols))

    def _shape(self, scene, fid):
        w, h = self.sizes[scene]
        if fid == FIELD_VELOCITY:
            return (h, w, 2), np.float32
        if fid == FIELD_MASK:
            return (h, w), np.uint8
        return (h, w), np.float32

    def download(self, scene, name):
        fid = _FIELDS[name]
        shape, dt = self._shape(scene, fid)
        out = np.empty(shape, dt)
        _lib.call("nx_batch_download", self._alive(), int(scene), fid,
                  out.ctypes.data_as(C.c_void_p), out.nbytes)
        return out

    def upload(self, scene, name, array):
        fid = _FIELDS[name]
        shape, dt = self._shape(scene, fid)
        a = np.ascontiguousarray(np.asarray(array, dtype=dt))
        if a.shape != shape:
            raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        _lib.call("nx_batch_upload", self._alive(), int(scene), fid,
                  a.ctypes.data_as(C.c_void_p), a.nbytes)

    def residuals(self):
        """Relative FP64 residual of every scene's current pressure."""
        n = len(self.sizes)
        out = (C.c_double * (6 * n))()
        _lib.call("nx_batch_residuals", self._alive(), out)
        return np.array([out[6 * s + 5] for s in range(n)])

    def synchronize(self):
        _lib.call("nx_batch_synchronize", self._alive())

    def set_timing(self, on=True):
        _lib.call("nx_batch_set_timing", self._alive(), 1 if on else 0)

    def collect_timing(self):
        ms = (C.c_double * len(_lib.TIME_TAGS))()
        n = (C.c_int * len(_lib.TIME_TAGS))()
        _lib.call("nx_batch_collect_timing", self._alive(), ms, n)
        return {tag: (ms[i], n[i]) for i, tag in enumerate(_lib.TIME_TAGS)}

    def destroy(self):
        if self._h is not None:
            _lib.call("nx_batch_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def random_rect_obstacles(rng, w, h, coverage):
    """Axis-aligned rectangles with sides U(0.05, 0.25) of the grid, added
    until the solid fraction reaches `coverage` (the batched scene spec)."""
    solid = np.zeros((h, w), bool)
    for _ in range(10000):
        if solid.mean() >= coverage:
            break
        rw = max(1, int(rng.uniform(0.05, 0.25) * w))
        rh = max(1, int(rng.uniform(0.05, 0.25) * h))
        x0 = int(rng.integers(0, max(1, w - rw)))
        y0 = int(rng.integers(0, max(1, h - rh)))
        solid[y0:y0 + rh, x0:x0 + rw] = True
    return solid


BENCHMARK_SIDES = (128, 256, 384, 512, 768, 1024)


def benchmark_scenes(n_scenes=64, iters=60):
    """The batched benchmark configuration: 64 scenes (seeds 3000..3063),
    side from {128..1024}, 1-5 Gaussian splats (radius 0.02*N, |F| ~
    U(50, 200) times dt, dye 1.0) applied every step, rectangle obstacles
    covering U(0, 0.15) of the cells.  Returns (batch, sides, solids,
    splats)."""
    sides, solids, splats = [], [], []
    for i in range(n_scenes):
        rng = np.random.default_rng(3000 + i)
        n = int(rng.choice(BENCHMARK_SIDES))
        sides.append(n)
        solids.append(random_rect_obstacles(rng, n, n, rng.uniform(0, 0.15)))
        for _ in range(int(rng.integers(1, 6))):
            f = rng.uniform(50, 200) / 60.0
            ang = rng.uniform(0, 2 * np.pi)
            splats.append((i, rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9),
                           f * np.cos(ang), f * np.sin(ang), 0.02 * n, 1.0))
    b = BatchedScenes(sides, dye=True)
    b.set_params(iterations=iters, vorticity=0.3, viscosity=0.0, bc_mode="free_slip")
    for i, s in enumerate(solids):
        b.set_obstacle_map(i, s)
    b.set_step_splats(splats)
    return b, sides, solids, splats
