"""``FluidSimulator``: the reference's Python API over the gfx950 HIP library.

Mirrors ``natrix.core.fluid_simulator.FluidSimulator``
(natrix/core/fluid_simulator.py:15-515): same constructor, properties,
defaults, validation messages, forcing/obstacle calls, ``update`` and
``destroy``.  Where the reference sets bgfx uniforms and dispatches compute
shaders, this class calls the C-ABI of ``libnatrix_hip.so`` through ctypes;
every operation runs as a HIP kernel on the simulator's stream.

Additions the reference lacks are keyword-only or separate methods: the dye
field, Gaussian splats, a persistent obstacle map, wall modes (free-slip,
inflow), the tolerance-driven pressure solve, field download/upload and
per-stage entry points.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .. import _lib
from .._lib import (BC_FREE_SLIP, BC_INFLOW, BC_REFERENCE, FIELD_CURL,
                    FIELD_DIVERGENCE, FIELD_DYE, FIELD_MASK, FIELD_PRESSURE,
                    FIELD_PRESSURE32, FIELD_VELOCITY, Params, Stats)

_FIELD_NAMES = {
    "velocity": FIELD_VELOCITY,
    "pressure": FIELD_PRESSURE,
    "divergence": FIELD_DIVERGENCE,
    "curl": FIELD_CURL,
    "vorticity": FIELD_CURL,
    "mask": FIELD_MASK,
    "obstacles": FIELD_MASK,
    "dye": FIELD_DYE,
    "pressure32": FIELD_PRESSURE32,
}

_STORAGE = {"f32": 0, "bf16": 1, "f16": 2}


def bf16_bits_to_f32(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def f32_to_bf16_bits(x):
    """Round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)

_BC_NAMES = {"reference": BC_REFERENCE, "free_slip": BC_FREE_SLIP,
             "inflow": BC_INFLOW}


class DeviceBuffer:
    """A borrowed view of a device field (the reference's
    ``get_velocity_buffer`` returns a bgfx handle; this returns the HIP
    pointer).  Valid until the next call that flips the field.  Exposes
    ``__cuda_array_interface__`` so ``torch.as_tensor(buf, device="cuda")``
    can wrap it without a copy."""

    def __init__(self, sim, field, ptr, nbytes, shape, dtype):
        self._sim = sim
        self.field = field
        self.ptr = ptr
        self.nbytes = nbytes
        self.shape = shape
        self.dtype = np.dtype(dtype)

    @property
    def __cuda_array_interface__(self):
        return {"shape": self.shape, "typestr": self.dtype.str,
                "data": (self.ptr, False), "version": 2, "strides": None}

    def numpy(self):
        return self._sim.download(self.field)


class FluidSimulator:
    VELOCITY_READ = 0
    VELOCITY_WRITE = 1

    PRESSURE_READ = 0
    PRESSURE_WRITE = 1

    _width = 512
    _height = 512

    # fluid_simulator.py:28-35
    _speed = 500.0
    _iterations = 50
    _dissipation = 1.0
    _vorticity = 0.0
    _viscosity = 0.1

    has_borders = True
    simulate = True

    def __init__(self, width: int, height: int, vertex_layout=None, *,
                 dye: bool | tuple = False, bc_mode="reference",
                 inflow: float = 1.0, fused: bool = True, storage: str = "f32"):
        """``vertex_layout`` is accepted for signature compatibility with the
        reference (a bgfx vertex layout) and ignored.  ``dye=True`` adds a
        dye field on the velocity grid, ``dye=(w, h)`` on its own grid.
        ``storage`` is "f32", "bf16" or "f16" for velocity, pressure,
        divergence and dye (arithmetic is float32 either way)."""
        self._width = int(width)
        self._height = int(height)
        self.vertex_layout = vertex_layout
        self._h = None
        if storage not in _STORAGE:
            raise ValueError(f"unknown storage {storage!r}")
        self.storage = storage
        h = C.c_void_p()
        _lib.call("nx_create_ex", self._width, self._height, _STORAGE[storage],
                  C.byref(h))
        self._h = h
        self._prm = Params()
        _lib.call("nx_default_params", C.byref(self._prm))
        self.bc_mode = bc_mode
        self.inflow = inflow
        self.fused = fused
        self.tol = 0.0
        self.max_iters = 1_000_000
        self.check_every = 256
        self.correction_sweeps = 8
        from ..autotune import lookup, lookup_smoother
        self.jacobi_tile = lookup(self._width, self._height, storage)
        if storage == "f32":
            self.smoother_tile = lookup_smoother(self._width, self._height)
        self._dye_speed = 500.0
        self._dye_dissipation = 1.0
        self._dye_in_step = True
        self._guess = None
        self._guess_tol = 1e-4
        self.last_solve = None
        self._use_graph = False
        self._dye_shape = None
        if dye:
            dw, dh = (self._width, self._height) if dye is True else dye
            self.enable_dye(dw, dh)

    # ------------------------------------------------------------ properties
    @property
    def width(self):
        return self._width

    @property
    def height(self):
        return self._height

    @property
    def speed(self):
        return self._speed

    @speed.setter
    def speed(self, value):
        if value > 0:
            self._speed = value
        else:
            raise ValueError("'Speed' should be greater than zero")

    @property
    def iterations(self):
        return self._iterations

    @iterations.setter
    def iterations(self, value):
        if value > 0:
            self._iterations = value
        else:
            raise ValueError("'Iterations' should be grater than zero")

    @property
    def dissipation(self):
        return self._dissipation

    @dissipation.setter
    def dissipation(self, value):
        if value > 0:
            self._dissipation = value
        else:
            raise ValueError("'Dissipation' should be grater than zero")

    @property
    def vorticity(self):
        return self._vorticity

    @vorticity.setter
    def vorticity(self, value):
        if value >= 0:
            self._vorticity = value
        else:
            raise ValueError("'Vorticity' should be grater or equal than zero")

    @property
    def viscosity(self):
        return self._viscosity

    @viscosity.setter
    def viscosity(self, value):
        if value >= 0.0:
            self._viscosity = value
        else:
            raise ValueError("'Viscosity' should be greater or equal than zero")

    @property
    def bc_mode(self):
        return self._bc_mode

    @bc_mode.setter
    def bc_mode(self, value):
        if isinstance(value, str):
            if value not in _BC_NAMES:
                raise ValueError(f"unknown bc_mode {value!r}")
            value = _BC_NAMES[value]
        if value not in (BC_REFERENCE, BC_FREE_SLIP, BC_INFLOW):
            raise ValueError(f"unknown bc_mode {value!r}")
        self._bc_mode = int(value)

    @property
    def use_graph(self):
        return self._use_graph

    @use_graph.setter
    def use_graph(self, on):
        self._use_graph = bool(on)
        _lib.call("nx_set_graph", self._h, int(self._use_graph))

    # ------------------------------------------------------------ internals
    def _push_params(self):
        p = self._prm
        p.speed = self._speed
        p.dissipation = self._dissipation
        p.vorticity = self._vorticity
        p.viscosity = self._viscosity
        p.iterations = int(self._iterations)
        p.has_borders = 1 if self.has_borders else 0
        p.bc_mode = self._bc_mode
        p.inflow = self.inflow
        p.dye_speed = self._dye_speed
        p.dye_dissipation = self._dye_dissipation
        p.dye_in_step = 1 if self._dye_in_step else 0
        p.fused = 1 if self.fused else 0
        tile = tuple(int(v) for v in self.jacobi_tile)
        p.jacobi_k, p.jacobi_rows, p.jacobi_waves, p.jacobi_cols = tile[:4]
        p.jacobi_seg = tile[4] if len(tile) > 4 else 0
        p.tol = float(self.tol)
        p.max_iters = int(self.max_iters)
        p.check_every = int(self.check_every)
        p.correction_sweeps = int(self.correction_sweeps)
        _lib.call("nx_set_params", self._h, C.byref(p))

    def _alive(self):
        if self._h is None:
            raise _lib.NatrixError("simulator has been destroyed")
        return self._h

    # ------------------------------------------------------------ reference API
    def get_velocity_buffer(self):
        """fluid_simulator.py:113-114: the current velocity buffer."""
        return self.field_buffer("velocity")

    def add_velocity(self, position: tuple, velocity: tuple, radius: float):
        """fluid_simulator.py:116-131 / shader.AddVelocity.comp."""
        if self.simulate:
            _lib.call("nx_add_velocity", self._alive(), position[0],
                      position[1], velocity[0], velocity[1], radius)

    def add_circle_obstacle(self, position: tuple, radius: float, static=False):
        """fluid_simulator.py:135-153 (position normalised, radius in cells)."""
        if self.simulate:
            _lib.call("nx_add_circle_obstacle", self._alive(), position[0],
                      position[1], radius, 1 if static else 0)

    def add_triangle_obstacle(self, p1: tuple, p2: tuple, p3: tuple,
                              static=False):
        """fluid_simulator.py:156-172 (points in normalised space)."""
        if self.simulate:
            _lib.call("nx_add_triangle_obstacle", self._alive(), p1[0], p1[1],
                      p2[0], p2[1], p3[0], p3[1], 1 if static else 0)

    def update(self, time_delta: float):
        """fluid_simulator.py:174-280 (+ the dye advection, if enabled).
        With a learned guess set (set_guess), the pressure is solved by the
        learned two-grid cycle to the guess's FP64 relative residual, and
        the step's solver statistics are kept in `last_solve`."""
        if self.simulate:
            if self._guess is not None:
                from ..learned.solver import solve_step
                self.last_solve = solve_step(self, self._guess, time_delta,
                                             tol=self._guess_tol)
                return
            self._push_params()
            _lib.call("nx_update", self._alive(), time_delta)

    def set_guess(self, guess, tol=1e-4):
        """Solve the pressure of every update() with `guess` (a
        natrix_amd.learned.solver.LearnedGuess) to FP64 relative residual
        <= tol (a positive number), by iterative refinement with learned
        cycles and Jacobi; guess=None restores the fixed-sweep (or
        `tol`-driven) Jacobi step and ignores `tol`.  The learned step needs
        float32 storage: a guess on a bf16/f16 simulator is refused here
        rather than by every later update()."""
        if guess is None:
            self._guess = None
            return
        if self.storage != "f32":
            raise ValueError("the learned step needs float32 storage "
                             f"(this simulator stores {self.storage})")
        if tol is None or isinstance(tol, bool) or not (float(tol) > 0 and
                                                        math.isfinite(float(tol))):
            raise ValueError("tol must be a positive finite number")
        self._guess = guess
        self._guess_tol = float(tol)

    def destroy(self):
        """fluid_simulator.py:476-515: release every device resource."""
        if self._h is not None:
            _lib.call("nx_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # ------------------------------------------------------------ extensions
    def enable_dye(self, width=None, height=None):
        """A scalar dye field advected by this simulator's velocity
        (the demo's SmoothParticlesArea, made part of the simulator)."""
        w, h = int(width or self._width), int(height or self._height)
        _lib.call("nx_enable_dye", self._alive(), w, h)
        self._dye_shape = (h, w)

    def add_dye(self, position: tuple, radius: float, strength: float):
        """Cone dye splat clamped to [0, 255] (shader.AddParticle.comp)."""
        if self.simulate:
            _lib.call("nx_add_dye", self._alive(), position[0], position[1],
                      strength, radius)

    def add_splats(self, splats):
        """Gaussian splats: iterable of (px, py, ax, ay, radius, dye) with
        final amplitudes (velocity += a * w, dye += dye * w)."""
        if not self.simulate:
            return
        arr = np.ascontiguousarray(np.asarray(splats, np.float32).reshape(-1, 6))
        _lib.call("nx_add_gaussian_splats", self._alive(),
                  arr.ctypes.data_as(C.POINTER(C.c_float)), arr.shape[0])

    def set_obstacle_map(self, solid):
        """Persistent obstacle map: (H, W) array, non-zero = solid."""
        m = np.ascontiguousarray(np.asarray(solid) != 0, dtype=np.uint8)
        if m.shape != (self._height, self._width):
            raise ValueError("obstacle map must have shape (height, width)")
        _lib.call("nx_set_obstacle_map", self._alive(),
                  m.ctypes.data_as(C.POINTER(C.c_uint8)))

    def _field_shape(self, fid):
        """(shape, stored dtype) of a field."""
        st = {"f32": np.float32, "bf16": np.uint16, "f16": np.float16}[self.storage]
        if fid == FIELD_VELOCITY:
            return (self._height, self._width, 2), st
        if fid == FIELD_MASK:
            return (self._height, self._width), np.uint8
        if fid == FIELD_DYE:
            return self.dye_shape, st
        if fid in (FIELD_CURL, FIELD_PRESSURE32):
            return (self._height, self._width), np.float32
        return (self._height, self._width), st

    def _widen(self, a):
        if a.dtype == np.uint16:
            return bf16_bits_to_f32(a)
        if a.dtype == np.float16:
            return a.astype(np.float32)
        return a

    def _narrow(self, a, dt):
        if dt == np.uint16:
            return f32_to_bf16_bits(a)
        return np.asarray(a, dtype=dt)

    def field_buffer_raw(self, fid):
        ptr = C.c_void_p()
        nbytes = C.c_size_t()
        _lib.call("nx_field_ptr", self._alive(), fid, C.byref(ptr),
                  C.byref(nbytes))
        return ptr.value, nbytes.value, nbytes.value // 4

    def field_buffer(self, name):
        fid = _FIELD_NAMES[name] if isinstance(name, str) else int(name)
        ptr, nbytes, _ = self.field_buffer_raw(fid)
        shape, dt = self._field_shape(fid)
        return DeviceBuffer(self, fid, ptr, nbytes, shape, dt)

    @property
    def dye_shape(self):
        if self._dye_shape is None:
            raise _lib.NatrixError("dye is not enabled")
        return self._dye_shape

    def dye_rgba8(self, device=False):
        """The dye as an RGBA8 image (demo/shaders/demo.ComputeShader.comp:9-19:
        every channel = clamp(dye, 0, 1) * 255 rounded): a (dye_h, dye_w, 4)
        uint8 numpy array, or with `device` a torch tensor on the GPU."""
        import torch
        h, w = self.dye_shape
        out = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        _lib.call("nx_dye_rgba8", self._alive(), C.c_void_p(out.data_ptr()))
        self.synchronize()
        return out if device else out.cpu().numpy()

    def quiver_rgba8(self, window_w, window_h, arrow_tile, device=False):
        """The demo's velocity arrows over a window_w x window_h framebuffer
        (demo/shaders/demo.QuiverFragmentShader.frag, one arrow per
        arrow_tile-pixel tile, row y at gl_FragCoord.y = y + 0.5): a
        (window_h, window_w, 4) uint8 image, white with the arrows in alpha,
        as numpy or with `device` a torch tensor on the GPU."""
        import torch
        out = torch.empty((int(window_h), int(window_w), 4), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        _lib.call("nx_quiver_rgba8", self._alive(), int(window_w), int(window_h),
                  float(arrow_tile), C.c_void_p(out.data_ptr()))
        self.synchronize()
        return out if device else out.cpu().numpy()

    def download(self, name):
        fid = _FIELD_NAMES[name] if isinstance(name, str) else int(name)
        ptr, nbytes, _ = self.field_buffer_raw(fid)
        shape, dt = self._field_shape(fid)
        out = np.empty(shape, dt)
        assert out.nbytes == nbytes
        _lib.call("nx_download", self._alive(), fid, out.ctypes.data_as(C.c_void_p),
                  out.nbytes)
        return out if fid == FIELD_MASK else self._widen(out)

    def upload(self, name, array):
        fid = _FIELD_NAMES[name] if isinstance(name, str) else int(name)
        ptr, nbytes, _ = self.field_buffer_raw(fid)
        shape, dt = self._field_shape(fid)
        a = np.ascontiguousarray(self._narrow(array, dt) if fid != FIELD_MASK
                                 else np.asarray(array, dtype=dt))
        if a.shape != tuple(shape):
            raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        _lib.call("nx_upload", self._alive(), fid, a.ctypes.data_as(C.c_void_p),
                  a.nbytes)

    def residual(self):
        """FP64 relative residual of the current pressure (mean-free over
        fluid cells).  Returns (rel, sums) with sums = (sum r, sum r^2,
        sum b, sum b^2, fluid cells)."""
        out = (C.c_double * 6)()
        _lib.call("nx_residual", self._alive(), out)
        return out[5], tuple(out[:5])

    def stats(self):
        st = Stats()
        _lib.call("nx_get_stats", self._alive(), C.byref(st))
        return {"sweeps": st.sweeps, "converged": bool(st.converged),
                "residual": st.residual, "residual_sweeps": st.residual_sweeps}

    def synchronize(self):
        _lib.call("nx_synchronize", self._alive())

    def stage(self, name, dt=0.0, sweeps=1):
        """Run one reference shader on the current buffers (parity tests)."""
        self._push_params()
        h = self._alive()
        if name == "boundaries":
            _lib.call("nx_stage_boundaries", h)
        elif name == "advect":
            _lib.call("nx_stage_advect", h, dt)
        elif name == "curl":
            _lib.call("nx_stage_curl", h)
        elif name == "confine":
            _lib.call("nx_stage_confine", h, dt)
        elif name == "viscosity":
            _lib.call("nx_stage_viscosity", h)
        elif name == "divergence":
            _lib.call("nx_stage_divergence", h)
        elif name == "clear_pressure":
            _lib.call("nx_stage_clear_pressure", h)
        elif name == "jacobi":
            _lib.call("nx_stage_jacobi", h, int(sweeps))
        elif name == "subtract_gradient":
            _lib.call("nx_stage_subtract_gradient", h)
        elif name == "clear_obstacles":
            _lib.call("nx_stage_clear_obstacles", h)
        elif name == "advect_dye":
            _lib.call("nx_stage_advect_dye", h, dt)
        elif name == "fused_vort_div":
            _lib.call("nx_stage_fused_vort_div", h, dt)
        else:
            raise ValueError(f"unknown stage {name!r}")

    @property
    def smoother_tile(self):
        """Launch shape of the fine Chebyshev smoother of the learned step:
        (passes per launch, rows, waves, cols) of the tile kernel, or with a
        fifth value seg > 0 the streaming kernel (nx_set_smoother_tile)."""
        t = (C.c_int * 5)()
        _lib.call("nx_get_smoother_tile", self._alive(), t)
        return tuple(t) if t[4] > 0 else tuple(t)[:4]

    @smoother_tile.setter
    def smoother_tile(self, tile):
        tile = tuple(int(v) for v in tile)
        _lib.call("nx_set_smoother_tile", self._alive(), (C.c_int * 5)(*tile, *([0] * (5 - len(tile)))),
                  len(tile))

    def jacobi_blocked(self, sweeps, k, rows, waves, cols=1, seg=0):
        """``sweeps`` Jacobi sweeps with the temporal-blocked kernel, k per
        launch, on a (64*cols) x (rows*waves) register tile (seg = 0), or
        with the streaming kernel whose workgroups each walk `seg` output
        rows of a (64*cols)-column strip in chunks of rows*waves rows."""
        _lib.call("nx_jacobi_blocked", self._alive(), int(sweeps), int(k),
                  int(rows), int(waves), int(cols), int(seg))

    # ------------------------------------------------------------ timing
    def set_timing(self, on=True):
        _lib.call("nx_set_timing", self._alive(), 1 if on else 0)

    def collect_timing(self):
        """Per kernel family: (total ms, launches) since the last collect,
        measured with HIP events on the simulator's stream."""
        ms = (C.c_double * len(_lib.TIME_TAGS))()
        n = (C.c_int * len(_lib.TIME_TAGS))()
        _lib.call("nx_collect_timing", self._alive(), ms, n)
        return {tag: (ms[i], n[i]) for i, tag in enumerate(_lib.TIME_TAGS)}

    def stream_handle(self):
        s = C.c_void_p()
        _lib.call("nx_get_stream", self._alive(), C.byref(s))
        return s.value
