"""CPU oracle of the reduced-precision (bfloat16 / float16 storage) step.

TEST INFRASTRUCTURE ONLY (see oracle/fluid.py).

The build stores velocity, pressure, divergence and dye in bfloat16 or
float16 and evaluates every stencil in float32; a value is rounded to the
storage format (nearest even) exactly when a kernel stores it.  This module
replays that dataflow with the float32 stage functions of oracle.fluid and a
rounding step at each store:

  splats -> round | walls+advection -> round | curl, confinement,
  [viscosity], divergence in float32 -> round velocity and b, p1 =
  round((0 - b32) * 0.25) | Jacobi launches of k sweeps in float32, p
  rounded after each launch | float32 residual correction: r = b - A p,
  e = Jacobi(0, r, kc), p32 = p + e | projection with p32 -> round | dye
  advection -> round.

There is no reference counterpart (the reference is float32 only); the
rounding functions are pinned against torch's and numpy's own conversions
and the dataflow, with identity rounding, against oracle.fluid.step.
"""
from __future__ import annotations

import numpy as np

from . import fluid as F


def round_bf16(x):
    """float32 -> bfloat16 (nearest even) -> float32, by bit arithmetic."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).reshape(np.shape(x))


def round_f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def rounder(storage):
    """Rounding at a store: "f32" is the float32 step, "bf16" / "f16" the
    reduced storage, and "f64" the identity in float64 (with float64 inputs
    and dt, step_mixed is then the FP64 twin of the same dataflow, the
    reference the reduced-storage error is measured against)."""
    return {"f32": lambda a: np.asarray(a, np.float32),
            "f64": lambda a: np.asarray(a, np.float64),
            "bf16": round_bf16, "f16": round_f16}[storage]


def jacobi_schedule(sweeps, k):
    out = []
    while sweeps > 0:
        out.append(min(k, sweeps))
        sweeps -= out[-1]
    return out


def correction(p, b, obs, sweeps):
    """p32 = p + e with r = b - ((((x1+x2)+y1)+y2) - 4p) and e = Jacobi
    sweeps of A e = r from 0 (float32)."""
    solid = F.solid_of(obs)
    x1, x2, y1, y2 = F._replaced_neighbours(p, solid)
    r = b - ((((x1 + x2) + y1) + y2) - np.float32(4) * p)
    e = F.jacobi(np.zeros_like(p), r, obs, sweeps)
    return p + e


def step_mixed(u, obs, dt, prm, storage, k, corr_sweeps=8, dye=None,
               persistent=None, splats=()):
    """One reduced-storage step.  Inputs are the widened stored values;
    returns (u, p_stored, p32, b_stored, dye)."""
    R = rounder(storage)
    d = np.float32
    o = obs.copy()
    if persistent is not None:
        o[persistent, 0] = 1
    if splats:
        u, dye = F.gaussian_splat(u, dye, splats)
        u = R(u)
        dye = None if dye is None else R(dye)
    v = F.enforce_boundaries(u, prm.bc_mode, prm.has_borders, prm.inflow)
    v = R(F.advect_velocity(v, o, dt, prm.speed, prm.dissipation))
    w = F.calc_vorticity(v)
    v = F.apply_vorticity(v, w, dt, prm.vorticity)
    if prm.viscosity > 0.0:
        a, rb = F.viscosity_coefficients(prm.viscosity)
        v = F.viscosity(v, a, rb)
    b32 = F.divergence(v, o)
    v = R(v)
    b = R(b32)
    p = R((d(0) - b32) * d(0.25))
    for kk in jacobi_schedule(prm.iterations - 1, k):
        p = R(F.jacobi(p, b, o, kk))
    p32 = correction(p, b, o, corr_sweeps)
    v = R(F.subtract_gradient(v, p32, o))
    if dye is not None:
        o_after = np.zeros_like(obs)
        if persistent is not None:
            o_after[persistent, 0] = 1
        dye = R(F.advect_dye(dye, v, o_after, dt, prm.dye_speed, prm.dye_dissipation))
    return v, p, p32, b, dye
