"""FP64 oracle solve on the GPU for a simulator's current Poisson system.

``solve(sim)`` runs FP64 conjugate gradients (csrc/nx_cg.hip) on the
simulator's current right-hand side and obstacle mask and returns the
mean-free FP64 pressure as a device array; ``pressure_error(sim, x)`` is the
relative L2 error of the simulator's current float32 pressure against it
(both mean-free over fluid cells, the operator's null space removed).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def solve(sim, tol=1e-10, max_iter=200000):
    W, H = sim.width, sim.height
    x = torch.empty(H * W, dtype=torch.float64, device="cuda")
    b = sim.field_buffer("divergence")
    m = sim.field_buffer("mask")
    out = (C.c_double * 2)()
    sim.synchronize()
    torch.cuda.synchronize()
    _lib.call("nx_fp64_solve", C.c_void_p(b.ptr), C.c_void_p(m.ptr), W, H,
              C.c_void_p(x.data_ptr()), float(tol), int(max_iter), out,
              C.c_void_p(sim.stream_handle()))
    return x.view(H, W), {"residual": out[0], "iterations": int(out[1])}


def pressure_error(sim, x, field="pressure"):
    W, H = sim.width, sim.height
    p = sim.field_buffer(field)
    m = sim.field_buffer("mask")
    rel = C.c_double()
    _lib.call("nx_fp64_compare", C.c_void_p(p.ptr), C.c_void_p(x.data_ptr()),
              C.c_void_p(m.ptr), W, H, C.byref(rel), C.c_void_p(sim.stream_handle()))
    return rel.value
