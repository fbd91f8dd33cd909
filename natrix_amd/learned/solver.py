"""Pressure solve with the learned guess; plain Jacobi finishes, is the
fallback, and judges the result.

The network runs on a coarse grid (`coarse` x `coarse`, the resolution it
was trained at) whose cells are f x f blocks of the fine grid, so its
receptive field spans the whole domain and it supplies the low-frequency
part of the pressure, which is what plain Jacobi converges slowly on
(its slowest mode decays like 1 - pi^2 / (4 N^2) per sweep).

One learned correction: restrict the current residual to the coarse grid
(f^2 x block mean, so the coarse problem is in coarse grid units), scale it
to unit RMS, run the U-Net, scale back, remove the fluid mean and add the
bilinear interpolation of the result to the pressure.  The first correction
acts on the first sweep from zero, i.e. it is the initial guess; the
following ones act on the residual left after `smooth` Jacobi sweeps (the
same network applied to the correction equation A e = r).  A correction is
kept only if, after its smoothing sweeps, the residual is below the
residual before it; otherwise the pressure is restored and the network is
not used again in that step.  Plain Jacobi then runs until the relative
residual is <= tol: the step's result always satisfies the same criterion
as the plain path.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _lib
from .unet import pack_input


class LearnedGuess:
    def __init__(self, net, coarse=128):
        self.net = net.eval()
        self.coarse = coarse

    def _factor(self, sim):
        W, H = sim.width, sim.height
        if W != H or W % self.coarse:
            raise ValueError("the learned guess needs a square grid divisible by "
                             f"{self.coarse}")
        return W // self.coarse

    @torch.no_grad()
    def correct(self, sim):
        """Add one learned correction to the simulator's pressure (on the
        simulator's stream)."""
        f = self._factor(sim)
        c = self.coarse
        rc = torch.empty(c * c, dtype=torch.float32, device="cuda")
        mc = torch.empty(c * c, dtype=torch.uint8, device="cuda")
        _lib.call("nx_restrict_residual", sim._alive(), f, C.c_void_p(rc.data_ptr()),
                  C.c_void_p(mc.data_ptr()))
        fluid = mc == 0
        n = fluid.sum().clamp_min(1)
        s = torch.sqrt((rc * rc * fluid).sum() / n).clamp_min(1e-30)
        x = pack_input((rc / s).view(1, c, c), (~fluid).view(1, c, c))
        e = self.net(x).view(-1)
        e = (e - (e * fluid).sum() / n) * s
        e = e * fluid
        _lib.call("nx_prolong_add", sim._alive(), f, C.c_void_p(e.data_ptr()), 1.0)
        return e


def _pressure_view(sim):
    buf = sim.field_buffer("pressure")
    return torch.as_tensor(buf, device="cuda")


def _solve(sim, tol, max_sweeps):
    st = _lib.Stats()
    _lib.call("nx_solve", sim._alive(), float(tol), int(max_sweeps), C.byref(st))
    return {"sweeps": st.sweeps, "converged": bool(st.converged), "residual": st.residual}


def _inner(sim, guess, target, smooth, max_cycles, chunk, max_sweeps, stall=0.9):
    """Solve the correction equation A e = r (current right-hand side, e from
    its first sweep) to relative residual `target`: learned corrections with
    smoothing first (each kept only if it lowers the residual), then plain
    Jacobi in chunks until the target, the sweep cap, or a stall (a chunk
    that reduces the residual by less than 1 - `stall`)."""
    kept, rejected = 0, False
    before, _ = sim.residual()
    start = sim.stats()["sweeps"]
    for _ in range(max_cycles if guess is not None else 0):
        backup = _pressure_view(sim).clone()
        guess.correct(sim)
        st = _solve(sim, target, smooth)
        if not (st["residual"] < before):
            _pressure_view(sim).copy_(backup)
            rejected = True
            break
        kept += 1
        before = st["residual"]
        if st["converged"]:
            return kept, rejected
    while True:
        st = _solve(sim, target, chunk)
        if st["converged"] or st["sweeps"] - start >= max_sweeps:
            break
        if st["residual"] > stall * before:
            break
        before = st["residual"]
    return kept, rejected


def solve_step(sim, guess, dt, tol=1e-4, inner_tol=0.03, smooth=512, max_cycles=4,
               max_outer=12, chunk=65536, max_inner_sweeps=4_000_000):
    """One fluid step whose pressure is solved to FP64 relative residual
    <= tol by iterative refinement (pressure held as a float32 pair, outer
    residual in FP64, float32 correction solves); with `guess` (a
    LearnedGuess) each correction solve starts with learned corrections,
    without it the correction solves are plain Jacobi from zero.  Returns
    the step's statistics."""
    sim._push_params()
    _lib.call("nx_step_begin", sim._alive(), float(dt))
    _lib.call("nx_refine_begin", sim._alive())
    stream = torch.cuda.ExternalStream(sim.stream_handle())
    rel = C.c_double()
    kept = 0
    fallback = False
    outer = 0
    converged = False
    with torch.cuda.stream(stream):
        for outer in range(max_outer + 1):
            _lib.call("nx_refine_residual", sim._alive(), C.byref(rel))
            if rel.value <= tol:
                converged = True
                break
            if outer == max_outer:
                break
            target = max(inner_tol, 0.5 * tol / rel.value)
            k, rej = _inner(sim, guess, target, smooth, max_cycles, chunk,
                            max_inner_sweeps)
            kept += k
            fallback |= rej
            _lib.call("nx_refine_accumulate", sim._alive())
    sweeps = sim.stats()["sweeps"]
    _lib.call("nx_step_finish", sim._alive())
    return {"sweeps": sweeps, "converged": converged, "residual": rel.value,
            "outer": outer, "learned_corrections": kept, "fallback": fallback}
