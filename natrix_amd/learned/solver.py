"""Pressure solve with the learned guess; plain Jacobi finishes, is the
fallback, and judges the result.

The network runs on a coarse grid (`coarse` x `coarse`, the resolution it
was trained at) whose cells are f x f blocks of the fine grid, so its
receptive field spans the whole domain and it supplies the low-frequency
part of the pressure, which is what plain Jacobi converges slowly on (its
slowest mode decays like 1 - pi^2 / (4 N^2) per sweep).

One learned cycle:
  1. restrict the current residual to the coarse grid (the block sum over
     fluid cells, so the coarse problem is in coarse grid units, with the
     open fraction of every coarse face as the coarse operator's weight)
     and scale it to unit RMS;
  2. run the U-Net and scale back: the coarse correction's first guess;
  3. Chebyshev-accelerated Jacobi on the coarse problem from that guess
     (tiled over 88 workgroups, 12 iterations per launch): removes the
     network's error at short coarse wavelengths, where its relative error
     is largest and which the fine sweeps would need tens of thousands of
     sweeps to undo;
  4. interpolate to the fine grid and add it scaled by the energy line
     search (the step along the correction that minimises the solve's
     energy functional: it never increases the error in the energy norm,
     and a useless, NaN or infinite correction gets step 0);
  5. `smooth` fine smoothing iterations: by default Chebyshev on
     [1.536 / f^2, 8] (nx_smooth_chebyshev, temporal-blocked k_cheb_tb,
     whose first launch also takes step 4's line-search step as it loads
     the pressure); with fine_lmin = 0, Jacobi sweeps ending in one damped
     sweep (omega = 1/2, nx_smooth).
Cycles repeat while they reduce the residual by a useful factor; then plain
Jacobi chunks run until the relative residual is <= the target, so the
step's result satisfies the same criterion as the plain path.  With the
native runner the cycles of a correction solve, decisions included, are one
C-ABI call (nx_learned_cycles), whose coarse-grid part (steps 2-3) replays
as a HIP graph; native_loop=False runs the same cycles from Python, one
nx_learned_correct and one smoothing call per cycle.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from .. import _lib
from .unet import UNet, pack_input


class NativeUNet:
    """The trained U-Net as a native runner (nx_unet_create): bfloat16 copies
    of the 14 convolution weights, float32 biases and head, kept alive here;
    the runner owns its activation buffers for one H x W coarse grid."""

    def __init__(self, net, H, W):
        convs = [m for blk in (net.e0, net.e1, net.e2, net.e3, net.d2, net.d1, net.d0)
                 for m in blk]
        self._w = [c.w.detach().to(torch.bfloat16).contiguous() for c in convs]
        self._b = [c.b.detach().float().contiguous() for c in convs]
        self._hw = net.hw.detach().float().contiguous()
        self._hb = net.hb.detach().float().contiguous()
        base = convs[0].w.shape[0]
        # nx_unet_create copies the weights on the null stream: they must be
        # written before it runs, whatever stream produced them
        torch.cuda.current_stream().synchronize()
        h = C.c_void_p()
        wp = (C.c_void_p * 14)(*[t.data_ptr() for t in self._w])
        bp = (C.c_void_p * 14)(*[t.data_ptr() for t in self._b])
        _lib.call("nx_unet_create", H, W, base, wp, bp, C.c_void_p(self._hw.data_ptr()),
                  C.c_void_p(self._hb.data_ptr()), C.byref(h))
        self.handle = h
        self.H, self.W = H, W

    def forward(self, x, out):
        """out[H*W] float32 = network(x[H, W, 8] bfloat16) on the current stream."""
        _lib.call("nx_unet_forward", self.handle, C.c_void_p(x.data_ptr()),
                  C.c_void_p(out.data_ptr()),
                  C.c_void_p(torch.cuda.current_stream().cuda_stream))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.call("nx_unet_destroy", h)
            except Exception:
                pass
            self.handle = None


def chebyshev_lmin(iterations, damping):
    """Lower end of the Chebyshev interval [lmin, 8] on which `iterations`
    iterations reduce every mode by `damping`: 1 / T_k(s) = damping with
    s = (8 + lmin) / (8 - lmin)."""
    s = math.cosh(math.acosh(1.0 / damping) / iterations)
    return 8.0 * (s - 1.0) / (s + 1.0)


class LearnedGuess:
    """`net` maps (coarse right-hand side, coarse mask) to the coarse
    solution; None gives the same cycle without a network (coarse smoothing
    from zero), for comparison.  `coarse_iters` Chebyshev iterations with
    `damping` on the modes above their interval's lower end (0 iterations:
    the network output as is).  A `UNet` runs natively (one C-ABI call per
    correction, nx_learned_correct) unless `native` is False; other modules
    run through PyTorch."""

    def __init__(self, net, coarse=128, coarse_iters=100, damping=1e-3, native=True,
                 use_network=True, mid=0, mid_iters=12):
        """use_network=False keeps the native runner but starts every coarse
        smoothing from zero (nx_unet_set_bypass): the same cycle, launches
        and timing path without the network's guess.  mid >= 2 (native
        runner only): a three-level cycle through a middle grid mid times
        coarser than the fine one, smoothed with mid_iters Chebyshev
        iterations of its face-weighted operator on [1.536 / (f / mid)^2,
        8] (nx_unet_set_levels); the fine interval's lower end is then
        1.536 / mid^2."""
        self.mid, self.mid_iters = int(mid), int(mid_iters)
        self.net = net.eval() if net is not None else None
        self.use_network = bool(use_network)
        if coarse < 8 or coarse % 8:
            raise ValueError(f"coarse grid side {coarse} is not a positive multiple of 8")
        self.coarse = coarse
        self.coarse_iters = coarse_iters
        self.lmin = chebyshev_lmin(coarse_iters, damping) if coarse_iters else 0.0
        self._native = native and isinstance(net, UNet)
        self._runners = {}

    def coarse_grid(self, W, H):
        """(f, Wc, Hc): f = ceil(max(W, H) / coarse), the smallest factor
        whose cells ceil(W / f) x ceil(H / f) fit in `coarse` x `coarse`;
        the network's grid Wc x Hc rounds both sides up to multiples of 8
        (the network pools three times).  Coarse cells beyond the fine
        grid are solid with a zero right-hand side; only the native runner
        takes such a padded grid."""
        c = self.coarse
        f = max(1, -(-max(W, H) // c))
        return f, -(-(-(-W // f)) // 8) * 8, -(-(-(-H // f)) // 8) * 8

    def fine_lmin(self, W, H):
        """Lower end of the fine Chebyshev smoother's interval, 1.536 / f^2
        (0.006 at f = 16): the modes of -L below it (wavelengths beyond
        about 4 f cells) are left to the coarse correction."""
        f = self.mid if self.mid > 1 else self.coarse_grid(W, H)[0]
        return 1.536 / (f * f)

    def _runner(self, wc, hc):
        if not self._native:
            return None
        r = self._runners.get((wc, hc))
        if r is None:
            r = self._runners[(wc, hc)] = NativeUNet(self.net, hc, wc)
            if not self.use_network:
                _lib.call("nx_unet_set_bypass", r.handle, 1)
        return r

    @torch.no_grad()
    def native(self, sim):
        """(f, runner) when corrections of `sim` run natively, else None."""
        f, wc, hc = self.coarse_grid(sim.width, sim.height)
        runner = self._runner(wc, hc)
        if runner is not None:
            self._levels(runner, f)
        return None if runner is None else (f, runner)

    def _levels(self, runner, f):
        """The runner's cycle shape for factor f: three levels with the
        middle interval's lower end 1.536 / (f / mid)^2, or two."""
        if self.mid > 1:
            if f % self.mid or f // self.mid < 2:
                raise ValueError(f"factor {f} is not a multiple >= 2 x of the middle factor "
                                 f"{self.mid}")
            f2 = f // self.mid
            _lib.call("nx_unet_set_levels", runner.handle, self.mid, self.mid_iters,
                      1.536 / (f2 * f2))
        else:
            _lib.call("nx_unet_set_levels", runner.handle, 0, 0, 0.0)

    def correct(self, sim, report=False, line_search=True):
        """Add one learned correction to the simulator's pressure (on the
        simulator's stream), scaled by the energy line search (or by 1
        without `line_search`).  With `report` returns (a, d.r, d.Ld)
        (synchronises); otherwise None."""
        f, wc, hc = self.coarse_grid(sim.width, sim.height)
        runner = self._runner(wc, hc)
        if runner is not None:
            self._levels(runner, f)
        elif self.mid > 1:
            raise ValueError("the three-level cycle runs only in the native runner")
        if runner is not None and line_search:
            # the native runner enqueues everything on the simulator's stream
            out = (C.c_double * 3)() if report else None
            _lib.call("nx_learned_correct", sim._alive(), runner.handle, f,
                      int(self.coarse_iters), float(self.lmin), out)
            return tuple(out) if report else None
        with torch.cuda.stream(torch.cuda.ExternalStream(sim.stream_handle())):
            return self._correct(sim, f, wc, hc, report, line_search)

    def _correct(self, sim, f, wc, hc, report, line_search):
        if -(-sim.width // f) != wc or -(-sim.height // f) != hc:
            raise ValueError(f"a {sim.width} x {sim.height} grid needs a padded coarse grid "
                             f"({wc} x {hc} for f = {f}), which only the native runner takes")
        rc = torch.empty(hc * wc, dtype=torch.float32, device="cuda")
        mc = torch.empty(hc * wc, dtype=torch.uint8, device="cuda")
        # Chebyshev coarse smoothing takes the face-weighted operator and its
        # block-sum restriction, as the native runner does
        faces = (torch.empty(2 * hc * wc, dtype=torch.float32, device="cuda")
                 if self.lmin > 0 else None)
        if faces is not None:
            _lib.call("nx_restrict_residual_faces", sim._alive(), f, C.c_void_p(rc.data_ptr()),
                      C.c_void_p(mc.data_ptr()), C.c_void_p(faces.data_ptr()))
        else:
            _lib.call("nx_restrict_residual", sim._alive(), f, C.c_void_p(rc.data_ptr()),
                      C.c_void_p(mc.data_ptr()))
        fluid = mc == 0
        n = fluid.sum().clamp_min(1)
        s = torch.sqrt((rc * rc * fluid).sum() / n).clamp_min(1e-30)
        if self.net is not None and self.use_network:
            x = pack_input((rc / s).view(1, hc, wc), (~fluid).view(1, hc, wc))
            e = (self.net(x).view(-1) * s).contiguous()
        else:
            e = torch.zeros_like(rc)
        if faces is not None:
            _lib.call("nx_coarse_smooth_faces", sim._alive(), wc, hc, C.c_void_p(e.data_ptr()),
                      C.c_void_p(rc.data_ptr()), C.c_void_p(mc.data_ptr()),
                      C.c_void_p(faces.data_ptr()), int(self.coarse_iters), float(self.lmin))
        else:
            _lib.call("nx_coarse_smooth", sim._alive(), wc, hc, C.c_void_p(e.data_ptr()),
                      C.c_void_p(rc.data_ptr()), C.c_void_p(mc.data_ptr()),
                      int(self.coarse_iters), float(self.lmin))
        e = (e - (e * fluid).sum() / n) * fluid
        if not line_search:
            _lib.call("nx_prolong_add", sim._alive(), f, C.c_void_p(e.data_ptr()),
                      C.c_void_p(mc.data_ptr()), 1.0)
            return None
        out = (C.c_double * 3)() if report else None
        _lib.call("nx_prolong_line_search", sim._alive(), f, C.c_void_p(e.data_ptr()),
                  C.c_void_p(mc.data_ptr()), out)
        return tuple(out) if report else None


def _solve(sim, tol, max_sweeps):
    st = _lib.Stats()
    _lib.call("nx_solve", sim._alive(), float(tol), int(max_sweeps), C.byref(st))
    return {"sweeps": st.sweeps, "converged": bool(st.converged), "residual": st.residual}


def _smooth(sim, target, sweeps, early_exit=False, lmin=0.0, tile=None):
    """`sweeps` fine smoothing iterations of the current system and the
    residual, in one synchronisation: Chebyshev iterations on [lmin, 8]
    (lmin > 0), or Jacobi sweeps ending in a damped one (omega = 1/2, which
    removes the near-checkerboard modes the plain sweep keeps); with
    `early_exit` plain Jacobi sweeps run as a tolerance solve that stops at
    relative residual <= target.  Also returns the last line search's
    step."""
    out = (C.c_double * 2)()
    if lmin > 0 and not early_exit:
        t = None if tile is None else (C.c_int * 4)(*(int(v) for v in tile))
        _lib.call("nx_smooth_chebyshev", sim._alive(), int(sweeps), float(lmin), t, out)
        return {"sweeps": sim.stats()["sweeps"], "converged": out[0] <= target,
                "residual": out[0], "alpha": out[1]}
    n = max(0, sweeps - 1)
    if early_exit:
        st = _solve(sim, target, max(1, n))
        if st["converged"]:
            st["alpha"] = None
            return st
        n = 0
    _lib.call("nx_smooth", sim._alive(), n, 0.5, out)
    return {"sweeps": sim.stats()["sweeps"], "converged": out[0] <= target,
            "residual": out[0], "alpha": out[1]}


def _native_cycles(sim, guess, nat, target, smooth, max_cycles, stall, trace, lmin, tile):
    """The learned cycles in one nx_learned_cycles call (the same GPU work
    and decisions as the loop in _inner).  Returns (kept, converged,
    abandoned, residual after the last cycle or None)."""
    f, runner = nat
    cap = 3 * max(1, max_cycles)
    rec = (C.c_double * cap)()
    res = (C.c_int * 4)()
    t = None if tile is None else (C.c_int * 4)(*(int(v) for v in tile))
    _lib.call("nx_learned_cycles", sim._alive(), runner.handle, f, int(guess.coarse_iters),
              float(guess.lmin), int(smooth), float(lmin), t, float(target), float(stall),
              int(max_cycles), rec, cap, res)
    n, kept, conv, abandoned = res[0], res[1], bool(res[2]), bool(res[3])
    if trace is not None:
        for c in range(n):
            trace.append(("cycle", round(rec[3 * c], 4), int(rec[3 * c + 1]), rec[3 * c + 2]))
    return kept, conv, abandoned, (rec[3 * (n - 1) + 2] if n else None)


def _inner(sim, guess, target, smooth, max_cycles, chunk, max_sweeps, stall=0.5,
           plain_stall=0.9, trace=None, lmin=0.0, tile=None, native_loop=True):
    """Solve the correction equation (current right-hand side, iterate
    from its first sweep) to relative residual `target`: learned cycles
    while each reduces the residual below `stall` x the previous one, then
    Jacobi chunks until the target, the sweep cap, or (after learned
    cycles) a chunk that reduces it by less than 1 - `plain_stall`.
    Returns (cycles whose line search took a nonzero step, True if the
    learned cycles were abandoned before the target)."""
    kept, abandoned = 0, False
    before = None
    start = sim.stats()["sweeps"]
    nat = guess.native(sim) if (guess is not None and native_loop) else None
    if nat is not None:
        kept, conv, abandoned, before = _native_cycles(sim, guess, nat, target, smooth,
                                                       max_cycles, stall, trace, lmin, tile)
        if conv:
            return kept, False
    for _ in range(max_cycles if (guess is not None and nat is None) else 0):
        guess.correct(sim)
        st = _smooth(sim, target, smooth, lmin=lmin, tile=tile)
        a = st["alpha"]
        if trace is not None:
            trace.append(("cycle", round(a, 4), st["sweeps"], st["residual"]))
        if st["converged"]:
            return kept + (a != 0.0), False
        if a == 0.0 or (before is not None and st["residual"] > stall * before):
            abandoned = True
            before = st["residual"]
            break
        kept += 1
        before = st["residual"]
    while True:
        st = (_smooth(sim, target, chunk, early_exit=True) if guess is not None
              else _solve(sim, target, chunk))
        if trace is not None:
            trace.append(("plain", st["sweeps"], st["residual"]))
        if st["converged"] or st["sweeps"] - start >= max_sweeps:
            break
        # the learned path's fallback chunks stop at a stall (the float32
        # floor of a correction solve); plain correction solves run to the
        # target or the sweep cap, as the reference's Jacobi would
        if guess is not None and before is not None and st["residual"] > plain_stall * before:
            break
        before = st["residual"]
    return kept, abandoned


def solve_step(sim, guess, dt, tol=1e-4, inner_tol=0.1, smooth=72, max_cycles=6,
               max_outer=12, chunk=65536, max_inner_sweeps=4_000_000, trace=None,
               fine_lmin=None, fine_tile=None, native_loop=True):
    """One fluid step whose pressure is solved to FP64 relative residual
    <= tol by iterative refinement (pressure held as a float32 pair, outer
    residual in FP64, float32 correction solves); with `guess` (a
    LearnedGuess) each correction solve starts with learned cycles, without
    it the correction solves are plain Jacobi from zero.  A learned cycle
    is the correction followed by `smooth` fine smoothing iterations:
    Chebyshev on [fine_lmin, 8] (None: the guess's 1.536 / f^2; `fine_tile`
    the kernel's tile, None for the library's default) or, with
    fine_lmin = 0, Jacobi sweeps ending in a damped one; with a native
    runner the cycles of a correction solve run in one nx_learned_cycles
    call (native_loop=False: the same cycles from Python).  Returns the
    step's statistics: sweeps (fine sweeps and Chebyshev iterations), converged, residual
    (FP64 relative), outer (refinement iterations), learned_corrections
    (cycles that moved the pressure), fallback (learned cycles abandoned
    for plain Jacobi in some correction solve).  `trace` (a list) receives
    one record per outer iteration, learned cycle and plain chunk."""
    if fine_lmin is None:
        lmin = guess.fine_lmin(sim.width, sim.height) if guess is not None else 0.0
    else:
        lmin = float(fine_lmin)
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
            if trace is not None:
                trace.append(("outer", outer, rel.value))
            if rel.value <= tol:
                converged = True
                break
            if outer == max_outer:
                break
            target = max(inner_tol, 0.5 * tol / rel.value)
            k, ab = _inner(sim, guess, target, smooth, max_cycles, chunk,
                           max_inner_sweeps, trace=trace, lmin=lmin, tile=fine_tile,
                           native_loop=native_loop)
            kept += k
            fallback |= ab
            _lib.call("nx_refine_accumulate", sim._alive())
    sweeps = sim.stats()["sweeps"]
    _lib.call("nx_step_finish", sim._alive())
    return {"sweeps": sweeps, "converged": converged, "residual": rel.value,
            "outer": outer, "learned_corrections": kept, "fallback": fallback}
