"""Pins the reduced-storage oracle (oracle/mixed.py) on the CPU: its
rounding functions against torch's own float32 -> bfloat16 / float16
conversions (round to nearest even), and its dataflow, with float32
(identity) rounding and no correction, against oracle.fluid.step."""
import numpy as np
import pytest
import torch

from oracle import fluid as F
from oracle import mixed as M


def test_round_bf16_matches_torch(rng):
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10,
                        np.float32([1.0, 1.00390625, 1.01171875, -3.5e-3, 0.0, 65504.0])])
    # exact ties: 1 + 2^-8 rounds to even (1.0); 1 + 3*2^-8 rounds up
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.round_bf16(x), ref)
    assert M.round_bf16(np.float32([1.00390625]))[0] == 1.0


def test_round_f16_matches_torch(rng):
    x = rng.standard_normal(20000).astype(np.float32) * 100
    ref = torch.from_numpy(x).to(torch.float16).to(torch.float32).numpy()
    assert np.array_equal(M.round_f16(x), ref)


def test_identity_rounding_reproduces_reference_step(rng):
    H, W = 24, 32
    u = rng.uniform(-0.5, 0.5, (H, W, 2)).astype(np.float32)
    obs = F.add_circle_obstacle(np.zeros((H, W, 2), np.float32), (0.5, 0.5), 3.0)
    prm = F.Params(vorticity=0.3, viscosity=0.2, iterations=11, bc_mode=F.BC_FREE_SLIP)
    dt = np.float32(1 / 60)
    q = rng.uniform(0, 1, (H, W)).astype(np.float32)
    v, p, p32, b, dye = M.step_mixed(u, obs, dt, prm, "f32", k=4, corr_sweeps=0, dye=q)
    ur, pr, _, qr, st = F.step(u, np.zeros((H, W), np.float32), obs, dt, prm, dye=q,
                               return_stages=True)
    assert np.array_equal(v, ur) and np.array_equal(p, pr) and np.array_equal(dye, qr)
    assert np.array_equal(p32, pr) and np.array_equal(b, st["divergence"])


def test_fp64_twin_is_the_float64_reference_step_plus_correction(rng):
    """With float64 inputs and "f64" (identity) rounding, step_mixed is the
    float64 step of oracle.fluid (no correction), and the correction
    launch adds exactly kc Jacobi sweeps of A e = b - A p from 0 in
    float64: the reference the reduced-storage errors are measured
    against."""
    H, W = 24, 32
    u = rng.uniform(-0.5, 0.5, (H, W, 2))
    obs = F.add_circle_obstacle(np.zeros((H, W, 2), np.float32), (0.5, 0.5), 3.0).astype(
        np.float64)
    prm = F.Params(vorticity=0.3, viscosity=0.0, iterations=11, bc_mode=F.BC_FREE_SLIP)
    q = rng.uniform(0, 1, (H, W))
    v, p, p32, b, dye = M.step_mixed(u, obs, 1 / 60, prm, "f64", k=4, corr_sweeps=0, dye=q)
    assert v.dtype == p32.dtype == dye.dtype == np.float64
    ur, pr, _, qr = F.step(u, np.zeros((H, W)), obs, 1 / 60, prm, dye=q)
    assert np.array_equal(v, ur) and np.array_equal(p32, pr) and np.array_equal(dye, qr)
    _, p8, p32c, b8, _ = M.step_mixed(u, obs, 1 / 60, prm, "f64", k=4, corr_sweeps=5, dye=q)
    solid = F.solid_of(obs)
    x1, x2, y1, y2 = F._replaced_neighbours(p8, solid)
    r = b8 - ((((x1 + x2) + y1) + y2) - 4.0 * p8)
    e = F.jacobi(np.zeros_like(p8), r, obs, 5)
    assert np.array_equal(p32c, p8 + e)


def test_correction_recovers_float32_residual(rng):
    # a bf16-stored Jacobi iterate carries ~2^-9 relative noise, which raises
    # its residual above the float32 iterate's; the float32 correction pass
    # (A e = r from 0) must bring it back to (at least) the float32 level
    H, W = 48, 48
    obs = np.zeros((H, W, 2), np.float32)
    b = (rng.standard_normal((H, W)) * 0.01).astype(np.float32)
    p = F.jacobi(np.zeros((H, W), np.float32), b, obs, 200)
    pb = M.round_bf16(p)
    bb = M.round_bf16(b)
    fixed = M.correction(pb, bb, obs, 8)
    r_f32 = F.relative_residual(p, b, obs)
    r_bf16 = F.relative_residual(pb, bb, obs)
    r_fix = F.relative_residual(fixed, bb, obs)
    assert r_bf16 > r_f32 * 1.05
    assert r_fix <= r_f32
