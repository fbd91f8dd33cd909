"""Pressure solve and full-step parity on the GPU.

* temporal-blocked Jacobi (k sweeps per launch, every tile shape) is bitwise
  equal to k single sweeps of the oracle;
* the FP64 residual matches the oracle's two-pass evaluation;
* the tolerance-driven solve stops at a residual <= tol, verified on the
  downloaded pressure by the oracle;
* fused and reference-order steps match the oracle step bitwise over
  10-step rollouts; a HIP-graph replay matches eager execution.
"""
import numpy as np
import pytest

from oracle import fluid as F
from oracle import poisson as P
from tests.helpers import (assert_bitwise, assert_close, blob_mask, f32,
                           make_sim, obs_from_mask, random_velocity)

pytestmark = pytest.mark.gpu

TILES = [(4, 8, 1), (4, 16, 1), (8, 8, 1), (8, 16, 1), (16, 8, 1), (2, 8, 2), (2, 16, 2),
         (4, 4, 2), (4, 8, 2), (4, 16, 2), (8, 4, 2), (8, 8, 2)]


@pytest.mark.parametrize("rows,waves,cols", TILES)
def test_jacobi_blocked_bitwise(gpu, rows, waves, cols):
    # odd and even widths and k: the float2 path needs both even
    for (H, W), k in [((100, 77), 1), ((100, 78), 4), ((100, 77), 5),
                      ((257, 130), 8), ((33, 200), 3), ((64, 256), 6)]:
        if 64 * cols - 2 * k < 8 or rows * waves - 2 * k < 8:
            continue
        rng = np.random.default_rng(H + W + k)
        m = blob_mask(rng, H, W)
        b = rng.standard_normal((H, W)).astype(f32)
        p = rng.standard_normal((H, W)).astype(f32)
        sim = make_sim(H, W)
        sim.upload("mask", m)
        sim.upload("divergence", b)
        sim.upload("pressure", p)
        sweeps = 3 * k + 2  # includes a shortened final launch
        sim.jacobi_blocked(sweeps, k, rows, waves, cols)
        ref = F.jacobi(p, b, obs_from_mask(m), sweeps)
        assert_bitwise(sim.download("pressure"), ref,
                       f"tb {rows}x{waves}x{cols} k={k} {H}x{W}")
        sim.destroy()


STREAM_TILES = [(4, 8, 2), (2, 16, 2), (4, 16, 2), (8, 8, 2), (8, 4, 2)]


@pytest.mark.parametrize("rows,waves,cols", STREAM_TILES)
def test_jacobi_streaming_bitwise(gpu, rows, waves, cols):
    """The streaming Jacobi kernel (a workgroup walks `seg` output rows of a
    column strip, carrying the two rows above each chunk per sweep) is
    bitwise equal to single sweeps of the oracle: odd and even widths and k,
    segments of one row, shorter and longer than a chunk, and longer than
    the grid, grids thinner than a chunk or a strip, k up to 16."""
    for (H, W), k, seg in [((100, 77), 1, 7), ((100, 78), 4, 33), ((100, 77), 5, 1),
                           ((257, 130), 8, 112), ((33, 200), 3, 1000), ((64, 256), 6, 48),
                           ((5, 300), 8, 16), ((300, 9), 16, 80), ((131, 128), 2, 131)]:
        rng = np.random.default_rng(H + W + k + seg)
        m = blob_mask(rng, H, W)
        b = rng.standard_normal((H, W)).astype(f32)
        p = rng.standard_normal((H, W)).astype(f32)
        sim = make_sim(H, W)
        sim.upload("mask", m)
        sim.upload("divergence", b)
        sim.upload("pressure", p)
        sweeps = 3 * k + 2  # includes a shortened final launch
        sim.jacobi_blocked(sweeps, k, rows, waves, cols, seg)
        ref = F.jacobi(p, b, obs_from_mask(m), sweeps)
        assert_bitwise(sim.download("pressure"), ref,
                       f"stream {rows}x{waves}x{cols} k={k} seg={seg} {H}x{W}")
        sim.destroy()


def test_residual_matches_oracle(gpu):
    H, W = 120, 90
    rng = np.random.default_rng(5)
    m = blob_mask(rng, H, W)
    obs = obs_from_mask(m)
    b = rng.standard_normal((H, W)).astype(f32)
    p = F.jacobi(np.zeros((H, W), f32), b, obs, 50)
    sim = make_sim(H, W)
    sim.upload("mask", m)
    sim.upload("divergence", b)
    sim.upload("pressure", p)
    rel, sums = sim.residual()
    ref = F.relative_residual(p, b, obs)
    assert abs(rel - ref) <= 1e-9 * ref
    one = F.residual_sums(p, b, obs)
    np.testing.assert_allclose(sums[:4], one[:4], rtol=1e-10)
    assert sums[4] == one[4]
    sim.destroy()


def test_residual_of_fp64_solution_is_small(gpu):
    H, W = 40, 48
    rng = np.random.default_rng(9)
    solid = np.zeros((H, W), bool)
    solid[10:20, 15:22] = True
    m = solid.astype(np.uint8)
    b = rng.standard_normal((H, W))
    b[solid] = 0
    b[~solid] -= b[~solid].mean()
    p = P.solve_fp64(b, solid)
    sim = make_sim(H, W)
    sim.upload("mask", m)
    sim.upload("divergence", b.astype(f32))
    sim.upload("pressure", p.astype(f32))
    rel, _ = sim.residual()
    assert rel < 1e-5  # float32 rounding of p and b only
    sim.destroy()


def _scene(H, W, seed, bc_mode=2):
    rng = np.random.default_rng(seed)
    u = random_velocity(rng, H, W, 0.5)
    m = np.zeros((H, W), np.uint8)
    obs = F.add_circle_obstacle(np.zeros((H, W, 2), f32), (0.4, 0.5), 0.1 * W)
    m[obs[..., 0] > 0] = 4  # persistent map
    return u, m


@pytest.mark.parametrize("fused,tile", [(True, None), (False, None), (True, (8, 4, 8, 2, 40))])
@pytest.mark.parametrize("bc_mode,nu", [(0, 0.1), (1, 0.0), (2, 0.0)])
def test_rollout_10_steps_bitwise(gpu, fused, tile, bc_mode, nu):
    H, W = 96, 128
    u, m = _scene(H, W, 21)
    sim = make_sim(H, W, dye=True, bc_mode=bc_mode, fused=fused)
    if tile is not None:
        sim.jacobi_tile = tile
    sim.vorticity = 0.3
    sim.viscosity = nu
    sim.iterations = 20
    sim.upload("velocity", u)
    sim.set_obstacle_map(m)
    q = np.zeros((H, W), f32)
    prm = F.Params(vorticity=0.3, viscosity=nu, iterations=20, bc_mode=bc_mode)
    dt = np.float32(1 / 60)
    pers = m != 0
    obs = np.zeros((H, W, 2), f32)
    p = np.zeros((H, W), f32)
    for step in range(10):
        splat = (0.25 + 0.05 * step, 0.5, 1.5, 0.3, 4.0, 1.0)
        sim.add_splats([splat])
        u, q = F.gaussian_splat(u, q, [tuple(np.float32(v) for v in splat)])
        # a dynamic obstacle re-added every step, as the demo does
        sim.add_circle_obstacle((0.75, 0.3), 5.0)
        obs = F.add_circle_obstacle(obs, (0.75, 0.3), 5.0)
        sim.update(dt)
        u, p, obs, q = F.step(u, p, obs, dt, prm, dye=q, persistent=pers)
    assert_bitwise(sim.download("velocity"), u, "rollout velocity")
    assert_bitwise(sim.download("pressure"), p, "rollout pressure")
    assert_bitwise(sim.download("dye"), q, "rollout dye")
    assert_bitwise(sim.download("mask"), m, "mask after clear")
    sim.destroy()


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("vel,dye", [((60, 80), (120, 160)), ((37, 50), (97, 131))])
def test_rollout_with_dye_at_its_own_resolution(gpu, fused, vel, dye):
    """The demo's use of the simulator: velocity at half the window
    (V = window / 2), dye at the framebuffer (P), including a ratio that is
    not an integer; velocity cone splats on V, dye cone splats on P, a
    dynamic obstacle every step; 10 update() steps bitwise equal to the
    oracle (simulation_demo.py:100-109, 220-237)."""
    H, W = vel
    Hp, Wp = dye
    u, m = _scene(H, W, 5)
    sim = make_sim(H, W, dye=(Wp, Hp), bc_mode=1, fused=fused)
    sim.vorticity = 0.3
    sim.iterations = 20
    sim.upload("velocity", u)
    sim.set_obstacle_map(m)
    q = np.zeros((Hp, Wp), f32)
    prm = F.Params(vorticity=0.3, iterations=20, bc_mode=1)
    dt = np.float32(1 / 60)
    pers = m != 0
    obs = np.zeros((H, W, 2), f32)
    p = np.zeros((H, W), f32)
    for step in range(10):
        pos = (0.2 + 0.04 * step, 0.55)
        sim.add_velocity(pos, (0.8, -0.2), 6.0)
        u = F.add_velocity(u, tuple(np.float32(v) for v in pos),
                           (np.float32(0.8), np.float32(-0.2)), np.float32(6.0))
        sim.add_dye(pos, 9.0, 120.0)
        q = F.add_dye(q, tuple(np.float32(v) for v in pos), np.float32(120.0), np.float32(9.0))
        sim.add_circle_obstacle((0.7, 0.35), 4.0)
        obs = F.add_circle_obstacle(obs, (0.7, 0.35), 4.0)
        sim.update(dt)
        u, p, obs, q = F.step(u, p, obs, dt, prm, dye=q, persistent=pers)
    assert_bitwise(sim.download("velocity"), u, "rollout velocity")
    assert_bitwise(sim.download("dye"), q, "rollout dye")
    assert np.abs(q).max() > 1.0  # the dye moved and accumulated
    sim.destroy()


def test_graph_replay_matches_eager(gpu):
    H, W = 80, 112
    u, m = _scene(H, W, 33)
    sims = []
    for graph in (False, True):
        sim = make_sim(H, W, dye=True, bc_mode=2)
        sim.vorticity = 0.3
        sim.viscosity = 0.0
        sim.iterations = 17
        sim.upload("velocity", u)
        sim.set_obstacle_map(m)
        sim.use_graph = graph
        sims.append(sim)
    for step in range(6):
        for sim in sims:
            sim.add_splats([(0.3, 0.5, 1.0, 0.5, 3.0, 1.0)])
            sim.update(1 / 60)
        # every replay checked against the eager result
        a, b = (s.download("velocity") for s in sims)
        assert_bitwise(b, a, f"graph replay step {step}")
        assert_bitwise(sims[1].download("dye"), sims[0].download("dye"), "graph dye")
    for s in sims:
        s.destroy()


@pytest.mark.parametrize("tile", [(8, 8, 8, 1), (8, 4, 8, 2, 24)])
def test_tolerance_solve_reaches_tol(gpu, tile):
    """Also with the streaming Jacobi kernel, whose last launch of a chunk
    writes the residual records per workgroup."""
    H, W = 64, 64
    u, m = _scene(H, W, 44)
    sim = make_sim(H, W, bc_mode=2)
    sim.vorticity = 0.3
    sim.viscosity = 0.0
    sim.upload("velocity", u)
    sim.set_obstacle_map(m)
    sim.tol = 1e-4
    sim.check_every = 64
    sim.jacobi_tile = tile
    sim.update(1 / 60)
    st = sim.stats()
    assert st["converged"] and st["residual"] <= 1e-4
    p = sim.download("pressure")
    b = sim.download("divergence")
    obs = obs_from_mask(m)
    rel = F.relative_residual(p, b, obs)
    assert rel <= 1e-4, rel
    assert abs(rel - st["residual"]) <= 1e-9
    # the right-hand side is the oracle's divergence centred on its fluid
    # mean, and p is Jacobi on that system from its first sweep
    dt = np.float32(1 / 60)
    v = F.enforce_boundaries(u, F.BC_INFLOW)
    v = F.advect_velocity(v, obs, dt, 500.0, 1.0)
    v = F.apply_vorticity(v, F.calc_vorticity(v), dt, 0.3)
    b0 = F.divergence(v, obs)
    fluid = m == 0
    mean = np.float32(b0[fluid].astype(np.float64).sum() / fluid.sum())
    bc = b0.copy()
    bc[fluid] = b0[fluid] - mean
    assert_bitwise(b, bc, "centred rhs")
    # solid cells keep b0, so (0 - bc) * 0.25 is the first sweep everywhere
    p1 = (np.float32(0) - bc) * np.float32(0.25)
    q = F.jacobi(p1, bc, obs, st["sweeps"] - 1)
    assert_bitwise(p, q, "tolerance solve vs oracle")
    # it stopped within two check intervals of the first iterate meeting tol
    r = p1
    k_first = None
    for k in range(1, st["sweeps"] + 1):
        if F.relative_residual(r, bc, obs) <= 1e-4:
            k_first = k
            break
        r = F.jacobi(r, bc, obs, 1)
    assert k_first is not None and st["sweeps"] - k_first <= 2 * 64 + 8
    sim.destroy()


def test_tolerance_solve_respects_cap(gpu):
    H, W = 64, 64
    u, m = _scene(H, W, 45)
    sim = make_sim(H, W, bc_mode=2)
    sim.upload("velocity", u)
    sim.set_obstacle_map(m)
    sim.viscosity = 0.0
    sim.tol = 1e-12
    sim.max_iters = 300
    sim.check_every = 64
    sim.update(1 / 60)
    st = sim.stats()
    assert not st["converged"] and st["sweeps"] == 300
    sim.destroy()


def test_velocity_buffer_interop_with_torch(gpu):
    import torch
    H, W = 32, 48
    rng = np.random.default_rng(1)
    u = random_velocity(rng, H, W)
    sim = make_sim(H, W)
    sim.upload("velocity", u)
    buf = sim.get_velocity_buffer()
    t = torch.as_tensor(buf, device="cuda")
    assert tuple(t.shape) == (H, W, 2)
    assert np.array_equal(t.cpu().numpy(), u)
    sim.destroy()


def test_simulate_false_freezes(gpu):
    H, W = 16, 16
    sim = make_sim(H, W)
    u = random_velocity(np.random.default_rng(2), H, W)
    sim.upload("velocity", u)
    sim.simulate = False
    sim.update(1 / 60)
    sim.add_velocity((0.5, 0.5), (1, 1), 4)
    sim.add_circle_obstacle((0.5, 0.5), 3)
    assert_bitwise(sim.download("velocity"), u, "frozen")
    assert np.all(sim.download("mask") == 0)
    sim.destroy()


def test_fp64_oracle_solve_matches_scipy(gpu):
    import torch
    from natrix_amd import fp64
    H, W = 48, 64
    rng = np.random.default_rng(12)
    solid = np.zeros((H, W), bool)
    solid[10:22, 30:37] = True
    solid[35:40, 5:9] = True
    b = rng.standard_normal((H, W)).astype(f32)
    sim = make_sim(H, W)
    sim.upload("mask", solid.astype(np.uint8))
    sim.upload("divergence", b)
    x, info = fp64.solve(sim, tol=1e-12)
    assert info["residual"] < 1e-11
    bb = b.astype(np.float64)
    bb[solid] = 0
    ref = P.solve_fp64(bb, solid)
    got = x.cpu().numpy()
    fl = ~solid
    got = got - got[fl].mean()
    assert np.abs(got[fl] - ref[fl]).max() < 1e-8 * np.abs(ref).max()
    # error metric: the exact solution rounded to float32 has a tiny error,
    # a Jacobi iterate a large one
    sim.upload("pressure", ref.astype(f32))
    assert fp64.pressure_error(sim, x) < 1e-6
    sim.upload("pressure", F.jacobi(np.zeros((H, W), f32), b, obs_from_mask(solid.astype(np.uint8)), 20))
    assert fp64.pressure_error(sim, x) > 0.1
    sim.destroy()
