"""Reduced-precision storage on the GPU: the bfloat16 / float16 fused step
is bitwise equal to the oracle's replay of the same dataflow
(oracle/mixed.py), and its error against the float64 rollout of the
reference algorithm is bounded."""
import numpy as np
import pytest

from oracle import fluid as F
from oracle import mixed as M
from tests.helpers import assert_bitwise, make_sim, random_velocity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("corr", [8, 0])
def test_mixed_step_matches_replay(gpu, storage, corr):
    H, W, steps, k = 80, 112, 4, 6
    rng = np.random.default_rng(17)
    R = M.rounder(storage)
    u = R(random_velocity(rng, H, W, 0.4))
    solid = np.zeros((H, W), bool)
    obs0 = F.add_circle_obstacle(np.zeros((H, W, 2), np.float32), (0.4, 0.5), 9.0)
    solid[obs0[..., 0] > 0] = True
    sim = make_sim(H, W, dye=True, bc_mode=2, storage=storage)
    sim.vorticity = 0.3
    sim.viscosity = 0.0
    sim.iterations = 31
    sim.jacobi_tile = (k, 4, 8, 2)
    sim.correction_sweeps = corr
    sim.upload("velocity", u)
    sim.set_obstacle_map(solid)
    prm = F.Params(vorticity=0.3, viscosity=0.0, iterations=31, bc_mode=F.BC_INFLOW)
    dt = np.float32(1 / 60)
    q = np.zeros((H, W), np.float32)
    obs = np.zeros((H, W, 2), np.float32)
    for s in range(steps):
        splat = (0.3, 0.5 + 0.02 * s, 1.2, -0.4, 5.0, 1.0)
        sim.add_splats([splat])
        sim.update(dt)
        u, p, p32, b, q = M.step_mixed(u, obs, dt, prm, storage, k, corr, dye=q,
                                       persistent=solid,
                                       splats=[tuple(np.float32(v) for v in splat)])
    assert_bitwise(sim.download("velocity"), u, f"{storage} velocity")
    assert_bitwise(sim.download("pressure"), p, f"{storage} stored pressure")
    assert_bitwise(sim.download("pressure32"), p32, f"{storage} corrected pressure")
    assert_bitwise(sim.download("dye"), q, f"{storage} dye")
    sim.destroy()


def test_mixed_error_against_fp64(gpu):
    """Relative L2 velocity error of short rollouts against the float64
    oracle: float32 storage < float16 < bfloat16, each within a bound set by
    its unit round-off (bf16 2^-9, f16 2^-12) times a modest amplification."""
    H, W, steps = 96, 96, 3
    rng = np.random.default_rng(23)
    u0 = random_velocity(rng, H, W, 0.2)
    solid = np.zeros((H, W), bool)
    solid[40:56, 30:38] = True
    prm = F.Params(vorticity=0.3, viscosity=0.0, iterations=60, speed=100.0,
                   bc_mode=F.BC_FREE_SLIP)
    errs = {}
    for storage in ("f32", "f16", "bf16"):
        u0r = M.rounder(storage)(u0)
        sim = make_sim(H, W, bc_mode=1, storage=storage)
        sim.vorticity = 0.3
        sim.viscosity = 0.0
        sim.speed = 100.0
        sim.iterations = 60
        sim.upload("velocity", u0r)
        sim.set_obstacle_map(solid)
        u64 = u0r.astype(np.float64)
        p64 = np.zeros((H, W))
        o64 = np.zeros((H, W, 2))
        for _ in range(steps):
            sim.update(1 / 60)
            u64, p64, o64, _ = F.step(u64, p64, o64, 1 / 60, prm, persistent=solid)
        got = sim.download("velocity").astype(np.float64)
        errs[storage] = float(np.linalg.norm(got - u64) / np.linalg.norm(u64))
        sim.destroy()
    print("velocity L2 error vs fp64:", errs)
    assert errs["f32"] < 1e-4
    assert errs["f16"] < 1e-2 and errs["bf16"] < 5e-2
    assert errs["f32"] < errs["f16"] < errs["bf16"]
