"""Per-stage parity: each HIP kernel against the oracle's transcription of
the same reference shader, on the same seeded inputs.

The library is built with FMA contraction off, so every stage whose
arithmetic is + - * / sqrt is expected to be bitwise equal to the float32
oracle.  The tolerance bar (max |diff| / max |ref| <= 1e-5 per field) is
asserted as well wherever bitwise equality is not claimed."""
import numpy as np
import pytest

from oracle import fluid as F
from tests.helpers import (assert_bitwise, assert_close, blob_mask, f32,
                           make_sim, obs_from_mask, random_velocity)

pytestmark = pytest.mark.gpu

SIZES = [(5, 7), (37, 53), (64, 64), (130, 97)]
DT = np.float32(1 / 60)


@pytest.fixture(params=SIZES, ids=lambda s: f"{s[0]}x{s[1]}")
def scene(request, gpu):
    H, W = request.param
    rng = np.random.default_rng(H * 1000 + W)
    u = random_velocity(rng, H, W)
    m = blob_mask(rng, H, W)
    sim = make_sim(H, W)
    sim.upload("velocity", u)
    sim.upload("mask", m)
    yield sim, u, m, rng
    sim.destroy()


@pytest.mark.parametrize("mode,borders", [(0, True), (0, False), (1, True), (2, True)])
def test_boundaries(gpu, mode, borders):
    for (H, W) in [(3, 3), (17, 29), (64, 66)]:
        rng = np.random.default_rng(7)
        u = random_velocity(rng, H, W)
        sim = make_sim(H, W, bc_mode=mode, inflow=0.75)
        sim.has_borders = borders
        sim.upload("velocity", u)
        sim.stage("boundaries")
        ref = F.enforce_boundaries(u, mode, borders, 0.75)
        assert_bitwise(sim.download("velocity"), ref, f"bc {mode} {H}x{W}")
        sim.destroy()


@pytest.mark.parametrize("speed,diss", [(500.0, 1.0), (37.5, 0.98)])
def test_advect(scene, speed, diss):
    sim, u, m, _ = scene
    sim.speed = speed
    sim.dissipation = diss
    sim.stage("advect", DT)
    ref = F.advect_velocity(u, obs_from_mask(m), DT, speed, diss)
    assert_bitwise(sim.download("velocity"), ref, "advect")


def test_advect_long_back_traces(gpu):
    """Velocities up to 3 per component (as unclamped Gaussian splats leave
    them) back-trace up to 25 cells at speed 500, past the grid's edge from
    many cells (corners clamped into the grid): bitwise equal to the
    oracle."""
    for (H, W) in [(37, 53), (130, 97), (96, 200)]:
        rng = np.random.default_rng(H * 7 + W)
        u = random_velocity(rng, H, W, 3.0)
        m = blob_mask(rng, H, W)
        sim = make_sim(H, W)
        sim.upload("velocity", u)
        sim.upload("mask", m)
        sim.speed = 500.0
        sim.stage("advect", DT)
        ref = F.advect_velocity(u, obs_from_mask(m), DT, 500.0, 1.0)
        assert_bitwise(sim.download("velocity"), ref, f"advect long back-traces {H}x{W}")
        sim.destroy()


def test_curl_and_confine(scene):
    sim, u, m, _ = scene
    sim.vorticity = 0.3
    sim.stage("curl")
    w = F.calc_vorticity(u)
    assert_bitwise(sim.download("curl"), w, "curl")
    sim.stage("confine", DT)
    ref = F.apply_vorticity(u, w, DT, 0.3)
    got = sim.download("velocity")
    assert_close(got, ref, 1e-5, "confine")
    assert_bitwise(got, ref, "confine")


def test_viscosity(scene):
    sim, u, m, _ = scene
    sim.viscosity = 0.5
    sim.stage("viscosity")
    a, rb = F.viscosity_coefficients(0.5)
    assert_bitwise(sim.download("velocity"), F.viscosity(u, a, rb), "viscosity")


def test_divergence(scene):
    sim, u, m, _ = scene
    sim.stage("divergence")
    assert_bitwise(sim.download("divergence"), F.divergence(u, obs_from_mask(m)),
                   "divergence")


def test_jacobi_single_sweeps(scene):
    sim, u, m, rng = scene
    H, W = m.shape
    b = rng.standard_normal((H, W)).astype(f32)
    p = rng.standard_normal((H, W)).astype(f32)
    sim.upload("divergence", b)
    sim.upload("pressure", p)
    sim.stage("jacobi", sweeps=7)
    ref = F.jacobi(p, b, obs_from_mask(m), 7)
    assert_bitwise(sim.download("pressure"), ref, "jacobi x7")


def test_clear_pressure_then_jacobi_from_zero(scene):
    sim, u, m, rng = scene
    H, W = m.shape
    b = rng.standard_normal((H, W)).astype(f32)
    sim.upload("divergence", b)
    sim.upload("pressure", np.ones((H, W), f32))
    sim.stage("clear_pressure")
    assert np.all(sim.download("pressure") == 0)
    sim.stage("jacobi", sweeps=3)
    ref = F.jacobi(np.zeros((H, W), f32), b, obs_from_mask(m), 3)
    assert_bitwise(sim.download("pressure"), ref, "jacobi from zero")


def test_subtract_gradient(scene):
    sim, u, m, rng = scene
    H, W = m.shape
    p = rng.standard_normal((H, W)).astype(f32)
    sim.upload("pressure", p)
    sim.stage("subtract_gradient")
    ref = F.subtract_gradient(u, p, obs_from_mask(m))
    assert_bitwise(sim.download("velocity"), ref, "subtract_gradient")


def test_clear_obstacles_keeps_persistent(gpu):
    H, W = 20, 30
    sim = make_sim(H, W)
    m = np.zeros((H, W), np.uint8)
    m[3:6, 4:9] = 1
    m[10, 10] = 2
    m[15, 2:5] = 4
    sim.upload("mask", m)
    sim.stage("clear_obstacles")
    assert_bitwise(sim.download("mask"), (m & 4), "clear obstacles")
    sim.destroy()


def test_add_velocity(scene):
    sim, u, m, _ = scene
    H, W = m.shape
    r = max(2.0, 0.3 * min(H, W))
    sim.add_velocity((0.41, 0.57), (0.9, -1.7), r)
    ref = F.add_velocity(u, (0.41, 0.57), (0.9, -1.7), r)
    assert_bitwise(sim.download("velocity"), ref, "add_velocity")


def test_circle_and_triangle_obstacles(gpu):
    for (H, W) in [(16, 16), (61, 83)]:
        sim = make_sim(H, W)
        obs = np.zeros((H, W, 2), f32)
        sim.add_circle_obstacle((0.3, 0.6), 0.2 * W)
        obs = F.add_circle_obstacle(obs, (0.3, 0.6), 0.2 * W)
        sim.add_triangle_obstacle((0.5, 0.1), (0.9, 0.4), (0.6, 0.8), static=True)
        obs = F.add_triangle_obstacle(obs, (0.5, 0.1), (0.9, 0.4), (0.6, 0.8), True)
        sim.add_triangle_obstacle((0.05, 0.9), (0.2, 0.3), (0.45, 0.95), static=False)
        obs = F.add_triangle_obstacle(obs, (0.05, 0.9), (0.2, 0.3), (0.45, 0.95), False)
        got = obs_from_mask(sim.download("mask"))
        assert_bitwise(got, obs, f"obstacles {H}x{W}")
        sim.destroy()


@pytest.mark.parametrize("dye_shape", [(64, 64), (128, 96), (50, 77)])
def test_advect_dye(gpu, dye_shape):
    H, W = 64, 64
    Hp, Wp = dye_shape
    rng = np.random.default_rng(Hp + Wp)
    sim = make_sim(H, W, dye=(Wp, Hp))
    u = random_velocity(rng, H, W)
    m = blob_mask(rng, H, W)
    q = rng.uniform(0, 2, (Hp, Wp)).astype(f32)
    sim.upload("velocity", u)
    sim.upload("mask", m)
    sim.upload("dye", q)
    sim._dye_dissipation = 0.98
    sim.stage("advect_dye", DT)
    ref = F.advect_dye(q, u, obs_from_mask(m), DT, 500.0, 0.98)
    assert_bitwise(sim.download("dye"), ref, f"advect_dye {dye_shape}")
    sim.destroy()


def test_add_dye(gpu):
    H, W = 48, 80
    rng = np.random.default_rng(3)
    sim = make_sim(H, W, dye=True)
    q = rng.uniform(0, 300, (H, W)).astype(f32)
    sim.upload("dye", q)
    sim.add_dye((0.6, 0.35), 9.5, 120.0)
    ref = F.add_dye(q, (0.6, 0.35), 120.0, 9.5)
    assert_bitwise(sim.download("dye"), ref, "add_dye")
    sim.destroy()


def test_gaussian_splats(gpu):
    H, W = 96, 128
    rng = np.random.default_rng(11)
    sim = make_sim(H, W, dye=True)
    u = random_velocity(rng, H, W, 0.2)
    q = rng.uniform(0, 1, (H, W)).astype(f32)
    sim.upload("velocity", u)
    sim.upload("dye", q)
    splats = [(0.3, 0.4, 2.5, -1.0, 5.12, 1.0), (0.7, 0.6, -0.5, 3.0, 3.0, 0.5)]
    splats += [(rng.uniform(), rng.uniform(), rng.normal(), rng.normal(), 4.0, 1.0)
               for _ in range(20)]  # more than one by-value batch
    sim.add_splats(splats)
    splats32 = [tuple(np.float32(v) for v in s) for s in splats]
    ur, qr = F.gaussian_splat(u, q, splats32)
    assert_close(sim.download("velocity"), ur, 1e-6, "splat velocity")
    assert_close(sim.download("dye"), qr, 1e-6, "splat dye")
    assert_bitwise(sim.download("velocity"), ur, "splat velocity")
    sim.destroy()


@pytest.mark.parametrize("nu", [0.0, 0.5])
def test_fused_vort_div_equals_stage_chain(scene, nu):
    sim, u, m, _ = scene
    sim.vorticity = 0.3
    sim.viscosity = nu
    sim.stage("fused_vort_div", DT)
    v = F.apply_vorticity(u, F.calc_vorticity(u), DT, 0.3)
    if nu > 0:
        a, rb = F.viscosity_coefficients(nu)
        v = F.viscosity(v, a, rb)
    obs = obs_from_mask(m)
    b = F.divergence(v, obs)
    assert_bitwise(sim.download("velocity"), v, "fused velocity")
    assert_bitwise(sim.download("divergence"), b, "fused divergence")
    p1 = F.jacobi(np.zeros_like(b), b, obs, 1)
    assert_bitwise(sim.download("pressure"), p1, "fused p1")


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_dye_rgba8_export(gpu, rng, storage):
    """The dye display pass: bitwise equal to the oracle's rgba8 conversion
    of the stored dye, on a dye grid of its own size."""
    from natrix_amd import FluidSimulator
    sim = FluidSimulator(40, 30, dye=(77, 50), storage=storage) if storage == "f32" else \
        FluidSimulator(40, 30, dye=True, storage=storage)
    h, w = sim.dye_shape
    q = rng.uniform(-0.5, 1.5, (h, w)).astype(np.float32)
    q[0, 0] = np.nan
    sim.upload("dye", q)
    stored = sim.download("dye")
    ref = F.dye_rgba8(stored)
    assert np.array_equal(sim.dye_rgba8(), ref)
    sim.destroy()


@pytest.mark.parametrize("storage", ["f32", "bf16"])
@pytest.mark.parametrize("vel,window,tile", [((37, 50), (148, 200), 16.0),
                                             ((60, 80), (120, 160), 12.0),
                                             ((64, 64), (97, 131), 20.0)])
def test_quiver_rgba8(gpu, rng, storage, vel, window, tile):
    """The demo's velocity arrows (demo.QuiverFragmentShader.frag): bitwise
    equal to the oracle's statement of the stored velocity, for grids whose
    window scale is and is not an integer and partial tiles at the edges."""
    from natrix_amd import FluidSimulator
    H, W = vel
    Hw, Ww = window
    sim = FluidSimulator(W, H, storage=storage)
    u = rng.uniform(-0.3, 0.3, (H, W, 2)).astype(np.float32)
    u[: H // 3] *= np.float32(30.0)  # arrows long enough to be clamped
    u[H // 2, :] = 0.0               # and tiles with no arrow
    sim.upload("velocity", u)
    stored = sim.download("velocity").astype(np.float32)
    img = sim.quiver_rgba8(Ww, Hw, tile)
    assert img.shape == (Hw, Ww, 4)
    assert np.array_equal(img, F.quiver_rgba8(stored, Ww, Hw, tile))
    assert (img[..., 3] > 0).any() and (img[..., 3] == 0).any()
    for bad in ((0, Hw, tile), (Ww, Hw, 0.0), (Ww, Hw, float("inf"))):
        with pytest.raises(ValueError):
            sim.quiver_rgba8(*bad)
    sim.destroy()
