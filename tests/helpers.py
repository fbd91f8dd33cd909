"""Shared helpers for the GPU parity tests: random scenes and comparisons
against the CPU oracle (oracle/fluid.py)."""
import numpy as np

from oracle import fluid as F

f32 = np.float32


def random_velocity(rng, H, W, scale=1.0):
    return (rng.uniform(-1, 1, (H, W, 2)) * scale).astype(f32)


def random_mask(rng, H, W, frac=0.1, static_frac=0.3):
    """u8 mask with dynamic (bit0) and static-triangle (bit1) cells."""
    m = np.zeros((H, W), np.uint8)
    solid = rng.random((H, W)) < frac
    stat = rng.random((H, W)) < static_frac
    m[solid & ~stat] = 1
    m[solid & stat] = 2
    return m


def blob_mask(rng, H, W, n=3):
    """A few disks plus scattered single cells (edges and corners included)."""
    obs = np.zeros((H, W, 2), f32)
    for _ in range(n):
        c = rng.uniform(0.15, 0.85, 2)
        obs = F.add_circle_obstacle(obs, c, rng.uniform(1.5, 0.12 * min(H, W) + 2))
    m = mask_from_obs(obs)
    pts = rng.integers(0, H * W, max(1, H * W // 200))
    m.ravel()[pts] |= 2
    m[0, 0] |= 1
    m[-1, -1] |= 1
    return m


def mask_from_obs(obs):
    m = np.zeros(obs.shape[:2], np.uint8)
    m[obs[..., 0] > 0] |= 1
    m[obs[..., 1] > 0] |= 2
    return m


def obs_from_mask(m):
    return F.obstacles_from_mask(m)


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.max(np.abs(ref))
    num = np.max(np.abs(got - ref))
    if den == 0:
        return float(num)
    return float(num / den)


def assert_bitwise(got, ref, what=""):
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    if not np.array_equal(got, ref):
        bad = np.argwhere(got != ref)
        i = tuple(bad[0])
        raise AssertionError(
            f"{what}: {len(bad)} of {got.size} values differ; first at {i}: "
            f"got {got[i]!r} ref {ref[i]!r}; rel err {rel_err(got, ref):.3e}")


def assert_close(got, ref, tol=1e-5, what=""):
    e = rel_err(got, ref)
    assert e <= tol, f"{what}: relative error {e:.3e} > {tol}"


def make_sim(H, W, **kw):
    from natrix_amd import FluidSimulator
    return FluidSimulator(W, H, **kw)
