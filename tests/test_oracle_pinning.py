"""Pins the CPU oracle (oracle/fluid.py, oracle/poisson.py).

The reference ships no tests, golden vectors or fixtures (its
tests/__init__.py is empty) and cannot run here (bgfx is absent), so each
oracle function is pinned by something other than itself: closed forms,
scipy routines, an independently assembled sparse operator, hand-computed
values on tiny grids, and invariants.  Each check is chosen so that a
plausible mistake (dropped term, wrong sign, swapped neighbour, transposed
axis) fails it.
"""
import math

import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import fluid as F
from oracle import poisson as P

f32 = np.float32


def _obs(H, W):
    return np.zeros((H, W, 2), f32)


# ------------------------------------------------------------ neighbours
def test_neighbours_hand_values():
    a = np.arange(12, dtype=f32).reshape(3, 4)  # a[y, x] = 4y + x
    L, R, B, T = F.neighbours(a)
    # cell (x=0, y=0): L clamps to itself, R = (1,0), B clamps, T = (0,1)
    assert (L[0, 0], R[0, 0], B[0, 0], T[0, 0]) == (0, 1, 0, 4)
    # cell (x=3, y=2): L = (2,2)=10, R clamps=11, B=(3,1)=7, T clamps=11
    assert (L[2, 3], R[2, 3], B[2, 3], T[2, 3]) == (10, 11, 7, 11)
    # interior (x=1, y=1) = 5
    assert (L[1, 1], R[1, 1], B[1, 1], T[1, 1]) == (4, 6, 1, 9)


# ------------------------------------------------------------ S1
def test_init_boundaries_ring(rng):
    u = rng.uniform(-1, 1, (6, 9, 2)).astype(f32)
    v = F.init_boundaries(u)
    ring = np.ones((6, 9), bool)
    ring[1:-1, 1:-1] = False
    assert np.all(v[ring] == 0)
    assert np.array_equal(v[~ring], u[~ring])


# ------------------------------------------------------------ S2
def test_advect_uniform_field_is_uniform():
    # a constant field resamples to itself; dissipation scales; clamp to 1
    u = np.zeros((8, 8, 2), f32)
    u[..., 0] = 0.75
    u[..., 1] = -0.5
    v = F.advect_velocity(u, _obs(8, 8), 0.5, 3.0, 1.0)
    assert np.all(v[..., 0] == f32(0.75)) and np.all(v[..., 1] == f32(-0.5))
    v = F.advect_velocity(u, _obs(8, 8), 0.5, 3.0, 2.0)
    assert np.all(v[..., 0] == 1.0) and np.all(v[..., 1] == -1.0)


def test_advect_matches_scipy_bilinear_fp64(rng):
    # FP64 twin vs scipy.ndimage.map_coordinates(order=1, mode='nearest'):
    # clamped-corner bilinear equals edge-clamped bilinear in exact arithmetic.
    H, W = 13, 17
    u = rng.uniform(-1, 1, (H, W, 2))
    dt, speed = 1.0 / 60.0, 150.0
    v = F.advect_velocity(u, np.zeros((H, W, 2)), dt, speed, 1.0)
    X, Y = np.meshgrid(np.arange(W, dtype=float), np.arange(H, dtype=float))
    fx = X - u[..., 0] * dt * speed
    fy = Y - u[..., 1] * dt * speed
    for c in range(2):
        ref = ndi.map_coordinates(u[..., c], [fy, fx], order=1, mode="nearest")
        np.testing.assert_allclose(v[..., c], np.clip(ref, -1, 1), atol=1e-12)


def test_advect_obstacle_zero_and_integer_shift(rng):
    # velocity (1, 0) with dt*speed = 1 (powers of two) is an exact shift by
    # one cell; solid cells give 0.
    u = np.zeros((5, 6, 2), f32)
    u[..., 0] = 1.0
    obs = _obs(5, 6)
    obs[2, 3, 1] = 1.0
    v = F.advect_velocity(u, obs, 0.5, 2.0, 1.0)
    assert np.all(v[2, 3] == 0)
    m = np.ones((5, 6), bool)
    m[2, 3] = False
    assert np.all(v[m][:, 0] == 1.0) and np.all(v[m][:, 1] == 0.0)


def test_advect_hand_value():
    # 2x2 grid, cell (1,1) with velocity (0.5, 0.25), dt*speed = 1:
    # back-trace to (0.5, 0.75) -> bilinear of the four cells
    u = np.array([[[0.0, 0.0], [0.2, 0.0]],
                  [[0.4, 0.0], [0.5, 0.25]]], f32)  # u[y, x]
    v = F.advect_velocity(u, _obs(2, 2), 1.0, 1.0, 1.0)
    # corners: lb=u[0,0].x=0, rb=u[0,1].x=0.2, lt=u[1,0].x=0.4, rt=u[1,1].x=0.5
    h1 = 0.4 * 0.5 + 0.5 * 0.5
    h2 = 0.0 * 0.5 + 0.2 * 0.5
    exp_x = h2 * 0.25 + h1 * 0.75
    assert abs(float(v[1, 1, 0]) - exp_x) < 1e-7
    # y component: only u[1,1].y = 0.25 non-zero: weight 0.5*0.75
    assert abs(float(v[1, 1, 1]) - 0.25 * 0.5 * 0.75) < 1e-7


# ------------------------------------------------------------ S3 / S4
def test_curl_linear_field():
    H, W = 7, 9
    a, b = 0.3, -0.7
    X, Y = np.meshgrid(np.arange(W, dtype=float), np.arange(H, dtype=float))
    u = np.stack([a * Y, b * X], -1)
    w = F.calc_vorticity(u)
    np.testing.assert_allclose(w[1:-1, 1:-1], b - a, atol=1e-12)
    # left column: R.y - L.y = b*1 - 0 (one-sided), curl = 0.5*(b - 2a)
    np.testing.assert_allclose(w[1:-1, 0], 0.5 * (b - 2 * a), atol=1e-12)


def test_confinement_irrotational_is_identity(rng):
    u = rng.uniform(-1, 1, (6, 7, 2)).astype(f32)
    w = np.zeros((6, 7), f32)
    v = F.apply_vorticity(u, w, 1 / 60, 0.3)
    assert np.array_equal(v, u)


def test_confinement_closed_form():
    # w = x (positive): |wR|-|wL| = 2 in the interior, |wT|-|wB| = 0
    # -> F = 0.5*(0, 2) = (0, 1), |F| = 1, times (s w, -s w) -> (0, -s x)
    H, W = 5, 8
    w = np.broadcast_to(np.arange(W, dtype=float)[None, :], (H, W)).copy()
    u = np.zeros((H, W, 2))
    s, dt = 0.3, 0.1
    v = F.apply_vorticity(u, w, dt, s)
    X = np.arange(1, W - 1, dtype=float)
    np.testing.assert_allclose(v[:, 1:-1, 1], np.broadcast_to(-s * X * dt, (H, W - 2)), rtol=1e-12)
    np.testing.assert_allclose(v[:, 1:-1, 0], 0, atol=1e-15)


def test_confinement_epsilon_floor():
    # tiny gradient: |F|^2 < eps -> F scaled by 1/sqrt(eps), not normalised
    H, W = 3, 5
    w = np.broadcast_to(1e-3 * np.arange(W, dtype=float)[None, :], (H, W)).copy()
    v = F.apply_vorticity(np.zeros((H, W, 2)), w, 1.0, 1.0)
    # interior: F = (0, 1e-3); scaled by 1/sqrt(2.4414e-4); times -w
    x = 2
    expect = -(1e-3 / math.sqrt(2.4414e-4)) * (1e-3 * x)
    assert abs(v[1, x, 1] - expect) < 1e-15


# ------------------------------------------------------------ S5
def test_viscosity_matches_convolution(rng):
    u = rng.uniform(-1, 1, (9, 11, 2))
    alpha, rbeta = 10.0, 1.0 / 14.0
    v = F.viscosity(u, alpha, rbeta)
    k = np.array([[0, 1, 0], [1, alpha, 1], [0, 1, 0]]) * rbeta
    for c in range(2):
        ref = ndi.convolve(u[..., c], k, mode="nearest")
        np.testing.assert_allclose(v[..., c], ref, atol=1e-12)


def test_viscosity_coefficients_float64_then_float32():
    a, rb = F.viscosity_coefficients(0.1)
    assert a == np.float32(10.0) and rb == np.float32(1.0 / 14.0)
    assert a.dtype == np.float32 and rb.dtype == np.float32


# ------------------------------------------------------------ S6
def test_divergence_linear_field_and_obstacle():
    H, W = 6, 7
    X, Y = np.meshgrid(np.arange(W, dtype=float), np.arange(H, dtype=float))
    u = np.stack([0.2 * X, -0.1 * Y], -1)
    d = F.divergence(u, np.zeros((H, W, 2)))
    np.testing.assert_allclose(d[1:-1, 1:-1], 0.1, atol=1e-12)
    obs = np.zeros((H, W, 2))
    obs[3, 4, 0] = 1.0  # right neighbour of (3,3) is solid
    d = F.divergence(u, obs)
    # cell (x=3, y=3): x2 := 0, x1 = 0.2*2 ; y terms -0.1*(4-2)
    assert abs(d[3, 3] - 0.5 * ((0 - 0.4) + (-0.4 + 0.2))) < 1e-12


# ------------------------------------------------------------ S8 / e6
def test_jacobi_hand_values():
    p = np.arange(9, dtype=f32).reshape(3, 3)
    b = np.full((3, 3), 1.0, f32)
    obs = _obs(3, 3)
    obs[1, 2, 0] = 1.0  # cell (x=2, y=1) solid
    q = F.jacobi(p, b, obs)
    # centre (1,1)=4: L=3, R solid->4, B=1, T=7 -> (3+4+1+7-1)/4 = 3.5
    assert q[1, 1] == f32(3.5)
    # corner (0,0)=0: L clamp 0, R 1, B clamp 0, T 3 -> (0+1+0+3-1)/4 = 0.75
    assert q[0, 0] == f32(0.75)
    # solid cell itself is swept too: (2,1)=5: L=4, R clamp 5, B 2, T 8 -> 4.5
    assert q[1, 2] == f32(4.5)


def test_jacobi_sweep_matches_independent_sparse_operator(rng):
    H, W = 9, 12
    solid = rng.random((H, W)) < 0.2
    obs = np.zeros((H, W, 2))
    obs[solid, 0] = 1
    A = P.assemble(solid)
    p = rng.standard_normal((H, W))
    b = rng.standard_normal((H, W))
    q = F.jacobi(p, b, obs)
    # Jacobi: p' = (A p + 4 p - b) / 4 ; fluid rows only (solid rows of the
    # assembled matrix drop solid columns, while the sweep reads them)
    ref = ((A @ p.ravel()) + 4 * p.ravel() - b.ravel()) / 4
    f = ~solid.ravel()
    np.testing.assert_allclose(q.ravel()[f], ref[f], atol=1e-12)


def test_jacobi_fixed_point_is_sparse_solution(rng):
    H, W = 8, 8
    solid = np.zeros((H, W), bool)
    solid[3:5, 3:5] = True
    obs = np.zeros((H, W, 2))
    obs[solid, 0] = 1
    b = rng.standard_normal((H, W))
    b[solid] = 0
    b[~solid] -= b[~solid].mean()
    p_ref = P.solve_fp64(b, solid)
    p = F.jacobi(np.zeros((H, W)), b, obs, 6000)
    pf = p[~solid] - p[~solid].mean()
    np.testing.assert_allclose(pf, p_ref[~solid], atol=1e-9)
    assert F.relative_residual(p_ref, b, obs) < 1e-12


def test_jacobi_zero_rhs_stays_zero():
    z = np.zeros((5, 5), f32)
    assert np.all(F.jacobi(z, z, _obs(5, 5), 7) == 0)


def test_sparse_operator_null_space(rng):
    solid = rng.random((7, 6)) < 0.25
    A = P.assemble(solid)
    f = ~solid.ravel()
    ones = np.ones(A.shape[0])
    r = A[f][:, f] @ ones[f]
    assert np.all(r == 0)  # A 1 = 0 on fluid rows
    Af = A[f][:, f]
    assert (abs(Af - Af.T)).max() == 0  # symmetric on fluid cells


# ------------------------------------------------------------ S9
def test_subtract_gradient_ramp_and_shift_invariance(rng):
    H, W = 6, 8
    X = np.broadcast_to(np.arange(W, dtype=float)[None, :], (H, W))
    p = 0.3 * X
    u = rng.uniform(-1, 1, (H, W, 2))
    v = F.subtract_gradient(u, p, np.zeros((H, W, 2)))
    np.testing.assert_allclose(v[:, 1:-1, 0], u[:, 1:-1, 0] - 0.3, atol=1e-12)
    np.testing.assert_allclose(v[..., 1], u[..., 1], atol=0)
    # constant shift of p leaves velocity unchanged (exact in float32 for
    # small integers)
    pf = rng.integers(-50, 50, (H, W)).astype(f32)
    uf = u.astype(f32)
    a = F.subtract_gradient(uf, pf, _obs(H, W))
    b = F.subtract_gradient(uf, pf + f32(64), _obs(H, W))
    assert np.array_equal(a, b)


def test_subtract_gradient_solid_neighbour():
    p = np.arange(9, dtype=float).reshape(3, 3)
    obs = np.zeros((3, 3, 2))
    obs[1, 2, 0] = 1
    v = F.subtract_gradient(np.zeros((3, 3, 2)), p, obs)
    # (1,1): x2 solid -> 4; x1 = 3 -> u.x = -0.5*(4-3)
    assert v[1, 1, 0] == -0.5
    assert v[1, 1, 1] == -0.5 * (7 - 1)


# ------------------------------------------------------------ F1..F4
def test_add_velocity_cone_and_clamp():
    H, W = 16, 16
    u = np.zeros((H, W, 2), f32)
    u[0, 0] = (3.0, -3.0)  # far away: clamped anyway
    v = F.add_velocity(u, (0.5, 0.5), (0.8, -0.4), 4.0)
    # centre (8, 8): distance 0 -> full value
    assert v[8, 8, 0] == f32(0.8) and v[8, 8, 1] == f32(-0.4)
    # distance 2 -> half value
    assert v[8, 10, 0] == f32(0.4)
    # distance 4 (= r) -> + 0 ; distance 5 -> untouched
    assert v[8, 12, 0] == 0 and v[8, 13, 0] == 0
    assert tuple(v[0, 0]) == (1.0, -1.0)


def test_circle_obstacle_matches_bruteforce():
    H, W = 20, 30
    obs = F.add_circle_obstacle(_obs(H, W), (0.37, 0.61), 5.3)
    cx, cy = 0.37 * W, 0.61 * H
    for y in range(H):
        for x in range(W):
            inside = math.hypot(cx - x, cy - y) <= 5.3
            assert bool(obs[y, x, 0]) == inside and obs[y, x, 1] == 0


def test_triangle_obstacle_bruteforce_and_static_channel():
    H, W = 24, 24
    p1, p2, p3 = (0.1, 0.1), (0.8, 0.2), (0.3, 0.9)
    dyn = F.add_triangle_obstacle(_obs(H, W), p1, p2, p3, static=False)
    sta = F.add_triangle_obstacle(_obs(H, W), p1, p2, p3, static=True)

    def inside(px, py):
        # barycentric test in float64 (points away from the edges)
        (x1, y1), (x2, y2), (x3, y3) = p1, p2, p3
        den = (y2 - y3) * (x1 - x3) + (x3 - x2) * (y1 - y3)
        a = ((y2 - y3) * (px - x3) + (x3 - x2) * (py - y3)) / den
        b = ((y3 - y1) * (px - x3) + (x1 - x3) * (py - y3)) / den
        return a, b, 1 - a - b

    checked = 0
    for y in range(H):
        for x in range(W):
            a, b, c = inside(x / W, y / H)
            if min(abs(a), abs(b), abs(c)) < 1e-3:
                continue
            ins = a > 0 and b > 0 and c > 0
            assert bool(dyn[y, x, 0]) == ins and dyn[y, x, 1] == 0
            assert bool(sta[y, x, 1]) == ins and sta[y, x, 0] == 0
            checked += 1
    assert checked > 500


def test_add_dye_clamp_inside_disk_only():
    q = np.full((10, 10), 300.0, f32)
    q[5, 5] = 250.0
    r = F.add_dye(q, (0.5, 0.5), 10.0, 3.0)
    assert r[5, 5] == 255.0          # clamped inside the disk
    assert r[0, 0] == 300.0          # outside the disk: untouched, no clamp
    assert r[5, 6] == 255.0


def test_advect_dye_integer_translation(rng):
    # dye grid == velocity grid, velocity (1, -1) * dt*speed=2 -> shift by 2
    H, W = 12, 10
    q = rng.uniform(0, 1, (H, W)).astype(f32)
    u = np.zeros((H, W, 2), f32)
    u[..., 0] = 1.0
    u[..., 1] = -1.0
    out = F.advect_dye(q, u, _obs(H, W), 1.0, 2.0, 1.0)
    assert np.array_equal(out[:H - 2, 2:], q[2:, :W - 2])
    out = F.advect_dye(q, u, _obs(H, W), 1.0, 2.0, 0.5)
    assert np.array_equal(out[:H - 2, 2:], q[2:, :W - 2] * f32(0.5))


def test_advect_dye_resampled_velocity_scaled():
    # dye grid twice the velocity grid: uniform velocity (0.25, 0) on V is
    # scaled by P/V = 2 -> 0.5 cells per unit dt*speed on the dye grid.
    u = np.zeros((4, 4, 2), f32)
    u[..., 0] = 0.25
    q = np.broadcast_to(np.arange(8, dtype=f32)[None, :], (8, 8)).copy()
    out = F.advect_dye(q, u, _obs(4, 4), 1.0, 2.0, 1.0)
    # back-trace by 0.25*2*2 = 1 dye cell: q(x-1) = x-1
    np.testing.assert_array_equal(out[:, 1:], q[:, :-1])


def test_advect_dye_obstacle_lookup_truncates():
    u = np.zeros((4, 4, 2), f32)
    obs = _obs(4, 4)
    obs[1, 2, 0] = 1.0
    q = np.ones((8, 8), f32)
    out = F.advect_dye(q, u, obs, 1.0, 1.0, 1.0)
    # dye cells (4..5, 2..3) map to velocity cell (2, 1)
    assert np.all(out[2:4, 4:6] == 0)
    assert out.sum() == 64 - 4


# ------------------------------------------------------------ extensions
def test_gaussian_weight_closed_forms():
    W = H = 64
    w = F.gaussian_weight(W, H, 0.5, 0.5, 4.0)
    assert w[32, 32] == 1.0
    assert abs(float(w[32, 36]) - math.exp(-1.0)) < 1e-7
    assert w[32, 28] == w[32, 36] == w[36, 32] == w[28, 32]


def test_gaussian_splat_amplitude():
    u = np.zeros((32, 32, 2), f32)
    q = np.zeros((32, 32), f32)
    u2, q2 = F.gaussian_splat(u, q, [(0.5, 0.5, 2.0, -1.0, 3.0, 1.0)])
    assert tuple(u2[16, 16]) == (2.0, -1.0) and q2[16, 16] == 1.0  # no clamp


def test_free_slip_preserves_tangential_flow():
    u = np.zeros((6, 8, 2), f32)
    u[..., 0] = 0.4  # flow along the bottom/top walls
    v = F.enforce_boundaries(u, F.BC_FREE_SLIP)
    assert np.all(v[0, 1:-1, 0] == f32(0.4)) and np.all(v[-1, 1:-1, 0] == f32(0.4))
    assert np.all(v[0, :, 1] == 0) and np.all(v[:, 0, 0] == 0)
    assert np.all(v[1:-1, 1:-1] == u[1:-1, 1:-1])


def test_inflow_column():
    u = np.full((6, 8, 2), 0.3, f32)
    v = F.enforce_boundaries(u, F.BC_INFLOW, inflow=1.0)
    assert np.all(v[:, 0, 0] == 1.0) and np.all(v[:, 0, 1] == 0.0)
    assert np.all(v[1:-1, -1] == v[1:-1, -2])


def test_reference_mode_without_borders_is_identity(rng):
    u = rng.uniform(-1, 1, (5, 5, 2)).astype(f32)
    assert np.array_equal(F.enforce_boundaries(u, F.BC_REFERENCE, False), u)


def test_relative_residual_closed_forms(rng):
    H, W = 10, 9
    solid = np.zeros((H, W), bool)
    solid[4:6, 4:6] = True
    obs = np.zeros((H, W, 2))
    obs[solid, 0] = 1
    b = rng.standard_normal((H, W))
    b[solid] = 0
    assert abs(F.relative_residual(np.zeros((H, W)), b, obs) - 1.0) < 1e-12
    p = P.solve_fp64(b, solid)
    assert F.relative_residual(p, b, obs) < 1e-12
    assert F.relative_residual(p + 123.0, b, obs) < 1e-11
    # a mean shift of b is invisible to the residual
    assert F.relative_residual(p, b + 5.0 * (~solid), obs) < 1e-11


def test_step_order_reference_sequence(rng):
    # the composite step equals the stages applied in the order of
    # fluid_simulator.py:180-280 (viscosity on, reference walls)
    H, W = 12, 14
    u = rng.uniform(-1, 1, (H, W, 2)).astype(f32)
    obs = F.add_circle_obstacle(_obs(H, W), (0.5, 0.5), 2.5)
    prm = F.Params(vorticity=0.7, viscosity=0.5, iterations=5)
    dt = 1 / 60
    un, pn, on, _ = F.step(u, np.zeros((H, W), f32), obs, dt, prm)
    v = F.init_boundaries(u)
    v = F.advect_velocity(v, obs, dt, 500.0, 1.0)
    v = F.apply_vorticity(v, F.calc_vorticity(v), dt, 0.7)
    a, rb = F.viscosity_coefficients(0.5)
    v = F.viscosity(v, a, rb)
    b = F.divergence(v, obs)
    p = F.jacobi(np.zeros((H, W), f32), b, obs, 5)
    v = F.subtract_gradient(v, p, obs)
    assert np.array_equal(un, v) and np.array_equal(pn, p)
    assert np.all(on == 0)


def test_one_pass_sums_agree_with_two_pass(rng):
    H, W = 16, 16
    obs = np.zeros((H, W, 2))
    obs[5:9, 6:8, 0] = 1
    b = rng.standard_normal((H, W)).astype(f32)
    p = F.jacobi(np.zeros((H, W), f32), b, obs.astype(f32), 30)
    one = F.relative_residual_from_sums(*F.residual_sums(p, b, obs))
    two = F.relative_residual(p, b, obs)
    assert abs(one - two) < 1e-9 * max(1.0, two)


def test_dye_rgba8_unorm_conversion():
    """rgba8 image store of a float: clamp to [0, 1], x255, round to nearest
    (ties up); all four channels equal (demo.ComputeShader.comp:18)."""
    q = np.array([[0.0, 1.0, 0.5, -3.0], [2.0, 255.0, 1.5 / 255, np.nan]], np.float32)
    got = F.dye_rgba8(q)
    assert got.shape == (2, 4, 4) and got.dtype == np.uint8
    want = np.array([[0, 255, 128, 0], [255, 255, 2, 0]], np.uint8)
    assert np.array_equal(got[..., 0], want)
    assert all(np.array_equal(got[..., c], got[..., 0]) for c in range(4))
    # every representable 8-bit level round-trips
    lv = (np.arange(256, dtype=np.float32) / np.float32(255.0))
    assert np.array_equal(F.dye_rgba8(lv)[..., 0], np.arange(256, dtype=np.uint8))


def test_quiver_statement_closed_forms():
    """The quiver statement (demo.QuiverFragmentShader.frag) on an 8 x 8
    velocity grid under a 128 x 128 window with 16-pixel tiles: tile centres
    sit at 8, 24, .., 120, so tile (i, j) samples fPos = (7.5 - i, 7.5 - j)
    (the window flip), and the arrow of velocity u is v = -u * 16 * 16 *
    0.4 pixels (the -WindowSize / VelocitySize scale), clamped to 8."""
    f = np.float32
    u = np.zeros((8, 8, 2), f)
    img = F.quiver_rgba8(u, 128, 128, 16.0)
    assert img.shape == (128, 128, 4)
    assert (img[..., :3] == 255).all() and (img[..., 3] == 0).all()  # no arrow

    # uniform +x flow: v = (-5.12, 0), tip (with the head) on the -x side.
    # Pixel offsets from the tile centre are odd multiples of 0.5.
    u[..., 0] = f(0.05)
    a = F.quiver_rgba8(u, 128, 128, 16.0)[..., 3].reshape(8, 16, 8, 16)
    # rows 0.5 from the axis on the tail side: shaft distance max(0.5,
    # |rx| - 5.12) = 0.5 -> alpha 0.5 -> 128 for rx = 0.5 .. 4.5
    assert (a[:, 7:9, :, 8:13] == 128).all()
    # rows 1.5 from the axis: the shaft is 1.5 away, only the head strokes
    # (from the tip to (-2.048, +-1.024)) can ink them, on the tip's side
    assert (a[:, [6, 9], :, 8:] == 0).all()
    assert (a[:, [6, 9], :, :8] > 0).any()

    # the y component is the bottom-left sample's (mix with (t, 0)): rows of
    # the grid alternate 0 and 5, tile row i samples grid row 7 - i exactly
    # at fy = 7.5 - i, so only tile rows with 7 - i odd carry arrows
    u[:] = 0
    u[1::2, :, 1] = f(5.0)
    a = F.quiver_rgba8(u, 128, 128, 16.0)[..., 3].reshape(8, 16, 8, 16)
    inked = a.reshape(8, -1).max(axis=1) > 0
    assert list(inked) == [(7 - i) % 2 == 1 for i in range(8)]

    # magnitudes beyond the clamp (tile / 2) give the same arrow, whose shaft
    # then spans the tile: alpha 128 at 0.5 from the axis, |rx| <= 7.5
    u[:] = 0
    u[..., 0] = f(50.0)
    big = F.quiver_rgba8(u, 128, 128, 16.0)
    u[..., 0] = f(500.0)
    assert np.array_equal(big, F.quiver_rgba8(u, 128, 128, 16.0))
    a = big[..., 3].reshape(8, 16, 8, 16)
    assert (a[:, 7:9, :, 9:] == 128).all()
