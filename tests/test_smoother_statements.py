"""Pins the float64 smoother statements used by the GPU tests on closed
forms.  Without obstacles, L is the Neumann 5-point Laplacian whose
eigenvectors are v_mn(x, y) = cos(pi m (x + 1/2) / W) cos(pi n (y + 1/2) / H)
with -L v = lam v, lam = 4 - 2 cos(pi m / W) - 2 cos(pi n / H).  Starting
from x* + v on the consistent problem r = L x*, k Jacobi sweeps leave the
error (1 - lam / 4)^k v and k Chebyshev iterations on [lmin, 8] leave
T_k((theta - lam) / delta) / T_k(theta / delta) v."""
import numpy as np
import pytest

from fractions import Fraction

from smoothers import cheb_ref, fma32, jacobi_ref, lap


def mode(W, H, m, n):
    x = np.arange(W) + 0.5
    y = np.arange(H) + 0.5
    return np.cos(np.pi * m * y[:, None] / H) * np.cos(np.pi * n * x[None, :] / W), \
        4 - 2 * np.cos(np.pi * n / W) - 2 * np.cos(np.pi * m / H)


def cheb_T(k, x):
    return np.cosh(k * np.arccosh(x)) if abs(x) >= 1 else np.cos(k * np.arccos(x))


@pytest.mark.parametrize("m,n", [(0, 1), (3, 5), (17, 30), (23, 31)])
def test_eigenmodes_damp_as_the_closed_forms(rng, m, n):
    H, W = 24, 32
    solid = np.zeros((H, W), bool)
    v, lam = mode(W, H, m, n)
    assert np.allclose(-lap(v, solid), lam * v, atol=1e-12)
    xs = rng.normal(size=(H, W))
    r = lap(xs, solid)
    k = 9
    got = jacobi_ref(xs + v, r, solid, k) - xs
    assert np.allclose(got, (1 - lam / 4) ** k * v, atol=1e-10)
    lmin = 0.3
    theta, delta = 0.5 * (8 + lmin), 0.5 * (8 - lmin)
    got = cheb_ref(xs + v, r, solid, k, lmin) - xs
    want = cheb_T(k, (theta - lam) / delta) / cheb_T(k, theta / delta) * v
    assert np.allclose(got, want, atol=1e-10)


def test_solid_neighbour_reads_the_centre():
    v = np.arange(12.0).reshape(3, 4)
    solid = np.zeros((3, 4), bool)
    solid[1, 2] = True
    got = lap(v, solid)
    # cell (1, 1): right neighbour (1, 2) is solid -> reads v[1, 1]
    assert got[1, 1] == (v[1, 0] + v[1, 1] + v[0, 1] + v[2, 1]) - 4 * v[1, 1]
    # corner (0, 0): clamped left and up neighbours are itself
    assert got[0, 0] == (v[0, 0] + v[0, 1] + v[0, 0] + v[1, 0]) - 4 * v[0, 0]


def test_restriction_of_a_blockwise_constant_residual():
    from smoothers import restrict_ref
    f, Hc, Wc = 4, 3, 5
    c = np.arange(Hc * Wc, dtype=np.float64).reshape(Hc, Wc) - 7.0
    b = np.kron(c, np.ones((f, f)))      # r = b for p = 0
    solid = np.zeros_like(b, bool)
    solid[0:4, 0:4] = True               # block (0, 0) entirely solid
    solid[5, 6] = True                   # one solid cell in block (1, 1)
    rc, mc = restrict_ref(np.zeros_like(b), b, solid, f)
    want = f * f * c
    want[0, 0] = 0.0
    assert np.allclose(rc, want)
    assert mc[0, 0] and mc.sum() == 1


def test_prolongation_reproduces_constant_and_linear_fields():
    from smoothers import prolong_ref
    f, Hc, Wc = 4, 6, 7
    H, W = Hc * f, Wc * f
    mc = np.zeros((Hc, Wc), bool)
    solid = np.zeros((H, W), bool)
    out = prolong_ref(np.full((Hc, Wc), 2.5), mc, f, H, W, solid)
    assert np.allclose(out, 2.5)
    Y, X = np.mgrid[0:Hc, 0:Wc]
    out = prolong_ref(0.3 * X - 1.7 * Y, mc, f, H, W, solid)
    y, x = np.mgrid[0:H, 0:W]
    lin = 0.3 * (x - 0.5 * (f - 1)) / f - 1.7 * (y - 0.5 * (f - 1)) / f
    inner = (x >= (f - 1) / 2) & (x <= f * (Wc - 1) + (f - 1) / 2) & \
        (y >= (f - 1) / 2) & (y <= f * (Hc - 1) + (f - 1) / 2)
    assert np.allclose(out[inner], lin[inner])
    # a solid coarse cell carries no weight: the constant survives next to it
    mc[2, 3] = True
    out = prolong_ref(np.where(mc, 99.0, 2.5), mc, f, H, W, solid)
    assert np.allclose(out, 2.5)


def round_f32(q):
    """A rational rounded to float32, to nearest with ties to even."""
    if q == 0:
        return np.float32(0.0)
    neg = q < 0
    q = -q if neg else q
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1
    if Fraction(2) ** (e + 1) <= q:
        e += 1
    e = max(e, -126)  # subnormals share the smallest exponent
    scale = Fraction(2) ** (23 - e)
    m = q * scale
    n = m.numerator // m.denominator
    rem = m - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    v = np.float32(float(Fraction(n) / scale))
    return -v if neg else v


def test_fma32_is_the_correctly_rounded_fused_multiply_add(rng):
    """fma32 against exact rational arithmetic: random operands over many
    binades, near-cancellations a b ~ -c, subnormal results, and products
    that land exactly halfway between two float32 values, where a tiny c
    decides the rounding (float64 evaluation rounds those twice and errs)."""
    a = (rng.standard_normal(3000) * 2.0 ** rng.integers(-30, 30, 3000)).astype(np.float32)
    b = (rng.standard_normal(3000) * 2.0 ** rng.integers(-30, 30, 3000)).astype(np.float32)
    c = (rng.standard_normal(3000) * 2.0 ** rng.integers(-60, 60, 3000)).astype(np.float32)
    c[:500] = (-(a[:500].astype(np.float64) * b[:500]) * (1 + 1e-7 * rng.standard_normal(500))).astype(np.float32)
    a[500:600] = np.float32(2.0 ** -70)
    b[500:600] = (rng.standard_normal(100) * 2.0 ** -70).astype(np.float32)
    u = np.float32(1 + 2.0 ** -12)
    ties = [(u, u, np.float32(s * 2.0 ** -80)) for s in (1, -1)] + \
        [(u, u, np.float32(0.0)), (np.float32(1 + 2.0 ** -12), np.float32(-(1 + 2.0 ** -12)),
                                   np.float32(1.0))]
    a = np.concatenate([a, [t[0] for t in ties]]).astype(np.float32)
    b = np.concatenate([b, [t[1] for t in ties]]).astype(np.float32)
    c = np.concatenate([c, [t[2] for t in ties]]).astype(np.float32)
    # the rounding reference itself: numpy's float64 -> float32 conversion
    # is correctly rounded
    xs = rng.standard_normal(2000) * 2.0 ** rng.integers(-140, 100, 2000)
    assert np.array_equal(np.array([round_f32(Fraction(float(x))) for x in xs], np.float32),
                          xs.astype(np.float32))
    got = fma32(a, b, c)
    want = np.array([round_f32(Fraction(float(x)) * Fraction(float(y)) + Fraction(float(z)))
                     for x, y, z in zip(a, b, c)], np.float32)
    assert np.array_equal(got, want)
    # the tie cases are ones where rounding the float64 sum directly errs
    naive = (a[-4:-2].astype(np.float64) * b[-4:-2] + c[-4:-2]).astype(np.float32)
    assert not np.array_equal(naive, want[-4:-2])


def test_faces_of_a_block_boundary_with_a_gap():
    """f = 4 on an 8 x 12 grid (2 x 3 coarse cells) with a wall in fine
    column 4 (the first column of block X = 1) open at rows 1 and 2."""
    from smoothers import restrict_faces_ref
    f = 4
    solid = np.zeros((8, 12), bool)
    solid[:, 4] = True
    solid[1:3, 4] = False
    b = np.arange(96, dtype=np.float64).reshape(8, 12)
    rc, mc, fe, fs = restrict_faces_ref(np.zeros_like(b), b, solid, f)
    # east faces: X = 0 | 1 crosses the wall (rows 1, 2 open in block Y = 0,
    # none in Y = 1); X = 1 | 2 is open; X = 2 is the grid's edge
    assert np.array_equal(fe, np.array([[0.5, 1.0, 0.0], [0.0, 1.0, 0.0]], np.float32))
    # south faces: rows 3 | 4, the wall's column shut in block X = 1
    assert np.array_equal(fs, np.array([[1.0, 0.75, 1.0], [0.0, 0.0, 0.0]], np.float32))
    # block sums over fluid cells (p = 0: r = b on them)
    want = np.where(solid, 0.0, b).reshape(2, 4, 3, 4).sum(axis=(1, 3))
    assert np.allclose(rc, want) and not mc.any()


def test_weighted_operator_reduces_to_the_plain_one(rng):
    """Faces from a fine mask made of whole solid blocks are 0 toward solid
    cells and 1 elsewhere: on the fluid cells the weighted float32 operator
    is then the plain one, bitwise."""
    from smoothers import lapc32, lapcw32, restrict_faces_ref
    f, Hc, Wc = 4, 9, 11
    mc = rng.random((Hc, Wc)) < 0.2
    solid = np.kron(mc, np.ones((f, f), bool))
    _, mc2, fe, fs = restrict_faces_ref(np.zeros(solid.shape), np.zeros(solid.shape), solid, f)
    assert np.array_equal(mc2, mc)
    assert set(np.unique(np.concatenate([fe.ravel(), fs.ravel()]))) <= {0.0, 1.0}
    v = rng.normal(size=(Hc, Wc)).astype(np.float32)
    got, want = lapcw32(v, fe, fs), lapc32(v, mc)
    assert np.array_equal(got[~mc], want[~mc])


def test_weighted_operator_is_symmetric(rng):
    """sum over faces of w (v_J - v_I) is a symmetric operator: u . L v =
    v . L u for partial weights too (to float32 round-off), which a face
    read from the wrong cell would break by order one."""
    from smoothers import lapcw32
    Hc, Wc = 13, 17
    fe = rng.choice([0.0, 0.25, 0.5, 1.0], size=(Hc, Wc)).astype(np.float32)
    fs = rng.choice([0.0, 0.25, 0.5, 1.0], size=(Hc, Wc)).astype(np.float32)
    u = rng.normal(size=(Hc, Wc)).astype(np.float32)
    v = rng.normal(size=(Hc, Wc)).astype(np.float32)
    a = float(np.dot(u.ravel().astype(np.float64), lapcw32(v, fe, fs).ravel()))
    b = float(np.dot(v.ravel().astype(np.float64), lapcw32(u, fe, fs).ravel()))
    assert abs(a - b) <= 1e-5 * np.linalg.norm(u) * np.linalg.norm(v) * 8
    # and the constant is in its null space
    assert np.abs(lapcw32(np.full((Hc, Wc), 3.0, np.float32), fe, fs)).max() == 0.0
