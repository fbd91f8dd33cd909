"""Pins the float64 smoother statements used by the GPU tests on closed
forms.  Without obstacles, L is the Neumann 5-point Laplacian whose
eigenvectors are v_mn(x, y) = cos(pi m (x + 1/2) / W) cos(pi n (y + 1/2) / H)
with -L v = lam v, lam = 4 - 2 cos(pi m / W) - 2 cos(pi n / H).  Starting
from x* + v on the consistent problem r = L x*, k Jacobi sweeps leave the
error (1 - lam / 4)^k v and k Chebyshev iterations on [lmin, 8] leave
T_k((theta - lam) / delta) / T_k(theta / delta) v."""
import numpy as np
import pytest

from smoothers import cheb_ref, jacobi_ref, lap


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
