"""Float64 statements of the learned solve's coarse smoothers, for the GPU
tests: the 5-point operator L of the Jacobi sweep (a solid or out-of-grid
neighbour reads the centre value; shader.Poisson.comp:26-37 with the clamp
of shader.common.sh:14-17), Jacobi sweeps e <- e + (L e - r) / 4, and the
Chebyshev iteration for -L e = -r on [lmin, 8] (three-term recurrence of
Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Algorithm
12.1)."""
import numpy as np


def lap(v, solid):
    H, W = v.shape
    P = np.pad(v, 1, mode="edge")
    S = np.pad(solid, 1, mode="constant", constant_values=True)
    acc = np.zeros_like(v)
    for dy, dx in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        nb = P[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        ns = S[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        acc = acc + np.where(ns, v, nb)
    return acc - 4.0 * v


def jacobi_ref(x, rhs, solid, iters):
    for _ in range(iters):
        x = x + (lap(x, solid) - rhs) * 0.25
    return x


def cheb_ref(x, rhs, solid, iters, lmin):
    theta, delta = 0.5 * (8 + lmin), 0.5 * (8 - lmin)
    sigma = theta / delta
    rho = 1 / sigma
    r = lap(x, solid) - rhs
    d = r / theta
    for _ in range(iters):
        ld = lap(d, solid)
        rho1 = 1 / (2 * sigma - rho)
        x = x + d
        r = r + ld
        d = rho1 * rho * d + 2 * rho1 / delta * r
        rho = rho1
    return x
