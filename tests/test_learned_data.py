"""Training problems of the learned guess (CPU): every right-hand side must
have a solution.

Obstacles can cut the 128^2 domain into several fluid regions; the
operator's null space then holds one constant per region, and a right-hand
side centred only on its global mean has no solution (FP64 conjugate
gradients diverge on it).  Seed 500469 is the first of the generator's
problems with more than one region."""
import numpy as np
from scipy import ndimage

from natrix_amd.learned.data import random_problem
from oracle.poisson import assemble, solve_fp64


def _relative_residual(b, solid, p):
    fluid = ~solid.ravel()
    r = (np.asarray(b, np.float64).ravel() - assemble(solid) @ p.ravel())[fluid]
    return np.linalg.norm(r) / np.linalg.norm(np.asarray(b, np.float64).ravel()[fluid])


def test_every_fluid_region_is_centred_and_the_problem_is_solvable():
    b, solid = random_problem(np.random.default_rng(500469), 128)
    lab, count = ndimage.label(~solid)
    assert count > 1
    rms = np.sqrt((b[~solid].astype(np.float64) ** 2).mean())
    for j in range(1, count + 1):
        assert abs(b[lab == j].astype(np.float64).mean()) <= 1e-6 * rms
    # the oracle solves each region with its own mean removed; with b already
    # centred per region, its solution satisfies A p = b for b as given
    p = solve_fp64(b, solid)
    assert _relative_residual(b, solid, p) <= 1e-6


def test_a_globally_centred_right_hand_side_has_no_solution():
    # the mistake the per-region centring prevents: the same problem with
    # only the global mean removed leaves a residual no solver can remove
    b, solid = random_problem(np.random.default_rng(500469), 128)
    lab, count = ndimage.label(~solid)
    g = b.astype(np.float64).copy()
    g[lab == 1] += 0.5
    g[~solid] -= g[~solid].mean()
    p = solve_fp64(g, solid)
    assert _relative_residual(g, solid, p) > 1e-2
