"""Training pairs for the learned pressure guess, produced by the project's
own solvers: FP64 conjugate gradients on the GPU, one workgroup per problem
(nx_fp64_solve_batch, make_dataset), converged to a relative residual of
1e-6 and stored as the float32 rounding of the FP64 solution; or, as the
first weights were made, the batched float32 Jacobi solver run to its
round-off floor (make_dataset_jacobi).

Each sample is a coarse Poisson problem (the grid the U-Net runs on): a
right-hand side b, a solid mask and the pressure p the solver converged to.
The right-hand sides are smooth random fields (sums of Gaussian blobs of
random width and sign plus low Fourier modes), like the restricted
divergence and residual fields the solver hands the network; masks are
disks (the cylinder scenes), rectangles, or empty.  b is centred on the
fluid mean of every connected fluid region (the operator's null space)
before solving, so every problem has a solution.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

from ..batch import BatchedScenes, random_rect_obstacles


def random_problem(rng, n):
    X = np.arange(n, dtype=np.float64)[None, :]
    Y = np.arange(n, dtype=np.float64)[:, None]
    kind = rng.random()
    solid = np.zeros((n, n), bool)
    if kind < 0.5:
        c = rng.uniform(0.2, 0.8, 2) * n
        r = rng.uniform(0.05, 0.2) * n
        solid = (X - c[0]) ** 2 + (Y - c[1]) ** 2 <= r * r
    elif kind < 0.75:
        solid = random_rect_obstacles(rng, n, n, rng.uniform(0.0, 0.15))
    b = np.zeros((n, n))
    for _ in range(int(rng.integers(2, 13))):
        c = rng.uniform(0, n, 2)
        s = np.exp(rng.uniform(np.log(1.5), np.log(n / 4)))
        b += rng.normal() * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * s * s))
    if rng.random() < 0.5:
        for _ in range(3):
            kx, ky = rng.integers(1, 6, 2)
            ph = rng.uniform(0, 2 * np.pi, 2)
            b += 0.3 * rng.normal() * np.cos(np.pi * kx * (X + 0.5) / n + ph[0]) * \
                np.cos(np.pi * ky * (Y + 0.5) / n + ph[1])
    b[solid] = 0.0
    fluid = ~solid
    # centred on every connected fluid region (4-connected, the stencil's
    # neighbourhood): obstacles can cut the domain in pieces, and the
    # operator's null space then holds one constant per piece, so a right-hand
    # side centred only on the global mean has no solution
    lab, count = ndimage.label(fluid)
    if count > 0:
        sums = ndimage.sum_labels(b, lab, np.arange(1, count + 1))
        cells = ndimage.sum_labels(fluid, lab, np.arange(1, count + 1))
        means = np.concatenate([[0.0], sums / cells])
        b = b - means[lab]
    b /= max(np.sqrt((b[fluid] ** 2).mean()), 1e-12)
    return b.astype(np.float32), solid


def make_dataset(count, n=128, seed=500000, tol=1e-6, max_iter=20000, per_batch=2048,
                 log=print):
    """Returns (b [count,n,n] f32, solid [count,n,n] bool, p [count,n,n] f32,
    residual [count]): problem i from seed + i, p the float32 rounding of
    its FP64 conjugate-gradient solution (fluid mean removed, 0 on solid
    cells), residual the FP64 solution's true relative residual."""
    import ctypes as C

    import torch

    from .. import _lib
    bs, ss, ps, rs = [], [], [], []
    done = 0
    while done < count:
        m = min(per_batch, count - done)
        probs = [random_problem(np.random.default_rng(seed + done + i), n) for i in range(m)]
        b = torch.from_numpy(np.stack([q[0] for q in probs])).cuda()
        mask = torch.from_numpy(np.stack([q[1] for q in probs]).astype(np.uint8)).cuda()
        work = torch.empty(m * n * n, dtype=torch.float64, device="cuda")
        xf = torch.empty(m, n, n, dtype=torch.float32, device="cuda")
        res = torch.empty(m, dtype=torch.float64, device="cuda")
        it = torch.empty(m, dtype=torch.int32, device="cuda")
        p = lambda t: C.c_void_p(t.data_ptr())
        _lib.call("nx_fp64_solve_batch", p(b), p(mask), n, m, p(work), p(xf), float(tol),
                  int(max_iter), p(res), p(it),
                  C.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        r = res.cpu().numpy()
        log(f"  data batch {done // per_batch}: FP64 CG residual median {np.median(r):.2e} "
            f"max {r.max():.2e}, iterations median {int(it.median())} max {int(it.max())}")
        bs.append(b.cpu().numpy())
        ss.append(np.stack([q[1] for q in probs]))
        ps.append(xf.cpu().numpy())
        rs.append(r)
        done += m
    return np.concatenate(bs), np.concatenate(ss), np.concatenate(ps), np.concatenate(rs)


def make_dataset_jacobi(count, n=128, seed=500000, tol=1e-6, max_sweeps=400000,
                        chunk_sweeps=20000, per_batch=1024, tile=None, log=print):
    """Returns (b [count,n,n] f32, solid [count,n,n] bool, p [count,n,n] f32,
    residual [count]) with p from Jacobi from zero on the GPU.

    Iteration stops when every scene is at `tol`, at `max_sweeps`, or when a
    chunk no longer lowers the median residual by 5% (the float32 round-off
    floor of the larger pressures); the reached residuals are returned."""
    bs, ss, ps, rs = [], [], [], []
    done = 0
    while done < count:
        m = min(per_batch, count - done)
        rng = [np.random.default_rng(seed + done + i) for i in range(m)]
        probs = [random_problem(r, n) for r in rng]
        batch = BatchedScenes([n] * m)
        if tile is not None:
            batch.set_params(jacobi_tile=tile)
        k, rows, waves, cols = batch.jacobi_tile
        for i, (b, solid) in enumerate(probs):
            batch.upload(i, "divergence", b)
            batch.upload(i, "pressure", np.zeros_like(b))
            batch.set_obstacle_map(i, solid)
        sweeps = 0
        res = batch.residuals()
        while sweeps < max_sweeps and res.max() > tol:
            batch.jacobi_blocked(chunk_sweeps, k, rows, waves, cols)
            sweeps += chunk_sweeps
            prev = np.median(res)
            res = batch.residuals()
            log(f"  data batch {done // per_batch}: {sweeps} sweeps, residual "
                f"median {np.median(res):.2e} max {res.max():.2e}")
            if np.median(res) > 0.95 * prev:
                break
        for i, (b, solid) in enumerate(probs):
            bs.append(b)
            ss.append(solid)
            ps.append(batch.download(i, "pressure"))
        rs.append(res)
        batch.destroy()
        done += m
    return np.stack(bs), np.stack(ss), np.stack(ps), np.concatenate(rs)
