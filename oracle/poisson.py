"""FP64 oracle for the pressure Poisson system.

TEST INFRASTRUCTURE ONLY (see oracle/fluid.py): imported by tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline, never by the
product path.

The matrix is assembled cell by cell from the stencil of
natrix/core/shaders/originals/shader.Poisson.comp:26-37 with an explicit loop
that is independent of the vectorised gathers in ``oracle.fluid``:

    (A p)_i = sum_{X in L,R,B,T} (p_X if X is an in-grid fluid cell else p_i) - 4 p_i

Out-of-grid neighbours clamp to the cell itself (common.sh:14-17) and solid
neighbours are replaced by the centre (Poisson.comp:32-35), so both drop out
of the row.  The Jacobi fixed point ``p = (sum nbr - b)/4`` is ``A p = b``.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csgraph
import scipy.sparse.linalg as spla


def assemble(solid):
    """Return the (n, n) CSR matrix of A over all cells (row-major).

    Solid rows use the same stencil (the reference sweeps them too); fluid
    rows never reference solid columns."""
    H, W = solid.shape
    rows, cols, vals = [], [], []
    for y in range(H):
        for x in range(W):
            i = y * W + x
            diag = -4.0
            for (nx, ny) in ((x - 1, y), (x + 1, y), (x, y - 1), (x, y + 1)):
                if 0 <= nx < W and 0 <= ny < H and not solid[ny, nx]:
                    rows.append(i)
                    cols.append(ny * W + nx)
                    vals.append(1.0)
                else:
                    diag += 1.0
            rows.append(i)
            cols.append(i)
            vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(H * W, H * W))


def solve_fp64(b, solid):
    """Mean-free FP64 solution of A p = b on the fluid cells.

    Per connected fluid component the right-hand side's mean is removed
    (the operator's null space is the constants on each component) and one
    cell is pinned to 0; the returned p has zero mean per component and is
    0 on solid cells."""
    H, W = solid.shape
    A = assemble(solid)
    fluid = ~solid.ravel()
    idx = np.flatnonzero(fluid)
    Aff = A[idx][:, idx].tocsr()
    bf = np.asarray(b, np.float64).ravel()[idx].copy()
    adj = (Aff - sp.diags(Aff.diagonal())).tocsr()
    ncomp, labels = csgraph.connected_components(adj, directed=False)
    p = np.zeros(len(idx))
    for c in range(ncomp):
        m = np.flatnonzero(labels == c)
        bc = bf[m] - bf[m].mean()
        if len(m) == 1:
            continue
        Ac = Aff[m][:, m].tocsc()
        keep = np.arange(1, len(m))
        sol = spla.spsolve(Ac[keep][:, keep], bc[keep])
        pc = np.concatenate([[0.0], sol])
        p[m] = pc - pc.mean()
    out = np.zeros(H * W)
    out[idx] = p
    return out.reshape(H, W)
