"""Float64 statements of the learned solve's grid transfers and coarse
smoothers, for the GPU tests: the 5-point operator L of the Jacobi sweep (a solid or out-of-grid
neighbour reads the centre value; shader.Poisson.comp:26-37 with the clamp
of shader.common.sh:14-17), Jacobi sweeps e <- e + (L e - r) / 4, and the
Chebyshev iteration for -L e = -r on [lmin, 8] (three-term recurrence of
Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Algorithm
12.1), the block-mean restriction of the residual and the bilinear
prolongation with weights renormalised over fluid coarse cells."""
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


def restrict_ref(p, b, solid, f):
    """(rc, mc): rc = f^2 x the mean of r = b - L p over the fluid cells of
    each f x f block (0 where the block has none), mc = 1 for such blocks."""
    H, W = p.shape
    Hc, Wc = -(-H // f), -(-W // f)
    r = np.where(solid, 0.0, b - lap(p, solid))
    fl = (~solid).astype(np.float64)
    pad_r = np.zeros((Hc * f, Wc * f))
    pad_c = np.zeros((Hc * f, Wc * f))
    pad_r[:H, :W] = r
    pad_c[:H, :W] = fl
    s = pad_r.reshape(Hc, f, Wc, f).sum(axis=(1, 3))
    c = pad_c.reshape(Hc, f, Wc, f).sum(axis=(1, 3))
    rc = np.where(c > 0, f * f * s / np.maximum(c, 1), 0.0)
    return rc, (c == 0)


def restrict_faces_ref(p, b, solid, f):
    """(rc, mc, fe, fs) of the face-weighted coarse operator: rc = the sum of
    r = b - L p over the fluid cells of each f x f block (0 where the block
    has none), mc as restrict_ref, and the float32 open fractions of each
    coarse cell's east face fe and south face fs: the number of the block's
    fine rows (columns) whose cells on both sides of the block boundary are
    fluid and inside the grid, over f; 0 on the grid's edge."""
    H, W = p.shape
    Hc, Wc = -(-H // f), -(-W // f)
    r = np.where(solid, 0.0, b - lap(p, solid))
    fl = ~solid
    pad_r = np.zeros((Hc * f, Wc * f))
    pad_c = np.zeros((Hc * f, Wc * f))
    pad_r[:H, :W] = r
    pad_c[:H, :W] = fl
    s = pad_r.reshape(Hc, f, Wc, f).sum(axis=(1, 3))
    c = pad_c.reshape(Hc, f, Wc, f).sum(axis=(1, 3))
    fe = np.zeros((Hc, Wc), np.float32)
    fs = np.zeros((Hc, Wc), np.float32)
    for Y in range(Hc):
        for X in range(Wc):
            xe, ys = f * X + f, f * Y + f
            if xe < W:
                n = sum(1 for y in range(f * Y, min(f * Y + f, H)) if fl[y, xe - 1] and fl[y, xe])
                fe[Y, X] = np.float32(n) / np.float32(f)
            if ys < H:
                n = sum(1 for x in range(f * X, min(f * X + f, W)) if fl[ys - 1, x] and fl[ys, x])
                fs[Y, X] = np.float32(n) / np.float32(f)
    return np.where(c > 0, s, 0.0), (c == 0), fe, fs


def prolong_ref(ec, mc, f, H, W, solid):
    """Bilinear interpolation of the coarse field at every fluid fine cell:
    coarse centres at fine coordinate f I + (f - 1) / 2, clamped to the
    coarse grid, weights renormalised over the fluid coarse cells (0 where
    none); solid fine cells get 0."""
    Hc, Wc = ec.shape
    out = np.zeros((H, W))
    for y in range(H):
        fy = min(max((y - 0.5 * (f - 1)) / f, 0.0), Hc - 1)
        y0 = int(fy)
        y1 = min(y0 + 1, Hc - 1)
        ty = fy - y0
        for x in range(W):
            if solid[y, x]:
                continue
            fx = min(max((x - 0.5 * (f - 1)) / f, 0.0), Wc - 1)
            x0 = int(fx)
            x1 = min(x0 + 1, Wc - 1)
            tx = fx - x0
            ws = [((1 - tx) * (1 - ty), y0, x0), (tx * (1 - ty), y0, x1),
                  ((1 - tx) * ty, y1, x0), (tx * ty, y1, x1)]
            num = sum(w * ec[j, i] for w, j, i in ws if not mc[j, i])
            den = sum(w for w, j, i in ws if not mc[j, i])
            out[y, x] = num / den if den > 0 else 0.0
    return out


def lap32(v, solid):
    """float32 L v in the kernels' summation order (((l + r) + u) + d) - 4c."""
    f32 = np.float32
    H, W = v.shape
    P = np.pad(v, 1, mode="edge")
    S = np.pad(solid, 1, mode="constant", constant_values=True)
    nb = []
    for dy, dx in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        n = P[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        s = S[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        nb.append(np.where(s, v, n))
    return (((nb[0] + nb[1]) + nb[2]) + nb[3]) - f32(4.0) * v


def jacobi32(x, rhs, solid, iters):
    """The coarse Jacobi sweeps' float32 operations: x + (L x - r) * 0.25."""
    for _ in range(iters):
        x = x + (lap32(x, solid) - rhs) * np.float32(0.25)
    return x


def fma32(a, b, c):
    """float32 fused multiply-add round(a b + c), elementwise: a b is exact
    in float64 (24 + 24 significant bits), the float64 sum is rounded to odd
    (its TwoSum error moves an even result one ulp towards the exact sum),
    and rounding that to float32 is then correctly rounded (53 >= 24 + 2
    bits; Boldo and Melquiond, "Emulation of FMA and correctly rounded sums:
    proved algorithms using rounding to odd", IEEE Trans. Comput. 2008)."""
    p = np.asarray(a, np.float64) * np.asarray(b, np.float64)
    c = np.asarray(c, np.float64)
    s = p + c
    t = s - p
    e = (p - (s - t)) + (c - t)
    even = (s.view(np.int64) & 1) == 0
    odd = (e != 0) & even & np.isfinite(s)
    s = np.where(odd, np.nextafter(s, np.where(e > 0, np.inf, -np.inf)), s)
    return s.astype(np.float32)


def lapc32(v, solid):
    """float32 L v in the Chebyshev kernels' summation order
    (((u + l) + r) + d) - 4c: both lane-shifted neighbours enter as the
    shifted operand of an addition."""
    f32 = np.float32
    H, W = v.shape
    P = np.pad(v, 1, mode="edge")
    S = np.pad(solid, 1, mode="constant", constant_values=True)
    nb = []
    for dy, dx in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        n = P[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        s = S[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        nb.append(np.where(s, v, n))
    return (((nb[2] + nb[0]) + nb[1]) + nb[3]) - f32(4.0) * v


def lapcw32(v, fe, fs):
    """float32 L v of the face-weighted coarse operator, sum over faces of
    w (v_J - v), in the Chebyshev kernels' order: a neighbour across a face
    of weight w enters as v (w = 0), as itself (w = 1) or as
    fma(w, v_J - v, v), then (((u + l) + r) + d) - 4 v.  Faces on the grid's
    edge count as shut whatever fe, fs hold there."""
    f32 = np.float32
    H, W = v.shape
    wr = fe.astype(f32).copy()
    wr[:, -1] = 0
    wd = fs.astype(f32).copy()
    wd[-1, :] = 0
    wl = np.zeros_like(wr)
    wl[:, 1:] = wr[:, :-1]
    wu = np.zeros_like(wd)
    wu[1:, :] = wd[:-1, :]
    P = np.pad(v, 1, mode="edge")

    def term(w, dy, dx):
        n = P[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        t = np.where(w == 0, v, n)
        part = (w > 0) & (w < 1)
        return np.where(part, fma32(w, (n - v).astype(f32), v), t)
    l, r, u, d = term(wl, 0, -1), term(wr, 0, 1), term(wu, -1, 0), term(wd, 1, 0)
    return (((u + l) + r) + d) - f32(4.0) * v


def cheb32(x, rhs, solid, iters, lmin, lapf=None):
    """The fine and coarse Chebyshev iterations' float32 operations, in the
    kernels' order: L v as lapc32 (the -4 v is one rounding, 4 v being
    exact) and d <- fma(cd, d, cr r)."""
    f32 = np.float32
    lmin = f32(lmin)
    theta = f32(0.5) * (f32(8.0) + lmin)
    delta = f32(0.5) * (f32(8.0) - lmin)
    sigma = theta / delta
    rho = f32(1.0) / sigma
    lapf = lapf or (lambda v: lapc32(v, solid))
    r = lapf(x) - rhs
    d = r / theta
    for _ in range(iters):
        ld = lapf(d)
        rho1 = f32(1.0) / (f32(2.0) * sigma - rho)
        cd = rho1 * rho
        cr = f32(2.0) * rho1 / delta
        x = x + d
        r = r + ld
        d = fma32(cd, d, cr * r)
        rho = rho1
    return x
