"""CPU oracle: a line-by-line NumPy restatement of Natrix's fluid step.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path
(``natrix_amd``) never calls into it and has no CPU fallback.

Every function is written from the reference's shader text and follows the
arithmetic order of the shader expression it cites, so that with float32
inputs each operation rounds exactly like a float32 GPU evaluation with FMA
contraction disabled.  Passing float64 arrays gives the FP64 twin with the
same operation order.

Layout: scalar fields are ``(H, W)`` arrays, velocity is ``(H, W, 2)``
(x-component first), obstacles in the reference's ``(H, W, 2)`` float layout
(``solid`` iff either channel is > 0).  Row-major ``i = y*W + x`` as in
``natrix/core/shaders/originals/shader.AddVelocity.comp:26``.

Pinning: every function here is pinned by ``tests/test_oracle_pinning.py``
with closed forms, library routines (scipy.ndimage / scipy.sparse) or
hand-computed tiny cases; none is checked only against itself.
"""
from __future__ import annotations

import numpy as np

# shader.ApplyVorticity.comp:34
VORTICITY_EPSILON = 2.4414e-4


def _s(v, dt):
    """Scalar ``v`` in the arithmetic dtype ``dt`` (a uniform upload)."""
    return dt.type(v)


def solid_of(obs):
    """``obstacle.x > 0.0f || obstacle.y > 0.0f`` (shader.AdvectVelocity.comp:30)."""
    return (obs[..., 0] > 0) | (obs[..., 1] > 0)


def obstacles_from_mask(mask):
    """Map the build's u8 mask to the reference float2 obstacle layout.

    bit0 -> channel x (dynamic obstacles, AddCircleObstacle.comp:30-34),
    bit1 -> channel y (static triangles, AddTriangleObstacle.comp:45),
    bit2 -> channel x (the persistent obstacle map, which the reference lacks).
    """
    m = np.asarray(mask, dtype=np.uint8)
    obs = np.zeros(m.shape + (2,), np.float32)
    obs[..., 0] = ((m & 1) != 0) | ((m & 4) != 0)
    obs[..., 1] = (m & 2) != 0
    return obs


def neighbours(a):
    """Clamp-to-edge L/R/B/T gathers of ``a`` (common.sh:9-19).

    L=(max(x-1,0),y), R=(min(x+1,W-1),y), B=(x,max(y-1,0)), T=(x,min(y+1,H-1)).
    """
    pad = [(1, 1), (1, 1)] + [(0, 0)] * (a.ndim - 2)
    p = np.pad(a, pad, mode="edge")
    L = p[1:-1, :-2]
    R = p[1:-1, 2:]
    B = p[:-2, 1:-1]
    T = p[2:, 1:-1]
    return L, R, B, T


def mix(a, b, t):
    """GLSL ``mix`` as ``a*(1-t) + b*t`` (the spec's form; one fixed choice)."""
    one = t.dtype.type(1)
    return a * (one - t) + b * t


# ---------------------------------------------------------------- S1
def init_boundaries(u):
    """S1: zero the one-cell ring in place (shader.InitBoundaries.comp:18-34)."""
    u = u.copy()
    u[0, :, :] = 0
    u[-1, :, :] = 0
    u[:, 0, :] = 0
    u[:, -1, :] = 0
    return u


def _bilinear(field, fx, fy):
    """Clamped-corner bilinear fetch used by AdvectVelocity.comp:38-48 and
    AdvectParticle.comp:23-34 / 58-69.

    ``field`` is (H, W) or (H, W, C).  Returns mix(h2, h1, delta.y) with
    h1 on the top (ceil y) row and h2 on the bottom (floor y) row.
    """
    H, W = field.shape[:2]
    dt = fx.dtype
    tx = np.clip(np.ceil(fx), 0, W - 1).astype(np.int64)
    ty = np.clip(np.ceil(fy), 0, H - 1).astype(np.int64)
    bx = np.clip(np.floor(fx), 0, W - 1).astype(np.int64)
    by = np.clip(np.floor(fy), 0, H - 1).astype(np.int64)
    dx = fx - bx.astype(dt)
    dy = fy - by.astype(dt)
    lt = field[ty, bx]
    rt = field[ty, tx]
    lb = field[by, bx]
    rb = field[by, tx]
    if field.ndim == 3:
        dx = dx[..., None]
        dy = dy[..., None]
    h1 = mix(lt, rt, dx)
    h2 = mix(lb, rb, dx)
    return mix(h2, h1, dy)


# ---------------------------------------------------------------- S2
def advect_velocity(u, obs, dt, speed, dissipation):
    """S2: semi-Lagrangian advection (shader.AdvectVelocity.comp:27-49)."""
    H, W = u.shape[:2]
    d = u.dtype
    dt_, sp = _s(dt, d), _s(speed, d)
    X = np.broadcast_to(np.arange(W, dtype=d)[None, :], (H, W))
    Y = np.broadcast_to(np.arange(H, dtype=d)[:, None], (H, W))
    # final_pos = (x - vel.x * _ElapsedTime * _Speed, y - vel.y * ...)
    fx = X - (u[..., 0] * dt_) * sp
    fy = Y - (u[..., 1] * dt_) * sp
    v = _bilinear(u, fx, fy) * _s(dissipation, d)
    v = np.clip(v, d.type(-1), d.type(1))
    v[solid_of(obs)] = 0
    return v


# ---------------------------------------------------------------- S3
def calc_vorticity(u):
    """S3: curl (shader.CalcVorticity.comp:21-26)."""
    L, R, B, T = neighbours(u)
    half = u.dtype.type(0.5)
    return half * ((R[..., 1] - L[..., 1]) - (T[..., 0] - B[..., 0]))


# ---------------------------------------------------------------- S4
def apply_vorticity(u, w, dt, scale):
    """S4: vorticity confinement (shader.ApplyVorticity.comp:27-39)."""
    d = u.dtype
    half = d.type(0.5)
    wL, wR, wB, wT = neighbours(w)
    fx = half * (np.abs(wT) - np.abs(wB))
    fy = half * (np.abs(wR) - np.abs(wL))
    mag = np.maximum(d.type(VORTICITY_EPSILON), fx * fx + fy * fy)
    inv = d.type(1) / np.sqrt(mag)  # inversesqrt (decision D7)
    fx = fx * inv
    fy = fy * inv
    s = _s(scale, d) * w  # _VorticityScale * vC, times vec2(1,-1)
    fx = fx * s
    fy = fy * (-s)
    dt_ = _s(dt, d)
    out = np.empty_like(u)
    out[..., 0] = u[..., 0] + fx * dt_
    out[..., 1] = u[..., 1] + fy * dt_
    return out


# ---------------------------------------------------------------- S5
def viscosity_coefficients(nu):
    """(alpha, rbeta) as fluid_simulator.py:327-336 computes them: in Python
    float64, then cast to float32 for the uniform upload."""
    alpha = 1.0 / nu
    rbeta = 1.0 / (4.0 + alpha)
    return np.float32(alpha), np.float32(rbeta)


def viscosity(u, alpha, rbeta):
    """S5: one diffusion Jacobi sweep (shader.Viscosity.comp:25-31)."""
    d = u.dtype
    L, R, B, T = neighbours(u)
    return ((((L + R) + B) + T) + u * _s(alpha, d)) * _s(rbeta, d)


# ---------------------------------------------------------------- S6
def divergence(u, obs):
    """S6: divergence with solid neighbours' component := 0
    (shader.Divergence.comp:24-40)."""
    sL, sR, sB, sT = neighbours(solid_of(obs))
    L, R, B, T = neighbours(u)
    z = u.dtype.type(0)
    x1 = np.where(sL, z, L[..., 0])
    x2 = np.where(sR, z, R[..., 0])
    y1 = np.where(sB, z, B[..., 1])
    y2 = np.where(sT, z, T[..., 1])
    return u.dtype.type(0.5) * ((x2 - x1) + (y2 - y1))


# ---------------------------------------------------------------- S8
def _replaced_neighbours(p, solid):
    """Pressure neighbours with a solid neighbour replaced by the centre value
    (shader.Poisson.comp:27-35, shader.SubtractGradient.comp:26-42)."""
    sL, sR, sB, sT = neighbours(solid)
    L, R, B, T = neighbours(p)
    return (np.where(sL, p, L), np.where(sR, p, R),
            np.where(sB, p, B), np.where(sT, p, T))


def jacobi(p, b, obs, sweeps=1):
    """S8: ``sweeps`` Jacobi pressure sweeps with ping-pong
    (shader.Poisson.comp:25-37, loop at fluid_simulator.py:251-255)."""
    solid = solid_of(obs)
    q = p.dtype.type(0.25)
    for _ in range(sweeps):
        x1, x2, y1, y2 = _replaced_neighbours(p, solid)
        p = ((((x1 + x2) + y1) + y2) - b) * q
    return p


# ---------------------------------------------------------------- S9
def subtract_gradient(u, p, obs):
    """S9: projection (shader.SubtractGradient.comp:26-46)."""
    x1, x2, y1, y2 = _replaced_neighbours(p, solid_of(obs))
    half = u.dtype.type(0.5)
    out = np.empty_like(u)
    out[..., 0] = u[..., 0] - half * (x2 - x1)
    out[..., 1] = u[..., 1] - half * (y2 - y1)
    return out


# ---------------------------------------------------------------- F1
def _splat_distance(W, H, px, py, d):
    """``distance(_Position * vec2(_Size), vec2(gid))`` (AddVelocity.comp:27-30)."""
    sx = _s(px, d) * d.type(W)
    sy = _s(py, d) * d.type(H)
    X = np.arange(W, dtype=d)[None, :]
    Y = np.arange(H, dtype=d)[:, None]
    dx = sx - X
    dy = sy - Y
    return np.sqrt(dx * dx + dy * dy)


def add_velocity(u, position, value, radius):
    """F1: cone splat then clamp of every cell (shader.AddVelocity.comp:26-35)."""
    H, W = u.shape[:2]
    d = u.dtype
    ln = _splat_distance(W, H, position[0], position[1], d)
    r = _s(radius, d)
    inside = ln <= r
    out = u.copy()
    for c in range(2):
        add = u[..., c] + (_s(value[c], d) * (r - ln)) / r
        out[..., c] = np.where(inside, add, u[..., c])
    return np.clip(out, d.type(-1), d.type(1))


# ---------------------------------------------------------------- F2
def add_circle_obstacle(obs, position, radius, static=False):
    """F2: ``distance <= _Radius`` -> (1,0) in both branches
    (shader.AddCircleObstacle.comp:24-36)."""
    H, W = obs.shape[:2]
    ln = _splat_distance(W, H, position[0], position[1], np.dtype(np.float32))
    out = obs.copy()
    inside = ln <= np.float32(radius)
    out[inside, 0] = 1
    out[inside, 1] = 0
    return out


# ---------------------------------------------------------------- F3
def _sign(p1x, p1y, p2x, p2y, p3x, p3y):
    """shader.AddTriangleObstacle.comp:19-22."""
    return ((p1x - p3x) * (p2y - p3y)) - ((p2x - p3x) * (p1y - p3y))


def add_triangle_obstacle(obs, p1, p2, p3, static=False):
    """F3: sign-test raster on normalised coordinates
    (shader.AddTriangleObstacle.comp:24-51)."""
    H, W = obs.shape[:2]
    f = np.float32
    X = np.broadcast_to(np.arange(W, dtype=f)[None, :], (H, W))
    Y = np.broadcast_to(np.arange(H, dtype=f)[:, None], (H, W))
    px = X / f(W)
    py = Y / f(H)
    v = [tuple(map(f, q)) for q in (p1, p2, p3)]
    b1 = _sign(px, py, v[0][0], v[0][1], v[1][0], v[1][1]) < 0
    b2 = _sign(px, py, v[1][0], v[1][1], v[2][0], v[2][1]) < 0
    b3 = _sign(px, py, v[2][0], v[2][1], v[0][0], v[0][1]) < 0
    inside = (b1 == b2) & (b2 == b3)
    out = obs.copy()
    out[inside, 0] = 0 if static else 1
    out[inside, 1] = 1 if static else 0
    return out


# ---------------------------------------------------------------- S11
def advect_dye(q, u, obs, dt, speed, dissipation):
    """S11: dye ("particles") advection on a dye grid P=(Wp,Hp) driven by the
    velocity grid V=(Wv,Hv) (demo/shaders/shader.AdvectParticle.comp:21-69).

    The scalar mixes at :67-69 take the x component of the vec2 mix
    (decision D14)."""
    Hp, Wp = q.shape
    Hv, Wv = u.shape[:2]
    d = q.dtype
    X = np.broadcast_to(np.arange(Wp, dtype=d)[None, :], (Hp, Wp))
    Y = np.broadcast_to(np.arange(Hp, dtype=d)[:, None], (Hp, Wp))
    nx = (X / d.type(Wp)) * d.type(Wv)
    ny = (Y / d.type(Hp)) * d.type(Hv)
    ox = nx.astype(np.int64)  # uint(fNormalisedPos.x): truncation, n >= 0
    oy = ny.astype(np.int64)
    solid = solid_of(obs)[oy, ox]
    vel = _bilinear(u, nx, ny)
    sx = d.type(Wp) / d.type(Wv)  # (_ParticleSize / _VelocitySize)
    sy = d.type(Hp) / d.type(Hv)
    vx = vel[..., 0] * sx
    vy = vel[..., 1] * sy
    dt_, sp = _s(dt, d), _s(speed, d)
    fx = X - (vx * dt_) * sp
    fy = Y - (vy * dt_) * sp
    out = _bilinear(q, fx, fy) * _s(dissipation, d)
    out[solid] = 0
    return out


# ---------------------------------------------------------------- F4
def dye_rgba8(q):
    """demo/shaders/demo.ComputeShader.comp:9-19: the dye value stored to an
    rgba8 image in all four channels, i.e. round(clamp(v, 0, 1) * 255) in
    float32 with ties rounded up; a NaN stores 0."""
    f = np.float32
    v = np.nan_to_num(q.astype(f), nan=0.0, posinf=1.0, neginf=0.0)
    v = np.minimum(np.maximum(v, f(0.0)), f(1.0))
    c = np.floor(v * f(255.0) + f(0.5)).astype(np.uint8)
    return np.repeat(c[..., None], 4, axis=-1)


def quiver_rgba8(u, window_w, window_h, tile):
    """demo/shaders/demo.QuiverFragmentShader.frag:14-70 for every pixel of
    a window_w x window_h framebuffer (row y holds gl_FragCoord.y = y + 0.5),
    in float32 in the shader's order: the velocity at the pixel's arrow-tile
    centre (_field: fPos = (1 - pos / WindowSize) * VelocitySize, corners
    clamped to the grid, mix(a, b, (t, 0)), which interpolates only the x
    component, scaled by -WindowSize / VelocitySize), times tile * 0.4, then
    the distance to that arrow (_vector: shaft and two head strokes, each the
    max of the distances across and along a segment) and colour (1, 1, 1,
    1 - clamp(dist, 0, 1)) stored to an rgba8 image (round to nearest, ties
    up).  u: [Hv, Wv, 2] float32.  Returns [window_h, window_w, 4] uint8."""
    f = np.float32
    Hv, Wv = u.shape[:2]
    T = f(tile)
    X = np.arange(window_w, dtype=f)[None, :] + f(0.5)
    Y = np.arange(window_h, dtype=f)[:, None] + f(0.5)
    X, Y = np.broadcast_arrays(X, Y)

    def centre(p):
        return (np.floor(p / T) + f(0.5)) * T

    def mix(a, b, t):
        return a * (f(1.0) - t) + b * t

    cx, cy = centre(X), centre(Y)
    Vw, Vh, Ww, Wh = f(Wv), f(Hv), f(window_w), f(window_h)
    fx = (f(1.0) - cx / Ww) * Vw
    fy = (f(1.0) - cy / Wh) * Vh
    trx = np.clip(np.ceil(fx), f(0), f(Vw - f(1))).astype(np.int64)
    tr_y = np.clip(np.ceil(fy), f(0), f(Vh - f(1))).astype(np.int64)
    blx = np.clip(np.floor(fx), f(0), f(Vw - f(1))).astype(np.int64)
    bly = np.clip(np.floor(fy), f(0), f(Vh - f(1))).astype(np.int64)
    dx = fx - blx.astype(f)
    dy = fy - bly.astype(f)
    lt, rt = u[tr_y, blx], u[tr_y, trx]
    lb, rb = u[bly, blx], u[bly, trx]
    z = f(0.0)
    h1 = (mix(lt[..., 0], rt[..., 0], dx), mix(lt[..., 1], rt[..., 1], z))
    h2 = (mix(lb[..., 0], rb[..., 0], dx), mix(lb[..., 1], rb[..., 1], z))
    m = (mix(h2[0], h1[0], dy), mix(h2[1], h1[1], z))
    vx = ((f(-1.0) * m[0]) * (Ww / Vw)) * T * f(0.4)
    vy = ((f(-1.0) * m[1]) * (Wh / Vh)) * T * f(0.4)
    px, py = X - cx, Y - cy

    def length(ax, ay):
        return np.sqrt(ax * ax + ay * ay)

    def line(p1x, p1y, p2x, p2y):
        ccx, ccy = (p1x + p2x) * f(0.5), (p1y + p2y) * f(0.5)
        ln = length(p2x - p1x, p2y - p1y)
        dirx, diry = (p2x - p1x) / ln, (p2y - p1y) / ln
        rx, ry = px - ccx, py - ccy
        d1 = np.abs(rx * diry + ry * (-dirx))
        d2 = np.abs(rx * dirx + ry * diry) - f(0.5) * ln
        return np.maximum(d1, d2)

    mag = length(vx, vy)
    arrow = mag > f(0.001)
    with np.errstate(divide="ignore", invalid="ignore"):
        ux, uy = vx / mag, vy / mag
        mc = np.clip(mag, f(0.0), T * f(0.5))
        ax, ay = ux * mc, uy * mc
        shaft = line(ax, ay, -ax, -ay)
        head = np.minimum(line(ax, ay, f(0.4) * ax + f(0.2) * (-ay), f(0.4) * ay + f(0.2) * ax),
                          line(ax, ay, f(0.4) * ax + f(0.2) * ay, f(0.4) * ay + f(0.2) * (-ax)))
    dist = np.where(arrow, np.minimum(shaft, head), f(1.0))
    alpha = f(1.0) - np.minimum(np.maximum(dist, f(0.0)), f(1.0))
    a8 = np.floor(alpha * f(255.0) + f(0.5)).astype(np.uint8)
    out = np.full((window_h, window_w, 4), 255, np.uint8)
    out[..., 3] = a8
    return out


def add_dye(q, position, value, radius):
    """F4: cone dye splat, clamp [0,255] inside the disk
    (demo/shaders/shader.AddParticle.comp:24-34)."""
    H, W = q.shape
    d = q.dtype
    ln = _splat_distance(W, H, position[0], position[1], d)
    r = _s(radius, d)
    add = np.clip(q + (_s(value, d) * (r - ln)) / r, d.type(0), d.type(255))
    return np.where(ln <= r, add, q)


# ---------------------------------------------------------------- extensions
def gaussian_weight(W, H, px, py, radius, d=np.dtype(np.float32)):
    """Gaussian splat weight (no reference counterpart; written spec):

        dx = px*W - x, dy = py*H - y          (float32, as F1)
        e  = -(dx*dx + dy*dy) / (r*r)         (float32)
        w  = float32(exp(float64(e)))

    The exponential is evaluated in float64 and rounded once, so any
    correctly-rounded-to-float64 exp gives the same float32 weight."""
    d = np.dtype(d)
    sx = _s(px, d) * d.type(W)
    sy = _s(py, d) * d.type(H)
    X = np.arange(W, dtype=d)[None, :]
    Y = np.arange(H, dtype=d)[:, None]
    dx = sx - X
    dy = sy - Y
    r = _s(radius, d)
    e = -(dx * dx + dy * dy) / (r * r)
    return np.exp(e.astype(np.float64)).astype(d)


def gaussian_splat(u, q, splats):
    """Apply Gaussian splats in order.  Each splat is a tuple
    ``(px, py, ax, ay, radius, dye)``: ``u += (ax, ay) * w`` and
    ``q += dye * w`` (amplitudes already multiplied by dt by the caller, in
    float32).  No clamp: the next advection clamps velocity to [-1, 1]
    (decision D15)."""
    u = u.copy()
    q = None if q is None else q.copy()
    H, W = u.shape[:2]
    d = u.dtype
    for (px, py, ax, ay, radius, dye) in splats:
        w = gaussian_weight(W, H, px, py, radius, d)
        u[..., 0] = u[..., 0] + _s(ax, d) * w
        u[..., 1] = u[..., 1] + _s(ay, d) * w
        if q is not None:
            q = q + _s(dye, d) * w
    return u, q


BC_REFERENCE, BC_FREE_SLIP, BC_INFLOW = 0, 1, 2


def enforce_boundaries(u, mode, has_borders=True, inflow=1.0):
    """Wall treatment applied where the reference runs InitBoundaries (start
    of the step, in place on the current velocity).

    * BC_REFERENCE: InitBoundaries.comp:18-34 when ``has_borders``.
    * BC_FREE_SLIP: normal component 0 and tangential copied from the interior
      neighbour on every wall; corners (0, 0).
    * BC_INFLOW: left column (inflow, 0); right column copies column W-2
      (zero-gradient outflow); bottom/top free-slip; corners (0, 0) except the
      two left corners, which take the inflow value.  Applied in the order
      bottom/top, right, left, so the left column wins.
    """
    u = u.copy()
    H, W = u.shape[:2]
    if mode == BC_REFERENCE:
        return init_boundaries(u) if has_borders else u
    if mode == BC_FREE_SLIP:
        u[1:-1, 0, 0] = 0
        u[1:-1, 0, 1] = u[1:-1, 1, 1]
        u[1:-1, -1, 0] = 0
        u[1:-1, -1, 1] = u[1:-1, -2, 1]
        u[0, 1:-1, 1] = 0
        u[0, 1:-1, 0] = u[1, 1:-1, 0]
        u[-1, 1:-1, 1] = 0
        u[-1, 1:-1, 0] = u[-2, 1:-1, 0]
        for (y, x) in ((0, 0), (0, W - 1), (H - 1, 0), (H - 1, W - 1)):
            u[y, x] = 0
        return u
    if mode == BC_INFLOW:
        # bottom/top free-slip (interior columns)
        u[0, 1:-1, 1] = 0
        u[0, 1:-1, 0] = u[1, 1:-1, 0]
        u[-1, 1:-1, 1] = 0
        u[-1, 1:-1, 0] = u[-2, 1:-1, 0]
        # right outflow: zero gradient, corners zero
        u[1:-1, -1, :] = u[1:-1, -2, :]
        u[0, -1] = 0
        u[-1, -1] = 0
        # left inflow, whole column
        u[:, 0, 0] = u.dtype.type(inflow)
        u[:, 0, 1] = 0
        return u
    raise ValueError(f"unknown boundary mode {mode}")


def residual_sums(p, b, obs):
    """FP64 sums for the relative residual of ``p`` (no reference counterpart).

    A p := sum of replaced neighbours - 4 p (the operator whose fixed point
    shader.Poisson.comp:37 iterates to), r := b - A p, over fluid cells only.
    Returns (sum r, sum r^2, sum b, sum b^2, n_fluid) in float64."""
    p = p.astype(np.float64)
    b = b.astype(np.float64)
    solid = solid_of(obs)
    x1, x2, y1, y2 = _replaced_neighbours(p, solid)
    r = b - ((((x1 + x2) + y1) + y2) - 4.0 * p)
    f = ~solid
    rf, bf = r[f], b[f]
    return (float(rf.sum()), float((rf * rf).sum()), float(bf.sum()),
            float((bf * bf).sum()), int(f.sum()))


def relative_residual_from_sums(sr, srr, sb, sbb, n):
    """||P r|| / ||P b|| where P removes the fluid-cell mean (the constant
    null space of the pure-Neumann operator)."""
    if n == 0:
        return 0.0
    num = max(srr - sr * sr / n, 0.0)
    den = max(sbb - sb * sb / n, 0.0)
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return float(np.sqrt(num / den))


def relative_residual(p, b, obs):
    """||P r|| / ||P b|| evaluated two-pass in FP64 (mean removed first), so
    it carries no cancellation error; the one-pass form above is what a
    streaming reduction computes."""
    p64 = p.astype(np.float64)
    b64 = b.astype(np.float64)
    solid = solid_of(obs)
    x1, x2, y1, y2 = _replaced_neighbours(p64, solid)
    r = (b64 - ((((x1 + x2) + y1) + y2) - 4.0 * p64))[~solid]
    bf = b64[~solid]
    if bf.size == 0:
        return 0.0
    r = r - r.mean()
    bf = bf - bf.mean()
    den = float(np.sqrt((bf * bf).sum()))
    num = float(np.sqrt((r * r).sum()))
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return num / den


# ---------------------------------------------------------------- the step
class Params:
    """Simulator parameters, defaults as fluid_simulator.py:28-35."""

    def __init__(self, **kw):
        self.speed = 500.0
        self.iterations = 50
        self.dissipation = 1.0
        self.vorticity = 0.0
        self.viscosity = 0.1
        self.has_borders = True
        self.bc_mode = BC_REFERENCE
        self.inflow = 1.0
        self.dye_speed = 500.0  # SmoothParticlesArea._speed, smooth_particles_area.py:23
        self.dye_dissipation = 1.0
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)


def step(u, p, obs, dt, prm, dye=None, persistent=None, return_stages=False):
    """One ``FluidSimulator.update(dt)`` (fluid_simulator.py:174-280) followed
    by the dye advection the demo runs after it (simulation_demo.py:222-223).

    ``obs`` is the dynamic obstacle buffer for this step; ``persistent`` an
    optional boolean map OR-ed in (the obstacle map the reference lacks).
    Returns (u, p, obs_after_clear, dye) and optionally a dict of stage
    outputs.  The dye sees the obstacle buffer after the end-of-step clear
    (decision D9)."""
    d = u.dtype
    stages = {}
    o = obs.copy()
    if persistent is not None:
        o[persistent, 0] = 1
    u = enforce_boundaries(u, prm.bc_mode, prm.has_borders, prm.inflow)
    stages["bc"] = u
    u = advect_velocity(u, o, dt, prm.speed, prm.dissipation)
    stages["advect"] = u
    w = calc_vorticity(u)
    stages["curl"] = w
    u = apply_vorticity(u, w, dt, prm.vorticity)
    stages["confine"] = u
    if prm.viscosity > 0.0:
        a, rb = viscosity_coefficients(prm.viscosity)
        u = viscosity(u, a, rb)
        stages["viscosity"] = u
    b = divergence(u, o)
    stages["divergence"] = b
    p = np.zeros_like(p)
    p = jacobi(p, b, o, prm.iterations)
    stages["pressure"] = p
    u = subtract_gradient(u, p, o)
    stages["project"] = u
    o_after = np.zeros_like(obs)
    if persistent is not None:
        o_after[persistent, 0] = 1
    if dye is not None:
        dye = advect_dye(dye, u, o_after, dt, prm.dye_speed, prm.dye_dissipation)
        stages["dye"] = dye
    out = (u, p, np.zeros_like(obs), dye)
    if return_stages:
        return out + (stages,)
    return out
