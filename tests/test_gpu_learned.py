"""The learned pressure solve's building blocks against NumPy statements of
their mathematics, and the solve itself: a useless or non-finite network
cannot spoil it (plain Jacobi still reaches the tolerance), and the trained
network reduces the work to the tolerance on an unseen scene.

Coarse smoother: Jacobi sweeps and the Chebyshev iteration for -L e = -r on
[lmin, 8], against the float64 statements of tests/smoothers.py (pinned on
eigenmodes by tests/test_smoother_statements.py).  Line search:
a = d.r / d.Ld minimises the energy phi(p) = 1/2 p.(-L)p + b.p along d.  Damped sweep: p + w (J(p) - p), the
float32 operations of the plain sweep (shader.Poisson.comp:26-37)."""
import ctypes as C

import numpy as np
import pytest
import torch

from natrix_amd import FluidSimulator, _lib

from smoothers import (cheb32, cheb_ref, jacobi32, jacobi_ref, lap, lapcw32, prolong_ref,
                       restrict_faces_ref, restrict_ref)

pytestmark = pytest.mark.gpu


def in_step_sim(rng, H, W, tile=None):
    sim = FluidSimulator(W, H)
    sim.viscosity = 0.0
    if tile is not None:
        sim.jacobi_tile = tile
    Y, X = np.mgrid[0:H, 0:W]
    sim.set_obstacle_map((X - 0.4 * W) ** 2 + (Y - 0.5 * H) ** 2 <= (0.15 * H) ** 2)
    sim._push_params()
    _lib.call("nx_step_begin", sim._alive(), 1 / 60)
    sim.synchronize()
    return sim


def dev(sim, name):
    return torch.as_tensor(sim.field_buffer(name), device="cuda")


@pytest.mark.parametrize("Wc,Hc", [(128, 128), (100, 72), (7, 128), (128, 13)])
@pytest.mark.parametrize("cheb,iters", [(False, 25), (True, 12), (True, 13), (True, 40),
                                        (True, 100)])
def test_coarse_smoother_matches_float64_statement(gpu, rng, Wc, Hc, cheb, iters):
    """Chebyshev with more than 12 iterations runs tiled (64 x 32 tiles with
    a 13-cell halo, 12 iterations per launch, a partial last launch for 13,
    40 and 100); 12 or fewer run in one resident workgroup."""
    from natrix_amd.learned.solver import chebyshev_lmin
    sim = in_step_sim(rng, 64, 64)
    solid = rng.random((Hc, Wc)) < 0.1
    e0 = rng.normal(size=(Hc, Wc)).astype(np.float32)
    rc = rng.normal(size=(Hc, Wc))
    # a consistent right-hand side, as the restriction of a mean-free
    # residual is: otherwise the operator's constant mode grows linearly
    rc[~solid] -= rc[~solid].mean()
    rc = rc.astype(np.float32)
    e = torch.from_numpy(e0).cuda().contiguous()
    r = torch.from_numpy(rc).cuda()
    m = torch.from_numpy(solid.astype(np.uint8)).cuda()
    lmin = chebyshev_lmin(iters, 1e-3) if cheb else 0.0
    torch.cuda.synchronize()
    _lib.call("nx_coarse_smooth", sim._alive(), Wc, Hc, C.c_void_p(e.data_ptr()),
              C.c_void_p(r.data_ptr()), C.c_void_p(m.data_ptr()), iters, lmin)
    sim.synchronize()
    extra = (lmin,) if cheb else ()
    got = e.cpu().numpy()
    # bitwise: the float32 statement of the kernels' own operations
    want32 = (cheb32 if cheb else jacobi32)(e0, rc, solid, iters, *extra)
    assert np.array_equal(got, want32)
    # and the float64 iteration, to float32 round-off accumulated over the
    # iterations (the float32 statement itself is 1.5e-5 off after 100)
    ref = (cheb_ref if cheb else jacobi_ref)(e0.astype(np.float64), rc.astype(np.float64),
                                            solid, iters, *extra)
    assert np.abs(got - ref).max() <= 5e-5 * np.abs(ref).max()
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


@pytest.mark.parametrize("Wc,Hc", [(128, 128), (100, 72), (7, 128), (128, 13)])
@pytest.mark.parametrize("iters", [5, 12, 13, 40, 100])
def test_weighted_coarse_smoother_matches_float32_statement(gpu, rng, Wc, Hc, iters):
    """The face-weighted coarse operator (always tiled, 12 iterations per
    launch) against the float32 statement of its operations, bitwise, with
    faces of every kind: shut (0, as toward a solid cell), open (1) and
    partial."""
    from natrix_amd.learned.solver import chebyshev_lmin
    sim = in_step_sim(rng, 64, 64)
    solid = rng.random((Hc, Wc)) < 0.1
    fe = rng.choice([0.0, 0.25, 0.5, 0.75, 1.0, 1.0, 1.0], size=(Hc, Wc)).astype(np.float32)
    fs = rng.choice([0.0, 0.25, 0.5, 0.75, 1.0, 1.0, 1.0], size=(Hc, Wc)).astype(np.float32)
    # faces of and toward solid cells are shut, as k_coarse_faces writes them
    fe[solid] = 0
    fe[:, :-1][solid[:, 1:]] = 0
    fs[solid] = 0
    fs[:-1, :][solid[1:, :]] = 0
    e0 = rng.normal(size=(Hc, Wc)).astype(np.float32)
    rc = rng.normal(size=(Hc, Wc))
    rc[~solid] -= rc[~solid].mean()
    rc = rc.astype(np.float32)
    e = torch.from_numpy(e0).cuda().contiguous()
    r = torch.from_numpy(rc).cuda()
    m = torch.from_numpy(solid.astype(np.uint8)).cuda()
    fw = torch.from_numpy(np.concatenate([fe.ravel(), fs.ravel()])).cuda()
    lmin = chebyshev_lmin(iters, 1e-3)
    torch.cuda.synchronize()
    _lib.call("nx_coarse_smooth_faces", sim._alive(), Wc, Hc, C.c_void_p(e.data_ptr()),
              C.c_void_p(r.data_ptr()), C.c_void_p(m.data_ptr()), C.c_void_p(fw.data_ptr()),
              iters, lmin)
    sim.synchronize()
    want = cheb32(e0, rc, solid, iters, lmin, lapf=lambda v: lapcw32(v, fe, fs))
    assert np.array_equal(e.cpu().numpy(), want)
    with pytest.raises(ValueError):  # Chebyshev only
        _lib.call("nx_coarse_smooth_faces", sim._alive(), Wc, Hc, C.c_void_p(e.data_ptr()),
                  C.c_void_p(r.data_ptr()), C.c_void_p(m.data_ptr()),
                  C.c_void_p(fw.data_ptr()), iters, 0.0)
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


def test_coarse_smoother_rejects_oversized_grids(gpu, rng):
    sim = in_step_sim(rng, 32, 32)
    t = torch.zeros(129 * 4, device="cuda")
    m = torch.zeros(129 * 4, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError):
        _lib.call("nx_coarse_smooth", sim._alive(), 129, 4, C.c_void_p(t.data_ptr()),
                  C.c_void_p(t.data_ptr()), C.c_void_p(m.data_ptr()), 5, 0.0)
    with pytest.raises(ValueError):
        _lib.call("nx_coarse_smooth", sim._alive(), 4, 4, C.c_void_p(t.data_ptr()),
                  C.c_void_p(t.data_ptr()), C.c_void_p(m.data_ptr()), 5, 8.0)
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


@pytest.mark.parametrize("N,f", [(64, 2), (96, 3), (128, 4), (128, 16), (80, 32)])
def test_restriction_matches_float64_statement(gpu, rng, N, f):
    sim = in_step_sim(rng, N, N)
    solid = sim.download("mask") != 0
    b = rng.normal(size=(N, N)).astype(np.float32)
    p = rng.normal(size=(N, N)).astype(np.float32)
    dev(sim, "divergence").copy_(torch.from_numpy(b))
    dev(sim, "pressure").copy_(torch.from_numpy(p))
    Wc = Hc = -(-N // f)
    rc = torch.empty(Hc * Wc, device="cuda")
    mc = torch.empty(Hc * Wc, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    _lib.call("nx_restrict_residual", sim._alive(), f, C.c_void_p(rc.data_ptr()),
              C.c_void_p(mc.data_ptr()))
    sim.synchronize()
    want, wmask = restrict_ref(p.astype(np.float64), b.astype(np.float64), solid, f)
    got = rc.cpu().numpy().reshape(Hc, Wc)
    assert np.array_equal(mc.cpu().numpy().reshape(Hc, Wc) != 0, wmask)
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


@pytest.mark.parametrize("N,f", [(64, 2), (96, 3), (128, 4), (128, 16), (80, 32), (100, 16)])
def test_restriction_with_faces_matches_statement(gpu, rng, N, f):
    """Block sums and the coarse faces' open fractions (the pair of the
    face-weighted coarse operator) on a grid with a circular obstacle, which
    cuts blocks and faces; 100 / 16 leaves a ragged last block."""
    sim = in_step_sim(rng, N, N)
    solid = sim.download("mask") != 0
    b = rng.normal(size=(N, N)).astype(np.float32)
    p = rng.normal(size=(N, N)).astype(np.float32)
    dev(sim, "divergence").copy_(torch.from_numpy(b))
    dev(sim, "pressure").copy_(torch.from_numpy(p))
    Wc = Hc = -(-N // f)
    rc = torch.empty(Hc * Wc, device="cuda")
    mc = torch.empty(Hc * Wc, dtype=torch.uint8, device="cuda")
    faces = torch.full((2 * Hc * Wc,), -1.0, device="cuda")
    torch.cuda.synchronize()
    _lib.call("nx_restrict_residual_faces", sim._alive(), f, C.c_void_p(rc.data_ptr()),
              C.c_void_p(mc.data_ptr()), C.c_void_p(faces.data_ptr()))
    sim.synchronize()
    want, wmask, fe, fs = restrict_faces_ref(p.astype(np.float64), b.astype(np.float64),
                                             solid, f)
    got = rc.cpu().numpy().reshape(Hc, Wc)
    assert np.array_equal(mc.cpu().numpy().reshape(Hc, Wc) != 0, wmask)
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
    fw = faces.cpu().numpy()
    assert np.array_equal(fw[:Hc * Wc].reshape(Hc, Wc), fe)
    assert np.array_equal(fw[Hc * Wc:].reshape(Hc, Wc), fs)
    assert ((fe > 0) & (fe < 1)).any()  # the obstacle cuts faces
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


@pytest.mark.parametrize("N,f", [(64, 4), (96, 3)])
def test_prolongation_matches_float64_statement(gpu, rng, N, f):
    sim = in_step_sim(rng, N, N)
    solid = sim.download("mask") != 0
    Hc = Wc = -(-N // f)
    ec = rng.normal(size=(Hc, Wc)).astype(np.float32)
    cm = rng.random((Hc, Wc)) < 0.15
    p = rng.normal(size=(N, N)).astype(np.float32)
    dev(sim, "pressure").copy_(torch.from_numpy(p))
    e = torch.from_numpy(ec).cuda().contiguous()
    m = torch.from_numpy(cm.astype(np.uint8)).cuda()
    torch.cuda.synchronize()
    _lib.call("nx_prolong_add", sim._alive(), f, C.c_void_p(e.data_ptr()),
              C.c_void_p(m.data_ptr()), 0.5)
    got = sim.download("pressure").astype(np.float64)
    want = p + 0.5 * prolong_ref(ec.astype(np.float64), cm, f, N, N, solid)
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


@pytest.mark.parametrize("H,W", [(72, 100), (131, 257)])
def test_refine_residual_matches_float64_statement(gpu, rng, H, W):
    """After nx_refine_begin and one nx_refine_accumulate of e the pressure
    pair is (e, 0) exactly; nx_refine_residual then makes the float32 of the
    FP64 residual b - L e (the operator's summation order) the right-hand
    side on fluid cells, 0 on solid ones, restarts the pressure at its first
    Jacobi sweep (0 - r) / 4 and returns the FP64 relative residual."""
    sim = in_step_sim(rng, H, W)
    solid = sim.download("mask") != 0
    b = rng.normal(size=(H, W)).astype(np.float32)
    e = rng.normal(size=(H, W)).astype(np.float32)
    dev(sim, "divergence").copy_(torch.from_numpy(b))
    torch.cuda.synchronize()
    _lib.call("nx_refine_begin", sim._alive())
    dev(sim, "pressure").copy_(torch.from_numpy(e))
    torch.cuda.synchronize()
    _lib.call("nx_refine_accumulate", sim._alive())
    rel = C.c_double()
    _lib.call("nx_refine_residual", sim._alive(), C.byref(rel))
    r64 = b.astype(np.float64) - lap(e.astype(np.float64), solid)
    want = np.where(solid, np.float32(0), r64.astype(np.float32))
    assert np.array_equal(sim.download("divergence"), want)
    assert np.array_equal(sim.download("pressure"), (np.float32(0) - want) * np.float32(0.25))
    rf, bf = r64[~solid], b.astype(np.float64)[~solid]
    n = rf.size
    want_rel = np.sqrt((np.sum(rf * rf) - np.sum(rf) ** 2 / n) /
                       (np.sum(bf * bf) - np.sum(bf) ** 2 / n))
    assert abs(rel.value - want_rel) <= 1e-12 * want_rel
    _lib.call("nx_refine_end", sim._alive())
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


def test_line_search_is_the_energy_minimiser(gpu, rng):
    H, W = 80, 96
    sim = in_step_sim(rng, H, W)
    solid = sim.download("mask") != 0
    b = rng.normal(size=(H, W)).astype(np.float32)
    p = rng.normal(size=(H, W)).astype(np.float32)
    d = rng.normal(size=(H, W)).astype(np.float32)
    d[solid] = 0
    dev(sim, "divergence").copy_(torch.from_numpy(b))
    dev(sim, "pressure").copy_(torch.from_numpy(p))
    dd = torch.from_numpy(d).cuda()
    torch.cuda.synchronize()
    out = (C.c_double * 3)()
    _lib.call("nx_line_search", sim._alive(), C.c_void_p(dd.data_ptr()), out)
    got = sim.download("pressure")
    fl = ~solid
    p64, b64, d64 = p.astype(np.float64), b.astype(np.float64), d.astype(np.float64)
    r = b64 - lap(p64, solid)
    dr = float((d64 * r)[fl].sum())
    dld = float((d64 * lap(d64, solid))[fl].sum())
    a = dr / dld
    assert abs(out[0] - a) <= 1e-9 * abs(a)
    assert abs(out[1] - dr) <= 1e-9 * abs(dr) and abs(out[2] - dld) <= 1e-9 * abs(dld)
    want = np.where(fl, p + np.float32(a) * d, p)
    assert np.array_equal(got, want.astype(np.float32))

    def phi(q):
        q = q.astype(np.float64)
        return float((0.5 * q * -lap(q, solid) + b64 * q)[fl].sum())
    # a is the minimiser: the energy is below both neighbours along d
    assert phi(got) < phi(p)
    assert phi(got) <= phi(p + np.float32(1.1 * a) * d)
    assert phi(got) <= phi(p + np.float32(0.9 * a) * d)
    # a non-finite direction gets step 0 and leaves the pressure as it was
    dd[0] = float("nan")
    torch.cuda.synchronize()
    _lib.call("nx_line_search", sim._alive(), C.c_void_p(dd.data_ptr()), out)
    assert out[0] == 0.0
    assert np.array_equal(sim.download("pressure"), got)
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


def test_damped_sweep_matches_float32_statement(gpu, rng):
    H, W = 72, 100
    sim = in_step_sim(rng, H, W)
    solid = sim.download("mask") != 0
    b = rng.normal(size=(H, W)).astype(np.float32)
    p = rng.normal(size=(H, W)).astype(np.float32)
    dev(sim, "divergence").copy_(torch.from_numpy(b))
    dev(sim, "pressure").copy_(torch.from_numpy(p))
    torch.cuda.synchronize()
    _lib.call("nx_damped_sweeps", sim._alive(), 3, 0.5)
    got = sim.download("pressure")
    w = np.float32(0.5)
    q = p
    for _ in range(3):
        P = np.pad(q, 1, mode="edge")
        S = np.pad(solid, 1, mode="edge")
        nb = []
        for dy, dx in ((0, -1), (0, 1), (-1, 0), (1, 0)):
            v = P[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
            s = S[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
            nb.append(np.where(s, q, v))
        jac = ((((nb[0] + nb[1]) + nb[2]) + nb[3]) - b) * np.float32(0.25)
        q = (q + w * (jac - q)).astype(np.float32)
    assert np.array_equal(got, q)
    with pytest.raises(ValueError):
        _lib.call("nx_damped_sweeps", sim._alive(), 1, 0.0)
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


def _tile(t):
    return None if t is None else (C.c_int * 4)(*t)


@pytest.mark.parametrize("H,W", [(72, 100), (131, 257)])
@pytest.mark.parametrize("tile,iters", [((8, 2, 16, 2), 1), ((8, 2, 16, 2), 7),
                                        ((8, 2, 16, 2), 8), ((8, 2, 16, 2), 40),
                                        ((5, 4, 8, 2), 40), ((3, 8, 8, 1), 17),
                                        ((12, 2, 16, 2), 30), (None, 40),
                                        ((10, 8, 16, 1), 25), ((6, 4, 16, 2), 33),
                                        ((9, 4, 16, 1), 29)])
def test_fine_chebyshev_matches_float32_statement(gpu, rng, H, W, tile, iters):
    """nx_smooth_chebyshev equals the float32 statement of the Chebyshev
    iteration (cheb32: the kernels' operation order and fused multiply-adds)
    bitwise, for every tile shape, odd and even passes per launch (float2
    and scalar accesses) and the default tile (None); the reported residual
    is the FP64 relative residual of the result."""
    lmin = 0.01
    sim = in_step_sim(rng, H, W)
    solid = sim.download("mask") != 0
    b = rng.normal(size=(H, W)).astype(np.float32)
    p = rng.normal(size=(H, W)).astype(np.float32)
    dev(sim, "divergence").copy_(torch.from_numpy(b))
    dev(sim, "pressure").copy_(torch.from_numpy(p))
    torch.cuda.synchronize()
    out = (C.c_double * 2)()
    _lib.call("nx_smooth_chebyshev", sim._alive(), iters, lmin, _tile(tile), out)
    got = sim.download("pressure")
    want = cheb32(p, b, solid, iters, lmin)
    assert np.array_equal(got, want)
    r = (b.astype(np.float64) - lap(got.astype(np.float64), solid))[~solid]
    bf = b.astype(np.float64)[~solid]
    n = r.size
    rel = np.sqrt((np.sum(r * r) - np.sum(r) ** 2 / n) / (np.sum(bf * bf) - np.sum(bf) ** 2 / n))
    assert abs(out[0] - rel) <= 1e-9 * rel
    for bad in ((0, lmin, None), (4, 0.0, None), (4, 8.0, None), (4, float("nan"), None),
                (4, lmin, _tile((4, 2, 8, 2))), (4, lmin, _tile((13, 2, 16, 2))),
                (4, lmin, _tile((0, 4, 8, 2)))):
        with pytest.raises(ValueError):
            _lib.call("nx_smooth_chebyshev", sim._alive(), *bad, None)
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


@pytest.mark.parametrize("H,W", [(72, 100), (131, 257), (5, 300), (300, 9)])
@pytest.mark.parametrize("cfg,iters", [((8, 4, 8, 2, 112), 40), ((8, 4, 8, 2, 1), 9),
                                       ((8, 4, 8, 2, 7), 17), ((5, 4, 8, 2, 40), 23),
                                       ((16, 2, 16, 2, 48), 40), ((8, 4, 16, 2, 80), 8),
                                       ((7, 8, 8, 2, 33), 1), ((16, 8, 8, 2, 1000), 50),
                                       ((2, 4, 8, 2, 16), 3), ((8, 8, 4, 2, 80), 20),
                                       ((6, 8, 4, 2, 33), 13)])
def test_fine_chebyshev_streaming_matches_float32_statement(gpu, rng, H, W, cfg, iters):
    """The streaming smoother (a workgroup walks a column strip over `seg`
    output rows, carrying the rows above each chunk per pass) equals the
    float32 statement bitwise: every shape, odd and even passes per launch,
    segments of one row to longer than the grid, grids thinner than a
    chunk or a strip; the reported residual is the FP64 relative residual
    of the result.  Unsupported streaming tiles are refused."""
    lmin = 0.01
    sim = in_step_sim(rng, H, W)
    solid = sim.download("mask") != 0
    b = rng.normal(size=(H, W)).astype(np.float32)
    p = rng.normal(size=(H, W)).astype(np.float32)
    dev(sim, "divergence").copy_(torch.from_numpy(b))
    dev(sim, "pressure").copy_(torch.from_numpy(p))
    torch.cuda.synchronize()
    _lib.call("nx_set_smoother_tile", sim._alive(), (C.c_int * 5)(*cfg), 5)
    got_cfg = (C.c_int * 5)()
    _lib.call("nx_get_smoother_tile", sim._alive(), got_cfg)
    assert tuple(got_cfg) == cfg
    out = (C.c_double * 2)()
    _lib.call("nx_smooth_chebyshev", sim._alive(), iters, lmin, None, out)
    got = sim.download("pressure")
    want = cheb32(p, b, solid, iters, lmin)
    assert np.array_equal(got, want)
    r = (b.astype(np.float64) - lap(got.astype(np.float64), solid))[~solid]
    bf = b.astype(np.float64)[~solid]
    n = r.size
    rel = np.sqrt((np.sum(r * r) - np.sum(r) ** 2 / n) / (np.sum(bf * bf) - np.sum(bf) ** 2 / n))
    assert abs(out[0] - rel) <= 1e-9 * rel
    for bad in ((8, 4, 8, 2, -1), (17, 4, 8, 2, 64), (8, 3, 8, 2, 64), (8, 4, 8, 1, 64),
                (0, 4, 8, 2, 64), (30, 4, 8, 2, 64)):
        with pytest.raises(ValueError):
            _lib.call("nx_set_smoother_tile", sim._alive(), (C.c_int * 5)(*bad), 5)
    with pytest.raises(ValueError):
        _lib.call("nx_set_smoother_tile", sim._alive(), (C.c_int * 5)(*cfg), 3)
    _lib.call("nx_get_smoother_tile", sim._alive(), got_cfg)
    assert tuple(got_cfg) == cfg  # a refused tile leaves the setting alone
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


def test_fine_chebyshev_damps_the_interval(gpu):
    """On a grid without obstacles the eigenvectors of -L are cosine modes
    with eigenvalue 4 sin^2(pi i / 2W) + 4 sin^2(pi j / 2H); n iterations
    scale a mode inside [lmin, 8] by at most 1 / T_n(sigma) and one below
    lmin by at most 1 (closed forms of the Chebyshev polynomial)."""
    H, W = 64, 96
    lmin, n = 0.05, 24
    sim = FluidSimulator(W, H)
    sim._push_params()
    _lib.call("nx_step_begin", sim._alive(), 1 / 60)
    sim.synchronize()
    Y, X = np.mgrid[0:H, 0:W]
    theta, delta = 0.5 * (8 + lmin), 0.5 * (8 - lmin)
    T = np.cosh(n * np.arccosh(theta / delta))
    for i, j in ((1, 0), (3, 2), (20, 10), (47, 31), (95, 63)):
        mode = np.cos(np.pi * i * (X + 0.5) / W) * np.cos(np.pi * j * (Y + 0.5) / H)
        lam = 4 * np.sin(np.pi * i / (2 * W)) ** 2 + 4 * np.sin(np.pi * j / (2 * H)) ** 2
        dev(sim, "divergence").zero_()
        dev(sim, "pressure").copy_(torch.from_numpy(mode.astype(np.float32)))
        torch.cuda.synchronize()
        _lib.call("nx_smooth_chebyshev", sim._alive(), n, lmin, None, None)
        got = sim.download("pressure").astype(np.float64)
        gain = np.sum(got * mode) / np.sum(mode * mode)
        # the closed form of the mode's factor: T_n((8 + lmin - 2 lam) / (8 - lmin)) / T_n(sigma)
        z = (8 + lmin - 2 * lam) / (8 - lmin)
        tz = np.cos(n * np.arccos(z)) if abs(z) <= 1 else np.sign(z) ** n * np.cosh(n * np.arccosh(abs(z)))
        assert abs(gain - tz / T) <= 1e-4 + 1e-3 * abs(tz / T), (i, j, gain, tz / T)
        if lam >= lmin:
            assert abs(gain) <= 1 / T + 1e-4
        assert abs(gain) <= 1 + 1e-4
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


def cylinder_sim(N, warm=3):
    sim = FluidSimulator(N, N, bc_mode="inflow")
    sim.vorticity = 0.3
    sim.viscosity = 0.0
    X = np.arange(N)[None, :]
    Y = np.arange(N)[:, None]
    sim.set_obstacle_map((X - 0.4 * N) ** 2 + (Y - 0.5 * N) ** 2 <= (0.1 * N) ** 2)
    for _ in range(warm):
        sim.update(1 / 60)
    return sim


class _Const(torch.nn.Module):
    def __init__(self, fn):
        super().__init__()
        self.fn = fn

    def forward(self, x):
        return self.fn(x.shape[:3], x.device)


@pytest.mark.parametrize("kind", ["nan", "huge"])
def test_learned_solve_survives_a_bad_network(gpu, kind):
    from natrix_amd.learned.solver import LearnedGuess, solve_step
    fn = {"nan": lambda s, dv: torch.full(s, float("nan"), device=dv),
          "huge": lambda s, dv: 1e6 * torch.randn(s, device=dv)}[kind]
    sim = cylinder_sim(128)
    st = solve_step(sim, LearnedGuess(_Const(fn), coarse=32), 1 / 60, tol=1e-4,
                    smooth=64, chunk=8192)
    assert st["converged"] and st["residual"] <= 1e-4
    if kind == "nan":
        # every line search took step 0: the network never moved the pressure
        assert st["learned_corrections"] == 0 and st["fallback"]
    # the step's pressure (ph + pl rounded to float32) is finite and the
    # oracle's residual of it is at the float32 floor of this grid
    from oracle import fluid as F
    p = sim.download("pressure")
    assert np.isfinite(p).all()
    res = F.relative_residual(p, sim.download("divergence"),
                              F.obstacles_from_mask(sim.download("mask")))
    assert res < 5e-4
    sim.destroy()


def test_trained_network_reduces_the_work_to_the_tolerance(gpu):
    """Unseen scene (cylinder at (0.4, 0.5) N, r = 0.1 N, 256^2): sweeps and
    time to FP64 relative residual <= 1e-4 with the trained network against
    the plain Jacobi correction solves and against the same cycle without the
    network (coarse smoothing from zero)."""
    import time
    from natrix_amd.learned.solver import LearnedGuess, solve_step
    from natrix_amd.learned.train import load
    net, meta = load()
    res = {}
    for name, guess in (("learned", LearnedGuess(net, meta.get("size", 128))),
                        ("no_net", LearnedGuess(None, meta.get("size", 128))),
                        ("plain", None)):
        sim = cylinder_sim(256, warm=5)
        solve_step(sim, guess, 1 / 60)  # first-call costs outside the clock
        sim.synchronize()
        t0 = time.perf_counter()
        st = solve_step(sim, guess, 1 / 60)
        sim.synchronize()
        res[name] = (st, time.perf_counter() - t0)
        assert st["converged"] and st["residual"] <= 1e-4, (name, st)
        sim.destroy()
    print({k: (v[0]["sweeps"], round(v[1] * 1e3, 2)) for k, v in res.items()})
    assert res["learned"][0]["sweeps"] * 3 <= res["plain"][0]["sweeps"]
    assert res["learned"][0]["sweeps"] <= res["no_net"][0]["sweeps"]


def test_native_unet_forward_equals_the_pytorch_forward(gpu):
    """The native runner launches the same kernels with the same K splits as
    the autograd wrapper: bitwise equal output on the trained weights."""
    from natrix_amd.learned.solver import NativeUNet
    from natrix_amd.learned.train import load
    from natrix_amd.learned.unet import pack_input
    net, meta = load()
    n = meta.get("size", 128)
    g = torch.Generator().manual_seed(11)
    rhs = torch.randn(1, n, n, generator=g).cuda()
    solid = (torch.rand(1, n, n, generator=g) < 0.1).cuda()
    x = pack_input(rhs, solid)
    with torch.no_grad():
        want = net(x).view(-1)
    runner = NativeUNet(net, n, n)
    got = torch.empty(n * n, device="cuda")
    runner.forward(x, got)
    torch.cuda.synchronize()
    assert torch.equal(got, want)


def test_native_correction_matches_the_pytorch_path(gpu):
    """One learned correction through nx_learned_correct and through the
    PyTorch path on the same state: the same step to within the float32
    rounding of the input normalisation (bf16 input)."""
    from natrix_amd.learned.solver import LearnedGuess
    from natrix_amd.learned.train import load
    net, meta = load()
    res = {}
    for native in (True, False):
        sim = cylinder_sim(256, warm=4)
        sim._push_params()
        _lib.call("nx_step_begin", sim._alive(), 1 / 60)
        p0 = sim.download("pressure")
        a = LearnedGuess(net, meta.get("size", 128), native=native).correct(sim, report=True)
        res[native] = (sim.download("pressure").astype(np.float64) - p0, a)
        _lib.call("nx_step_finish", sim._alive())
        sim.destroy()
    d_n, a_n = res[True]
    d_p, a_p = res[False]
    assert np.linalg.norm(d_n - d_p) <= 1e-2 * np.linalg.norm(d_p)
    assert abs(a_n[0] - a_p[0]) <= 1e-2 * abs(a_p[0])


def test_native_correction_graph_replays_equal_the_eager_call(gpu):
    """The first call runs the coarse pipeline eagerly, the second captures
    and launches it as a HIP graph, later calls replay it: from the same
    state every call gives the same pressure, bitwise, with eager work and
    synchronisations between the replays."""
    from natrix_amd.learned.solver import LearnedGuess
    from natrix_amd.learned.train import load
    net, meta = load()
    g = LearnedGuess(net, meta.get("size", 128))
    sim = cylinder_sim(256, warm=4)
    sim._push_params()
    _lib.call("nx_step_begin", sim._alive(), 1 / 60)
    p0 = sim.download("pressure")
    outs = []
    for k in range(5):
        sim.upload("pressure", p0)
        g.correct(sim)
        outs.append(sim.download("pressure"))
        # eager work on the same stream between replays
        _lib.call("nx_damped_sweeps", sim._alive(), 2, 0.5)
        sim.residual()
    assert not np.array_equal(outs[0], p0)
    for k in range(1, 5):
        assert np.array_equal(outs[k], outs[0]), k
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


def test_set_guess_routes_update_through_the_learned_step(gpu):
    """FluidSimulator.set_guess makes update() the learned step: the same
    result, bitwise, as solve_step on a twin simulator, and None restores
    the fixed-sweep step."""
    from natrix_amd.learned.solver import LearnedGuess, solve_step
    from natrix_amd.learned.train import load
    net, meta = load()
    g = LearnedGuess(net, meta.get("size", 128))
    a, b = cylinder_sim(256, warm=3), cylinder_sim(256, warm=3)
    a.set_guess(g, tol=1e-4)
    for _ in range(2):
        a.update(1 / 60)
        st = solve_step(b, g, 1 / 60, tol=1e-4)
    assert a.last_solve["converged"] and a.last_solve == st
    assert np.array_equal(a.download("velocity"), b.download("velocity"))
    a.set_guess(None)
    b.update(1 / 60)
    a.update(1 / 60)
    assert np.array_equal(a.download("velocity"), b.download("velocity"))
    for bad in (0.0, None, float("nan"), float("inf"), -1e-4):
        with pytest.raises(ValueError):
            a.set_guess(g, tol=bad)
    a.set_guess(None, tol=None)  # no guess: tol is ignored
    a.destroy()
    b.destroy()
    from natrix_amd import FluidSimulator
    for st in ("bf16", "f16"):
        m = FluidSimulator(64, 64, storage=st)
        with pytest.raises(ValueError, match="float32 storage"):
            m.set_guess(g)
        m.update(1 / 60)  # the simulator still steps
        m.destroy()


@pytest.mark.parametrize("W,H,grid", [(640, 360, (5, 128, 72)), (1000, 700, (8, 128, 88))])
def test_learned_step_on_a_window_shaped_grid(gpu, W, H, grid):
    """640 x 360 (the demo's velocity grid for a 1280 x 720 window): the
    coarse grid is 128 x 72 (f = 5, a factor that is not a power of two).
    1000 x 700: 125 x 88 coarse cells, padded to the network's 128 x 88
    grid.  The learned step reaches the tolerance with a third of plain
    Jacobi's sweeps or fewer."""
    from natrix_amd import FluidSimulator
    from natrix_amd.learned.solver import LearnedGuess, solve_step
    from natrix_amd.learned.train import load
    net, meta = load()
    g = LearnedGuess(net, meta.get("size", 128))
    assert g.coarse_grid(W, H) == grid
    res = {}
    for name, guess in (("learned", g), ("plain", None)):
        sim = FluidSimulator(W, H, bc_mode="inflow")
        sim.vorticity, sim.viscosity = 0.3, 0.0
        X = np.arange(W)[None, :]
        Y = np.arange(H)[:, None]
        sim.set_obstacle_map((X - 200) ** 2 + (Y - 170) ** 2 <= 36 ** 2)
        for _ in range(5):
            sim.update(1 / 60)
        st = solve_step(sim, guess, 1 / 60, tol=1e-4)
        assert st["converged"] and st["residual"] <= 1e-4, (name, st)
        res[name] = st["sweeps"]
        sim.destroy()
    print(res)
    assert res["learned"] * 3 <= res["plain"]


def test_padded_coarse_grid_correction_matches_the_pytorch_statement(gpu):
    """A 300 x 200 grid restricts onto 100 x 67 coarse cells, which the
    native runner pads to the network's 104 x 72 grid with solid cells of
    zero right-hand side.  The same correction stated in PyTorch from the
    public entry points (restriction onto the 100 x 67 grid, padding, the
    network, coarse smoothing on the padded grid, centring, interpolation
    and line search from the 100 x 67 part) gives the same step to within
    the float32 rounding of the input normalisation (bf16 input).  The
    PyTorch path itself refuses a padded grid."""
    from natrix_amd.learned.solver import LearnedGuess, pack_input
    from natrix_amd.learned.train import load
    net, meta = load()
    W, H = 300, 200
    g = LearnedGuess(net, meta.get("size", 128))
    f, wc, hc = g.coarse_grid(W, H)
    assert (f, wc, hc) == (3, 104, 72)
    wv, hv = -(-W // f), -(-H // f)

    def scene():
        sim = FluidSimulator(W, H, bc_mode="inflow")
        sim.vorticity, sim.viscosity = 0.3, 0.0
        X = np.arange(W)[None, :]
        Y = np.arange(H)[:, None]
        sim.set_obstacle_map((X - 90) ** 2 + (Y - 100) ** 2 <= 20 ** 2)
        for _ in range(4):
            sim.update(1 / 60)
        sim._push_params()
        _lib.call("nx_step_begin", sim._alive(), 1 / 60)
        return sim, sim.download("pressure")

    sim, p0 = scene()
    a_n = g.correct(sim, report=True)
    d_n = sim.download("pressure").astype(np.float64) - p0
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()

    sim, p0 = scene()
    with torch.no_grad(), torch.cuda.stream(torch.cuda.ExternalStream(sim.stream_handle())):
        rv = torch.empty(hv * wv, device="cuda")
        mv = torch.empty(hv * wv, dtype=torch.uint8, device="cuda")
        _lib.call("nx_restrict_residual", sim._alive(), f, C.c_void_p(rv.data_ptr()),
                  C.c_void_p(mv.data_ptr()))
        rc = torch.zeros(hc, wc, device="cuda")
        mc = torch.ones(hc, wc, dtype=torch.uint8, device="cuda")
        rc[:hv, :wv] = rv.view(hv, wv)
        mc[:hv, :wv] = mv.view(hv, wv)
        rc, mc = rc.view(-1), mc.view(-1)
        fluid = mc == 0
        n = fluid.sum()
        s = torch.sqrt((rc * rc * fluid).sum() / n)
        x = pack_input((rc / s).view(1, hc, wc), (~fluid).view(1, hc, wc))
        e = (net(x).view(-1) * s).contiguous()
        _lib.call("nx_coarse_smooth", sim._alive(), wc, hc, C.c_void_p(e.data_ptr()),
                  C.c_void_p(rc.data_ptr()), C.c_void_p(mc.data_ptr()),
                  int(g.coarse_iters), float(g.lmin))
        e = (e - (e * fluid).sum() / n) * fluid
        ev = e.view(hc, wc)[:hv, :wv].contiguous()
        mcv = mc.view(hc, wc)[:hv, :wv].contiguous()
        out = (C.c_double * 3)()
        _lib.call("nx_prolong_line_search", sim._alive(), f, C.c_void_p(ev.data_ptr()),
                  C.c_void_p(mcv.data_ptr()), out)
        a_s = tuple(out)
    d_s = sim.download("pressure").astype(np.float64) - p0
    with pytest.raises(ValueError):
        LearnedGuess(net, meta.get("size", 128), native=False).correct(sim)
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()
    assert a_s[0] > 0
    assert np.linalg.norm(d_n - d_s) <= 1e-2 * np.linalg.norm(d_s)
    assert abs(a_n[0] - a_s[0]) <= 1e-2 * abs(a_s[0])


@pytest.mark.parametrize("W,H", [(8, 8), (40, 24), (130, 70), (1024, 64), (333, 517)])
def test_learned_step_on_small_odd_and_narrow_grids(gpu, W, H):
    """The learned step reaches FP64 relative residual <= 1e-4 on grids the
    coarse-grid rule handles at its edges: the network's smallest grid
    (8 x 8, f = 1), odd sizes padded to multiples of 8, a 16:1 aspect ratio
    (a 128 x 8 coarse grid) and a portrait grid; the returned residual is
    the FP64 residual the simulator reports for its pressure."""
    from natrix_amd.learned.solver import LearnedGuess, solve_step
    from natrix_amd.learned.train import load
    net, meta = load()
    g = LearnedGuess(net, meta.get("size", 128))
    f, wc, hc = g.coarse_grid(W, H)
    assert wc % 8 == 0 and hc % 8 == 0 and wc <= 128 and hc <= 128
    assert -(-W // f) <= wc and -(-H // f) <= hc
    sim = FluidSimulator(W, H, bc_mode="inflow")
    sim.vorticity, sim.viscosity = 0.3, 0.0
    X = np.arange(W)[None, :]
    Y = np.arange(H)[:, None]
    r = max(1.0, 0.12 * min(W, H))
    sim.set_obstacle_map((X - 0.35 * W) ** 2 + (Y - 0.5 * H) ** 2 <= r * r)
    for _ in range(3):
        sim.update(1 / 60)
    for _ in range(2):
        st = solve_step(sim, g, 1 / 60, tol=1e-4)
        assert st["converged"] and st["residual"] <= 1e-4, st
    sim.destroy()


@pytest.mark.parametrize("fine_lmin", [None, 0.0])
def test_native_cycle_loop_equals_the_python_loop(gpu, fine_lmin):
    """nx_learned_cycles enqueues the same corrections and smoothing calls
    and takes the same decisions as the Python loop it replaces: the step's
    pressure, statistics and per-cycle trace are identical, with Chebyshev
    and with Jacobi fine smoothing."""
    from natrix_amd.learned.solver import LearnedGuess, solve_step
    from natrix_amd.learned.train import load
    net, meta = load()
    res = {}
    for native_loop in (True, False):
        sim = cylinder_sim(256, warm=4)
        g = LearnedGuess(net, meta.get("size", 128))
        tr = []
        st = solve_step(sim, g, 1 / 60, tol=1e-4, trace=tr, fine_lmin=fine_lmin,
                        native_loop=native_loop)
        assert st["converged"] and st["residual"] <= 1e-4, st
        res[native_loop] = (sim.download("pressure"), st, tr)
        sim.destroy()
    (p1, s1, t1), (p2, s2, t2) = res[True], res[False]
    assert np.array_equal(p1, p2)
    assert s1 == s2
    assert t1 == t2
    assert sum(1 for t in t1 if t[0] == "cycle") >= 2


def test_native_cycle_loop_arguments(gpu):
    """nx_learned_cycles rejects bad arguments before enqueueing anything
    and returns at once for zero cycles.  Regression cases: a record
    buffer smaller than 3 * max_cycles, an unsupported Chebyshev tile and a
    fine interval bound >= 8 are rejected up front (the records were once
    written with no length, the tile read and the interval checked only by
    the first smoothing call), and the binding refuses a Python integer in
    a pointer slot instead of handing it to the library as an address."""
    from natrix_amd.learned.solver import LearnedGuess
    from natrix_amd.learned.train import load
    net, meta = load()
    sim = cylinder_sim(128, warm=2)
    g = LearnedGuess(net, meta.get("size", 128))
    f, runner = g.native(sim)
    sim._push_params()
    _lib.call("nx_step_begin", sim._alive(), 1 / 60)
    before = sim.download("pressure")
    res = (C.c_int * 4)()
    rec = (C.c_double * 3)()
    bad_tile = (C.c_int * 4)(8, 3, 8, 2)
    args = [f, 100, float(g.lmin), 80, 0.3, None, 0.1, 0.5, 1, rec, 3, res]
    # factor, coarse iterations, smoothing iterations, fine interval, tile,
    # stall factor, cycle count, record capacity (too small for 2 cycles)
    for i, bad in ((0, 0), (1, -1), (3, 0), (4, 8.0), (4, float("nan")), (5, bad_tile),
                   (7, 0.0), (8, -1), (10, 2), (10, -1)):
        a = list(args)
        a[i] = bad
        with pytest.raises(ValueError):
            _lib.call("nx_learned_cycles", sim._alive(), runner.handle, *a)
    a = list(args)
    a[8] = 2  # two cycles need six record slots
    with pytest.raises(ValueError):
        _lib.call("nx_learned_cycles", sim._alive(), runner.handle, *a)
    with pytest.raises(ValueError):
        _lib.call("nx_learned_cycles", sim._alive(), runner.handle, *args[:-1], None)
    # an integer where a pointer belongs never reaches the library
    for i in (5, 9):
        a = list(args)
        a[i] = 4096
        with pytest.raises(C.ArgumentError):
            _lib.call("nx_learned_cycles", sim._alive(), runner.handle, *a)
    assert np.array_equal(sim.download("pressure"), before)
    a = list(args)
    a[8] = 0
    _lib.call("nx_learned_cycles", sim._alive(), runner.handle, *a)
    assert list(res) == [0, 0, 0, 0]
    assert np.array_equal(sim.download("pressure"), before)
    _lib.call("nx_step_finish", sim._alive())
    sim.destroy()


@pytest.mark.parametrize("n,count", [(128, 6), (37, 5)])
def test_fp64_batch_solve_matches_single_grid_cg(gpu, n, count):
    """nx_fp64_solve_batch (one workgroup per problem, the training
    targets) reaches the tolerance on every problem and agrees with the
    single-grid FP64 conjugate gradients of nx_fp64_solve on each (both
    mean-free over fluid cells); the float32 output is
    that solution rounded (within one float32 ulp of its magnitude, plus
    the condition number times the residuals for the difference of two FP64
    solves)."""
    from natrix_amd.learned.data import random_problem
    probs = [random_problem(np.random.default_rng(910000 + i), n) for i in range(count)]
    b = torch.from_numpy(np.stack([q[0] for q in probs])).cuda()
    m = torch.from_numpy(np.stack([q[1] for q in probs]).astype(np.uint8)).cuda()
    work = torch.empty(count * n * n, dtype=torch.float64, device="cuda")
    xf = torch.empty(count, n, n, dtype=torch.float32, device="cuda")
    res = torch.empty(count, dtype=torch.float64, device="cuda")
    its = torch.empty(count, dtype=torch.int32, device="cuda")
    p = lambda t: C.c_void_p(t.data_ptr())
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call("nx_fp64_solve_batch", p(b), p(m), n, count, p(work), p(xf), 1e-11, 20000,
              p(res), p(its), s)
    torch.cuda.synchronize()
    assert float(res.max()) <= 1e-10
    assert int(its.min()) > 0
    for i in range(count):
        x = torch.empty(n * n, dtype=torch.float64, device="cuda")
        out = (C.c_double * 2)()
        _lib.call("nx_fp64_solve", p(b[i]), p(m[i]), n, n, p(x), 1e-12, 100000, out, s)
        torch.cuda.synchronize()
        fluid = (m[i] == 0).view(-1)
        ref = x - x[fluid].mean()
        ref = torch.where(fluid, ref, torch.zeros_like(ref))
        got = xf[i].view(-1).double()
        assert torch.all(got[~fluid] == 0)
        err = float((got - ref).abs().max())
        # two FP64 solves to relative residuals 1e-11 and 1e-12 differ by at
        # most cond(A) (about 1.3e4 at 128^2) times 1e-10 relative
        assert err <= (2e-6 + 2.0 ** -23) * float(ref.abs().max()), (i, err)
    with pytest.raises(ValueError):
        _lib.call("nx_fp64_solve_batch", p(b), p(m), 129, count, p(work), p(xf), 1e-10, 100,
                  p(res), p(its), s)
