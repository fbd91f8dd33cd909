"""The allocation-free FP64 solve (nx_fp64_solve_noalloc) against the
allocating one (nx_fp64_solve): the same kernels on the library's static
device scratch instead of hipMalloc'd buffers, so the solution and the
reported residual and iteration count are bitwise equal.  Two calls in a
row, the second on a smaller grid (the scratch keeps the first call's
contents, which the solve must not depend on), and a grid beyond the
scratch's 512 x 512 cells refused before anything runs."""
import ctypes as C

import numpy as np
import pytest
import torch

from natrix_amd import _lib

pytestmark = pytest.mark.gpu


def _problem(rng, H, W):
    Y, X = np.mgrid[0:H, 0:W]
    solid = (X - 0.4 * W) ** 2 + (Y - 0.55 * H) ** 2 <= (0.15 * min(W, H)) ** 2
    b = rng.standard_normal((H, W)).astype(np.float32)
    b[solid] = 0.0
    return (torch.from_numpy(b).cuda(), torch.from_numpy(solid.astype(np.uint8)).cuda())


def _solve(name, b, m, H, W, tol, max_iter):
    x = torch.full((H * W,), float("nan"), dtype=torch.float64, device="cuda")
    out = (C.c_double * 2)()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call(name, C.c_void_p(b.data_ptr()), C.c_void_p(m.data_ptr()), W, H,
              C.c_void_p(x.data_ptr()), float(tol), int(max_iter), out, s)
    torch.cuda.synchronize()
    return x.cpu().numpy(), (out[0], out[1])


def test_noalloc_fp64_solve_equals_allocating_solve(gpu):
    rng = np.random.default_rng(42)
    for (H, W), tol in (((300, 410), 1e-10), ((96, 130), 1e-9)):
        b, m = _problem(rng, H, W)
        x_ref, o_ref = _solve("nx_fp64_solve", b, m, H, W, tol, 50000)
        x_na, o_na = _solve("nx_fp64_solve_noalloc", b, m, H, W, tol, 50000)
        assert o_ref[0] <= tol * 10 and o_ref[1] > 0
        assert o_na == o_ref, (H, W, o_na, o_ref)
        assert np.array_equal(x_na, x_ref), (H, W)
    b, m = _problem(rng, 513, 512)
    with pytest.raises(ValueError, match="exceeds its scratch"):
        _solve("nx_fp64_solve_noalloc", b, m, 513, 512, 1e-8, 10)
