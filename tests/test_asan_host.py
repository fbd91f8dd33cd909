"""Host AddressSanitizer run of the C-ABI's host-side paths (CPU only).

The library is rebuilt with the host pass instrumented
(natrix_amd.asan_build) and loaded in a child process that preloads the
sanitizer runtime.  The child drives what runs on the host without a GPU:
parameter defaults and validation with every field out of range, the
argument checks of the learned-cycle, smoothing, U-Net, convolution,
batched-solve, batch and copy-probe entry points (null handles, bad
shapes, too-small record buffers), the thread-local error text from several
threads, and grid-size limits.  Any out-of-bounds or use-after-free host
access aborts the child with a sanitizer report, which fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import ctypes as C, sys, threading
sys.path.insert(0, ROOT)
from natrix_amd import _lib
lib = C.CDLL(LIB)
for name, args in _lib.SIGNATURES.items():
    fn = getattr(lib, name)
    fn.argtypes = args
    fn.restype = _lib._RESTYPE.get(name, C.c_int)
ARG = _lib.NX_ERR_ARG

def err():
    return lib.nx_last_error().decode()

p = _lib.Params()
assert lib.nx_default_params(C.byref(p)) == 0
assert lib.nx_check_params(C.byref(p)) == 0
bad = dict(speed=0.0, dissipation=0.0, vorticity=-1.0, viscosity=-1.0, iterations=0,
           bc_mode=7, jacobi_k=0, jacobi_rows=3, jacobi_cols=5, jacobi_seg=-2, tol=-1.0,
           max_iters=0, check_every=0, correction_sweeps=17, dye_speed=0.0,
           dye_dissipation=-1.0)
for field, value in bad.items():
    q = _lib.Params()
    lib.nx_default_params(C.byref(q))
    setattr(q, field, value)
    assert lib.nx_check_params(C.byref(q)) == ARG, field
    assert err(), field
# streaming Jacobi tile: supported shape accepted, k beyond 16 refused
q = _lib.Params(); lib.nx_default_params(C.byref(q))
q.jacobi_k, q.jacobi_rows, q.jacobi_waves, q.jacobi_cols, q.jacobi_seg = 8, 4, 8, 2, 112
assert lib.nx_check_params(C.byref(q)) == 0
q.jacobi_k = 17
assert lib.nx_check_params(C.byref(q)) == ARG

h = C.c_void_p()
for w, hh in ((0, 5), (5, 0), (1 << 15, 1 << 15)):
    assert lib.nx_create(w, hh, C.byref(h)) == ARG and not h.value
assert lib.nx_create_ex(8, 8, 9, C.byref(h)) == ARG
assert lib.nx_set_params(None, C.byref(p)) == ARG

res = (C.c_int * 4)()
rec = (C.c_double * 3)()
tile = (C.c_int * 4)(8, 4, 8, 2)
assert lib.nx_learned_cycles(None, None, 16, 100, 0.01, 72, 0.006, tile, 0.1, 0.5, 1, rec,
                             3, res) == ARG
assert lib.nx_smooth_chebyshev(None, 72, 0.006, None, None) == ARG
assert lib.nx_smooth(None, 10, 0.5, None) == ARG
t5 = (C.c_int * 5)(8, 4, 8, 2, 112)
assert lib.nx_set_smoother_tile(None, t5, 5) == ARG
assert lib.nx_get_smoother_tile(None, t5) == ARG

ptrs = (C.c_void_p * 14)(*([1] * 14))
u = C.c_void_p()
for H, W, base in ((7, 128, 32), (128, 136, 32), (64, 64, 24), (64, 64, 0)):
    assert lib.nx_unet_create(H, W, base, ptrs, ptrs, C.c_void_p(1), C.c_void_p(1),
                              C.byref(u)) == ARG
    assert not u.value
assert lib.nx_unet_create(64, 64, 32, None, ptrs, None, None, C.byref(u)) == ARG
assert lib.nx_unet_set_bypass(None, 1) == ARG
one = C.c_void_p(1)
assert lib.nx_conv3x3_fwd_split(one, one, one, one, None, 2, 1, 16, 16, 64, 64, 1,
                                None) == ARG
assert lib.nx_conv3x3_fwd_split(one, one, one, one, one, 99, 1, 16, 16, 64, 64, 1,
                                None) == ARG
assert lib.nx_fp64_solve_batch(one, one, 129, 4, one, one, 1e-6, 10, one, one, None) == ARG
assert lib.nx_fp64_solve_batch(one, one, 64, 0, one, one, 1e-6, 10, one, one, None) == ARG
assert lib.nx_fp64_solve(None, one, 8, 8, one, 1e-6, 10, None, None) == ARG
g = C.c_double()
assert lib.nx_copy_bandwidth_variant(15, 1, 0, C.byref(g)) == ARG
assert lib.nx_copy_bandwidth_variant(1 << 20, 1, 9, C.byref(g)) == ARG
widths = (C.c_int * 3)(64, 2, 64)
b = C.c_void_p()
assert lib.nx_batch_create(3, widths, widths, 0, C.byref(b)) == ARG and not b.value
assert lib.nx_batch_create(0, widths, widths, 0, C.byref(b)) == ARG

# the error text is per thread
def worker(i, out):
    q = _lib.Params()
    lib.nx_default_params(C.byref(q))
    q.iterations = 0 if i % 2 else 5
    q.speed = 1.0 if i % 2 else 0.0
    for _ in range(200):
        lib.nx_check_params(C.byref(q))
    out[i] = err()
out = [None] * 8
ts = [threading.Thread(target=worker, args=(i, out)) for i in range(8)]
[t.start() for t in ts]
[t.join() for t in ts]
assert len(set(out[0::2])) == 1 and len(set(out[1::2])) == 1 and out[0] != out[1], out
print("asan child ok")
'''


@pytest.fixture(scope="module")
def asan_lib():
    from natrix_amd import asan_build
    try:
        rt = asan_build.asan_runtime()
    except RuntimeError as e:
        pytest.skip(str(e))
    return asan_build.build(), rt


def test_host_paths_under_address_sanitizer(asan_lib):
    lib, rt = asan_lib
    env = dict(os.environ)
    env["LD_PRELOAD"] = rt
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    code = f"ROOT = {ROOT!r}\nLIB = {lib!r}\n" + CHILD
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "asan child ok" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]
