"""Host AddressSanitizer build of the C-ABI library, for CPU-side checks.

    python -m natrix_amd.asan_build      # -> natrix_amd/_build_asan/libnatrix_hip_asan.so

Every csrc/*.hip file is compiled as in natrix_amd.build, with the host pass
instrumented by AddressSanitizer (the device code is built as usual; GPU
sanitizers are not used).  The library links the sanitizer runtime as a
shared object, so a process that loads it must preload that runtime
(``asan_runtime()``).  tests/test_asan_host.py loads it in a child process
and drives the argument checks and error paths of the entry points that run
on the host only (no GPU is needed or touched), so an out-of-bounds host
access there aborts the child with a report.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

from .build import FLAGS, HIPCC, sources

HERE = os.path.dirname(os.path.abspath(__file__))
OBJ = os.path.join(HERE, "_build_asan")
LIB = os.path.join(OBJ, "libnatrix_hip_asan.so")
# host-only instrumentation: each sanitizer flag applies to the host pass
HOST_ASAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]


def asan_runtime():
    """Path of the shared AddressSanitizer runtime of the ROCm clang."""
    root = os.path.dirname(os.path.dirname(HIPCC))
    hits = sorted(glob.glob(os.path.join(root, "llvm", "lib", "clang", "*", "lib", "linux",
                                         "libclang_rt.asan-x86_64.so")))
    if not hits:
        raise RuntimeError("no libclang_rt.asan-x86_64.so under " + root)
    return hits[-1]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    r = subprocess.run([HIPCC] + FLAGS + HOST_ASAN + ["-c", src, "-o", obj],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc (host ASan) failed for {src}:\n{r.stdout}{r.stderr}")
    return obj


def build():
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + HOST_ASAN +
                       ["-shared-libasan", "-o", tmp] + objs, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link (host ASan) failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
    sys.stdout.flush()
