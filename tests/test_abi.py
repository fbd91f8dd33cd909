"""CPU checks of the C-ABI boundary: the library loads, exports every
function include/natrix.h declares, its parameter structs have the C
layout, and parameter validation matches the reference's setters
(natrix/core/fluid_simulator.py:58-111).  No kernel is launched."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from natrix_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "natrix.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|long long|const char\*)\s+(nx_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = _declared()
    assert len(names) >= 35
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes table covers the whole header and nothing else
    assert sorted(_lib.SIGNATURES) == names


def test_timing_tags_match_c_header():
    """The timing arrays the Python side passes to nx_collect_timing hold
    NX_TIME_COUNT entries, one per tag, in the header's order."""
    txt = open(HEADER).read()
    tags = dict((m[0], int(m[1])) for m in re.findall(r"^#define (NX_TIME_\w+) (\d+)", txt, re.M))
    assert tags.pop("NX_TIME_COUNT") == len(_lib.TIME_TAGS)
    assert sorted(tags.values()) == list(range(len(_lib.TIME_TAGS)))
    assert tags["NX_TIME_JACOBI"] == _lib.TIME_TAGS.index("jacobi")
    assert tags["NX_TIME_CHEB"] == _lib.TIME_TAGS.index("chebyshev")


def test_struct_layout_matches_c_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "natrix.h"
#define F(s, m) printf(#m " %zu\n", offsetof(s, m));
int main(void) {
  printf("params %zu\n", sizeof(nx_params));
  printf("stats %zu\n", sizeof(nx_stats));
  F(nx_params, speed) F(nx_params, iterations) F(nx_params, dye_in_step)
  F(nx_params, jacobi_waves) F(nx_params, tol) F(nx_params, max_iters)
  F(nx_params, check_every) F(nx_stats, residual) F(nx_stats, residual_sweeps)
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        cpath = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(cpath, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), cpath, "-o", exe],
                       check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict(l.split() for l in out.strip().splitlines())
    assert int(got["params"]) == C.sizeof(_lib.Params)
    assert int(got["stats"]) == C.sizeof(_lib.Stats)
    for f in ("speed", "iterations", "dye_in_step", "jacobi_waves", "tol",
              "max_iters", "check_every"):
        assert int(got[f]) == getattr(_lib.Params, f).offset, f
    for f in ("residual", "residual_sweeps"):
        assert int(got[f]) == getattr(_lib.Stats, f).offset, f


def _defaults():
    p = _lib.Params()
    _lib.call("nx_default_params", C.byref(p))
    return p


def test_defaults_match_reference():
    p = _defaults()
    # fluid_simulator.py:28-35
    assert (p.speed, p.iterations, p.dissipation, p.vorticity) == (500.0, 50, 1.0, 0.0)
    assert abs(p.viscosity - 0.1) < 1e-7 and p.has_borders == 1
    assert p.fused == 1 and p.tol == 0.0


@pytest.mark.parametrize("field,value,msg", [
    ("speed", 0.0, "'Speed' should be greater than zero"),
    ("iterations", 0, "'Iterations' should be grater than zero"),
    ("dissipation", -1.0, "'Dissipation' should be grater than zero"),
    ("vorticity", -0.5, "'Vorticity' should be grater or equal than zero"),
    ("viscosity", -0.1, "'Viscosity' should be greater or equal than zero"),
])
def test_c_validation_messages(field, value, msg):
    p = _defaults()
    assert _lib.lib().nx_check_params(C.byref(p)) == 0
    setattr(p, field, value)
    with pytest.raises(ValueError, match=re.escape(msg)):
        _lib.check(_lib.lib().nx_check_params(C.byref(p)))


def test_c_validation_jacobi_tile_and_solver():
    p = _defaults()
    p.jacobi_k = 70
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().nx_check_params(C.byref(p)))
    p = _defaults()
    p.tol = 1e-4
    p.fused = 0
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().nx_check_params(C.byref(p)))


def test_null_arguments_are_errors_not_crashes():
    lib = _lib.lib()
    assert lib.nx_destroy(None) == 0
    assert lib.nx_update(None, C.c_float(0.1)) == _lib.NX_ERR_ARG
    assert lib.nx_set_params(None, None) == _lib.NX_ERR_ARG
    assert lib.nx_create(0, 5, C.byref(C.c_void_p())) == _lib.NX_ERR_ARG
    assert "grid size" in _lib.last_error()


def test_python_property_validation_matches_reference():
    from natrix_amd.core.fluid_simulator import FluidSimulator
    s = FluidSimulator.__new__(FluidSimulator)  # properties only, no device
    assert (s.speed, s.iterations, s.dissipation, s.vorticity, s.viscosity) == \
        (500.0, 50, 1.0, 0.0, 0.1)
    assert s.has_borders is True and s.simulate is True
    cases = [("speed", 0, "'Speed' should be greater than zero"),
             ("iterations", -3, "'Iterations' should be grater than zero"),
             ("dissipation", 0.0, "'Dissipation' should be grater than zero"),
             ("vorticity", -1, "'Vorticity' should be grater or equal than zero"),
             ("viscosity", -1e-3, "'Viscosity' should be greater or equal than zero")]
    for name, bad, msg in cases:
        with pytest.raises(ValueError, match=re.escape(msg)):
            setattr(s, name, bad)
    s.vorticity = 0
    s.viscosity = 0.0
    s.speed = 250
    assert (s.vorticity, s.viscosity, s.speed) == (0, 0.0, 250)


def _header_arity():
    """name -> number of parameters of every function include/natrix.h
    declares (declarations may span lines; `(void)` has none)."""
    txt = open(HEADER).read()
    out = {}
    for m in re.finditer(r"^(?:int|long long|const char\*)\s+(nx_\w+)\(([^;]*?)\);", txt, re.M | re.S):
        params = " ".join(m.group(2).split())
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_binding_arity_matches_header():
    """Every ctypes signature has as many arguments as the C declaration:
    a stale table would shift arguments into the wrong slots."""
    ar = _header_arity()
    assert sorted(ar) == sorted(_lib.SIGNATURES)
    for name, args in _lib.SIGNATURES.items():
        assert len(args) == ar[name], name


def test_learned_cycles_binding_refuses_integers_as_pointers():
    """The tile, record and result slots of nx_learned_cycles are typed
    pointers: ctypes raises before the call when a Python integer lands
    there (a c_void_p slot would pass it on as an address).  Also the
    null-simulator check returns NX_ERR_ARG without touching the GPU."""
    lib = _lib.lib()
    res = (C.c_int * 4)()
    rec = (C.c_double * 3)()
    good = [None, None, 16, 100, 0.01, 72, 0.006, None, 0.1, 0.5, 1, rec, 3, res]
    for slot in (7, 11, 13):
        a = list(good)
        a[slot] = 4096
        with pytest.raises(C.ArgumentError):
            lib.nx_learned_cycles(*a)
    assert lib.nx_learned_cycles(*good) == _lib.NX_ERR_ARG
    assert "bad learned cycle arguments" in _lib.last_error()
