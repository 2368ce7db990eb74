"""CPU-side checks of the C-ABI boundary: the library loads and exports every
entry point include/rcurves.h declares (no device calls)."""
import ctypes
import os
import re

from rcurves import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "rcurves.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rc_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    for name in ("rc_ctx_create", "rc_energy", "rc_gradient", "rc_sobolev_step"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_lib.EXPORTS) == set(declared_symbols())


def test_default_options_match_reference_defaults():
    o = _lib.StepOpts()
    _lib.lib().rc_default_step_opts(ctypes.byref(o))
    assert (o.params.alpha, o.params.beta, o.params.epsilon) == (3.0, 6.0, 1e-10)
    assert (o.c1, o.shrink, o.tau0, o.tau_min) == (1e-4, 0.5, 1.0, 1e-12)
    assert (o.tol_abs, o.tol_rel, o.max_iter) == (1e-4, 1e-4, 8)
    assert o.projection_frozen == 1
    assert (o.gradient_mode, o.fd_h) == (_lib.MODE_ANALYTICAL, 1e-6)  # optimizer.ts:18


def test_struct_sizes_match_header_layout():
    # rc_block: 2 ints, 3 + 1 + 3 doubles, pointer, 2 ints
    assert ctypes.sizeof(_lib.Block) == 8 + 56 + 8 + 8
    assert ctypes.sizeof(_lib.StepStats) == 8 * 8 + 6 * 4
    assert ctypes.sizeof(_lib.KernelProfile) == 8 + 8 + 4 + 4
    # rc_step_opts: params, 2 ints, 1 double, 2 ints, 7 doubles, 2 ints, int + pad, double
    assert ctypes.sizeof(_lib.StepOpts) == 24 + 8 + 8 + 8 + 56 + 8 + 8 + 8
