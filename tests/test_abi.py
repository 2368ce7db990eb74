"""CPU-side checks of the C-ABI boundary: the library loads and exports every
entry point include/rcurves.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

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


def _c_layout(struct, fields):
    """sizeof and offsetof of a header struct, from a C probe compiled here."""
    import subprocess
    import tempfile

    body = [f'printf("%zu\\n", sizeof({struct}));']
    body += [f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    src = ("#include <stddef.h>\n#include <stdio.h>\n#include \"rcurves.h\"\n"
           "int main(void) {\n" + "\n".join(body) + "\nreturn 0;\n}\n")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        with open(c, "w") as f:
            f.write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return [int(x) for x in out]


@pytest.mark.parametrize("struct,cls", [
    ("rc_block", "Block"), ("rc_penalties", "Penalties"), ("rc_step_opts", "StepOpts"),
    ("rc_step_stats", "StepStats"), ("rc_timings", "Timings"),
    ("rc_kernel_profile", "KernelProfile"), ("rc_params", "Params"),
])
def test_struct_layouts_match_header(struct, cls):
    """Every ctypes mirror has the header struct's size and field offsets."""
    py = getattr(_lib, cls)
    names = [f[0] for f in py._fields_]
    layout = _c_layout(struct, names)
    assert layout[0] == ctypes.sizeof(py), (struct, layout[0], ctypes.sizeof(py))
    for name, off in zip(names, layout[1:]):
        assert getattr(py, name).offset == off, (struct, name, off)
