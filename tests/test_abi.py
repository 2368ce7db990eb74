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
    ("rc_step_stats", "StepStats"), ("rc_phase", "Phase"), ("rc_timings", "Timings"),
    ("rc_kernel_profile", "KernelProfile"), ("rc_params", "Params"),
    ("rc_mg_level", "MgLevel"), ("rc_mg_opts", "MgOpts"), ("rc_mg_stats", "MgStats"),
])
def test_struct_layouts_match_header(struct, cls):
    """Every ctypes mirror has the header struct's size and field offsets."""
    py = getattr(_lib, cls)
    names = [f[0] for f in py._fields_]
    layout = _c_layout(struct, names)
    assert layout[0] == ctypes.sizeof(py), (struct, layout[0], ctypes.sizeof(py))
    for name, off in zip(names, layout[1:]):
        assert getattr(py, name).offset == off, (struct, name, off)


def test_phase_ledger_keys():
    """SobolevPhaseKey (phaseTimings.ts:9-20) minus the keys that do not fire
    on the device: bHigh and bLow are one fused Gram kernel timed as
    assembleA, and expand never fires (as in the reference's flat path)."""
    reference = ["dE", "energy", "bHigh", "bLow", "assembleA", "expand", "saddle", "factor",
                 "projection", "lineSearch", "step"]
    fused = {"bHigh", "bLow", "expand"}
    assert [k for k, _ in _lib.Timings._fields_] == [
        "dE", "energy", "assembleA", "saddle", "factor", "projection", "lineSearch", "step"]
    assert {k for k, _ in _lib.Timings._fields_} == set(reference) - fused
    # PhaseSample {ms, calls} per key
    assert [f[0] for f in _lib.Phase._fields_][:2] == ["ms", "calls"]
    assert all(t is _lib.Phase for _, t in _lib.Timings._fields_)


def test_ts_binding_layouts_match_header():
    """ts/sobolevFfi.ts packs rc_step_opts, rc_step_stats, rc_timings and
    rc_block by byte offset (no Bun runtime here to run it): its constants
    must equal a compiled probe of the header."""
    import re

    src = open(os.path.join(ROOT, "repulsive-test2_amd", "ts", "sobolevFfi.ts")).read()

    def const(name):
        return int(re.search(rf"const {name} = (\d+);", src).group(1))

    def table(name):
        body = re.search(rf"const {name} = \{{(.*?)\}} as const;", src, re.S).group(1)
        return {k: int(v) for k, v in re.findall(r"(\w+): (\d+)", body)}

    sizes = {"rc_step_opts": "OPTS_BYTES", "rc_step_stats": "STATS_BYTES",
             "rc_timings": "TIMINGS_BYTES", "rc_phase": "PHASE_BYTES", "rc_block": "BLOCK_BYTES"}
    for struct, name in sizes.items():
        assert _c_layout(struct, [])[0] == const(name), struct
    params = {"alpha", "beta", "epsilon"}
    for struct, name in (("rc_step_opts", "OPT"), ("rc_step_stats", "STAT"), ("rc_block", "BLK")):
        tab = table(name)
        fields = [f for f in tab if f not in params]
        lay = _c_layout(struct, fields)
        for f, off in zip(fields, lay[1:]):
            assert tab[f] == off, (struct, f, tab[f], off)
        if name == "OPT":
            p = _c_layout("rc_params", ["alpha", "beta", "epsilon"])
            assert [tab[k] for k in ("alpha", "beta", "epsilon")] == p[1:]
    # the ledger keys in the order of rc_timings
    keys = re.search(r"const PHASE_KEYS = \[(.*?)\] as const;", src, re.S).group(1)
    assert re.findall(r"'(\w+)'", keys) == [k for k, _ in _lib.Timings._fields_]
