"""ctypes binding of librcurves.so (the C ABI declared in include/rcurves.h).

The library is the only compute path: if it is missing this module raises
instead of falling back to any CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("RCURVES_LIB", os.path.join(PKG_DIR, "librcurves.so"))

RC_OK, RC_SINGULAR, RC_NONFINITE, RC_BAD_ARG, RC_HIP_ERROR, RC_UNSUPPORTED, RC_DIVERGED = range(7)
STATUS_NAMES = {0: "ok", 1: "singular", 2: "nonfinite", 3: "bad_arg", 4: "hip_error", 5: "unsupported",
                6: "diverged"}
MG_MAXLEVELS, MG_MAXHIST = 16, 64
BLOCK_BARYCENTER, BLOCK_TOTAL_LENGTH, BLOCK_EDGE_LENGTHS, BLOCK_POINT = range(4)
PRECISION_FP64, PRECISION_MIXED = 0, 1
MODE_ANALYTICAL, MODE_FINITE_DIFF = 0, 1
REASONS = {
    0: None,
    1: "zero_or_nonfinite_gradient_norm",
    2: "not_a_descent_direction",
    3: "armijo_failed",
    4: "projection_failed",
    5: "tau_below_min",
    6: "singular_system",
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class RcError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Params(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("epsilon", C.c_double)]


class Block(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("vertex", C.c_int32),
        ("x0", C.c_double * 3),
        ("L0", C.c_double),
        ("target", C.c_double * 3),
        ("ell0", _dp),
        ("ell0_len", C.c_int32),
        ("_pad", C.c_int32),
    ]


class Penalties(C.Structure):
    _fields_ = [
        ("w_length", C.c_double),
        ("w_diff", C.c_double),
        ("w_field", C.c_double),
        ("X", C.c_double * 3),
    ]


class StepOpts(C.Structure):
    _fields_ = [
        ("params", Params),
        ("has_energy_before", C.c_int32),
        ("precision", C.c_int32),
        ("energy_before", C.c_double),
        ("projection_frozen", C.c_int32),
        ("max_iter", C.c_int32),
        ("tol_abs", C.c_double),
        ("tol_rel", C.c_double),
        ("c1", C.c_double),
        ("shrink", C.c_double),
        ("tau0", C.c_double),
        ("tau_min", C.c_double),
        ("refine_tol", C.c_double),
        ("refine_max_iter", C.c_int32),
        ("collect_timings", C.c_int32),
        ("gradient_mode", C.c_int32),
        ("fd_h", C.c_double),
        ("saddle_refine", C.c_int32),
        ("refine_columns", C.c_int32),
        ("saddle_refine_tol", C.c_double),
        ("report_ir_residual", C.c_int32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("energy", C.c_double),
        ("energy_before", C.c_double),
        ("tau", C.c_double),
        ("residual", C.c_double),
        ("gradient_l2_norm", C.c_double),
        ("slope", C.c_double),
        ("phi_norm", C.c_double),
        ("refine_residual", C.c_double),
        ("accepted", C.c_int32),
        ("converged", C.c_int32),
        ("projection_iterations", C.c_int32),
        ("reason", C.c_int32),
        ("trials", C.c_int32),
        ("refine_iters", C.c_int32),
        ("saddle_ir_residual", C.c_double),
        ("saddle_ir_correction", C.c_double),
        ("saddle_ir_steps", C.c_int32),
        ("fp64_fallback", C.c_int32),
    ]


class Phase(C.Structure):
    _fields_ = [("ms", C.c_double), ("calls", C.c_int32), ("_pad", C.c_int32)]


class Timings(C.Structure):
    _fields_ = [(k, Phase) for k in
                ("dE", "energy", "assembleA", "saddle", "factor", "projection", "lineSearch", "step")]


class KernelProfile(C.Structure):
    _fields_ = [("syrk_ms", C.c_double), ("syrk_flop", C.c_double),
                ("syrk_launches", C.c_int32), ("_pad", C.c_int32)]


class MgLevel(C.Structure):
    _fields_ = [("n", C.c_int32), ("k", C.c_int32), ("C", C.POINTER(C.c_double)),
                ("Gp", C.POINTER(C.c_double)), ("parent", C.POINTER(C.c_int32)),
                ("weight", C.POINTER(C.c_double)), ("A", C.POINTER(C.c_double))]


class MgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_cycles", C.c_int32), ("nu_pre", C.c_int32),
                ("nu_post", C.c_int32), ("init_nested", C.c_int32)]


class MgStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("used_direct_cleanup", C.c_int32),
                ("nhist", C.c_int32), ("_pad", C.c_int32), ("residual", C.c_double),
                ("history", C.c_double * MG_MAXHIST)]


EXPORTS = (
    "rc_default_step_opts", "rc_ctx_create", "rc_ctx_destroy", "rc_last_error", "rc_energy",
    "rc_gradient", "rc_gradient_fd", "rc_step", "rc_energy_bh", "rc_gradient_bh",
    "rc_bct_matvec",
    "rc_bct_apply_b", "rc_assemble_a", "rc_penalties_eval", "rc_solve_constrained_gradient",
    "rc_constraint_rows", "rc_sobolev_step", "rc_set_positions", "rc_get_positions",
    "rc_sobolev_step_resident", "rc_kernel_profile_enable", "rc_kernel_profile_read",
    "rc_energy_noalloc", "rc_mg_solve", "rc_gram_apply",
)

_lib = None


def lib():
    """Load librcurves.so (raises when the HIP extension is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RcError(RC_UNSUPPORTED, f"HIP extension not built: {LIB_PATH} is missing "
                                      "(run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.rc_default_step_opts.argtypes = [C.POINTER(StepOpts)]
    L.rc_default_step_opts.restype = None
    L.rc_ctx_create.argtypes = [_ip, C.c_int32, C.c_int32, C.POINTER(vp)]
    L.rc_ctx_destroy.argtypes = [vp]
    L.rc_ctx_destroy.restype = None
    L.rc_last_error.argtypes = [vp]
    L.rc_last_error.restype = C.c_char_p
    L.rc_energy.argtypes = [vp, _dp, C.POINTER(Params), _dp]
    L.rc_energy_noalloc.restype = C.c_int
    L.rc_energy_noalloc.argtypes = [vp, C.c_int32, vp, C.c_int32, C.POINTER(Params), vp, vp]
    L.rc_gradient.argtypes = [vp, _dp, C.POINTER(Params), _dp, _dp]
    L.rc_gradient_fd.argtypes = [vp, _dp, C.POINTER(Params), C.c_double, _dp]
    L.rc_energy_bh.argtypes = [vp, _dp, C.POINTER(Params), C.c_double, C.c_double, C.c_int32,
                               _dp, C.POINTER(C.c_int64)]
    L.rc_gradient_bh.argtypes = [vp, _dp, C.POINTER(Params), C.c_double, C.c_double, C.c_int32,
                                 C.c_double, _dp]
    L.rc_bct_matvec.argtypes = [vp, _dp, C.c_double, _dp, C.c_int32, C.c_int32, C.c_double,
                                C.c_double, C.c_int32, _dp, C.POINTER(C.c_int64)]
    L.rc_bct_apply_b.argtypes = [vp, _dp, C.c_double, _dp, C.c_int32, C.c_double, C.c_double,
                                 C.c_int32, _dp]
    L.rc_step.argtypes = [vp, _dp, C.POINTER(Params), C.c_int32, C.c_double, C.c_double, _dp,
                          _dp, _dp]
    L.rc_assemble_a.argtypes = [vp, _dp, C.POINTER(Params), _dp]
    L.rc_penalties_eval.argtypes = [vp, _dp, C.POINTER(Penalties), _dp, _dp]
    L.rc_solve_constrained_gradient.argtypes = [vp, _dp, _dp, C.POINTER(Block), C.c_int32,
                                                C.POINTER(StepOpts), _dp, _dp, _dp]
    L.rc_constraint_rows.argtypes = [vp, C.POINTER(Block), C.c_int32]
    L.rc_constraint_rows.restype = C.c_int32
    L.rc_sobolev_step.argtypes = [vp, _dp, C.POINTER(Block), C.c_int32, C.POINTER(Penalties),
                                  C.POINTER(StepOpts), _dp, C.POINTER(StepStats), _dp,
                                  C.POINTER(Timings)]
    L.rc_set_positions.argtypes = [vp, _dp]
    L.rc_get_positions.argtypes = [vp, _dp]
    L.rc_sobolev_step_resident.argtypes = [vp, C.POINTER(Block), C.c_int32, C.POINTER(Penalties),
                                           C.POINTER(StepOpts), C.POINTER(StepStats),
                                           C.POINTER(Timings)]
    L.rc_mg_solve.argtypes = [vp, _dp, C.POINTER(Params), C.POINTER(Block), C.c_int32,
                              C.POINTER(MgLevel), C.c_int32, _dp, C.POINTER(MgOpts), _dp,
                              C.POINTER(MgStats)]
    L.rc_gram_apply.argtypes = [vp, _dp, C.POINTER(Params), _dp, _dp]
    L.rc_kernel_profile_enable.argtypes = [vp, C.c_int32]
    L.rc_kernel_profile_read.argtypes = [vp, C.POINTER(KernelProfile)]
    for name in ("rc_ctx_create", "rc_energy", "rc_gradient", "rc_gradient_fd", "rc_step",
                 "rc_energy_bh", "rc_gradient_bh", "rc_bct_matvec", "rc_bct_apply_b",
                 "rc_assemble_a", "rc_penalties_eval",
                 "rc_solve_constrained_gradient", "rc_sobolev_step", "rc_set_positions",
                 "rc_get_positions", "rc_sobolev_step_resident", "rc_kernel_profile_enable",
                 "rc_kernel_profile_read", "rc_mg_solve", "rc_gram_apply"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class Context:
    """One device topology (the worker's topology cache entry). Owns the
    rc_ctx; destroyed on close() or garbage collection."""

    def __init__(self, edges, num_vertices: int):
        self.edges = np.ascontiguousarray(np.asarray(edges, dtype=np.int32).reshape(-1, 2))
        self.V = int(num_vertices)
        self.E = len(self.edges)
        self._h = C.c_void_p()
        L = lib()
        st = L.rc_ctx_create(iptr(self.edges), self.E, self.V, C.byref(self._h))
        if st != RC_OK:
            msg = L.rc_last_error(self._h).decode() if self._h else "rc_ctx_create failed"
            if self._h:
                L.rc_ctx_destroy(self._h)
                self._h = C.c_void_p()
            raise RcError(st, msg)

    @property
    def handle(self):
        return self._h

    def check(self, st: int):
        if st != RC_OK:
            raise RcError(st, lib().rc_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().rc_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
