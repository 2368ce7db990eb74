"""ctypes view of the CPU FP64 oracle (oracle/rc_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library. tests/, the smoke
check in __graft_entry__ and the CPU-baseline leg of bench.py are the only
callers; the product path never imports this module.

Vertices are (V, 3) float64 arrays (the reference's Vec3[]), edges (E, 2) int32.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librc_oracle.so")

BARYCENTER, TOTAL_LENGTH, EDGE_LENGTHS, POINT = 0, 1, 2, 3
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
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("epsilon", C.c_double),
        ("has_energy_before", C.c_int32),
        ("projection_frozen", C.c_int32),
        ("energy_before", C.c_double),
        ("factor_lu", C.c_int32),
        ("max_iter", C.c_int32),
        ("tol_abs", C.c_double),
        ("tol_rel", C.c_double),
        ("c1", C.c_double),
        ("shrink", C.c_double),
        ("tau0", C.c_double),
        ("tau_min", C.c_double),
        ("penalties", C.POINTER(Penalties)),
        ("gradient_mode", C.c_int32),
        ("_pad2", C.c_int32),
        ("fd_h", C.c_double),
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
        ("accepted", C.c_int32),
        ("converged", C.c_int32),
        ("projection_iterations", C.c_int32),
        ("reason", C.c_int32),
        ("trials", C.c_int32),
        ("_pad", C.c_int32),
    ]

    def as_dict(self) -> dict:
        d = {f[0]: getattr(self, f[0]) for f in self._fields_ if not f[0].startswith("_")}
        d["reason"] = REASONS[self.reason]
        d["accepted"] = bool(self.accepted)
        d["converged"] = bool(self.converged)
        if self.projection_iterations < 0:
            d["projection_iterations"] = None
        return d


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        i, d = C.c_int, C.c_double
        L.rco_energy.restype = d
        L.rco_energy.argtypes = [i, i, _dp, _ip, d, d, d]
        L.rco_energy_compensated.restype = d
        L.rco_energy_compensated.argtypes = [i, i, _dp, _ip, d, d, d]
        L.rco_gradient.argtypes = [i, i, _dp, _ip, d, d, d, _dp]
        L.rco_gradient_fd.argtypes = [i, i, _dp, _ip, d, d, d, d, _dp]
        L.rco_bct_matvec.argtypes = [i, i, _dp, _ip, d, _dp, i, i, d, d, i, _dp,
                                     C.POINTER(C.c_int64)]
        L.rco_bct_apply_b.argtypes = [i, i, _dp, _ip, d, _dp, i, d, d, i, _dp]
        L.rco_gradient_bh_fd.argtypes = [i, i, _dp, _ip, d, d, d, d, d, i, d, _dp]
        L.rco_energy_bh.restype = d
        L.rco_energy_bh.argtypes = [i, i, _dp, _ip, d, d, d, d, d, i, C.POINTER(C.c_int64)]
        L.rco_step.argtypes = [i, i, _dp, _ip, d, d, d, i, d, d, _dp, _dp, _dp]
        for nm in ("rco_assemble_b_high", "rco_assemble_b_low", "rco_assemble_a", "rco_assemble_a_ld"):
            getattr(L, nm).argtypes = [i, i, _dp, _ip, d, d, d, _dp]
        L.rco_barycenter_target.argtypes = [i, i, _dp, _ip, _dp]
        L.rco_total_length.restype = d
        L.rco_total_length.argtypes = [i, i, _dp, _ip]
        L.rco_constraint_rows.argtypes = [i, i, i, C.POINTER(Block)]
        L.rco_eval_constraints.argtypes = [i, i, _dp, _ip, i, C.POINTER(Block), _dp, _dp, _dp, _ip]
        L.rco_penalties_active.argtypes = [C.POINTER(Penalties)]
        L.rco_penalty_energy.restype = d
        L.rco_penalty_energy.argtypes = [i, i, _dp, _ip, C.POINTER(Penalties)]
        L.rco_penalty_gradient.argtypes = [i, i, _dp, _ip, C.POINTER(Penalties), _dp]
        L.rco_ldlt_factor.argtypes = [_dp, i, _ip]
        L.rco_ldlt_solve.argtypes = [_dp, _ip, i, _dp]
        L.rco_lu_factor.argtypes = [_dp, i, _ip]
        L.rco_lu_solve.argtypes = [_dp, _ip, i, _dp, _dp]
        L.rco_solve_saddle_from_a.argtypes = [_dp, i, _dp, i, _dp, _dp, i, _dp, _dp, _dp]
        L.rco_l2_curve_norm.restype = d
        L.rco_l2_curve_norm.argtypes = [i, i, _dp, _dp, _ip]
        L.rco_solve_constrained_gradient.argtypes = [
            i, i, _dp, _ip, d, d, d, _dp, i, C.POINTER(Block), i, _dp, _dp, _dp]
        L.rco_line_search.argtypes = [
            i, i, _dp, _ip, _dp, _dp, i, C.POINTER(Block), C.POINTER(StepOpts), _dp,
            C.POINTER(StepStats)]
        L.rco_sobolev_step.argtypes = [
            i, i, _dp, _ip, i, C.POINTER(Block), C.POINTER(StepOpts), _dp, _dp,
            C.POINTER(StepStats)]
        L.rco_default_opts.argtypes = [C.POINTER(StepOpts)]
        L.rco_saddle_residual_dd.argtypes = [_dp, i, _dp, i, _dp, _dp, _dp, _dp, _dp, _dp]
        L.rco_set_threads.argtypes = [i]
        L.rco_threads.restype = i
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Threads of the pair loops (0 = OpenMP default); results do not depend on it."""
    lib().rco_set_threads(int(n))


def threads() -> int:
    return lib().rco_threads()


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _verts(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1, 3))


def _edges(e):
    return np.ascontiguousarray(np.asarray(e, dtype=np.int32).reshape(-1, 2))


def energy(v, e, alpha=3.0, beta=6.0, eps=1e-10, compensated=False):
    v, e = _verts(v), _edges(e)
    f = lib().rco_energy_compensated if compensated else lib().rco_energy
    return f(len(v), len(e), _d(v), _i(e), alpha, beta, eps)


def gradient(v, e, alpha=3.0, beta=6.0, eps=1e-10):
    v, e = _verts(v), _edges(e)
    g = np.zeros_like(v)
    lib().rco_gradient(len(v), len(e), _d(v), _i(e), alpha, beta, eps, _d(g))
    return g


def gradient_fd(v, e, alpha=3.0, beta=6.0, eps=1e-10, h=1e-6):
    v, e = _verts(v), _edges(e)
    g = np.zeros_like(v)
    lib().rco_gradient_fd(len(v), len(e), _d(v), _i(e), alpha, beta, eps, h, _d(g))
    return g


def energy_bh(v, e, theta, theta_t=None, leaf_size=8, alpha=3.0, beta=6.0, eps=1e-10):
    """Barnes-Hut energy and (visits, approximated edge refs, direct pairs)."""
    v, e = _verts(v), _edges(e)
    st = (C.c_int64 * 3)()
    tt = 1.5 * theta if theta_t is None else theta_t
    en = lib().rco_energy_bh(len(v), len(e), _d(v), _i(e), alpha, beta, eps, theta, tt,
                             leaf_size, st)
    return en, (st[0], st[1], st[2])


def gradient_bh_fd(v, e, theta, theta_t=None, leaf_size=8, h=1e-6, alpha=3.0, beta=6.0,
                   eps=1e-10):
    v, e = _verts(v), _edges(e)
    g = np.zeros_like(v)
    tt = 1.5 * theta if theta_t is None else theta_t
    lib().rco_gradient_bh_fd(len(v), len(e), _d(v), _i(e), alpha, beta, eps, theta, tt,
                             leaf_size, h, _d(g))
    return g


def bct_matvec(v, e, psi, kind, theta, theta_t=None, leaf_size=8, eps=1e-10):
    """K psi through the block cluster tree; kind 'high' | 'low'."""
    v, e = _verts(v), _edges(e)
    psi = np.ascontiguousarray(np.asarray(psi, dtype=np.float64))
    nc = 1 if psi.ndim == 1 else psi.shape[1]
    phi = np.zeros_like(psi)
    st = (C.c_int64 * 4)()
    tt = 1.5 * theta if theta_t is None else theta_t
    lib().rco_bct_matvec(len(v), len(e), _d(v), _i(e), eps, _d(psi), nc,
                         0 if kind == "high" else 1, theta, tt, leaf_size, _d(phi), st)
    return phi, (st[0], st[1], st[2], st[3])


def bct_apply_b(v, e, u, kind, theta, theta_t=None, leaf_size=8, eps=1e-10):
    v, e = _verts(v), _edges(e)
    u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    out = np.zeros(len(v))
    tt = 1.5 * theta if theta_t is None else theta_t
    lib().rco_bct_apply_b(len(v), len(e), _d(v), _i(e), eps, _d(u), 0 if kind == "high" else 1,
                          theta, tt, leaf_size, _d(out))
    return out


MODE_ANALYTICAL, MODE_FINITE_DIFF = 0, 1


def step(v, e, mode, step_size, alpha=3.0, beta=6.0, eps=1e-10, h=1e-6):
    """Raw L2 step: (vertices, energy at them, dE at the input)."""
    v, e = _verts(v), _edges(e)
    out = np.zeros_like(v)
    fld = np.zeros_like(v)
    en = C.c_double(0)
    lib().rco_step(len(v), len(e), _d(v), _i(e), alpha, beta, eps, mode, step_size, h, _d(out),
                   C.byref(en), _d(fld))
    return out, en.value, fld


def assemble(kind, v, e, alpha=3.0, beta=6.0, eps=1e-10):
    v, e = _verts(v), _edges(e)
    n = len(v)
    M = np.zeros((n, n))
    getattr(lib(), {"B": "rco_assemble_b_high", "B0": "rco_assemble_b_low", "A": "rco_assemble_a",
                    "A_ld": "rco_assemble_a_ld"}[kind])(
        n, len(e), _d(v), _i(e), alpha, beta, eps, _d(M))
    return M


def barycenter_target(v, e):
    v, e = _verts(v), _edges(e)
    x0 = np.zeros(3)
    lib().rco_barycenter_target(len(v), len(e), _d(v), _i(e), _d(x0))
    return x0


def total_length(v, e):
    v, e = _verts(v), _edges(e)
    return lib().rco_total_length(len(v), len(e), _d(v), _i(e))


@dataclass
class ConstraintSet:
    """POD mirror of the reference ConstraintSet (constraintSet.ts:39-52)."""

    blocks: list = field(default_factory=list)
    _keep: list = field(default_factory=list)

    def barycenter(self, x0):
        b = Block(kind=BARYCENTER)
        b.x0[:] = list(map(float, x0))
        self.blocks.append(b)
        return self

    def total_length(self, L0):
        self.blocks.append(Block(kind=TOTAL_LENGTH, L0=float(L0)))
        return self

    def edge_lengths(self, ell0):
        arr = np.ascontiguousarray(np.asarray(ell0, dtype=np.float64))
        self._keep.append(arr)
        self.blocks.append(Block(kind=EDGE_LENGTHS, ell0=_d(arr), ell0_len=len(arr)))
        return self

    def point(self, vertex, target):
        b = Block(kind=POINT, vertex=int(vertex))
        b.target[:] = list(map(float, target))
        self.blocks.append(b)
        return self

    def array(self):
        arr = (Block * max(1, len(self.blocks)))(*self.blocks)
        return arr, len(self.blocks)


def eval_constraints(v, e, cs: ConstraintSet):
    v, e = _verts(v), _edges(e)
    arr, nb = cs.array()
    k = lib().rco_constraint_rows(len(v), len(e), nb, arr)
    phi = np.zeros(max(k, 1))
    Cm = np.zeros((max(k, 1), 3 * len(v)))
    scales = np.zeros(max(nb, 1))
    rows = np.zeros(max(nb, 1), dtype=np.int32)
    lib().rco_eval_constraints(len(v), len(e), _d(v), _i(e), nb, arr, _d(phi), _d(Cm), _d(scales), _i(rows))
    return phi[:k], Cm[:k], scales[:nb], rows[:nb]


def make_penalties(w_length=0.0, w_diff=0.0, w_field=0.0, X=(0.0, 0.0, 0.0)):
    p = Penalties(w_length=w_length, w_diff=w_diff, w_field=w_field)
    p.X[:] = list(map(float, X))
    return p


def penalty_energy(v, e, pen):
    v, e = _verts(v), _edges(e)
    return lib().rco_penalty_energy(len(v), len(e), _d(v), _i(e), C.byref(pen))


def penalty_gradient(v, e, pen):
    v, e = _verts(v), _edges(e)
    g = np.zeros_like(v)
    lib().rco_penalty_gradient(len(v), len(e), _d(v), _i(e), C.byref(pen), _d(g))
    return g


def l2_curve_norm(values, v, e):
    values, v, e = _verts(values), _verts(v), _edges(e)
    return lib().rco_l2_curve_norm(len(v), len(e), _d(values), _d(v), _i(e))


def ldlt(K):
    K = np.ascontiguousarray(np.array(K, dtype=np.float64))
    n = K.shape[0]
    ipiv = np.zeros(max(n, 1), dtype=np.int32)
    rc = lib().rco_ldlt_factor(_d(K), n, _i(ipiv))
    return rc, K, ipiv[:n]


def ldlt_solve(m, ipiv, b):
    m = np.ascontiguousarray(m)
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    x = np.ascontiguousarray(np.array(b, dtype=np.float64))
    lib().rco_ldlt_solve(_d(m), _i(ipiv), m.shape[0], _d(x))
    return x


def saddle_residual_dd(A, Cm, r_top, r_bot, x, lam):
    """rhs - K z of the reference saddle system in double-double arithmetic
    (rounded once): (rho_top[3V], rho_bot[k]). A (V, V), Cm (k, 3V), x and
    r_top coordinate-block 3V vectors."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    V = A.shape[0]
    Cm = np.ascontiguousarray(np.asarray(Cm, dtype=np.float64).reshape(-1, 3 * V))
    k = Cm.shape[0]
    rt = np.ascontiguousarray(np.asarray(r_top, dtype=np.float64).reshape(3 * V))
    rb = np.ascontiguousarray(np.asarray(r_bot, dtype=np.float64).reshape(max(k, 1))[:max(k, 1)]
                              if k else np.zeros(1))
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(3 * V))
    lam = np.ascontiguousarray(np.asarray(lam, dtype=np.float64).reshape(-1)) if k else np.zeros(1)
    ot = np.zeros(3 * V)
    ob = np.zeros(max(k, 1))
    lib().rco_saddle_residual_dd(_d(A), V, _d(Cm) if k else _d(np.zeros(1)), k, _d(rt), _d(rb),
                                 _d(x), _d(lam), _d(ot), _d(ob))
    return ot, ob[:k]


def default_opts(**kw) -> StepOpts:
    o = StepOpts()
    lib().rco_default_opts(C.byref(o))
    pen = kw.pop("penalties", None)
    for k, val in kw.items():
        setattr(o, k, val)
    if pen is not None:
        o._pen = pen  # keep alive
        o.penalties = C.pointer(pen)
    return o


def solve_constrained_gradient(v, e, dE, cs: ConstraintSet, alpha=3.0, beta=6.0, eps=1e-10, factor_lu=0):
    v, e, dE = _verts(v), _edges(e), _verts(dE)
    arr, nb = cs.array()
    k = lib().rco_constraint_rows(len(v), len(e), nb, arr)
    gt = np.zeros_like(v)
    lam = np.zeros(max(k, 1))
    res = C.c_double(0)
    rc = lib().rco_solve_constrained_gradient(len(v), len(e), _d(v), _i(e), alpha, beta, eps, _d(dE),
                                             nb, arr, factor_lu, _d(gt), _d(lam), C.byref(res))
    return rc, gt, lam[:k], res.value


def line_search(v, e, dE, gt, cs: ConstraintSet, opts: StepOpts):
    v, e, dE, gt = _verts(v), _edges(e), _verts(dE), _verts(gt)
    arr, nb = cs.array()
    out = np.zeros_like(v)
    st = StepStats()
    lib().rco_line_search(len(v), len(e), _d(v), _i(e), _d(dE), _d(gt), nb, arr, C.byref(opts),
                          _d(out), C.byref(st))
    return out, st.as_dict()


def sobolev_step(v, e, cs: ConstraintSet, opts: StepOpts):
    v, e = _verts(v), _edges(e)
    arr, nb = cs.array()
    out = np.zeros_like(v)
    gt = np.zeros_like(v)
    st = StepStats()
    lib().rco_sobolev_step(len(v), len(e), _d(v), _i(e), nb, arr, C.byref(opts), _d(out), _d(gt),
                           C.byref(st))
    return out, gt, st.as_dict()
