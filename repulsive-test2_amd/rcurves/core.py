"""Host-side mirror of the reference's src/core energy / gradient / Sobolev-step
API over the HIP C ABI (librcurves.so).

Names, argument meaning and return shapes follow the reference:
  calculate_disjoint_pairs  src/core/tangentPointEnergy.ts:33-47
  calculate_energy          src/core/tangentPointEnergy.ts:50-111
  gradient_analytical       src/core/tangentPointEnergy.ts:139-392
  assemble_a                src/core/sobolev/innerProduct.ts:325-350
  barycenter_block & co.    src/core/sobolev/constraintSet.ts:87-301
  barycenter_target         src/core/sobolev/constraints.ts:23-52
  penalty_energy/_gradient  src/core/sobolev/penalties.ts:139-241
  solve_constrained_gradient_set_frozen  src/core/sobolev/gradient.ts:84-115
  sobolev_step_set          src/core/optimizer.ts:231-440
  sobolev_step              src/core/optimizer.ts:455-469

Vertices are (V, 3) arrays (Vec3[]), edges (E, 2) int arrays (Edge[]). The
disjoint-pair lists are accepted for signature parity and checked for shape
only: the device derives disjointness from edge endpoints, exactly as
calculate_disjoint_pairs defines it.
"""
from __future__ import annotations

import ctypes as C
from collections import OrderedDict
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib as L

DEFAULTS = {"alpha": 3.0, "beta": 6.0, "epsilon": 1e-10, "h": 1e-6}  # optimizer.ts:18
SOBOLEV_CONVERGENCE_TOL = 1e-4  # optimizer.ts:78

_CTX_CACHE: "OrderedDict[tuple, L.Context]" = OrderedDict()
_CTX_CACHE_MAX = 4


def _verts(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1, 3))


def _edges(e) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(e, dtype=np.int32).reshape(-1, 2))


def context_for(edges, num_vertices: int) -> L.Context:
    """Device topology for (edges, V), cached like the worker's per-topology
    cache (src/worker/solverWorker.ts:32-43)."""
    e = _edges(edges)
    key = (int(num_vertices), e.tobytes())
    ctx = _CTX_CACHE.get(key)
    if ctx is not None:
        _CTX_CACHE.move_to_end(key)
        return ctx
    ctx = L.Context(e, num_vertices)
    _CTX_CACHE[key] = ctx
    while len(_CTX_CACHE) > _CTX_CACHE_MAX:
        _, old = _CTX_CACHE.popitem(last=False)
        old.close()
    return ctx


def calculate_disjoint_pairs(edges) -> list:
    """Ordered disjoint partner lists (tangentPointEnergy.ts:33-47). Host-only
    topology helper; the device never materialises these lists."""
    e = _edges(edges)
    a1, a2 = e[:, 0][:, None], e[:, 1][:, None]
    b1, b2 = e[:, 0][None, :], e[:, 1][None, :]
    ok = (a1 != b1) & (a1 != b2) & (a2 != b1) & (a2 != b2)
    np.fill_diagonal(ok, False)
    return [np.nonzero(ok[i])[0].tolist() for i in range(len(e))]


def _check_pairs(disjoint_pairs, E: int) -> None:
    if disjoint_pairs is not None and len(disjoint_pairs) != E:
        raise ValueError(f"disjointPairs has {len(disjoint_pairs)} rows, expected {E}")


def _params(alpha, beta, epsilon) -> L.Params:
    return L.Params(float(alpha), float(beta), float(epsilon))


def calculate_energy(vertices, edges, disjoint_pairs, alpha, beta, epsilon) -> float:
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    _check_pairs(disjoint_pairs, ctx.E)
    out = C.c_double()
    ctx.check(L.lib().rc_energy(ctx.handle, L.dptr(v), C.byref(_params(alpha, beta, epsilon)),
                                C.byref(out)))
    return out.value


def gradient_analytical(vertices, edges, disjoint_pairs, alpha, beta, epsilon,
                        with_energy: bool = False):
    """dE/dgamma (ascent orientation). with_energy=True also returns E from the
    same device pass (the forward/backward pair)."""
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    _check_pairs(disjoint_pairs, ctx.E)
    g = np.zeros_like(v)
    e = C.c_double()
    ctx.check(L.lib().rc_gradient(ctx.handle, L.dptr(v), C.byref(_params(alpha, beta, epsilon)),
                                  L.dptr(g), C.byref(e) if with_energy else None))
    return (g, e.value) if with_energy else g


def gradient_finite_diff(vertices, edges, disjoint_pairs, alpha, beta, epsilon, h) -> np.ndarray:
    """gradientFiniteDiff (tangentPointEnergy.ts:114-136): the forward
    difference (E(x + h e_vd) - E(x)) / h for every vertex and coordinate."""
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    _check_pairs(disjoint_pairs, ctx.E)
    g = np.zeros_like(v)
    ctx.check(L.lib().rc_gradient_fd(ctx.handle, L.dptr(v),
                                     C.byref(_params(alpha, beta, epsilon)), float(h), L.dptr(g)))
    return g


def bh_energy(vertices, edges, alpha, beta, epsilon, theta=0.5, leaf_size=8, theta_t=None):
    """Barnes-Hut energy (the reference's stage-2 oracle bh_energy,
    oracle/tpe_stage2_oracle.py:532-596): (energy, {visits,
    approximated_edge_refs, direct_pairs}); theta_t defaults to 1.5 theta."""
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    tt = 1.5 * theta if theta_t is None else theta_t
    out = C.c_double()
    st = (C.c_int64 * 3)()
    ctx.check(L.lib().rc_energy_bh(ctx.handle, L.dptr(v), C.byref(_params(alpha, beta, epsilon)),
                                   float(theta), float(tt), int(leaf_size), C.byref(out), st))
    return out.value, {"visits": st[0], "approximated_edge_refs": st[1], "direct_pairs": st[2]}


def bh_gradient_fd(vertices, edges, alpha, beta, epsilon, theta=0.5, leaf_size=8, theta_t=None,
                   h=1e-6):
    """Forward difference of the Barnes-Hut energy (the stage-2 oracle's
    finite_difference_gradient of bh_energy)."""
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    tt = 1.5 * theta if theta_t is None else theta_t
    g = np.zeros_like(v)
    ctx.check(L.lib().rc_gradient_bh(ctx.handle, L.dptr(v), C.byref(_params(alpha, beta, epsilon)),
                                     float(theta), float(tt), int(leaf_size), float(h), L.dptr(g)))
    return g


_KINDS = {"high": 0, "low": 1}


def bct_matvec_kernel(vertices, edges, epsilon, psi, kind, theta=0.5, leaf_size=8, theta_t=None):
    """Block-cluster-tree product K psi of the 'high' or 'low' edge kernel
    (stage-2 oracle bct_matvec_kernel, oracle/tpe_stage2_oracle.py:649-709):
    (phi, {leaves, admissible_blocks, direct_blocks, direct_pairs})."""
    if kind not in _KINDS:
        raise L.RcError(L.RC_BAD_ARG, f"unknown kernel kind {kind!r}")
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    psi = np.ascontiguousarray(np.asarray(psi, dtype=np.float64))
    nc = 1 if psi.ndim == 1 else psi.shape[1]
    phi = np.zeros_like(psi)
    st = (C.c_int64 * 4)()
    tt = 1.5 * theta if theta_t is None else theta_t
    ctx.check(L.lib().rc_bct_matvec(ctx.handle, L.dptr(v), float(epsilon), L.dptr(psi), nc,
                                    _KINDS[kind], float(theta), float(tt), int(leaf_size),
                                    L.dptr(phi), st))
    return phi, {"leaves": st[0], "admissible_blocks": st[1], "direct_blocks": st[2],
                 "direct_pairs": st[3]}


def bct_apply_b_scalar(vertices, edges, epsilon, u, kind, theta=0.5, leaf_size=8, theta_t=None):
    """B u through the block cluster tree (stage-2 oracle bct_apply_B_scalar,
    oracle/tpe_stage2_oracle.py:763-795)."""
    if kind not in _KINDS:
        raise L.RcError(L.RC_BAD_ARG, f"unknown kernel kind {kind!r}")
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    out = np.zeros(len(v))
    tt = 1.5 * theta if theta_t is None else theta_t
    ctx.check(L.lib().rc_bct_apply_b(ctx.handle, L.dptr(v), float(epsilon), L.dptr(u),
                                     _KINDS[kind], float(theta), float(tt), int(leaf_size),
                                     L.dptr(out)))
    return out


_MODES = {"analytical": L.MODE_ANALYTICAL, "finiteDiff": L.MODE_FINITE_DIFF}


def _mode(mode: str) -> int:
    if mode not in _MODES:
        raise L.RcError(L.RC_BAD_ARG, f"unknown gradient mode {mode!r}")
    return _MODES[mode]


@dataclass
class StepOptions:
    """StepOptions of the raw L2 step (optimizer.ts:20-36)."""

    mode: str
    step_size: float
    alpha: Optional[float] = None
    beta: Optional[float] = None
    epsilon: Optional[float] = None
    h: Optional[float] = None
    collect_field: bool = False


def step(vertices, edges, disjoint_pairs, opts: StepOptions) -> dict:
    """step (optimizer.ts:40-68): one raw gradient-descent step. Pure: returns
    new arrays {vertices, energy[, field]}; field (the dE taken at the input)
    only when opts.collect_field."""
    mode = _mode(opts.mode)
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    _check_pairs(disjoint_pairs, ctx.E)
    a = DEFAULTS["alpha"] if opts.alpha is None else opts.alpha
    b = DEFAULTS["beta"] if opts.beta is None else opts.beta
    eps = DEFAULTS["epsilon"] if opts.epsilon is None else opts.epsilon
    h = DEFAULTS["h"] if opts.h is None else opts.h
    out = np.zeros_like(v)
    fld = np.zeros_like(v) if opts.collect_field else None
    e = C.c_double()
    ctx.check(L.lib().rc_step(ctx.handle, L.dptr(v), C.byref(_params(a, b, eps)), mode,
                              float(opts.step_size), float(h), L.dptr(out), C.byref(e),
                              L.dptr(fld) if fld is not None else None))
    res = {"vertices": out, "energy": e.value}
    if opts.collect_field:
        res["field"] = fld
    return res


class TangentPointEnergy:
    """Forward/backward pair: forward() evaluates E, backward() the adjoint
    gradient of the same energy (recomputing the pair terms, no E^2 tape)."""

    def __init__(self, edges, num_vertices, alpha=3.0, beta=6.0, epsilon=1e-10):
        self.edges = _edges(edges)
        self.V = int(num_vertices)
        self.params = (alpha, beta, epsilon)
        self._saved = None

    def forward(self, vertices) -> float:
        self._saved = _verts(vertices).copy()
        return calculate_energy(self._saved, self.edges, None, *self.params)

    def backward(self, grad_output: float = 1.0) -> np.ndarray:
        if self._saved is None:
            raise RuntimeError("backward() before forward()")
        return grad_output * gradient_analytical(self._saved, self.edges, None, *self.params)


def assemble_a(vertices, edges, disjoint_pairs, alpha, beta, epsilon) -> np.ndarray:
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    _check_pairs(disjoint_pairs, ctx.E)
    A = np.zeros((len(v), len(v)))
    ctx.check(L.lib().rc_assemble_a(ctx.handle, L.dptr(v), C.byref(_params(alpha, beta, epsilon)),
                                    L.dptr(A)))
    return A


# --------------------------------------------------------------------------
# constraint sets (constraintSet.ts) and penalties (penalties.ts)

@dataclass
class ConstraintBlock:
    kind: str  # 'barycenter' | 'totalLength' | 'edgeLengths' | 'point'
    x0: Optional[np.ndarray] = None
    L0: Optional[float] = None
    ell0: Optional[np.ndarray] = None
    vertex: Optional[int] = None
    target: Optional[np.ndarray] = None


def barycenter_block(x0) -> ConstraintBlock:
    return ConstraintBlock("barycenter", x0=np.asarray(x0, float))


def total_length_block(L0) -> ConstraintBlock:
    return ConstraintBlock("totalLength", L0=float(L0))


def edge_lengths_block(ell0) -> ConstraintBlock:
    return ConstraintBlock("edgeLengths", ell0=np.ascontiguousarray(np.asarray(ell0, float)))


def point_block(vertex_index: int, target) -> ConstraintBlock:
    return ConstraintBlock("point", vertex=int(vertex_index), target=np.asarray(target, float))


def assert_valid_constraint_set(cs) -> None:
    """constraintSet.ts:333-344."""
    kinds = {b.kind for b in cs}
    if "totalLength" in kinds and "edgeLengths" in kinds:
        raise ValueError("ConstraintSet: totalLength and edgeLengths are mutually exclusive")


def _blocks_array(cs):
    keep = []
    arr = (L.Block * max(1, len(cs)))()
    for i, b in enumerate(cs):
        blk = arr[i]
        if b.kind == "barycenter":
            blk.kind = L.BLOCK_BARYCENTER
            blk.x0[:] = [float(x) for x in b.x0]
        elif b.kind == "totalLength":
            blk.kind = L.BLOCK_TOTAL_LENGTH
            blk.L0 = float(b.L0)
        elif b.kind == "edgeLengths":
            blk.kind = L.BLOCK_EDGE_LENGTHS
            keep.append(b.ell0)
            blk.ell0 = L.dptr(b.ell0)
            blk.ell0_len = len(b.ell0)
        elif b.kind == "point":
            blk.kind = L.BLOCK_POINT
            blk.vertex = int(b.vertex)
            blk.target[:] = [float(x) for x in b.target]
        else:
            raise ValueError(f"unknown constraint kind {b.kind!r}")
    return arr, len(cs), keep


def _length_vec(vertices, edges):
    v, e = _verts(vertices), _edges(edges)
    return np.linalg.norm(v[e[:, 1]] - v[e[:, 0]], axis=1)


def barycenter_target(vertices, edges) -> np.ndarray:
    """constraints.ts:23-52 (host-side target setup, run once per run start)."""
    v, e = _verts(vertices), _edges(edges)
    total = 0.0
    acc = np.zeros(3)
    for (i1, i2) in e:
        p1, p2 = v[i1], v[i2]
        d = p2 - p1
        ell = float(np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
        total += ell
        acc += ell * 0.5 * (p1 + p2)
    if total < 1e-14:
        return v.mean(axis=0)
    return acc / total


def total_length(vertices, edges) -> float:
    return float(sum(_length_vec(vertices, edges)))


def edge_lengths(vertices, edges) -> np.ndarray:
    return _length_vec(vertices, edges)


@dataclass
class PenaltyConfig:
    """PenaltyConfig (penalties.ts:45-49): weights absent/0 = off."""

    total_length: float = 0.0
    length_diff: float = 0.0
    field_weight: float = 0.0
    field_X: tuple = (0.0, 0.0, 0.0)

    def to_c(self) -> L.Penalties:
        p = L.Penalties(float(self.total_length), float(self.length_diff), float(self.field_weight))
        p.X[:] = [float(x) for x in self.field_X]
        return p

    def active(self) -> bool:
        n = float(np.linalg.norm(self.field_X))
        wf = self.field_weight if n >= 1e-14 else 0.0
        return self.total_length != 0 or self.length_diff != 0 or wf != 0


def penalty_energy_and_gradient(vertices, edges, config: PenaltyConfig):
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    pen = config.to_c()
    e = C.c_double()
    g = np.zeros_like(v)
    ctx.check(L.lib().rc_penalties_eval(ctx.handle, L.dptr(v), C.byref(pen), C.byref(e), L.dptr(g)))
    return e.value, g


# --------------------------------------------------------------------------
# solves and steps

@dataclass
class SobolevStepOptions:
    """SobolevStepOptions (optimizer.ts:95-175) plus the line-search and
    projection knobs the reference exposes (lineSearch.ts:120-145, 333-371)."""

    mode: str = "analytical"      # 'analytical' | 'finiteDiff' (dE of the step)
    alpha: float = 3.0
    beta: float = 6.0
    epsilon: float = 1e-10
    h: float = 1e-6               # forward-difference step (optimizer.ts:18)
    collect_timings: bool = False
    collect_field: bool = False
    energy_before: Optional[float] = None
    penalties: Optional[PenaltyConfig] = None
    projection_mode: str = "reassemble"  # optimizer.ts:144-160 default; the app sets 'frozen' 
    factor_mode: str = "ldlt"     # accepted; the device always factors A + mu 11^T (Cholesky)
    precision: str = "fp64"       # 'fp64' (plain) | 'mixed'
    tol_abs: float = 1e-4
    tol_rel: float = 1e-4
    max_iter: int = 8
    c1: float = 1e-4
    shrink: float = 0.5
    tau0: float = 1.0
    tau_min: float = 1e-12
    refine_tol: float = 1e-12
    refine_max_iter: int = 200
    # double-double refinement of the saddle solution (not a reference knob:
    # the FP64 solve alone stops ~kappa u_64 from the exact solution)
    saddle_refine: int = 1
    refine_columns: int = 0       # also refine A_hat^-1 C^T (exact frozen projection)
    saddle_refine_tol: float = 1e-6
    report_ir_residual: bool = False  # also form the refined solution's compensated residual

    def to_c(self) -> L.StepOpts:
        o = L.StepOpts()
        L.lib().rc_default_step_opts(C.byref(o))
        o.params = L.Params(self.alpha, self.beta, self.epsilon)
        o.has_energy_before = 0 if self.energy_before is None else 1
        o.energy_before = 0.0 if self.energy_before is None else float(self.energy_before)
        o.precision = L.PRECISION_MIXED if self.precision == "mixed" else L.PRECISION_FP64
        if self.projection_mode not in ("frozen", "reassemble"):
            raise L.RcError(L.RC_BAD_ARG, f"unknown projection mode {self.projection_mode!r}")
        o.projection_frozen = 1 if self.projection_mode == "frozen" else 0
        o.max_iter, o.tol_abs, o.tol_rel = int(self.max_iter), float(self.tol_abs), float(self.tol_rel)
        o.c1, o.shrink, o.tau0, o.tau_min = self.c1, self.shrink, self.tau0, self.tau_min
        o.refine_tol, o.refine_max_iter = float(self.refine_tol), int(self.refine_max_iter)
        o.collect_timings = 1 if self.collect_timings else 0
        o.gradient_mode = _mode(self.mode)
        o.fd_h = float(self.h)
        o.saddle_refine = int(self.saddle_refine)
        o.refine_columns = int(self.refine_columns)
        o.saddle_refine_tol = float(self.saddle_refine_tol)
        o.report_ir_residual = 1 if self.report_ir_residual else 0
        return o


def solve_constrained_gradient_set_frozen(vertices, edges, disjoint_pairs, alpha, beta, epsilon,
                                          dE, cs, precision: str = "fp64", saddle_refine: int = 1,
                                          refine_tol: float = 1e-12, refine_columns: int = 0):
    """Returns (g_tilde (V,3), lambda (k,), residual)."""
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    _check_pairs(disjoint_pairs, ctx.E)
    arr, nb, keep = _blocks_array(cs)
    opts = SobolevStepOptions(alpha=alpha, beta=beta, epsilon=epsilon, precision=precision,
                              saddle_refine=saddle_refine, refine_tol=refine_tol,
                              refine_columns=refine_columns).to_c()
    k = L.lib().rc_constraint_rows(ctx.handle, arr, nb)
    gt = np.zeros_like(v)
    lam = np.zeros(max(k, 1))
    res = C.c_double()
    d = _verts(dE)
    ctx.check(L.lib().rc_solve_constrained_gradient(ctx.handle, L.dptr(v), L.dptr(d), arr, nb,
                                                    C.byref(opts), L.dptr(gt), L.dptr(lam),
                                                    C.byref(res)))
    del keep
    return gt, lam[:k], res.value


def _stats_dict(st: L.StepStats) -> dict:
    out = {
        "tau": st.tau,
        "residual": st.residual,
        "gradientL2Norm": st.gradient_l2_norm,
        "projectionIterations": None if st.projection_iterations < 0 else st.projection_iterations,
    }
    if st.reason:
        out["reason"] = L.REASONS[st.reason]
    return out


def sobolev_step_set(vertices, edges, disjoint_pairs, cs, opts: SobolevStepOptions) -> dict:
    """One constrained Sobolev descent step (optimizer.ts:231-440). Pure:
    returns new arrays, never raises for a singular saddle (folded into
    stats['reason'] like the reference)."""
    v = _verts(vertices)
    ctx = context_for(edges, len(v))
    _check_pairs(disjoint_pairs, ctx.E)
    arr, nb, keep = _blocks_array(cs)
    o = opts.to_c()
    pen = opts.penalties.to_c() if opts.penalties is not None else None
    out = np.zeros_like(v)
    gt = np.zeros_like(v) if opts.collect_field else None
    st = L.StepStats()
    tm = L.Timings() if opts.collect_timings else None
    ctx.check(L.lib().rc_sobolev_step(ctx.handle, L.dptr(v), arr, nb,
                                      C.byref(pen) if pen is not None else None, C.byref(o),
                                      L.dptr(out), C.byref(st),
                                      L.dptr(gt) if gt is not None else None,
                                      C.byref(tm) if tm is not None else None))
    del keep
    res = {
        "vertices": out,
        "energy": st.energy,
        "accepted": bool(st.accepted),
        "converged": bool(st.converged),
        "stats": _stats_dict(st),
        "diagnostics": {"trials": st.trials, "slope": st.slope, "phiNorm": st.phi_norm,
                        "energyBefore": st.energy_before, "refineIters": st.refine_iters,
                        "refineResidual": st.refine_residual,
                        "saddleRefineSteps": st.saddle_ir_steps,
                        "saddleResidualDD": st.saddle_ir_residual,
                        "saddleCorrection": st.saddle_ir_correction,
                        "fp64Fallback": bool(st.fp64_fallback)},
    }
    if tm is not None:
        # only the phases that fired, like the reference's Partial ledger
        res["timings"] = {k: {"ms": getattr(tm, k).ms, "calls": getattr(tm, k).calls}
                          for k, _ in L.Timings._fields_ if getattr(tm, k).calls > 0}
    if opts.collect_field:
        res["descentField"] = None if res["stats"].get("reason") == "singular_system" else gt
    return res


def sobolev_step(vertices, edges, disjoint_pairs, x0, opts: SobolevStepOptions) -> dict:
    """Back-compat barycenter-only signature (optimizer.ts:455-469)."""
    return sobolev_step_set(vertices, edges, disjoint_pairs, [barycenter_block(x0)], opts)
