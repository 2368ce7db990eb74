"""Multigrid saddle solves over the device path: the reference's stage-2
oracle mg_saddles / mg_projected_solve (oracle/tpe_stage2_oracle.py:796-1198).

The host side is this module, as in the reference (Python): the level
hierarchy (curve coarsening by red-black selection along vertex chains,
build_coarse_level :816-937), each level's constraint rows and value
(eval_constraints :319-346, with coarse edge-length targets summed along the
fine edge paths and pins mapped to their coarse vertices) and the projector's
(C C^T)^+ (projector :957-966, pinv with rcond 1e-12). The solve itself —
Gram assembly, Galerkin coarse operators, projected-CG smoothing, V-cycles,
nested initial guess, divergence guard and dense cleanup — runs on the device
(rc_mg_solve in librcurves.so). Vectors are coordinate blocks
[x.., y.., z..] (flatten_block :54-56).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from . import core

TANGENT_TOL = 1e-14


@dataclass
class ConstraintSpec:
    """ConstraintSpec (tpe_stage2_oracle.py:205-210): rows stack in this
    order: barycenter, total length, edge lengths, points."""

    barycenter: bool = True
    total_length: bool = False
    edge_lengths: bool = False
    points: Tuple[Tuple[int, Tuple[float, float, float]], ...] = ()


@dataclass
class Level:
    vertices: np.ndarray
    edges: np.ndarray
    orig_vertices: List[int]
    edge_paths: List[List[int]]
    parent: Optional[np.ndarray] = None   # (n_finer, 2) coarse parents, -1 = none
    weight: Optional[np.ndarray] = None   # (n_finer, 2)
    C: Optional[np.ndarray] = None
    phi: Optional[np.ndarray] = None
    Gp: Optional[np.ndarray] = None


def _adjacency(n: int, edges) -> List[List[Tuple[int, int]]]:
    adj: List[List[Tuple[int, int]]] = [[] for _ in range(n)]
    for ei, (a, b) in enumerate(edges):
        a, b = int(a), int(b)
        adj[a].append((b, ei))
        adj[b].append((a, ei))
    return adj


def build_coarse_level(level: Level, pinned_orig: set, min_vertices: int = 8) -> Optional[Level]:
    """One coarsening step (tpe_stage2_oracle.py:816-937): vertices of degree
    != 2 and pinned vertices are anchors; along every chain from an anchor the
    even positions stay (black), the odd ones go (white), the chain's last
    vertex before an anchor stays; closed anchor-free cycles alternate, the
    last vertex of an odd cycle stays. A white vertex is the average of its
    two black neighbours (weight 1/2 each, or 1 with one); coarse edges join
    black vertices along the fine chains, each carrying its fine edge path."""
    V, E = level.vertices, level.edges
    n = len(V)
    if n <= min_vertices:
        return None
    adj = _adjacency(n, E)
    deg = [len(x) for x in adj]
    anchors = {i for i in range(n) if deg[i] != 2}
    for i, orig in enumerate(level.orig_vertices):
        if orig in pinned_orig:
            anchors.add(i)
    black = set(anchors)
    visited = set()
    for a in sorted(anchors):
        for nb, ei in adj[a]:
            if ei in visited:
                continue
            path = [a, nb]
            visited.add(ei)
            prev, cur = a, nb
            while cur not in anchors and deg[cur] == 2:
                nxt = [(x, j) for x, j in adj[cur] if x != prev]
                if not nxt:
                    break
                x, j = nxt[0]
                if j in visited:
                    break
                path.append(x)
                visited.add(j)
                prev, cur = cur, x
            for pos, v in enumerate(path):
                if pos % 2 == 0 or v in anchors:
                    black.add(v)
            if path[-1] in anchors:
                black.add(path[-1])
                if len(path) >= 2 and path[-2] not in black:
                    black.add(path[-2])
    for start in range(n):
        for nb, ei in adj[start]:
            if ei in visited:
                continue
            cycle = [start]
            prev, cur = start, nb
            visited.add(ei)
            while cur != start:
                cycle.append(cur)
                nxt = [(x, j) for x, j in adj[cur] if x != prev]
                if not nxt:
                    break
                x, j = nxt[0]
                if j in visited and x != start:
                    break
                if j not in visited:
                    visited.add(j)
                prev, cur = cur, x
            for pos, v in enumerate(cycle):
                if pos % 2 == 0:
                    black.add(v)
            if len(cycle) % 2 == 1 and len(cycle) > 1:
                black.add(cycle[-1])
    if len(black) == n:
        return None
    black_list = sorted(black)
    cidx = {v: i for i, v in enumerate(black_list)}
    parent = np.full((n, 2), -1, dtype=np.int32)
    weight = np.zeros((n, 2))
    for v in range(n):
        if v in cidx:
            parent[v, 0] = cidx[v]
            weight[v, 0] = 1.0
        else:
            bn = [x for x, _ in adj[v] if x in cidx]
            if len(bn) >= 2:
                parent[v] = (cidx[bn[0]], cidx[bn[1]])
                weight[v] = (0.5, 0.5)
            elif len(bn) == 1:
                parent[v, 0] = cidx[bn[0]]
                weight[v, 0] = 1.0
            else:
                return None
    coarse_edges, edge_paths = [], []
    seen, vfine = set(), set()
    for bi in black_list:
        for nb, ei in adj[bi]:
            if ei in vfine:
                continue
            pedges = [ei]
            prev, cur = bi, nb
            vfine.add(ei)
            while cur not in cidx:
                nxt = [(x, j) for x, j in adj[cur] if x != prev]
                if not nxt:
                    break
                x, j = nxt[0]
                if j in vfine:
                    break
                pedges.append(j)
                vfine.add(j)
                prev, cur = cur, x
            if cur in cidx and cur != bi:
                a, b = cidx[bi], cidx[cur]
                key = (min(a, b), max(a, b))
                if key not in seen:
                    seen.add(key)
                    coarse_edges.append([a, b])
                    path: List[int] = []
                    for pe in pedges:
                        path.extend(level.edge_paths[pe])
                    edge_paths.append(path)
    if not coarse_edges:
        return None
    return Level(V[black_list].copy(), np.asarray(coarse_edges, dtype=np.int32),
                 [level.orig_vertices[i] for i in black_list], edge_paths, parent, weight)


def build_hierarchy(vertices, edges, spec: ConstraintSpec, max_levels: int = 16,
                    min_vertices: int = 8) -> List[Level]:
    """build_hierarchy (tpe_stage2_oracle.py:945-954)."""
    v = np.asarray(vertices, float)
    e = np.asarray(edges, np.int32)
    pinned = {int(p[0]) for p in spec.points}
    levels = [Level(v.copy(), e.copy(), list(range(len(v))), [[i] for i in range(len(e))])]
    while len(levels) < max_levels:
        nxt = build_coarse_level(levels[-1], pinned, min_vertices=min_vertices)
        if nxt is None:
            break
        levels.append(nxt)
    return levels


def _geometry(v, e):
    d = v[e[:, 1]] - v[e[:, 0]]
    ell = np.linalg.norm(d, axis=1)
    T = np.zeros_like(d)
    m = ell >= TANGENT_TOL
    T[m] = d[m] / ell[m, None]
    return ell, T, 0.5 * (v[e[:, 0]] + v[e[:, 1]])


def level_constraints(level: Level, spec: ConstraintSpec, targets: dict):
    """eval_constraints (tpe_stage2_oracle.py:247-346) on one level: Phi and
    C (k x 3n, coordinate blocks)."""
    v, e = level.vertices, level.edges
    n = len(v)
    ell, T, mid = _geometry(v, e)
    phis, rows = [], []
    if spec.barycenter:
        x0 = np.asarray(targets["x0"], float)
        phi = np.zeros(3)
        Cb = np.zeros((3, 3 * n))
        for I, (a, b) in enumerate(e):
            q = mid[I] - x0
            phi += ell[I] * q
            for r in range(3):
                for c in range(3):
                    Cb[r, c * n + a] += -T[I, c] * q[r] + (0.5 * ell[I] if r == c else 0.0)
                    Cb[r, c * n + b] += T[I, c] * q[r] + (0.5 * ell[I] if r == c else 0.0)
        phis.append(phi)
        rows.append(Cb)
    if spec.total_length:
        Cl = np.zeros((1, 3 * n))
        for I, (a, b) in enumerate(e):
            for c in range(3):
                Cl[0, c * n + a] += T[I, c]
                Cl[0, c * n + b] -= T[I, c]
        phis.append(np.array([targets["L0"] - float(np.sum(ell))]))
        rows.append(Cl)
    if spec.edge_lengths:
        ell0 = np.asarray(targets["ell0"], float)
        tg = np.array([float(np.sum(ell0[p])) for p in level.edge_paths])
        Ce = np.zeros((len(e), 3 * n))
        for I, (a, b) in enumerate(e):
            for c in range(3):
                Ce[I, c * n + a] += T[I, c]
                Ce[I, c * n + b] -= T[I, c]
        phis.append(tg - ell)
        rows.append(Ce)
    if spec.points:
        vmap = {o: i for i, o in enumerate(level.orig_vertices)}
        pp, pr = [], []
        for orig, target in spec.points:
            if orig not in vmap:
                continue
            i = vmap[orig]
            pp.extend((v[i] - np.asarray(target, float)).tolist())
            for c in range(3):
                row = np.zeros(3 * n)
                row[c * n + i] = 1.0
                pr.append(row)
        if pr:
            phis.append(np.asarray(pp))
            rows.append(np.vstack(pr))
    if not phis:
        return np.zeros(0), np.zeros((0, 3 * n))
    return np.concatenate(phis), np.vstack(rows)


def prepare_levels(levels: List[Level], spec: ConstraintSpec, targets: dict) -> None:
    """The host half of prepare_levels (tpe_stage2_oracle.py:969-999): each
    level's constraints and the projector's (C C^T)^+ (the Galerkin operators
    are formed on the device)."""
    for lv in levels:
        lv.phi, lv.C = level_constraints(lv, spec, targets)
        k = lv.C.shape[0]
        lv.Gp = (np.linalg.pinv(lv.C @ lv.C.T, rcond=1e-12) if k else np.zeros((0, 0)))


def _blocks_for(spec: ConstraintSpec, targets: dict) -> list:
    cs = []
    if spec.barycenter:
        cs.append(core.barycenter_block(targets["x0"]))
    if spec.total_length:
        cs.append(core.total_length_block(targets["L0"]))
    if spec.edge_lengths:
        cs.append(core.edge_lengths_block(targets["ell0"]))
    for v, t in spec.points:
        cs.append(core.point_block(int(v), t))
    return cs


@dataclass
class MgOptions:
    tol: float = 1e-10
    max_cycles: int = 50
    nu_pre: int = 2
    nu_post: int = 2
    init: str = "nested"   # 'nested' | 'zero'


class MGDivergenceError(RuntimeError):
    """MGDivergenceError (tpe_stage2_oracle.py:1061-1072)."""

    def __init__(self, msg, residual_history=None):
        super().__init__(msg)
        self.residual_history = residual_history or []


def mg_projected_solve(vertices, edges, levels: List[Level], rhs, spec: ConstraintSpec,
                       targets: dict, opts: MgOptions = MgOptions(),
                       coarse_operators: Optional[Dict[int, np.ndarray]] = None,
                       alpha=3.0, beta=6.0, epsilon=1e-10) -> dict:
    """mg_projected_solve (tpe_stage2_oracle.py:1109-1155) on the device.
    coarse_operators: level -> n x n operator replacing the Galerkin product
    (the reference's coarse_op='rediscretize')."""
    if opts.init not in ("nested", "zero"):
        raise ValueError(f"unknown init: {opts.init!r}")
    v = core._verts(vertices)
    ctx = core.context_for(edges, len(v))
    nl = len(levels)
    arr = (L.MgLevel * nl)()
    keep = []
    for i, lv in enumerate(levels):
        Cm = np.ascontiguousarray(lv.C, float)
        Gp = np.ascontiguousarray(lv.Gp, float)
        keep += [Cm, Gp]
        arr[i].n = len(lv.vertices)
        arr[i].k = Cm.shape[0]
        arr[i].C = L.dptr(Cm) if Cm.size else None
        arr[i].Gp = L.dptr(Gp) if Gp.size else None
        if i > 0:
            par = np.ascontiguousarray(lv.parent, np.int32)
            w = np.ascontiguousarray(lv.weight, float)
            keep += [par, w]
            arr[i].parent = par.ctypes.data_as(C.POINTER(C.c_int32))
            arr[i].weight = L.dptr(w)
        if coarse_operators and i in coarse_operators:
            A = np.ascontiguousarray(coarse_operators[i], float)
            keep.append(A)
            arr[i].A = L.dptr(A)
    o = L.MgOpts(float(opts.tol), int(opts.max_cycles), int(opts.nu_pre), int(opts.nu_post),
                 1 if opts.init == "nested" else 0)
    cs = _blocks_for(spec, targets)
    barr, nb, bkeep = core._blocks_array(cs)
    b = np.ascontiguousarray(np.asarray(rhs, float))
    x = np.zeros_like(b)
    st = L.MgStats()
    status = L.lib().rc_mg_solve(ctx.handle, L.dptr(v), C.byref(core._params(alpha, beta, epsilon)),
                                 barr, nb, arr, nl, L.dptr(b), C.byref(o), L.dptr(x), C.byref(st))
    del keep, bkeep
    hist = [st.history[i] for i in range(st.nhist)]
    if status == L.RC_DIVERGED:
        raise MGDivergenceError(L.lib().rc_last_error(ctx.handle).decode(), hist)
    ctx.check(status)
    return {"x": x, "iterations": st.iterations, "residual": st.residual,
            "residual_history": hist, "used_direct_cleanup": bool(st.used_direct_cleanup)}


def gram_apply(vertices, edges, u, alpha=3.0, beta=6.0, epsilon=1e-10) -> np.ndarray:
    """A u per coordinate block (u: 3V, coordinate blocks) on the device."""
    v = core._verts(vertices)
    ctx = core.context_for(edges, len(v))
    uu = np.ascontiguousarray(np.asarray(u, float))
    y = np.zeros_like(uu)
    ctx.check(L.lib().rc_gram_apply(ctx.handle, L.dptr(v), C.byref(core._params(alpha, beta, epsilon)),
                                    L.dptr(uu), L.dptr(y)))
    return y


def _blocks(x):
    x = np.asarray(x, float)
    n = x.size // 3
    return np.stack([x[:n], x[n:2 * n], x[2 * n:]], axis=1)


def _flat(v):
    v = np.asarray(v, float)
    return np.concatenate([v[:, 0], v[:, 1], v[:, 2]])


def mg_saddles(vertices, edges, alpha, beta, epsilon, spec: ConstraintSpec, targets: dict,
               dE_block, opts: MgOptions = MgOptions()) -> dict:
    """mg_saddles (tpe_stage2_oracle.py:1157-1198): the gradient solve with
    rhs dE and the projection solve (z = C^+ (-Phi), rhs A z, x = z - y), each
    by multigrid and compared with the dense saddle solve (the library's FP64
    solve of the same system)."""
    v = np.asarray(vertices, float)
    e = np.asarray(edges, np.int32)
    levels = build_hierarchy(v, e, spec)
    prepare_levels(levels, spec, targets)
    C0, phi0, Gp0 = levels[0].C, levels[0].phi, levels[0].Gp
    k = C0.shape[0]
    dE_block = np.asarray(dE_block, float)
    g = mg_projected_solve(v, e, levels, dE_block, spec, targets, opts, alpha=alpha, beta=beta,
                           epsilon=epsilon)
    cs = _blocks_for(spec, targets)
    gd, _, _ = core.solve_constrained_gradient_set_frozen(v, e, None, alpha, beta, epsilon,
                                                          _blocks(dE_block), cs)
    gd = _flat(gd)
    z = C0.T @ (Gp0 @ (-phi0)) if k else np.zeros(3 * len(v))
    rhs_p = gram_apply(v, e, z, alpha, beta, epsilon)
    y = mg_projected_solve(v, e, levels, rhs_p, spec, targets, opts, alpha=alpha, beta=beta,
                           epsilon=epsilon)
    x_proj = z - y["x"]
    yd, _, _ = core.solve_constrained_gradient_set_frozen(v, e, None, alpha, beta, epsilon,
                                                          _blocks(rhs_p), cs)
    xd_proj = z - _flat(yd)

    def saddle_res(x, top, bottom):
        Ax = gram_apply(v, e, x, alpha, beta, epsilon)
        if k:
            lam = np.linalg.lstsq(C0 @ C0.T, C0 @ (top - Ax), rcond=None)[0]
            r_top = Ax + C0.T @ lam - top
            r_bot = C0 @ x - bottom
            rhs = np.concatenate([top, bottom])
            return float(np.linalg.norm(np.concatenate([r_top, r_bot])) /
                         max(1.0, np.linalg.norm(rhs)))
        return float(np.linalg.norm(Ax - top) / max(1.0, np.linalg.norm(top)))

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(1.0, np.linalg.norm(b)))

    return {
        "levels": [{"n_vertices": len(lv.vertices), "n_edges": len(lv.edges),
                    "n_constraints": int(lv.C.shape[0])} for lv in levels],
        "gradient": {"x": g["x"], "projected_residual": g["residual"],
                     "saddle_residual": saddle_res(g["x"], dE_block, np.zeros(k)),
                     "iterations": g["iterations"], "residual_history": g["residual_history"],
                     "used_direct_cleanup": g["used_direct_cleanup"],
                     "dense_rel_diff": rel(g["x"], gd)},
        "projection": {"x": x_proj, "projected_residual": y["residual"],
                       "saddle_residual": saddle_res(x_proj, np.zeros(3 * len(v)), -phi0),
                       "iterations": y["iterations"], "residual_history": y["residual_history"],
                       "used_direct_cleanup": y["used_direct_cleanup"],
                       "dense_rel_diff": rel(x_proj, xd_proj),
                       "phi_norm": float(np.linalg.norm(phi0))},
    }
