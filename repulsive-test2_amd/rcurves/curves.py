"""Seeded curve generators for the Sobolev-step workloads.

Each returns (vertices (V,3) float64, edges (E,2) int32). Seeds use
numpy.random.default_rng(seed) (PCG64) and Box-Muller-free standard normals;
the seed is recorded by the caller beside the result.

Workloads (sizes as named by the step benchmark):
  trefoil_jittered   closed trefoil V=E=n, vertices as src/core/fixtures.ts:10-23
                     plus N(0, sigma^2) per coordinate
  torus_knot         closed (p,q) torus knot, R, r, uniform parameter spacing,
                     plus N(0, (0.02 * mean edge length)^2) per coordinate
  fourier_curve      closed curve sum_{k=1..6} a_k cos kt + b_k sin kt,
                     a_k, b_k in R^3 ~ N(0, 1/k^2), resampled to uniform arc length
  theta_network      components of two degree-3 junctions joined by three
                     Fourier arcs each
  stress_grid        the first 200 points of a jittered 6 x 6 x 6 grid joined to
                     their grid neighbours (the app's 'stress' preset,
                     src/core/testConfigs.ts:102-150)
  random_graph       random points in a ball of radius 2 and random distinct
                     edges, isolated vertices joined to their nearest neighbour
                     (the app's 'random' preset, src/core/testConfigs.ts:152-224)
"""
from __future__ import annotations

import numpy as np


def closed_chain(n: int) -> np.ndarray:
    i = np.arange(n, dtype=np.int32)
    return np.stack([i, (i + 1) % n], axis=1).astype(np.int32)


def trefoil(n: int) -> tuple[np.ndarray, np.ndarray]:
    """src/core/fixtures.ts:10-23 (deterministic trefoil)."""
    t = 2 * np.pi * np.arange(n) / n
    v = np.stack([np.sin(t) + 2 * np.sin(2 * t), np.cos(t) - 2 * np.cos(2 * t), -np.sin(3 * t)], 1)
    return v.astype(np.float64), closed_chain(n)


def trefoil_jittered(n: int = 128, sigma: float = 0.01, seed: int = 1):
    v, e = trefoil(n)
    rng = np.random.default_rng(seed)
    return v + sigma * rng.standard_normal(v.shape), e


def mean_edge_length(v: np.ndarray, e: np.ndarray) -> float:
    return float(np.linalg.norm(v[e[:, 1]] - v[e[:, 0]], axis=1).mean())


def torus_knot(p: int, q: int, n: int, R: float = 2.0, r: float = 1.0,
               jitter: float = 0.02, seed: int = 1):
    t = 2 * np.pi * np.arange(n) / n
    rad = R + r * np.cos(q * t)
    v = np.stack([rad * np.cos(p * t), rad * np.sin(p * t), r * np.sin(q * t)], 1)
    e = closed_chain(n)
    if jitter:
        rng = np.random.default_rng(seed)
        v = v + jitter * mean_edge_length(v, e) * rng.standard_normal(v.shape)
    return v, e


def _fourier_coeffs(rng, kmax=6):
    k = np.arange(1, kmax + 1, dtype=np.float64)
    a = rng.standard_normal((kmax, 3)) / k[:, None]
    b = rng.standard_normal((kmax, 3)) / k[:, None]
    return k, a, b


def _fourier_eval(k, a, b, t):
    ph = np.outer(t, k)
    return np.cos(ph) @ a + np.sin(ph) @ b


def _uniform_arclength(k, a, b, n, closed, t0=0.0, t1=2 * np.pi, dense=64):
    m = n * dense
    ts = np.linspace(t0, t1, m + 1)
    pts = _fourier_eval(k, a, b, ts)
    seg = np.linalg.norm(np.diff(pts, axis=0), axis=1)
    s = np.concatenate([[0.0], np.cumsum(seg)])
    if closed:
        targets = s[-1] * np.arange(n) / n
    else:
        targets = s[-1] * np.arange(n + 1) / n
    tt = np.interp(targets, s, ts)
    return _fourier_eval(k, a, b, tt)


def fourier_curve(n: int = 8192, seed: int = 1, kmax: int = 6):
    rng = np.random.default_rng(seed)
    k, a, b = _fourier_coeffs(rng, kmax)
    v = _uniform_arclength(k, a, b, n, closed=True)
    return np.ascontiguousarray(v), closed_chain(n)


def theta_network(total_edges: int = 16384, components: int = 8, seed: int = 1, kmax: int = 6):
    """Each component: junctions P, Q joined by three arcs. Every arc is a
    Fourier curve over t in [0, pi] (open), shifted so its ends meet P and Q.
    Arc interior vertices are numbered consecutively along the arc."""
    rng = np.random.default_rng(seed)
    per_comp = total_edges // components
    per_arc = [per_comp // 3 + (1 if i < per_comp % 3 else 0) for i in range(3)]
    verts, edges = [], []
    base = 0
    for c in range(components):
        P = rng.standard_normal(3) * 0.5 + np.array([4.0 * c, 0.0, 0.0])
        Q = P + np.array([0.0, 2.0, 0.0]) + 0.3 * rng.standard_normal(3)
        comp_v = [P, Q]
        idxP, idxQ = base, base + 1
        comp_e = []
        nxt = base + 2
        for arc in range(3):
            m = per_arc[arc]
            k, a, b = _fourier_coeffs(rng, kmax)
            a *= 0.5
            b *= 0.5
            pts = _uniform_arclength(k, a, b, m, closed=False, t0=0.0, t1=np.pi)
            # affine blend so the arc runs exactly from P to Q
            s = np.linspace(0.0, 1.0, m + 1)[:, None]
            pts = pts - (1 - s) * pts[0] - s * pts[-1] + (1 - s) * P + s * Q
            interior = pts[1:-1]
            ids = list(range(nxt, nxt + len(interior)))
            comp_v.extend(interior)
            chain = [idxP] + ids + [idxQ]
            for u, w in zip(chain[:-1], chain[1:]):
                comp_e.append((u, w))
            nxt += len(interior)
        verts.extend(comp_v)
        edges.extend(comp_e)
        base = nxt
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64))
    e = np.ascontiguousarray(np.asarray(edges, dtype=np.int32))
    return v, e


def seeded_pins(V: int, count: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed + 1000)
    return np.sort(rng.choice(V, size=count, replace=False)).astype(np.int32)


def stress_grid(seed: int = 1, grid: int = 6, spacing: float = 0.5, count: int = 200):
    """testConfigs.ts:102-150: points (x, y, z) of a grid x grid x grid
    lattice in x-major order up to `count`, jitter U(-0.1, 0.1) per
    coordinate, an edge to each existing +x, +y, +z neighbour (in that order,
    vertices in index order)."""
    rng = np.random.default_rng(seed)
    idx = {}
    verts = []
    for x in range(grid):
        for y in range(grid):
            for z in range(grid):
                if len(verts) >= count:
                    break
                j = (rng.random(3) - 0.5) * 0.2
                verts.append([(x - grid / 2) * spacing + j[0], (y - grid / 2) * spacing + j[1],
                              (z - grid / 2) * spacing + j[2]])
                idx[(x, y, z)] = len(verts) - 1
    edges = []
    for x in range(grid):
        for y in range(grid):
            for z in range(grid):
                i = idx.get((x, y, z))
                if i is None:
                    continue
                for nb in ((x + 1, y, z), (x, y + 1, z), (x, y, z + 1)):
                    if nb in idx:
                        edges.append((i, idx[nb]))
    return (np.ascontiguousarray(np.asarray(verts, dtype=np.float64)),
            np.ascontiguousarray(np.asarray(edges, dtype=np.int32)))


def random_graph(num_vertices: int = 20, num_edges: int = 30, seed: int = 1):
    """testConfigs.ts:152-224: points uniform in a ball of radius 2, up to
    num_edges distinct undirected edges from 10 num_edges random draws,
    then every vertex left without an edge joined to its nearest neighbour."""
    rng = np.random.default_rng(seed)
    th = rng.random(num_vertices) * 2 * np.pi
    ph = np.arccos(2 * rng.random(num_vertices) - 1)
    r = rng.random(num_vertices) ** (1 / 3) * 2
    v = np.stack([r * np.sin(ph) * np.cos(th), r * np.sin(ph) * np.sin(th), r * np.cos(ph)], 1)
    seen, edges = set(), []
    attempts = 0
    while len(edges) < num_edges and attempts < num_edges * 10:
        a, b = (int(x) for x in rng.integers(0, num_vertices, 2))
        if a != b:
            key = (min(a, b), max(a, b))
            if key not in seen:
                seen.add(key)
                edges.append((a, b))
        attempts += 1
    used = {x for ed in edges for x in ed}
    for q in range(num_vertices):
        if q in used:
            continue
        d = np.sum((v - v[q]) ** 2, axis=1)
        d[q] = np.inf
        nb = int(np.argmin(d))
        edges.append((q, nb))
        used.update((q, nb))
    return (np.ascontiguousarray(v.astype(np.float64)),
            np.ascontiguousarray(np.asarray(edges, dtype=np.int32)))


def connected_components(V: int, e: np.ndarray) -> int:
    parent = list(range(V))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in np.asarray(e):
        ra, rb = find(int(a)), find(int(b))
        if ra != rb:
            parent[ra] = rb
    return len({find(x) for x in range(V)})
