"""Writes multigrid saddle-solve vectors of the reference's stage-2 oracle.

The stage-2 multigrid (oracle/tpe_stage2_oracle.py:796-1198 of the
reference: build_hierarchy, prepare_levels with Galerkin coarse operators,
mg_projected_solve with nested iteration, its divergence guard, mg_saddles)
has one golden in the reference, at six edges where no level is coarsened
(crossing_stage2_golden.json), so its vectors are produced here by running
that oracle once, in a container where the reference checkout is mounted.
Inputs are the topology classes of the reference's own gate script
(oracle/check_mg_stage2.py: the helix fixture subdivided (open chain,
n = 59), a synthetic open chain (n = 61), the junction-y fixture subdivided
twice (n = 50), the knot fixture (closed)) plus a closed trefoil (n = 96),
with the gate script's probe right-hand side and constraint sets: barycenter
only, barycenter + total length, and barycenter + total length + two pins.
Outputs: each level's size, kept vertices and prolongation, the solves'
iterates, iteration counts and residual histories (nested and zero initial
guess), the 'rediscretize' divergence-guard outcome, and mg_saddles'
gradient and projection. Only these numbers are written
(tests/golden/mg_goldens.json), no reference code.

Usage: python tests/golden/make_mg_goldens.py [/root/reference]
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ALPHA, BETA, EPS = 3.0, 6.0, 1e-10


def subdivide(V, E):
    # oracle/check_mg_stage2.py:subdivide
    pts = [v for v in V]
    newE = []
    for a, b in E:
        m = 0.5 * (V[int(a)] + V[int(b)])
        mi = len(pts)
        pts.append(m)
        newE.append([int(a), mi])
        newE.append([mi, int(b)])
    return np.asarray(pts), np.asarray(newE, dtype=int)


def open_chain(n):
    # oracle/check_mg_stage2.py:open_chain
    t = np.linspace(0.0, 1.0, n)
    V = np.stack([t, 0.15 * np.sin(2 * np.pi * t), 0.1 * np.cos(3 * np.pi * t)], axis=1)
    E = np.asarray([[i, i + 1] for i in range(n - 1)], dtype=int)
    return V, E


def trefoil(n):
    t = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    v = np.stack([np.sin(t) + 2.0 * np.sin(2.0 * t), np.cos(t) - 2.0 * np.cos(2.0 * t),
                  -np.sin(3.0 * t)], axis=1) * 0.25
    e = np.asarray([[i, (i + 1) % n] for i in range(n)], dtype=int)
    return v, e


def probe(n3):
    # oracle/check_mg_stage2.py:probe_rhs before its projection
    return np.sin(np.arange(1, n3 + 1) * 0.7) + 0.3 * np.cos(np.arange(n3) * 1.3)


def main(ref):
    sys.path.insert(0, os.path.join(ref, "oracle"))
    import tpe_stage2_oracle as O  # noqa: E402  (the reference oracle, run here only)

    fx = json.load(open(os.path.join(HERE, "reference_goldens.json")))["fixtures"]

    def fixture(name):
        return np.asarray(fx[name]["vertices"], float), np.asarray(fx[name]["edges"], int)

    shapes = {
        "helix-sub": subdivide(*fixture("helix")),
        "chain61": open_chain(61),
        "junction-y-sub2": subdivide(*subdivide(*fixture("junction-y"))),
        "knot": fixture("knot"),
        "trefoil96": trefoil(96),
    }
    out = {"shapes": {}, "cases": []}
    for name, (V, E) in shapes.items():
        out["shapes"][name] = {"vertices": V.tolist(), "edges": E.tolist()}
    specs = {
        "bary": lambda V: O.ConstraintSpec(),
        "bary-length": lambda V: O.ConstraintSpec(total_length=True),
        "bary-length-pins": lambda V: O.ConstraintSpec(
            total_length=True, points=((3, tuple(V[3] + 0.01)), (len(V) // 2, tuple(V[len(V) // 2])))),
    }
    for name, (V, E) in shapes.items():
        for sname, mk in specs.items():
            spec = mk(V)
            targets = {"x0": O.barycenter_target(V, E)}
            L0, ell0 = O.length_targets(V, E)
            targets["L0"] = L0
            targets["ell0"] = ell0
            levels = O.build_hierarchy(V, E, spec)
            O.prepare_levels(levels, ALPHA, BETA, EPS, spec, targets)
            lv = []
            for li, L in enumerate(levels):
                d = {"n": len(L.vertices), "edges": np.asarray(L.edges).tolist(),
                     "orig": list(map(int, L.orig_vertices)), "k": int(L.C.shape[0]),
                     "phi": np.asarray(L.phi).tolist()}
                if li > 0:
                    d["P"] = np.asarray(L.P_to_fine).tolist()
                lv.append(d)
            Pi, _ = O.projector(levels[0].C)
            rhs = Pi @ probe(3 * len(V))
            case = {"shape": name, "spec": sname,
                    "points": [[int(p[0]), list(p[1])] for p in spec.points],
                    "x0": targets["x0"].tolist(), "L0": float(L0), "levels": lv,
                    "rhs": rhs.tolist()}
            for init in ("nested", "zero"):
                res = O.mg_projected_solve(levels, rhs, tol=1e-10, init=init)
                case[f"solve_{init}"] = {"x": np.asarray(res["x"]).tolist(),
                                         "iterations": int(res["iterations"]),
                                         "residual": float(res["residual"]),
                                         "history": list(map(float, res["residual_history"])),
                                         "cleanup": bool(res["used_direct_cleanup"])}
            if sname == "bary":
                lr = O.build_hierarchy(V, E, spec)
                O.prepare_levels(lr, ALPHA, BETA, EPS, spec, targets, coarse_op="rediscretize")
                try:
                    res = O.mg_projected_solve(lr, rhs, tol=1e-10)
                    case["rediscretize"] = {"raised": False, "residual": float(res["residual"]),
                                            "iterations": int(res["iterations"])}
                except O.MGDivergenceError as err:
                    case["rediscretize"] = {"raised": True,
                                            "history": list(map(float, err.residual_history))}
                dE = O.finite_difference_gradient(
                    lambda X: O.calculate_energy(X, E, ALPHA, BETA, EPS), V)
                mg = O.mg_saddles(V, E, ALPHA, BETA, EPS, spec, targets, O.flatten_block(dE))
                case["dE"] = O.flatten_block(dE).tolist()
                case["mg_saddles"] = {
                    part: {"x": np.asarray(mg[part]["x"]).tolist(),
                           "iterations": int(mg[part]["iterations"]),
                           "projected_residual": float(mg[part]["projected_residual"]),
                           "saddle_residual": float(mg[part]["saddle_residual"]),
                           "dense_rel_diff": float(mg[part]["dense_rel_diff"]),
                           "cleanup": bool(mg[part]["used_direct_cleanup"])}
                    for part in ("gradient", "projection")}
            out["cases"].append(case)
            print(name, sname, [d["n"] for d in lv], case["solve_nested"]["iterations"],
                  case["solve_zero"]["iterations"], case.get("rediscretize"), flush=True)
    with open(os.path.join(HERE, "mg_goldens.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
