"""Writes Barnes-Hut energy and block-cluster-tree matvec vectors of the
reference's stage-2 oracle.

The Barnes-Hut energy (oracle/tpe_stage2_oracle.py:456-596 of the reference:
edge BVH, admissibility with separate spatial and tangent thresholds, cluster
term) has no golden in the reference that admits a cluster (its only golden,
crossing_stage2_golden.json, approximates nothing at six edges), so its
vectors are produced here by running that oracle once, in a container where
the reference checkout is mounted. Inputs are the closed trefoils of the
reference's own gate script (oracle/check_bh_stage2.py: scale 0.25, N = 60
and 120) and three of its golden fixtures (a junction, two linked rings, a
knot); outputs are the energy and the traversal counts, and for the
block-cluster-tree products of the two edge kernels (bct_matvec_kernel,
bct_apply_B_scalar, :637-795; equally without a golden that admits a block)
the products of fixed densities and the block counts, and the Barnes-Hut
differential (its forward difference, :133-141, 1289-1290). Only these
numbers are written (tests/golden/bh_goldens.json), no reference code.

Usage: python tests/golden/make_bh_goldens.py [/root/reference]
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ALPHA, BETA, EPS = 3.0, 6.0, 1e-10


def trefoil(n):
    # oracle/check_bh_stage2.py:trefoil
    t = np.linspace(0.0, 2.0 * np.pi, n, endpoint=False)
    v = np.stack([np.sin(t) + 2.0 * np.sin(2.0 * t), np.cos(t) - 2.0 * np.cos(2.0 * t),
                  -np.sin(3.0 * t)], axis=1) * 0.25
    e = np.asarray([[i, (i + 1) % n] for i in range(n)], dtype=int)
    return v, e


def main(ref):
    sys.path.insert(0, os.path.join(ref, "oracle"))
    import tpe_stage2_oracle as O  # noqa: E402  (the reference oracle, run here only)

    fixtures = json.load(open(os.path.join(HERE, "reference_goldens.json")))["fixtures"]
    shapes = {"trefoil60": trefoil(60), "trefoil120": trefoil(120)}
    for name in ("junction-y", "linked-rings", "knot"):
        fx = fixtures[name]
        shapes[name] = (np.asarray(fx["vertices"], float), np.asarray(fx["edges"], int))
    settings = [(0.0, None, 2), (0.0, None, 8), (0.5, 0.75, 4), (0.25, 0.5, 4), (0.5, None, 8),
                (1.0, 1.5, 2)]
    cases = []
    for name, (v, e) in shapes.items():
        exact = O.calculate_energy(v, e, ALPHA, BETA, EPS)
        for theta, theta_t, leaf in settings:
            eb, st = O.bh_energy(v, e, ALPHA, BETA, EPS, theta=theta, leaf_size=leaf,
                                 theta_t=theta_t)
            cases.append({"shape": name, "theta": theta,
                          "theta_t": 1.5 * theta if theta_t is None else theta_t,
                          "leaf_size": leaf, "energy": eb, "exact_energy": exact,
                          "visits": st["visits"],
                          "approximated_edge_refs": st["approximated_edge_refs"],
                          "direct_pairs": st["direct_pairs"]})
    # block-cluster-tree matvecs of the edge kernels (tpe_stage2_oracle.py:598-795):
    # K psi for a scalar and a 3-vector density, and the Gram-operator apply B u
    bct = []
    for name, (v, e) in shapes.items():
        m, n = len(e), len(v)
        i = np.arange(m)
        psi1 = np.sin(i * 0.61) + 0.2
        psi3 = np.stack([np.sin(0.37 * i), np.cos(0.53 * i), np.sin(0.11 * i) + 0.3], axis=1)
        u = np.sin(0.7 * np.arange(n)) + 0.1 * np.arange(n) / n
        for theta, theta_t, leaf in [(0.0, None, 2), (0.5, 0.75, 4), (0.25, 0.5, 4),
                                     (0.5, None, 8)]:
            tt = 1.5 * theta if theta_t is None else theta_t
            for kind in ("high", "low"):
                phi1, st1 = O.bct_matvec_kernel(v, e, EPS, psi1, kind, theta, leaf, theta_t=tt)
                phi3, _ = O.bct_matvec_kernel(v, e, EPS, psi3, kind, theta, leaf, theta_t=tt)
                bu, _ = O.bct_apply_B_scalar(v, e, EPS, u, kind, theta, leaf, theta_t=tt)
                bct.append({"shape": name, "theta": theta, "theta_t": tt, "leaf_size": leaf,
                            "kind": kind, "psi1": psi1.tolist(), "psi3": psi3.tolist(),
                            "u": u.tolist(), "phi1": phi1.tolist(), "phi3": phi3.tolist(),
                            "Bu": bu.tolist(), "stats": {k: int(x) for k, x in st1.items()}})
    # Barnes-Hut differential: the oracle's forward difference of bh_energy
    # (finite_difference_gradient, h = 1e-6; tpe_stage2_oracle.py:133-141,1289-1290)
    grads = []
    for name in ("trefoil60", "linked-rings", "knot"):
        v, e = shapes[name]
        for theta, theta_t, leaf in [(0.5, 0.75, 4), (0.0, 0.0, 8)]:
            g = O.finite_difference_gradient(
                lambda V, th=theta, tt=theta_t, lf=leaf: O.bh_energy(
                    V, e, ALPHA, BETA, EPS, th, lf, theta_t=tt)[0], v)
            grads.append({"shape": name, "theta": theta, "theta_t": theta_t, "leaf_size": leaf,
                          "h": 1e-6, "dE": g.tolist()})
    out = {"alpha": ALPHA, "beta": BETA, "epsilon": EPS, "bh_gradients": grads,
           "shapes": {k: {"vertices": v.tolist(), "edges": e.tolist()}
                      for k, (v, e) in shapes.items()},
           "cases": cases, "bct": bct}
    with open(os.path.join(HERE, "bh_goldens.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(cases)} energy cases, {len(bct)} matvec cases, {len(grads)} gradients")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
