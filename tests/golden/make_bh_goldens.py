"""Writes Barnes-Hut energy vectors of the reference's stage-2 oracle.

The Barnes-Hut energy (oracle/tpe_stage2_oracle.py:456-596 of the reference:
edge BVH, admissibility with separate spatial and tangent thresholds, cluster
term) has no golden in the reference that admits a cluster (its only golden,
crossing_stage2_golden.json, approximates nothing at six edges), so its
vectors are produced here by running that oracle once, in a container where
the reference checkout is mounted. Inputs are the closed trefoils of the
reference's own gate script (oracle/check_bh_stage2.py: scale 0.25, N = 60
and 120) and three of its golden fixtures (a junction, two linked rings, a
knot); outputs are the energy and the traversal counts. Only these numbers
are written (tests/golden/bh_goldens.json), no reference code.

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
    out = {"alpha": ALPHA, "beta": BETA, "epsilon": EPS,
           "shapes": {k: {"vertices": v.tolist(), "edges": e.tolist()}
                      for k, (v, e) in shapes.items()},
           "cases": cases}
    with open(os.path.join(HERE, "bh_goldens.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
