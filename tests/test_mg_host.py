"""Host half of the multigrid saddle solver (rcurves.mg): the level hierarchy
(curve coarsening, kept vertices, prolongation, coarse edges) and each
level's constraint value and rows, against the reference's stage-2 oracle
(build_hierarchy / prepare_levels / eval_constraints,
oracle/tpe_stage2_oracle.py:796-999) on the topology classes of its gate
script (open chains, a junction network, closed curves) with barycenter,
total-length and pin constraints (tests/golden/mg_goldens.json, written by
tests/golden/make_mg_goldens.py). No device call."""
import json
import os

import numpy as np
import pytest

from rcurves import mg

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def mg_goldens():
    with open(os.path.join(HERE, "golden", "mg_goldens.json")) as f:
        return json.load(f)


def spec_of(case):
    return mg.ConstraintSpec(total_length=case["spec"] != "bary",
                             points=tuple((p[0], tuple(p[1])) for p in case["points"]))


def dense_P(level):
    nf = level.parent.shape[0]
    nc = len(level.vertices)
    P = np.zeros((nf, nc))
    for a in range(nf):
        for h in range(2):
            if level.parent[a, h] >= 0:
                P[a, level.parent[a, h]] += level.weight[a, h]
    return P


def test_hierarchy_and_level_constraints_match_reference(mg_goldens):
    for case in mg_goldens["cases"]:
        sh = mg_goldens["shapes"][case["shape"]]
        V = np.asarray(sh["vertices"], float)
        E = np.asarray(sh["edges"], np.int32)
        spec = spec_of(case)
        levels = mg.build_hierarchy(V, E, spec)
        mg.prepare_levels(levels, spec, {"x0": np.asarray(case["x0"]), "L0": case["L0"]})
        assert len(levels) == len(case["levels"]), case["shape"]
        for li, (lv, gd) in enumerate(zip(levels, case["levels"])):
            tag = (case["shape"], case["spec"], li)
            assert len(lv.vertices) == gd["n"], tag
            assert lv.edges.tolist() == gd["edges"], tag
            assert lv.orig_vertices == gd["orig"], tag
            assert lv.C.shape[0] == gd["k"], tag
            np.testing.assert_allclose(lv.phi, gd["phi"], rtol=1e-12, atol=1e-14)
            if li > 0:
                assert np.array_equal(dense_P(lv), np.asarray(gd["P"])), tag
            # the projector: Pi C^T = 0 and Gp is the pseudo-inverse of C C^T
            if lv.C.shape[0]:
                G = lv.C @ lv.C.T
                np.testing.assert_allclose(G @ lv.Gp @ G, G, rtol=1e-9, atol=1e-12)


def test_pins_are_kept_on_every_level(mg_goldens):
    for case in mg_goldens["cases"]:
        if not case["points"]:
            continue
        sh = mg_goldens["shapes"][case["shape"]]
        levels = mg.build_hierarchy(np.asarray(sh["vertices"]), np.asarray(sh["edges"], np.int32),
                                    spec_of(case))
        for lv in levels:
            for p in case["points"]:
                assert p[0] in lv.orig_vertices


def test_edge_length_targets_sum_along_paths():
    """Coarse edge-length rows take the sum of their fine edges' targets
    (edge_lengths_phi_C with edge_map, tpe_stage2_oracle.py:280-296): on a
    uniformly sampled open chain the coarse Phi is the chord defect."""
    n = 33
    t = np.linspace(0.0, 1.0, n)
    V = np.stack([t, 0.1 * np.sin(3 * t), np.zeros(n)], 1)
    E = np.asarray([[i, i + 1] for i in range(n - 1)], np.int32)
    spec = mg.ConstraintSpec(barycenter=False, edge_lengths=True)
    ell0 = np.linalg.norm(V[E[:, 1]] - V[E[:, 0]], axis=1)
    levels = mg.build_hierarchy(V, E, spec)
    mg.prepare_levels(levels, spec, {"ell0": ell0})
    assert len(levels) > 1
    np.testing.assert_allclose(levels[0].phi, 0.0, atol=1e-15)
    for lv in levels[1:]:
        for I, (a, b) in enumerate(lv.edges):
            chord = np.linalg.norm(lv.vertices[b] - lv.vertices[a])
            assert lv.phi[I] == pytest.approx(ell0[lv.edge_paths[I]].sum() - chord, abs=1e-14)
            assert lv.phi[I] >= -1e-15
