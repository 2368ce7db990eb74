"""Multigrid saddle solves on the device (rc_mg_solve through rcurves.mg)
against the reference's stage-2 oracle (oracle/tpe_stage2_oracle.py:796-1198)
on its gate script's topology classes (tests/golden/mg_goldens.json):

- mg_projected_solve with nested and zero initial guesses: the iterate to
  1e-8 (two converged runs of the oracle itself differ by up to 3e-10), the
  cycle count within one (a residual crossing the 1e-10 target is rounding-
  sensitive), the dense cleanup taken in the same cases, the first residuals
  to 1e-6;
- the divergence guard: with rediscretized coarse operators (the Gram
  matrix assembled on each coarse curve) the open chains and the junction
  network raise MGDivergenceError exactly where the oracle does;
- mg_saddles: the gradient and projection solves and their distance to the
  dense solve;
- at V = 2048 (no reference vectors at that size): convergence, the dense
  solve's agreement, and the projection's constraint residual.
"""
import json
import os
import sys
import time

import numpy as np
import pytest

from rcurves import core, curves, mg

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def mg_goldens():
    with open(os.path.join(HERE, "golden", "mg_goldens.json")) as f:
        return json.load(f)


def setup(g, case):
    sh = g["shapes"][case["shape"]]
    V = np.asarray(sh["vertices"], float)
    E = np.asarray(sh["edges"], np.int32)
    spec = mg.ConstraintSpec(total_length=case["spec"] != "bary",
                             points=tuple((p[0], tuple(p[1])) for p in case["points"]))
    targets = {"x0": np.asarray(case["x0"]), "L0": case["L0"]}
    levels = mg.build_hierarchy(V, E, spec)
    mg.prepare_levels(levels, spec, targets)
    return V, E, spec, targets, levels


def test_mg_projected_solve_matches_reference(mg_goldens):
    worst = 0.0
    for case in mg_goldens["cases"]:
        V, E, spec, targets, levels = setup(mg_goldens, case)
        for init in ("nested", "zero"):
            gd = case[f"solve_{init}"]
            log("mg solve", case["shape"], case["spec"], init)
            r = mg.mg_projected_solve(V, E, levels, case["rhs"], spec, targets,
                                      mg.MgOptions(init=init))
            e = rel(r["x"], gd["x"])
            worst = max(worst, e)
            log("  iterations", r["iterations"], gd["iterations"], "x err", e,
                "cleanup", r["used_direct_cleanup"], gd["cleanup"])
            assert e <= 1e-8
            assert abs(r["iterations"] - gd["iterations"]) <= 1
            assert r["used_direct_cleanup"] == gd["cleanup"]
            assert r["residual"] <= 100 * 1e-10
            n = min(3, len(gd["history"]), len(r["residual_history"]))
            np.testing.assert_allclose(r["residual_history"][:n], gd["history"][:n], rtol=1e-6)
    log("worst iterate error", worst)


def test_divergence_guard_with_rediscretized_coarse_operators(mg_goldens):
    for case in mg_goldens["cases"]:
        if "rediscretize" not in case:
            continue
        V, E, spec, targets, levels = setup(mg_goldens, case)
        ops = {li: core.assemble_a(lv.vertices, lv.edges, None, 3.0, 6.0, 1e-10)
               for li, lv in enumerate(levels) if li > 0}
        log("rediscretize", case["shape"], case["rediscretize"])
        if case["rediscretize"]["raised"]:
            with pytest.raises(mg.MGDivergenceError) as ei:
                mg.mg_projected_solve(V, E, levels, case["rhs"], spec, targets,
                                      coarse_operators=ops)
            hist = ei.value.residual_history
            gh = case["rediscretize"]["history"]
            assert len(hist) == len(gh)
            np.testing.assert_allclose(hist, gh, rtol=1e-6)
        else:
            r = mg.mg_projected_solve(V, E, levels, case["rhs"], spec, targets,
                                      coarse_operators=ops)
            assert r["residual"] <= 100 * 1e-10
            assert abs(r["iterations"] - case["rediscretize"]["iterations"]) <= 1


def test_mg_saddles_matches_reference(mg_goldens):
    for case in mg_goldens["cases"]:
        if "mg_saddles" not in case:
            continue
        V, E, spec, targets, _ = setup(mg_goldens, case)
        targets = dict(targets, ell0=None)
        log("mg_saddles", case["shape"])
        r = mg.mg_saddles(V, E, 3.0, 6.0, 1e-10, spec, targets, case["dE"])
        for part in ("gradient", "projection"):
            gd = case["mg_saddles"][part]
            e = rel(r[part]["x"], gd["x"])
            log(" ", part, "x err", e, "iterations", r[part]["iterations"], gd["iterations"],
                "dense diff", r[part]["dense_rel_diff"], gd["dense_rel_diff"],
                "saddle residual", r[part]["saddle_residual"])
            assert e <= 1e-8
            assert abs(r[part]["iterations"] - gd["iterations"]) <= 1
            assert r[part]["used_direct_cleanup"] == gd["cleanup"]
            assert r[part]["dense_rel_diff"] <= max(10 * gd["dense_rel_diff"], 1e-9)
            assert r[part]["saddle_residual"] <= 1e-8


def test_mg_saddles_at_v2048():
    v, e = curves.trefoil_jittered(2048, 0.002, seed=5)
    x0 = core.barycenter_target(v, e)
    L0 = core.total_length(v, e)
    spec = mg.ConstraintSpec(total_length=True)
    targets = {"x0": x0, "L0": L0 * 1.001, "ell0": None}
    dE = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
    flat = np.concatenate([dE[:, 0], dE[:, 1], dE[:, 2]])
    t = time.perf_counter()
    r = mg.mg_saddles(v, e, 3.0, 6.0, 1e-10, spec, targets, flat,
                      mg.MgOptions(max_cycles=60))
    log("V=2048 mg_saddles", time.perf_counter() - t, "s; levels",
        [lv["n_vertices"] for lv in r["levels"]])
    for part in ("gradient", "projection"):
        p = r[part]
        log(" ", part, "iterations", p["iterations"], "residual", p["projected_residual"],
            "cleanup", p["used_direct_cleanup"], "dense diff", p["dense_rel_diff"],
            "saddle residual", p["saddle_residual"])
        assert p["projected_residual"] <= 100 * 1e-10
        assert p["dense_rel_diff"] <= 1e-4
    assert r["projection"]["saddle_residual"] <= 1e-6
