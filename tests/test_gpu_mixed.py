"""Mixed form (FP32 factor with FP16x3 MFMA trailing updates, FP64 PCG, one
refinement step with compensated residuals) against the plain FP64 form
(also refined) and the golden fixtures.
Gates: energy rel err <= 1e-10, direction (g~) rel err <= 1e-8 against FP64;
g~ / lambda <= 1e-9 against the goldens. Both forms against the exact
solution of the assembled system: tests/test_gpu_refine.py."""
import sys

import numpy as np
import pytest

from oracle import rc_oracle as rco
from rcurves import core, curves

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["crossing", "helix", "knot"])
def test_mixed_constrained_gradient_goldens(goldens, name):
    fx = goldens["fixtures"][name]
    gd = goldens["goldens"][name]
    cs = [core.barycenter_block(gd["x0_barycenter_target"])]
    gt, lam, res = core.solve_constrained_gradient_set_frozen(
        fx["vertices"], fx["edges"], None, 3.0, 6.0, 1e-10, gd["dE"], cs, precision="mixed")
    log(name, rel(gt, gd["g_tilde"]), res)
    assert rel(gt, gd["g_tilde"]) <= 1e-9
    assert rel(lam, gd["lambda"]) <= 1e-9
    assert res <= 1e-10


@pytest.mark.parametrize("V,maker", [
    (1024, "torus"), (2048, "fourier"), (4096, "torus27"), (8192, "fourier"),
])
def test_mixed_vs_fp64_step(V, maker):
    """Sizes past the FP32 Cholesky limit (kappa u_f32 > 1) included: the
    (2,7) torus knot at V = 4096 and the Fourier curve at V = 8192, where
    the unrefined FP64 solve alone is ~2e-8 from the exact direction."""
    if maker == "torus":
        v, e = curves.torus_knot(3, 5, V, jitter=0.02, seed=1)
    elif maker == "torus27":
        v, e = curves.torus_knot(2, 7, V, jitter=0.02, seed=1)
    else:
        v, e = curves.fourier_curve(V, seed=1)
    x0 = core.barycenter_target(v, e)
    L0 = core.total_length(v, e)
    cs = [core.barycenter_block(x0), core.total_length_block(L0)]
    pen = core.PenaltyConfig(field_weight=10.0, field_X=(0.5, 1.0, 0.0))
    r64 = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(penalties=pen, collect_field=True,
                                                                             projection_mode="frozen"))
    rmx = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
        penalties=pen, collect_field=True, precision="mixed", projection_mode="frozen"))
    de = abs(rmx["energy"] - r64["energy"]) / abs(r64["energy"])
    dd = rel(rmx["descentField"], r64["descentField"])
    log(maker, V, "energy", de, "direction", dd, "tau", r64["stats"]["tau"], rmx["stats"]["tau"],
        "pcg", rmx["diagnostics"]["refineIters"], rmx["diagnostics"]["refineResidual"])
    assert rmx["stats"]["tau"] == r64["stats"]["tau"]
    assert de <= 1e-10
    assert dd <= 1e-8
    assert rmx["stats"]["residual"] <= 1e-10


def test_mixed_step_matches_oracle_trefoil():
    v, e = curves.trefoil_jittered(128, 0.01, seed=1)
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    r = core.sobolev_step_set(v, e, None, [core.barycenter_block(x0), core.total_length_block(L0)],
                              core.SobolevStepOptions(precision="mixed", projection_mode="frozen"))
    out, _, st = rco.sobolev_step(v, e, rco.ConstraintSet().barycenter(x0).total_length(L0),
                                  rco.default_opts())
    assert r["stats"]["tau"] == st["tau"]
    assert abs(r["energy"] - st["energy"]) <= 1e-12 * abs(st["energy"])
    assert rel(r["vertices"], out) <= 1e-9
