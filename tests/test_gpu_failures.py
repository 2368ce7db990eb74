"""Failure contracts of the Sobolev step on the device, each against the
oracle's answer for the same input (the oracle restates the reference's
guards and throw rules):

- a non-finite vertex: the factorisation meets a NaN pivot, the reference's
  "treating as singular" throw (linsolve.ts:236-280), folded into
  reason 'singular_system' with the input echoed (optimizer.ts:329-356);
- a point block whose vertex is out of range: NaN Phi and zero C rows
  (constraintSet.ts:271-301, constraintSetM2.test.ts:222-230), so the saddle
  system has zero rows and is singular;
- a length block with no frozen target (dispatch.ts:221-230 NaN backstop) or
  a per-edge target vector of the wrong length (constraintSetM2.test.ts:
  209-220): NaN Phi. The reference's comments expect 'projection_failed',
  but its per-block rule (lineSearch.ts:232-245) rejects only when
  `blockNorm > tol`, which is false for NaN: the candidate counts as
  converged at iteration 0 with a NaN Phi norm and the Armijo test decides.
  Device and oracle both follow the code, not the comment;
- zero-length edges (the 1e-14 guards of tangentPointEnergy.ts:173-211,
  constraintSet.ts:135-141, penalties.test.ts:272-285): no NaN anywhere.
  On a closed curve the coincident endpoints of the two neighbouring
  (disjoint) edges meet at distance 0, where the epsilon-regularised kernel
  gives (eps^alpha / eps^beta) = 1e30-sized terms: energy, gradient and A are
  finite and equal to the oracle's to the rounding of the cross products at
  the coincidence (1e-7), the assembled A is numerically indefinite and the
  step is rejected (by both, for reasons that depend on rounding).
"""
import sys

import numpy as np
import pytest

from oracle import rc_oracle as rco
from rcurves import core, curves

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def curve(n=48, seed=4):
    return curves.trefoil_jittered(n, 0.01, seed=seed)


@pytest.mark.parametrize("projection", ["frozen", "reassemble"])
def test_nan_vertex_is_singular_system(projection):
    v, e = curve()
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    v = v.copy()
    v[7, 1] = np.nan
    log("nan vertex", projection)
    cs = [core.barycenter_block(x0), core.total_length_block(L0)]
    r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(projection_mode=projection))
    _, _, st = rco.sobolev_step(v, e, rco.ConstraintSet().barycenter(x0).total_length(L0),
                                rco.default_opts(projection_frozen=int(projection == "frozen")))
    assert st["reason"] == "singular_system"
    assert r["stats"]["reason"] == "singular_system"
    assert r["accepted"] is False and r["converged"] is False
    assert r["stats"]["tau"] == 0.0
    assert np.isnan(r["stats"]["residual"]) and np.isnan(r["stats"]["gradientL2Norm"])
    assert r["stats"]["projectionIterations"] is None
    assert np.array_equal(r["vertices"], v, equal_nan=True)
    # the echoed objective is recomputed at the input (NaN here) unless E0 is given
    assert np.isnan(r["energy"]) and np.isnan(st["energy"])
    r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
        projection_mode=projection, energy_before=12.5))
    assert r["stats"]["reason"] == "singular_system" and r["energy"] == 12.5


def test_out_of_range_pin_block_is_singular_system():
    v, e = curve()
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    bad = len(v) + 5
    log("out-of-range pin", bad)
    cs = [core.barycenter_block(x0), core.total_length_block(L0),
          core.point_block(bad, [0.0, 0.0, 0.0])]
    r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(projection_mode="frozen"))
    _, _, st = rco.sobolev_step(
        v, e, rco.ConstraintSet().barycenter(x0).total_length(L0).point(bad, [0.0, 0.0, 0.0]),
        rco.default_opts())
    assert st["reason"] == "singular_system"
    assert r["stats"]["reason"] == st["reason"]
    assert r["accepted"] is False
    assert np.array_equal(r["vertices"], v)


@pytest.mark.parametrize("projection", ["frozen", "reassemble"])
def test_missing_length_target_converges_with_nan_phi(projection):
    v, e = curve()
    x0 = rco.barycenter_target(v, e)
    log("missing L0", projection)
    frozen = int(projection == "frozen")
    r = core.sobolev_step_set(v, e, None,
                              [core.barycenter_block(x0), core.total_length_block(np.nan)],
                              core.SobolevStepOptions(projection_mode=projection))
    out, _, st = rco.sobolev_step(v, e, rco.ConstraintSet().barycenter(x0).total_length(np.nan),
                                  rco.default_opts(projection_frozen=frozen))
    log("device", r["stats"], r["diagnostics"]["trials"], "oracle", st)
    assert st["accepted"] and np.isnan(st["phi_norm"]) and st["projection_iterations"] == 0
    assert r["accepted"] is True
    assert np.isnan(r["diagnostics"]["phiNorm"])
    assert r["stats"]["projectionIterations"] == 0
    assert r["stats"]["tau"] == st["tau"]
    assert r["diagnostics"]["trials"] == st["trials"]
    # the gradient solve itself is sound: C is finite, only Phi is NaN
    assert r["stats"]["residual"] <= 1e-10
    assert rel(r["vertices"], out) <= 1e-9
    assert abs(r["energy"] - st["energy"]) <= 1e-12 * abs(st["energy"])


def test_short_edge_length_targets_converge_with_nan_phi():
    v, e = curve()
    x0 = rco.barycenter_target(v, e)
    log("short ell0")
    r = core.sobolev_step_set(v, e, None,
                              [core.barycenter_block(x0), core.edge_lengths_block([1.0])],
                              core.SobolevStepOptions(projection_mode="frozen"))
    out, _, st = rco.sobolev_step(v, e, rco.ConstraintSet().barycenter(x0).edge_lengths([1.0]),
                                  rco.default_opts())
    assert st["accepted"] and np.isnan(st["phi_norm"])
    assert r["accepted"] is True and np.isnan(r["diagnostics"]["phiNorm"])
    assert r["stats"]["tau"] == st["tau"]
    assert r["diagnostics"]["trials"] == st["trials"]
    assert rel(r["vertices"], out) <= 1e-9


def degenerate_curve():
    """A closed trefoil with vertex 11 moved onto vertex 10: edge (10, 11) has
    length exactly 0."""
    v, e = curve(64, seed=6)
    v = v.copy()
    v[11] = v[10]
    return v, e


def test_zero_length_edge_energy_gradient_gram_match_oracle():
    """Finite values, equal to the oracle's to the rounding of the cross
    products at the coincidence. The device forms each numerator
    |e_I x (g_i - g_j)| (|T_I x (p_a - q_b)| in the Gram kernel) once per
    (I, j), from the first endpoint of I, with fused multiply-adds; the
    reference forms it per endpoint from separately rounded products. Equal in
    exact arithmetic, and to rounding for separated points, but where g_i and
    g_j coincide the reference's numerator is exactly 0 and the device's a
    rounding residue of order u |e_I|^2 ~ 1e-18, which stands beside
    eps = 1e-10: the eps^-3-sized terms move by ~1e-7 relative (measured
    1.8e-7, 2.2e-7 and 1.1e-7 for E, dE and A on this curve). Gate 1e-6."""
    v, e = degenerate_curve()
    log("zero-length edge: energy / gradient / A")
    E = core.calculate_energy(v, e, None, 3.0, 6.0, 1e-10)
    g = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
    A = core.assemble_a(v, e, None, 3.0, 6.0, 1e-10)
    assert np.isfinite(E) and np.all(np.isfinite(g)) and np.all(np.isfinite(A))
    Eo = rco.energy(v, e)
    go = rco.gradient(v, e)
    Ao = rco.assemble("A", v, e)
    errs = (abs(E - Eo) / abs(Eo), rel(g, go), np.linalg.norm(A - Ao) / np.linalg.norm(Ao))
    log("E, g, A relative errors", errs)
    assert max(errs) <= 1e-6


@pytest.mark.parametrize("length", ["total", "perEdge"])
def test_zero_length_edge_step_is_rejected_without_nan(length):
    v, e = degenerate_curve()
    x0 = rco.barycenter_target(v, e)
    log("zero-length edge step", length)
    if length == "total":
        L0 = rco.total_length(v, e)
        cs = [core.barycenter_block(x0), core.total_length_block(L0)]
        ocs = rco.ConstraintSet().barycenter(x0).total_length(L0)
    else:
        ell0 = core.edge_lengths(v, e)
        assert ell0[10] == 0.0
        cs = [core.barycenter_block(x0), core.edge_lengths_block(ell0)]
        ocs = rco.ConstraintSet().barycenter(x0).edge_lengths(ell0)
    r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(projection_mode="frozen"))
    _, _, st = rco.sobolev_step(v, e, ocs, rco.default_opts())
    log("device", r["stats"], "oracle", st)
    rejected = {"singular_system", "not_a_descent_direction"}
    assert not st["accepted"] and st["reason"] in rejected
    assert r["accepted"] is False and r["stats"]["reason"] in rejected
    assert np.array_equal(r["vertices"], v)
    # the echoed energy, to the bound of the energy test above
    assert abs(r["energy"] - st["energy"]) <= 1e-6 * abs(st["energy"])


def test_zero_length_edge_contributes_zero_penalty():
    v = np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]])
    e = np.array([[0, 1]], np.int32)
    log("degenerate edge penalties")
    cfg = core.PenaltyConfig(1.0, 1.0, 1.0, (0.0, 0.0, 1.0))
    E, g = core.penalty_energy_and_gradient(v, e, cfg)
    assert E == 0.0
    assert np.all(g == 0.0)
    pen = rco.make_penalties(1.0, 1.0, 1.0, (0.0, 0.0, 1.0))
    assert rco.penalty_energy(v, e, pen) == 0.0
    assert np.all(rco.penalty_gradient(v, e, pen) == 0.0)
