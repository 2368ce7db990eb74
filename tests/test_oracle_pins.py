"""Pins of the oracle's composed step and of its compensated energy.

rco_sobolev_step composes pieces pinned one by one in test_oracle_golden.py
(energy, dE, Gram matrices, constraints, penalties, saddle solve, line
search); here the composition itself is pinned against the reference's
one-step goldens: the goldens' dE is the reference oracle's forward
difference (h = finite_difference_h), so the composed step in finiteDiff
mode (optimizer.ts:278-282) must reproduce every golden's line_search_step
(tau and projection iterations exact; vertices, g~ and energy within the
measured resolution of the forward difference), with the penalties added to
dE and to the Armijo energy (optimizer.ts:283-303), and the composition is
bit-identical to the chain of its separately pinned pieces.
The glue of sobolevStepSet (optimizer.ts:305-367) — the convergence echo
below |g~|_L2h = 1e-4, E0 reuse and the singular fold — is pinned by the
reference's own rules on constructed cases (optimizer-sobolev.test.ts:
66-105, energyReuse.test.ts).
"""
import numpy as np
import pytest

from oracle import rc_oracle as rco

STAGE1 = ["crossing", "junction-y", "helix", "linked-rings", "knot"]
SET_GOLDENS = [
    ("crossing-length", "crossing"), ("crossing-length-frozen", "crossing"),
    ("linked-rings-length", "linked-rings"), ("linked-rings-length-frozen", "linked-rings"),
    ("knot-length-frozen", "knot"), ("crossing-point", "crossing"),
    ("crossing-point-frozen", "crossing"), ("linked-rings-point", "linked-rings"),
    ("crossing-edgelengths", "crossing"), ("crossing-edgelengths-frozen", "crossing"),
    ("junction-y-edgelengths", "junction-y"), ("junction-y-edgelengths-frozen", "junction-y"),
    ("crossing-bary-pen-length", "crossing"), ("helix-bary-pen-field", "helix"),
    ("junction-y-bary-pen-diff", "junction-y"), ("crossing-length-pen-combo", "crossing"),
]


def frob_rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def build_set(gd):
    cs = rco.ConstraintSet()
    for kind in gd.get("block_kinds", ["barycenter"]):
        if kind == "barycenter":
            cs.barycenter(gd["x0_barycenter_target"])
        elif kind == "totalLength":
            cs.total_length(gd["L0_total_length_target"])
        elif kind == "edgeLengths":
            cs.edge_lengths(gd["ell0_edge_length_targets"])
        elif kind == "point":
            cs.point(gd["pin_vertex_index"], gd["pin_target"])
        else:
            raise AssertionError(kind)
    return cs


CASES = [(n, n, True) for n in STAGE1] + [(g, f, False) for g, f in SET_GOLDENS]


def step_kwargs(gd, stage1):
    kw = dict(gradient_mode=rco.MODE_FINITE_DIFF, fd_h=gd["finite_difference_h"],
              projection_frozen=1 if gd["projection_mode"] == "frozen" else 0)
    if stage1:
        # lineSearch.test.ts:32 pins the stage-1 projection tolerance 1e-10
        kw.update(tol_abs=1e-10, tol_rel=1e-10)
    p = gd.get("penalties")
    if p:
        kw["penalties"] = rco.make_penalties(p["w_length"], p["w_diff"], p["w_field"], p["X"])
    return kw


@pytest.mark.parametrize("gname,fname,stage1", CASES)
def test_composed_step_equals_its_pieces(goldens, gname, fname, stage1):
    """sobolevStepSet = dE (forward difference + penalty gradient,
    optimizer.ts:278-294) -> constrained solve -> line search with the
    penalties: the composed oracle step is bit-identical to the chain of its
    separately pinned pieces on the same dE."""
    fx = goldens["fixtures"][fname]
    gd = goldens["goldens"][gname]
    v, e = fx["vertices"], fx["edges"]
    cs = build_set(gd)
    kw = step_kwargs(gd, stage1)
    out, gt, st = rco.sobolev_step(v, e, cs, rco.default_opts(**kw))
    dE = rco.gradient_fd(v, e, h=gd["finite_difference_h"])
    if "penalties" in kw:
        dE = dE + rco.penalty_gradient(v, e, kw["penalties"])
    rc, gt2, lam, res = rco.solve_constrained_gradient(v, e, dE, cs)
    assert rc == 0 and np.array_equal(gt, gt2)
    out2, st2 = rco.line_search(v, e, dE, gt2, cs, rco.default_opts(**kw))
    assert np.array_equal(out, out2)
    for k in ("tau", "energy", "energy_before", "projection_iterations", "accepted"):
        assert st[k] == st2[k], k


@pytest.mark.parametrize("gname,fname,stage1", CASES)
def test_composed_step_reproduces_line_search_goldens(goldens, gname, fname, stage1):
    """The composed step against each golden's line_search_step: tau and
    projection iterations exact. This dE and the golden's are both forward
    differences (h = 1e-6) of energies summed in different orders, so they
    differ at the difference's rounding resolution, ulp(E) / h relative
    (1.8e-10 .. 6.4e-9 on these fixtures, measured below); the step, g~ and
    the energy at the step move with it, and are gated against that
    measured difference (a dropped penalty gradient, a backward difference or
    a wrong set order moves them by orders of magnitude more)."""
    fx = goldens["fixtures"][fname]
    gd = goldens["goldens"][gname]
    v, e = fx["vertices"], fx["edges"]
    cs = build_set(gd)
    kw = step_kwargs(gd, stage1)
    out, gt, st = rco.sobolev_step(v, e, cs, rco.default_opts(**kw))
    dE = rco.gradient_fd(v, e, h=gd["finite_difference_h"])
    if "penalties" in kw:
        dE = dE + rco.penalty_gradient(v, e, kw["penalties"])
    res = max(frob_rel(dE, gd["dE"]), 1e-10)
    assert res <= 1e-8
    ls = gd["line_search_step"]
    assert st["accepted"] == ls["accepted"]
    assert st["converged"] is False
    assert st["tau"] == ls["tau"]
    assert st["projection_iterations"] == ls["projection_iterations"]
    assert frob_rel(out, ls["vertices"]) <= 4 * res
    assert abs(st["energy"] - ls["energy_after"]) <= 4 * res * abs(ls["energy_after"])
    # g~ = K^-1 [dE; 0] amplifies the dE difference by the saddle's
    # conditioning on these fixtures (at most ~7x measured)
    assert frob_rel(gt, gd["g_tilde"]) <= 20 * res
    e0 = gd.get("objective_energy_initial", gd["energy"])
    assert abs(st["energy_before"] - e0) <= 1e-12 * abs(e0)
    # the backward difference must not reproduce it
    kw["fd_h"] = -gd["finite_difference_h"]
    out_b, _, st_b = rco.sobolev_step(v, e, cs, rco.default_opts(**kw))
    assert st_b["tau"] != ls["tau"] or frob_rel(out_b, ls["vertices"]) > 4 * res or \
        st_b["projection_iterations"] != ls["projection_iterations"]


def regular_polygon(n, r=1.0):
    t = 2 * np.pi * np.arange(n) / n
    v = np.stack([r * np.cos(t), r * np.sin(t), np.zeros(n)], 1)
    e = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
    return v, e


def test_convergence_echo_at_a_critical_point():
    """A regular polygon is a critical point of the energy under
    [barycenter, total length]: |g~|_L2h < 1e-4 returns the input unchanged,
    accepted false, converged true, tau 0, projection iterations 0, energy
    E0 (optimizer.ts:355-367)."""
    v, e = regular_polygon(32)
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    cs = rco.ConstraintSet().barycenter(x0).total_length(L0)
    out, gt, st = rco.sobolev_step(v, e, cs, rco.default_opts())
    assert st["converged"] is True and st["accepted"] is False
    assert st["gradient_l2_norm"] < 1e-4
    assert st["tau"] == 0 and st["projection_iterations"] == 0
    assert np.array_equal(out, v)
    assert st["energy"] == rco.energy(v, e)
    # a provided E0 is echoed as the energy (E0 reuse, optimizer.ts:357-361)
    out, _, st = rco.sobolev_step(v, e, cs, rco.default_opts(has_energy_before=1, energy_before=123.5))
    assert st["energy"] == 123.5 and np.array_equal(out, v)
    # off the critical point the same set steps (guards a vacuous echo)
    w = v.copy()
    w[3] += (0.05, -0.02, 0.03)
    _, _, st = rco.sobolev_step(w, e, cs, rco.default_opts())
    assert st["converged"] is False and st["gradient_l2_norm"] > 1e-4


def test_energy_before_is_the_armijo_reference():
    """E0 reuse: a supplied energyBefore replaces the objective at gamma0 in
    the Armijo test (lineSearch.ts:446-451); supplying the true E0 changes
    nothing, a far larger one makes tau = 1 pass trivially."""
    fx_v, fx_e = regular_polygon(24)
    rng = np.random.default_rng(3)
    v = fx_v + 0.02 * rng.standard_normal(fx_v.shape)
    cs = rco.ConstraintSet().barycenter(rco.barycenter_target(v, fx_e))
    out0, _, st0 = rco.sobolev_step(v, fx_e, cs, rco.default_opts())
    out1, _, st1 = rco.sobolev_step(v, fx_e, cs, rco.default_opts(
        has_energy_before=1, energy_before=st0["energy_before"]))
    assert np.array_equal(out0, out1) and st0["tau"] == st1["tau"]
    assert st1["energy_before"] == st0["energy_before"]
    _, _, st2 = rco.sobolev_step(v, fx_e, cs, rco.default_opts(
        has_energy_before=1, energy_before=st0["energy_before"] + 1e6))
    assert st2["accepted"] and st2["tau"] == 1.0


def test_singular_system_folds_into_the_step():
    """An isolated vertex leaves a zero row in A and C: the saddle solve
    throws in the reference and the step folds it (optimizer.ts:331-353):
    input echoed, accepted false, reason singular_system, tau 0."""
    v, e = regular_polygon(16)
    rng = np.random.default_rng(5)
    v = v + 0.01 * rng.standard_normal(v.shape)
    v = np.vstack([v, [[3.0, 0.0, 0.0]]])   # vertex 16 has no edge
    cs = rco.ConstraintSet().barycenter(rco.barycenter_target(v, e))
    out, _, st = rco.sobolev_step(v, e, cs, rco.default_opts())
    assert st["reason"] == "singular_system"
    assert st["accepted"] is False and st["converged"] is False and st["tau"] == 0
    assert np.array_equal(out, v)
    assert st["energy"] == rco.energy(v, e)


def test_energy_compensated():
    """rco_energy_compensated sums the same pair terms as rco_energy with a
    Neumaier-compensated accumulator: it agrees with the plain sum within the
    plain sum's rounding, and it does not depend on the order of the terms
    (reversing the edge order permutes the terms) to within a few ulps,
    which the plain left-to-right sum does not guarantee."""
    from rcurves import curves
    v, e = curves.fourier_curve(512, seed=2)
    e_rev = np.ascontiguousarray(np.asarray(e)[::-1])
    P = len(e) * (len(e) - 3)
    plain, comp = rco.energy(v, e), rco.energy(v, e, compensated=True)
    plain_r, comp_r = rco.energy(v, e_rev), rco.energy(v, e_rev, compensated=True)
    u = np.finfo(float).eps / 2
    assert abs(plain - comp) <= P * u * abs(comp)
    assert abs(comp - comp_r) <= 4 * u * abs(comp)
    assert abs(plain - plain_r) > abs(comp - comp_r)
