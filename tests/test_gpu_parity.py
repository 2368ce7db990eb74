"""Device parity through the C ABI: librcurves.so against the CPU oracle and
the reference's golden fixtures. FP64 form: energy and gradient <= 1e-12
relative, Gram matrix <= 1e-12 Frobenius-relative, g~/lambda <= 1e-9 (the
reference's gradient.test.ts gate), step tau / projection iterations exact,
energy <= 1e-12, vertices <= 1e-9 (lineSearch.test.ts gates)."""
import sys
import time

import numpy as np
import pytest

import rcurves
from oracle import rc_oracle as rco
from rcurves import core, curves

pytestmark = pytest.mark.gpu

STAGE1 = ["crossing", "junction-y", "helix", "linked-rings", "knot"]


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def workloads(goldens):
    out = []
    for name in STAGE1:
        fx = goldens["fixtures"][name]
        out.append((name, np.asarray(fx["vertices"], float), np.asarray(fx["edges"], np.int32)))
    v, e = curves.trefoil_jittered(128, 0.01, seed=1)
    out.append(("trefoil128", v, e))
    v, e = curves.torus_knot(3, 5, 1024, jitter=0.02, seed=1)
    out.append(("torus35_1024", v, e))
    return out


def test_energy_gradient_goldens(goldens):
    for case in goldens["energy_gradient"]:
        log("case", case["name"])
        v, e = case["vertices"], case["edges"]
        a, b, eps = case["alpha"], case["beta"], case["epsilon"]
        E = core.calculate_energy(v, e, None, a, b, eps)
        g = core.gradient_analytical(v, e, None, a, b, eps)
        assert abs(E - case["energy"]) <= 1e-12 * abs(case["energy"])
        assert rel(g, case["gradient"]) <= 1e-12


def test_energy_gradient_vs_oracle(goldens):
    for name, v, e in workloads(goldens):
        log("workload", name, len(v), len(e))
        E = core.calculate_energy(v, e, None, 3.0, 6.0, 1e-10)
        g, E2 = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10, with_energy=True)
        Eo = rco.energy(v, e)
        go = rco.gradient(v, e)
        assert abs(E - Eo) <= 1e-12 * abs(Eo), (name, E, Eo)
        assert E2 == E
        assert rel(g, go) <= 1e-12, (name, rel(g, go))
        # translation invariance of the energy: the gradient sums to zero
        assert np.linalg.norm(g.sum(axis=0)) <= 1e-10 * np.linalg.norm(g)


def test_gradient_bit_reproducible():
    v, e = curves.fourier_curve(2048, seed=3)
    g1 = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
    g2 = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
    assert np.array_equal(g1, g2)


def test_gram_matrix_vs_oracle(goldens):
    for name, v, e in workloads(goldens)[:6]:
        log("gram", name)
        A = core.assemble_a(v, e, None, 3.0, 6.0, 1e-10)
        Ao = rco.assemble("A", v, e)
        assert rel(A, Ao) <= 1e-12, (name, rel(A, Ao))
        assert np.array_equal(A, A.T)
        assert np.linalg.norm(A @ np.ones(len(v))) <= 1e-10 * np.linalg.norm(A)
        if name in goldens["goldens"]:
            assert rel(A, goldens["goldens"][name]["A"]) <= 1e-12


def test_gram_matrix_theta_network():
    # degree-3 junctions: incidence lists of 3, several components
    v, e = curves.theta_network(total_edges=768, components=2, seed=2)
    A = core.assemble_a(v, e, None, 3.0, 6.0, 1e-10)
    Ao = rco.assemble("A", v, e)
    assert rel(A, Ao) <= 1e-12
    assert np.array_equal(A, A.T)


def build_set(gd):
    cs = []
    ocs = rco.ConstraintSet()
    for kind in gd["block_kinds"]:
        if kind == "barycenter":
            cs.append(core.barycenter_block(gd["x0_barycenter_target"]))
            ocs.barycenter(gd["x0_barycenter_target"])
        elif kind == "totalLength":
            cs.append(core.total_length_block(gd["L0_total_length_target"]))
            ocs.total_length(gd["L0_total_length_target"])
        elif kind == "edgeLengths":
            cs.append(core.edge_lengths_block(gd["ell0_edge_length_targets"]))
            ocs.edge_lengths(gd["ell0_edge_length_targets"])
        elif kind == "point":
            cs.append(core.point_block(gd["pin_vertex_index"], gd["pin_target"]))
            ocs.point(gd["pin_vertex_index"], gd["pin_target"])
    return cs, ocs


SET_GOLDENS = [
    ("crossing-length-frozen", "crossing"), ("linked-rings-length-frozen", "linked-rings"),
    ("knot-length-frozen", "knot"), ("crossing-point-frozen", "crossing"),
    ("crossing-edgelengths-frozen", "crossing"), ("junction-y-edgelengths-frozen", "junction-y"),
    ("crossing-bary-pen-length", "crossing"), ("helix-bary-pen-field", "helix"),
    ("junction-y-bary-pen-diff", "junction-y"), ("crossing-length-pen-combo", "crossing"),
]


@pytest.mark.parametrize("gname,fname", SET_GOLDENS + [(n, n) for n in STAGE1])
def test_constrained_gradient_goldens(goldens, gname, fname):
    log("solve", gname)
    fx = goldens["fixtures"][fname]
    gd = goldens["goldens"][gname]
    if "block_kinds" not in gd:
        gd = dict(gd, block_kinds=["barycenter"])
    cs, _ = build_set(gd)
    gt, lam, res = core.solve_constrained_gradient_set_frozen(
        fx["vertices"], fx["edges"], None, 3.0, 6.0, 1e-10, gd["dE"], cs)
    assert rel(gt, gd["g_tilde"]) <= 1e-9, rel(gt, gd["g_tilde"])
    assert rel(lam, gd["lambda"]) <= 1e-9, rel(lam, gd["lambda"])
    assert res <= 1e-10


@pytest.mark.parametrize("V", [128, 256, 1024])
def test_constrained_gradient_multiblock_vs_oracle(V):
    """Several 64-row factor blocks (the goldens have V <= 50, one block):
    g~ and lambda against the oracle's full-saddle LDL^T with the same dE."""
    v, e = curves.torus_knot(3, 5, V, jitter=0.02, seed=1)
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    pins = curves.seeded_pins(V, 2, seed=1)
    cs = [core.barycenter_block(x0), core.total_length_block(L0)]
    ocs = rco.ConstraintSet().barycenter(x0).total_length(L0)
    for p in pins:
        cs.append(core.point_block(int(p), v[p]))
        ocs.point(int(p), v[p])
    dE = rco.gradient(v, e)
    rc, gto, lamo, reso = rco.solve_constrained_gradient(v, e, dE, ocs)
    assert rc == 0
    for prec in ("fp64", "mixed"):
        gt, lam, res = core.solve_constrained_gradient_set_frozen(v, e, None, 3.0, 6.0, 1e-10, dE,
                                                                 cs, precision=prec)
        log(V, prec, rel(gt, gto), rel(lam, lamo), res, reso)
        assert rel(gt, gto) <= 1e-9
        assert rel(lam, lamo) <= 1e-8
        assert res <= 1e-10


def pen_objects(gd):
    p = gd.get("penalties")
    if not p:
        return None, None
    return (core.PenaltyConfig(p["w_length"], p["w_diff"], p["w_field"], tuple(p["X"])),
            rco.make_penalties(p["w_length"], p["w_diff"], p["w_field"], p["X"]))


@pytest.mark.parametrize("gname,fname", SET_GOLDENS)
def test_step_vs_oracle_step(goldens, gname, fname):
    """One full sobolevStepSet (analytic dE, frozen projection) device vs oracle."""
    log("step", gname)
    fx = goldens["fixtures"][fname]
    gd = goldens["goldens"][gname]
    cs, ocs = build_set(gd)
    pen, open_ = pen_objects(gd)
    opts = core.SobolevStepOptions(penalties=pen, collect_field=True, projection_mode="frozen")
    r = core.sobolev_step_set(fx["vertices"], fx["edges"], None, cs, opts)
    kw = {"penalties": open_} if open_ is not None else {}
    out, gto, st = rco.sobolev_step(fx["vertices"], fx["edges"], ocs, rco.default_opts(**kw))
    assert r["accepted"] == st["accepted"]
    assert r["stats"]["tau"] == st["tau"]
    assert r["stats"]["projectionIterations"] == st["projection_iterations"]
    assert abs(r["energy"] - st["energy"]) <= 1e-12 * abs(st["energy"])
    assert rel(r["vertices"], out) <= 1e-9
    assert rel(r["descentField"], gto) <= 1e-9
    assert r["stats"]["residual"] <= 1e-10


def test_trajectory_c1_trefoil_20_steps():
    """Config 1: trefoil V=E=128 + N(0, 0.01^2), bary + total length, 20 steps;
    device FP64 vs oracle: tau sequence exact, positions <= 1e-9 each step."""
    v, e = curves.trefoil_jittered(128, 0.01, seed=1)
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    cs = [core.barycenter_block(x0), core.total_length_block(L0)]
    ocs = rco.ConstraintSet().barycenter(x0).total_length(L0)
    vg, vo = v.copy(), v.copy()
    eg = eo = None
    for step in range(20):
        r = core.sobolev_step_set(vg, e, None, cs, core.SobolevStepOptions(energy_before=eg, projection_mode="frozen"))
        o = rco.default_opts() if eo is None else rco.default_opts(has_energy_before=1, energy_before=eo)
        out, _, st = rco.sobolev_step(vo, e, ocs, o)
        log("step", step, r["stats"]["tau"], st["tau"], r["energy"], st["energy"])
        assert r["accepted"] == st["accepted"] and r["stats"]["tau"] == st["tau"]
        assert r["stats"]["projectionIterations"] == st["projection_iterations"]
        assert rel(r["vertices"], out) <= 1e-9
        assert abs(r["energy"] - st["energy"]) <= 1e-12 * abs(st["energy"])
        vg, vo = r["vertices"], out
        eg, eo = r["energy"], st["energy"]


def test_full_size_directional_derivatives():
    """E = 8192 Fourier curve: central FD (h = 1e-6) of the device energy on
    32 seeded unit directions against the adjoint gradient."""
    v, e = curves.fourier_curve(8192, seed=1)
    g = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
    rng = np.random.default_rng(11)
    h = 1e-6
    worst = 0.0
    for i in range(32):
        u = rng.standard_normal(v.shape)
        u /= np.linalg.norm(u)
        ep = core.calculate_energy(v + h * u, e, None, 3.0, 6.0, 1e-10)
        em = core.calculate_energy(v - h * u, e, None, 3.0, 6.0, 1e-10)
        fd = (ep - em) / (2 * h)
        an = float(np.sum(g * u))
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-3 * np.linalg.norm(g)))
    log("worst FD rel err", worst)
    assert worst <= 1e-5


def test_isolated_vertex_is_singular_system():
    v, e = curves.trefoil_jittered(32, 0.01, seed=2)
    v = np.vstack([v, [[10.0, 0.0, 0.0]]])
    x0 = rco.barycenter_target(v, e)
    r = core.sobolev_step_set(v, e, None, [core.barycenter_block(x0)], core.SobolevStepOptions())
    assert r["accepted"] is False and r["converged"] is False
    assert r["stats"]["reason"] == "singular_system"
    assert r["stats"]["projectionIterations"] is None
    assert np.array_equal(r["vertices"], v)


REASSEMBLE_GOLDENS = [
    ("crossing-length", "crossing"), ("linked-rings-length", "linked-rings"),
    ("crossing-point", "crossing"), ("linked-rings-point", "linked-rings"),
    ("crossing-edgelengths", "crossing"), ("junction-y-edgelengths", "junction-y"),
    ("crossing-bary-pen-length", "crossing"), ("helix-bary-pen-field", "helix"),
    ("junction-y-bary-pen-diff", "junction-y"), ("crossing-length-pen-combo", "crossing"),
]


@pytest.mark.parametrize("gname,fname", REASSEMBLE_GOLDENS + [(n, n) for n in STAGE1])
def test_step_reassemble_vs_oracle_step(goldens, gname, fname):
    """projectionMode 'reassemble' (lineSearch.ts:263-276, the function default
    of sobolevStepSet and the mode of these goldens): A and C rebuilt and the
    saddle refactored at every projection iterate. Device vs the oracle's
    reassemble step, whose line search the goldens pin."""
    log("step reassemble", gname)
    fx = goldens["fixtures"][fname]
    gd = goldens["goldens"][gname]
    if "block_kinds" not in gd:
        gd = dict(gd, block_kinds=["barycenter"])
    cs, ocs = build_set(gd)
    pen, open_ = pen_objects(gd)
    r = core.sobolev_step_set(fx["vertices"], fx["edges"], None, cs,
                              core.SobolevStepOptions(penalties=pen, collect_field=True))
    kw = {"penalties": open_} if open_ is not None else {}
    out, gto, st = rco.sobolev_step(fx["vertices"], fx["edges"], ocs,
                                    rco.default_opts(projection_frozen=0, **kw))
    assert r["accepted"] == st["accepted"]
    assert r["stats"]["tau"] == st["tau"]
    assert r["stats"]["projectionIterations"] == st["projection_iterations"]
    assert abs(r["energy"] - st["energy"]) <= 1e-12 * abs(st["energy"])
    assert rel(r["vertices"], out) <= 1e-9
    assert rel(r["descentField"], gto) <= 1e-9


def test_reassemble_and_frozen_differ_where_the_reference_says():
    # optimizer.ts:144-160: on junction-heavy fixtures the frozen projection's
    # tau = 1 trial can fail where reassembling projects it (the goldens:
    # junction-y-edgelengths tau 1.0 reassembled, 0.5 frozen); the device
    # follows the oracle in each mode
    import json
    import os
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "reference_goldens.json")))
    fx = G["fixtures"]["junction-y"]
    gd = G["goldens"]["junction-y-edgelengths"]
    cs, ocs = build_set(gd)
    taus = {}
    for mode, frozen in (("reassemble", 0), ("frozen", 1)):
        r = core.sobolev_step_set(fx["vertices"], fx["edges"], None, cs,
                                  core.SobolevStepOptions(projection_mode=mode))
        _, _, st = rco.sobolev_step(fx["vertices"], fx["edges"], ocs,
                                    rco.default_opts(projection_frozen=frozen))
        log(mode, r["stats"], st["tau"], st["projection_iterations"])
        assert r["stats"]["tau"] == st["tau"]
        assert r["stats"]["projectionIterations"] == st["projection_iterations"]
        taus[mode] = r["stats"]["tau"]
    assert taus["reassemble"] >= taus["frozen"]


@pytest.mark.parametrize("V", [128, 320])
def test_step_reassemble_multiblock_vs_oracle(V):
    """Several 64-row factor blocks, barycenter + length + 2 pins, jittered
    torus knot: tau and projection iterations exact, positions 1e-9."""
    v, e = curves.torus_knot(3, 5, V, jitter=0.02, seed=3)
    x0 = rco.barycenter_target(v, e)
    # targets off the curve: two reassembled corrections (frozen: two as well)
    L0 = rco.total_length(v, e) * 1.1
    pins = curves.seeded_pins(V, 2, seed=2)
    cs = [core.barycenter_block(x0), core.total_length_block(L0)]
    ocs = rco.ConstraintSet().barycenter(x0).total_length(L0)
    for p in pins:
        cs.append(core.point_block(int(p), v[p] + 0.01))
        ocs.point(int(p), v[p] + 0.01)
    log("reassemble multiblock", V)
    r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(collect_field=True))
    out, gto, st = rco.sobolev_step(v, e, ocs, rco.default_opts(projection_frozen=0))
    log(r["stats"], st["tau"], st["projection_iterations"])
    assert st["projection_iterations"] == 2
    assert r["accepted"] == st["accepted"]
    assert r["stats"]["tau"] == st["tau"]
    assert r["stats"]["projectionIterations"] == st["projection_iterations"]
    assert abs(r["energy"] - st["energy"]) <= 1e-11 * abs(st["energy"])
    assert rel(r["vertices"], out) <= 1e-9


def test_reassemble_mixed_precision_is_unsupported():
    v, e = curves.trefoil_jittered(64, 0.01, seed=1)
    x0 = rco.barycenter_target(v, e)
    with pytest.raises(core.L.RcError, match="reassemble"):
        core.sobolev_step_set(v, e, None, [core.barycenter_block(x0)],
                              core.SobolevStepOptions(precision="mixed"))


@pytest.mark.parametrize("E", [64, 128, 256, 512])
def test_per_edge_lengths_large_k_vs_oracle(E):
    """Per-edge length set (constraintSet.ts:206-250) beyond the LDS Schur
    system: k = E + 3 rows, the Schur system factored in global memory and
    3 + 3k right-hand sides carried by the factorisation. Jittered trefoil
    with every target 1% longer; frozen and reassembled projection against
    the oracle's full-saddle LDL^T: tau and projection iterations exact."""
    v, e = curves.trefoil_jittered(E, 0.01, seed=5)
    x0 = rco.barycenter_target(v, e)
    ell0 = core.edge_lengths(v, e) * 1.01
    cs = [core.barycenter_block(x0), core.edge_lengths_block(ell0)]
    ocs = rco.ConstraintSet().barycenter(x0).edge_lengths(ell0)
    for mode, frozen in (("frozen", 1), ("reassemble", 0)):
        log("per-edge", E, mode)
        r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
            projection_mode=mode, collect_field=True))
        out, gto, st = rco.sobolev_step(v, e, ocs, rco.default_opts(projection_frozen=frozen))
        log("  ", r["stats"], st["tau"], st["projection_iterations"], rel(r["descentField"], gto),
            rel(r["vertices"], out), abs(r["energy"] - st["energy"]) / abs(st["energy"]))
        assert r["accepted"] == st["accepted"]
        assert r["stats"]["tau"] == st["tau"]
        assert r["stats"]["projectionIterations"] == st["projection_iterations"]
        assert rel(r["descentField"], gto) <= 1e-9
        assert rel(r["vertices"], out) <= 1e-9
        assert abs(r["energy"] - st["energy"]) <= 1e-12 * abs(st["energy"])
        assert r["stats"]["residual"] <= 1e-10


def test_per_edge_lengths_at_the_row_limit():
    """k = 3 + 1021 = 1024 rows, the device limit (the oracle's full saddle
    at this size takes minutes): residual of the saddle solve, g~ tangent to
    the constraints (C g~ = 0 to rounding), and the accepted curve meets every
    block's projection tolerance. One row more is refused."""
    E = 1021
    v, e = curves.trefoil_jittered(E, 0.01, seed=6)
    x0 = rco.barycenter_target(v, e)
    ell0 = core.edge_lengths(v, e) * 1.005
    cs = [core.barycenter_block(x0), core.edge_lengths_block(ell0)]
    log("per-edge at the row limit", E)
    t0 = time.perf_counter()
    r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
        projection_mode="frozen", collect_field=True))
    log("  ", r["stats"], f"{time.perf_counter() - t0:.2f}s")
    assert r["stats"]["residual"] <= 1e-10
    ocs = rco.ConstraintSet().barycenter(x0).edge_lengths(ell0)
    _, C, _, _ = rco.eval_constraints(v, e, ocs)
    gt = np.asarray(r["descentField"])
    assert np.linalg.norm(C @ gt.T.ravel()) <= 1e-9 * np.linalg.norm(C) * np.linalg.norm(gt)
    assert r["accepted"]
    out = r["vertices"]
    L = core.total_length(out, e)
    assert np.linalg.norm(core.edge_lengths(out, e) - ell0) <= max(1e-4, 1e-4 * max(1.0, L))
    with pytest.raises(core.L.RcError, match="exceed"):
        v2, e2 = curves.trefoil_jittered(E + 1, 0.01, seed=6)
        core.sobolev_step_set(v2, e2, None,
                              [core.barycenter_block(x0),
                               core.edge_lengths_block(core.edge_lengths(v2, e2))],
                              core.SobolevStepOptions(projection_mode="frozen"))


def test_per_edge_lengths_large_k_mixed_form():
    """The mixed form's PCG on the 3 + 3k right-hand sides of a per-edge set
    beyond the LDS Schur system (k = 131; its workspace is sized by k)."""
    E = 128
    v, e = curves.trefoil_jittered(E, 0.01, seed=7)
    x0 = rco.barycenter_target(v, e)
    cs = [core.barycenter_block(x0), core.edge_lengths_block(core.edge_lengths(v, e))]
    ocs = rco.ConstraintSet().barycenter(x0).edge_lengths(core.edge_lengths(v, e))
    dE = rco.gradient(v, e)
    rc, gto, lamo, _ = rco.solve_constrained_gradient(v, e, dE, ocs)
    assert rc == 0
    for prec in ("fp64", "mixed"):
        gt, lam, res = core.solve_constrained_gradient_set_frozen(v, e, None, 3.0, 6.0, 1e-10, dE,
                                                                 cs, precision=prec)
        log("per-edge k=131", prec, rel(gt, gto), rel(lam, lamo), res)
        assert rel(gt, gto) <= 1e-9
        assert rel(lam, lamo) <= 1e-8
        assert res <= 1e-10
