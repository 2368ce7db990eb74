"""Pins the CPU oracle against the reference's own golden vectors.

Gates mirror the reference's tests: test/golden.test.ts (bit-identical energy
and analytic gradient), test/sobolev/innerProduct.test.ts + lowOrder.test.ts
(B, B0, A at 1e-12 Frobenius-relative), gradient.test.ts (g~, lambda at 1e-9,
residual 1e-10), lineSearch.test.ts / constraintSetFlow.test.ts /
frozenProjection.test.ts / penalties.test.ts (tau and projection iterations
exact, energy 1e-12, vertices 1e-9).
"""
import numpy as np
import pytest

from oracle import rc_oracle as rco

STAGE1 = ["crossing", "junction-y", "helix", "linked-rings", "knot"]


def frob_rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_energy_gradient_bit_identical(goldens):
    # test/golden.test.ts:29-48 — Object.is equality on every value.
    for case in goldens["energy_gradient"]:
        v, e = case["vertices"], case["edges"]
        E = rco.energy(v, e, case["alpha"], case["beta"], case["epsilon"])
        g = rco.gradient(v, e, case["alpha"], case["beta"], case["epsilon"])
        assert E == case["energy"], (case["name"], E, case["energy"])
        assert np.array_equal(g, np.asarray(case["gradient"], float)), case["name"]


@pytest.mark.parametrize("name", STAGE1)
def test_stage1_energy_metric_constraints(goldens, name):
    fx = goldens["fixtures"][name]
    gd = goldens["goldens"][name]
    v, e = fx["vertices"], fx["edges"]
    a, b, eps = fx["alpha"], fx["beta"], fx["epsilon"]
    E = rco.energy(v, e, a, b, eps)
    assert abs(E - gd["energy"]) / abs(gd["energy"]) <= 1e-12
    for key in ("B", "B0", "A"):
        M = rco.assemble(key, v, e, a, b, eps)
        assert frob_rel(M, gd[key]) <= 1e-12, key
        assert np.array_equal(M, M.T), key  # explicit symmetrisation is exact
    A = rco.assemble("A", v, e, a, b, eps)
    # constant nullspace (innerProduct.test.ts:92-98)
    assert np.linalg.norm(A @ np.ones(len(v))) <= 1e-10 * np.linalg.norm(A)
    x0 = rco.barycenter_target(v, e)
    assert np.allclose(x0, gd["x0_barycenter_target"], rtol=0, atol=1e-15)
    cs = rco.ConstraintSet().barycenter(gd["x0_barycenter_target"])
    phi, C, _, _ = rco.eval_constraints(v, e, cs)
    assert np.allclose(phi, gd["Phi_barycenter"], rtol=0, atol=1e-14)
    assert frob_rel(C, gd["C_barycenter"]) <= 1e-14


@pytest.mark.parametrize("name", STAGE1)
def test_stage1_analytic_gradient_vs_fd_golden(goldens, name):
    # The goldens' dE is the oracle's forward FD (h=1e-6): O(h) truncation error.
    fx = goldens["fixtures"][name]
    gd = goldens["goldens"][name]
    g = rco.gradient(fx["vertices"], fx["edges"], fx["alpha"], fx["beta"], fx["epsilon"])
    assert frob_rel(g, gd["dE"]) <= 1e-4


@pytest.mark.parametrize("name", STAGE1)
def test_stage1_fd_gradient_vs_fd_golden(goldens, name):
    # gradientFiniteDiff (tangentPointEnergy.ts:114-136) against the goldens'
    # forward difference at the same h: they agree to the resolution of a
    # difference of two full energies, one ulp(E) / h per quantum.
    fx = goldens["fixtures"][name]
    gd = goldens["goldens"][name]
    v, e = fx["vertices"], fx["edges"]
    a, b, eps = fx["alpha"], fx["beta"], fx["epsilon"]
    h = gd["finite_difference_h"]
    g = rco.gradient_fd(v, e, a, b, eps, h)
    q = float(np.spacing(abs(rco.energy(v, e, a, b, eps)))) / h
    assert np.abs(g - np.asarray(gd["dE"])).max() <= 8 * q
    # a backward difference or a missing 1/h fails it by orders of magnitude
    assert np.abs(-rco.gradient_fd(v, e, a, b, eps, -h) - np.asarray(gd["dE"])).max() > 1e3 * q


def test_raw_step_composition_oracle():
    # optimizer.test.ts:21-42: v - stepSize * grad, energy at the new vertices,
    # input untouched; field is the dE taken
    v = np.array([[-1, 0, 0.3], [1, 0, 0.3], [0, -1, -0.3], [0, 1, -0.3]], float)
    e = np.array([[0, 1], [2, 3]], np.int32)
    keep = v.copy()
    out, en, fld = rco.step(v, e, rco.MODE_ANALYTICAL, 0.001)
    g = rco.gradient(v, e)
    assert np.array_equal(out, v - 0.001 * g)
    assert en == rco.energy(v - 0.001 * g, e)
    assert np.array_equal(fld, g) and np.array_equal(v, keep)
    out, en, fld = rco.step(v, e, rco.MODE_FINITE_DIFF, 0.001)
    assert np.array_equal(fld, rco.gradient_fd(v, e))
    assert np.array_equal(out, v - 0.001 * fld)


@pytest.mark.parametrize("name", STAGE1)
def test_stage1_solve_and_line_search(goldens, name):
    fx = goldens["fixtures"][name]
    gd = goldens["goldens"][name]
    v, e = fx["vertices"], fx["edges"]
    cs = rco.ConstraintSet().barycenter(gd["x0_barycenter_target"])
    for lu in (0, 1):
        rc, gt, lam, res = rco.solve_constrained_gradient(v, e, gd["dE"], cs, factor_lu=lu)
        assert rc == 0
        assert frob_rel(gt, gd["g_tilde"]) <= 1e-9
        assert frob_rel(lam, gd["lambda"]) <= 1e-9
        assert res <= 1e-10
    nrm = rco.l2_curve_norm(gd["g_tilde"], v, e)
    assert abs(nrm - gd["gradient_l2_norm"]) / gd["gradient_l2_norm"] <= 1e-12
    # lineSearch.test.ts pins the stage-1 projection tolerance 1e-10 (reassemble).
    opts = rco.default_opts(tol_abs=1e-10, tol_rel=1e-10, projection_frozen=0)
    out, st = rco.line_search(v, e, gd["dE"], gd["g_tilde"], cs, opts)
    ls = gd["line_search_step"]
    assert st["accepted"] == ls["accepted"]
    assert st["tau"] == ls["tau"]
    assert st["projection_iterations"] == ls["projection_iterations"]
    assert abs(st["energy"] - ls["energy_after"]) / abs(ls["energy_after"]) <= 1e-12
    assert frob_rel(out, ls["vertices"]) <= 1e-9
    assert st["energy"] <= st["energy_before"] - 1e-4 * st["tau"] * st["slope"]


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


def build_set(gd):
    cs = rco.ConstraintSet()
    for kind in gd["block_kinds"]:
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


def golden_penalties(gd):
    p = gd.get("penalties")
    if not p:
        return None
    return rco.make_penalties(p["w_length"], p["w_diff"], p["w_field"], p["X"])


@pytest.mark.parametrize("gname,fname", SET_GOLDENS)
def test_constraint_set_goldens(goldens, gname, fname):
    fx = goldens["fixtures"][fname]
    gd = goldens["goldens"][gname]
    v, e = fx["vertices"], fx["edges"]
    cs = build_set(gd)
    phi, C, scales, rows = rco.eval_constraints(v, e, cs)
    assert np.allclose(phi, gd["Phi_stacked"], rtol=0, atol=1e-13)
    assert frob_rel(C, gd["C_stacked"]) <= 1e-14
    assert np.allclose(scales, gd["block_scales"], rtol=1e-14, atol=0)
    assert list(rows) == gd["block_row_counts"]
    pen = golden_penalties(gd)
    if pen is not None:
        assert abs(rco.penalty_energy(v, e, pen) - gd["penalty_energy_initial"]) <= 1e-12 * max(
            1, abs(gd["penalty_energy_initial"]))
        assert frob_rel(rco.penalty_gradient(v, e, pen).T.ravel(), gd["dE_penalty_flat"]) <= 1e-12
    rc, gt, lam, res = rco.solve_constrained_gradient(v, e, gd["dE"], cs)
    assert rc == 0
    assert frob_rel(gt, gd["g_tilde"]) <= 1e-9
    assert frob_rel(lam, gd["lambda"]) <= 1e-9
    assert res <= 1e-10
    frozen = 1 if gd["projection_mode"] == "frozen" else 0
    kw = {}
    if pen is not None:
        kw["penalties"] = pen
    opts = rco.default_opts(projection_frozen=frozen, **kw)
    out, st = rco.line_search(v, e, gd["dE"], gd["g_tilde"], cs, opts)
    ls = gd["line_search_step"]
    assert st["accepted"] == ls["accepted"]
    assert st["tau"] == ls["tau"]
    assert st["projection_iterations"] == ls["projection_iterations"]
    assert abs(st["energy"] - ls["energy_after"]) / abs(ls["energy_after"]) <= 1e-12
    assert frob_rel(out, ls["vertices"]) <= 1e-9


def test_gradient_central_fd():
    # test_gradient.ts:25-82: |fd - an| <= 1e-6 + 1e-5 * max|an|, h = 1e-6.
    rng = np.random.default_rng(7)
    t = 2 * np.pi * np.arange(24) / 24
    v = np.stack([np.cos(t), np.sin(t), 0.3 * np.sin(3 * t)], 1) + 0.01 * rng.standard_normal((24, 3))
    e = np.stack([np.arange(24), (np.arange(24) + 1) % 24], 1)
    g = rco.gradient(v, e)
    h = 1e-6
    fd = np.zeros_like(v)
    for i in range(len(v)):
        for c in range(3):
            vp, vm = v.copy(), v.copy()
            vp[i, c] += h
            vm[i, c] -= h
            fd[i, c] = (rco.energy(vp, e) - rco.energy(vm, e)) / (2 * h)
    assert np.all(np.abs(fd - g) <= 1e-6 + 1e-5 * np.abs(g).max())


def test_ldlt_exchange_kat():
    # ldlt.test.ts:238-249: the 2x2 exchange KAT yields ipiv = [-2, -2].
    rc, m, ipiv = rco.ldlt([[0.0, 1.0], [1.0, 0.0]])
    assert rc == 0 and list(ipiv) == [-2, -2]
    x = rco.ldlt_solve(m, ipiv, [3.0, 5.0])
    assert np.allclose(x, [5.0, 3.0], rtol=0, atol=1e-15)


def test_ldlt_reconstruction_random():
    # ldlt.test.ts:210-236 property: reconstructed solve residual <= 1e-12.
    rng = np.random.default_rng(3)
    for n in (5, 17, 40):
        M = rng.standard_normal((n, n))
        K = M + M.T
        rc, m, ipiv = rco.ldlt(K)
        assert rc == 0
        b = rng.standard_normal(n)
        x = rco.ldlt_solve(m, ipiv, b)
        assert np.linalg.norm(K @ x - b) <= 1e-12 * np.linalg.norm(K, 2) * np.linalg.norm(x)


def test_bh_energy_oracle_vs_reference_vectors():
    # Barnes-Hut energy (tpe_stage2_oracle.py:456-596): the reference oracle's
    # own outputs (tests/golden/make_bh_goldens.py), 5 curves x 6 settings incl.
    # theta = 0 (exact) and clustering at leaf 2/4/8: energy 1e-12 and the
    # traversal counts exactly (they pin the tree and the admissibility rule)
    import json
    import os
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "bh_goldens.json")))
    assert len(G["cases"]) == 30
    clustered = 0
    for c in G["cases"]:
        sh = G["shapes"][c["shape"]]
        en, st = rco.energy_bh(sh["vertices"], sh["edges"], c["theta"], c["theta_t"],
                               c["leaf_size"])
        assert abs(en - c["energy"]) <= 1e-12 * abs(c["energy"]), c
        assert st == (c["visits"], c["approximated_edge_refs"], c["direct_pairs"]), c
        if c["theta"] == 0.0:
            assert st[1] == 0
            assert abs(en - rco.energy(sh["vertices"], sh["edges"])) <= 1e-12 * abs(en)
        clustered += st[1] > 0
    assert clustered >= 10


def test_bct_products_oracle_vs_reference_vectors():
    # block-cluster-tree products (tpe_stage2_oracle.py:598-795): K psi for a
    # scalar and a 3-vector density and the Gram apply B u, both kernels, four
    # admissibility settings on five curves: 1e-12 (Frobenius-relative) and the
    # block counts exactly
    import json
    import os
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "bh_goldens.json")))
    admitted = 0
    for c in G["bct"]:
        sh = G["shapes"][c["shape"]]
        args = (c["kind"], c["theta"], c["theta_t"], c["leaf_size"])
        p1, st = rco.bct_matvec(sh["vertices"], sh["edges"], c["psi1"], *args)
        p3, _ = rco.bct_matvec(sh["vertices"], sh["edges"], c["psi3"], *args)
        bu = rco.bct_apply_b(sh["vertices"], sh["edges"], c["u"], *args)
        assert frob_rel(p1, c["phi1"]) <= 1e-12 and frob_rel(p3, c["phi3"]) <= 1e-12, c["shape"]
        assert frob_rel(bu, c["Bu"]) <= 1e-12
        s = c["stats"]
        assert st == (s["leaves"], s["admissible_blocks"], s["direct_blocks"], s["direct_pairs"])
        admitted += st[1] > 0
    assert admitted >= 10


def test_bh_gradient_oracle_vs_reference_vectors():
    # forward difference of the Barnes-Hut energy (tpe_stage2_oracle.py:
    # 133-141, 1289-1290): agreement to the resolution of a difference of two
    # energies (quanta of ulp(E)/h), a handful of quanta here
    import json
    import os
    G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "bh_goldens.json")))
    for c in G["bh_gradients"]:
        sh = G["shapes"][c["shape"]]
        g = rco.gradient_bh_fd(sh["vertices"], sh["edges"], c["theta"], c["theta_t"],
                               c["leaf_size"], c["h"])
        q = float(np.spacing(rco.energy(sh["vertices"], sh["edges"]))) / c["h"]
        assert np.abs(g - np.asarray(c["dE"])).max() <= 8 * q, c["shape"]
