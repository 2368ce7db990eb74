"""Forward-difference gradient, the raw L2 step and their dispatch paths,
through the C ABI.

- gradientFiniteDiff (tangentPointEnergy.ts:114-136): the device forms
  E(x + h e_vd) - E(x) from the per-pair differences of the pairs that the
  move changes; the reference differences two full energies. Both are the
  same quotient; they differ by the rounding of the full-energy difference
  (quanta of one unit in the last place of E divided by h, accumulated over
  the P pair terms of each sum). The gate is 2 sqrt(P) quanta per component
  against the reference's own FD goldens (oracle/golden/*.json "dE",
  h = 1e-6) and against the oracle; at sizes the oracle cannot reach, the
  first-order convergence to the adjoint gradient in h.
- step (optimizer.ts:40-68) mirrors test/optimizer.test.ts:21-42: the output
  is exactly v - stepSize * dE with dE from the same library, the energy is
  exactly the energy at the output, the input is not mutated.
- dispatch 'raw' and the worker's finiteDiff field mirror
  test/optimizer-sobolev.test.ts:109-131 and test/worker-solver.test.ts:389-430
  (bit-identical to the direct calls).
"""
import sys
import time
from fractions import Fraction

import numpy as np
import pytest

from oracle import rc_oracle as rco
from rcurves import core, curves, dispatch

pytestmark = pytest.mark.gpu

STAGE1 = ["crossing", "junction-y", "helix", "linked-rings", "knot"]
H = 1e-6


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fd_quantum(E, h=H):
    return float(np.spacing(abs(E))) / abs(h)


def fd_noise(E, E_edges, h=H):
    """Rounding resolution of the reference's full-energy forward difference:
    E0 and E1 are left-to-right sums of the P ordered pair terms whose
    rounding errors do not cancel in E1 - E0; 2 sqrt(P) quanta bound it
    (a wrong term, sign or coordinate, or the O(h) of another difference
    formula, is orders of magnitude larger)."""
    P = max(E_edges * (E_edges - 3), 1)
    return 2.0 * np.sqrt(P) * fd_quantum(E, h)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", STAGE1)
def test_fd_gradient_vs_reference_goldens(goldens, name):
    fx, gd = goldens["fixtures"][name], goldens["goldens"][name]
    v, e = np.asarray(fx["vertices"], float), np.asarray(fx["edges"], np.int32)
    a, b, eps = fx["alpha"], fx["beta"], fx["epsilon"]
    h = gd.get("finite_difference_h", H)
    log("fd golden", name, len(v), len(e))
    g = core.gradient_finite_diff(v, e, None, a, b, eps, h)
    E = rco.energy(v, e, a, b, eps)
    tol = fd_noise(E, len(e), h)
    assert np.abs(g - np.asarray(gd["dE"])).max() <= tol, (name, np.abs(g - gd["dE"]).max(), tol)
    go = rco.gradient_fd(v, e, a, b, eps, h)
    assert np.abs(g - go).max() <= tol


def test_fd_gradient_vs_oracle_larger():
    cases = [("trefoil128", *curves.trefoil_jittered(128, 0.01, seed=1)),
             ("theta_net_192", *curves.theta_network(192, components=2, seed=2))]
    for name, v, e in cases:
        log("fd oracle", name, len(v), len(e))
        g = core.gradient_finite_diff(v, e, None, 3.0, 6.0, 1e-10, H)
        go = rco.gradient_fd(v, e, h=H)
        tol = fd_noise(rco.energy(v, e), len(e))
        assert np.abs(g - go).max() <= tol, (name, np.abs(g - go).max(), tol)
        # the forward difference is O(h) from the analytic gradient
        ga = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
        assert rel(g, ga) <= 1e-4


def test_fd_gradient_generic_exponents():
    v, e = curves.trefoil_jittered(64, 0.01, seed=4)
    log("fd generic exponents")
    g = core.gradient_finite_diff(v, e, None, 2.0, 4.5, 1e-10, H)
    go = rco.gradient_fd(v, e, 2.0, 4.5, 1e-10, H)
    tol = fd_noise(rco.energy(v, e, 2.0, 4.5, 1e-10), len(e))
    assert np.abs(g - go).max() <= tol, (np.abs(g - go).max(), tol)


@pytest.mark.parametrize("n", [1024, 8192])
def test_fd_gradient_against_adjoint_at_size(n):
    # size-independent checks: the forward difference approaches the adjoint
    # gradient at first order in h (halving h halves the error, and the
    # Richardson combination 2 g(h/2) - g(h) removes most of it), and it is
    # bit-reproducible run to run
    v, e = curves.fourier_curve(n, seed=1)
    log("fd vs adjoint", n)
    t0 = time.perf_counter()
    g = core.gradient_finite_diff(v, e, None, 3.0, 6.0, 1e-10, H)
    t1 = time.perf_counter()
    log(f"  fd gradient {1e3 * (t1 - t0):.1f} ms (incl. transfers)")
    g2 = core.gradient_finite_diff(v, e, None, 3.0, 6.0, 1e-10, H / 2)
    ga = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
    r1, r2, rr = rel(g, ga), rel(g2, ga), rel(2 * g2 - g, ga)
    log(f"  rel err h {r1:.3e}  h/2 {r2:.3e}  richardson {rr:.3e}")
    assert r1 <= 0.1
    assert 1.6 <= r1 / r2 <= 2.4
    assert rr <= 0.2 * r1
    assert np.array_equal(g, core.gradient_finite_diff(v, e, None, 3.0, 6.0, 1e-10, H))


def test_fd_gradient_rejects_bad_h():
    v, e = curves.trefoil_jittered(32, 0.01, seed=1)
    for h in (0.0, float("nan"), float("inf")):
        with pytest.raises(core.L.RcError):
            core.gradient_finite_diff(v, e, None, 3.0, 6.0, 1e-10, h)


@pytest.mark.parametrize("mode", ["analytical", "finiteDiff"])
def test_raw_step_exact_composition(mode):
    # optimizer.test.ts:21-36 with the reference's 4-vertex two-edge fixture
    # and a jittered trefoil
    quad = (np.array([[-1, 0, 0.3], [1, 0, 0.3], [0, -1, -0.3], [0, 1, -0.3]], float),
            np.array([[0, 1], [2, 3]], np.int32))
    for v, e in (quad, curves.trefoil_jittered(128, 0.01, seed=1)):
        log("raw step", mode, len(v))
        keep = v.copy()
        s = 0.001
        if mode == "analytical":
            g = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
        else:
            g = core.gradient_finite_diff(v, e, None, 3.0, 6.0, 1e-10, H)
        expected = v - s * g
        out = core.step(v, e, None, core.StepOptions(mode=mode, step_size=s, collect_field=True))
        bad = np.argwhere(out["vertices"] != expected)
        if len(bad):
            i, c = bad[0]
            log("  mismatches", len(bad), "field equal", np.array_equal(out["field"], g),
                "first", i, c, repr(v[i, c]), repr(g[i, c]), repr(out["field"][i, c]),
                repr(out["vertices"][i, c]), repr(expected[i, c]),
                "exact-fma", repr(float(Fraction(v[i, c]) - Fraction(s) * Fraction(g[i, c]))))
        assert np.array_equal(out["field"], g)
        assert np.array_equal(out["vertices"], expected)
        assert out["energy"] == core.calculate_energy(expected, e, None, 3.0, 6.0, 1e-10)
        assert np.array_equal(out["field"], g)
        assert np.array_equal(v, keep)
        assert "field" not in core.step(v, e, None, core.StepOptions(mode=mode, step_size=s))


@pytest.mark.parametrize("mode", [0, 1])
def test_raw_step_vs_oracle(mode):
    v, e = curves.trefoil_jittered(128, 0.01, seed=2)
    s = 0.002
    log("raw step vs oracle", mode)
    out = core.step(v, e, None, core.StepOptions(mode=["analytical", "finiteDiff"][mode],
                                                 step_size=s))
    vo, eo, go = rco.step(v, e, mode, s)
    tol = 1e-12 * np.linalg.norm(go) if mode == 0 else fd_noise(rco.energy(v, e), len(e))
    dv = out["vertices"] - vo
    assert np.abs(dv).max() <= s * tol + 4 * np.spacing(np.abs(vo).max())
    # the energies differ by the rounding of the energy plus, to first order,
    # dE . (the difference of the two outputs)
    assert abs(out["energy"] - eo) <= 1e-12 * abs(eo) + 2 * np.linalg.norm(go) * np.linalg.norm(dv)


def test_sobolev_step_finite_diff_vs_oracle(goldens):
    fx, gd = goldens["fixtures"]["knot"], goldens["goldens"]["knot"]
    v, e = np.asarray(fx["vertices"], float), np.asarray(fx["edges"], np.int32)
    x0 = gd["x0_barycenter_target"]
    log("sobolev step finiteDiff, knot")
    r = core.sobolev_step_set(v, e, None, [core.barycenter_block(x0)],
                              core.SobolevStepOptions(mode="finiteDiff", collect_field=True,
                                                      projection_mode="frozen"))
    cs = rco.ConstraintSet().barycenter(x0)
    vo, gto, sto = rco.sobolev_step(v, e, cs, rco.default_opts(gradient_mode=1, fd_h=H))
    assert r["accepted"] == sto["accepted"] and r["stats"]["tau"] == sto["tau"]
    assert r["stats"]["projectionIterations"] == sto["projection_iterations"]
    # g~ inherits the forward difference's rounding (about 1e-9 relative here)
    assert rel(r["descentField"], gto) <= 1e-7
    assert abs(r["energy"] - sto["energy"]) <= 1e-9 * abs(sto["energy"])
    assert np.abs(r["vertices"] - vo).max() <= 1e-8
    # the analytic-mode step differs from it by the O(h) of the difference
    ra = core.sobolev_step_set(v, e, None, [core.barycenter_block(x0)],
                               core.SobolevStepOptions(collect_field=True, projection_mode="frozen"))
    assert rel(ra["descentField"], r["descentField"]) <= 1e-4


def test_dispatch_raw_matches_step_exactly():
    v, e = curves.trefoil_jittered(128, 0.01, seed=3)
    log("dispatch raw")
    w = dispatch.SolverWorker()
    w.handle({"type": "topology", "graphVersion": 7, "edges": e})
    for mode in ("analytical", "finiteDiff"):
        resp = w.handle({"type": "step", "graphVersion": 7,
                         "args": {"vertices": v, "x0": [0, 0, 0], "descentMode": "raw",
                                  "mode": mode, "stepSize": 1e-3, "collectField": True}})
        assert resp["type"] == "result" and resp["graphVersion"] == 7
        r = resp["result"]
        exp = core.step(v, e, None, core.StepOptions(mode=mode, step_size=1e-3,
                                                     collect_field=True))
        assert np.array_equal(r["vertices"], exp["vertices"])
        assert r["energy"] == exp["energy"]
        assert r["accepted"] is True and r["converged"] is False
        assert r["stats"] is None and r["timings"] is None
        assert np.array_equal(r["descentField"], exp["field"])


def test_worker_field_modes():
    v, e = curves.trefoil_jittered(128, 0.01, seed=5)
    x0 = core.barycenter_target(v, e)
    log("worker field modes")
    w = dispatch.SolverWorker()
    w.handle({"type": "topology", "graphVersion": 13, "edges": e})
    d = core.DEFAULTS

    def field(mode, dmode):
        resp = w.handle({"type": "field", "graphVersion": 13,
                         "args": {"vertices": v, "mode": mode, "descentMode": dmode, "x0": x0}})
        assert resp["type"] == "fieldResult" and resp["graphVersion"] == 13
        return resp["field"]

    fd = core.gradient_finite_diff(v, e, None, d["alpha"], d["beta"], d["epsilon"], d["h"])
    assert np.array_equal(field("finiteDiff", "raw"), fd)
    an = core.gradient_analytical(v, e, None, d["alpha"], d["beta"], d["epsilon"])
    assert np.array_equal(field("analytical", "raw"), an)
    gt, _, _ = core.solve_constrained_gradient_set_frozen(
        v, e, None, d["alpha"], d["beta"], d["epsilon"], fd, [core.barycenter_block(x0)])
    assert np.array_equal(field("finiteDiff", "sobolev"), gt)
