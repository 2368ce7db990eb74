"""The host boundary on the device: dispatchDescentStep (dispatch.ts:185-290)
against the oracle step for the same arguments, its pre-toggle legacy branch,
the pin rules of store-pins.test.ts:116-200, the TangentPointEnergy
forward/backward pair, and the phase ledger (phaseTimings.ts:9-78): keys,
call counts and nesting for accepted, converged, singular and reassembled
steps."""
import sys

import numpy as np
import pytest

from oracle import rc_oracle as rco
from rcurves import TangentPointEnergy, core, curves, dispatch

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def crossing(goldens):
    fx = goldens["fixtures"]["crossing"]
    return np.asarray(fx["vertices"], float), np.asarray(fx["edges"], np.int32)


def same_step(r, out, st):
    """lineSearch.test.ts gates: accepted and tau and projection iterations
    exact, energy <= 1e-12, vertices <= 1e-9."""
    assert r["accepted"] == st["accepted"]
    assert r["stats"]["tau"] == st["tau"]
    assert r["stats"]["projectionIterations"] == st["projection_iterations"]
    assert abs(r["energy"] - st["energy"]) <= 1e-12 * abs(st["energy"])
    assert rel(r["vertices"], out) <= 1e-9


@pytest.mark.parametrize("projection", ["frozen", "reassemble"])
def test_dispatch_sobolev_matches_oracle_step(goldens, projection):
    v, e = crossing(goldens)
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    target = [v[0][0] + 0.05, v[0][1] - 0.05, v[0][2] + 0.05]
    log("dispatch vs oracle", projection)
    args = {"descentMode": "sobolev", "vertices": v, "edges": e, "disjointPairs": None,
            "mode": "analytical", "stepSize": 1e-3, "x0": x0, "sobolevL0": L0,
            "barycenterConstraint": True, "lengthMode": "total", "projectionMode": projection,
            "pins": [{"vertexIndex": 0, "target": target, "enabled": True}],
            "penalties": {"totalLength": 0.5, "field": {"weight": 2.0, "X": [0.0, 0.0, 1.0]}},
            "collectField": True}
    r = dispatch.dispatch_descent_step(args)
    pen = rco.make_penalties(0.5, 0.0, 2.0, (0.0, 0.0, 1.0))
    out, gto, st = rco.sobolev_step(
        v, e, rco.ConstraintSet().barycenter(x0).total_length(L0).point(0, target),
        rco.default_opts(projection_frozen=int(projection == "frozen"), penalties=pen))
    log("device", r["stats"], "oracle", st)
    same_step(r, out, st)
    assert rel(r["descentField"], gto) <= 1e-9
    assert r["timings"] is None


def test_dispatch_legacy_branch_matches_oracle_barycenter_step(goldens):
    """No toggles: sobolevStep(x0) with the mode alone (dispatch.ts:187-210):
    reassembled projection, no penalties, E0 recomputed; timings and
    descentField null whatever the caller asked for."""
    v, e = crossing(goldens)
    x0 = rco.barycenter_target(v, e)
    log("legacy dispatch")
    args = {"descentMode": "sobolev", "vertices": v, "edges": e, "disjointPairs": None,
            "mode": "analytical", "stepSize": 1e-3, "x0": x0,
            "collectTimings": True, "collectField": True, "energyBefore": 123.0,
            "projectionMode": "frozen", "penalties": {"totalLength": 3.0}}
    r = dispatch.dispatch_descent_step(args)
    out, _, st = rco.sobolev_step(v, e, rco.ConstraintSet().barycenter(x0),
                                  rco.default_opts(projection_frozen=0))
    log("device", r["stats"], "oracle", st)
    same_step(r, out, st)
    assert r["timings"] is None and r["descentField"] is None


def test_dispatch_pins_rules_match_direct_steps(goldens):
    """store-pins.test.ts:116-200: an enabled in-range pin appends a point
    block after the length block; disabled, out-of-range and absent pins add
    nothing."""
    v, e = crossing(goldens)
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    target = [v[0][0] + 0.1, v[0][1] - 0.2, v[0][2] + 0.3]
    base = {"descentMode": "sobolev", "vertices": v, "edges": e, "disjointPairs": None,
            "mode": "analytical", "stepSize": 1e-3, "x0": x0, "sobolevL0": L0,
            "sobolevEll0": core.edge_lengths(v, e), "barycenterConstraint": True,
            "lengthMode": "total"}
    opts = core.SobolevStepOptions(mode="analytical")
    log("pins rules")
    with_pin = dispatch.dispatch_descent_step(
        dict(base, pins=[{"vertexIndex": 0, "target": target, "enabled": True}]))
    ref = core.sobolev_step_set(v, e, None, [core.barycenter_block(x0), core.total_length_block(L0),
                                             core.point_block(0, target)], opts)
    assert with_pin["energy"] == ref["energy"]
    assert np.array_equal(with_pin["vertices"], ref["vertices"])
    no_pin = core.sobolev_step_set(v, e, None,
                                   [core.barycenter_block(x0), core.total_length_block(L0)], opts)
    for pins in ([{"vertexIndex": 0, "target": target, "enabled": False}],
                 [{"vertexIndex": 999, "target": target, "enabled": True}], None):
        args = dict(base) if pins is None else dict(base, pins=pins)
        got = dispatch.dispatch_descent_step(args)
        assert got["energy"] == no_pin["energy"]
        assert np.array_equal(got["vertices"], no_pin["vertices"])


def test_tangent_point_energy_forward_backward_match_oracle():
    v, e = curves.torus_knot(2, 3, 256, jitter=0.02, seed=3)
    log("TangentPointEnergy")
    tpe = TangentPointEnergy(e, len(v))
    E = tpe.forward(v)
    Eo = rco.energy(v, e)
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    go = rco.gradient(v, e)
    assert rel(tpe.backward(), go) <= 1e-12
    assert rel(tpe.backward(2.5), 2.5 * go) <= 1e-12
    # backward uses the positions saved by forward, not later edits of them
    v2 = v.copy()
    tpe.forward(v2)
    v2[:] = 0.0
    assert rel(tpe.backward(), go) <= 1e-12
    tpe2 = TangentPointEnergy(e, len(v), alpha=2.0, beta=4.5, epsilon=1e-8)
    assert abs(tpe2.forward(v) - rco.energy(v, e, 2.0, 4.5, 1e-8)) <= 1e-12 * abs(
        rco.energy(v, e, 2.0, 4.5, 1e-8))
    assert rel(tpe2.backward(), rco.gradient(v, e, 2.0, 4.5, 1e-8)) <= 1e-12


KEYS = {"dE", "energy", "assembleA", "saddle", "factor", "projection", "lineSearch", "step"}


def ledger_ok(tm):
    assert set(tm) <= KEYS
    for k, s in tm.items():
        assert s["calls"] >= 1 and s["ms"] >= 0.0, (k, s)
    assert tm["step"]["calls"] == 1 and tm["dE"]["calls"] == 1
    # sub-phases inside their parents (device event time, 5 % slack for
    # event resolution)
    step = tm["step"]["ms"] * 1.05 + 0.05
    for k in tm:
        assert tm[k]["ms"] <= step, (k, tm)
    if "factor" in tm:
        assert tm["factor"]["ms"] <= tm["saddle"]["ms"] * 1.05 + 0.05
    if "lineSearch" in tm:
        inner = tm.get("projection", {"ms": 0.0})["ms"]
        assert inner <= tm["lineSearch"]["ms"] * 1.05 + 0.05


def test_ledger_accepted_frozen_step():
    v, e = curves.trefoil_jittered(256, 0.01, seed=8)
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    cs = [core.barycenter_block(x0), core.total_length_block(L0), core.point_block(3, v[3])]
    log("ledger frozen")
    for e0 in (None, 1.0e6):
        r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
            projection_mode="frozen", collect_timings=True, energy_before=e0))
        tm = r["timings"]
        log(r["stats"], r["diagnostics"]["trials"], tm)
        assert r["accepted"]
        ledger_ok(tm)
        assert set(tm) == KEYS
        trials = r["diagnostics"]["trials"]
        assert tm["projection"]["calls"] == trials and tm["projection"]["ms"] > 0.0
        # every trial's projection converged here: each trial evaluates an energy
        assert tm["energy"]["calls"] == trials + (1 if e0 is None else 0)
        assert tm["assembleA"]["calls"] == 1 and tm["factor"]["calls"] == 1
        assert tm["saddle"]["calls"] >= 1 + r["stats"]["projectionIterations"]
        assert tm["lineSearch"]["calls"] == 1


def test_ledger_reassembled_projection():
    v, e = curves.trefoil_jittered(128, 0.01, seed=9)
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    log("ledger reassemble")
    r = core.sobolev_step_set(v, e, None, [core.barycenter_block(x0), core.total_length_block(L0)],
                              core.SobolevStepOptions(projection_mode="reassemble",
                                                      collect_timings=True))
    tm = r["timings"]
    log(r["stats"], tm)
    assert r["accepted"]
    ledger_ok(tm)
    corr = tm["saddle"]["calls"] - 1
    assert corr >= r["stats"]["projectionIterations"] >= 1
    # every correction reassembles A and refactors (lineSearch.ts:263-276)
    assert tm["assembleA"]["calls"] == 1 + corr
    assert tm["factor"]["calls"] == 1 + corr


def test_ledger_converged_and_singular_steps():
    v, e = curves.trefoil_jittered(32, 0.01, seed=2)
    log("ledger singular")
    vi = np.vstack([v, [[10.0, 0.0, 0.0]]])
    r = core.sobolev_step_set(vi, e, None, [core.barycenter_block(rco.barycenter_target(vi, e))],
                              core.SobolevStepOptions(collect_timings=True))
    assert r["stats"]["reason"] == "singular_system"
    tm = r["timings"]
    ledger_ok(tm)
    # the reference's factor throws inside both wraps: neither is recorded,
    # nor is the line search it never reaches; the echoed energy is
    assert set(tm) == {"step", "dE", "energy", "assembleA"}
    log("ledger converged")
    # a regular polygon under [barycenter, total length]: dE is a multiple of
    # the length row, g~ vanishes and the step converges (optimizer.ts:355-367)
    n = 64
    t = 2 * np.pi * np.arange(n) / n
    vc = np.stack([np.cos(t), np.sin(t), np.zeros(n)], 1)
    ec = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    cs = [core.barycenter_block([0.0, 0.0, 0.0]), core.total_length_block(core.total_length(vc, ec))]
    for e0 in (None, 1.0):
        r = core.sobolev_step_set(vc, ec, None, cs,
                                  core.SobolevStepOptions(collect_timings=True, energy_before=e0))
        log(r["stats"], r["converged"], r["timings"])
        assert r["converged"] is True
        tm = r["timings"]
        ledger_ok(tm)
        assert "lineSearch" not in tm and "projection" not in tm
        assert tm["saddle"]["calls"] == 1 and tm["factor"]["calls"] == 1
        # the echo recomputes E0 unless it is given
        assert ("energy" in tm) == (e0 is None)
