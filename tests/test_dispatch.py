"""Host side of the solver-worker protocol and the descent dispatch (no device
calls): constraint-set construction order and validation
(src/core/dispatch.ts:213-242, src/core/sobolev/constraintSet.ts:333-344) and
the protocol errors of src/worker/solverWorker.ts:34-98 (a step or field
request before any topology message posts {type:'error'} containing
'before topology'; test/worker-solver.test.ts cases of the same names)."""
import numpy as np
import pytest

from rcurves import core, dispatch


def fixture():
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.2]], float)
    e = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    return v, e


def test_set_order_barycenter_length_pins():
    v, e = fixture()
    pins = [{"vertexIndex": 2, "target": [1, 1, 0], "enabled": True},
            {"vertexIndex": 1, "target": [1, 0, 0], "enabled": False},
            {"vertexIndex": 9, "target": [0, 0, 0], "enabled": True},   # out of range: skipped
            {"vertexIndex": 0, "target": [0, 0, 0], "enabled": True}]
    cs = dispatch.build_constraint_set({"vertices": v, "edges": e, "x0": [0.5, 0.5, 0.05],
                                        "sobolevL0": 4.0, "barycenterConstraint": True,
                                        "lengthMode": "total", "pins": pins})
    assert [b.kind for b in cs] == ["barycenter", "totalLength", "point", "point"]
    assert [b.vertex for b in cs[2:]] == [2, 0]
    assert cs[1].L0 == 4.0


def test_set_per_edge_lengths_and_disabled_blocks():
    v, e = fixture()
    ell0 = core.edge_lengths(v, e)
    cs = dispatch.build_constraint_set({"vertices": v, "edges": e, "x0": [0, 0, 0],
                                        "barycenterConstraint": False, "lengthMode": "perEdge",
                                        "sobolevEll0": ell0})
    assert [b.kind for b in cs] == ["edgeLengths"]
    np.testing.assert_array_equal(cs[0].ell0, ell0)
    cs = dispatch.build_constraint_set({"vertices": v, "edges": e, "x0": [0, 0, 0],
                                        "lengthConstraint": False})
    assert [b.kind for b in cs] == ["barycenter"]


def test_length_blocks_are_mutually_exclusive():
    v, e = fixture()
    with pytest.raises(ValueError, match="mutually exclusive"):
        core.assert_valid_constraint_set([core.total_length_block(1.0),
                                          core.edge_lengths_block(core.edge_lengths(v, e))])


def test_worker_step_before_topology_is_an_error():
    w = dispatch.SolverWorker()
    v, _ = fixture()
    resp = w.handle({"type": "step", "graphVersion": 1, "args": {"vertices": v, "x0": [0, 0, 0]}})
    assert resp["type"] == "error" and "before topology" in resp["message"]


def test_worker_field_before_topology_is_an_error():
    w = dispatch.SolverWorker()
    v, _ = fixture()
    resp = w.handle({"type": "field", "graphVersion": 1, "args": {"vertices": v, "x0": [0, 0, 0]}})
    assert resp["type"] == "error" and "before topology" in resp["message"]


def test_worker_topology_message_has_no_reply_and_is_cached():
    w = dispatch.SolverWorker()
    _, e = fixture()
    assert w.handle({"type": "topology", "graphVersion": 4, "edges": e}) is None
    assert w.topology["graphVersion"] == 4
    np.testing.assert_array_equal(w.topology["edges"], e)


def test_raw_step_with_unknown_gradient_mode_is_an_error():
    # the mode is checked on the host before any device work
    v, e = fixture()
    w = dispatch.SolverWorker()
    w.handle({"type": "topology", "graphVersion": 1, "edges": e})
    out = w.handle({"type": "step", "graphVersion": 1,
                    "args": {"vertices": v, "x0": [0, 0, 0], "descentMode": "raw",
                             "mode": "centralDiff", "stepSize": 1e-3}})
    assert out["type"] == "error" and "unknown gradient mode" in out["message"]


def test_legacy_branch_passes_the_mode_alone(monkeypatch):
    """dispatch.ts:187-210: with no M-set toggle the barycenter-only
    sobolevStep runs with {mode} only (default projection, no penalties, no
    E0 reuse, nothing collected) and returns null timings and field."""
    seen = {}

    def fake_step(vertices, edges, disjoint_pairs, x0, opts):
        seen["x0"], seen["opts"] = x0, opts
        return {"vertices": vertices, "energy": 1.5, "accepted": True, "converged": False,
                "stats": {"tau": 1.0}, "timings": {"step": {"ms": 1.0, "calls": 1}},
                "descentField": np.zeros_like(vertices)}

    monkeypatch.setattr(core, "sobolev_step", fake_step)
    monkeypatch.setattr(core, "sobolev_step_set", lambda *a, **k: pytest.fail("toggle path taken"))
    v, e = fixture()
    out = dispatch.dispatch_descent_step({
        "descentMode": "sobolev", "vertices": v, "edges": e, "x0": [0.5, 0.5, 0.0],
        "mode": "finiteDiff", "collectTimings": True, "collectField": True,
        "energyBefore": 3.0, "projectionMode": "frozen", "penalties": {"totalLength": 2.0},
        "precision": "mixed"})
    o = seen["opts"]
    default = core.SobolevStepOptions()
    assert o.mode == "finiteDiff"
    assert o.projection_mode == default.projection_mode == "reassemble"
    assert o.penalties is None and o.energy_before is None
    assert not o.collect_timings and not o.collect_field
    assert o.precision == default.precision
    assert seen["x0"] == [0.5, 0.5, 0.0]
    assert out["timings"] is None and out["descentField"] is None
    assert out["energy"] == 1.5 and out["accepted"] is True and out["stats"] == {"tau": 1.0}


def test_toggle_present_takes_the_set_path(monkeypatch):
    """Any one toggle (even False / 'none') selects the ConstraintSet path."""
    calls = []
    monkeypatch.setattr(core, "sobolev_step", lambda *a, **k: pytest.fail("legacy path taken"))
    monkeypatch.setattr(core, "sobolev_step_set",
                        lambda v, e, dp, cs, opts: calls.append(([b.kind for b in cs], opts)) or
                        {"vertices": v, "energy": 0.0, "accepted": False, "converged": False,
                         "stats": {}})
    v, e = fixture()
    base = {"descentMode": "sobolev", "vertices": v, "edges": e, "x0": [0, 0, 0],
            "mode": "analytical", "sobolevL0": 4.0}
    for extra, kinds in (({"barycenterConstraint": False}, ["totalLength"]),
                         ({"lengthConstraint": False}, ["barycenter"]),
                         ({"lengthMode": "none"}, ["barycenter"])):
        out = dispatch.dispatch_descent_step(dict(base, **extra))
        assert calls[-1][0] == kinds
        assert out["timings"] is None and out["descentField"] is None


def test_absent_descent_mode_is_the_raw_step(monkeypatch):
    """dispatch.ts:185-188 takes the Sobolev branch only for 'sobolev'."""
    seen = {}

    def fake_raw(vertices, edges, dp, opts):
        seen["opts"] = opts
        return {"vertices": vertices, "energy": 2.0}

    monkeypatch.setattr(core, "step", fake_raw)
    v, e = fixture()
    out = dispatch.dispatch_descent_step({"vertices": v, "edges": e, "x0": [0, 0, 0],
                                          "mode": "analytical", "stepSize": 1e-3})
    assert seen["opts"].step_size == 1e-3
    assert out["stats"] is None and out["timings"] is None and out["descentField"] is None
    assert out["accepted"] is True and out["converged"] is False


def test_tangent_point_energy_backward_needs_forward():
    tpe = core.TangentPointEnergy(fixture()[1], 4)
    with pytest.raises(RuntimeError, match="before forward"):
        tpe.backward()
