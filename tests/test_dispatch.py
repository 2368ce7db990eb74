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
