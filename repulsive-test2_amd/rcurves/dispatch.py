"""Descent dispatch and the solver-worker message protocol over the device path.

Mirrors src/core/dispatch.ts:185-290 (set construction order barycenter ->
length -> enabled in-range pins, rank-rule check, sobolev step) and
src/worker/solverWorker.ts:34-98 (topology cache per graphVersion, step
before topology -> error message, graphVersion echoed untouched).
"""
from __future__ import annotations

from typing import Any, Optional

import numpy as np

from . import core


def build_constraint_set(args: dict) -> list:
    cs = []
    if args.get("barycenterConstraint", True):
        cs.append(core.barycenter_block(args["x0"]))
    length_mode = args.get("lengthMode")
    if length_mode is None:
        length_mode = "total" if args.get("lengthConstraint", True) else "none"
    if length_mode == "total":
        L0 = args.get("sobolevL0")
        cs.append(core.total_length_block(float("nan") if L0 is None else L0))
    elif length_mode == "perEdge":
        ell0 = args.get("sobolevEll0")
        E = len(args["edges"])
        cs.append(core.edge_lengths_block(np.full(E, np.nan) if ell0 is None else ell0))
    V = len(args["vertices"])
    for pin in args.get("pins") or []:
        if pin.get("enabled") and 0 <= pin["vertexIndex"] < V:
            cs.append(core.point_block(pin["vertexIndex"], pin["target"]))
    core.assert_valid_constraint_set(cs)
    return cs


def dispatch_descent_step(args: dict) -> dict:
    """dispatchDescentStep for descentMode 'sobolev' (the device path)."""
    if args.get("descentMode", "sobolev") != "sobolev":
        raise core.L.RcError(core.L.RC_UNSUPPORTED, "raw descent mode is not provided by the device path")
    if ("barycenterConstraint" not in args and "lengthConstraint" not in args
            and "lengthMode" not in args):
        cs = [core.barycenter_block(args["x0"])]
    else:
        cs = build_constraint_set(args)
    pen = args.get("penalties")
    if isinstance(pen, dict):
        f = pen.get("field") or {}
        pen = core.PenaltyConfig(pen.get("totalLength", 0.0), pen.get("lengthDiff", 0.0),
                                 f.get("weight", 0.0), tuple(f.get("X", (0.0, 0.0, 0.0))))
    opts = core.SobolevStepOptions(
        mode=args.get("mode", "analytical"),
        collect_timings=bool(args.get("collectTimings", False)),
        collect_field=bool(args.get("collectField", False)),
        energy_before=args.get("energyBefore"),
        penalties=pen,
        projection_mode=args.get("projectionMode", "frozen"),
        precision=args.get("precision", "fp64"),
    )
    r = core.sobolev_step_set(args["vertices"], args["edges"], args.get("disjointPairs"), cs, opts)
    return {
        "vertices": r["vertices"],
        "energy": r["energy"],
        "accepted": r["accepted"],
        "converged": r["converged"],
        "stats": r["stats"],
        "timings": r.get("timings"),
        "descentField": r.get("descentField"),
    }


class SolverWorker:
    """Stateless compute server holding only a topology cache
    (solverWorker.ts:27-43). handle() returns the response message."""

    def __init__(self):
        self.topology: Optional[dict] = None

    def handle(self, msg: dict) -> Optional[dict]:
        try:
            if msg["type"] == "topology":
                edges = np.asarray(msg["edges"], dtype=np.int32)
                self.topology = {"graphVersion": msg["graphVersion"], "edges": edges}
                return None
            if msg["type"] == "field":
                if self.topology is None:
                    raise RuntimeError("solverWorker: field received before topology")
                a = msg["args"]
                cs = [core.barycenter_block(a["x0"])]
                r = core.sobolev_step_set(a["vertices"], self.topology["edges"], None, cs,
                                          core.SobolevStepOptions(collect_field=True))
                return {"type": "fieldResult", "graphVersion": msg["graphVersion"],
                        "field": r.get("descentField")}
            if self.topology is None:
                raise RuntimeError("solverWorker: step received before topology")
            args: dict[str, Any] = dict(msg["args"])
            args["edges"] = self.topology["edges"]
            args["disjointPairs"] = None
            result = dispatch_descent_step(args)
            return {"type": "result", "graphVersion": msg["graphVersion"], "result": result}
        except Exception as err:  # any throw becomes a protocol error (solverWorker.ts:89-98)
            return {"type": "error", "message": str(err)}
