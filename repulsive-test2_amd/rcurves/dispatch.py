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


def compute_arrow_field(vertices, edges, mode: str, descent_mode: str, x0) -> Optional[np.ndarray]:
    """computeArrowField (src/core/arrowField.ts:24-74): dE by mode at the
    reference defaults; descentMode 'sobolev' maps it to the barycenter-only
    g~ (None when the saddle system is singular), 'raw' returns dE."""
    d = core.DEFAULTS
    if mode == "analytical":
        dE = core.gradient_analytical(vertices, edges, None, d["alpha"], d["beta"], d["epsilon"])
    else:
        dE = core.gradient_finite_diff(vertices, edges, None, d["alpha"], d["beta"], d["epsilon"],
                                       d["h"])
    if descent_mode == "sobolev":
        try:
            gt, _, _ = core.solve_constrained_gradient_set_frozen(
                vertices, edges, None, d["alpha"], d["beta"], d["epsilon"], dE,
                [core.barycenter_block(x0)])
            return gt
        except core.L.RcError as err:
            if err.status in (core.L.RC_SINGULAR, core.L.RC_NONFINITE):
                return None
            raise
    return dE


def dispatch_descent_step(args: dict) -> dict:
    """dispatchDescentStep (dispatch.ts:185-290): 'sobolev' -> the constrained
    Sobolev step, anything else -> step() with the frame loop's arguments."""
    if args.get("descentMode") != "sobolev":
        r = core.step(args["vertices"], args["edges"], args.get("disjointPairs"),
                      core.StepOptions(mode=args.get("mode", "analytical"),
                                       step_size=args["stepSize"],
                                       collect_field=bool(args.get("collectField", False))))
        return {
            "vertices": r["vertices"],
            "energy": r["energy"],
            "accepted": True,
            "converged": False,
            "stats": None,
            "timings": None,
            "descentField": r.get("field"),
        }
    if (args.get("barycenterConstraint") is None and args.get("lengthConstraint") is None
            and args.get("lengthMode") is None):
        # pre-toggle call shape (dispatch.ts:187-210): the barycenter-only
        # sobolevStep with the mode alone (default projection, no penalties,
        # no E0 reuse), never timings or a descent field
        r = core.sobolev_step(args["vertices"], args["edges"], args.get("disjointPairs"),
                              args["x0"], core.SobolevStepOptions(mode=args.get("mode", "analytical")))
        return {
            "vertices": r["vertices"],
            "energy": r["energy"],
            "accepted": r["accepted"],
            "converged": r["converged"],
            "stats": r["stats"],
            "timings": None,
            "descentField": None,
        }
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
        projection_mode=args.get("projectionMode") or "reassemble",
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
                fld = compute_arrow_field(a["vertices"], self.topology["edges"],
                                          a.get("mode", "analytical"),
                                          a.get("descentMode"), a["x0"])
                return {"type": "fieldResult", "graphVersion": msg["graphVersion"], "field": fld}
            if self.topology is None:
                raise RuntimeError("solverWorker: step received before topology")
            args: dict[str, Any] = dict(msg["args"])
            args["edges"] = self.topology["edges"]
            args["disjointPairs"] = None
            result = dispatch_descent_step(args)
            return {"type": "result", "graphVersion": msg["graphVersion"], "result": result}
        except Exception as err:  # any throw becomes a protocol error (solverWorker.ts:89-98)
            return {"type": "error", "message": str(err)}
