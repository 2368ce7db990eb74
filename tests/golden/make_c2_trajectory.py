"""Writes the oracle's 50-step trajectories of the C2 workload (CPU only):
- tests/golden/c2_oracle_trajectory.npz: the oracle's sobolevStepSet as the
  reference computes it (full saddle Bunch-Kaufman LDL^T, unrefined);
- tests/golden/c2_exact_trajectory.npz (argument `exact`): the same step
  with the saddle solution carried to the exact solution of the assembled
  system (tests/exact_saddle.py: Cholesky through the Schur reduction plus
  double-double residual refinement), then the oracle's line search.
Along this trajectory the unrefined solve sits 1e-11 .. 2e-10 from the exact
direction, and the short Armijo steps (tau down to 1/32) amplify such
differences over tens of steps.

C2: closed (3,5) torus knot, V = E = 1024, R = 2, r = 1, uniform parameter,
seeded Gaussian jitter of 0.02 x mean edge length (curves.torus_knot, seed 1);
set [barycenter, total length, 8 pins at seeded vertices (seeded_pins, seed 1)
with their initial positions as targets], all frozen at the start;
penalties totalLength w = 1, lengthDiff w = 10; frozen projection; FP64; each
step's E0 is the previous step's energy (the app's energyBefore reuse).

Stored per step: accepted, tau, projection iterations, energy; the vertices
after steps 10, 20, 30, 40, 50; and the sha256 of the initial curve.
    python tests/golden/make_c2_trajectory.py [exact]
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (ROOT, os.path.join(ROOT, "repulsive-test2_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

sys.path.insert(0, os.path.join(ROOT, "tests"))

from exact_saddle import ExactSaddle, aos, blocks  # noqa: E402
from oracle import rc_oracle as rco  # noqa: E402
from rcurves import curves  # noqa: E402

STEPS = 50


def exact_step(v, e, cs, pen, opts):
    """The oracle's step with the saddle solved exactly (analytic dE +
    penalty gradient, optimizer.ts:283-294; no convergence echo is reached
    on this workload)."""
    dE = rco.gradient(v, e) + rco.penalty_gradient(v, e, pen)
    A = rco.assemble("A", v, e)
    _, Cm, _, _ = rco.eval_constraints(v, e, cs)
    x, _, _ = ExactSaddle(A, Cm).solve(blocks(dE))
    g = aos(x, len(v))
    assert rco.l2_curve_norm(g, v, e) >= 1e-4
    return rco.line_search(v, e, dE, g, cs, opts)


def main(out, exact=False):
    v, e = curves.torus_knot(3, 5, 1024, jitter=0.02, seed=1)
    sha = hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
    pins = curves.seeded_pins(1024, 8, seed=1)
    targets = v[pins].copy()
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    pen = rco.make_penalties(1.0, 10.0, 0.0, (0.0, 0.0, 0.0))
    cs = rco.ConstraintSet().barycenter(x0).total_length(L0)
    for p, t in zip(pins, targets):
        cs.point(int(p), t)
    acc, tau, iters, en, snaps = [], [], [], [], []
    eo = None
    t0 = time.time()
    for step in range(STEPS):
        o = (rco.default_opts(penalties=pen) if eo is None else
             rco.default_opts(penalties=pen, has_energy_before=1, energy_before=eo))
        if exact:
            v, st = exact_step(v, e, cs, pen, o)
        else:
            v, _, st = rco.sobolev_step(v, e, cs, o)
        acc.append(st["accepted"])
        tau.append(st["tau"])
        iters.append(-1 if st["projection_iterations"] is None else st["projection_iterations"])
        en.append(st["energy"])
        eo = st["energy"]
        if (step + 1) % 10 == 0:
            snaps.append(v.copy())
        print(step, st["tau"], st["projection_iterations"], st["energy"],
              f"{time.time() - t0:.0f}s", flush=True)
    np.savez(out, accepted=np.array(acc), tau=np.array(tau), projection_iterations=np.array(iters),
             energy=np.array(en), vertices_every10=np.array(snaps), sha_initial=sha)


if __name__ == "__main__":
    ex = len(sys.argv) > 1 and sys.argv[1] == "exact"
    main(os.path.join(ROOT, "tests", "golden",
                      "c2_exact_trajectory.npz" if ex else "c2_oracle_trajectory.npz"), exact=ex)
