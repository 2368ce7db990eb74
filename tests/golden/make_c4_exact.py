"""Writes the extended-precision reference of the first Sobolev step of the
E = 8192 benchmark workload (closed Fourier curve V = E = 8192, seed 1,
[barycenter, total length] frozen at the initial curve, field penalty w = 10,
X = (0.5, 1, 0), frozen projection): tests/golden/c4_exact_step1.npz.

Run on a machine with an MI355X (it reads the device-assembled A through the C ABI):
    python tests/golden/make_c4_exact.py [out.npz]

Contents:
- g_exact (V, 3), lambda_exact (k): the exact solution, rounded to FP64, of
  the saddle system [[I3 (x) A, C^T],[C, 0]] [g; lambda] = [dE; 0] for the
  device's A (bit-reproducible), the oracle's C and dE = the device gradient
  + penalty gradient: LAPACK Cholesky through the Schur reduction, refined
  with the oracle's double-double residual until the correction is below
  1e-16 (tests/exact_saddle.py);
- energy_ref, tau_ref, vertices_ref: the FP64-form step with one refinement
  step, whose direction matches g_exact to ~5e-16 (checked here);
- sha256 of the input vertices, edges, A and dE, so a later run can tell
  whether it reproduces the same system.
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (ROOT, os.path.join(ROOT, "repulsive-test2_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

from exact_saddle import ExactSaddle, aos, blocks  # noqa: E402
from oracle import rc_oracle as rco  # noqa: E402
from rcurves import core, curves  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def c4_system(V=8192):
    v, e = curves.fourier_curve(V, seed=1)
    x0 = core.barycenter_target(v, e)
    L0 = core.total_length(v, e)
    cs = [core.barycenter_block(x0), core.total_length_block(L0)]
    pen = core.PenaltyConfig(field_weight=10.0, field_X=(0.5, 1.0, 0.0))
    g = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
    _, pg = core.penalty_energy_and_gradient(v, e, pen)
    dE = g + pg
    A = core.assemble_a(v, e, None, 3.0, 6.0, 1e-10)
    _, Cm, _, _ = rco.eval_constraints(v, e, rco.ConstraintSet().barycenter(x0).total_length(L0))
    return v, e, cs, pen, dE, A, Cm


def main(out):
    V = 8192
    v, e, cs, pen, dE, A, Cm = c4_system(V)
    x, lam, hist = ExactSaddle(A, Cm).solve(blocks(dE))
    print("refinement corrections", hist, flush=True)
    g_exact = aos(x, V)
    r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
        penalties=pen, collect_field=True, projection_mode="frozen", saddle_refine=1))
    d = float(np.linalg.norm(r["descentField"] - g_exact) / np.linalg.norm(g_exact))
    print("FP64 form (1 refinement step) direction vs exact", d, flush=True)
    assert d <= 1e-13, d
    np.savez(out, g_exact=g_exact, lambda_exact=lam, energy_ref=r["energy"],
             tau_ref=r["stats"]["tau"], vertices_ref=r["vertices"],
             sha_vertices=sha(v), sha_edges=sha(np.asarray(e, dtype=np.int32)), sha_A=sha(A),
             sha_dE=sha(dE))
    print("wrote", out, "energy_ref", r["energy"], flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else
         os.path.join(ROOT, "tests", "golden", "c4_exact_step1.npz"))
