"""Sobolev-step benchmark on one MI355X.

Workload (the step benchmark's E=8192 configuration): closed Fourier curve,
V = E = 8192, coordinates sum_{k=1..6} a_k cos kt + b_k sin kt with
a_k, b_k ~ N(0, 1/k^2) (seed 1), resampled to uniform arc length; constraint
set [barycenter, total length] frozen at the initial curve; field penalty
w = 10, X = (0.5, 1, 0); alpha = 3, beta = 6, eps = 1e-10; frozen projection.

A step is one sobolevStepSet on the device-resident curve (dE + penalty,
Gram assembly, factorisation, solves, Schur system, residual, Armijo line
search with back-projection), continuing the flow from the previous step.
Timed: exactly --steps steps after --warmup untimed steps, each ending in the
library's stream synchronisation (the single host round trip of a step).

Prints one JSON line. value = steps/s of the chosen form (mixed when built);
the plain FP64 form is measured in the same run for the speedup and the
accuracy gates (energy and direction relative error of the mixed form
against FP64 at the same curve).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "repulsive-test2_amd"))

import numpy as np  # noqa: E402

METRIC = "Sobolev steps/s at E=8192: mixed+adjoint vs FP64 GPU, energy rel err <=1e-10"
PEAK_FP64_TFLOPS = 78.6       # MI355X dense FP64 (vector and MFMA)
PEAK_HBM_GBS = 8000.0
WORKLOAD = ("closed Fourier curve V=E=8192 (k=1..6, N(0,1/k^2), seed 1, uniform arc length); "
            "[barycenter, totalLength] frozen; field penalty w=10 X=(0.5,1,0); alpha=3 beta=6")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(V):
    from rcurves import core, curves
    v, e = curves.fourier_curve(V, seed=1)
    x0 = core.barycenter_target(v, e)
    L0 = core.total_length(v, e)
    cs = [core.barycenter_block(x0), core.total_length_block(L0)]
    pen = core.PenaltyConfig(field_weight=10.0, field_X=(0.5, 1.0, 0.0))
    return v, e, cs, pen


def run_form(ctx, v0, cs, pen, precision, steps, warmup):
    from rcurves import _lib as L
    from rcurves import core
    arr, nb, keep = core._blocks_array(cs)
    penc = pen.to_c()
    opts = core.SobolevStepOptions(precision=precision, collect_timings=True).to_c()
    ctx.check(L.lib().rc_set_positions(ctx.handle, L.dptr(np.ascontiguousarray(v0))))
    st = L.StepStats()
    tm = L.Timings()
    e_prev = None
    phase = {k: 0.0 for k, _ in L.Timings._fields_}
    trials = []

    def one(record):
        nonlocal e_prev
        if e_prev is not None:
            opts.has_energy_before = 1
            opts.energy_before = e_prev
        ctx.check(L.lib().rc_sobolev_step_resident(ctx.handle, arr, nb, C.byref(penc),
                                                   C.byref(opts), C.byref(st), C.byref(tm)))
        e_prev = st.energy
        if record:
            for k, _ in L.Timings._fields_:
                phase[k] += getattr(tm, k)
            trials.append(st.trials)

    for i in range(warmup):
        one(False)
        log(f"[{precision}] warmup {i}: accepted={st.accepted} tau={st.tau} E={st.energy:.12g} "
            f"trials={st.trials} refine={st.refine_iters}")
    t0 = time.perf_counter()
    for i in range(steps):
        one(True)
    dt = time.perf_counter() - t0
    log(f"[{precision}] {steps} steps in {dt:.4f}s; last accepted={st.accepted} tau={st.tau} "
        f"E={st.energy:.12g} trials={trials}")
    del keep
    return dt / steps, {k: v / steps for k, v in phase.items()}, trials


def accuracy(v, e, cs, pen):
    """Mixed vs FP64 at the same curve: step energy and g~ (direction)."""
    from rcurves import core
    r64 = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
        penalties=pen, collect_field=True))
    rmx = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
        penalties=pen, collect_field=True, precision="mixed"))
    g64, gmx = r64["descentField"], rmx["descentField"]
    if g64 is None or gmx is None:
        return {"accuracy_error": "singular saddle: " + str(r64["stats"]) + " / " + str(rmx["stats"])}
    return {
        "energy_rel_err": abs(rmx["energy"] - r64["energy"]) / abs(r64["energy"]),
        "direction_rel_err": float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64)),
        "vertices_rel_err": float(np.linalg.norm(rmx["vertices"] - r64["vertices"]) /
                                  np.linalg.norm(r64["vertices"])),
        "tau_fp64": r64["stats"]["tau"], "tau_mixed": rmx["stats"]["tau"],
        "refine_iters": rmx["diagnostics"]["refineIters"],
        "refine_residual": rmx["diagnostics"]["refineResidual"],
    }


def cpu_baseline():
    """The oracle (FP64 C restatement of the reference, 1 thread) on a bounded
    sample, scaled to E = 8192: energy, dE and assembly measured at E = 2048
    and scaled by (8192/2048)^2; the reference's saddle LDL^T factor measured
    at V = 1024 (n = 3V+4 = 3076) and scaled by (24580/3076)^3."""
    from oracle import rc_oracle as rco
    from rcurves import curves
    v, e = curves.fourier_curve(2048, seed=1)
    t = time.perf_counter(); rco.energy(v, e); t_en = time.perf_counter() - t
    t = time.perf_counter(); rco.gradient(v, e); t_de = time.perf_counter() - t
    t = time.perf_counter(); rco.assemble("A", v, e); t_as = time.perf_counter() - t
    v1, e1 = curves.fourier_curve(1024, seed=1)
    A = rco.assemble("A", v1, e1)
    n = 3 * 1024 + 4
    K = np.zeros((n, n))
    for b in range(3):
        K[b * 1024:(b + 1) * 1024, b * 1024:(b + 1) * 1024] = A
    rng = np.random.default_rng(0)
    Cm = rng.standard_normal((4, 3 * 1024))
    K[3072:, :3072] = Cm
    K[:3072, 3072:] = Cm.T
    t = time.perf_counter(); rco.ldlt(K); t_fa = time.perf_counter() - t
    s2 = (8192 / 2048) ** 2
    s3 = (3 * 8192 + 4) ** 3 / n ** 3
    step = (t_en + t_de + t_as) * s2 + t_fa * s3
    return {
        "value": 1.0 / step, "unit": "steps/s", "cores": 1, "kind": "port",
        "sample": (f"oracle C FP64, 1 thread: energy {t_en:.2f}s + dE {t_de:.2f}s + assembleA "
                   f"{t_as:.2f}s at E=2048 (x{s2:.0f}), reference-structure LDL^T {t_fa:.2f}s at "
                   f"n={n} (x{s3:.0f}); extrapolated step {step:.0f}s at E=8192"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--V", type=int, default=8192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the Sobolev step has no multi-GPU path in this build "
              "(see DESIGN.md for how it would shard). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from rcurves import _lib as L
    log("building workload")
    v, e, cs, pen = make_workload(args.V)
    ctx = L.Context(e, len(v))
    t64, ph64, tr64 = run_form(ctx, v, cs, pen, "fp64", args.steps, args.warmup)
    mixed_ok = False
    acc = {}
    if not args.fp64_only:
        try:
            tmx, phmx, trmx = run_form(ctx, v, cs, pen, "mixed", args.steps, args.warmup)
            mixed_ok = True
            acc = accuracy(v, e, cs, pen)
        except L.RcError as err:
            log("mixed form unavailable:", err)
    if mixed_ok:
        t_sel, ph_sel, prec = tmx, phmx, "mixed"
    else:
        t_sel, ph_sel, prec = t64, ph64, "fp64"
    V = args.V
    # dominant phase of the headline form and its roofline
    fac_flops = V ** 3 / 3.0
    fac_ms = ph64["factor"]
    roof = {
        "bound": "mfma", "kernel": "FP64 blocked Cholesky of A + mu 11^T (k_potrf_panel + "
                                   "k_syrk_mfma), fp64 form",
        "achieved": fac_flops / (fac_ms * 1e-3) / 1e12, "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s", "traffic": None,
    }
    roof["frac"] = roof["achieved"] / roof["peak"]
    out = {
        "metric": METRIC, "value": 1.0 / t_sel, "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_sel * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f64 pairs/refinement + f32/f16-mfma factor",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "V": V, "E": V, "form": prec,
                   "fp64_steps_per_s": 1.0 / t64,
                   "speedup_vs_fp64": (t64 / t_sel) if mixed_ok else 1.0,
                   "phase_ms": {k: round(x, 4) for k, x in ph_sel.items()},
                   "phase_ms_fp64": {k: round(x, 4) for k, x in ph64.items()},
                   "trials_per_step": tr64, **acc},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, bounded sample)")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
