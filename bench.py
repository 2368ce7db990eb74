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

Both forms refine the saddle solution once with compensated (double-double)
residuals. Prints one JSON line. value = steps/s of the headline form: the
faster of the forms whose timed steps are all accepted and whose first step
passes the accuracy gates (energy relative error <= 1e-10 and direction
relative error <= 1e-8) against the committed extended-precision reference
of that step (tests/golden/c4_exact_step1.npz); both forms are reported.
roofline: the dominant kernel of the FP64 step, the trailing update
k_syrk128, timed with HIP events around each of its launches on the
library's stream during the timed steps; achieved = algorithmic flops
(128 m (m + 1) per launch, m trailing rows) / summed launch time.
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


def run_form(ctx, v0, cs, pen, precision, steps, warmup, kprof=False, projection="frozen"):
    """Warm up, then time exactly `steps` resident steps from v0. With kprof,
    HIP events bracket every trailing-update launch of the timed steps."""
    from rcurves import _lib as L
    from rcurves import core
    arr, nb, keep = core._blocks_array(cs)
    penc = pen.to_c()
    opts = core.SobolevStepOptions(precision=precision, collect_timings=True,
                                   projection_mode=projection).to_c()
    ctx.check(L.lib().rc_set_positions(ctx.handle, L.dptr(np.ascontiguousarray(v0))))
    st = L.StepStats()
    tm = L.Timings()
    e_prev = None
    phase = {k: 0.0 for k, _ in L.Timings._fields_}
    trials, accepted = [], []

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
                phase[k] += getattr(tm, k).ms
            trials.append(st.trials)
            accepted.append(st.accepted)

    for i in range(warmup):
        one(False)
        log(f"[{precision}{'' if projection == 'frozen' else '/' + projection}] warmup {i}: accepted={st.accepted} tau={st.tau} E={st.energy:.12g} "
            f"trials={st.trials} refine={st.refine_iters}")
    prof = None
    if kprof:
        ctx.check(L.lib().rc_kernel_profile_enable(ctx.handle, 1))
    t0 = time.perf_counter()
    for i in range(steps):
        one(True)
    dt = time.perf_counter() - t0
    if kprof:
        kp = L.KernelProfile()
        ctx.check(L.lib().rc_kernel_profile_read(ctx.handle, C.byref(kp)))
        ctx.check(L.lib().rc_kernel_profile_enable(ctx.handle, 0))
        prof = {"ms": kp.syrk_ms, "flop": kp.syrk_flop, "launches": kp.syrk_launches}
    log(f"[{precision}{'' if projection == 'frozen' else '/' + projection}] {steps} steps in {dt:.4f}s; last accepted={st.accepted} tau={st.tau} "
        f"E={st.energy:.12g} trials={trials}")
    del keep
    return {"t": dt / steps, "phase": {k: v / steps for k, v in phase.items()}, "trials": trials,
            "accepted": accepted, "kprof": prof}


REFERENCE = os.path.join(ROOT, "tests", "golden", "c4_exact_step1.npz")


def accuracy(v, e, cs, pen):
    """First step of the workload in each form against the committed
    extended-precision reference (exact solution of the device-assembled
    saddle system, tests/golden/make_c4_exact.py): direction (g~), step
    energy, vertices and the post-step constraint violation |Phi|; the
    unrefined FP64 solve for comparison."""
    import hashlib
    from rcurves import core
    try:
        fx = np.load(REFERENCE)
    except OSError:
        return {"accuracy_error": "reference fixture missing"}
    if str(fx["sha_vertices"]) != hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest():
        return {"accuracy_error": "workload differs from the committed reference (V != 8192?)"}
    g_ref, e_ref, v_ref = fx["g_exact"], float(fx["energy_ref"]), fx["vertices_ref"]
    out = {"reference": "exact saddle solution of the assembled system (double-double refined), "
                        "tests/golden/c4_exact_step1.npz"}
    for name, prec, nref in (("fp64", "fp64", 1), ("mixed", "mixed", 1), ("fp64_unrefined", "fp64", 0)):
        r = core.sobolev_step_set(v, e, None, cs, core.SobolevStepOptions(
            penalties=pen, collect_field=True, precision=prec, projection_mode="frozen",
            saddle_refine=nref, report_ir_residual=True))
        g = r["descentField"]
        if g is None:
            out[name] = {"error": "singular saddle: " + str(r["stats"])}
            continue
        out[name] = {
            "energy_rel_err": abs(r["energy"] - e_ref) / abs(e_ref),
            "direction_rel_err": float(np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)),
            "vertices_rel_err": float(np.linalg.norm(r["vertices"] - v_ref) / np.linalg.norm(v_ref)),
            "constraint_violation": r["diagnostics"]["phiNorm"],
            "tau": r["stats"]["tau"],
            "saddle_residual_dd": r["diagnostics"]["saddleResidualDD"],
            "pcg_iters": r["diagnostics"]["refineIters"],
        }
        out[name]["passes_gates"] = bool(r["accepted"] and out[name]["energy_rel_err"] <= 1e-10
                                         and out[name]["direction_rel_err"] <= 1e-8)
    return out


def cpu_baseline():
    """The oracle (FP64 C restatement of the reference, 1 thread) on a bounded
    sample, scaled to E = 8192: energy, dE and assembly measured at E = 4096
    and scaled by (8192/4096)^2; the reference's saddle LDL^T (the structure
    of its linsolve, n = 3V + k) measured at n = 3076 and n = 6148 and
    extrapolated to n = 24580 with the exponent the two sizes give."""
    from oracle import rc_oracle as rco
    from rcurves import curves
    rco.set_threads(1)
    v, e = curves.fourier_curve(4096, seed=1)
    t = time.perf_counter(); rco.energy(v, e); t_en = time.perf_counter() - t
    t = time.perf_counter(); rco.gradient(v, e); t_de = time.perf_counter() - t
    t = time.perf_counter(); rco.assemble("A", v, e); t_as = time.perf_counter() - t
    fac = []
    for Vs in (1024, 2048):
        v1, e1 = curves.fourier_curve(Vs, seed=1)
        A = rco.assemble("A", v1, e1)
        n = 3 * Vs + 4
        K = np.zeros((n, n))
        for b in range(3):
            K[b * Vs:(b + 1) * Vs, b * Vs:(b + 1) * Vs] = A
        rng = np.random.default_rng(0)
        Cm = rng.standard_normal((4, 3 * Vs))
        K[3 * Vs:, :3 * Vs] = Cm
        K[:3 * Vs, 3 * Vs:] = Cm.T
        t = time.perf_counter(); rco.ldlt(K); fac.append((n, time.perf_counter() - t))
    (n1, t1), (n2, t2) = fac
    p = float(np.log(t2 / t1) / np.log(n2 / n1))
    nfull = 3 * 8192 + 4
    t_fa = t2 * (nfull / n2) ** p
    s2 = (8192 / 4096) ** 2
    step = (t_en + t_de + t_as) * s2 + t_fa
    try:
        host = len(os.sched_getaffinity(0))
    except AttributeError:
        host = os.cpu_count()
    return {
        "value": 1.0 / step, "unit": "steps/s", "cores": 1, "kind": "port",
        "host_cores_available": host,
        "sample": (f"oracle C FP64, 1 thread of {host}: energy {t_en:.2f}s + dE {t_de:.2f}s + "
                   f"assembleA {t_as:.2f}s at E=4096 (x{s2:.0f}); reference-structure LDL^T "
                   f"{t1:.2f}s at n={n1} and {t2:.2f}s at n={n2} (exponent {p:.2f}), "
                   f"extrapolated {t_fa:.0f}s at n={nfull}; step {step:.0f}s at E=8192"),
    }


# Algorithmic FP64 work per ordered disjoint edge pair (sqrt, divide and
# reciprocal counted as one flop): energy ~150, adjoint gradient ~480, Gram
# assembly ~250 (the executed instruction counts of the kernels, Newton
# refinements included, are higher: ~188 for the energy and ~342 for the
# adjoint).
FLOP_PER_PAIR = {"energy": 150.0, "dE": 480.0, "assembleA": 250.0}


def kernel_rates(v, e, phase_fp64, ctx):
    """Per-kernel milliseconds (HIP events of the timed FP64 steps for dE and
    assembleA, which run concurrently on two streams; host-timed rc_energy
    calls for a full energy evaluation) and the rates they imply: algorithmic
    TFLOP/s and, for the assembly, the GB/s of its V x V FP64 write."""
    from rcurves import _lib as L
    from rcurves import core
    E, V = len(e), len(v)
    P = E * (E - 3)   # ordered disjoint pairs of a closed curve
    p = core._params(3.0, 6.0, 1e-10)
    out = C.c_double()
    vv = np.ascontiguousarray(v)
    ctx.check(L.lib().rc_energy(ctx.handle, L.dptr(vv), C.byref(p), C.byref(out)))
    t = time.perf_counter()
    for _ in range(5):
        ctx.check(L.lib().rc_energy(ctx.handle, L.dptr(vv), C.byref(p), C.byref(out)))
    en_ms = (time.perf_counter() - t) / 5 * 1e3
    res = {"ordered_pairs": P}
    for k, ms in (("energy", en_ms), ("dE", phase_fp64["dE"]), ("assembleA", phase_fp64["assembleA"])):
        res[k] = {"ms": round(ms, 4),
                  "tflops_algorithmic": round(P * FLOP_PER_PAIR[k] / (ms * 1e-3) / 1e12, 2)}
    res["energy"]["timing"] = "host, 5 rc_energy calls incl. 196 KB position upload"
    res["assembleA"]["write_gbs"] = round(8.0 * V * V / (res["assembleA"]["ms"] * 1e-3) / 1e9, 1)
    # in a step the assembly runs on the second stream beside the energy and
    # the adjoint gradient: both phase times are wall times of kernels that
    # share the GPU, so the rates are lower bounds of the kernels' own
    for k in ("dE", "assembleA"):
        res[k]["timing"] = "step phase (HIP events), concurrent with the other of dE / assembleA"
    res["dE_plus_assembleA_overlapped_ms"] = round(
        phase_fp64["step"] - phase_fp64["saddle"] - phase_fp64["lineSearch"] - phase_fp64["energy"], 4)
    return res


def measured_traffic():
    """HBM bytes per launch of k_syrk128 from the committed counter summary
    (FETCH_SIZE / WRITE_SIZE passes scaled by a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", "k_syrk128_fourier_v8192_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch k_syrk128:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def copy_calibration(mib):
    """One device-to-device hipMemcpy of `mib` MiB (a plain copy kernel of
    known traffic, for scaling the FETCH_SIZE / WRITE_SIZE counters)."""
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--V", type=int, default=8192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--mixed-only", action="store_true",
                    help="time only the mixed form (profiling; the line reports it as measured)")
    ap.add_argument("--no-kernel-profile", action="store_true",
                    help="time the FP64 steps without the per-launch events of the roofline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="also run one device-to-device copy of this size (counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the Sobolev step has no multi-GPU path. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from rcurves import _lib as L
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    log("building workload")
    v, e, cs, pen = make_workload(args.V)
    ctx = L.Context(e, len(v))
    if args.mixed_only:
        rmx = run_form(ctx, v, cs, pen, "mixed", args.steps, args.warmup)
        print(json.dumps({"metric": METRIC, "form": "mixed", "value": 1.0 / rmx["t"],
                          "unit": "steps/s", "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": rmx["t"] * 1e3,
                          "phase_ms": {k: round(x, 4) for k, x in rmx["phase"].items()}}),
              flush=True)
        ctx.close()
        return
    r64 = run_form(ctx, v, cs, pen, "fp64", args.steps, args.warmup,
                   kprof=not args.no_kernel_profile)
    rmx = None
    acc = accuracy(v, e, cs, pen) if not args.fp64_only else {}
    if not args.fp64_only:
        try:
            rmx = run_form(ctx, v, cs, pen, "mixed", args.steps, args.warmup)
        except L.RcError as err:
            log("mixed form unavailable:", err)
    # the reference function's default projection: the saddle refactored at
    # every projection iterate (a few steps; not the headline form)
    rre = None
    if not args.fp64_only:
        rre = run_form(ctx, v, cs, pen, "fp64", min(args.steps, 3), 1, projection="reassemble")
    # headline form: the faster of the two that pass the accuracy gates
    # against the reference (the FP64 form if neither does)
    fp64_ok = bool(acc.get("fp64", {}).get("passes_gates")) and all(r64["accepted"])
    mixed_ok = (rmx is not None and all(rmx["accepted"])
                and bool(acc.get("mixed", {}).get("passes_gates")))
    sel, prec = ((rmx, "mixed") if mixed_ok and (not fp64_ok or rmx["t"] < r64["t"])
                 else (r64, "fp64"))
    V = args.V
    kp = r64["kprof"]
    achieved = kp["flop"] / (kp["ms"] * 1e-3) / 1e12 if kp and kp["ms"] > 0 else None
    roof = {
        "bound": "mfma", "kernel": "k_syrk128 (FP64 trailing update of the Cholesky factor, "
                                   "v_mfma_f64_16x16x4_f64), fp64 form",
        "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": (achieved / PEAK_FP64_TFLOPS) if achieved else None,
        "traffic": measured_traffic() if V == 8192 else None,
        "launches": kp["launches"] if kp else 0,
        "avg_launch_us": (kp["ms"] * 1e3 / kp["launches"]) if kp and kp["launches"] else None,
    }
    out = {
        "metric": METRIC, "value": 1.0 / sel["t"], "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sel["t"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f64 pairs/refinement + f32/f16-mfma factor",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "V": V, "E": V, "form": prec,
                   "fp64_steps_per_s": 1.0 / r64["t"],
                   "mixed_steps_per_s": (1.0 / rmx["t"]) if rmx else None,
                   "mixed_accepted": rmx["accepted"] if rmx else None,
                   "fp64_passes_gates": fp64_ok,
                   "mixed_passes_gates": mixed_ok,
                   "speedup_mixed_vs_fp64": (r64["t"] / rmx["t"]) if rmx else None,
                   "reassemble_steps_per_s": (1.0 / rre["t"]) if rre else None,
                   "reassemble_phase_ms": ({k: round(x, 4) for k, x in rre["phase"].items()}
                                           if rre else None),
                   "phase_ms": {k: round(x, 4) for k, x in sel["phase"].items()},
                   "phase_ms_fp64": {k: round(x, 4) for k, x in r64["phase"].items()},
                   "trials_per_step": r64["trials"], "accuracy": acc},
        "roofline": roof,
    }
    out["config"]["kernels"] = kernel_rates(v, e, r64["phase"], ctx)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, bounded sample)")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
