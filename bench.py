"""Benchmark of the ADDA FFT-matvec + Krylov hot path on one MI355X.

Workload (the grid-256 configuration): homogeneous sphere on a 128^3 box (1,099,136 dipoles, padded FFT grid 256^3,
3 complex components), refractive index 1.5+0.01i, size parameter ~40 (-size 80 with lambda = 2 pi), LDR
polarizability, plane wave along z, Y polarization. Synthetic by construction (the particle is generated, nothing is
read).

One step = one complete time-to-solution run of the FP32-mixed solver (FP32 inner BiCGStab, FP64 defect correction
to ||r||/||b|| <= 1e-10, then Cext), inputs (incident field, D-matrix) already resident in HBM. The plain FP64
BiCGStab solve is timed beside it, as are QMR_CS in both forms, the matvec throughput of both precisions, and the
per-pass kernel times of the matvec (HIP events on the operator's stream) that give the roofline figures. The matvec
radix schedules and tiles are picked by the autotuner first (its time is reported, not counted).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--box 128]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "a-dda_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "time-to-solution (s) & matvecs/s, grid 256, FP32-mixed vs FP64, Cext rel err ≤1e-6"
HBM_PEAK_GBS = 8000.0
PASS_NAMES = ["P1 scatter+FFT-x", "P2 FFT-z", "P3 FFT-y.D.FFT-y^-1", "P4 inverse FFT-z", "P5 inverse FFT-x+gather"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--box", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-qmr", action="store_true")
    ap.add_argument("--matvec-only", action="store_true",
                    help="profiling mode: autotune, 20 matvecs per precision and one 52.8 MB device copy (the "
                         "calibration kernel for HBM counters), no solves; prints the matvec figures only")
    ap.add_argument("--solve-only", choices=["fp64", "mixed"],
                    help="profiling mode: autotune, then one solve (--method, to --eps) in the given form; prints its "
                         "time")
    ap.add_argument("--batched-solve", choices=["fp64", "mixed"],
                    help="only one batched solve of the 64-RHS workload (after an untimed one), in this form: the run "
                         "a kernel trace of a batched solve is taken from")
    ap.add_argument("--methods", action="store_true",
                    help="per-iteration cost of every Krylov method on this workload in FP64: 200 iterations each, "
                         "device-resident recurrences against the host form (one host sync per inner product)")
    ap.add_argument("--method", default="bicgstab", help="solver of --solve-only / --batched-solve (ADDA's -iter names)")
    ap.add_argument("--eps", type=float, default=1e-10, help="tolerance of --solve-only (default 1e-10)")
    ap.add_argument("--inner-method", default=None,
                    help="Krylov method of the inner solves of --solve-only mixed (default: automatic, BiCGStab for QMR_CS)")
    ap.add_argument("--inner-eps", type=float, default=0.0,
                    help="relative tolerance of the FP32 inner solves of the mixed forms (default: the solver's, 1e-4)")
    ap.add_argument("--m", type=float, nargs=2, default=[1.5, 0.01], metavar=("RE", "IM"),
                    help="refractive index of the sphere (default 1.5 0.01; the 640^3 workload uses 1.33 0.001)")
    ap.add_argument("--fp64-steps", type=int, default=3,
                    help="timed plain FP64 BiCGStab solves (after one untimed), averaged for tts_fp64_s")
    ap.add_argument("--qmr-max-inner", type=int, default=2000,
                    help="inner-iteration cap of the mixed QMR_CS form forced to FP32 QMR_CS inner solves (they stall "
                         "near a 1e-3 true residual while the quasi-residual keeps falling, so an uncapped step runs on "
                         "without gain)")
    ap.add_argument("--qmr-max-inner-v64", type=int, default=3000,
                    help="the same cap for QMR_CS with FP64 vectors around FP32 transforms")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched-RHS and adjoint-gradient workloads")
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "a-dda_amd", "tune_cache_mi355x.jsonl"),
                    help="autotuner cache (JSON lines): configurations found there are reused without timing, new "
                         "ones are tuned and appended (default: the cache the autotuner wrote on an MI355X for these "
                         "workloads; pass an empty string to always tune)")
    args = ap.parse_args()
    args.tune_cache = args.tune_cache or None
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import adda_amd as A
    if A.device_count() < 1:
        raise RuntimeError("no HIP device visible")
    if args.batched_solve:
        batched_solve_only(A, args)
        return

    lam = 2 * np.pi
    pos, mat, box = A.make_particle("sphere", args.box)
    nd = len(pos)
    size = 80.0 * args.box / 128
    dpl = lam * (nd / (np.pi / 6)) ** (1 / 3) / size  # -size with volume correction, src/make_particle.c:2192-2196
    m = args.m[0] + 1j * args.m[1]
    t0 = time.perf_counter()
    pr = A.Problem(pos, box, m, dpl=dpl, fp32=True)
    t_dmat = time.perf_counter() - t0
    log(f"particle: {nd} dipoles, box {box}, grid {pr.grid}, dpl {dpl:.4f}, x = {2*np.pi/lam*(3*nd/(4*np.pi))**(1/3)*lam/dpl:.3f}; "
        f"operator + FP64 D-matrix build {t_dmat:.3f} s")
    prop, pol = np.array([0.0, 0, 1]), np.array([0.0, 1, 0])
    pr.set_polarization(prop, pol)
    op = pr.op
    t0 = time.perf_counter()
    tune = {name: op.autotune(prec, reps=3, cache=args.tune_cache) for prec, name in ((A.F64, "fp64"), (A.F32, "fp32"))}
    t_tune = time.perf_counter() - t0
    log(f"autotune {t_tune:.2f} s: " + "; ".join(f"{k}: {v}" for k, v in tune.items()))

    if args.methods:
        for meth in ("bicgstab", "qmr", "cgnr", "bicg", "csym", "qmr2", "bcgs2"):
            row = {"mode": "methods", "method": meth, "dipoles": nd, "grid": list(pr.grid), "iterations": 200}
            for hs, key in ((False, "device_us_per_iter"), (True, "host_us_per_iter")):
                op.sync()
                t = time.perf_counter()
                res = pr.solve(meth, eps=1e-14, maxiter=200, host_scalars=hs, hist_len=4)
                op.sync()
                row[key] = round((time.perf_counter() - t) / max(res["niter"], 1) * 1e6, 1)
                row["matvecs"] = res["nmatvec"]
            print(json.dumps(row), flush=True)
        pr.close()
        return

    if args.solve_only:
        prec = A.PREC_FP64 if args.solve_only == "fp64" else A.PREC_MIXED
        op.sync()
        t = time.perf_counter()
        # verbose bit 1: a progress line about every 10 s, so that a long solve (the 640^3 sphere) shows it is alive
        res = pr.solve(args.method, eps=args.eps, precision=prec, verbose=2, recalc_resid=True, inner_eps=args.inner_eps,
                       inner_method=args.inner_method)
        cext = pr.cext()
        dt = time.perf_counter() - t
        print(json.dumps({"mode": "solve-only", "form": args.solve_only, "method": args.method, "eps": args.eps,
                          "m": args.m, "dipoles": nd, "grid": list(pr.grid), "s": round(dt, 4), "rc": res["rc"],
                          "iterations": res["niter"], "matvecs": res["nmatvec"], "fp64_matvecs": res["nmatvec_fp64"],
                          "outer": res["nouter"], "true_resid": res["true_resid"], "cext": cext, "config": tune}),
              flush=True)
        pr.close()
        return

    # ---- matvec throughput and per-pass timing (HIP events on the operator's stream) --------------------------------
    mv = {}
    for prec, name in ((A.F64, "fp64"), (A.F32, "fp32")):
        a, r = op.alloc(prec), op.alloc(prec)
        A.check(A.hip().adda_hip_vec_convert(op.h, a, prec, pr.einc, A.F64))
        op.profile_matvec(a, r, prec, nrep=3)
        passes, tot = op.profile_matvec(a, r, prec, nrep=20)
        op.free(a)
        op.free(r)
        bytes_pass = [op.matvec_bytes(prec, i) for i in range(5)]
        mv[name] = dict(ms=tot, passes=passes, bytes=bytes_pass, total_bytes=op.matvec_bytes(prec, 5))
        log(f"matvec {name}: {tot*1e3:.1f} us ({1e3/tot:.1f} matvecs/s), "
            f"{mv[name]['total_bytes']/tot/1e6:.0f} GB/s algorithmic; passes (us): "
            + ", ".join(f"{p*1e3:.1f}" for p in passes))

    def run(method, precision, **kw):
        op.sync()
        t = time.perf_counter()
        res = pr.solve(method, eps=1e-10, precision=precision, **kw)
        cext = pr.cext()
        dt = time.perf_counter() - t
        return dt, cext, res

    if args.matvec_only:
        a, b = op.alloc(A.F64), op.alloc(A.F64)
        A.check(A.hip().adda_hip_vec_convert(op.h, b, A.F64, a, A.F64))  # plain copy of 3*nd*16 bytes
        op.sync()
        print(json.dumps({"mode": "matvec-only", "copy_bytes": 3 * nd * 16, "hbm": A.hbm_rates(), "matvec": {
            k: {"us": round(v["ms"] * 1e3, 2), "pass_us": [round(x * 1e3, 2) for x in v["passes"]],
                "bytes": v["total_bytes"], "pass_bytes": v["bytes"]} for k, v in mv.items()}, "config": tune}),
              flush=True)
        pr.close()
        return

    # ---- the drop-in boundary on host vectors (adda_hip_matvec_host: ADDA's MatVec signature, PCIe both ways) ------
    xh = op.download(pr.einc)
    op.matvec_host(xh)
    op.sync()
    t = time.perf_counter()
    nh = 10
    for _ in range(nh):
        op.matvec_host(xh)
    t_host = (time.perf_counter() - t) / nh
    log(f"matvec on host vectors (PCIe in and out): {t_host * 1e3:.2f} ms, device-resident FP64 {mv['fp64']['ms']:.3f} ms")

    # ---- plain FP64 BiCGStab (reference solve): one untimed, then the mean of --fp64-steps -------------------------
    run("bicgstab", A.PREC_FP64)
    t64s = []
    for _ in range(max(1, args.fp64_steps)):
        t64, cext64, r64 = run("bicgstab", A.PREC_FP64)
        t64s.append(t64)
    t64 = float(np.mean(t64s))
    log(f"FP64 BiCGStab: {t64:.3f} s (mean of {len(t64s)}: " + ", ".join(f"{t:.3f}" for t in t64s) +
        f"), {r64['niter']} iterations, {r64['nmatvec']} matvecs, Cext {cext64:.10g}")

    # ---- timed steps: FP32-mixed BiCGStab time-to-solution ---------------------------------------------------------
    for _ in range(args.warmup):
        run("bicgstab", A.PREC_MIXED)
    op.sync()
    tstart = time.perf_counter()
    mixed = []
    for _ in range(args.steps):
        mixed.append(run("bicgstab", A.PREC_MIXED))
    op.sync()
    elapsed = time.perf_counter() - tstart
    tmix = elapsed / args.steps
    _, cextmix, rmix = mixed[-1]
    err = abs(cextmix - cext64) / abs(cext64)
    log(f"mixed BiCGStab: {tmix:.3f} s/solve, {rmix['niter']} FP32 iterations, {rmix['nmatvec']} FP32 + "
        f"{rmix['nmatvec_fp64']} FP64 matvecs, {rmix['nouter']} outer steps, true residual {rmix['true_resid']:.2e}, "
        f"Cext {cextmix:.10g} (rel err vs FP64 {err:.2e}), speedup {t64/tmix:.2f}x")

    extra = {}
    if not args.no_qmr:
        tq64, cq64, rq64 = run("qmr", A.PREC_FP64)
        # the mixed form with its automatic inner method: FP32 BiCGStab inner solves (FP32 QMR_CS loses its symmetric
        # Lanczos basis on this operator, below), FP64 defect correction; a rejected correction finishes in FP64 QMR_CS
        tqm, cqm, rqm = run("qmr", A.PREC_MIXED)
        extra["qmr"] = dict(tts_fp64_s=round(tq64, 4), tts_mixed_s=round(tqm, 4), speedup=round(tq64 / tqm, 3),
                            inner_method="bicgstab (automatic for QMR_CS)",
                            converged_fp64=rq64["rc"] == 0, converged_mixed=rqm["rc"] == 0,
                            fp64_matvecs_mixed=rqm["nmatvec_fp64"],
                            iterations_fp64=rq64["niter"], iterations_mixed=rqm["niter"], outer_mixed=rqm["nouter"],
                            true_resid_mixed=rqm["true_resid"],
                            cext_rel_err=abs(cqm - cq64) / abs(cq64), cext_vs_bicgstab_fp64=abs(cq64 - cext64) / cext64)
        log(f"QMR_CS: FP64 {tq64:.3f} s ({rq64['niter']} it), mixed {tqm:.3f} s ({rqm['niter']} it, "
            f"{rqm['nouter']} outer, true residual {rqm['true_resid']:.2e}), Cext rel err {extra['qmr']['cext_rel_err']:.2e}")
        # the same with the inner solves forced to FP32 QMR_CS, capped: the measured limit of that form
        tqq, cqq, rqq = run("qmr", A.PREC_MIXED, max_inner=args.qmr_max_inner, inner_method="qmr")
        extra["qmr"]["inner_qmr_cs"] = dict(
            tts_s=round(tqq, 4), speedup=round(tq64 / tqq, 3), converged=rqq["rc"] == 0, iterations=rqq["niter"],
            outer=rqq["nouter"], fp64_matvecs=rqq["nmatvec_fp64"], true_resid=rqq["true_resid"],
            max_inner=args.qmr_max_inner, cext_rel_err=abs(cqq - cq64) / abs(cq64))
        log(f"QMR_CS mixed with FP32 QMR_CS inner solves (cap {args.qmr_max_inner}): {tqq:.3f} s ({rqq['niter']} it, "
            f"{rqq['nouter']} outer), speedup {tq64 / tqq:.2f}x")
        # FP64 Krylov vectors and recurrences around FP32 transforms (the matvec's FFTs and D-matrix product in FP32),
        # QMR_CS inner solves
        tqv, cqv, rqv = run("qmr", A.PREC_MIXED_V64, max_inner=args.qmr_max_inner_v64, inner_method="qmr")
        extra["qmr"]["fp64_vectors_fp32_transforms"] = dict(
            tts_s=round(tqv, 4), speedup=round(tq64 / tqv, 3), converged=rqv["rc"] == 0, iterations=rqv["niter"],
            outer=rqv["nouter"], true_resid=rqv["true_resid"], max_inner=args.qmr_max_inner_v64,
            cext_rel_err=abs(cqv - cq64) / abs(cq64))
        log(f"QMR_CS with FP64 vectors around FP32 transforms: {tqv:.3f} s ({rqv['niter']} it, {rqv['nouter']} outer, "
            f"true residual {rqv['true_resid']:.2e}), speedup {tq64 / tqv:.2f}x")
        tbv, cbv, rbv = run("bicgstab", A.PREC_MIXED_V64)
        extra["bicgstab_fp64_vectors_fp32_transforms"] = dict(
            tts_s=round(tbv, 4), speedup=round(t64 / tbv, 3), converged=rbv["rc"] == 0, iterations=rbv["niter"],
            outer=rbv["nouter"], true_resid=rbv["true_resid"], cext_rel_err=abs(cbv - cext64) / abs(cext64))
        log(f"BiCGStab with FP64 vectors around FP32 transforms: {tbv:.3f} s ({rbv['niter']} it), speedup "
            f"{t64 / tbv:.2f}x")

    extra["matvec_host_pcie"] = {
        "what": "adda_hip_matvec_host (ADDA's MatVec on host vectors: 52.8 MB copied in and out over PCIe per call), "
                "FP64, wall clock, mean of 10 calls", "ms": round(t_host * 1e3, 3),
        "device_resident_ms": round(mv["fp64"]["ms"], 4), "ratio": round(t_host * 1e3 / mv["fp64"]["ms"], 2)}
    if not args.no_extra:
        extra["batched_rhs"] = bench_batched(A, args)
        extra["adjoint_gradient"] = bench_adjoint(A)

    # ---- roofline: dominant kernel of the timed workload (FP32 matvec pass P3) --------------------------------------
    dom = mv["fp32"]
    k = int(np.argmax(dom["passes"]))
    achieved = dom["bytes"][k] / (dom["passes"][k] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": PASS_NAMES[k] + " (fp32)", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(k, tune["fp32"]), "algorithmic_bytes": dom["bytes"][k]}
    matvec_roof = {name: {"GBps": round(v["total_bytes"] / (v["ms"] * 1e-3) / 1e9, 1),
                          "frac": round(v["total_bytes"] / (v["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "bytes_per_matvec": v["total_bytes"], "us_per_matvec": round(v["ms"] * 1e3, 2),
                          "pass_us": [round(p * 1e3, 2) for p in v["passes"]]} for name, v in mv.items()}

    # ---- what a plain device copy reaches on this GPU (context for the fractions of the 8 TB/s peak above) ----------
    extra["hbm_copy"] = A.hbm_rates()

    # ---- CPU baseline: the oracle's FP64 FFT matvec on one host core, bounded sample of this workload ---------------
    cpu = None
    if not args.no_cpu_baseline:
        from cases import Case
        oc = Case(pos, box, m, dpl=dpl)
        dmat = op.dmatrix()  # the oracle's matvec needs only D; building it on the CPU would add minutes
        x = op.download(pr.einc)
        t = time.perf_counter()
        oc.matvec_fft(dmat, pr.cc_sqrt, x)
        tc = time.perf_counter() - t
        cpu = {"value": round(1.0 / tc, 5), "unit": "matvecs/s", "cores": 1, "kind": "port",
               "what": "the oracle's FP64 FFT matvec: a textbook recursive mixed-radix FFT over the full padded grid (no zero-plane "
                       "skipping), not FFTW -- ADDA's FFTW build cannot be built without FFTW3",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "host_cpus_usable": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
               "sample": f"one FP64 FFT matvec of this workload (grid {pr.grid[0]}^3) by the oracle (1 thread, "
                         f"{tc:.1f} s); the FP64 BiCGStab solve ({r64['nmatvec']} matvecs) extrapolates to "
                         f"{tc * r64['nmatvec']:.0f} s"}
        log(f"CPU baseline: {tc:.2f} s per matvec")
    pr.close()

    out = {
        "metric": METRIC, "value": round(tmix, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(tmix * 1e3, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 inner (fp32 complex) + f64 outer", "data": "synthetic",
        "config": {"workload": f"sphere {box[0]}^3 box ({nd} dipoles), FFT grid {pr.grid[0]}x{pr.grid[1]}x{pr.grid[2]}, "
                               f"m=1.5+0.01i, size parameter ~40, LDR, BiCGStab to 1e-10 (FP32-mixed); plain FP64 timed "
                               f"beside; QMR_CS in both forms under extra",
                   "dipoles": nd, "grid": list(pr.grid), "eps": 1e-10, "solver": "bicgstab"},
        "tts_fp64_s": round(t64, 4), "tts_fp64_runs_s": [round(t, 4) for t in t64s], "tts_mixed_s": round(tmix, 4),
        "tts_mixed_runs_s": [round(m[0], 4) for m in mixed], "speedup_mixed_vs_fp64": round(t64 / tmix, 3),
        "cext_fp64": cext64, "cext_mixed": cextmix, "cext_rel_err": err,
        "iterations": {"fp64": r64["niter"], "mixed_fp32": rmix["niter"], "mixed_outer": rmix["nouter"]},
        "matvecs_per_s": {"fp64": round(1e3 / mv["fp64"]["ms"], 1), "fp32": round(1e3 / mv["fp32"]["ms"], 1)},
        "matvec_roofline": matvec_roof, "dmatrix_build_s": round(t_dmat, 3), "autotune_s": round(t_tune, 3),
        "matvec_config": tune,
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(out), flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measured_traffic(pass_index, cfg):
    """HBM bytes per launch of the FP32 pass kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (calibrated against a copy of known size), when it was taken with the same tuned FP32 configuration."""
    try:
        with open(os.path.join(ROOT, "profiles", "matvec_grid256_kernel_trace.txt")) as f:
            prof_cfg = [l for l in f if l.strip().startswith("fp32 x=")][0].split("fp32", 1)[1].split("(")[0].strip()
        if prof_cfg != cfg.split("(")[0].strip():
            return None
        with open(os.path.join(ROOT, "profiles", "matvec_grid256_hbm_traffic.txt")) as f:
            for line in f:
                if line.startswith(f"P{pass_index + 1} ") and "<float" in line:
                    return float(line.split()[-3])
    except (OSError, ValueError, IndexError):
        pass
    return None


def batched_workload(A):
    """The batched multi-RHS workload: ellipsoid box 64x52x32 (semi-axes 1:0.8:0.5), 32 seeded orientations x 2
    polarizations; returns (pos, box, props, pols, pars, pers)."""
    from adda_rot import rotate_zyz
    from cases import orientations, PROP_Z, POL_X, POL_Y
    pos, _, box = A.make_particle("ellipsoid", 64, 0.8, 0.5)
    props, pols, pars, pers = [], [], [], []
    for a, b, g in orientations(32):
        prop, pY, pX = (rotate_zyz(a, b, g, v) for v in (PROP_Z, POL_Y, POL_X))
        props += [prop, prop]
        pols += [pY, pX]
        pars += [pY, pY]
        pers += [pX, pX]
    return pos, box, props, pols, pars, pers


def batched_solve_only(A, args):
    """One batched solve of the 64-RHS workload in one form (--batched-solve), after an untimed one."""
    pos, box, props, pols, _, _ = batched_workload(A)
    K = len(props)
    bp = A.BatchProblem(pos, box, 1.33 + 0.001j, np.array(props), np.array(pols), fp32=True, grid=A.fast_grid(box))
    try:
        for prec in (A.F64, A.F32):
            bp.op.autotune_batch(K, prec, reps=3, cache=args.tune_cache)
        kw = {"precision": A.PREC_MIXED} if args.batched_solve == "mixed" else {}
        bp.solve(args.method, eps=1e-10, **kw)
        bp.op.sync()
        t = time.perf_counter()
        st = bp.solve(args.method, eps=1e-10, **kw)
        bp.op.sync()
        dt = time.perf_counter() - t
        print(json.dumps({"mode": "batched-solve", "form": args.batched_solve, "method": args.method, "rhs": K,
                          "s": round(dt, 4), "iterations_max": max(s["niter"] for s in st),
                          "matvecs": sum(s["nmatvec"] for s in st), "fp64_matvecs": sum(s["nmatvec_fp64"] for s in st),
                          "outer_max": max(s["nouter"] for s in st),
                          "max_true_resid": max(s["true_resid"] for s in st)}), flush=True)
    finally:
        bp.close()


def bench_batched(A, args):
    """Batched multi-RHS workload: ellipsoid on a 64^3 grid (semi-axes 1:0.8:0.5, box 64x52x32), m = 1.33+0.001i,
    default dpl 10|m|, 32 seeded orientations x 2 polarizations = 64 right-hand sides solved together to 1e-10 (FP64
    and FP32-mixed, HIP-graph replay), against the same 64 solved one after another (ADDA's orientation loop). The FFT
    grid is adda_hip_fastgrid's (128 along y, where fftFit gives 108), for the batched and the sequential solves alike."""
    # pars, pers: the scattering plane of each orientation's beam frame, for both polarizations
    pos, box, props, pols, pars, pers = batched_workload(A)
    K = len(props)
    m = 1.33 + 0.001j
    grid = A.fast_grid(box)
    bp = A.BatchProblem(pos, box, m, np.array(props), np.array(pols), fp32=True, grid=grid)
    out = {"workload": f"ellipsoid box {box[0]}x{box[1]}x{box[2]} ({len(pos)} dipoles), grid "
                       f"{bp.grid[0]}x{bp.grid[1]}x{bp.grid[2]}, m=1.33+0.001i, 32 orientations x 2 polarizations = "
                       f"{K} RHS, eps 1e-10", "rhs": K}
    try:
        op = bp.op
        out["autotune"] = {n: op.autotune_batch(K, prec, reps=3, cache=args.tune_cache)
                           for prec, n in ((A.F64, "fp64"), (A.F32, "fp32"))}
        for prec, n in ((A.F64, "fp64"), (A.F32, "fp32")):
            a, r = op.alloc_batch(K, prec), op.alloc_batch(K, prec)
            op.profile_matvec_batch(a, r, K, prec, nrep=2)
            passes, tot = op.profile_matvec_batch(a, r, K, prec, nrep=10)
            op.free(a)
            op.free(r)
            by = op.matvec_bytes_batch(K, prec)
            out[f"batched_matvec_{n}"] = {"us": round(tot * 1e3, 1), "rhs_matvecs_per_s": round(K * 1e3 / tot, 1),
                                          "GBps": round(by / (tot * 1e-3) / 1e9, 1),
                                          "pass_us": [round(x * 1e3, 1) for x in passes]}
        res = {}
        for method in ("bicgstab", "qmr"):
            variants = [("fp64", {}), ("mixed", {"precision": A.PREC_MIXED}), ("fp64_nograph", {"graph": False})]
            if method == "qmr":  # QMR_CS's FP32 Lanczos vectors are fragile: also FP64 vectors around FP32 transforms
                variants.append(("mixed_v64", {"precision": A.PREC_MIXED_V64}))
            for name, kw in variants:
                bp.solve(method, eps=1e-10, **kw)  # warm-up (graph capture, allocations)
                op.sync()
                t = time.perf_counter()
                st = bp.solve(method, eps=1e-10, **kw)
                cx = list(bp.cext_all())
                res[(method, name)] = (time.perf_counter() - t, cx, st)
            c64 = np.array(res[(method, "fp64")][1])
            cm = np.array(res[(method, "mixed")][1])
            out[method] = {"tts_batched_fp64_s": round(res[(method, "fp64")][0], 4),
                           "tts_batched_fp64_no_graph_s": round(res[(method, "fp64_nograph")][0], 4),
                           "tts_batched_mixed_s": round(res[(method, "mixed")][0], 4),
                           "iterations_fp64_max": max(s["niter"] for s in res[(method, "fp64")][2]),
                           "iterations_mixed_max": max(s["niter"] for s in res[(method, "mixed")][2]),
                           "max_true_resid_mixed": max(s["true_resid"] for s in res[(method, "mixed")][2]),
                           "cext_max_rel_err_mixed_vs_fp64": float(np.max(np.abs(cm - c64) / np.abs(c64)))}
            if (method, "mixed_v64") in res:
                tv, cv, sv = res[(method, "mixed_v64")]
                out[method]["fp64_vectors_fp32_transforms"] = {
                    "tts_batched_s": round(tv, 4), "iterations_max": max(s["niter"] for s in sv),
                    "max_true_resid": max(s["true_resid"] for s in sv),
                    "cext_max_rel_err_vs_fp64": float(np.max(np.abs(np.array(cv) - c64) / np.abs(c64)))}
        # post-processing of the last (FP64 QMR_CS) solution: Cext, Cabs and the 181-angle scattering-plane fields of
        # every RHS, one RHS at a time (ADDA's orientation loop) and in one batched evaluation
        bp.solve("qmr", eps=1e-10)
        op.sync()
        t = time.perf_counter()
        one = [(bp.cext(k), bp.cabs(k), bp.eplane(k, pars[k], pers[k], 181)) for k in range(K)]
        t_one = time.perf_counter() - t
        t = time.perf_counter()
        ce, ca, ep = bp.cext_all(), bp.cabs_all(), bp.eplane_all(pars, pers, 181)
        t_all = time.perf_counter() - t
        out["postprocessing"] = {
            "what": "Cext + Cabs + E in the 181-angle scattering plane, for all RHS", "per_rhs_s": round(t_one, 4),
            "batched_s": round(t_all, 4), "speedup": round(t_one / t_all, 2),
            "batched_equals_per_rhs": bool(all(ce[k] == one[k][0] and ca[k] == one[k][1] and
                                               np.array_equal(ep[k], one[k][2]) for k in range(K)))}
        # the same RHS one after another on one operator (ADDA's loop over orientations and polarizations)
        pr = A.Problem(pos, box, m, fp32=True, grid=grid)
        try:
            pr.op.autotune(A.F64, reps=3, cache=args.tune_cache)
            t = time.perf_counter()
            cs = []
            for k in range(K):
                pr.set_polarization(props[k], pols[k])
                pr.solve("qmr", eps=1e-10)
                cs.append(pr.cext())
            t_seq = time.perf_counter() - t
        finally:
            pr.close()
        out["qmr"]["tts_sequential_fp64_s"] = round(t_seq, 4)
        out["qmr"]["speedup_batched_fp64_vs_sequential"] = round(t_seq / out["qmr"]["tts_batched_fp64_s"], 3)
        out["qmr"]["speedup_batched_mixed_vs_sequential_fp64"] = round(t_seq / out["qmr"]["tts_batched_mixed_s"], 3)
        out["qmr"]["cext_max_rel_err_batched_vs_sequential"] = float(np.max(np.abs(np.array(cs) - np.array(
            res[("qmr", "fp64")][1])) / np.abs(np.array(cs))))
        log(f"batched RHS ({K}): " + json.dumps({k: out[k] for k in ("bicgstab", "qmr", "postprocessing")}))
    finally:
        bp.close()
    return out


def bench_adjoint(A):
    """Adjoint-gradient workload: 48^3 cube, per-voxel m (Re ~ U(1.2,1.8), Im ~ U(0,0.05), seed 20261204), QMR_CS
    to 1e-10: forward solve + the two batched adjoint solves + assembly, FP64 and FP32-mixed."""
    from cases import per_voxel_m, PROP_Z, POL_Y
    pos, _, box = A.make_particle("box", 48)
    nd = len(pos)
    pr = A.Problem(pos, box, per_voxel_m(nd), mat=np.arange(nd, dtype=np.int32), fp32=True)
    out = {"workload": f"cube {box[0]}^3 ({nd} voxels, one material each), grid {pr.grid[0]}^3, QMR_CS to 1e-10"}
    try:
        pr.set_polarization(PROP_Z, POL_Y)
        g = {}
        for name, prec in (("fp64", A.PREC_FP64), ("mixed", A.PREC_MIXED)):
            pr.solve("qmr", eps=1e-10, precision=prec)
            pr.adjoint_gradient("qmr", eps=1e-10, precision=prec)  # warm-up
            pr.op.sync()
            t0 = time.perf_counter()
            r = pr.solve("qmr", eps=1e-10, precision=prec)
            t1 = time.perf_counter()
            ge, ga, st = pr.adjoint_gradient("qmr", eps=1e-10, precision=prec)
            t2 = time.perf_counter()
            g[name] = (ge, ga)
            out[name] = {"forward_s": round(t1 - t0, 4), "adjoint_s": round(t2 - t1, 4),
                         "forward_iterations": r["niter"], "adjoint_iterations": [s["niter"] for s in st]}
        out["grad_max_rel_diff_mixed_vs_fp64"] = {
            q: float(np.abs(g["mixed"][i] - g["fp64"][i]).max() / np.abs(g["fp64"][i]).max())
            for i, q in enumerate(("cext", "cabs"))}
        log("adjoint: " + json.dumps(out))
    finally:
        pr.close()
    return out


if __name__ == "__main__":
    main()
