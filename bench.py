#!/usr/bin/env python3
"""Benchmark of the multifrontal numeric phase on one MI355X.

Workload (one "step" = one numeric factorization + refined solve of the batch of right-hand
sides, inputs already resident in HBM):
    3D Poisson 48^3 (n = 110592) + Diagonal(uniform[0, 0.1]), AMD ordering (host analysis,
    timed separately), SCALE_SUM, 8 RHS randn(seed 3), iterative refinement until the normwise
    backward error is <= 1e-14.
Both GPU paths run the same workload: the plain FP64 path and the mixed path (FP32 fronts on
MFMA with FP32 accumulation, FP64 residual and update).  `value` is the mixed path's
factor+solve seconds per step; the FP64 path, their ratio, the front TFLOP/s, the roofline of
the dominant kernel (Schur gemm, from HIP events in a profiled pass) and a CPU baseline (the
oracle's C restatement with OpenBLAS on the host cores) are reported beside it.

A same-pattern batch (512 matrices on the poisson3d(16) pattern, 64 RHS each, FP32 fronts +
FP64 refinement) is measured after it and reported under "batched".

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--k 48] [--nrhs 8] [--batch 512,16,64] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "pureumfpack.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from pureumfpack_hip import _lib  # noqa: E402
from pureumfpack_hip import interface as I  # noqa: E402
from pureumfpack_hip import matrices as M  # noqa: E402

METRIC = "factor+solve s & front TFLOP/s, mixed+refine vs FP64 GPU, bwd err≤1e-14"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}   # MI355X dense MFMA TFLOP/s
CLASSES = ["fused_small", "asm_cols", "panel_getrf", "panel_update_gemm", "schur_gemm", "fwd_solve",
           "bwd_solve", "schur_gemm_nf1024"]


class Hip:
    """Minimal HIP runtime binding for device buffers (no torch needed)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def download(self, p, shape, dtype=np.float64, order="F"):
        out = np.empty(shape, dtype=dtype, order=order)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self, p):
        self.h.hipFree(p)


def run_gpu(hip, A, S, Rs, B, precision, steps, warmup, tol, max_refine, graph=False):
    lib = _lib.load()
    n, nrhs = B.shape
    t0 = time.perf_counter()
    plan = I._Plan(A, S, precision=precision, use_graph=graph)
    t_plan = time.perf_counter() - t0
    dAx, dRs, dB = hip.upload(A.data), hip.upload(Rs), hip.upload(np.asfortranarray(B))
    dX = hip.upload(np.zeros((n, nrhs), order="F"))
    info = ctypes.c_int64()
    rep = _lib.mf_solve_report()

    def step():
        a = time.perf_counter()
        rc = lib.mf_factor_device(plan.h, dAx, dRs, 1, ctypes.byref(info))
        if rc:
            raise RuntimeError(lib.mf_last_error().decode())
        b = time.perf_counter()
        rc = lib.mf_solve_device(plan.h, nrhs, dB, n, dX, n, max_refine, tol, ctypes.byref(rep))
        if rc:
            raise RuntimeError(lib.mf_last_error().decode())
        c = time.perf_counter()
        return b - a, c - b

    for _ in range(warmup):
        step()
    hip.sync()
    fac, sol = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        f, s = step()
        fac.append(f)
        sol.append(s)
    hip.sync()
    total = time.perf_counter() - t0
    X = hip.download(dX, (n, nrhs))
    res = {
        "seconds": total / steps, "factor_s": float(np.mean(fac)), "solve_s": float(np.mean(sol)),
        "refine_iterations": int(rep.iterations), "backward_error": float(rep.backward_error),
        "residual_2": float(rep.residual_2), "plan_s": t_plan,
    }
    st = plan.stats()
    res["factor_flops"] = st["factor_flops"]
    res["front_tflops"] = st["factor_flops"] / res["factor_s"] / 1e12
    res["launches_per_factor"] = st["launches"]
    res["front_bytes"] = st["front_bytes"]
    res["store_bytes"] = st["store_bytes"]
    res["arena_bytes"] = st["arena_bytes"]
    res["resident_front_bytes"] = st["resident_front_bytes"]
    # profiled pass (HIP events around every launch) for the per-class roofline figures
    lib.mf_plan_set_profiling(plan.h, 1)
    step()
    ms = np.zeros(8); cnt = np.zeros(8, np.int64); fl = np.zeros(8); by = np.zeros(8)
    lib.mf_plan_get_profile(plan.h, ms.ctypes.data_as(_lib.c_f64p), cnt.ctypes.data_as(_lib.c_i64p),
                            fl.ctypes.data_as(_lib.c_f64p), by.ctypes.data_as(_lib.c_f64p))
    lib.mf_plan_set_profiling(plan.h, 0)
    res["profile"] = {CLASSES[c]: {"ms": float(ms[c]), "launches": int(cnt[c]), "gflop": float(fl[c] / 1e9),
                                   "gbytes": float(by[c] / 1e9),
                                   "tflops": float(fl[c] / (ms[c] * 1e-3) / 1e12) if ms[c] > 0 else 0.0,
                                   "gbps": float(by[c] / (ms[c] * 1e-3) / 1e9) if ms[c] > 0 else 0.0}
                      for c in range(8)}
    for p in (dAx, dRs, dB, dX):
        hip.free(p)
    del plan
    return res, X


def rel_error(X, Xref):
    """max over columns of ||x - x_ref||_inf / ||x_ref||_inf"""
    X = np.asarray(X).reshape(Xref.shape)
    return float((np.abs(X - Xref).max(axis=0) / np.abs(Xref).max(axis=0)).max())


def run_batch(hip, k, nmat, nrhs, precision, steps, warmup, tol, max_refine, graph=True, relax=False,
              solve_path="auto"):
    """Same-pattern batch: nmat matrices with the poisson3d(k) pattern and unsymmetric values
    (diag 6 + U[0,1], off-diagonals -U[0.5,1.5]), nrhs randn right-hand sides each; one step =
    mf_refactor_batch_device + mf_solve_batch_device refined to bwd err <= tol."""
    lib = _lib.load()
    P, Ax = M.poisson3d_batch_values(k, nmat, seed=4)
    n = P.shape[0]
    t0 = time.perf_counter()
    S = I.symbolic_mf(P, relax=relax)
    B = I.BatchedLU(P, nmat, precision=precision, use_graph=graph, symbolic=S, solve_path=solve_path)
    Rs = np.stack([I.row_scaling(B.matrix(v), I.SCALE_SUM) for v in Ax])
    t_plan = time.perf_counter() - t0
    rhs = np.random.default_rng(3).standard_normal((nmat, nrhs, n))   # per matrix n x nrhs, column-major
    dAx, dRs, dB = hip.upload(Ax), hip.upload(Rs), hip.upload(rhs)
    dX = hip.upload(np.zeros_like(rhs))
    info = np.zeros(nmat, np.int64)
    rep = _lib.mf_solve_report()
    plan = B._plan

    def step():
        a = time.perf_counter()
        rc = lib.mf_refactor_batch_device(plan.h, nmat, dAx, dRs, info.ctypes.data_as(_lib.c_i64p))
        if rc:
            raise RuntimeError(lib.mf_last_error().decode())
        b = time.perf_counter()
        rc = lib.mf_solve_batch_device(plan.h, nmat, nrhs, dB, dX, max_refine, tol, ctypes.byref(rep))
        if rc:
            raise RuntimeError(lib.mf_last_error().decode())
        return b - a, time.perf_counter() - b

    for _ in range(warmup):
        step()
    hip.sync()
    fac, sol = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        f, s_ = step()
        fac.append(f)
        sol.append(s_)
    hip.sync()
    total = time.perf_counter() - t0
    X = hip.download(dX, rhs.shape, order="C")
    rep_bwd, rep_it = rep.backward_error, rep.iterations
    # independent check of every matrix's backward error (host, FP64)
    worst = 0.0
    for m in range(nmat):
        A = B.matrix(Ax[m])
        R = rhs[m].T - A @ X[m].T
        den = abs(A).sum(axis=1).max() * np.abs(X[m]).max(axis=1) + np.abs(rhs[m]).max(axis=1)
        worst = max(worst, float((np.abs(R).max(axis=0) / den).max()))
    st = plan.stats()
    del B, plan
    # the same batch through the FP64 path once: each result's error against the FP64 solve
    B64 = I.BatchedLU(P, nmat, precision="fp64", use_graph=False, symbolic=S)
    rc = lib.mf_refactor_batch_device(B64._plan.h, nmat, dAx, dRs, info.ctypes.data_as(_lib.c_i64p))
    rc = rc or lib.mf_solve_batch_device(B64._plan.h, nmat, nrhs, dB, dX, 1, 0.0, ctypes.byref(rep))
    if rc:
        raise RuntimeError(lib.mf_last_error().decode())
    X64 = hip.download(dX, rhs.shape, order="C")
    err64 = max(rel_error(X[m].T, X64[m].T) for m in range(nmat))
    del B64
    for p_ in (dAx, dRs, dB, dX):
        hip.free(p_)
    fac_s = float(np.mean(fac))
    return {"workload": f"{nmat} x poisson3d-{k} pattern (n={n}), unsymmetric values, {nrhs} RHS each, "
                        f"{precision} fronts + FP64 refinement to bwd err <= {tol:g}, hipGraph={int(graph)}"
                        + (", relaxed amalgamation" if relax else ""),
            "nsuper": int(S.nsuper), "nlevels": int(st["nlevels"]),
            "seconds": total / steps, "factor_s": fac_s, "solve_s": float(np.mean(sol)),
            "matrices_per_s": nmat * steps / total, "refine_iterations": int(rep_it),
            "backward_error": float(rep_bwd), "backward_error_host_check": worst, "error_vs_fp64": err64,
            "front_tflops": st["factor_flops"] * nmat / fac_s / 1e12, "plan_s": t_plan,
            "launches_per_factor": st["launches"]}


def run_convdiff(hip, k, steps, warmup, relax=False):
    """Unsymmetric 3D convection-diffusion (upwind, velocity (1, 0.5, 0.25), cell Peclet 5),
    COLAMD ordering, BF16 Schur complements with per-tile power-of-two scaling, FP32
    getrf/trsm, FP64 refinement to backward error <= 1e-12; 1 RHS randn(5)."""
    A = M.convdiff3d(k)
    t0 = time.perf_counter()
    S = I.symbolic_mf(A, ordering="colamd", relax=relax)
    t_an = time.perf_counter() - t0
    Rs = I.row_scaling(A, I.SCALE_SUM)
    B = np.random.default_rng(5).standard_normal((A.shape[0], 1))
    res, X = run_gpu(hip, A, S, Rs, B, "bf16", steps, warmup, 1e-12, 60, graph=False)
    res.pop("profile", None)
    r64, X64 = run_gpu(hip, A, S, Rs, B, "fp64", 1, 0, 0.0, 1, graph=False)
    res.update({"workload": f"convdiff3d-{k} (n={A.shape[0]}), COLAMD, SCALE_SUM, 1 RHS, bf16 Schur + FP32 "
                            f"getrf/trsm, refine to bwd err <= 1e-12" + (", relaxed amalgamation" if relax else ""),
                "analysis_s": t_an, "nlevels": int(I.plan_estimate(A, S, precision="bf16")["nlevels"]),
                "nnz_LU": int(S.predicted_fill()), "nsuper": int(S.nsuper),
                "error_vs_fp64": rel_error(X, X64), "fp64_seconds": r64["seconds"],
                "fp64_backward_error": r64["backward_error"]})
    return res


def cpu_baseline(A, S, Rs, B, refine):
    """The oracle's C restatement of the reference multifrontal path (FP64, dense fronts through
    OpenBLAS), timed on the same matrix: numeric factor + solve(F, b; refine) per RHS column.
    Timed with 1 BLAS thread (the reference's published setting) and with every host core the
    process may use; the faster run is reported, both are kept."""
    from oracle import numeric as ON

    class _Q:
        qf = S.qf

    flat = (S.sstart, S.upd_ptr, S.upd_idx, S.child_ptr, S.child_idx)
    runs = {}
    for threads in sorted({1, len(os.sched_getaffinity(0))}):
        ON.use_openblas(threads=threads)
        t0 = time.perf_counter()
        F = ON.splu_mf(A, S=_Q, flat=flat, backend=1)
        t1 = time.perf_counter()
        for j in range(B.shape[1]):
            ON.solve(F, B[:, j], refine=refine)
        t2 = time.perf_counter()
        runs[threads] = {"value": t2 - t0, "factor_s": t1 - t0, "solve_s": t2 - t1}
        print(f"# cpu baseline {threads} thread(s): {t2 - t0:.2f} s", file=sys.stderr, flush=True)
        del F
    best = min(runs, key=lambda t: runs[t]["value"])
    return {"value": runs[best]["value"], "unit": "s", "cores": best, "kind": "port",
            "factor_s": runs[best]["factor_s"], "solve_s": runs[best]["solve_s"],
            "by_threads": {str(t): r for t, r in runs.items()},
            "sample": f"full workload: FP64 factor + {B.shape[1]} x solve(F, b; refine={refine})"}


def measured_traffic(k, path):
    """HBM bytes per Schur-gemm launch from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE calibrated against a copy kernel), when it covers this workload and path."""
    f = os.path.join(ROOT, "profiles", f"pmc_traffic_poisson3d{k}_{path}.md")
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("schur_gemm_nf1024_traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--k", type=int, default=48)
    ap.add_argument("--nrhs", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-14)
    ap.add_argument("--max-refine", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--paths", default="fp64,fp32,bf16,fp64+relax,fp32+relax,bf16+relax",
                    help="precisions to run; '+relax' uses the relaxed supernode amalgamation")
    ap.add_argument("--graph", type=int, default=1, help="replay the factor schedule and solve sweeps as hipGraphs")
    ap.add_argument("--tune", type=int, default=1, help="tune the gemm tile policy on the plan's Schur launches")
    ap.add_argument("--batch", default="512,16,64", help="nmat,k,nrhs of the same-pattern batch run ('' = skip)")
    ap.add_argument("--convdiff", type=int, default=64, help="grid of the conv-diff COLAMD + BF16 run (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many bytes (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        sec = ctypes.c_double()
        rc = _lib.load().mf_copy_probe(args.copy_probe, ctypes.byref(sec))
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "seconds": sec.value, "rc": rc}))
        sys.exit(0 if rc == 0 else 1)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the numeric phase runs on a single MI355X "
              "(multi-GPU sharding is a design, not a build target).", file=sys.stderr)
        sys.exit(2)

    A = M.poisson3d_perturbed(args.k, seed=3)
    n = A.shape[0]
    B = np.random.default_rng(3).standard_normal((n, args.nrhs))
    t0 = time.perf_counter()
    S = I.symbolic_mf(A)
    t_an = time.perf_counter() - t0
    paths = args.paths.split(",")
    S_relax = I.symbolic_mf(A, relax=True) if any(p.endswith("+relax") for p in paths) else None
    Rs = I.row_scaling(A, I.SCALE_SUM)
    hip = Hip()
    # tile policy per precision from the plan's own Schur launches: the cost model predicts, a
    # measurement confirms every non-default pick (a few hundred timings, under a second)
    tuning = {}
    if args.tune:
        from pureumfpack_hip import tuning as T
        for prec in sorted({p.split("+")[0] for p in paths} - {"bf16"}):
            t0 = time.perf_counter()
            tu = T.PlanTuner(S_relax if S_relax is not None else S, prec)
            pol = tu.with_model(seed=0)
            T.set_policy(prec, pol)
            tuning[prec] = {"policy_schur_by_bucket": pol[1].tolist(), "timings": tu.timings,
                            "pairs": len(tu.pairs()), "tune_s": time.perf_counter() - t0,
                            "schur_launches_ms_default": tu.policy_time(np.zeros_like(pol)) * 1e3,
                            "schur_launches_ms_tuned": tu.policy_time(pol) * 1e3}
    results, Xs = {}, {}
    for path in paths:
        prec, relaxed = path.split("+")[0], path.endswith("+relax")
        results[path], Xs[path] = run_gpu(hip, A, S_relax if relaxed else S, Rs, B, prec, args.steps, args.warmup,
                                          args.tol, args.max_refine, graph=bool(args.graph))
        results[path]["analysis"] = "relaxed amalgamation" if relaxed else "the reference's supernodes"
        print(f"# {path}: {json.dumps(results[path]['seconds'])} s/step, bwd {results[path]['backward_error']:.2e}, "
              f"{results[path]['refine_iterations']} refine", file=sys.stderr, flush=True)
    if "fp64" in Xs:
        for path in results:
            results[path]["error_vs_fp64"] = rel_error(Xs[path], Xs["fp64"])
    # `value`: the fastest mixed-precision path that reaches the backward-error target
    mixed = [p for p in results if not p.startswith("fp64") and results[p]["backward_error"] <= args.tol]
    main_p = min(mixed, key=lambda p: results[p]["seconds"]) if mixed else list(results)[0]
    R = results[main_p]
    prec_main = main_p.split("+")[0]
    same_fp64 = "fp64+relax" if main_p.endswith("+relax") else "fp64"
    # dominant MFMA kernel: the Schur complements of the fronts with nf >= 1024 (the FP32 gemm;
    # BF16 paths run k_gemm_bf16 there)
    prof = R["profile"]["schur_gemm_nf1024"]
    peak = PEAK[prec_main]
    roof = {"bound": "mfma",
            "kernel": ("k_gemm_bf16" if prec_main == "bf16" else "k_gemm<T, 64, 64, 2>") +
                      " (Schur complements, fronts nf >= 1024)",
            "achieved": prof["tflops"],
            "peak": peak, "unit": "TFLOP/s", "frac": prof["tflops"] / peak,
            "traffic": measured_traffic(args.k, main_p.replace("+", "_")),
            "algorithmic_bytes_per_launch": prof["gbytes"] * 1e9 / max(1, prof["launches"]),
            "launch_us": prof["ms"] * 1e3 / max(1, prof["launches"]), "path": main_p}
    analysis = S_relax if main_p.endswith("+relax") else S
    out = {
        "metric": METRIC, "value": R["seconds"], "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": R["seconds"] * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": prec_main, "data": "synthetic",
        "config": {"workload": f"poisson3d-{args.k} + Diagonal(U[0,0.1]), {args.nrhs} RHS randn(3), AMD, SCALE_SUM, "
                               f"refine to bwd err <= {args.tol:g}", "n": n, "nnz": int(A.nnz), "nrhs": args.nrhs,
                   "nsuper": int(analysis.nsuper), "nnz_LU": int(analysis.predicted_fill()), "path": main_p,
                   "analysis": R["analysis"]},
        "front_tflops": R["front_tflops"], "analysis_s": t_an,
        "paths": results,
        "speedup_vs_fp64": (results["fp64"]["seconds"] / R["seconds"]) if "fp64" in results else None,
        "speedup_vs_fp64_same_analysis": (results[same_fp64]["seconds"] / R["seconds"]) if same_fp64 in results
        else None,
        "roofline": roof,
        "tuning": tuning,
    }
    if args.batch:
        nmat, kb, nrb = (int(v) for v in args.batch.split(","))
        for key, relaxed in (("batched", False), ("batched_relaxed", True)):
            out[key] = run_batch(hip, kb, nmat, nrb, "fp32", args.steps, args.warmup, args.tol, args.max_refine,
                                 graph=bool(args.graph), relax=relaxed)
            print(f"# {key}: {out[key]['seconds']:.4f} s/step, {out[key]['matrices_per_s']:.0f} matrices/s, "
                  f"bwd {out[key]['backward_error_host_check']:.2e}", file=sys.stderr, flush=True)
    if args.convdiff:
        for key, relaxed, steps, warmup in (("convdiff_colamd_bf16_relaxed", True, args.steps, args.warmup),
                                            ("convdiff_colamd_bf16", False, 2, 1)):
            out[key] = run_convdiff(hip, args.convdiff, steps, warmup, relax=relaxed)
            cd = out[key]
            print(f"# {key} {args.convdiff}^3: {cd['seconds']:.4f} s/step, bwd {cd['backward_error']:.2e}, "
                  f"{cd['refine_iterations']} refine", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, S, Rs, B, refine=1)
    # accuracy of every leg against its target: a timing of wrong answers is no timing
    failed = [f"{p}: bwd {r['backward_error']:.2e}" for p, r in results.items() if r["backward_error"] > args.tol]
    for key in ("batched", "batched_relaxed"):
        if key in out and out[key]["backward_error_host_check"] > args.tol:
            failed.append(f"{key}: bwd {out[key]['backward_error_host_check']:.2e}")
    for key in ("convdiff_colamd_bf16", "convdiff_colamd_bf16_relaxed"):
        if key in out and out[key]["backward_error"] > 1e-12:
            failed.append(f"{key}: bwd {out[key]['backward_error']:.2e}")
    out["accuracy_ok"] = not failed
    if failed:
        print("# ACCURACY CHECK FAILED: " + "; ".join(failed), file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
