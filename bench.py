#!/usr/bin/env python
"""Benchmark of the cluster-coupling hot path on one MI355X.

Workload: N = 1024 spheres, lmax = 6 (98,304 unknowns), radii U[0.15, 0.35],
permittivity U[2, 6] + i U[0, 0.2], centres in a cube of side 16 with surface gap
>= 0.05, k0 = 1.3, four plane-wave right-hand sides (+-z x +-helicity).  One step =
the whole job for one cluster: build the coupling (c64 and FP64 pair harmonic
tables), build the learned preconditioner, and solve the four columns by
mixed-precision GMRES(50) (inner solves on the c64-precision coupling, FP16 MFMA) to a true
FP64 relative residual <= 1e-8.  value = solves/s = 4 K / (time of K steps).

Also measured in the same run: the same solve on the dense stored c64 coupling
(assembly + streaming matvec, with their HBM rooflines), the plain FP64
block-Jacobi GMRES(50) on the same GPU (stored c128 coupling and FP64 pair-table
GEMMs), the per-kernel rooflines (HIP events on the library's stream), the dense LU at N = 512,
lmax = 4 (c64 with FP64 refinement and c128; MFMA TFLOP/s of the factorisation),
the cross-section gradient at N = 256, lmax = 4 through the LU adjoint, and the
reference algorithm on the host cores: the FP64 oracle's matrix-free operator
(OpenMP) under batched unpreconditioned GMRES for the same 4 columns, timed whole.
The oracle then checks the headline answer (true residual, solution, cross
sections, coupling rows) in "oracle_check".

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-lu] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s @N=1024,lmax=6 to 1e-8 resid: learned-PC vs plain; %HBM/MFMA peak"
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
FP16_MFMA_PEAK_TFLOPS = 2500.0
FP64_MFMA_PEAK_TFLOPS = 78.6


def subset_modes(pkg, lmax: int) -> int:
    """Sum over harmonic indices t of the source modes a that some plan term (b, a, t)
    uses: the GEMM columns per helicity and right-hand side."""
    starts, nsame, w, hidx = pkg.plan_terms(lmax)
    H = lmax * (lmax + 2)
    pairs = set()
    for e in range(H * H):
        for k in range(starts[e], starts[e + 1]):
            pairs.add((int(hidx[k]), e // H))
    return len(pairs)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rhs_columns(cl, torch):
    cols = [cl.plane_wave([0, 0, 1], 1), cl.plane_wave([0, 0, 1], 0), cl.plane_wave([0, 0, -1], 1),
            cl.plane_wave([0, 0, -1], 0)]
    return torch.stack(cols, dim=1)


def run_method(cl, B, torch, *, mixed, pc, steps, warmup, rtol, restart, label, stored=True):
    """Time `steps` whole-job steps (storage or tables + preconditioner + solve).  FP64
    solves (mixed = False) run on the stored c128 coupling, or matrix-free with
    stored = False."""
    conv = None
    x = None
    for w in range(warmup):
        cl.release()
        x, conv = cl.solve(B, rtol=rtol, restart=restart, max_iterations=2000, mixed_precision=mixed,
                           preconditioner=pc, convergence="full")
        log(f"{label} warmup {w + 1}/{warmup}: iterations {[c.iterations for c in conv]}")
    # storage left by an earlier method (up to the 77 GB c128 coupling) is freed before
    # the clock starts; every step then recomputes the coupling (store or tables) and
    # the preconditioner into the memory the previous step used (reset, not release:
    # re-allocating a 38.65 GB store is memory management, not the job)
    cl.release()
    torch.cuda.synchronize()
    cl.timing_reset()
    cl.timing(True)
    t0 = time.perf_counter()
    for s in range(steps):
        ts = time.perf_counter()
        cl.reset()
        if not mixed and stored:
            cl.assemble("c128")
        x, conv = cl.solve(B, rtol=rtol, restart=restart, max_iterations=2000, mixed_precision=mixed,
                           preconditioner=pc, convergence="full")
        torch.cuda.synchronize()
        log(f"{label} step {s + 1}/{steps}: {time.perf_counter() - ts:.3f} s, iterations "
            f"{[c.iterations for c in conv]}, refinements {[c.refinements for c in conv]}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cl.timing(False)
    kernels = {}
    for name in ("assemble", "matvec_stored", "coupling_mf", "coupling_mf32", "gemm_c64", "gemm_split16",
                 "gemm_split16_hi", "gemm_f64", "pc_apply", "pc_factor"):
        ms, count = cl.timing_get(name)
        if count:
            kernels[name] = {"total_ms": ms, "launches": count, "avg_ms": ms / count}
    return x, conv, elapsed, kernels


_WRITE_CEILING = None


def write_ceiling(torch, gib=16):
    """The write-only HBM rate this device reaches: the HIP runtime's fill kernel
    (hipMemsetAsync, called on torch's current stream) over a 16 GiB buffer, best of 5
    after a warm-up.  The dense assembly is a write stream, so this, not the 8 TB/s
    read-and-write peak, is what it can approach."""
    global _WRITE_CEILING
    if _WRITE_CEILING is None:
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemsetAsync.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
        hip.hipMemsetAsync.restype = ctypes.c_int
        nbytes = gib << 30
        buf = torch.empty(nbytes // 16, dtype=torch.complex128, device="cuda")
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        if hip.hipMemsetAsync(ctypes.c_void_p(buf.data_ptr()), 0, nbytes, stream) != 0:
            raise RuntimeError("hipMemsetAsync failed")
        best = float("inf")
        for _ in range(5):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            if hip.hipMemsetAsync(ctypes.c_void_p(buf.data_ptr()), 0, nbytes, stream) != 0:
                raise RuntimeError("hipMemsetAsync failed")
            b.record()
            b.synchronize()
            best = min(best, a.elapsed_time(b))
        del buf
        torch.cuda.empty_cache()
        _WRITE_CEILING = {"gbs": nbytes / (best * 1e-3) / 1e9, "bytes": nbytes, "ms": best,
                          "kernel": "hipMemsetAsync of a 16 GiB buffer (the runtime's fill kernel), best of 5"}
    return _WRITE_CEILING


def cpu_info():
    """Host CPU model, logical CPUs and the threads the OpenMP oracle runs with."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", affinity))
    return model, os.cpu_count() or 1, affinity, threads


def cpu_baseline(lmax, k0, radii, eps, pos, A, gpu, rtol, restart, oracle):
    """The reference algorithm on the host cores: the oracle's matrix-free FP64
    operator (cluster/iterative.rs:187-232 restated, OpenMP) under batched,
    unpreconditioned GMRES (linalg/gmres.rs:119-185 restated) for the same 4 columns
    to the same tolerance, timed whole.  The oracle then checks the GPU's answer (not
    timed): the true residual of the GPU solution by one more oracle apply, the
    solution and cross sections against the oracle's, and the FP64 and c64-precision
    coupling applies on 8 destination spheres."""
    model, nproc, affinity, threads = cpu_info()
    mf = oracle.MatrixFreeCluster(lmax, k0, radii, eps, pos)
    t0 = time.perf_counter()
    X, conv = mf.solve(A, rtol=rtol, restart=restart, max_iterations=2000)
    t_solve = time.perf_counter() - t0
    its = [c.iterations for c in conv]
    log(f"cpu baseline: oracle GMRES {t_solve:.1f} s, iterations {its}")
    out = {
        "value": A.shape[1] / t_solve,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": model, "nproc": nproc, "affinity_cpus": affinity,
        "sample": (f"the full job on the host: oracle FP64 matrix-free GMRES({restart}) of all {A.shape[1]} "
                   f"columns to {rtol:g} true residual, {its} iterations, {t_solve:.1f} s "
                   f"({threads} OpenMP threads on {model}, {affinity} of {nproc} CPUs); measured, not extrapolated"),
    }
    b = mf.rhs(A)
    r = b - mf.apply(gpu["x"])
    res = [float(np.linalg.norm(r[:, c]) / np.linalg.norm(b[:, c])) for c in range(A.shape[1])]
    sca, ext = mf.cross_sections(A, X)
    cs = max(max(abs(gpu["sca"][c] - sca[c]) / abs(sca[c]), abs(gpu["ext"][c] - ext[c]) / abs(ext[c]))
             for c in range(A.shape[1]))
    rows = gpu["rows"]
    ref_rows = mf.coupling(gpu["x"], rows=rows)[rows[0] * mf.M:rows[1] * mf.M]
    scale = np.linalg.norm(ref_rows)
    out["oracle_check"] = {
        "gpu_true_relative_residual": res,
        "solution_rel_error": float(np.linalg.norm(gpu["x"] - X) / np.linalg.norm(X)),
        "cross_section_rel_error": float(cs),
        "coupling_rel_error_fp64": float(np.linalg.norm(gpu["cx64"] - ref_rows) / scale),
        "coupling_rel_error_c64_inner": float(np.linalg.norm(gpu["cx16"] - ref_rows) / scale),
        "coupling_rows": f"destination spheres {rows[0]}..{rows[1] - 1}, the 4 headline solution columns",
        "oracle_iterations": its,
        "note": "the GPU headline solution and cross sections against the FP64 oracle's solve of the same system",
    }
    log(f"oracle check: residuals {res}, cross sections {cs:.2e}")
    return out


def dense_lu(pkg, torch, seed):
    """Dense LU at N = 512, lmax = 4 (24,576 unknowns), 4 plane waves: c64 factors on
    MFMA with FP64 refinement against c128 factors, both on this GPU."""
    lmax, radii, eps, pos = pkg.workload("n512_l4", seed=seed)
    cl = pkg.IterativeSphereCluster(lmax, 1.3, radii, eps, pos)
    B = rhs_columns(cl, torch)
    n = cl.n
    flops = 8.0 / 3.0 * n ** 3
    out = {"workload": f"N={cl.N}, lmax={lmax} ({n} unknowns), 4 plane waves, rtol 1e-12",
           "algorithmic_flops": flops}
    xs = {}
    for prec, peak in (("c64", 157.3), ("c128", 78.6)):
        cl.timing_reset()
        cl.timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = pkg.SphereClusterFactor(cl, prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x, conv = f.solve(B)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        cl.timing(False)
        lu_ms = cl.timing_get("lu_factor")[0]
        asm_ms = cl.timing_get("assemble_operator")[0]
        tf = flops / (lu_ms * 1e-3) / 1e12
        out[prec] = {"factor_ms": 1e3 * (t1 - t0), "assemble_operator_ms": asm_ms, "lu_kernels_ms": lu_ms,
                     "tflops": tf, "peak_tflops": peak, "frac": tf / peak,
                     "solve_ms_4rhs": 1e3 * (t2 - t1), "refinements": [c.refinements for c in conv],
                     "true_relative_residual": [c.residual_norm / c.rhs_norm for c in conv]}
        xs[prec] = x
        log(f"dense LU {prec}: factor {1e3 * (t1 - t0):.0f} ms ({tf:.1f} TFLOP/s), solve {1e3 * (t2 - t1):.0f} ms")
        if prec == "c128":
            # the cluster's T-matrix about the origin (lmax 4): 48 columns solved with the
            # factors, then one GEMM
            t3 = time.perf_counter()
            Tg, gconv = f.global_tmatrix()
            torch.cuda.synchronize()
            out["global_tmatrix_ms"] = 1e3 * (time.perf_counter() - t3)
            out["global_tmatrix_columns"] = len(gconv)
            log(f"global T-matrix (lmax {lmax}, {len(gconv)} columns): {out['global_tmatrix_ms']:.1f} ms")
        f.close()
    out["c64_vs_c128_solution_rel_diff"] = float(torch.linalg.norm(xs["c64"] - xs["c128"]) /
                                                 torch.linalg.norm(xs["c128"]))
    return out


def interacting_tmatrix(pkg, torch):
    """The reference's dense benchmark (docs/performance/evidence.md:152-176,
    scripts/benchmark_cluster.py:281-284): chains of spheres with radii linspace(0.15,
    0.25), spacing 0.8, eps 4+0.1i, k0 1.3; particle coefficients, translations and the
    full interacting T-matrix X = (I - T C)^{-1} T (all n columns), here c128 factors
    and FP64-refined forward solves.  Timed the reference's way
    (scripts/benchmark_cluster.py:1395-1405): one warm-up call, a batch of calls per
    sample sized to about 20 ms, 7 samples, the median per-call time.  Published
    4-thread CPU medians beside."""
    import statistics

    published = {(8, 3): 1.96, (16, 3): 9.35, (8, 5): 12.69, (32, 3): 43.41}
    out = []
    for (N, lmax), ref_ms in published.items():
        radii = np.linspace(0.15, 0.25, N)
        pos = np.column_stack([np.arange(N) * 0.8, np.zeros((N, 2))])
        eps = np.full(N, 4 + 0.1j)

        def call():
            cl = pkg.IterativeSphereCluster(lmax, 1.3, radii, eps, pos)
            f = pkg.SphereClusterFactor(cl, "c128")
            X, _ = f.solve(torch.eye(cl.n, dtype=torch.complex128, device="cuda"))
            torch.cuda.synchronize()
            f.close()
            del cl
            return X

        call()  # warm-up
        t0 = time.perf_counter()
        X = call()
        batch = max(1, int(np.ceil(0.02 / (time.perf_counter() - t0))))
        samples = []
        for _ in range(7):
            t0 = time.perf_counter()
            for _ in range(batch):
                X = call()
            samples.append((time.perf_counter() - t0) / batch)
        ms = 1e3 * statistics.median(samples)
        out.append({"spheres": N, "lmax": lmax, "n": int(X.shape[0]), "ms": ms, "min_ms": 1e3 * min(samples),
                    "calls_per_sample": batch, "reference_cpu_4_threads_ms": ref_ms, "speedup": ref_ms / ms})
        log(f"interacting T-matrix N={N} lmax={lmax} (n={X.shape[0]}): {ms:.2f} ms median of 7 x {batch} calls "
            f"(reference CPU, 4 threads: {ref_ms} ms)")
    return out


def illumination_case(n):
    """scripts/benchmark_illumination.py:55-86 ("weak"): spheres on a cube grid of
    spacing 1.3, radii linspace(0.12, 0.18), eps 2.2+0.02i, k0 1.3, one plane wave of
    direction (0.13, -0.11, 1) / norm and polarisation 0."""
    side = int(np.ceil(n ** (1 / 3)))
    grid = np.column_stack(np.unravel_index(np.arange(n), (side, side, side)))
    direction = np.array([0.13, -0.11, 1.0])
    return np.linspace(0.12, 0.18, n), np.full(n, 2.2 + 0.02j), grid * 1.3, direction / np.linalg.norm(direction)


def reference_illumination(pkg, torch):
    """The reference's requested-illumination benchmark (docs/performance/
    large-problems.md:27-50): lmax 1, one incident column, GMRES tolerance 1e-10
    (restart 30).  "Fresh" = cluster set-up (Mie, coupling or tables, factor) plus the
    solve; "pullback" = gradients with respect to k0, radii, permittivities, positions
    and the incident field from a recorded solve.  Here: c128 LU ("selected LU": the
    requested column; "full T": all n columns, then X a), or FP64 matrix-free GMRES.
    Published 4-thread CPU seconds beside (9950X)."""
    published = {(512, "full T"): (0.940, 1.885), (512, "selected LU"): (0.381, 0.764),
                 (512, "matrix-free"): (0.544, 0.711), (1024, "matrix-free"): (2.192, 2.826)}
    out = []
    for (n, method), (ref_fresh, ref_pull) in published.items():
        radii, eps, pos, direction = illumination_case(n)
        cl0 = pkg.IterativeSphereCluster(1, 1.3, radii, eps, pos)
        a = cl0.plane_wave(direction, 0)  # input construction, excluded as in the reference
        del cl0
        fresh, pull = [], []
        for rep in range(3):  # the first is the warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cl = pkg.IterativeSphereCluster(1, 1.3, radii, eps, pos)
            if method == "matrix-free":
                x, ctx, conv = cl.record(a, rtol=1e-10, restart=30)
            else:
                f = pkg.SphereClusterFactor(cl, "c128", rtol=1e-10)
                if method == "full T":
                    X, _ = f.solve(torch.eye(cl.n, dtype=torch.complex128, device="cuda"))
                    x = X @ a
                x2, ctx, conv = f.record(a)
                if method == "selected LU":
                    x = x2
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            ctx.pullback(torch.ones_like(x))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if rep:
                fresh.append(t1 - t0)
                pull.append(t2 - t1)
            if method != "matrix-free":
                f.close()
            del cl
        fs, ps = min(fresh), min(pull)
        out.append({"spheres": n, "lmax": 1, "method": method, "fresh_s": fs, "pullback_s": ps,
                    "reference_cpu_4_threads_fresh_s": ref_fresh, "reference_cpu_4_threads_pullback_s": ref_pull,
                    "speedup_fresh": ref_fresh / fs, "speedup_pullback": ref_pull / ps})
        log(f"illumination N={n} lmax=1 {method}: fresh {1e3 * fs:.2f} ms, pullback {1e3 * ps:.2f} ms "
            f"(reference CPU, 4 threads: {ref_fresh} s, {ref_pull} s)")
    return out


def gradient_lu(pkg, torch, seed):
    """sigma_sca of a +z positive-helicity plane wave and its gradient with respect to
    all 768 position components and 256 radii (N = 256, lmax = 4, 12,288 unknowns),
    the adjoint reusing the c128 LU factors."""
    lmax, radii, eps, pos = pkg.workload("n256_l4_grad", seed=seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f = pkg.sphere_cluster_factor(lmax, 1.3, radii, eps, pos)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sca, gr, gp, info = pkg.scattering_gradient(f, [0, 0, 1], 1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    f.close()
    # translation invariance: the position gradient sums to zero
    drift = float(np.abs(gp.sum(axis=0)).max() / np.abs(gp).max())
    log(f"gradient (LU): factor {1e3 * (t1 - t0):.0f} ms, solve + adjoint + pullback {1e3 * (t2 - t1):.0f} ms")
    return {"workload": f"N={radii.size}, lmax={lmax} ({f.cluster.n} unknowns), c128 LU",
            "factor_ms": 1e3 * (t1 - t0), "solve_adjoint_pullback_ms": 1e3 * (t2 - t1),
            "scattering_cross_section": sca, "position_gradient_sum_rel": drift}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="n1024_l6")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--pc", default="learned")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--bj-group", type=int, default=4)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--restart", type=int, default=50)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-reference", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lu", action="store_true")
    ap.add_argument("--dense-benchmark-only", action="store_true",
                    help="run only the reference's dense benchmark (interacting T-matrices of sphere chains, "
                         "n = 240 to 960) and print its results as one JSON line (for profiling)")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--operator", choices=("tables", "stored"), default="tables",
                    help="c64 coupling of the headline inner solves: pair tables (GEMMs) or the dense store")
    ap.add_argument("--relax", type=float, default=None,
                    help="relaxed inner applies: threshold of the allowed operator error at which a split-FP16 "
                         "inner apply multiplies the hi halves only (0 off; default the library's)")
    ap.add_argument("--inner", choices=("split16", "fp32"), default="split16",
                    help="arithmetic of the c64-precision table GEMMs: scaled FP16 hi + lo halves on FP16 MFMA, "
                         "or c64 tables on FP32 MFMA")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="bytes of a plain device copy run first (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: the path is built for one GPU's share "
                                   "(pair-table row blocks per GPU are a design, not code), so --gpus must be 1"}))
        return 2

    import torch

    from __graft_entry__ import load_package

    pkg = load_package()
    torch.cuda.set_device(0)
    if args.copy_calibration > 0:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch, divided
        # by the size, is the counter scale applied to the workload's kernels
        src = torch.empty(args.copy_calibration // 4, dtype=torch.float32, device="cuda").uniform_()
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {src.numel() * 4} bytes x 3")
        del src, dst
    if args.dense_benchmark_only:
        print(json.dumps({"interacting_tmatrix": interacting_tmatrix(pkg, torch)}))
        return 0
    lmax, radii, eps, pos = pkg.workload(args.workload, seed=args.seed)
    k0 = 1.3
    N = radii.size
    log(f"workload {args.workload}: N={N} lmax={lmax} n={N * 2 * lmax * (lmax + 2)}")
    cl = pkg.IterativeSphereCluster(lmax, k0, radii, eps, pos)
    cl.set_inner_gemm(args.inner)
    if args.relax is not None:
        cl.set_inner_relaxation(args.relax)
    cl.set_group_size(args.bj_group)
    if args.pc == "learned":
        import importlib

        learned = importlib.import_module("treams_rs_amd.learned")
        mlp = learned.MLP.load(args.weights or learned.DEFAULT_WEIGHTS)
        learned.attach(cl, mlp)
    n, H = cl.n, lmax * (lmax + 2)
    nh = N * H
    B = rhs_columns(cl, torch)
    ncols = B.shape[1]

    def measure(operator, label):
        """Learned-PC (or --pc) mixed-precision job on the c64-precision coupling held
        as pair tables (split-FP16 or FP32 MFMA GEMMs, --inner) or as the dense
        helicity-split store (streaming matvec); with its rooflines."""
        cl.release()
        xm, cv, el, kern = run_method(cl, B, torch, mixed="matrix_free" if operator == "tables" else True,
                                      pc=args.pc, steps=args.steps, warmup=args.warmup, rtol=args.rtol,
                                      restart=args.restart, label=label)
        out = {"value": ncols * args.steps / el, "unit": "solves/s", "ms_per_step": 1e3 * el / args.steps,
               "iterations": [c.iterations for c in cv], "refinements": [c.refinements for c in cv],
               "true_relative_residual": [c.residual_norm / c.rhs_norm for c in cv], "kernels": kern}
        roofs = {}
        gemm = kern.get("gemm_split16", {})
        if gemm:
            cmacs = float(N) * N * 2 * ncols * subset_modes(pkg, lmax)
            achieved = 8.0 * cmacs / (gemm["avg_ms"] * 1e-3) / 1e12
            executed = 24.0 * cmacs / (gemm["avg_ms"] * 1e-3) / 1e12
            roofs["gemm"] = {
                "bound": "mfma",
                "kernel": "k_zgemm_subsets_split16 (c64-precision pair-table GEMMs on FP16 MFMA, scaled hi + lo "
                          "halves, 4 RHS, 4M form)",
                "achieved": achieved, "peak": FP16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP16_MFMA_PEAK_TFLOPS, "traffic": None, "flops_per_launch": 8.0 * cmacs,
                "avg_ms": gemm["avg_ms"], "executed_fp16_tflops": executed,
                "executed_frac": executed / FP16_MFMA_PEAK_TFLOPS,
                "flop_count": "8 real flops per complex multiply-add over the nonzero (harmonic, source mode) "
                              "columns; the split 4M form executes 24 FP16 MFMA flops per complex multiply-add "
                              "(4 real products x 3 half products)"}
            add_traffic(roofs["gemm"], "gemm_split16_traffic_n1024_l6.json")
        gemm = kern.get("gemm_split16_hi", {})
        if gemm:
            cmacs = float(N) * N * 2 * ncols * subset_modes(pkg, lmax)
            achieved = 8.0 * cmacs / (gemm["avg_ms"] * 1e-3) / 1e12
            roofs["gemm_hi"] = {
                "bound": "mfma",
                "kernel": "k_zgemm_subsets_split16<hi only> (the relaxed inner applies: split-FP16 pair tables, "
                          "hi halves only, 4 RHS, 4M form)",
                "achieved": achieved, "peak": FP16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP16_MFMA_PEAK_TFLOPS, "traffic": None, "flops_per_launch": 8.0 * cmacs,
                "avg_ms": gemm["avg_ms"], "launches_per_job": gemm["launches"] / args.steps,
                "flop_count": "8 real flops per complex multiply-add over the nonzero (harmonic, source mode) "
                              "columns, which the hi-only 4M form also executes (4 real products x 1 half product)"}
            add_traffic(roofs["gemm_hi"], "gemm_split16_hi_traffic_n1024_l6.json")
        gemm = kern.get("gemm_f64", {})
        if gemm:
            cmacs = float(N) * N * 2 * ncols * subset_modes(pkg, lmax)
            achieved = 8.0 * cmacs / (gemm["avg_ms"] * 1e-3) / 1e12
            roofs["gemm_f64"] = {
                "bound": "mfma",
                "kernel": "k_zgemm_subsets<double> (FP64 pair-table GEMMs of the refinement residuals and the "
                          "certificate, 4 RHS, 3M form on v_mfma_f64_16x16x4f64)",
                "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_MFMA_PEAK_TFLOPS, "traffic": None, "flops_per_launch": 8.0 * cmacs,
                "avg_ms": gemm["avg_ms"], "launches_per_job": gemm["launches"] / args.steps,
                "flop_count": "8 real flops per complex multiply-add over the nonzero (harmonic, source mode) "
                              "columns; the 3M form executes 6 on MFMA"}
            add_traffic(roofs["gemm_f64"], "gemm_f64_traffic_n1024_l6.json")
        gemm = kern.get("gemm_c64", {})
        if gemm:
            cmacs = float(N) * N * 2 * ncols * subset_modes(pkg, lmax)
            achieved = 8.0 * cmacs / (gemm["avg_ms"] * 1e-3) / 1e12
            roofs["gemm"] = {
                "bound": "mfma", "kernel": "k_zgemm_subsets<float> (c64 pair-table GEMMs, 4 RHS, 3M form)",
                "achieved": achieved, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_MFMA_PEAK_TFLOPS, "traffic": None, "flops_per_launch": 8.0 * cmacs,
                "avg_ms": gemm["avg_ms"],
                "flop_count": "8 real flops per complex multiply-add over the nonzero (harmonic, source mode) "
                              "columns; the 3M form executes 6 on MFMA"}
            add_traffic(roofs["gemm"], "gemm_c64_traffic_n1024_l6.json")
        mv = kern.get("matvec_stored", {})
        if mv:
            achieved = c64_bytes / (mv["avg_ms"] * 1e-3) / 1e9
            roofs["matvec"] = {
                "bound": "hbm", "kernel": "k_matvec_stored<float2> (dense c64 coupling, 4 RHS)",
                "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": None, "bytes_per_launch": c64_bytes, "avg_ms": mv["avg_ms"]}
            add_traffic(roofs["matvec"], "matvec_traffic_n1024_l6.json")
        asm = kern.get("assemble", {})
        if asm:
            achieved = c64_bytes / (asm["avg_ms"] * 1e-3) / 1e9
            wc = write_ceiling(torch)
            roofs["assembly"] = {"bound": "hbm", "kernel": "k_assemble (dense c64 coupling, written once)",
                                 "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": achieved / HBM_PEAK_GBS, "bytes_per_launch": c64_bytes,
                                 "avg_ms": asm["avg_ms"], "write_only_ceiling": wc,
                                 "frac_of_write_only_ceiling": achieved / wc["gbs"]}
        return xm, cv, out, roofs

    def add_traffic(roof, name):
        tpath = os.path.join(ROOT, "profiles", name)
        if args.workload == "n1024_l6" and os.path.exists(tpath):
            with open(tpath) as fh:
                t = json.load(fh)
            roof["traffic"] = t["traffic_bytes"]
            roof["traffic_source"] = t.get("source", f"profiles/{name}")

    c64_bytes = 2 * nh * nh * 8
    # the headline: inner solves on the c64-precision coupling held as pair harmonic
    # tables (split-FP16 MFMA GEMMs by default) unless --operator stored; FP64
    # refinement and the true-residual certificate on the FP64 tables
    x, conv, head, roofs = measure(args.operator, "mixed")
    elapsed = head["ms_per_step"] * 1e-3 * args.steps
    value = head["value"]
    kernels = head["kernels"]
    # the roofline of the kernel that takes the most time per job; the other GEMMs beside it
    if args.operator == "tables":
        timed = {"gemm_f64": "gemm_f64", "gemm": "gemm_split16", "gemm_hi": "gemm_split16_hi"}
        timed = {k: kernels[v]["total_ms"] for k, v in timed.items() if k in roofs and v in kernels}
        if "gemm_c64" in kernels and "gemm" in roofs:
            timed["gemm"] = kernels["gemm_c64"]["total_ms"]
        dominant = max(timed, key=timed.get)
        roof = dict(roofs[dominant])
        roof["kernel_ms_per_job"] = timed[dominant] / args.steps
    else:
        roof = roofs.get("matvec")
    iterations = head["iterations"]
    x_mixed = x
    tables = args.operator == "tables"
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (("c64-precision coupling (pair tables as scaled FP16 hi + lo halves on FP16 MFMA"
                   if args.inner == "split16" else "c64 coupling (pair tables on FP32 MFMA")
                  if tables else "c64 coupling (dense store") +
                 ") with FP32 accumulate for the inner solves; c128 tables, Krylov vectors, residuals and refinement",
        "data": "synthetic",
        "config": {"workload": f"N={N} spheres, lmax={lmax} ({n} unknowns), radii U[0.15,0.35], eps U[2,6]+iU[0,0.2], "
                               f"cube side 16, gap 0.05, k0 1.3, 4 plane waves (+-z x +-helicity), "
                               f"GMRES({args.restart}) to {args.rtol:g} true FP64 relative residual",
                   "coupling": (("pair harmonic tables (N^2 (2 lmax + 1)^2 values) as scaled FP16 hi + lo "
                                 "halves (1.4 GB) applied by FP16 MFMA GEMMs (4M form, 3 half products per real "
                                 "product)" if args.inner == "split16" else
                                 "c64 pair harmonic tables (N^2 (2 lmax + 1)^2 values, 1.4 GB) applied by FP32 MFMA "
                                 "GEMMs") + "; the dense c64 store (38.65 GB, helicity-split) is measured in "
                                "stored_dense_c64") if tables else
                               "dense c64 store (38.65 GB, helicity-split), streaming matvec",
                   "rhs": ncols, "seed": args.seed, "preconditioner": args.pc, "restart": args.restart},
        "roofline": roof,
        "rooflines_inner_and_residual": {k: roofs[k] for k in ("gemm_f64", "gemm", "gemm_hi") if k in roofs},
        "kernels": kernels,
        "iterations": iterations,
        "refinements": head["refinements"],
        "true_relative_residual": head["true_relative_residual"],
    }
    if "assembly" in roofs:
        result["assembly_roofline"] = roofs["assembly"]
    # what the oracle checks later, while the cluster exists: cross sections of the
    # headline solution, and its FP64 and c64-precision coupling on 8 destination spheres
    rows = (N // 2, N // 2 + 8)
    gpu_check = {"x": x_mixed.cpu().numpy(), "rows": rows, "sca": [], "ext": []}
    for c in range(ncols):
        s_, e_ = cl.cross_sections(B[:, c], x_mixed[:, c])
        gpu_check["sca"].append(s_)
        gpu_check["ext"].append(e_)
    sl = slice(rows[0] * cl.modes_per_sphere, rows[1] * cl.modes_per_sphere)
    gpu_check["cx64"] = cl.coupling(x_mixed)[sl].cpu().numpy()
    gpu_check["cx16"] = cl.coupling(x_mixed, c64=True)[sl].cpu().numpy()

    # iterations of the headline's mixed solve with each preconditioner (outside the
    # timed steps): none, block-Jacobi, the near-field form without the network (alpha
    # = 1, an MLP with a zero last layer) and with the trained network
    if args.pc == "learned":
        pcs = {}
        for name, pc, net in (("none", "none", None), ("block_jacobi", "block_jacobi", None),
                              ("near_field_alpha1", "learned", learned.MLP(seed=1)), ("learned", "learned", mlp)):
            cl.reset()
            if net is not None:
                learned.attach(cl, net)
            torch.cuda.synchronize()
            tp0 = time.perf_counter()
            _, cv = cl.solve(B, rtol=args.rtol, restart=args.restart, max_iterations=2000,
                             mixed_precision="matrix_free" if args.operator == "tables" else True, preconditioner=pc,
                             convergence="full")
            torch.cuda.synchronize()
            pcs[name] = {"iterations": max(c.iterations for c in cv), "refinements": max(c.refinements for c in cv),
                         "ms": 1e3 * (time.perf_counter() - tp0)}
        learned.attach(cl, mlp)
        result["preconditioner_iterations"] = pcs
        result["preconditioner_iterations_note"] = (
            "one mixed-precision job each, same operator, tolerance and restart; ms includes the "
            "preconditioner's set-up (block-Jacobi factors, or near pairs, features and network)")
        log(f"preconditioner iterations: {pcs}")

    # near fields of the headline solution: 128 x 128 points on the plane y = 0 across
    # the cluster, 4 columns (one thread per point, N M waves each)
    g = np.linspace(-10.0, 10.0, 128)
    pts = torch.tensor([[u, 0.0, w] for u in g for w in g], dtype=torch.float64)
    cl.field(x_mixed, pts[:16])  # warm
    torch.cuda.synchronize()
    tf0 = time.perf_counter()
    E = cl.field(x_mixed, pts)
    torch.cuda.synchronize()
    tfield = time.perf_counter() - tf0
    waves = pts.shape[0] * N * cl.modes_per_sphere
    result["fields"] = {"points": int(pts.shape[0]), "columns": ncols, "ms": 1e3 * tfield,
                        "waves_per_s": waves / tfield, "finite": bool(torch.isfinite(E).all()),
                        "note": "outgoing-wave field of the headline solution on a 128 x 128 grid at y = 0"}
    log(f"fields: {pts.shape[0]} points x {N} spheres x {cl.modes_per_sphere} modes x {ncols} columns in "
        f"{1e3 * tfield:.1f} ms")

    ref_iterations = None
    x_dense = None
    if not args.no_variants:
        other = "stored" if tables else "tables"
        xv, convv, var, vroofs = measure(other, "mixed-" + other)
        var["rooflines"] = vroofs
        if "assembly" in vroofs:
            result["assembly_roofline"] = vroofs["assembly"]
        result["stored_dense_c64" if tables else "pair_tables_c64"] = var
        x_dense = xv
    if not args.no_plain:
        plains = {}
        for stored, key in ((True, "plain_fp64_block_jacobi"), (False, "plain_fp64_block_jacobi_matrix_free")):
            cl.release()
            # timed as the headline: the same warm-up and step counts
            x64, conv64, el64, k64 = run_method(cl, B, torch, mixed=False, pc="block_jacobi", steps=args.steps,
                                                warmup=args.warmup, rtol=args.rtol, restart=args.restart,
                                                label="plain-fp64-bj" + ("" if stored else "-mf"), stored=stored)
            plains[key] = ncols * args.steps / el64
            result[key] = {"value": ncols * args.steps / el64, "unit": "solves/s", "seconds": el64,
                           "ms_per_step": 1e3 * el64 / args.steps, "steps": args.steps,
                           "iterations": [c.iterations for c in conv64], "kernels": k64,
                           "operator": "stored c128 coupling" if stored else "FP64 matrix-free (pair-table GEMMs)"}
        plain = max(plains.values())
        result["speedup_vs_plain_fp64_block_jacobi"] = value / plain
        result["speedup_note"] = "against the faster of the two plain FP64 block-Jacobi variants"
        # cross sections of the mixed solution against the FP64 solution
        errs = []
        torch.cuda.synchronize()
        tx = time.perf_counter()
        for c in range(ncols):
            s1, e1 = cl.cross_sections(B[:, c], x_mixed[:, c])
            s2, e2 = cl.cross_sections(B[:, c], x64[:, c])
            errs.append(max(abs(s1 - s2) / abs(s2), abs(e1 - e2) / abs(e2)))
        result["cross_sections_ms_per_column"] = 1e3 * (time.perf_counter() - tx) / (2 * ncols)
        result["cross_section_rel_error_vs_fp64"] = max(errs)
        result["solution_rel_error_vs_fp64"] = float(torch.linalg.norm(x_mixed - x64) / torch.linalg.norm(x64))
        if x_dense is not None:
            errs = []
            for c in range(ncols):
                s1, e1 = cl.cross_sections(B[:, c], x_dense[:, c])
                s2, e2 = cl.cross_sections(B[:, c], x64[:, c])
                errs.append(max(abs(s1 - s2) / abs(s2), abs(e1 - e2) / abs(e2)))
            result["stored_dense_c64" if tables else "pair_tables_c64"]["cross_section_rel_error_vs_fp64"] = max(errs)
        cl.release()
    if not args.no_reference:
        # the reference algorithm on the GPU: unpreconditioned matrix-free FP64 GMRES
        t0 = time.perf_counter()
        xr, convr = cl.solve(B, rtol=args.rtol, restart=args.restart, max_iterations=2000,
                             convergence="full")
        torch.cuda.synchronize()
        elr = time.perf_counter() - t0
        ref_iterations = max(c.iterations for c in convr)
        result["reference_algorithm_gpu"] = {"value": ncols / elr, "unit": "solves/s", "seconds": elr,
                                             "iterations": [c.iterations for c in convr],
                                             "note": "unpreconditioned FP64 matrix-free GMRES, as the reference"}
        log(f"reference algorithm on GPU: {elr:.2f} s, iterations {[c.iterations for c in convr]}")
    if not args.no_lu:
        del cl
        result["dense_lu"] = dense_lu(pkg, torch, args.seed)
        result["interacting_tmatrix"] = interacting_tmatrix(pkg, torch)
        result["reference_illumination"] = reference_illumination(pkg, torch)
        result["gradient_lu"] = gradient_lu(pkg, torch, args.seed)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        result["cpu_baseline"] = cpu_baseline(lmax, k0, radii, eps, pos, B.cpu().numpy(), gpu_check, args.rtol,
                                              args.restart, oracle)
        result["oracle_check"] = result["cpu_baseline"].pop("oracle_check")
        log(f"cpu baseline: {result['cpu_baseline']['value']:.4g} solves/s")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
