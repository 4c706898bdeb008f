"""Benchmark: time to 1e-8 relative residual of the 8M-cell pressure solve.

Workload (one step = one full solve, device-resident inputs):
  200^3 hex grid of the unit cube (8,000,000 cells, 55,760,000 nonzeros),
  two-phase pressure-Poisson matrix (rho = 1 in a seeded sphere of radius
  0.25, 1000 outside, harmonic-averaged faces), FP64 CSR with int32 indices,
  seeded mean-free U(-1,1) RHS, x0 = 0, Jacobi-preconditioned CG stopping at
  ||b - A x||_2 / ||b||_2 < 1e-8.

Prints ONE JSON line.  `roofline` is the SpMV kernel (the dominant one):
algorithmic bytes per launch 12*nnz + 20*n + 4 over the mean launch duration
measured with HIP events on the library's stream during the timed solves.
`cpu_baseline` is the oracle (the C restatement of Native_Mod, 1 thread)
timed on a bounded number of CG iterations of the same system and scaled to
the GPU's iteration count.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "t-flows_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _time_oracle(s, iters, omp):
    from oracle import oracle

    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    x = np.zeros(s.n)
    t0 = time.perf_counter()
    oracle.solve("cg", "jacobi", A, x, s.b.copy(), iters, 1e-30, criterion="relres", omp=omp)
    return time.perf_counter() - t0


def cpu_baseline(s, gpu_iters, seconds_budget=15.0):
    """Oracle CG (FP64) on the same system for a bounded number of iterations,
    1 thread (the reference's default OMP=no build) and all host cores (its
    OMP=yes analogue); returns the projected time to the GPU's iteration
    count."""
    probe = _time_oracle(s, 2, False) / 2.0
    iters = int(max(3, min(200, seconds_budget / max(probe, 1e-6))))
    dt = _time_oracle(s, iters, False)
    per_iter = dt / iters
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    probe_omp = _time_oracle(s, 2, True) / 2.0
    iters_omp = int(max(3, min(400, seconds_budget / 2 / max(probe_omp, 1e-6))))
    dt_omp = _time_oracle(s, iters_omp, True)
    return {
        "value": per_iter * gpu_iters,
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle Jacobi-CG (C restatement of Native_Mod Cg.f90, FP64, 1 thread) on the "
                   f"same 200^3 system for {iters} iterations ({dt:.1f} s incl. normalisation "
                   f"and initial residual), {per_iter * 1e3:.1f} ms/iteration scaled to the GPU's "
                   f"{gpu_iters} iterations"),
        "ms_per_iteration": per_iter * 1e3,
        "all_cores": {
            "value": dt_omp / iters_omp * gpu_iters, "cores": cores,
            "sample": f"same oracle built with OpenMP, {iters_omp} iterations in {dt_omp:.1f} s",
            "ms_per_iteration": dt_omp / iters_omp * 1e3,
        },
    }


TRAFFIC = os.path.join(ROOT, "profiles", "spmv_traffic_twophase200.json")


def traffic_per_launch(kernel_prefix):
    """HBM bytes per launch of the kernel, from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by its copy-kernel calibration)."""
    if not os.path.exists(TRAFFIC):
        return None, None
    with open(TRAFFIC) as f:
        t = json.load(f)
    k = t["kernels"].get(kernel_prefix)
    if not k:
        return None, None
    return k["fetch_kib"] * 1024 * t["fetch_scale"] + k["write_kib"] * 1024 * t["write_scale"], TRAFFIC


def calibrate():
    """Known-traffic copy (1 GiB read + 1 GiB write) for counter calibration."""
    import tflows_gpu as G

    nbytes = 1 << 30
    src, dst = G.DeviceArray(nbytes, np.uint8), G.DeviceArray(nbytes, np.uint8)
    ms = G.stream_copy(dst, src, nbytes, reps=5)
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=200, help="grid side (default 200 -> 8M cells)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the known-traffic copy kernel (counter calibration)")
    args = ap.parse_args()
    if args.calibrate:
        calibrate()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(the solver is single-GPU; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    import tflows_cases
    import tflows_gpu as G

    N = args.n
    t0 = time.perf_counter()
    s = tflows_cases.hex_pressure(N, rho_out=1000.0)
    log(f"bench: generated 200^3-class system n={s.n} nnz={s.nnz} in {time.perf_counter() - t0:.1f} s")
    nat = G.Native(s.row, s.col, s.dia)
    info = nat.info()
    dv, db = G.DeviceArray(s.val), G.DeviceArray(s.b)
    dx = G.DeviceArray(np.zeros(s.n))
    zeros = np.zeros(s.n)

    opt = G.Options()
    opt.precision = G.FP64
    opt.criterion = G.CRIT_RELRES
    opt.timing = 1
    opt.true_residual = 1
    tol, miter = 1e-8, 50000

    def step():
        dx.upload(zeros)
        G.lib().tfg_synchronize(nat.h)
        t = time.perf_counter()
        it, res, rep = nat.solve_native_device("cg", "jacobi", dv, dx, db, miter, tol,
                                               options=opt, report=True)
        return time.perf_counter() - t, it, res, rep

    for w in range(args.warmup):
        dt, it, res, rep = step()
        log(f"bench: warmup {w}: {dt:.3f} s, {it} iterations, res {res:.3e}")
    times, reps = [], []
    for k in range(args.steps):
        dt, it, res, rep = step()
        times.append(dt)
        reps.append((it, res, rep))
        log(f"bench: step {k}: {dt:.3f} s, {it} iterations, relres {res:.3e}, "
            f"true {rep['true_relres']:.3e}, spmv {rep['spmv_ms'] * 1e3:.1f} us "
            f"({rep['spmv_samples']} samples)")
    its = [r[0] for r in reps]
    spmv_ms = float(np.mean([r[2]["spmv_ms"] for r in reps]))
    spmv_bytes = 12 * s.nnz + 20 * s.n + 4
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    total = float(np.sum(times))
    ms_per_step = total / args.steps * 1e3
    iter_bytes = spmv_bytes + 88 * s.n
    solve_gbs = iter_bytes * its[-1] / (times[-1]) / 1e9

    traffic, traffic_src = traffic_per_launch("k_spmv_pq")
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline leg (oracle, 1 thread)...")
        cpu = cpu_baseline(s, its[-1], args.cpu_seconds)

    out = {
        "metric": "time to 1e-8 rel. residual, 8M-cell pressure solve; SpMV HBM GB/s vs 8 TB/s",
        "value": total / args.steps,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"two-phase (1:1000) pressure-Poisson, {N}^3 hex grid, Jacobi-CG FP64, "
                        "relres 1e-8, device-resident",
            "n": s.n, "nnz": s.nnz, "index": "int32", "solver": "cg", "prec": "jacobi",
            "tol": tol, "iterations": its[-1], "true_relres": reps[-1][2]["true_relres"],
            "spmv_engine": G.Native.SPMV_KINDS[nat.info()["spmv_kind"]],
            "spmv_grid": nat.info()["spmv_grid"],
            "solve_GBps_algorithmic": solve_gbs, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": os.path.relpath(traffic_src, ROOT) if traffic_src else None,
            "kernel": "k_spmv_pq (FP64 CSR SpMV with fused p.q)",
            "bytes_per_launch": spmv_bytes,
            "launch_us": spmv_ms * 1e3,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
