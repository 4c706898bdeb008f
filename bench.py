"""Benchmark: time to 1e-8 relative residual of the 8M-cell pressure solve.

Workload (one step = one full solve, device-resident inputs):
  200^3 hex grid of the unit cube (8,000,000 cells, 55,760,000 nonzeros),
  two-phase pressure-Poisson matrix (rho = 1 in a seeded sphere of radius
  0.25, 1000 outside, harmonic-averaged faces), FP64 CSR with int32 indices,
  seeded mean-free U(-1,1) RHS, x0 = 0, Jacobi-preconditioned CG stopping at
  ||b - A x||_2 / ||b||_2 < 1e-8.

Prints ONE JSON line.  `roofline` is the SpMV kernel (the dominant one):
the bytes its engine must move per launch in its own matrix format (for the
symmetric-diagonal engine 8 (K + 1) n + 16 n: the diagonal and K upper
diagonals once, p read, q written) over the mean launch duration measured
with HIP events on the library's stream during the timed solves; the CSR
formula of the metric (12 nnz + 20 n + 4) is reported beside it.
`cpu_baseline` is the oracle (the C restatement of Native_Mod, 1 thread)
solving the same system to convergence, timed; `parity` compares its
iterations and solution with the GPU's.
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


def _oracle_full_solve(s, omp, progress, tol=1e-8, sum_order=0):
    """The oracle's Jacobi-CG on system s from x0 = 0 to relres tol; returns
    (seconds, niter, x)."""
    from oracle import oracle

    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    x = np.zeros(s.n)
    t0 = time.perf_counter()
    it, _ = oracle.solve("cg", "jacobi", A, x, s.b.copy(), 50000, tol, criterion="relres",
                         omp=omp, progress=progress, sum_order=sum_order)
    return time.perf_counter() - t0, it, x


FP32_TOLS = (1e-8, 5e-9, 2e-9, 1e-9, 5e-10)
KNN_FP32_TOLS = (1e-10, 5e-11, 2e-11, 1e-11, 5e-12)
# Equal error: the FP32 form's energy-norm error against the reference
# solution within 1 % of the FP64 GPU CG's (the 2-norm error is reported
# beside it; see energy_error).
SAME_ERROR_SLACK = 1.01


def fp32_leg(G, nat, s, dv, db, dx, tol, miter):
    """The FP32 refinement form (FP32 sweeps, FP64 true residual and
    solution) on the headline system at the headline tolerance and tighter
    ones; each solve timed twice, the second kept.  Returns
    {tol: (seconds, iterations, refinements, true relres, x)}."""
    runs = {}
    opt = G.Options()
    opt.precision = G.FP32_IR
    opt.criterion = G.CRIT_RELRES
    opt.true_residual = 1
    for t in FP32_TOLS:
        for rep_ in range(2):
            dx.upload(np.zeros(s.n))
            G.lib().tfg_synchronize(nat.h)
            t0 = time.perf_counter()
            it, res, rep = nat.solve_native_device("cg", "jacobi", dv, dx, db, miter, t,
                                                   options=opt, report=True)
            dt = time.perf_counter() - t0
        runs[t] = (dt, it, rep["refinements"], rep["true_relres"], dx.download(), rep["deflated"])
        log(f"bench: FP32 form tol {t:.0e}: {dt:.3f} s, {it} FP32 iterations, "
            f"{rep['refinements']} refinements, true relres {rep['true_relres']:.2e}, "
            f"deflated {rep['deflated']}")
    return runs


def energy_error(A, x, x_ref):
    """||x - x_ref||_A / ||x_ref||_A: the error CG minimises.  On these
    nearly singular pressure systems the 2-norm error is dominated by the
    near-null modes, which a relative residual pins only weakly (two FP64
    reference solves at relres 1e-12 of the 200^3 system differ by 5e-7 in
    the 2-norm); the energy norm weighs those modes by their tiny
    eigenvalues, so it compares solvers and not the reference's own noise."""
    e = x - x_ref
    return float(np.sqrt(max(e @ (A @ e), 0.0) / (x_ref @ (A @ x_ref))))


def fp32_summary(runs, x64, x_ref, t64, tol, A):
    """Errors of the FP64 CG and of the FP32 form against the reference
    solution (2-norm and energy norm), and the FP32 form's speed-up at the
    loosest tolerance whose energy-norm error is no worse than FP64 CG's at
    the headline one."""
    def rel(a):
        return float(np.linalg.norm(a - x_ref) / np.linalg.norm(x_ref))

    err64, en64 = rel(x64), energy_error(A, x64, x_ref)
    table = {}
    for t, v in runs.items():
        en = energy_error(A, v[4], x_ref)
        table[f"{t:.0e}"] = {"seconds": v[0], "iterations": v[1], "refinements": v[2],
                             "true_relres": v[3], "x_rel_err_vs_reference": rel(v[4]),
                             "energy_err_vs_reference": en,
                             "energy_error_ratio_vs_fp64_cg": en / en64,
                             "speedup_vs_fp64_cg": t64 / v[0],
                             "deflated": bool(v[5])}
    ok = [t for t in FP32_TOLS if table[f"{t:.0e}"]["energy_err_vs_reference"] <= SAME_ERROR_SLACK * en64]
    pick = ok[0] if ok else None
    out = {"fp64_cg_x_rel_err_vs_reference": err64, "fp64_cg_energy_err_vs_reference": en64,
           "fp64_cg_seconds": t64, "runs": table,
           "same_error_rule": f"energy-norm error <= {SAME_ERROR_SLACK} x the FP64 GPU CG's",
           "tolerance_at_fp64_error": pick,
           "speedup_at_same_error": (t64 / runs[pick][0]) if pick is not None else None,
           "speedup_at_same_tolerance": t64 / runs[tol][0]}
    log(f"bench: FP32 form: FP64 CG error {err64:.2e} (energy {en64:.2e}); "
        + "; ".join(f"tol {k}: {v['seconds']:.3f} s err {v['x_rel_err_vs_reference']:.2e} "
                    f"energy {v['energy_err_vs_reference']:.2e}" for k, v in table.items())
        + (f"; same energy error at tol {pick:.0e}: {t64 / runs[pick][0]:.2f}x" if pick else ""))
    return out


def cpu_baseline(s, x_gpu, gpu_iters):
    """The oracle (the C restatement of Native_Mod's Cg.f90) solves the same
    system to convergence on the host, timed once each: built with OpenMP on
    every host core (the reference's OMP=yes analogue), then single-threaded
    (its default OMP=no build, the reported value).  Both solutions are
    compared with the GPU's: the 1-thread oracle is the parity reference
    (iterations, relative x difference).  The OpenMP build, whose sums run
    in another order, run again with the normalisation factor fn one ulp
    below and above the reference's (the other sum of the diagonal it comes
    from), gives the ensemble of equally valid runs whose largest
    difference is the parity envelope at this size."""
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    log(f"bench: CPU baseline: oracle Jacobi-CG to relres 1e-8 on {cores} threads (OpenMP build)")
    t_omp, it_omp, x_omp = _oracle_full_solve(s, True, 500)
    log(f"bench: CPU baseline: OpenMP oracle {t_omp:.1f} s, {it_omp} iterations")
    log("bench: reference solution: OpenMP oracle Jacobi-CG to relres 1e-12")
    t_ref, it_ref, x_ref = _oracle_full_solve(s, True, 500, tol=1e-12)
    log(f"bench: reference solution: {t_ref:.1f} s, {it_ref} iterations")
    ens = [(it_omp, x_omp)]
    for so in (7, 8):
        log(f"bench: parity ensemble: OpenMP oracle with fn one ulp {'below' if so == 7 else 'above'}")
        _, it_e, x_e = _oracle_full_solve(s, True, 500, sum_order=so)
        ens.append((it_e, x_e))
    log("bench: CPU baseline: oracle Jacobi-CG to relres 1e-8, 1 thread")
    t1, it1, x1 = _oracle_full_solve(s, False, 250)
    log(f"bench: CPU baseline: 1-thread oracle {t1:.1f} s, {it1} iterations")

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    runs = [(it1, x1)] + ens
    err = rel(x_gpu, x1)
    env = max(rel(a[1], b[1]) for a in runs for b in runs if a is not b)
    spread = max(abs(a[0] - b[0]) for a in runs for b in runs)
    cpu = {
        "value": t1, "unit": "s", "cores": 1, "kind": "port",
        "sample": (f"the whole solve, measured: oracle Jacobi-CG (C restatement of Native_Mod "
                   f"Cg.f90, FP64, 1 thread) on the same 200^3 system from x0 = 0 to relres "
                   f"1e-8, {it1} iterations"),
        "iterations": it1, "ms_per_iteration": t1 / max(it1, 1) * 1e3,
        "all_cores": {"value": t_omp, "cores": cores, "iterations": it_omp,
                      "sample": "the same whole solve, oracle built with OpenMP"},
    }
    parity = {
        "reference": "1-thread oracle (C restatement of Cg.f90), same system and inputs",
        "oracle_iterations": it1, "gpu_iterations": gpu_iters,
        "iterations_diff": gpu_iters - it1,
        "x_rel_err_gpu_vs_oracle": err,
        "oracle_openmp_iterations": it_omp,
        "ensemble": "serial oracle, OpenMP oracle, OpenMP oracle with fn -1 / +1 ulp",
        "ensemble_iterations": [r[0] for r in runs],
        "iterations_spread": spread,
        "x_rel_envelope": env,
        "x_rel_err_gpu_vs_oracle_over_envelope": err / env if env > 0 else None,
        # the parity rule of the tests (tests/parity_bar.py): iterations
        # within max(2, 1.2 x the spread), x within max(1e-10, 1.2 x the
        # envelope) of equally valid runs
        "bar": "iterations within max(2, 1.2 x spread); x within max(1e-10, 1.2 x envelope)",
        "within_bar": bool(abs(gpu_iters - it1) <= max(2, 1.2 * spread)
                           and err <= max(1e-10, 1.2 * env)),
        "reference_solution": {"solver": "OpenMP oracle Jacobi-CG", "relres": 1e-12,
                               "iterations": it_ref},
        "x_rel_err_gpu_vs_reference": rel(x_gpu, x_ref),
        "x_rel_err_oracle_vs_reference": rel(x1, x_ref),
    }
    log(f"bench: parity: gpu {gpu_iters} vs oracle {it1} iterations (ensemble "
        f"{[r[0] for r in runs]}); x rel err {err:.2e} (envelope {env:.2e})")
    return cpu, parity, x_ref


TRAFFIC = os.path.join(ROOT, "profiles", "spmv_traffic_twophase200.json")


def format_bytes(kind, nnz, n, sym_diagonals=0):
    """Bytes one SpMV must move in the matrix format the engine streams (the
    CSR formula counts int32 columns and row pointers; the coded engine
    streams a 1-byte offset code per nonzero instead of the column, the
    row-pattern engine a 1-byte pattern id per row instead of columns and row
    pointers, the symmetric engine the diagonal and its K upper diagonals,
    zero-padded: 8 (K + 1) n).  All add 8 B per row of p read and 8 B of q
    written."""
    if kind == 8:
        return 8 * (sym_diagonals + 1) * n + 16 * n
    if kind == 7:
        return 8 * nnz + 17 * n
    if kind == 6:
        return 9 * nnz + 20 * n + 4
    return 12 * nnz + 20 * n + 4


FORMAT_FORMULA = {8: "8 (K + 1) n + 16 n (diagonal and K upper diagonals, p read, q written)",
                  7: "8 nnz + 17 n", 6: "9 nnz + 20 n + 4", 5: "12 nnz + 20 n + 4",
                  0: "12 nnz + 20 n + 4", 1: "12 nnz + 20 n + 4", 2: "12 nnz + 20 n + 4",
                  3: "12 nnz + 20 n + 4", 4: "12 nnz + 20 n + 4"}


def spmv_roofline(kind, nnz, n, sym_diagonals, launch_ms, traffic=None):
    """Roofline of one SpMV launch: the bytes the engine that ran must move
    in its own matrix format (the symmetric engine stores the values once)
    over the launch duration, against the HBM peak; the same launch priced by
    the metric's CSR formula (12 nnz + 20 n + 4) beside it, which can exceed
    the peak because the engine's format moves fewer bytes than CSR."""
    fmt = format_bytes(kind, nnz, n, sym_diagonals)
    csr = 12 * nnz + 20 * n + 4
    sec = launch_ms * 1e-3
    return {
        "bound": "hbm", "achieved": fmt / sec / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": fmt / sec / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
        "bytes_per_launch": fmt, "bytes_formula": FORMAT_FORMULA[kind],
        "launch_us": launch_ms * 1e3,
        "traffic_GBps": traffic / sec / 1e9 if traffic else None,
        "traffic_frac": traffic / sec / 1e9 / HBM_PEAK_GBS if traffic else None,
        "csr_equivalent": {"bytes_per_launch": csr, "GBps": csr / sec / 1e9,
                           "frac": csr / sec / 1e9 / HBM_PEAK_GBS},
    }


def practical_floor(G, read_bytes, write_bytes, launch_ms):
    """The launch against what this GPU's HBM reaches in practice: a 1 GiB
    read sweep and a 1 GiB write sweep (tfg_stream_bench, 16-byte accesses,
    mean of 5 after a warm-up) give the ceilings; the floor is the engine's
    format bytes at them (reads at the read rate, writes at the write
    rate)."""
    nbytes = 1 << 30
    a, b = G.DeviceArray(nbytes, np.uint8), G.DeviceArray(nbytes, np.uint8)
    rd = nbytes / (G.stream_bench(b, a, nbytes, 1, reps=5) * 1e-3) / 1e9
    wr = nbytes / (G.stream_bench(b, None, nbytes, 2, reps=5) * 1e-3) / 1e9
    floor_us = (read_bytes / rd + write_bytes / wr) / 1e3
    return {"read_GBps": rd, "write_GBps": wr, "format_read_bytes": read_bytes,
            "format_write_bytes": write_bytes, "floor_us": floor_us,
            "frac": floor_us / (launch_ms * 1e3)}


# SpMV engine -> text of its template argument in the kernel's full name
ENGINE_TAG = {0: "StreamEngine<2040, false, double, false>", 1: "StreamEngine<4088,",
              2: "WaveEngine<4,", 3: "WaveEngine<8,", 4: "WaveEngine<16,",
              5: "StreamEngine<2040, true, double, false>",
              6: "StreamEngine<2040, true, double, true>", 7: "PatternEngine<2040,",
              8: "k_spmv_pq_sym"}


def traffic_per_launch(kernel_prefix, kind=None):
    """HBM bytes per launch of the kernel, from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by its copy-kernel calibration).
    None when the summary was measured on another SpMV engine than `kind`."""
    if not os.path.exists(TRAFFIC):
        return None, None
    with open(TRAFFIC) as f:
        t = json.load(f)
    k = t["kernels"].get(kernel_prefix)
    if not k:
        return None, None
    if kind is not None and ENGINE_TAG[kind] not in k.get("name", ""):
        return None, None
    return k["fetch_kib"] * 1024 * t["fetch_scale"] + k["write_kib"] * 1024 * t["write_scale"], TRAFFIC


def _kernel_key(name):
    """Short kernel name: 'void tfg::k_spmv_pq<...>(...)' -> 'k_spmv_pq'."""
    s = name.split("tfg::", 1)[-1] if "tfg::" in name else name
    for c in "<(":
        s = s.split(c, 1)[0]
    return s.strip()


def _counter_means(path, counter):
    """Mean value per dispatch of one rocprofv3 counter, per kernel, over the
    dispatches that did work (launches after convergence return at once)."""
    import csv

    rows, names = {}, {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            key = _kernel_key(r["Kernel_Name"])
            rows.setdefault(key, []).append(float(r["Counter_Value"]))
            names.setdefault(key, r["Kernel_Name"])
    out = {}
    for k, v in rows.items():
        v = np.asarray(v)
        busy = v[v >= 0.5 * v.max()] if v.max() > 0 else v
        out[k] = {"mean_kib": float(busy.mean()), "dispatches": int(busy.size),
                  "idle_dispatches": int(v.size - busy.size), "name": names[k]}
    return out


def summarize_counters(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, out_path, commands):
    """Write the HBM-traffic summary bench.py reads: FETCH_SIZE / WRITE_SIZE
    (KiB per dispatch) of every kernel of a solve, and the factors that map
    each counter onto true bytes, from the copy kernel of known size (1 GiB
    read + 1 GiB written per launch)."""
    fetch, write = _counter_means(fetch_csv, "FETCH_SIZE"), _counter_means(write_csv, "WRITE_SIZE")
    cf = _counter_means(cal_fetch_csv, "FETCH_SIZE")["k_copy16"]["mean_kib"]
    cw = _counter_means(cal_write_csv, "WRITE_SIZE")["k_copy16"]["mean_kib"]
    gib_kib = (1 << 30) / 1024.0
    doc = {
        "what": "HBM traffic per dispatch from rocprofv3 counters, 200^3 two-phase Jacobi-CG solve",
        "commands": commands,
        "fetch_scale": gib_kib / cf, "write_scale": gib_kib / cw,
        "calibration": {"kernel": "k_copy16", "bytes_read": 1 << 30, "bytes_written": 1 << 30,
                        "fetch_kib": cf, "write_kib": cw},
        "kernels": {k: {"fetch_kib": fetch[k]["mean_kib"],
                        "write_kib": write.get(k, {}).get("mean_kib", 0.0),
                        "dispatches": fetch[k]["dispatches"],
                        "name": fetch[k]["name"][:160]} for k in sorted(fetch)},
    }
    with open(out_path, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print(json.dumps({k: round((v["fetch_kib"] * doc["fetch_scale"] + v["write_kib"] * doc["write_scale"])
                               * 1024 / 1e6, 1) for k, v in doc["kernels"].items()}), flush=True)


def summarize_stats(stats_csv, out_path, command):
    """Markdown table of a rocprofv3 --stats kernel summary."""
    import csv

    with open(stats_csv) as f:
        rows = list(csv.DictReader(f))
    lines = [f"# Kernel statistics: {command['workload']}", "", "Command:", "", "```",
             command["cmd"], "```", "",
             "| kernel | calls | total ms | avg us | min us | max us | % |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        lines.append(f"| `{r['Name'][:110]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.1f} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                     f"{float(r['MaxNs']) / 1e3:.1f} | {r['Percentage']} |")
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _oracle_ms_per_iter(s, solver, prec, iters, criterion="relres"):
    """Oracle time per iteration (1 thread) of a solver on system s."""
    from oracle import oracle

    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    x = np.zeros(s.n)
    t0 = time.perf_counter()
    try:
        oracle.solve(solver, prec, A, x, s.b.copy(), iters, 1e-300, criterion=criterion)
    except ArithmeticError:
        pass
    return (time.perf_counter() - t0) / iters * 1e3


def bench_upwind(args):
    """Non-symmetric workload: 120^3 first-order upwind convection-diffusion
    (1.73M cells), Jacobi, ||r||/||b|| < 1e-8, with the reference's 'bicg'
    (the CG recurrence), textbook BiCG, CGS and BiCGSTAB."""
    import tflows_cases
    import tflows_gpu as G

    N = 120 if args.n == 200 else args.n
    s = tflows_cases.upwind(N)
    nat = G.Native(s.row, s.col, s.dia)
    dv, db = G.DeviceArray(s.val), G.DeviceArray(s.b)
    dx = G.DeviceArray(np.zeros(s.n))
    runs = {}
    spmv_bytes = 12 * s.nnz + 20 * s.n + 4
    # reference solution: FP64 BiCGSTAB to relres 1e-13
    ro = G.Options()
    ro.criterion = G.CRIT_RELRES
    dx.upload(np.zeros(s.n))
    nat.solve_native_device("bcgs", "jacobi", dv, dx, db, 5000, 1e-13, options=ro)
    x_ref = dx.download()
    for name, solver, textbook, prec32 in (("reference_bicg", "bicg", 0, 0), ("bicg", "bicg", 1, 0),
                                           ("cgs", "cgs", 0, 0), ("bicgstab", "bcgs", 0, 0),
                                           ("bicg_fp32_ir", "bicg", 1, 1),
                                           ("cgs_fp32_ir", "cgs", 0, 1),
                                           ("bicgstab_fp32_ir", "bcgs", 0, 1)):
        opt = G.Options()
        opt.criterion = G.CRIT_RELRES
        opt.timing = 1
        opt.true_residual = 1
        opt.textbook_bicg = textbook
        opt.precision = G.FP32_IR if prec32 else G.FP64
        best = None
        for _ in range(args.steps + args.warmup):
            dx.upload(np.zeros(s.n))
            t = time.perf_counter()
            try:
                it, res, rep = nat.solve_native_device(solver, "jacobi", dv, dx, db, 2000, 1e-8,
                                                       options=opt, report=True)
                status = "ok"
            except G.TFlowsGpuError as e:
                it, res, rep, status = -1, float("nan"), {"true_relres": float("nan"),
                                                          "spmv_ms": 0.0}, str(e)
            dt = time.perf_counter() - t
            if best is None or dt < best[0]:
                best = (dt, it, res, rep, status, dx.download())
        dt, it, res, rep, status, xs = best
        runs[name] = {"seconds": dt, "iterations": it, "relres": res,
                      "true_relres": rep["true_relres"], "status": status,
                      "converged": bool(it >= 0 and res < 1e-8),
                      "x_rel_err_vs_reference": float(np.linalg.norm(xs - x_ref) /
                                                      np.linalg.norm(x_ref)),
                      "spmv_GBps": spmv_bytes / (rep["spmv_ms"] * 1e-3) / 1e9 if rep["spmv_ms"] else None}
        log(f"bench upwind{N}: {name}: {dt:.3f} s, {it} iterations, relres {res:.2e}, "
            f"true {rep['true_relres']:.2e} ({status})")
    cpu = None
    for m in ("bicg", "cgs", "bicgstab"):
        r32, r64 = runs[f"{m}_fp32_ir"], runs[m]
        r32["speedup_vs_fp64"] = r64["seconds"] / r32["seconds"]
        r32["error_ratio_vs_fp64"] = r32["x_rel_err_vs_reference"] / max(r64["x_rel_err_vs_reference"], 1e-300)
    if not args.no_cpu:
        cpu = {k: _oracle_ms_per_iter(s, sv, "jacobi", 20) for k, sv in
               (("reference_bicg", "bicg"), ("bicg", "bicg_textbook"), ("cgs", "cgs"),
                ("bicgstab", "bcgs"))}
    best_name = min((k for k in runs if runs[k]["converged"]), key=lambda k: runs[k]["seconds"])
    out = {
        "metric": f"time to 1e-8 rel. residual, {N}^3 upwind convection-diffusion (non-symmetric)",
        "value": runs[best_name]["seconds"], "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"upwind convection-diffusion {N}^3, Jacobi, relres 1e-8",
                   "n": s.n, "nnz": s.nnz, "fastest": best_name},
        "runs": runs,
        "cpu_baseline": None if cpu is None else {
            "unit": "ms per iteration", "cores": 1, "kind": "port",
            "sample": "oracle (C restatement), 20 iterations of each method, 1 thread",
            "ms_per_iteration": cpu},
    }
    print(json.dumps(out), flush=True)


def bench_knn_ir(args):
    """Unstructured workload: 4M-point kNN Laplacian (jittered 200x200x100
    lattice, 10..18 neighbours, random permutation), Jacobi CG to relres
    1e-10 in FP64, and FP32 sweeps (restart every 50) inside FP64
    refinement; error of each against an FP64 CG solution at relres 1e-13."""
    import tflows_cases
    import tflows_gpu as G

    t0 = time.perf_counter()
    s = tflows_cases.knn(200, 200, 100)
    log(f"bench knn: n={s.n} nnz={s.nnz} generated in {time.perf_counter() - t0:.1f} s")
    Acsr = s.scipy().tocsr()
    nat = G.Native(s.row, s.col, s.dia)
    dv, db = G.DeviceArray(s.val), G.DeviceArray(s.b)
    dx = G.DeviceArray(np.zeros(s.n))

    def run(precision, tol, inner_max=0, miter=100000):
        opt = G.Options()
        opt.precision = precision
        opt.criterion = G.CRIT_RELRES
        opt.true_residual = 1
        opt.inner_max = inner_max
        best = None
        for _ in range(args.steps + args.warmup):
            dx.upload(np.zeros(s.n))
            t = time.perf_counter()
            it, res, rep = nat.solve_native_device("cg", "jacobi", dv, dx, db, miter, tol,
                                                   options=opt, report=True)
            log(f"bench knn: precision {precision} tol {tol} inner_max {inner_max}: "
                f"{time.perf_counter() - t:.3f} s, {it} iterations, relres {res:.2e}")
            dt = time.perf_counter() - t
            if best is None or dt < best[0]:
                best = (dt, it, res, rep, dx.download())
        return best

    t = time.perf_counter()
    nat.reorder(G.ORDER_RCM)
    t_order = time.perf_counter() - t
    ref = run(G.FP64, 1e-13)
    out_runs = {}
    # FP64 CG with the internal RCM ordering; FP32 sweeps in FP64 refinement
    # on it (refinement whenever the inner residual fell by 1e-3, at 1e-10
    # and tighter, to find the tolerance of equal error, and the
    # restart-every-50 form of the configuration, capped at 20000 inner
    # iterations); then FP64 CG in the caller's random numbering
    legs = [("fp64_cg_rcm", G.FP64, 1e-10, 0, G.ORDER_RCM)]
    legs += [(f"fp32_ir_rcm_{t:.0e}", G.FP32_IR, t, 0, G.ORDER_RCM) for t in KNN_FP32_TOLS]
    legs += [("fp32_ir_restart50_rcm", G.FP32_IR, 1e-10, 50, G.ORDER_RCM),
             ("fp64_cg", G.FP64, 1e-10, 0, G.ORDER_NATURAL)]
    for name, prec, tol, im, order in legs:
        if order == G.ORDER_NATURAL:
            nat.reorder(order)
        dt, it, res, rep, x = run(prec, tol, im, 100000 if prec == G.FP64 else 20000)
        err = float(np.linalg.norm(x - ref[4]) / np.linalg.norm(ref[4]))
        out_runs[name] = {"seconds": dt, "tol": tol, "iterations": it,
                          "refinements": rep["refinements"],
                          "relres": res, "true_relres": rep["true_relres"],
                          "converged": bool(rep["true_relres"] < 1e-10),
                          "error_vs_fp64_solution": err,
                          "energy_error_vs_fp64_solution": energy_error(Acsr, x, ref[4])}
        log(f"bench knn: {name} tol {tol}: {dt:.3f} s, {it} iterations, true {rep['true_relres']:.2e}, "
            f"error {err:.2e}")
    base = out_runs["fp64_cg_rcm"]
    for v in out_runs.values():
        v["speedup_vs_fp64_cg_rcm"] = base["seconds"] / v["seconds"]
        v["error_ratio_vs_fp64_cg_rcm"] = v["error_vs_fp64_solution"] / base["error_vs_fp64_solution"]
        v["energy_error_ratio_vs_fp64_cg_rcm"] = (v["energy_error_vs_fp64_solution"] /
                                                  base["energy_error_vs_fp64_solution"])
    # equal error: the loosest FP32 tolerance whose 2-norm and energy-norm
    # errors are both within 1 % of FP64 CG's at 1e-10 (the FP32 form stops
    # with more error in the smooth modes at one tolerance)
    same = [f"fp32_ir_rcm_{t:.0e}" for t in KNN_FP32_TOLS
            if max(out_runs[f"fp32_ir_rcm_{t:.0e}"]["error_ratio_vs_fp64_cg_rcm"],
                   out_runs[f"fp32_ir_rcm_{t:.0e}"]["energy_error_ratio_vs_fp64_cg_rcm"])
            <= SAME_ERROR_SLACK]
    same_name = same[0] if same else None
    # the value is the fastest form that reached 1e-10 (true residual)
    ok = {k: v for k, v in out_runs.items() if v["converged"]}
    best_name = min(ok, key=lambda k: ok[k]["seconds"]) if ok else "fp64_cg"
    out = {
        "metric": "time to 1e-10 rel. residual, 4M-point kNN Laplacian, FP32 sweeps in FP64 refinement vs FP64 CG",
        "value": out_runs[best_name]["seconds"], "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": False,
        "dtype": "f64" if best_name.startswith("fp64") else "f32+f64", "data": "synthetic",
        "best": best_name,
        "config": {"workload": "kNN 200x200x100 jittered lattice, k 10..18, permuted; Jacobi; relres 1e-10",
                   "n": s.n, "nnz": s.nnz},
        "rcm_speedup": out_runs["fp64_cg"]["seconds"] / out_runs["fp64_cg_rcm"]["seconds"],
        "ordering_s": t_order,
        "reference_solution": {"solver": "FP64 GPU CG", "tol": 1e-13, "iterations": ref[1],
                               "seconds": ref[0]},
        "fp32_speedup_at_same_tolerance": out_runs["fp32_ir_rcm_1e-10"]["speedup_vs_fp64_cg_rcm"],
        "same_error_rule": f"2-norm and energy-norm errors <= {SAME_ERROR_SLACK} x FP64 CG's at 1e-10",
        "fp32_at_same_error": same_name,
        "fp32_speedup_at_same_error": (out_runs[same_name]["speedup_vs_fp64_cg_rcm"]
                                       if same_name else None),
        "runs": out_runs,
    }
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """Batched workload: 256 two-phase pressure systems, sides U{20..80},
    density ratios logU(1, 1000), own seeded RHS each, relres 1e-8 per
    system, solved as one batch (masks, hipGraph chunks): FP32 sweeps inside
    FP64 refinement (the configuration's form) and plain FP64 CG, against the
    256 FP64 solves one after another through the single-system call (skipped
    with --no-sequential)."""
    import tflows_cases
    import tflows_gpu as G

    count = 256
    sides, ratios = tflows_cases.batch_params(count, 20, 80, 1000.0)
    t0 = time.perf_counter()
    systems = [tflows_cases.hex_pressure(int(sd), rho_out=float(rt), seed=7000 + k)
               for k, (sd, rt) in enumerate(zip(sides, ratios))]
    total = sum(s.n for s in systems)
    log(f"bench batch: {count} systems, {total} cells, {sum(s.nnz for s in systems)} nonzeros "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    bt = G.Batch([(s.row, s.col, s.dia) for s in systems])
    val = np.concatenate([s.val for s in systems])
    b = np.concatenate([s.b for s in systems])
    dv, db = G.DeviceArray(val), G.DeviceArray(b)
    dx = G.DeviceArray(np.zeros(total))

    def batched(precision):
        opt = G.Options()
        opt.criterion = G.CRIT_RELRES
        opt.precision = precision
        best = None
        for _ in range(args.steps + args.warmup):
            dx.upload(np.zeros(total))
            t = time.perf_counter()
            niter, fin, rep = bt.solve("cg", "jacobi", dv, dx, db, 20000, 1e-8, options=opt,
                                       device=True)
            dt = time.perf_counter() - t
            if best is None or dt < best[0]:
                best = (dt, niter.copy(), fin.copy(), rep, bt.times(), dx.download())
        return best

    def relres(x):
        worst = 0.0
        for k, s in enumerate(systems):
            xs = x[bt.offsets[k]:bt.offsets[k + 1]]
            worst = max(worst, float(np.linalg.norm(s.b - s.scipy() @ xs) / np.linalg.norm(s.b)))
        return worst

    legs = {}
    for name, prec in (("fp32_ir", G.FP32_IR), ("fp64", G.FP64)):
        dt, niter, fin, rep, done_ms, x = batched(prec)
        # algorithmic bytes per iteration in CSR format: FP64 CG at
        # 12 nnz + 108 n (the single-system CG iteration); FP32 sweeps at
        # 8 nnz + 56 n (FP32 value and int32 column per nonzero; row start,
        # p gathered and q written per row; p update 24 n, r update 20 n)
        per = (lambda s: 12 * s.nnz + 108 * s.n) if prec == G.FP64 else \
              (lambda s: 8 * s.nnz + 56 * s.n)
        bbytes = float(sum(int(it) * per(s) for it, s in zip(niter, systems)))
        legs[name] = {
            "device_ms": rep["solve_ms"], "wall_s": dt,
            "solves_per_s": count / (rep["solve_ms"] * 1e-3),
            "iterations_total": int(niter.sum()), "iterations_max": int(niter.max()),
            "iterations_median": float(np.median(niter)),
            "p50_system_ms": float(np.percentile(done_ms, 50)),
            "p99_system_ms": float(np.percentile(done_ms, 99)),
            "all_converged": bool(np.all(fin < 1e-8)),
            "worst_true_relres": relres(x),
            "GBps_algorithmic": bbytes / (rep["solve_ms"] * 1e-3) / 1e9,
            "autotuner": bt.tuning(prec),
        }
        log(f"bench batch {name}: {dt:.3f} s wall, device {rep['solve_ms']:.1f} ms, "
            f"{legs[name]['solves_per_s']:.1f} solves/s, max niter {niter.max()}, p50/p99 "
            f"{legs[name]['p50_system_ms']:.0f}/{legs[name]['p99_system_ms']:.0f} ms, worst true "
            f"relres {legs[name]['worst_true_relres']:.2e}")
    seq = None
    if not args.no_sequential:
        # one after another through the single-system drop-in (device-resident)
        seq_ms, seq_it = [], []
        for k, s in enumerate(systems):
            nat = G.Native(s.row, s.col, s.dia)
            sv, sb, sx = G.DeviceArray(s.val), G.DeviceArray(s.b), G.DeviceArray(np.zeros(s.n))
            o = G.Options()
            o.criterion = G.CRIT_RELRES
            it, res, r = nat.solve_native_device("cg", "jacobi", sv, sx, sb, 20000, 1e-8, options=o,
                                                 report=True)
            seq_ms.append(r["solve_ms"])
            seq_it.append(it)
            nat.close()
        seq_ms = np.array(seq_ms)
        seq_done = np.cumsum(seq_ms)
        seq = {"device_ms": float(seq_ms.sum()),
               "solves_per_s": count / (seq_ms.sum() * 1e-3),
               "p50_system_ms": float(np.percentile(seq_ms, 50)),
               "p99_system_ms": float(np.percentile(seq_ms, 99)),
               "p50_completion_ms": float(np.percentile(seq_done, 50)),
               "p99_completion_ms": float(np.percentile(seq_done, 99)),
               "iterations_total": int(np.sum(seq_it))}
    head = legs["fp32_ir"]
    out = {
        "metric": "batched solves/s: 256 two-phase pressure systems (sides 20..80), FP32 Jacobi CG in FP64 refinement to relres 1e-8",
        "value": head["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "256 hex pressure systems, sides U{20..80}, rho ratio logU(1,1000)",
                   "cells": int(total), "nnz": int(sum(s.nnz for s in systems)),
                   "blocks": bt.info()["blocks"]},
        "batch_fp32_ir": head, "batch_fp64": legs["fp64"], "sequential_fp64": seq,
        "fp32_vs_fp64_batch": legs["fp64"]["device_ms"] / head["device_ms"],
        "speedup_vs_sequential_fp64": (seq["device_ms"] / head["device_ms"]) if seq else None,
    }
    print(json.dumps(out), flush=True)


def bench_assembly(args):
    """Pressure solve fed on the device: 200^3 face data (fc, f, v_m, fluxes)
    assembled into A % val and b by the face kernel, against uploading the
    host-built A % val and b (what a host-assembled drop-in call copies)."""
    import tflows_cases
    import tflows_gpu as G

    N = args.n
    reps = max(3, args.steps)
    timings = {}
    for order in ("shuffled", "natural"):
        t0 = time.perf_counter()
        f = tflows_cases.hex_faces(N, shuffle=order == "shuffled")
        log(f"bench assembly: {N}^3, {f.n_faces} faces ({order} order), generated in "
            f"{time.perf_counter() - t0:.1f} s")
        nat = G.Native(f.row, f.col, f.dia)
        faces = G.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
        d = {k: G.DeviceArray(getattr(f, k)) for k in ("fc", "fw", "vm", "vflux")}
        dval, db = G.DeviceArray(np.zeros(len(f.col))), G.DeviceArray(np.zeros(f.n))
        faces.assemble(d["fc"], d["fw"], d["vm"], d["vflux"], dval, db, singular=True)
        t = time.perf_counter()
        for _ in range(reps):
            faces.assemble(d["fc"], d["fw"], d["vm"], d["vflux"], dval, db, singular=True)
        timings[order] = (time.perf_counter() - t) / reps
        log(f"bench assembly: {order} face order: {timings[order] * 1e3:.2f} ms")
    t_asm = timings["natural"]
    # Rhie-Chow fluxes of the same faces (grid order), the step before
    rf = tflows_cases.rhie_chow_fields(f, N, N, N)
    rc = G.RhieChow(rf, choi=True, gu=True)
    dflux = G.DeviceArray(f.vflux)
    faces.rhie_chow(rc, dflux)
    t = time.perf_counter()
    for _ in range(reps):
        faces.rhie_chow(rc, dflux)
    t_rc = (time.perf_counter() - t) / reps
    log(f"bench assembly: Rhie-Chow fluxes (Choi and Gu corrections) {t_rc * 1e3:.2f} ms")
    hval, hb = np.zeros(len(f.col)), np.zeros(f.n)
    hval[:] = dval.download()
    hb[:] = db.download()
    t = time.perf_counter()
    for _ in range(reps):
        dval.upload(hval)
        db.upload(hb)
    t_up = (time.perf_counter() - t) / reps
    bytes_up = hval.nbytes + hb.nbytes
    log(f"bench assembly: device assembly {t_asm * 1e3:.2f} ms, upload {t_up * 1e3:.2f} ms "
        f"({bytes_up / t_up / 1e9:.1f} GB/s)")
    print(json.dumps({
        "metric": "pressure-system assembly on the device vs upload of A % val and b, 200^3",
        "value": t_asm * 1e3, "unit": "ms", "n_gpus": 1, "steps": reps, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"hex {N}^3 face data, pressure matrix + rhs, faces in grid order",
                   "cells": f.n, "faces": f.n_faces, "nnz": len(f.col)},
        "shuffled_faces_ms": timings["shuffled"] * 1e3,
        "rhie_chow_fluxes_ms": t_rc * 1e3,
        "upload_ms": t_up * 1e3, "upload_bytes": bytes_up, "speedup_vs_upload": t_up / t_asm,
    }), flush=True)


def bench_momentum(args):
    """Momentum system of one component fed on the device: 200^3 face and
    cell data (MINMOD advection with the implicit upwind part, BDF2 inertia,
    cross diffusion, forces, under-relaxation) assembled into A % val, b and
    v_m by the face kernel, against uploading A % val and b, and against the
    oracle's serial face loops (1 thread) on the same data."""
    import tflows_cases
    import tflows_gpu as G

    N = args.n
    kw = dict(component=2, adv="minmod", blend_matrix=True, td="parabolic")
    reps = max(3, args.steps)
    timings = {}
    for order in ("shuffled", "natural"):
        t0 = time.perf_counter()
        f = tflows_cases.hex_faces(N, shuffle=order == "shuffled")
        mf = tflows_cases.momentum_fields(f)
        log(f"bench momentum: {N}^3, {f.n_faces} faces ({order} order), generated in "
            f"{time.perf_counter() - t0:.1f} s")
        nat = G.Native(f.row, f.col, f.dia)
        faces = G.Faces(nat, f.c1, f.c2, f.pos, f.diag_bc, f.b_region)
        m = G.Momentum(mf, **kw)
        dval, db, dvm = (G.DeviceArray(np.zeros(k)) for k in (len(f.col), f.n, f.n))
        faces.momentum(m, dval, db, dvm)
        t = time.perf_counter()
        for _ in range(reps):
            faces.momentum(m, dval, db, dvm)
        timings[order] = (time.perf_counter() - t) / reps
        log(f"bench momentum: {order} face order: {timings[order] * 1e3:.2f} ms")
    t_asm = timings["natural"]
    hval, hb = dval.download(), db.download()
    t = time.perf_counter()
    for _ in range(reps):
        dval.upload(hval)
        db.upload(hb)
    t_up = (time.perf_counter() - t) / reps
    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        t = time.perf_counter()
        val_o, b_o, _ = oracle.momentum_assemble(f, mf, **kw)
        t_cpu = time.perf_counter() - t
        same = bool(np.array_equal(val_o, hval) and np.array_equal(b_o, hb))
        cpu = {"value": t_cpu * 1e3, "unit": "ms", "cores": 1, "kind": "port",
               "sample": "oracle (C restatement of the face loops), the whole system once",
               "bit_identical": same}
        log(f"bench momentum: oracle {t_cpu * 1e3:.0f} ms, bit-identical {same}")
    log(f"bench momentum: device assembly {t_asm * 1e3:.2f} ms (faces in grid order), "
        f"upload {t_up * 1e3:.2f} ms")
    print(json.dumps({
        "metric": "momentum-system assembly (one component) on the device vs upload of A % val and b, 200^3",
        "value": t_asm * 1e3, "unit": "ms", "n_gpus": 1, "steps": reps, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"hex {N}^3 face data, momentum matrix + rhs + v_m, MINMOD, "
                               "implicit upwind blend, BDF2, faces in grid order",
                   "cells": f.n, "faces": f.n_faces, "nnz": len(f.col)},
        "shuffled_faces_ms": timings["shuffled"] * 1e3,
        "upload_ms": t_up * 1e3, "speedup_vs_upload": t_up / t_asm, "cpu_baseline": cpu,
    }), flush=True)


def bench_icc(args):
    """The reference's default preconditioner on the headline system: ICC
    (DIC, level-scheduled, bit-exact sweeps) against Jacobi, CG to relres
    1e-8 on the 200^3 two-phase system, in FP64 and in the FP32-sweep form
    (ICC's z from the FP64 sweeps of the FP32 residual)."""
    import tflows_cases
    import tflows_gpu as G

    N = args.n
    s = tflows_cases.hex_pressure(N, rho_out=1000.0)
    nat = G.Native(s.row, s.col, s.dia)
    dv, db = G.DeviceArray(s.val), G.DeviceArray(s.b)
    dx = G.DeviceArray(np.zeros(s.n))
    runs = {}
    for name, prec, precision in (("icc", "icc", G.FP64), ("jacobi", "jacobi", G.FP64),
                                  ("icc_fp32_ir", "icc", G.FP32_IR),
                                  ("jacobi_fp32_ir", "jacobi", G.FP32_IR)):
        opt = G.Options()
        opt.criterion = G.CRIT_RELRES
        opt.true_residual = 1
        opt.precision = precision
        dx.upload(np.zeros(s.n))
        t = time.perf_counter()
        it, res, rep = nat.solve_native_device("cg", prec, dv, dx, db, 50000, 1e-8, options=opt,
                                               report=True)
        dt = time.perf_counter() - t
        runs[name] = {"seconds": dt, "iterations": it, "refinements": rep["refinements"],
                      "relres": res, "true_relres": rep["true_relres"],
                      "ms_per_iteration": dt / max(it, 1) * 1e3}
        log(f"bench icc: {name}: {dt:.3f} s, {it} iterations, true relres "
            f"{rep['true_relres']:.2e}")
    print(json.dumps({
        "metric": f"time to 1e-8 rel. residual with ICC vs Jacobi, {N}^3 two-phase pressure",
        "value": runs["icc"]["seconds"], "unit": "s", "n_gpus": 1, "steps": 1,
        "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"two-phase (1:1000) pressure-Poisson {N}^3, CG relres 1e-8",
                   "n": s.n, "nnz": s.nnz}, "runs": runs}), flush=True)


def bench_poisson40(args):
    """The small case the reference runs on a CPU: 40^3 Poisson (64,000
    cells, cell 1 pinned), Jacobi CG to 1e-8 with the reference's NRMS (norm
    absent and norm = 1) and with relres; GPU against the oracle, both timed,
    iteration counts side by side."""
    import tflows_cases
    import tflows_gpu as G
    from oracle import oracle

    s = tflows_cases.hex_pressure(40)
    nat = G.Native(s.row, s.col, s.dia)
    runs = {}
    for name, norm, crit in (("nrms", None, 0), ("nrms_norm1", 1.0, 0), ("relres", None, 1)):
        opt = G.Options()
        opt.criterion = crit
        best = None
        for _ in range(max(3, args.steps)):
            x = np.zeros(s.n)
            t = time.perf_counter()
            it, res = nat.solve_native("cg", "jacobi", s.val.copy(), x, s.b.copy(), 1000, 1e-8,
                                       norm=norm, options=opt)
            dt = time.perf_counter() - t
            best = (dt, it, res, x) if best is None or dt < best[0] else best
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        xo = np.zeros(s.n)
        t = time.perf_counter()
        it_o, res_o = oracle.solve("cg", "jacobi", A, xo, s.b.copy(), 1000, 1e-8, norm=norm,
                                   criterion="nrms" if crit == 0 else "relres")
        dt_o = time.perf_counter() - t
        runs[name] = {"gpu_seconds": best[0], "gpu_iterations": best[1], "gpu_res": best[2],
                      "oracle_seconds": dt_o, "oracle_iterations": it_o, "oracle_res": res_o,
                      "x_rel_diff": float(np.linalg.norm(best[3] - xo) / np.linalg.norm(xo))}
        log(f"bench poisson40 {name}: gpu {best[0] * 1e3:.1f} ms {best[1]} it; "
            f"oracle {dt_o * 1e3:.1f} ms {it_o} it")
    print(json.dumps({
        "metric": "40^3 Poisson Jacobi-CG to 1e-8 (host arrays, drop-in call), GPU vs 1-thread oracle",
        "value": runs["nrms"]["gpu_seconds"], "unit": "s", "n_gpus": 1, "steps": max(3, args.steps),
        "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "hex Poisson 40^3, pinned cell 1, Jacobi, tol 1e-8", "n": s.n,
                   "nnz": s.nnz}, "runs": runs}), flush=True)


def calibrate():
    """Known-traffic copy (1 GiB read + 1 GiB write) for counter calibration,
    the HBM ceilings of the three access mixes: 1 GiB copied, read, and
    written (16-byte accesses, 4 workgroups per CU, mean of 5 launches after
    one warm-up), and the cost of a grid-wide barrier against a kernel
    boundary."""
    import tflows_gpu as G

    nbytes = 1 << 30
    src, dst = G.DeviceArray(nbytes, np.uint8), G.DeviceArray(nbytes, np.uint8)
    ms = G.stream_copy(dst, src, nbytes, reps=5)
    ms_read = G.stream_bench(dst, src, nbytes, 1, reps=5)
    ms_write = G.stream_bench(dst, None, nbytes, 2, reps=5)
    # phase separation: a grid-wide barrier in one cooperative launch against
    # a kernel boundary, 1000 of each, at the grid sizes of the solve's kernels
    sync = {}
    for grid in (64, 128, 256, 768, 2048):
        try:
            b, k = G.sync_bench(grid, 1000)
            sync[str(grid)] = {"grid_barrier_us": b, "kernel_boundary_us": k}
        except G.TFlowsGpuError as e:   # not resident at once on this part
            sync[str(grid)] = {"error": str(e)}
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "ms": ms, "GBps": 2 * nbytes / (ms * 1e-3) / 1e9,
                      "read_ms": ms_read, "read_GBps": nbytes / (ms_read * 1e-3) / 1e9,
                      "write_ms": ms_write, "write_GBps": nbytes / (ms_write * 1e-3) / 1e9,
                      "peak_GBps": HBM_PEAK_GBS, "sync_workgroups": sync}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=200, help="grid side (default 200 -> 8M cells)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batch workload: skip the one-after-another comparison leg")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the known-traffic copy kernel (counter calibration)")
    ap.add_argument("--workload", default="twophase200",
                    choices=["twophase200", "upwind120", "knn4m", "batch256", "assembly200", "momentum200",
                             "icc200", "poisson40"],
                    help="the headline solve (default) or one of the other measured workloads")
    ap.add_argument("--summarize-counters", nargs=4, metavar=("FETCH", "WRITE", "CAL_FETCH", "CAL_WRITE"),
                    help="write the traffic summary from four rocprofv3 counter_collection.csv files")
    ap.add_argument("--summarize-stats", metavar="KERNEL_STATS_CSV",
                    help="write a markdown table of a rocprofv3 kernel_stats.csv")
    ap.add_argument("--out", help="output path of --summarize-*")
    ap.add_argument("--command", default="", help="command line recorded in the summary")
    ap.add_argument("--label", default="200^3 two-phase Jacobi-CG FP64 solve (bench.py --steps 1 --warmup 0)",
                    help="workload named in the title of a --summarize-stats table")
    args = ap.parse_args()
    if args.calibrate:
        calibrate()
        return
    if args.summarize_counters:
        summarize_counters(*args.summarize_counters, args.out or TRAFFIC, args.command.split(" ;; "))
        return
    if args.summarize_stats:
        summarize_stats(args.summarize_stats, args.out,
                        {"workload": args.label, "cmd": args.command})
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(the solver runs on one GPU; the design notes describe how it would shard by rows)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "twophase200":
        {"upwind120": bench_upwind, "knn4m": bench_knn_ir, "batch256": bench_batch,
         "assembly200": bench_assembly, "momentum200": bench_momentum, "icc200": bench_icc,
         "poisson40": bench_poisson40}[args.workload](args)
        return

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
    x_gpu = dx.download()
    # the FP32 refinement form on the same system, outside the timed steps
    fp32_runs = fp32_leg(G, nat, s, dv, db, dx, tol, miter)
    # the drop-in call with host arrays (val, b and x0 uploaded, x
    # downloaded: PCIe inside the call), once, outside the timed steps
    opt_h = G.Options()
    opt_h.criterion = G.CRIT_RELRES
    xh = np.zeros(s.n)
    t = time.perf_counter()
    it_h, _ = nat.solve_native("cg", "jacobi", s.val, xh, s.b, miter, tol, options=opt_h)
    host_s = time.perf_counter() - t
    log(f"bench: drop-in call with host arrays: {host_s:.3f} s, {it_h} iterations")
    spmv_ms = float(np.mean([r[2]["spmv_ms"] for r in reps]))
    kind = nat.info()["spmv_kind"]
    sym_ran = kind == 8 and nat.info()["sym_state"] == 1
    total = float(np.sum(times))
    ms_per_step = total / args.steps * 1e3
    iter_bytes = 12 * s.nnz + 20 * s.n + 4 + 88 * s.n
    solve_gbs = iter_bytes * its[-1] / (times[-1]) / 1e9

    spmv_kernel = "k_spmv_pq_sym" if sym_ran else "k_spmv_pq"
    traffic, traffic_src = traffic_per_launch(spmv_kernel, None if sym_ran else kind)
    cpu = parity = fp32 = None
    if not args.no_cpu:
        cpu, parity, x_ref = cpu_baseline(s, x_gpu, its[-1])
        fp32 = fp32_summary(fp32_runs, x_gpu, x_ref, total / args.steps, tol, s.scipy().tocsr())

    # the product against the measured HBM ceilings: its format bytes (all
    # but q, which it writes, are read) at the read and write sweep rates
    fmt_bytes = format_bytes(kind if (kind != 8 or sym_ran) else 7, s.nnz, s.n,
                             nat.info()["sym_diagonals"])
    practical = practical_floor(G, fmt_bytes - 8 * s.n, 8 * s.n, spmv_ms)
    log(f"bench: product at {practical['frac']:.2f} of its practical floor "
        f"({practical['floor_us']:.1f} us at {practical['read_GBps']:.0f} GB/s read, "
        f"{practical['write_GBps']:.0f} GB/s write)")
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
            "solve_GBps_algorithmic": solve_gbs,
            "drop_in_host_arrays_s": host_s, "parallelism": "single GPU",
        },
        "roofline": dict(
            spmv_roofline(kind if (kind != 8 or sym_ran) else 7, s.nnz, s.n,
                          nat.info()["sym_diagonals"], spmv_ms, traffic),
            practical=practical,
            traffic_source=os.path.relpath(traffic_src, ROOT) if traffic_src else None,
            kernel=spmv_kernel + " (FP64 SpMV with fused p.q; engine "
            + G.Native.SPMV_KINDS[kind] + ")",
            launch_timing="mean of sampled in-solve launches, hipEvents on the library's stream"),
        "cpu_baseline": cpu,
        "parity": parity,
        "fp32_refinement": fp32,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
