"""Batched solves (tfg_batch_*) against the oracle, system by system.

Each system of a batch must behave as its own Solve_Native call: iteration
count within +-2 of the oracle's, x within 1e-10 relative of the oracle's or
within the spread that reordering the oracle's global sums causes
(parity_bar.py).
"""
import numpy as np
import pytest

import parity_bar

pytestmark = pytest.mark.gpu


def _systems(cases, specs):
    out = []
    for kind, a, b in specs:
        if kind == "hex":
            out.append(cases.hex_pressure(a, rho_out=b))
        elif kind == "knn":
            out.append(cases.knn(a, a, a))
        else:
            out.append(cases.upwind(a))
    return out


def _concat(systems, attr):
    return np.ascontiguousarray(np.concatenate([getattr(s, attr) for s in systems]))


def _oracle(oracle, s, prec, miter, tol, norm=None, criterion="nrms", sum_order=0, x0=None):
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    x = np.zeros(s.n) if x0 is None else x0.copy()
    it, res = oracle.solve("cg", prec, A, x, s.b.copy(), miter, tol, norm=norm,
                           criterion=criterion, sum_order=sum_order)
    return it, res, x


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


SPECS = [("hex", 6, 1.0), ("hex", 12, 1000.0), ("hex", 9, 30.0), ("knn", 7, 0),
         ("hex", 16, 300.0), ("upwind", 8, 0), ("hex", 4, 1.0), ("hex", 14, 5.0)]


def _batch(gpu, systems):
    return gpu.Batch([(s.row, s.col, s.dia) for s in systems])


@pytest.mark.parametrize("prec,crit,norm", [("jacobi", 0, None), ("none", 0, None),
                                            ("jacobi", 1, None), ("jacobi", 0, 1.0)])
def test_batch_parity(cases, oracle, gpu, prec, crit, norm):
    systems = _systems(cases, SPECS)
    bt = _batch(gpu, systems)
    print("batch", bt.info(), prec, crit, norm, flush=True)
    x = np.zeros(int(bt.offsets[-1]))
    opt = gpu.Options()
    opt.criterion = crit
    miter, tol = 3000, 1e-8
    norms = None if norm is None else np.full(len(systems), norm)
    niter, fin, rep = bt.solve("cg", prec, _concat(systems, "val"), x, _concat(systems, "b"),
                               miter, tol, norm=norms, options=opt)
    cname = "nrms" if crit == 0 else "relres"
    for k, s in enumerate(systems):
        xs = x[bt.offsets[k]:bt.offsets[k + 1]]
        it_o, res_o, x_o = _oracle(oracle, s, prec, miter, tol, norm, cname)
        # unpreconditioned CG on the 1:1000 system: reordering the sums alone
        # moves the iteration count by more than 2; the bar widens likewise
        env, spread = parity_bar.envelope(
            lambda so: _oracle(oracle, s, prec, miter, tol, norm, cname, sum_order=so)[::2],
            it_o, x_o)
        print(f"  system {k} {s.name}: res {fin[k]:.3e} (oracle {res_o:.3e})", flush=True)
        parity_bar.check(f"batch {prec} {cname} norm={norm} system {k} {s.name}", niter[k], xs,
                         it_o, x_o, env, spread)
        if it_o < miter:
            assert fin[k] < tol
    assert rep["iterations"] == int(niter.max())


def test_graph_and_kernels_paths_agree(cases, gpu):
    systems = _systems(cases, SPECS[:5])
    bt = _batch(gpu, systems)
    val, b = _concat(systems, "val"), _concat(systems, "b")
    out = []
    for path in (gpu.PATH_GRAPH, gpu.PATH_KERNELS, gpu.PATH_GRAPH):
        opt = gpu.Options()
        opt.path = path
        x = np.zeros(int(bt.offsets[-1]))
        niter, fin, _ = bt.solve("cg", "jacobi", val, x, b, 2000, 1e-9, options=opt)
        out.append((niter.copy(), fin.copy(), x))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])
        assert np.array_equal(o[2], out[0][2])


def test_edge_systems(cases, oracle, gpu):
    """A zero right-hand side, an exact initial guess and the iteration cap,
    side by side with normal systems."""
    systems = _systems(cases, [("hex", 8, 10.0), ("hex", 6, 1.0), ("hex", 7, 100.0),
                               ("hex", 10, 1.0)])
    systems[1].b[:] = 0.0
    bt = _batch(gpu, systems)
    x0 = np.zeros(int(bt.offsets[-1]))
    # exact guess for system 2: a converged oracle solution
    it, _, xs2 = _oracle(oracle, systems[2], "jacobi", 5000, 1e-14)
    x0[bt.offsets[2]:bt.offsets[3]] = xs2
    x = x0.copy()
    niter, fin, _ = bt.solve("cg", "jacobi", _concat(systems, "val"), x, _concat(systems, "b"),
                             5, 1e-8)
    assert niter[1] == 0 and fin[1] == 0.0
    assert niter[2] == 0
    assert niter[0] == 5 and niter[3] == 5       # capped
    for k in (0, 3):
        it_o, res_o, x_o = _oracle(oracle, systems[k], "jacobi", 5, 1e-8)
        assert it_o == 5 and abs(fin[k] - res_o) <= 1e-8 * res_o
        assert _rel(x[bt.offsets[k]:bt.offsets[k + 1]], x_o) < 1e-12


def test_device_pointers_and_determinism(cases, gpu):
    systems = _systems(cases, SPECS[:4])
    bt = _batch(gpu, systems)
    val, b = _concat(systems, "val"), _concat(systems, "b")
    n = int(bt.offsets[-1])
    xh = np.zeros(n)
    r1 = bt.solve("cg", "jacobi", val, xh, b, 1000, 1e-9)
    dv, db, dx = gpu.DeviceArray(val), gpu.DeviceArray(b), gpu.DeviceArray(np.zeros(n))
    r2 = bt.solve("cg", "jacobi", dv, dx, db, 1000, 1e-9, device=True)
    xd = dx.download()
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
    assert np.array_equal(xh, xd)
    t = bt.times()
    assert np.all(t > 0) and np.all(t <= r2[2]["solve_ms"] * 1.01 + 1e-3)


def test_batch_many_medium(cases, gpu):
    """64 two-phase systems of sides 10..30 to relres 1e-8, checked on each
    system's FP64 true residual computed on the host."""
    sides, ratios = cases.batch_params(64, 10, 30, 1000.0)
    systems = [cases.hex_pressure(int(sd), rho_out=float(rt), seed=1000 + k)
               for k, (sd, rt) in enumerate(zip(sides, ratios))]
    bt = _batch(gpu, systems)
    opt = gpu.Options()
    opt.criterion = gpu.CRIT_RELRES
    x = np.zeros(int(bt.offsets[-1]))
    niter, fin, rep = bt.solve("cg", "jacobi", _concat(systems, "val"), x, _concat(systems, "b"),
                               5000, 1e-8, options=opt)
    print(f"64 systems: max niter {niter.max()}, {rep['solve_ms']:.1f} ms", flush=True)
    for k, s in enumerate(systems):
        xs = x[bt.offsets[k]:bt.offsets[k + 1]]
        r = s.b - s.scipy() @ xs
        assert niter[k] < 5000 and fin[k] < 1e-8
        assert np.linalg.norm(r) / np.linalg.norm(s.b) < 1e-7


def test_batch_diagonal_products_equal_csr(cases, gpu, monkeypatch):
    """Systems on a few diagonals (hex grids) take their products from
    zero-padded diagonals, the others (kNN; upwind, whose values are not
    symmetric) from the CSR rows, in one batch.  Forcing every system onto
    the CSR product (TFG_BATCH_SPMV=csr) gives the same bits: niter, fin_res
    and x.  A hex system whose val is made asymmetric by one ulp falls back
    to the CSR product for that solve, again with the same bits."""
    systems = _systems(cases, SPECS)
    bt = _batch(gpu, systems)
    val, b = _concat(systems, "val"), _concat(systems, "b")
    row0 = systems[1].row - 1
    i = systems[1].n // 2
    j = next(t for t in range(row0[i], row0[i + 1]) if systems[1].col[t] - 1 < i)
    val_asym = val.copy()
    off = int(sum(s.nnz for s in systems[:1]))
    val_asym[off + j] = np.nextafter(val_asym[off + j], 0.0)
    for v in (val, val_asym):
        out = []
        for mode in ("", "csr"):
            monkeypatch.setenv("TFG_BATCH_SPMV", mode)
            x = np.zeros(int(bt.offsets[-1]))
            niter, fin, _ = bt.solve("cg", "jacobi", v, x, b, 3000, 1e-9)
            out.append((niter.copy(), fin.copy(), x))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
        assert np.array_equal(out[0][2], out[1][2])


# ---- FP32 sweeps inside FP64 refinement (TFG_FP32_IR), per system ---------

def _fp32_opt(gpu, criterion=1):
    opt = gpu.Options()
    opt.precision = gpu.FP32_IR
    opt.criterion = criterion
    return opt


def _deflates_in_batch(s):
    """The batch deflates a system automatically (as the single-system form
    does) when its values are symmetric bit for bit and its row sums add to
    less than 1e-2 of the diagonal's sum."""
    A = s.scipy()
    rows = np.asarray(A.sum(axis=1)).ravel()
    return bool(abs(A - A.T).max() == 0.0 and 0.0 < rows.sum() < 1e-2 * A.diagonal().sum())


FP32_SPECS = [("hex", 10, 1000.0), ("hex", 7, 1.0), ("knn", 7, 0), ("hex", 12, 30.0),
              ("upwind", 7, 0)]


@pytest.mark.parametrize("miter", [5, 8, 24, 61])
def test_batch_fp32_steps_match_oracle(cases, oracle, gpu, miter):
    """A fixed number of FP32 iterations, each system against the oracle's
    restatement of the single-system form (the same refinement decisions
    every 8 iterations): the same count, x within 1e-10 (only the order of
    the FP64 sums differs), the last true residual within FP64 rounding."""
    systems = _systems(cases, FP32_SPECS)
    bt = _batch(gpu, systems)
    x = np.zeros(int(bt.offsets[-1]))
    niter, fin, _ = bt.solve("cg", "jacobi", _concat(systems, "val"), x, _concat(systems, "b"),
                             miter, 1e-30, options=_fp32_opt(gpu))
    for k, s in enumerate(systems):
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        xo = np.zeros(s.n)
        it_o, res_o, nref_o = oracle.ir_solve("jacobi", A, xo, s.b.copy(), miter, 1e-30,
                                              criterion="relres", delta=1e-3, check=8,
                                              deflate=_deflates_in_batch(s))
        xs = x[bt.offsets[k]:bt.offsets[k + 1]]
        err = _rel(xs, xo)
        print(f"batch FP32 miter {miter} system {k} {s.name}: gpu {niter[k]} res {fin[k]:.4e}; "
              f"oracle {it_o} res {res_o:.4e} ({nref_o} refinements); x rel diff {err:.2e}",
              flush=True)
        assert niter[k] == it_o == miter
        # the two last true residuals differ by at most ||A (x - x_o)|| / ||b||
        # (exact) plus each one's FP64 rounding
        rows = s.row - 1
        g = (np.diff(rows) + 2) * 2.0 ** -53 * (
            np.abs(s.b) + np.add.reduceat(np.abs(s.val * xs[s.col - 1]), rows[:-1]))
        bn = np.linalg.norm(s.b)
        dres = np.linalg.norm(s.scipy() @ (xs - xo)) / bn
        assert abs(fin[k] - res_o) <= dres + 2.0 * np.linalg.norm(g) / bn
        assert err <= 1e-10


def test_batch_fp32_converged(cases, oracle, gpu):
    """FP32 sweeps to relres 1e-8 and 1e-10: every system meets the
    tolerance on its FP64 true residual recomputed on the host, its error
    against a tight FP64 oracle solution is within 10x that of FP64 CG at
    the same tolerance or of the tolerance (the FP32 form stops with more
    error in the smooth modes), and the batch's count equals the oracle
    restatement's within the spread of equally valid orders."""
    systems = _systems(cases, [("hex", 14, 1000.0), ("hex", 9, 1.0), ("knn", 8, 0),
                               ("hex", 16, 100.0), ("hex", 5, 3.0)])
    bt = _batch(gpu, systems)
    val, b = _concat(systems, "val"), _concat(systems, "b")
    for tol in (1e-8, 1e-10):
        x = np.zeros(int(bt.offsets[-1]))
        niter, fin, _ = bt.solve("cg", "jacobi", val, x, b, 20000, tol, options=_fp32_opt(gpu))
        x64 = np.zeros(int(bt.offsets[-1]))
        o64 = gpu.Options()
        o64.criterion = 1
        bt.solve("cg", "jacobi", val, x64, b, 20000, tol, options=o64)
        for k, s in enumerate(systems):
            sl = slice(bt.offsets[k], bt.offsets[k + 1])
            xs = x[sl]
            true = np.linalg.norm(s.b - s.scipy() @ xs) / np.linalg.norm(s.b)
            A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
            xe = np.zeros(s.n)
            oracle.solve("cg", "jacobi", A, xe, s.b.copy(), 50000, 1e-14, criterion="relres")
            err, err64 = _rel(xs, xe), _rel(x64[sl], xe)
            env = max(_rel(_oracle(oracle, s, "jacobi", 50000, 1e-14, criterion="relres",
                                   sum_order=so)[2], xe) for so in parity_bar.ORDERS)
            xo = np.zeros(s.n)
            it_o, _, _ = oracle.ir_solve("jacobi", oracle.Matrix(s.row, s.col, s.dia, s.val.copy()),
                                         xo, s.b.copy(), 20000, tol, criterion="relres",
                                         delta=1e-3, check=8, deflate=_deflates_in_batch(s))
            print(f"batch FP32 tol {tol:.0e} system {k} {s.name}: {niter[k]} FP32 iterations "
                  f"(oracle form {it_o}), true relres {true:.2e}, error {err:.2e} (FP64 CG "
                  f"{err64:.2e}, reordered oracle {env:.2e})", flush=True)
            assert fin[k] < tol and true < 1.01 * tol
            # on a well-conditioned system FP64 CG's error at one tolerance
            # lies far below it, while the FP32 form's is set by the FP32
            # rounding of its last correction, about the tolerance itself
            # (the oracle's restatement: 4.3e-10 at 1e-10 on the 9^3
            # Poisson system, FP64 CG 1.8e-11)
            assert err <= 10.0 * max(err64, env, tol)
            assert abs(int(niter[k]) - it_o) <= max(16, it_o // 10)


def test_batch_fp32_edges(cases, gpu):
    """b = 0, an exact initial guess, the cap and miter = 0 beside normal
    systems; no preconditioner and the NRMS criterion."""
    systems = _systems(cases, [("hex", 8, 10.0), ("hex", 6, 1.0), ("hex", 7, 100.0)])
    systems[1].b[:] = 0.0
    bt = _batch(gpu, systems)
    val, b = _concat(systems, "val"), _concat(systems, "b")
    x = np.zeros(int(bt.offsets[-1]))
    niter, fin, _ = bt.solve("cg", "none", val, x, b, 7, 1e-12, options=_fp32_opt(gpu, 0))
    assert niter[1] == 0 and fin[1] == 0.0 and np.all(x[bt.offsets[1]:bt.offsets[2]] == 0.0)
    assert niter[0] == 7 and niter[2] == 7
    x = np.zeros(int(bt.offsets[-1]))
    niter, fin, _ = bt.solve("cg", "jacobi", val, x, b, 0, 1e-8, options=_fp32_opt(gpu))
    assert np.all(niter == 0) and np.all(x == 0.0)
    # converged with NRMS and the x range (no norm)
    x = np.zeros(int(bt.offsets[-1]))
    niter, fin, _ = bt.solve("cg", "jacobi", val, x, b, 5000, 1e-9, options=_fp32_opt(gpu, 0))
    for k in (0, 2):
        assert niter[k] < 5000 and fin[k] < 1e-9


def test_batch_fp32_256_systems(cases, gpu):
    """The configuration of the batched benchmark: 256 two-phase systems of
    sides 20..80 with density ratios logU(1, 1000), FP32 sweeps in FP64
    refinement to relres 1e-8; each system's FP64 true residual is
    recomputed on the host."""
    sides, ratios = cases.batch_params(256, 20, 80, 1000.0)
    systems = [cases.hex_pressure(int(sd), rho_out=float(rt), seed=7000 + k)
               for k, (sd, rt) in enumerate(zip(sides, ratios))]
    bt = _batch(gpu, systems)
    n = int(bt.offsets[-1])
    dv, db = gpu.DeviceArray(_concat(systems, "val")), gpu.DeviceArray(_concat(systems, "b"))
    dx = gpu.DeviceArray(np.zeros(n))
    niter, fin, rep = bt.solve("cg", "jacobi", dv, dx, db, 20000, 1e-8, options=_fp32_opt(gpu),
                               device=True)
    x = dx.download()
    print(f"256 systems, {n} cells: max FP32 iterations {niter.max()}, "
          f"{rep['solve_ms']:.1f} ms", flush=True)
    worst = 0.0
    for k, s in enumerate(systems):
        xs = x[bt.offsets[k]:bt.offsets[k + 1]]
        true = np.linalg.norm(s.b - s.scipy() @ xs) / np.linalg.norm(s.b)
        worst = max(worst, true)
        assert niter[k] < 20000 and fin[k] < 1e-8 and true < 1.01e-8, (k, niter[k], fin[k], true)
    print(f"worst true relres {worst:.3e}", flush=True)


def test_batch_rows_per_thread_keeps_bits(cases, gpu, monkeypatch):
    """Two rows per thread in the vector kernels (two row blocks per
    workgroup, each with its own reduction slots) gives the bits of one row
    per thread, with systems stopping at different iterations, in FP64 and
    the FP32 form."""
    systems = _systems(cases, [("hex", 12, 100.0), ("knn", 7, 0), ("hex", 20, 1000.0),
                               ("hex", 5, 1.0), ("hex", 17, 10.0)])
    val, b = _concat(systems, "val"), _concat(systems, "b")
    monkeypatch.setenv("TFG_BATCH_TUNE", "0")
    for precision in (gpu.FP64, gpu.FP32_IR):
        out = []
        for rows in ("1", "2"):
            monkeypatch.setenv("TFG_BATCH_ROWS", rows)
            bt = _batch(gpu, systems)
            print(f"precision {precision} rows {rows}: {bt.info()}", flush=True)
            opt = gpu.Options()
            opt.precision = precision
            opt.criterion = 1
            x = np.zeros(int(bt.offsets[-1]))
            niter, fin, _ = bt.solve("cg", "jacobi", val, x, b, 3000, 1e-9, options=opt)
            out.append((niter, fin, x))
            bt.close()
        assert len(set(out[0][0].tolist())) > 1        # the systems stop at different iterations
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
        assert np.array_equal(out[0][2], out[1][2])


def test_batch_autotuner_keeps_bits(cases, gpu, monkeypatch):
    """The autotuner's probes (graph length, diagonal or CSR products per
    size class, rows per thread) leave the caller's x0 alone and every
    choice gives the same bits: a tuned solve equals one with the defaults
    (TFG_BATCH_TUNE=0), in FP64 and in the FP32 form, from a nonzero x0."""
    systems = _systems(cases, [("hex", 12, 100.0), ("knn", 7, 0), ("hex", 20, 1000.0),
                               ("hex", 33, 10.0), ("hex", 9, 1.0)])
    val, b = _concat(systems, "val"), _concat(systems, "b")
    n = sum(s.n for s in systems)
    x0 = np.random.default_rng(3).uniform(-1.0, 1.0, n)
    for precision in (gpu.FP64, gpu.FP32_IR):
        out = []
        for tune in ("1", "0"):
            monkeypatch.setenv("TFG_BATCH_TUNE", tune)
            bt = _batch(gpu, systems)
            opt = gpu.Options()
            opt.precision = precision
            opt.criterion = 1
            x = x0.copy()
            niter, fin, _ = bt.solve("cg", "jacobi", val, x, b, 3000, 1e-9, options=opt)
            t = bt.tuning(precision)
            print(f"precision {precision} TFG_BATCH_TUNE={tune}: {t}, max niter {niter.max()}",
                  flush=True)
            if tune == "1":
                assert t["probes"] >= 3 and t["chunk"] in (8, 16, 32)
                assert 0 <= t["diag_classes"] <= 7 and t["best_probe_ms"] > 0.0
                assert t["rows_per_thread"] in (1, 2)
            else:
                assert t["probes"] == 0 and t["diag_classes"] == 7 and t["rows_per_thread"] == 1
            out.append((niter, fin, x))
            bt.close()
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
        assert np.array_equal(out[0][2], out[1][2])
