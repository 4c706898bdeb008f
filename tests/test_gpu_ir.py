"""FP32 sweeps inside FP64 refinement (TFG_FP32_IR; an extension: the
reference has no mixed-precision solver) through the C ABI.

* The FP32 product of the row-excess form is bit-exact against the oracle's
  restatement (oracle.lap32_form / lap32_spmv) with every engine that runs
  it: the symmetric FP32 storage (four rows per thread, one row per thread)
  and the CSR stream, coded and pattern engines.
* Step for step: a fixed number of FP32 iterations agrees with the oracle's
  restatement of the same form (oracle.ir_solve), which makes the same
  refinement decisions; only the order of the global sums differs.
* Converged: the stopping test is taken on the FP64 TRUE residual, so the
  requested tolerance holds for ||b - A x|| / ||b|| recomputed on the host;
  the error against a tightly converged FP64 oracle solve is stated for
  every run and bounded by the error the FP64 solve reaches at the same
  tolerance, or by the spread that reordering the FP64 oracle's own sums
  causes, whichever is larger (parity_bar.py: on these nearly singular
  systems that spread is the FP64 rounding floor of the near-null mode).
"""
import numpy as np
import pytest

import parity_bar

pytestmark = pytest.mark.gpu


def _oracle_x(oracle, s, tol=1e-14, sum_order=0):
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    x = np.zeros(s.n)
    it, _ = oracle.solve("cg", "jacobi", A, x, s.b.copy(), 50000, tol, criterion="relres",
                         sum_order=sum_order)
    return x, it


def _opt(gpu, criterion=1, inner_max=0, inner_tol=0.0, deflation=0):
    opt = gpu.Options()
    opt.precision = gpu.FP32_IR
    opt.criterion = criterion
    opt.inner_max = inner_max
    opt.inner_tol = inner_tol
    opt.deflation = deflation
    opt.true_residual = 1
    return opt


def _ir(gpu, s, tol, prec="jacobi", miter=20000, norm=None, nat=None, b=None, x0=None, **kw):
    nat = nat or gpu.Native(s.row, s.col, s.dia)
    x = np.zeros(s.n) if x0 is None else x0.copy()
    it, res, rep = nat.solve_native("cg", prec, s.val.copy(), x, (s.b if b is None else b).copy(),
                                    miter, tol, norm=norm, options=_opt(gpu, **kw), report=True)
    return it, res, x, rep


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _deflates(s):
    """The library's automatic choice: symmetric values and the sum of the
    row sums below 1e-2 of the diagonal's sum."""
    A = s.scipy()
    sym = abs(A - A.T).max() == 0.0
    rows = np.asarray(A.sum(axis=1)).ravel()
    return bool(sym and 0.0 < rows.sum() < 1e-2 * A.diagonal().sum())


# ---- the FP32 product ------------------------------------------------------

PRODUCT_CASES = [
    # name, generator, forced FP64 engine (the FP32 product follows it)
    ("hex24x20x16_quad", lambda c: c.hex_pressure(24, 20, 16, rho_out=1000.0), None),
    ("hex9x7x5_rows", lambda c: c.hex_pressure(9, 7, 5, rho_out=30.0), None),
    ("hex10x8x6_rows", lambda c: c.hex_pressure(10, 8, 6), None),
    ("grid1x30x20_quad", lambda c: c.hex_pressure(1, 30, 20, rho_out=5.0), None),
    ("knn10_stream", lambda c: c.knn(10, 10, 10), None),
    ("hex12_stage", lambda c: c.hex_pressure(12, rho_out=100.0), 0),
    ("hex12_coded", lambda c: c.hex_pressure(12, rho_out=100.0), 6),
    ("hex12_pattern", lambda c: c.hex_pressure(12, rho_out=100.0), 7),
    ("upwind10_asymmetric", lambda c: c.upwind(10), None),
]


@pytest.mark.parametrize("name,make,kind", PRODUCT_CASES, ids=[c[0] for c in PRODUCT_CASES])
def test_fp32_product_bitexact(cases, oracle, gpu, name, make, kind):
    s = make(cases)
    print(f"FP32 product {name} n={s.n} nnz={s.nnz}", flush=True)
    nat = gpu.Native(s.row, s.col, s.dia)
    if kind is not None:
        nat.spmv_select(kind)
    rng = np.random.default_rng(7)
    fn = 1.0 / (s.val[s.dia - 1].sum() / s.n)
    for x in (rng.uniform(-1, 1, s.n).astype(np.float32),
              np.full(s.n, 0.75, dtype=np.float32),     # a constant: y = s x exactly
              rng.normal(size=s.n).astype(np.float32) + 100.0):
        dv, dx = gpu.DeviceArray(s.val), gpu.DeviceArray(x)
        dy = gpu.DeviceArray(np.zeros(s.n, dtype=np.float32))
        nat.spmv_fp32(dv, fn, dx, dy)
        y = dy.download()
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        v32 = oracle.lap32_form(A, fn)
        y_ref = oracle.lap32_spmv(A, v32, x)
        assert np.array_equal(y, y_ref), (name, np.abs(y - y_ref).max())


# ---- step for step against the oracle's restatement ----------------------

@pytest.mark.parametrize("make,prec", [(lambda c: c.hex_pressure(16, rho_out=1000.0), "jacobi"),
                                       (lambda c: c.hex_pressure(12, 10, 8), "none"),
                                       (lambda c: c.knn(8, 8, 8), "jacobi"),
                                       (lambda c: c.hex_pressure(18, 17, 20, rho_out=1000.0),
                                        "icc"),
                                       (lambda c: c.knn(8, 8, 8), "icc")],
                         ids=["twophase16_jacobi", "poisson_none", "knn8_jacobi",
                              "twophase18x17x20_icc_tiles", "knn8_icc_levels"])
@pytest.mark.parametrize("miter", [5, 8, 24, 61])
def test_fp32_steps_match_oracle(cases, oracle, gpu, make, prec, miter):
    s = make(cases)
    it, res, x, rep = _ir(gpu, s, 1e-30, prec=prec, miter=miter)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    xo = np.zeros(s.n)
    it_o, res_o, nref_o = oracle.ir_solve(prec, A, xo, s.b.copy(), miter, 1e-30,
                                          criterion="relres", delta=1e-3, check=8,
                                          deflate=_deflates(s))
    err = _rel(x, xo)
    print(f"{s.name} {prec} miter {miter}: gpu {it} it {rep['refinements']} refinements "
          f"res {res:.4e} deflated {rep['deflated']}; oracle {it_o} it {nref_o} refinements "
          f"res {res_o:.4e}; x rel diff {err:.2e}", flush=True)
    assert it == it_o == miter
    assert rep["refinements"] == nref_o
    assert rep["deflated"] == int(_deflates(s))
    # the last true residual: apart by what the two x make apart (their
    # FP64 sums run in different orders), plus each side's own FP64
    # rounding (which dominates once the residual is ~1e-12 of b)
    dx = float(np.linalg.norm(s.scipy() @ (x - xo)) / np.linalg.norm(s.b))
    assert abs(res - res_o) <= dx + 1e-6 * res_o + 2.0 * _relres_rounding(s, x)
    # the same FP32 iterates; only the order of the FP64 sums differs
    # (measured: at most 1e-11)
    assert err <= 1e-10


# ---- converged solves ----------------------------------------------------

CONVERGED = [
    ("poisson40", lambda c: c.hex_pressure(40), 1e-10),
    ("twophase32", lambda c: c.hex_pressure(32, rho_out=1000.0), 1e-8),
    ("twophase24_tight", lambda c: c.hex_pressure(24, rho_out=1000.0), 1e-10),
    ("knn14", lambda c: c.knn(14, 14, 14), 1e-10),
]


@pytest.mark.parametrize("name,make,tol", CONVERGED, ids=[c[0] for c in CONVERGED])
def test_fp32_converged_error(cases, oracle, gpu, name, make, tol):
    s = make(cases)
    it, res, x, rep = _ir(gpu, s, tol)
    true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
    xe, _ = _oracle_x(oracle, s)
    # FP64 CG at the same tolerance on the GPU (tree sums), and the spread
    # of the FP64 oracle's own reordered sums around its tight solution
    nat = gpu.Native(s.row, s.col, s.dia)
    o64 = gpu.Options()
    o64.criterion = 1
    x64 = np.zeros(s.n)
    it64, _ = nat.solve_native("cg", "jacobi", s.val.copy(), x64, s.b.copy(), 50000, tol,
                               options=o64)
    env = max(_rel(_oracle_x(oracle, s, sum_order=so)[0], xe) for so in parity_bar.ORDERS)
    err, err64 = _rel(x, xe), _rel(x64, xe)
    # At one tolerance the FP32 form stops later in the near-null modes than
    # FP64 CG (measured: up to 6x FP64 CG's error); the bench states the
    # speed-up at equal error, with the FP32 form run to tighter tolerances.
    bar = 10.0 * max(err64, env)
    print(f"FP32 {name}: {it} FP32 iterations (FP64 CG {it64}), {rep['refinements']} "
          f"refinements, deflated {rep['deflated']}, relres {res:.2e}, true {true:.2e}; error "
          f"vs FP64 oracle {err:.2e}, FP64 CG's {err64:.2e}, reordered-oracle spread {env:.2e}, "
          f"err/bar {err / bar:.2f}", flush=True)
    assert res < tol and true < 1.01 * tol
    assert abs(rep["true_relres"] - true) <= 1e-3 * true
    assert rep["refinements"] >= 1
    assert err <= bar


ICC_CONVERGED = [
    ("twophase24_icc", lambda c: c.hex_pressure(24, rho_out=1000.0), 1e-8),
    ("twophase20_icc_tight", lambda c: c.hex_pressure(20, rho_out=1000.0), 1e-10),
    ("knn12_icc", lambda c: c.knn(12, 12, 12), 1e-10),
]


@pytest.mark.parametrize("name,make,tol", ICC_CONVERGED, ids=[c[0] for c in ICC_CONVERGED])
def test_fp32_icc_converged(cases, oracle, gpu, name, make, tol):
    """The FP32 form with the reference's default preconditioner: converged
    on the FP64 true residual, with an error against the tight FP64 solution
    within ten times FP64 ICC-CG's at the same tolerance (or the reordered
    oracle's spread), and fewer iterations than the FP32 form with Jacobi."""
    s = make(cases)
    it, res, x, rep = _ir(gpu, s, tol, prec="icc")
    it_j, _, _, _ = _ir(gpu, s, tol, prec="jacobi")
    true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
    xe, _ = _oracle_x(oracle, s)
    nat = gpu.Native(s.row, s.col, s.dia)
    o64 = gpu.Options()
    o64.criterion = 1
    x64 = np.zeros(s.n)
    it64, _ = nat.solve_native("cg", "icc", s.val.copy(), x64, s.b.copy(), 50000, tol,
                               options=o64)
    env = max(_rel(_oracle_x(oracle, s, sum_order=so)[0], xe) for so in parity_bar.ORDERS)
    err, err64 = _rel(x, xe), _rel(x64, xe)
    bar = 10.0 * max(err64, env)
    print(f"FP32 ICC {name}: {it} FP32 iterations (FP64 ICC-CG {it64}, FP32 Jacobi {it_j}), "
          f"{rep['refinements']} refinements, deflated {rep['deflated']}, relres {res:.2e}, "
          f"true {true:.2e}; error {err:.2e}, FP64 ICC-CG's {err64:.2e}, spread {env:.2e}, "
          f"err/bar {err / bar:.2f}", flush=True)
    assert res < tol and true < 1.01 * tol
    assert rep["refinements"] >= 1
    assert it < it_j
    assert err <= bar


def _residual_rounding(s, x):
    """Per row, the first-order bound on the FP64 rounding of b_i - A_i x:
    (len_i + 2) u (|b_i| + sum_j |a_ij x_j|), u = 2^-53 (len_i + 1 rounded
    terms, and the library's scaling by fn rounds each once more)."""
    rows = s.row - 1
    aabs = np.add.reduceat(np.abs(s.val * x[s.col - 1]), rows[:-1])
    return (np.diff(rows) + 2) * 2.0 ** -53 * (np.abs(s.b) + aabs)


def _relres_rounding(s, x):
    """That bound for ||b - A x||_2 / ||b||_2."""
    return float(np.linalg.norm(_residual_rounding(s, x)) / np.linalg.norm(s.b))


def _nrms_extended(s, x, norm):
    """NRMS of b - A x computed in extended precision, and the bound on how
    far an FP64 computation of the same residual may be from it: row i sums
    len_i + 1 rounded terms (and the library's scaling by fn rounds each once
    more), so |fl(r_i) - r_i| <= (len_i + 2) u (|b_i| + sum_j |a_ij x_j|)
    with u = 2^-53, to first order."""
    ld = np.longdouble
    rows = s.row - 1
    prod = s.val.astype(ld) * x.astype(ld)[s.col - 1]
    ax = np.add.reduceat(prod, rows[:-1])
    aabs = np.add.reduceat(np.abs(prod), rows[:-1])
    r = s.b.astype(ld) - ax
    d = s.val[s.dia - 1].astype(ld)
    lens = np.diff(rows).astype(ld)
    g = (lens + 2) * ld(2.0) ** -53 * (np.abs(s.b).astype(ld) + aabs)
    rng = ld(norm) if norm is not None else ld(x.max() - x.min())
    nr = np.sqrt(np.sum((r / d) ** 2)) / rng
    bound = np.sqrt(np.sum((g / d) ** 2)) / rng
    return float(nr), float(bound)


def test_fp32_nrms_criterion(cases, oracle, gpu):
    s = cases.hex_pressure(24, rho_out=100.0)
    for norm in (None, 1.0):
        it, res, x, rep = _ir(gpu, s, 1e-9, criterion=0, norm=norm)
        # fin_res is the NRMS of the FP64 true residual (NRMS is scale-free,
        # so the unscaled system gives the same measure); at ~1e-12 of b the
        # residual is mostly cancellation, so it is compared with the
        # extended-precision residual within the FP64 rounding bound
        nr, bound = _nrms_extended(s, x, norm)
        print(f"NRMS norm={norm}: {it} iterations, res {res:.4e}, extended {nr:.4e}, "
              f"FP64 rounding bound {bound:.1e}", flush=True)
        assert res < 1e-9 and abs(nr - res) <= bound + 1e-10 * res


def test_fp32_conventions(cases, oracle, gpu):
    s = cases.hex_pressure(10)
    # b = 0: no iteration, fin_res 0, x untouched
    it, res, x, rep = _ir(gpu, s, 1e-8, b=np.zeros(s.n))
    assert it == 0 and res == 0.0 and np.all(x == 0.0)
    # exact x0: no iteration
    xe, _ = _oracle_x(oracle, s, 1e-15)
    b = s.scipy() @ xe
    it, res, x, rep = _ir(gpu, s, 1e-8, b=b, x0=xe)
    assert it == 0 and res < 1e-8
    # the cap, inside a chunk and at a chunk end: niter = miter, fin_res the
    # measure of the FP64 true residual at the cap
    for miter in (0, 7, 16):
        it, res, x, rep = _ir(gpu, s, 1e-14, miter=miter)
        true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
        assert it == miter and res > 1e-14 and abs(res - true) <= 1e-9 * true


def test_fp32_deflation_modes(cases, gpu):
    s = cases.hex_pressure(20, rho_out=1000.0)
    runs = {}
    for mode in (0, 1, 2):
        it, res, x, rep = _ir(gpu, s, 1e-8, deflation=mode)
        true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
        runs[mode] = (it, rep["deflated"], true)
        print(f"deflation mode {mode}: {it} iterations, deflated {rep['deflated']}, true {true:.2e}",
              flush=True)
        assert true < 1.01e-8
    assert runs[0][1] == 1 and runs[1][1] == 0 and runs[2][1] == 1
    assert runs[0][0] == runs[2][0]
    # on a non-symmetric matrix the automatic choice does not deflate (CG
    # on it, the reference's "bicg", need not converge: only the choice is
    # checked)
    u = cases.upwind(10)
    it, res, x, rep = _ir(gpu, u, 1e-8, miter=40)
    assert rep["deflated"] == 0


def test_fp32_forced_refinements(cases, oracle, gpu):
    """inner_max: a refinement at least every inner_max iterations (at the
    first chunk end after it), as in the oracle."""
    s = cases.knn(10, 10, 10)
    it, res, x, rep = _ir(gpu, s, 1e-10, inner_max=16)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    xo = np.zeros(s.n)
    it_o, res_o, nref_o = oracle.ir_solve("jacobi", A, xo, s.b.copy(), 20000, 1e-10,
                                          criterion="relres", delta=1e-3, check=8, force=16,
                                          deflate=_deflates(s))
    print(f"forced every 16: gpu {it} it {rep['refinements']} ref; oracle {it_o} it {nref_o} ref",
          flush=True)
    assert res < 1e-10 and rep["refinements"] >= it // 16
    assert abs(it - it_o) <= 8


TEXTBOOK = {"bicg_textbook": ("bicg", 1), "cgs": ("cgs", 0), "bcgs": ("bcgs", 0)}


def _textbook(gpu, s, method, tol, precision, prec="jacobi", miter=2000, nat=None, crit=1):
    nat = nat or gpu.Native(s.row, s.col, s.dia)
    opt = _opt(gpu, criterion=crit) if precision else gpu.Options()
    opt.criterion = crit
    opt.true_residual = 1
    opt.textbook_bicg = TEXTBOOK[method][1]
    x = np.zeros(s.n)
    it, res, rep = nat.solve_native(TEXTBOOK[method][0], prec, s.val.copy(), x, s.b.copy(), miter,
                                    tol, options=opt, report=True)
    return it, res, x, rep


@pytest.mark.parametrize("method", ["bicg_textbook", "cgs", "bcgs"])
@pytest.mark.parametrize("prec", ["jacobi", "none"])
@pytest.mark.parametrize("tol", [1e-8, 1e-10])
def test_fp32_textbook_converged(cases, gpu, method, prec, tol):
    """BiCG, CGS and BiCGSTAB with FP32 sweeps in FP64 refinement on the
    non-symmetric upwind system: the FP64 true residual meets tol, and the
    error against the direct solution is no worse than the FP64 method's at
    the same tolerance (or than FP64 rounding of the solve)."""
    s = cases.upwind(16)
    xd = np.linalg.solve(s.scipy().toarray(), s.b)
    it, res, x, rep = _textbook(gpu, s, method, tol, True, prec)
    it64, res64, x64, _ = _textbook(gpu, s, method, tol, False, prec)
    true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
    err, err64 = _rel(x, xd), _rel(x64, xd)
    print(f"FP32 {method}/{prec} tol {tol:.0e}: {it} FP32 iterations ({rep['refinements']} "
          f"refinements; FP64 {it64}), true relres {true:.2e}, error {err:.2e} (FP64 {err64:.2e})",
          flush=True)
    assert res < tol and true < 1.01 * tol
    # at one tolerance the FP32 sweeps may stop with more error than the
    # FP64 method (measured up to 3.3x on this system)
    assert err <= max(10.0 * err64, 1e-13)


def test_fp32_textbook_conventions(cases, gpu):
    s = cases.upwind(8)
    nat = gpu.Native(s.row, s.col, s.dia)
    for method in TEXTBOOK:
        opt = _opt(gpu)
        opt.textbook_bicg = TEXTBOOK[method][1]
        x = np.zeros(s.n)
        it, res = nat.solve_native(TEXTBOOK[method][0], "jacobi", s.val.copy(), x, np.zeros(s.n),
                                   50, 1e-8, options=opt)
        assert it == 0 and res == 0.0 and np.all(x == 0.0)
        x = np.zeros(s.n)
        it, res = nat.solve_native(TEXTBOOK[method][0], "jacobi", s.val.copy(), x, s.b.copy(),
                                   3, 1e-14, options=opt)
        true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
        assert it == 3 and abs(res - true) <= 1e-9 * true


def test_fp32_textbook_upwind120(cases, gpu):
    """The 1.7M-cell upwind system: each method with FP32 sweeps reaches
    relres 1e-8 on the FP64 true residual."""
    s = cases.upwind(120)
    nat = gpu.Native(s.row, s.col, s.dia)
    for method in TEXTBOOK:
        it, res, x, rep = _textbook(gpu, s, method, 1e-8, True, nat=nat)
        it64, _, _, rep64 = _textbook(gpu, s, method, 1e-8, False, nat=nat)
        print(f"upwind120 FP32 {method}: {it} iterations, {rep['refinements']} refinements, "
              f"true {rep['true_relres']:.2e}, {rep['solve_ms']:.1f} ms (FP64 {it64} iterations, "
              f"{rep64['solve_ms']:.1f} ms)", flush=True)
        assert res < 1e-8 and rep["true_relres"] < 1.01e-8


def test_fp32_twophase_200(cases, gpu):
    """The 8M-cell two-phase system: the FP64 true residual meets 1e-8."""
    s = cases.hex_pressure(200, rho_out=1000.0)
    print(f"full size n={s.n} nnz={s.nnz}", flush=True)
    nat = gpu.Native(s.row, s.col, s.dia)
    dv, dx, db = gpu.DeviceArray(s.val), gpu.DeviceArray(np.zeros(s.n)), gpu.DeviceArray(s.b)
    it, res, rep = nat.solve_native_device("cg", "jacobi", dv, dx, db, 20000, 1e-8,
                                           options=_opt(gpu), report=True)
    x = dx.download()
    true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
    print(f"  FP32: {it} iterations, {rep['refinements']} refinements, deflated "
          f"{rep['deflated']}, res {res:.3e}, true {true:.3e}, {rep['solve_ms']:.1f} ms",
          flush=True)
    assert res < 1e-8 and true < 1.01e-8 and rep["deflated"] == 1
