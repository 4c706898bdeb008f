"""Pins of the CPU oracle (the C restatement of Native_Mod) by mathematics.

The reference holds no golden vectors, known-answer tests or fixtures for
this path, and its Fortran build needs a prebuilt third-party library
(METIS, through Grid_Mod -> Metis_Mod), so it is not compiled here.  Each
routine of the oracle is instead checked against something the mathematics
fixes, chosen so that a plausible slip (wrong sign, dropped term, transposed
operand, wrong index, wrong step length) fails:

  spmv / residual  scipy's CSR product and dense numpy products
  nrms             hand-computed special cases
  cg (recurrence)  the Galerkin characterisation of CG: x_k minimises the
                   A-norm error over x0 + K_k(M^-1 A, M^-1 r0), computed by a
                   dense projection; exact-arithmetic iteration counts on
                   matrices with few distinct eigenvalues
  prec icc         the DIC identity diag((D+L) D^-1 (D+U)) = diag(A) and the
                   dense solve with that M; per rank (the rows' ranks given),
                   each rank's own DIC on its matrix with buffer cells, bit
                   for bit, and the identity on A without cross-rank terms
  sum order 11     closed forms of fn and NRMS summed per rank
  bicg             bit-identical to cg (Bicg.f90 runs the same recurrence)
  conventions      early exits, iteration cap, optional norm (Cg.f90)
plus the committed fixtures of tests/golden (made by tests/golden/make_golden.py
with scipy/numpy, independent of the oracle).
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _mat(oracle, s):
    return oracle.Matrix(s.row, s.col, s.dia, s.val)


def _csr_from_dense(D):
    """1-based Matrix_Type arrays (diagonal first, rest ascending) of a dense
    matrix's pattern."""
    n = D.shape[0]
    row, col, val = [1], [], []
    for i in range(n):
        cols = [i] + sorted(j for j in range(n) if j != i and D[i, j] != 0.0)
        col += [c + 1 for c in cols]
        val += [D[i, c] for c in cols]
        row.append(row[-1] + len(cols))
    row = np.array(row, np.int32)
    return row, np.array(col, np.int32), row[:-1].copy(), np.array(val)


# ---------------------------------------------------------------- products --

@pytest.mark.parametrize("make", [lambda c: c.hex_pressure(9, 7, 5, rho_out=100.0),
                                  lambda c: c.upwind(7), lambda c: c.knn(6, 5, 4)])
def test_spmv_and_residual_vs_scipy(cases, oracle, make):
    s = make(cases)
    A = s.scipy()
    rng = np.random.default_rng(1)
    p = rng.standard_normal(s.n)
    M = _mat(oracle, s)
    q = oracle.spmv(M, p)
    assert np.allclose(q, A @ p, rtol=0, atol=1e-13 * np.abs(A).max() * np.abs(p).max())
    r = oracle.residual(M, s.b, p)
    assert np.allclose(r, s.b - A @ p, rtol=0, atol=1e-13 * np.abs(A).max() * np.abs(p).max())


def test_spmv_dense_brute_force(oracle):
    rng = np.random.default_rng(5)
    D = rng.standard_normal((9, 9)) * (rng.random((9, 9)) < 0.4)
    np.fill_diagonal(D, 5.0)
    row, col, dia, val = _csr_from_dense(D)
    M = oracle.Matrix(row, col, dia, val)
    x = rng.standard_normal(9)
    assert np.allclose(oracle.spmv(M, x), D @ x, atol=1e-14)
    assert np.allclose(oracle.residual(M, x, x), x - D @ x, atol=1e-14)


# -------------------------------------------------------------------- NRMS --

def test_nrms_special_cases(oracle):
    # 2x2 with diagonal (2, 4): r = (6, 16) -> r/a = (3, 4) -> rms 5
    D = np.array([[2.0, -1.0], [-1.0, 4.0]])
    row, col, dia, val = _csr_from_dense(D)
    M = oracle.Matrix(row, col, dia, val)
    r = np.array([6.0, 16.0])
    assert oracle.nrms(M, r, np.array([1.0, 11.0])) == pytest.approx(0.5, rel=1e-15)   # range 10
    assert oracle.nrms(M, r, np.array([3.0, 3.0]), norm=2.5) == pytest.approx(2.0, rel=1e-15)
    # x constant and norm absent: the range is floored at TINY = 1e-15
    assert oracle.nrms(M, r, np.array([3.0, 3.0])) == pytest.approx(5.0e15, rel=1e-15)
    # norm = 0 is floored too
    assert oracle.nrms(M, r, np.array([0.0, 1.0]), norm=0.0) == pytest.approx(5.0e15, rel=1e-15)


# ------------------------------------------------------------ CG recurrence --

def _galerkin(A, Minv, b, x0, k):
    """x_k of exact-arithmetic PCG: minimiser of the A-norm error over
    x0 + span{Minv r0, (Minv A) Minv r0, ...} (k vectors)."""
    r0 = b - A @ x0
    v = Minv @ r0
    K = []
    for _ in range(k):
        K.append(v)
        v = Minv @ (A @ v)
    Q, _ = np.linalg.qr(np.array(K).T)
    y = np.linalg.solve(Q.T @ A @ Q, Q.T @ r0)
    return x0 + Q @ y


@pytest.mark.parametrize("prec", ["jacobi", "none", "icc"])
def test_cg_iterates_are_galerkin(cases, oracle, prec):
    s = cases.hex_pressure(6, 5, 4, rho_out=20.0)
    A = s.scipy().toarray()
    if prec == "jacobi":
        Minv = np.diag(1.0 / np.diag(A))
    elif prec == "none":
        Minv = np.eye(s.n)
    else:
        L = np.tril(A, -1)
        D = np.zeros(s.n)
        for i in range(s.n):
            D[i] = A[i, i] - np.sum(L[i, :i] ** 2 / D[:i])
        Mm = (np.diag(D) + L) @ np.diag(1.0 / D) @ (np.diag(D) + L.T)
        Minv = np.linalg.inv(Mm)
    x0 = np.linspace(-0.5, 0.5, s.n)
    for k in range(1, 7):
        x = x0.copy()
        it, res = oracle.solve("cg", prec, _mat(oracle, s), x, s.b.copy(), k, 1e-300)
        assert it == k
        xg = _galerkin(A, Minv, s.b, x0, k)
        assert np.linalg.norm(x - xg) <= 1e-9 * np.linalg.norm(xg), (prec, k)
        # fin_res is the NRMS of the k-th residual (Normalized_Root_Mean_Square.f90)
        rk = s.b - A @ x
        nr = np.sqrt(np.sum((rk / np.diag(A)) ** 2)) / (x.max() - x.min())
        assert res == pytest.approx(nr, rel=1e-8)


def _complete_graph(m, shift):
    D = -np.ones((m, m))
    np.fill_diagonal(D, m - 1 + shift)
    return D


def test_exact_arithmetic_iteration_counts(oracle):
    """K_m Laplacian + I has eigenvalues {1, 1+m}: CG stops after exactly 2
    iterations.  Two such blocks of sizes 5 and 9 give 3 distinct eigenvalues
    of A (unpreconditioned: 3 iterations) but 4 distinct eigenvalues of
    D^-1 A (Jacobi: 4 iterations)."""
    D = _complete_graph(7, 1.0)
    row, col, dia, val = _csr_from_dense(D)
    b = np.arange(7, dtype=float) - 2.5
    for prec in ("none", "jacobi"):
        x = np.zeros(7)
        it, res = oracle.solve("cg", prec, oracle.Matrix(row, col, dia, val), x, b.copy(), 50, 1e-10)
        assert it == 2, prec
        assert np.allclose(D @ x, b, atol=1e-12)
    B = sla.block_diag(_complete_graph(5, 1.0), _complete_graph(9, 1.0))
    row, col, dia, val = _csr_from_dense(B)
    b = np.sin(np.arange(14.0)) + 0.1
    got = {}
    for prec in ("none", "jacobi"):
        x = np.zeros(14)
        got[prec] = oracle.solve("cg", prec, oracle.Matrix(row, col, dia, val), x, b.copy(), 50,
                                 1e-10)[0]
        assert np.allclose(B @ x, b, atol=1e-11)
    assert got == {"none": 3, "jacobi": 4}


def test_first_step_closed_form(cases, oracle):
    """x_1 = alpha z0 with alpha = (z0.r0)/(z0.A z0), z0 = D^-1 b (x0 = 0)."""
    s = cases.hex_pressure(5, 4, 3, rho_out=7.0)
    A = s.scipy().toarray()
    z = s.b / np.diag(A)
    alpha = (z @ s.b) / (z @ A @ z)
    x = np.zeros(s.n)
    oracle.solve("cg", "jacobi", _mat(oracle, s), x, s.b.copy(), 1, 1e-300)
    assert np.allclose(x, alpha * z, rtol=1e-12, atol=1e-15 * np.abs(alpha * z).max())


def test_converged_solution(cases, oracle):
    s = cases.hex_pressure(10, rho_out=50.0)
    xe = np.linalg.solve(s.scipy().toarray(), s.b)
    x = np.zeros(s.n)
    it, res = oracle.solve("cg", "jacobi", _mat(oracle, s), x, s.b.copy(), 5000, 1e-13,
                           criterion="relres")
    assert res < 1e-13
    assert np.linalg.norm(x - xe) <= 1e-9 * np.linalg.norm(xe)
    assert np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b) < 2e-13


# ---------------------------------------------------------------- ICC (DIC) --

def test_icc_dic_identity(cases, oracle):
    s = cases.hex_pressure(5, 4, 4, rho_out=30.0)
    A = s.scipy().toarray()
    M = _mat(oracle, s)
    d, dinv = oracle.prec_form(M, "icc")
    # Prec_Form.f90:68-78 stores the pivot in d_inv and its inverse in d
    assert np.allclose(d * dinv, 1.0, rtol=1e-15)
    L = np.tril(A, -1)
    P = np.diag(dinv)
    Mm = (P + L) @ np.diag(1.0 / dinv) @ (P + L.T)
    assert np.allclose(np.diag(Mm), np.diag(A), rtol=1e-13)
    # off-diagonal pattern of M equals A's on the stencil (fill-in elsewhere)
    mask = A != 0
    assert np.allclose(Mm[mask], A[mask], rtol=1e-12, atol=1e-15)
    rhs = np.cos(np.arange(s.n))
    z = oracle.prec_solve(M, d, dinv, rhs, "icc")
    assert np.allclose(Mm @ z, rhs, atol=1e-12 * np.abs(rhs).max())


def test_icc_per_rank_is_each_ranks_own_dic(cases, oracle):
    """With the rows' ranks given, "icc" on the global matrix is the DIC each
    rank forms on its own matrix (ni < nt, buffer columns never read by the
    recurrences, Prec_Solve.f90:89), to the bit; the preconditioner is the
    DIC of A without the couplings between ranks (block diagonal), checked
    by the DIC identity on that matrix."""
    import ranks as rank_split
    nx = 6
    s = cases.hex_pressure(nx, 5, 4, rho_out=30.0)
    rks = rank_split.split(s, nx, local_order=True)
    parts = rank_split.parts(rks)
    M = _mat(oracle, s)
    d, dinv = oracle.prec_form(M, "icc", parts=parts)
    rhs = np.cos(np.arange(s.n) * 0.7) + 0.1
    z = oracle.prec_solve(M, d, dinv, rhs, "icc", parts=parts)
    for rk in rks:
        ni, g = rk["ni"], rk["glob"]
        Ml = oracle.Matrix(rk["row"], rk["col"], rk["dia"], rk["val"], ni=ni)
        dl, dinvl = oracle.prec_form(Ml, "icc")
        assert np.array_equal(dl[:ni], d[g[:ni]]) and np.array_equal(dinvl[:ni], dinv[g[:ni]])
        rl = rhs[g].copy()
        rl[ni:] = 1e300                   # buffer cells: never read
        zl = oracle.prec_solve(Ml, dl, dinvl, rl, "icc")
        assert np.array_equal(zl[:ni], z[g[:ni]])
    # the DIC identity on A without the couplings between ranks
    A = s.scipy().toarray()
    same = parts[:, None] == parts[None, :]
    Ab = np.where(same, A, 0.0)
    L = np.tril(Ab, -1)
    P = np.diag(dinv)
    Mm = (P + L) @ np.diag(1.0 / dinv) @ (P + L.T)
    assert np.allclose(np.diag(Mm), np.diag(A), rtol=1e-13)
    assert np.all(Mm[~same] == 0.0)
    mask = Ab != 0
    assert np.allclose(Mm[mask], Ab[mask], rtol=1e-12, atol=1e-15)
    assert np.allclose(Mm @ z, rhs, atol=1e-12 * np.abs(rhs).max())
    # one rank: the ordinary DIC, to the bit
    d1, dinv1 = oracle.prec_form(M, "icc", parts=np.zeros(s.n, dtype=np.int32))
    d0, dinv0 = oracle.prec_form(M, "icc")
    assert np.array_equal(d1, d0) and np.array_equal(dinv1, dinv0)
    assert not np.array_equal(d, d0)


def test_sum_order_per_rank(cases, oracle):
    """Sum order 11 sums each rank's rows serially and adds the rank totals
    (a two-rank allreduce): checked through fn (Cg.f90:62-76) and the first
    NRMS of a solve that stops before iterating, both closed forms; with
    one rank it is the reference's serial order, to the bit."""
    s = cases.hex_pressure(7, 4, 3, rho_out=10.0)
    n0 = 40
    parts = (np.arange(s.n) >= n0).astype(np.int32)
    dg = s.val[s.dia - 1]

    def serial(v):
        t = 0.0
        for e in v.tolist():
            t = t + e
        return t

    def first_res(order, p):
        x = np.zeros(s.n)
        _, res = oracle.solve("cg", "none", _mat(oracle, s), x, s.b.copy(), 0, 1e-30, norm=1.0,
                              sum_order=order, parts=p)
        return res

    fn = 1.0 / ((serial(dg[:n0]) + serial(dg[n0:])) / s.n)
    r, dd = s.b * fn, dg * fn
    t = (r * r) / (dd * dd)
    assert first_res(11, parts) == np.sqrt(serial(t[:n0]) + serial(t[n0:]))
    assert first_res(11, np.zeros(s.n, dtype=np.int32)) == first_res(0, None)


def test_jacobi_and_none_prec(cases, oracle):
    s = cases.hex_pressure(4, 3, 3, rho_out=9.0)
    M = _mat(oracle, s)
    d, dinv = oracle.prec_form(M, "jacobi")
    diag = s.scipy().diagonal()
    assert np.array_equal(d, diag) and np.array_equal(dinv, 1.0 / diag)
    r = np.arange(s.n, dtype=float)
    assert np.array_equal(oracle.prec_solve(M, d, dinv, r, "jacobi"), r * (1.0 / diag))
    d0, _ = oracle.prec_form(M, "none")
    assert np.all(d0 == 1.0)
    assert np.array_equal(oracle.prec_solve(M, d0, d0, r, "none"), r)


# ------------------------------------------------------------------ BiCG ----

@pytest.mark.parametrize("make,norm", [(lambda c: c.upwind(9), 1.0),
                                       (lambda c: c.hex_pressure(8, rho_out=100.0), None)])
@pytest.mark.parametrize("prec", ["jacobi", "none", "icc"])
def test_bicg_equals_cg_bitwise(cases, oracle, make, norm, prec):
    s = make(cases)
    xa, xb = np.zeros(s.n), np.zeros(s.n)
    ra = oracle.solve("cg", prec, _mat(oracle, s), xa, s.b.copy(), 60, 1e-12, norm=norm)
    rb = oracle.solve("bicg", prec, _mat(oracle, s), xb, s.b.copy(), 60, 1e-12, norm=norm)
    assert ra == rb and np.array_equal(xa, xb)


# ------------------------------------------------------------ conventions ---

def test_zero_rhs_exits_with_zero_residual(cases, oracle):
    s = cases.hex_pressure(5)
    x0 = np.linspace(0.0, 1.0, s.n)
    x = x0.copy()
    assert oracle.solve("cg", "jacobi", _mat(oracle, s), x, np.zeros(s.n), 100, 1e-8) == (0, 0.0)
    assert np.array_equal(x, x0)


def test_exact_start_exits_at_zero_iterations(cases, oracle):
    s = cases.hex_pressure(5)
    xe = np.linalg.solve(s.scipy().toarray(), s.b)
    x = xe.copy()
    it, res = oracle.solve("cg", "jacobi", _mat(oracle, s), x, s.b.copy(), 100, 1e-8)
    assert it == 0 and 0.0 < res < 1e-8 and np.array_equal(x, xe)


@pytest.mark.parametrize("miter", [0, -3])
def test_no_iterations(cases, oracle, miter):
    """Fortran `do iter = 1, miter` with miter < 1 runs no pass; iter - 1 = 0."""
    s = cases.hex_pressure(5)
    A = s.scipy().toarray()
    x = np.zeros(s.n)
    it, res = oracle.solve("cg", "jacobi", _mat(oracle, s), x, s.b.copy(), miter, 1e-8)
    assert it == 0 and np.all(x == 0.0)
    # fin_res = NRMS(b - A x0) with x range 0 -> TINY floor
    assert res == pytest.approx(np.sqrt(np.sum((s.b / np.diag(A)) ** 2)) / 1e-15, rel=1e-12)


def test_values_round_trip(cases, oracle):
    """The solve scales A and b by fn in place and divides back
    (Cg.f90:77-84, 245-250): entries come back within an ulp or two."""
    s = cases.hex_pressure(6, rho_out=100.0)
    M = _mat(oracle, s)
    b = s.b.copy()
    oracle.solve("cg", "jacobi", M, np.zeros(s.n), b, 20, 1e-8)
    assert np.all(np.abs(M.val - s.val) <= 2 * np.spacing(np.abs(s.val)))
    assert np.all(np.abs(b - s.b) <= 2 * np.spacing(np.abs(s.b)))


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6, 7, 8, 9, 10])
def test_reordered_sums_stay_close(cases, oracle, order):
    """Each other summation order (the rounding envelope) solves the same
    problem, and differs from the reference order only by rounding."""
    s = cases.hex_pressure(12, rho_out=10.0)
    xa, xb = np.zeros(s.n), np.zeros(s.n)
    ia, _ = oracle.solve("cg", "jacobi", _mat(oracle, s), xa, s.b.copy(), 1000, 1e-9)
    ib, _ = oracle.solve("cg", "jacobi", _mat(oracle, s), xb, s.b.copy(), 1000, 1e-9,
                         sum_order=order)
    assert abs(ia - ib) <= 1
    assert np.linalg.norm(xa - xb) <= 1e-9 * np.linalg.norm(xa)
    assert not np.array_equal(xa, xb)


@pytest.mark.parametrize("order,ulps", [(7, -1), (8, 1)])
def test_fn_ulp_orders(cases, oracle, order, ulps):
    """Orders 7 and 8 are the reference's order with fn one ulp away: the
    scaled values are the reference's with that fn (checked through the
    first residual of a one-iteration solve with no preconditioner, whose
    NRMS is a closed form in fn)."""
    s = cases.hex_pressure(5, rho_out=10.0)
    fn = 1.0 / (sum(s.val[s.dia - 1].tolist()) / s.n)  # serial sum, Cg.f90:62-76
    d = s.val[s.dia - 1]
    for o, f in ((0, fn), (order, np.nextafter(fn, np.inf if ulps > 0 else 0.0))):
        x = np.zeros(s.n)
        _, res = oracle.solve("cg", "none", _mat(oracle, s), x, s.b.copy(), 0, 1e-30, norm=1.0,
                              sum_order=o)
        # r0 = b fn (x0 = 0); NRMS = sqrt(sum (r_i / (a_ii fn))^2) / 1
        r = s.b * f
        dd = d * f
        exp = np.sqrt(sum(((r * r) / (dd * dd)).tolist()))
        assert res == exp, (o, res, exp)


# ------------------------------------------------------------- fixtures -----

def _load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_golden_fixtures(cases, oracle):
    """Fixtures written by tests/golden/make_golden.py (scipy/numpy only)."""
    g = _load_golden("cg_small.json")
    for case in g["cases"]:
        s = cases.hex_pressure(*case["grid"], rho_out=case["rho_out"], seed=case["seed"])
        assert s.nnz == case["nnz"]
        assert abs(float(np.sum(s.val)) - case["val_sum"]) <= 1e-12 * abs(case["val_sum"]) + 1e-12
        x = np.zeros(s.n)
        it, res = oracle.solve("cg", case["prec"], _mat(oracle, s), x, s.b.copy(), case["miter"],
                               1e-300)
        xg = np.array(case["x_galerkin"])
        assert np.linalg.norm(x - xg) <= 1e-9 * np.linalg.norm(xg)
        xs = np.array(case["x_solve"])
        x = np.zeros(s.n)
        oracle.solve("cg", case["prec"], _mat(oracle, s), x, s.b.copy(), 5000, 1e-14,
                     criterion="relres")
        assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)
