"""The oracle's restatement of the FP32 refinement form (no counterpart in
the reference), pinned by what the mathematics fixes:

* the row-excess product is A x rewritten: on integer-valued matrices and
  vectors, where FP32 arithmetic is exact, it equals the dense product; a
  vector constant over the cells gives s_i x exactly;
* the row sums are compensated: a row whose entries cancel gets its exact
  sum, where a plain left-to-right sum keeps a rounding error;
* the solve converges to the dense solution (jacobi, icc, none), stops on
  the FP64 true residual, counts iterations as documented, and its coarse
  correction leaves the residual orthogonal to the constant vector;
* its first step is the FP64 method's first iterate alpha M^-1 r0 (ICC:
  the DIC sweeps of the widened FP32 residual);
* planted mistakes (a wrong sign in the product's differences, a plain
  FP32 diagonal instead of the row sum) fail these checks.
"""
import numpy as np
import pytest


def _mat(oracle, s):
    return oracle.Matrix(s.row, s.col, s.dia, s.val.copy())


def _int_system(n=30, seed=3):
    """A random symmetric integer matrix on a ring-plus-chords pattern,
    diagonal first, columns ascending (1-based CSR)."""
    rng = np.random.default_rng(seed)
    nb = {i: set() for i in range(n)}
    for i in range(n):
        for j in ((i + 1) % n, (i + 7) % n):
            nb[i].add(j)
            nb[j].add(i)
    w = {}
    for i in range(n):
        for j in nb[i]:
            if (j, i) in w:
                w[(i, j)] = w[(j, i)]
            else:
                w[(i, j)] = -float(rng.integers(1, 9))
    row, col, val, dia = [1], [], [], []
    for i in range(n):
        dia.append(len(col) + 1)
        col.append(i + 1)
        val.append(-sum(w[(i, j)] for j in nb[i]) + float(rng.integers(0, 3)))
        for j in sorted(nb[i]):
            col.append(j + 1)
            val.append(w[(i, j)])
        row.append(len(col) + 1)
    return (np.array(row, np.int32), np.array(col, np.int32), np.array(dia, np.int32),
            np.array(val))


def test_lap32_product_is_the_matrix_product(oracle):
    row, col, dia, val = _int_system()
    A = oracle.Matrix(row, col, dia, val)
    n = len(row) - 1
    dense = np.zeros((n, n))
    for i in range(n):
        for j in range(row[i] - 1, row[i + 1] - 1):
            dense[i, col[j] - 1] = val[j]
    v32 = oracle.lap32_form(A, 1.0)
    rng = np.random.default_rng(1)
    for _ in range(3):
        x = rng.integers(-20, 21, n).astype(np.float32)
        y = oracle.lap32_spmv(A, v32, x)
        assert np.array_equal(y.astype(np.float64), dense @ x.astype(np.float64))
    # the diagonal slot holds the row sum; a constant gives s_i c exactly
    s = dense.sum(axis=1)
    assert np.array_equal(v32[dia - 1].astype(np.float64), s)
    c = np.full(n, 0.3, dtype=np.float32)
    assert np.array_equal(oracle.lap32_spmv(A, v32, c), (s.astype(np.float32) * np.float32(0.3)))
    # a planted sign error in the differences, or the plain diagonal in the
    # diagonal slot, breaks the identity
    x = rng.integers(-20, 21, n).astype(np.float32)
    y = oracle.lap32_spmv(A, v32, x).astype(np.float64)
    wrong_sign = 2 * v32[dia - 1] * x - y          # s x - sum a (x_j - x_i)
    assert not np.array_equal(wrong_sign, dense @ x)
    v_plain = v32.copy()
    v_plain[dia - 1] = val[dia - 1].astype(np.float32)
    assert not np.array_equal(oracle.lap32_spmv(A, v_plain, x).astype(np.float64), dense @ x)


def test_row_sums_are_compensated(oracle):
    # one row whose entries cancel to below their rounding: 1e16 + 1 - 1e16
    # = 1 exactly; a left-to-right sum gives 0 (or 2)
    row = np.array([1, 4, 6, 8], np.int32)
    col = np.array([1, 2, 3, 1, 2, 1, 3], np.int32)
    dia = np.array([1, 4, 6], np.int32)
    val = np.array([1.0, 1e16, -1e16, 1e16, 5.0, -1e16, 7.0])
    A = oracle.Matrix(row, col, dia, val)
    v32 = oracle.lap32_form(A, 1.0)
    assert v32[0] == 1.0
    assert (1.0 + 1e16) - 1e16 != 1.0   # what a plain sum would give


def test_ir_solve_converges_to_the_dense_solution(cases, oracle):
    for s, deflate in ((cases.hex_pressure(8, rho_out=1000.0), 1),
                       (cases.hex_pressure(7, 6, 5), 0),
                       (cases.knn(6, 6, 6), 1)):
        xd = np.linalg.solve(s.scipy().toarray(), s.b)
        for prec in ("jacobi", "none", "icc"):
            x = np.zeros(s.n)
            it, res, nref = oracle.ir_solve(prec, _mat(oracle, s), x, s.b.copy(), 5000, 1e-10,
                                            criterion="relres", delta=1e-3, check=8,
                                            deflate=deflate)
            true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
            # fin_res is the FP64 true residual's measure (recomputed on the
            # host in another order: ~1e-14 of b apart); niter a multiple of
            # the check interval (or the cap)
            assert res < 1e-10 and true < 1.01e-10 and abs(res - true) <= 1e-2 * true
            assert it % 8 == 0 and nref >= 1
            assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd), (s.name, prec)


@pytest.mark.parametrize("prec", ["icc", "jacobi"])
def test_ir_first_step_is_the_fp64_methods(cases, oracle, prec):
    """One FP32 iteration and its refinement give x1 = alpha z0 with
    z0 = M^-1 r0: the FP64 method's first iterate (Cg.f90 with one
    iteration) to FP32 accuracy, and far from the other preconditioner's
    (a missing sweep or the Jacobi z in place of the DIC fails)."""
    s = cases.hex_pressure(9, 7, 6, rho_out=30.0)
    x = np.zeros(s.n)
    oracle.ir_solve(prec, _mat(oracle, s), x, s.b.copy(), 1, 1e-30, criterion="relres", check=1)
    firsts = {}
    for p in ("icc", "jacobi"):
        xf = np.zeros(s.n)
        oracle.solve("cg", p, _mat(oracle, s), xf, s.b.copy(), 1, 1e-30, criterion="relres")
        firsts[p] = xf
    other = "jacobi" if prec == "icc" else "icc"
    err = np.linalg.norm(x - firsts[prec]) / np.linalg.norm(firsts[prec])
    far = np.linalg.norm(x - firsts[other]) / np.linalg.norm(firsts[other])
    assert err < 1e-6 and far > 1e-2, (err, far)


def test_ir_solve_conventions(cases, oracle):
    s = cases.hex_pressure(6)
    x = np.zeros(s.n)
    assert oracle.ir_solve("jacobi", _mat(oracle, s), x, np.zeros(s.n), 100, 1e-8,
                           criterion="relres") == (0, 0.0, 0)
    for miter in (0, 5):
        x = np.zeros(s.n)
        it, res, nref = oracle.ir_solve("jacobi", _mat(oracle, s), x, s.b.copy(), miter, 1e-30,
                                        criterion="relres", check=8)
        true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
        assert it == miter and abs(res - true) <= 1e-9 * true
        assert nref == (1 if miter else 0)


def test_coarse_correction_orthogonalises_the_residual(cases, oracle):
    """With deflation the start's coarse correction makes the residual sum
    to zero (w'r = 0), and the deflated FP32 directions keep it so: after 8
    iterations the FP64 residual's sum is rounding-sized, where without
    deflation it is not."""
    s = cases.hex_pressure(10, rho_out=1000.0)
    sums = {}
    for deflate in (0, 1):
        x = np.zeros(s.n)
        oracle.ir_solve("jacobi", _mat(oracle, s), x, s.b.copy(), 8, 1e-30, criterion="relres",
                        check=8, deflate=deflate)
        r = s.b - s.scipy() @ x
        sums[deflate] = abs(r.sum()) / np.abs(s.b).sum()
    assert sums[1] < 1e-6 * sums[0]


def test_deflation_resolves_the_near_null_mode(cases, oracle):
    """Without deflation the FP32 sweeps leave an error along the near-null
    (constant) mode that the residual cannot see; with it the error meets
    FP64 CG's at the same tolerance."""
    s = cases.hex_pressure(16, rho_out=1000.0)
    xe = np.zeros(s.n)
    oracle.solve("cg", "jacobi", _mat(oracle, s), xe, s.b.copy(), 50000, 1e-14, criterion="relres")
    x64 = np.zeros(s.n)
    oracle.solve("cg", "jacobi", _mat(oracle, s), x64, s.b.copy(), 50000, 1e-9, criterion="relres")
    err = {}
    for deflate in (0, 1):
        x = np.zeros(s.n)
        it, res, nref = oracle.ir_solve("jacobi", _mat(oracle, s), x, s.b.copy(), 20000, 1e-9,
                                        criterion="relres", delta=1e-3, check=8, deflate=deflate)
        assert res < 1e-9
        err[deflate] = np.linalg.norm(x - xe) / np.linalg.norm(xe)
    e64 = np.linalg.norm(x64 - xe) / np.linalg.norm(xe)
    assert err[1] <= 3 * e64
    assert err[0] > 3 * err[1]
