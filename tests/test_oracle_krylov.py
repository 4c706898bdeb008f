"""Pins of the oracle's textbook methods (BiCG, CGS, BiCGSTAB).

Two independent checks, each able to catch a wrong alpha, beta, omega, sign
or preconditioner placement:
  * scipy.sparse.linalg's bicg / cgs / bicgstab (library implementations of
    the same Templates algorithms) run for k = 1..5 iterations from the same
    x0 with the same Jacobi M, compared iterate by iterate;
  * the polynomial characterisation that exact arithmetic fixes, computed
    densely: BiCG's residual r_k = phi_k(B) r0 with B = A M^-1 and the
    Petrov-Galerkin condition r_k _|_ M^-T K_k(A^T M^-T, r~0); CGS's residual
    is phi_k(B)^2 r0 and BiCGSTAB's psi_k(B) phi_k(B) r0 with
    psi_k(t) = prod(1 - omega_i t), phi_k now the BiCG polynomial of B under
    the plain bilinear form (r~0^T B^m r0 moments).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sl

KMAX = 5


def _system(cases):
    s = cases.upwind(6)
    return s, s.scipy().toarray()


def _oracle_iterate(oracle, s, method, prec, k, omegas=None):
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    x = np.zeros(s.n)
    it, _ = oracle.solve(method, prec, A, x, s.b.copy(), k, 1e-300, criterion="relres",
                         omegas=omegas)
    assert it == k
    return x


def _scipy_iterate(method, Ad, b, dinv, k):
    n = len(b)
    M = sl.LinearOperator((n, n), matvec=lambda v: v * dinv, rmatvec=lambda v: v * dinv,
                          dtype=np.float64)
    f = {"bicg_textbook": sl.bicg, "cgs": sl.cgs, "bcgs": sl.bicgstab}[method]
    x, info = f(Ad, b, x0=np.zeros(n), rtol=0.0, atol=0.0, maxiter=k, M=M)
    assert info == k
    return x


@pytest.mark.parametrize("method", ["bicg_textbook", "cgs", "bcgs"])
@pytest.mark.parametrize("prec", ["jacobi", "none"])
def test_iterates_match_scipy(cases, oracle, method, prec):
    s, Ad = _system(cases)
    dinv = 1.0 / np.diag(Ad) if prec == "jacobi" else np.ones(s.n)
    for k in range(1, KMAX + 1):
        xo = _oracle_iterate(oracle, s, method, prec, k)
        xs = _scipy_iterate(method, Ad, s.b, dinv, k)
        rel = np.linalg.norm(xo - xs) / np.linalg.norm(xs)
        assert rel < 1e-10, (method, prec, k, rel)


def _phi(B, r0, w0, k, M_inv=None):
    """Coefficients c_1..c_k of r_k = r0 - sum_j c_j B^j r0 with
    (w0^T [M^-1] B^(i+j) r0) moments: the BiCG residual polynomial."""
    V = [r0]
    for _ in range(2 * k):
        V.append(B @ V[-1])
    left = w0 if M_inv is None else M_inv @ w0
    mom = np.array([left @ v for v in V])
    G = np.array([[mom[i + j] for j in range(1, k + 1)] for i in range(k)])
    return np.linalg.solve(G, mom[:k])


def _poly_apply(B, coeffs, v):
    """sum_m coeffs[m] B^m v."""
    out = np.zeros_like(v)
    w = v.copy()
    for c in coeffs:
        out += c * w
        w = B @ w
    return out


@pytest.mark.parametrize("prec", ["jacobi", "none"])
def test_polynomial_characterisation(cases, oracle, prec):
    s, Ad = _system(cases)
    n = s.n
    Minv = np.diag(1.0 / np.diag(Ad)) if prec == "jacobi" else np.eye(n)
    B = Ad @ Minv
    r0 = s.b.copy()                 # x0 = 0
    for k in range(1, 5):
        # BiCG (Templates' preconditioned form): r_k = phi_k(B) r0 with the
        # M^-1-weighted moments; x_k = M^-1 sum_j c_j B^(j-1) r0
        c = _phi(B, r0, r0, k, M_inv=Minv)
        xk = Minv @ _poly_apply(B, np.concatenate([c, [0.0]]), r0)
        xo = _oracle_iterate(oracle, s, "bicg_textbook", prec, k)
        assert np.linalg.norm(xo - xk) / np.linalg.norm(xk) < 1e-8, ("bicg", k)
        # CGS and BiCGSTAB: phi_k under the plain moments r0^T B^m r0
        c = _phi(B, r0, r0, k)
        phi = np.concatenate([[1.0], -c])
        r_cgs = _poly_apply(B, np.convolve(phi, phi), r0)
        x_cgs = np.linalg.solve(Ad, r0 - r_cgs)
        xo = _oracle_iterate(oracle, s, "cgs", prec, k)
        assert np.linalg.norm(xo - x_cgs) / np.linalg.norm(x_cgs) < 1e-8, ("cgs", k)
        om = np.zeros(KMAX)
        xo = _oracle_iterate(oracle, s, "bcgs", prec, k, omegas=om)
        # omega is reported for the normalised system (A*fn); with M = I its
        # B is fn*B, so omega*fn multiplies the unscaled B (Jacobi: B = A D^-1
        # is scale-free)
        fn = 1.0 / np.mean(np.diag(Ad)) if prec == "none" else 1.0
        psi = np.array([1.0])
        for w in om[:k]:
            psi = np.convolve(psi, [1.0, -w * fn])
        r_bs = _poly_apply(B, np.convolve(psi, phi), r0)
        x_bs = np.linalg.solve(Ad, r0 - r_bs)
        assert np.linalg.norm(xo - x_bs) / np.linalg.norm(x_bs) < 1e-8, ("bcgs", k)


@pytest.mark.parametrize("method", ["bicg_textbook", "cgs", "bcgs"])
def test_converges_to_dense_solution(cases, oracle, method):
    s, Ad = _system(cases)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    x = np.zeros(s.n)
    it, res = oracle.solve(method, "jacobi", A, x, s.b.copy(), 500, 1e-12, criterion="relres")
    xd = np.linalg.solve(Ad, s.b)
    assert it < 500 and res < 1e-12
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_bicg_textbook_equals_cg_on_symmetric(cases, oracle):
    """On a symmetric matrix with r~0 = r0 and symmetric M, textbook BiCG's
    shadow sequence equals the primal one: same iterates as CG (up to the
    rounding of the transposed product's values, which are equal here)."""
    s = cases.hex_pressure(6)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    xb, xc = np.zeros(s.n), np.zeros(s.n)
    ib, rb = oracle.solve("bicg_textbook", "jacobi", A, xb, s.b.copy(), 50, 1e-10)
    ic, rc = oracle.solve("cg", "jacobi", A, xc, s.b.copy(), 50, 1e-10)
    assert ib == ic and np.array_equal(xb, xc) and rb == rc


def test_conventions(cases, oracle):
    s, _ = _system(cases)
    for method in ("bicg_textbook", "cgs", "bcgs"):
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        x = np.zeros(s.n)
        assert oracle.solve(method, "jacobi", A, x, np.zeros(s.n), 10, 1e-8) == (0, 0.0)
        x = np.zeros(s.n)
        it, res = oracle.solve(method, "jacobi", A, x, s.b.copy(), 0, 1e-8)
        assert it == 0 and res > 0
        x = np.zeros(s.n)
        it, _ = oracle.solve(method, "jacobi", A, x, s.b.copy(), 3, 1e-30)
        assert it == 3
        # values restored bit for bit within the reference's 1-ulp scaling
        assert np.allclose(A.val, s.val, rtol=4e-16, atol=0)


def test_transpose_positions(cases, oracle):
    s = cases.knn(4, 4, 4)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    t = oracle.transpose_positions(A)
    S = s.scipy()
    rows = np.repeat(np.arange(s.n), np.diff(s.row))
    cols = s.col - 1
    assert np.array_equal(rows[t], cols) and np.array_equal(cols[t], rows)
    assert S.shape == (s.n, s.n)
