"""FP32 sweeps inside FP64 iterative refinement (an extension: the reference
has no mixed-precision solver).  Every run states its error against a
tightly converged FP64 oracle solve of the same system.  The bars: the
requested tolerance on the FP64 TRUE residual; and, since restarted FP32
sweeps leave more error in the near-null (smooth) modes of these pinned
pressure systems than FP64 CG does at the same residual, that a tolerance
1000x tighter reaches the error FP64 CG reaches at the nominal one ("the
same FP64-oracle error")."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_x(oracle, s, tol=1e-13):
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    x = np.zeros(s.n)
    it, _ = oracle.solve("cg", "jacobi", A, x, s.b.copy(), 50000, tol, criterion="relres")
    return x, it


def _ir(gpu, s, tol, criterion=1, inner_max=0, inner_tol=0.0, prec="jacobi", miter=20000,
        norm=None, nat=None, b=None):
    nat = nat or gpu.Native(s.row, s.col, s.dia)
    opt = gpu.Options()
    opt.precision = gpu.FP32_IR
    opt.criterion = criterion
    opt.inner_max = inner_max
    opt.inner_tol = inner_tol
    opt.true_residual = 1
    x = np.zeros(s.n)
    it, res, rep = nat.solve_native("cg", prec, s.val.copy(), x, (s.b if b is None else b).copy(),
                                    miter, tol, norm=norm, options=opt, report=True)
    return it, res, x, rep


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("make,tol", [(lambda c: c.hex_pressure(40), 1e-10),
                                      (lambda c: c.hex_pressure(32, rho_out=1000.0), 1e-8),
                                      (lambda c: c.knn(14, 14, 14), 1e-10),
                                      (lambda c: c.upwind(16), 1e-8)],
                         ids=["poisson40", "twophase32", "knn14", "upwind16"])
def test_ir_reaches_fp64_tolerance(cases, oracle, gpu, make, tol):
    s = make(cases)
    it, res, x, rep = _ir(gpu, s, tol)
    true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
    if "upwind" in s.name:   # non-symmetric: reference = dense direct solve
        xe = np.linalg.solve(s.scipy().toarray(), s.b)
        err, err64, it64 = _rel(x, xe), None, None
    else:
        xe, _ = _oracle_x(oracle, s)
        x64, it64 = _oracle_x(oracle, s, tol)
        err, err64 = _rel(x, xe), _rel(x64, xe)
    print(f"{s.name}: FP32 iterations {it} (FP64 oracle {it64}), refinements "
          f"{rep['refinements']}, relres {res:.2e}, true {true:.2e}, error vs FP64 oracle "
          f"{err:.2e} (FP64 CG at the same tol: {err64})", flush=True)
    assert res < tol and true < 1.01 * tol
    assert abs(rep["true_relres"] - true) <= 1e-3 * true
    assert rep["refinements"] >= 1
    if err64 is not None:
        it2, _, x2, rep2 = _ir(gpu, s, tol * 1e-3)
        err2 = _rel(x2, xe)
        print(f"  at tol {tol * 1e-3:.0e}: {it2} FP32 iterations, error {err2:.2e}", flush=True)
        assert err2 <= 2 * err64 + 1e-12


def test_ir_inner_restart_every_50(cases, oracle, gpu):
    """Inner CG restarted every 50 iterations (the unstructured case's recipe)."""
    s = cases.knn(12, 12, 12)
    it, res, x, rep = _ir(gpu, s, 1e-10, inner_max=50)
    true = np.linalg.norm(s.b - s.scipy() @ x) / np.linalg.norm(s.b)
    xe, _ = _oracle_x(oracle, s)
    x64, _ = _oracle_x(oracle, s, 1e-10)
    err, err64 = _rel(x, xe), _rel(x64, xe)
    print(f"restart 50: {it} iterations, {rep['refinements']} refinements, true {true:.2e}, "
          f"err {err:.2e} (FP64 CG at 1e-10: {err64:.2e})")
    assert true < 1.01e-10 and rep["refinements"] >= it // 50
    assert err < 1e-6


def test_ir_nrms_criterion(cases, oracle, gpu):
    s = cases.hex_pressure(24, rho_out=100.0)
    it, res, x, rep = _ir(gpu, s, 1e-9, criterion=0)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    r = s.b - s.scipy() @ x
    nr = np.sqrt(np.sum((r / s.scipy().diagonal()) ** 2)) / (x.max() - x.min())
    # the residual is ~1e-11 of b here, so its rounding is ~1e-5 relative
    assert res < 1e-9 and abs(nr - res) <= 1e-4 * res + 1e-20
    assert oracle.nrms(A, r, x) == pytest.approx(nr, rel=1e-9)


def test_ir_zero_rhs_and_cap(cases, gpu):
    s = cases.hex_pressure(10)
    it, res, x, rep = _ir(gpu, s, 1e-8, b=np.zeros(s.n))
    assert it == 0 and res == 0.0 and np.all(x == 0.0)
    it, res, x, rep = _ir(gpu, s, 1e-14, miter=7)
    assert it == 7 and res > 1e-14
