"""Internal reverse Cuthill-McKee ordering (tfg_matrix_reorder) on a
randomly numbered kNN matrix: the caller's numbering is kept at every call;
products, residuals and the Jacobi step are bit-identical to the oracle's
(each row keeps its entries' order), solves meet the same bar as every
other engine (only the dot products' order changes), and the orderings an
ordering-dependent method needs are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _knn(cases):
    return cases.knn(16, 14, 12)


def test_products_bitexact(cases, oracle, gpu):
    s = _knn(cases)
    print(f"reorder case knn n={s.n} nnz={s.nnz}", flush=True)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    p = np.random.default_rng(2).standard_normal(s.n)
    nat = gpu.Native(s.row, s.col, s.dia)
    nat.reorder(gpu.ORDER_RCM)
    dv, dp, db = gpu.DeviceArray(s.val), gpu.DeviceArray(p), gpu.DeviceArray(s.b)
    dq = gpu.DeviceArray(np.zeros(s.n))
    nat.spmv(dv, dp, dq)
    assert np.array_equal(dq.download(), oracle.spmv(A, p))
    nat.residual(dv, db, dp, dq)
    assert np.array_equal(dq.download(), oracle.residual(A, s.b, p))
    nat.prec_apply("jacobi", dv, dp, dq)
    d, dinv = oracle.prec_form(A, "jacobi")
    assert np.array_equal(dq.download(), oracle.prec_solve(A, d, dinv, p, "jacobi"))


@pytest.mark.parametrize("solver,crit", [("cg", 0), ("cg", 1), ("bicg", 0)])
def test_solve_parity(cases, oracle, gpu, solver, crit):
    s = _knn(cases)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    xo = np.zeros(s.n)
    it_o, res_o = oracle.solve(solver, "jacobi", A, xo, s.b.copy(), 2000, 1e-9,
                               criterion="relres" if crit else "nrms")
    nat = gpu.Native(s.row, s.col, s.dia)
    nat.reorder(gpu.ORDER_RCM)
    o = gpu.Options()
    o.criterion = crit
    x = np.zeros(s.n)
    it, res = nat.solve_native(solver, "jacobi", s.val.copy(), x, s.b.copy(), 2000, 1e-9,
                               options=o)
    print(f"  {solver} crit {crit}: gpu {it} its {res:.3e}; oracle {it_o} its {res_o:.3e}",
          flush=True)
    assert abs(it - it_o) <= 2 and res < 1e-9
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    # the device-pointer path gives the same bits as the host-pointer one
    dv, db, dx = gpu.DeviceArray(s.val), gpu.DeviceArray(s.b), gpu.DeviceArray(np.zeros(s.n))
    it2, res2 = nat.solve_native_device(solver, "jacobi", dv, dx, db, 2000, 1e-9, options=o)
    assert (it2, res2) == (it, res) and np.array_equal(dx.download(), x)


def test_fp32_refinement_and_textbook_methods(cases, oracle, gpu):
    s = _knn(cases)
    nat = gpu.Native(s.row, s.col, s.dia)
    nat.reorder(gpu.ORDER_RCM)
    o = gpu.Options()
    o.precision = gpu.FP32_IR
    o.criterion = gpu.CRIT_RELRES
    x = np.zeros(s.n)
    it, res = nat.solve_native("cg", "jacobi", s.val.copy(), x, s.b.copy(), 5000, 1e-10,
                               options=o)
    r = s.b - s.scipy() @ x
    assert res < 1e-10 and np.linalg.norm(r) <= 1e-9 * np.linalg.norm(s.b)
    u = cases.upwind(12)
    nu = gpu.Native(u.row, u.col, u.dia)
    nu.reorder(gpu.ORDER_RCM)
    for method in ("cgs", "bcgs"):
        A = oracle.Matrix(u.row, u.col, u.dia, u.val)
        xo = np.zeros(u.n)
        it_o, _ = oracle.solve(method, "jacobi", A, xo, u.b.copy(), 200, 1e-10, norm=1.0)
        xg = np.zeros(u.n)
        itg, resg = nu.solve_native(method, "jacobi", u.val.copy(), xg, u.b.copy(), 200, 1e-10,
                                    norm=1.0)
        assert abs(itg - it_o) <= 2 and resg < 1e-10
        assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)


def test_caller_order_paths_and_back_to_natural(cases, oracle, gpu):
    """ICC (ordering-dependent) runs in the caller's order on a reordered
    handle and gives the bits of a handle without ordering; dropping the
    ordering gives the natural-order solve's bits back."""
    s = _knn(cases)
    nat = gpu.Native(s.row, s.col, s.dia)
    assert nat.info()["ordering"] == gpu.ORDER_NATURAL
    x0 = np.zeros(s.n)
    ref = nat.solve_native("cg", "jacobi", s.val.copy(), x0, s.b.copy(), 500, 1e-8)
    xi0 = np.zeros(s.n)
    ref_icc = nat.solve_native("cg", "icc", s.val.copy(), xi0, s.b.copy(), 300, 1e-8)
    nat.reorder(gpu.ORDER_RCM)
    assert nat.info()["ordering"] == gpu.ORDER_RCM
    xi = np.zeros(s.n)
    assert nat.solve_native("cg", "icc", s.val.copy(), xi, s.b.copy(), 300, 1e-8) == ref_icc
    assert np.array_equal(xi, xi0)
    with pytest.raises(gpu.TFlowsGpuError) as e:
        nat.reorder(7)
    assert e.value.code == gpu.TFG_EINVAL
    nat.reorder(gpu.ORDER_NATURAL)
    x1 = np.zeros(s.n)
    assert nat.solve_native("cg", "jacobi", s.val.copy(), x1, s.b.copy(), 500, 1e-8) == ref
    assert np.array_equal(x1, x0)


def test_automatic_ordering_of_a_random_numbering(cases, oracle, gpu):
    """tfg_matrix_create orders a large randomly numbered matrix by itself,
    not a structured one; the solve meets the bar either way."""
    s = cases.knn(44, 42, 40)        # 73920 rows, randomly numbered
    nat = gpu.Native(s.row, s.col, s.dia)
    print(f"auto-order case n={s.n} nnz={s.nnz}: {nat.info()}", flush=True)
    assert nat.info()["ordering"] == gpu.ORDER_RCM
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    xo = np.zeros(s.n)
    it_o, _ = oracle.solve("cg", "jacobi", A, xo, s.b.copy(), 3000, 1e-9)
    x = np.zeros(s.n)
    it, res = nat.solve_native("cg", "jacobi", s.val.copy(), x, s.b.copy(), 3000, 1e-9)
    assert abs(it - it_o) <= 2 and res < 1e-9
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    h = cases.hex_pressure(42, 42, 40)
    assert gpu.Native(h.row, h.col, h.dia).info()["ordering"] == gpu.ORDER_NATURAL
