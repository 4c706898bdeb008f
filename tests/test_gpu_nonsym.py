"""Textbook BiCG / CGS / BiCGSTAB on the GPU against the oracle.

Same bar as the FP64 CG drop-in: iteration count within +-2 of the oracle's,
x within 1e-10 relative of the oracle's, or within the spread that
reordering the oracle's own global sums causes (parity_bar.py; the GPU's dot
products use a different, fixed order; every product row and element-wise
update is the same arithmetic).
"""
import numpy as np
import pytest

import parity_bar

pytestmark = pytest.mark.gpu

X_TOL = 1e-10
GPU_NAME = {"bicg_textbook": "bicg", "cgs": "cgs", "bcgs": "bcgs"}


def _opt(gpu, method, criterion=0):
    o = gpu.Options()
    o.criterion = criterion
    o.true_residual = 1
    o.textbook_bicg = 1 if method == "bicg_textbook" else 0
    return o


def _gpu(gpu, s, method, prec, miter, tol, norm=None, criterion=0, nat=None):
    nat = nat or gpu.Native(s.row, s.col, s.dia)
    x = np.zeros(s.n)
    it, res, rep = nat.solve_native(GPU_NAME[method], prec, s.val.copy(), x, s.b.copy(), miter,
                                    tol, norm=norm, options=_opt(gpu, method, criterion),
                                    report=True)
    return it, res, x, rep


def _oracle(oracle, s, method, prec, miter, tol, norm=None, criterion="nrms", sum_order=0):
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    x = np.zeros(s.n)
    it, res = oracle.solve(method, prec, A, x, s.b.copy(), miter, tol, norm=norm,
                           criterion=criterion, sum_order=sum_order)
    return it, res, x


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = [
    # name, generator, norm, miter, tol, criterion
    ("upwind16_norm1", lambda c: c.upwind(16), 1.0, 500, 1e-8, "nrms"),
    ("upwind20_relres", lambda c: c.upwind(20), None, 500, 1e-10, "relres"),
    ("upwind24_nrms", lambda c: c.upwind(24), None, 500, 1e-8, "nrms"),
]


@pytest.mark.parametrize("method", ["bicg_textbook", "cgs", "bcgs"])
@pytest.mark.parametrize("prec", ["jacobi", "none"])
@pytest.mark.parametrize("name,make,norm,miter,tol,crit", CASES, ids=[c[0] for c in CASES])
def test_parity(cases, oracle, gpu, method, prec, name, make, norm, miter, tol, crit):
    s = make(cases)
    print(f"case {name} {method}/{prec} n={s.n} nnz={s.nnz}", flush=True)
    ci = 0 if crit == "nrms" else 1
    it_o, res_o, x_o = _oracle(oracle, s, method, prec, miter, tol, norm, crit)
    env, spread = parity_bar.envelope(
        lambda so: _oracle(oracle, s, method, prec, miter, tol, norm, crit, sum_order=so)[::2],
        it_o, x_o)
    it_g, res_g, x_g, rep = _gpu(gpu, s, method, prec, miter, tol, norm, ci)
    print(f"  oracle niter={it_o} res={res_o:.3e}; gpu niter={it_g} res={res_g:.3e}; "
          f"true relres {rep['true_relres']:.2e}", flush=True)
    assert it_o < miter
    parity_bar.check(f"{method} {prec} {name}", it_g, x_g, it_o, x_o, env, spread)
    assert res_g < tol


@pytest.mark.parametrize("method", ["bicg_textbook", "cgs", "bcgs"])
def test_first_iterations(cases, oracle, gpu, method):
    """After 1, 2, 3 iterations the GPU iterate equals the oracle's up to the
    order of the global sums (no stopping test involved)."""
    s = cases.upwind(12)
    nat = gpu.Native(s.row, s.col, s.dia)
    for k in (1, 2, 3):
        it_o, _, x_o = _oracle(oracle, s, method, "jacobi", k, 1e-300)
        it_g, _, x_g, _ = _gpu(gpu, s, method, "jacobi", k, 1e-300, nat=nat)
        assert it_o == it_g == k
        assert _rel(x_g, x_o) < 1e-12, (method, k, _rel(x_g, x_o))


def test_reference_bicg_unchanged(cases, gpu):
    """Without textbook_bicg, 'bicg' stays the reference's (the CG recurrence)."""
    s = cases.upwind(10)
    nat = gpu.Native(s.row, s.col, s.dia)
    xa, xb = np.zeros(s.n), np.zeros(s.n)
    ra = nat.solve_native("bicg", "jacobi", s.val.copy(), xa, s.b.copy(), 30, 1e-8, norm=1.0)
    rb = nat.solve_native("cg", "jacobi", s.val.copy(), xb, s.b.copy(), 30, 1e-8, norm=1.0)
    assert ra == rb and np.array_equal(xa, xb)


def test_deterministic(cases, gpu):
    s = cases.upwind(14)
    nat = gpu.Native(s.row, s.col, s.dia)
    out = []
    for _ in range(2):
        x = np.zeros(s.n)
        r = nat.solve_native("bcgs", "jacobi", s.val.copy(), x, s.b.copy(), 200, 1e-10, norm=1.0)
        out.append((r, x))
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])


def test_conventions(cases, oracle, gpu):
    s = cases.upwind(8)
    nat = gpu.Native(s.row, s.col, s.dia)
    for method in ("bicg_textbook", "cgs", "bcgs"):
        opt = _opt(gpu, method)
        x = np.zeros(s.n)
        assert nat.solve_native(GPU_NAME[method], "jacobi", s.val.copy(), x, np.zeros(s.n), 10,
                                1e-8, options=opt) == (0, 0.0)
        for miter in (0, 1, 4):
            x = np.zeros(s.n)
            it_g, res_g = nat.solve_native(GPU_NAME[method], "jacobi", s.val.copy(), x,
                                           s.b.copy(), miter, 1e-30, options=opt)
            it_o, res_o, _ = _oracle(oracle, s, method, "jacobi", miter, 1e-30)
            assert it_g == it_o == miter
            assert abs(res_g - res_o) <= 1e-10 * res_o


def test_unsupported_combinations(cases, gpu):
    s = cases.upwind(6)
    nat = gpu.Native(s.row, s.col, s.dia)
    o = gpu.Options()
    o.precision = gpu.FP32_IR
    # ICC runs with cg / bicg (FP64 or the FP32 form), not the textbook methods
    with pytest.raises(gpu.TFlowsGpuError) as e:
        nat.solve_native("cgs", "icc", s.val.copy(), np.zeros(s.n), s.b.copy(), 5, 1e-8,
                         options=o)
    assert e.value.code == gpu.TFG_ENOTSUP
    with pytest.raises(gpu.TFlowsGpuError) as e:
        nat.solve_native("bcgs", "icc", s.val.copy(), np.zeros(s.n), s.b.copy(), 5, 1e-8)
    assert e.value.code == gpu.TFG_ENOTSUP


def test_textbook_bicg_needs_symmetric_pattern(gpu):
    # 3x3, diagonal first; entry (1,2) without (2,1)
    row = np.array([1, 3, 4, 5], dtype=np.int32)
    col = np.array([1, 2, 2, 3], dtype=np.int32)
    dia = np.array([1, 3, 4], dtype=np.int32)
    val = np.array([4.0, -1.0, 4.0, 4.0])
    nat = gpu.Native(row, col, dia)
    o = gpu.Options()
    o.textbook_bicg = 1
    with pytest.raises(gpu.TFlowsGpuError) as e:
        nat.solve_native("bicg", "jacobi", val, np.zeros(3), np.ones(3), 5, 1e-8, options=o)
    assert e.value.code == gpu.TFG_ENOTSUP
    # CGS and BiCGSTAB need no transpose
    x = np.zeros(3)
    it, res = nat.solve_native("cgs", "jacobi", val.copy(), x, np.ones(3), 20, 1e-12,
                               options=gpu.Options())
    assert np.allclose(np.array([[4, -1, 0], [0, 4, 0], [0, 0, 4.0]]) @ x, 1.0, atol=1e-10)


def test_upwind120_relres(cases, gpu):
    """The 1.7M-cell upwind convection-diffusion system with all three
    methods to ||r||/||b|| < 1e-8, checked on the FP64 true residual."""
    s = cases.upwind(120)
    nat = gpu.Native(s.row, s.col, s.dia)
    for method in ("bicg_textbook", "cgs", "bcgs"):
        it, res, x, rep = _gpu(gpu, s, method, "jacobi", 2000, 1e-8, criterion=1, nat=nat)
        print(f"upwind120 {method}: {it} iterations, relres {res:.2e}, "
              f"true {rep['true_relres']:.2e}, {rep['solve_ms']:.1f} ms", flush=True)
        assert it < 2000 and res < 1e-8 and rep["true_relres"] < 1.01e-8
