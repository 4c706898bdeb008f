"""The allocation-free solve (tfg_solve_native_noalloc; no counterpart in the
reference) through the C ABI, against the existing device-pointer solve
(tfg_solve_native_ex) on the same handle.

Both run the same kernels with the same SpMV engine and launch shapes, so
the bar is bitwise: the same niter, fin_res and x bits.  Two calls in a row
share the GPU's one scratch array, the second on a smaller system, so
contents the first call left behind (partial sums, the symmetric engine's
diagonals and their zero gaps) must not reach the second.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(gpu, s, solver, prec, miter, tol, criterion=0, norm=None, kind=None, nat=None):
    nat = nat or gpu.Native(s.row, s.col, s.dia)
    if kind is not None:
        nat.spmv_select(kind)
    dv, db = gpu.DeviceArray(s.val), gpu.DeviceArray(s.b)
    opt = gpu.Options()
    opt.criterion = criterion
    opt.true_residual = 1
    dx = gpu.DeviceArray(np.zeros(s.n))
    it_e, res_e, rep_e = nat.solve_native_device(solver, prec, dv, dx, db, miter, tol, norm=norm,
                                                 options=opt, report=True)
    x_e = dx.download()
    dx2 = gpu.DeviceArray(np.zeros(s.n))
    it_n, res_n, rep_n = nat.solve_native_noalloc(solver, prec, dv, dx2, db, miter, tol, norm=norm,
                                                  options=opt, report=True)
    x_n = dx2.download()
    info = nat.info()
    print(f"NOALLOC n={s.n} nnz={s.nnz} {solver}/{prec} engine {info['spmv_kind']}: "
          f"ex {it_e} {res_e:.3e} noalloc {it_n} {res_n:.3e} true relres "
          f"{rep_n['true_relres']:.3e} max|dx| {np.abs(x_n - x_e).max():.1e}", flush=True)
    assert it_n == it_e and res_n == res_e
    assert np.array_equal(x_n, x_e)
    assert rep_n["true_relres"] == rep_e["true_relres"]
    assert np.array_equal(dv.download(), s.val) and np.array_equal(db.download(), s.b)
    return nat, it_n, x_n


def test_noalloc_matches_existing_solve_two_calls(cases, gpu):
    # larger system first, symmetric engine (its diagonals in the scratch)
    big = cases.hex_pressure(40, 36, 32, rho_out=1000.0)
    _pair(gpu, big, "cg", "jacobi", 3000, 1e-8, criterion=1, kind=8)
    # then a smaller one on the same engine: stale diagonals from the first
    # call sit where this one's gaps must read zero
    small = cases.hex_pressure(18, 14, 11, rho_out=50.0)
    nat, it, x = _pair(gpu, small, "cg", "jacobi", 2000, 1e-8, kind=8)
    assert nat.info()["sym_state"] == 1
    # and the default engine, NRMS with norm, no preconditioner, reference bicg
    smaller = cases.hex_pressure(12, 10, 9)
    _pair(gpu, smaller, "bicg", "none", 1000, 1e-8, norm=1.0)


def test_noalloc_other_engines_and_caps(cases, gpu):
    s = cases.knn(10, 9, 8)
    for kind in (5, 0):
        _pair(gpu, s, "cg", "jacobi", 1000, 1e-8, kind=kind)
    # the iteration cap inside a chunk, and miter = 0
    p = cases.hex_pressure(16)
    _pair(gpu, p, "cg", "jacobi", 7, 1e-12)
    _pair(gpu, p, "cg", "jacobi", 0, 1e-8)


def test_noalloc_refusals(cases, gpu):
    s = cases.hex_pressure(8)
    nat = gpu.Native(s.row, s.col, s.dia)
    dv, db, dx = gpu.DeviceArray(s.val), gpu.DeviceArray(s.b), gpu.DeviceArray(np.zeros(s.n))
    for solver, prec, prec_opt in (("cg", "icc", None), ("cgs", "jacobi", None),
                                   ("cg", "jacobi", "fp32")):
        opt = gpu.Options()
        if prec_opt:
            opt.precision = gpu.FP32_IR
        with pytest.raises(gpu.TFlowsGpuError) as e:
            nat.solve_native_noalloc(solver, prec, dv, dx, db, 100, 1e-8, options=opt)
        assert e.value.code == gpu.TFG_ENOTSUP
    with pytest.raises(gpu.TFlowsGpuError) as e:
        nat.solve_native_noalloc("gmres", "jacobi", dv, dx, db, 100, 1e-8)
    assert e.value.code == gpu.TFG_EINVAL
    # a system beyond the scratch: refused with the bytes it needs, before
    # any kernel runs
    big = cases.hex_pressure(100)
    nb = gpu.Native(big.row, big.col, big.dia)
    bv, bb, bx = gpu.DeviceArray(big.val), gpu.DeviceArray(big.b), gpu.DeviceArray(np.zeros(big.n))
    with pytest.raises(gpu.TFlowsGpuError) as e:
        nb.solve_native_noalloc("cg", "jacobi", bv, bx, bb, 100, 1e-8)
    assert e.value.code == gpu.TFG_EINVAL and "bytes of scratch" in str(e.value)
    assert np.all(bx.download() == 0.0)
