"""The Fortran host binding (t-flows_amd/fortran: Tflows_Gpu_Mod, an
iso_c_binding module keeping Solve_Native's argument list) driving the GPU
solver, against the oracle and against the same solve through Python."""
import os
import subprocess

import numpy as np
import pytest

import parity_bar

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "t-flows_amd", "fortran", "solve_driver")


def _write(d, s, x0):
    with open(os.path.join(d, "dims.txt"), "w") as f:
        f.write(f"{s.n} {s.n} {s.nnz}\n")
    for name, a in (("row", s.row), ("col", s.col), ("dia", s.dia)):
        a.astype(np.int32).tofile(os.path.join(d, f"{name}.bin"))
    for name, a in (("val", s.val), ("b", s.b), ("x0", x0)):
        a.astype(np.float64).tofile(os.path.join(d, f"{name}.bin"))


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no amdflang)")
@pytest.mark.parametrize("solver,prec,norm", [("cg", "jacobi", None), ("bicg", "jacobi", 1.0),
                                              ("cg", "none", None)])
def test_fortran_driver(cases, oracle, gpu, tmp_path, solver, prec, norm):
    s = cases.hex_pressure(20, 16, 12, rho_out=10.0)
    x0 = np.zeros(s.n)
    _write(str(tmp_path), s, x0)
    args = [DRIVER, str(tmp_path), solver, prec, "2000", "1e-8"] + ([str(norm)] if norm else [])
    print("running", " ".join(args), flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    niter, fin_res = out.stdout.split()
    niter, fin_res = int(niter), float(fin_res)
    x = np.fromfile(os.path.join(str(tmp_path), "x.bin"), dtype=np.float64)
    # the same call through the Python mirror gives the same bits
    nat = gpu.Native(s.row, s.col, s.dia)
    xp = x0.copy()
    it_p, res_p = nat.solve_native(solver, prec, s.val.copy(), xp, s.b.copy(), 2000, 1e-8, norm=norm)
    assert (niter, fin_res) == (it_p, res_p) and np.array_equal(x, xp)
    # and agrees with the oracle: +-2 iterations, x within 1e-10 relative or
    # the spread that reordering the oracle's sums causes by itself

    def solve(so):
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        xo = x0.copy()
        it, _ = oracle.solve(solver, prec, A, xo, s.b.copy(), 2000, 1e-8, norm=norm, sum_order=so)
        return it, xo

    it_o, xo = solve(0)
    env, spread = parity_bar.envelope(solve, it_o, xo)
    parity_bar.check(f"fortran {solver} {prec} norm={norm}", niter, x, it_o, xo, env, spread)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no amdflang)")
def test_fortran_unknown_solver_stops(cases, tmp_path):
    """An unknown solver name is fatal, as Message % Error is in the reference."""
    s = cases.hex_pressure(4)
    _write(str(tmp_path), s, np.zeros(s.n))
    out = subprocess.run([DRIVER, str(tmp_path), "gmres", "jacobi", "10", "1e-8"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode != 0 and "Unknown native solver" in out.stdout


def _run_driver(d, solver, prec, norm, mode, miter=2000, tol=1e-8):
    args = [DRIVER, d, solver, prec, str(miter), str(tol), str(norm) if norm else "-", mode]
    print("running", " ".join(args), flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    x = np.fromfile(os.path.join(d, "x.bin"), dtype=np.float64)
    return int(f[0]), float(f[1]), (float(f[2]) if len(f) > 2 else None), x


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no amdflang)")
@pytest.mark.parametrize("solver,prec,norm", [("cg", "jacobi", None), ("bicg", "icc", 1.0)])
def test_fortran_device_resident(cases, oracle, tmp_path, solver, prec, norm):
    """Tflows_Gpu_Solve_Device (tfg_solve_native_ex on device addresses from
    tfg_device_alloc): the same bits as the host-array call; and the FP32
    sweeps inside FP64 refinement through the same entry point, converged,
    with an error against the oracle's tightly converged solution no worse
    than ten times the FP64 run's."""
    s = cases.hex_pressure(20, 16, 12, rho_out=10.0)
    _write(str(tmp_path), s, np.zeros(s.n))
    d = str(tmp_path)
    it_h, res_h, _, x_h = _run_driver(d, solver, prec, norm, "host")
    it_d, res_d, true_d, x_d = _run_driver(d, solver, prec, norm, "device")
    assert (it_d, res_d) == (it_h, res_h) and np.array_equal(x_d, x_h)
    assert true_d < 1e-6
    A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
    xo = np.zeros(s.n)
    oracle.solve(solver, prec, A, xo, s.b.copy(), 5000, 1e-14, criterion="relres")
    err64 = parity_bar.rel(x_h, xo)
    it_f, res_f, true_f, x_f = _run_driver(d, solver, prec, norm, "fp32")
    err32 = parity_bar.rel(x_f, xo)
    print(f"  fp32 form: {it_f} FP32 iterations, fin_res {res_f:.3e}, true relres {true_f:.3e}, "
          f"x err {err32:.2e} (FP64 {err64:.2e})", flush=True)
    assert it_f < 2000 and res_f < 1e-8
    assert err32 <= 10 * max(err64, 1e-12)
