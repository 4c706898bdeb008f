"""The reference's default preconditioner, incomplete Cholesky (DIC), on the
GPU: z = M^-1 r bit-identical to the oracle's serial sweeps, and ICC-CG
solves at the FP64 parity bar."""
import numpy as np
import pytest

import parity_bar

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# box grids run the tiled sweeps (tiles of up to 8x8x16: 40x33x18 has 5x5x2
# tiles, partial along y and z; 17x19x21 has 3x3x2, partial along every
# axis; 1x50x37 is a 2-D grid of 7x3 tiles);
# kNN runs the row-level sweeps
@pytest.mark.parametrize("which", ["poisson12", "twophase10x8x6", "knn8", "upwind9", "line7",
                                   "twophase40x33x18", "twophase17x19x21", "plane1x50x37"])
def test_prec_apply_bitexact(cases, oracle, gpu, which):
    s = {"poisson12": lambda: cases.hex_pressure(12),
         "twophase10x8x6": lambda: cases.hex_pressure(10, 8, 6, rho_out=1000.0),
         "knn8": lambda: cases.knn(8, 8, 8),
         "upwind9": lambda: cases.upwind(9),
         "line7": lambda: cases.hex_pressure(1, 1, 7),
         "twophase40x33x18": lambda: cases.hex_pressure(40, 33, 18, rho_out=1000.0),
         "twophase17x19x21": lambda: cases.hex_pressure(17, 19, 21, rho_out=1000.0),
         "plane1x50x37": lambda: cases.hex_pressure(1, 50, 37, rho_out=100.0)}[which]()
    print(f"icc apply case {s.name} n={s.n} nnz={s.nnz}", flush=True)
    r = np.random.default_rng(5).standard_normal(s.n)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    nat = gpu.Native(s.row, s.col, s.dia)
    tiles = {"poisson12": 4, "twophase10x8x6": 2, "knn8": 0, "upwind9": 4, "line7": 1,
             "twophase40x33x18": 50, "twophase17x19x21": 18,
             "plane1x50x37": 21}[which]
    assert nat.info()["icc_tiles"] == tiles
    dv, dr, dz = gpu.DeviceArray(s.val), gpu.DeviceArray(r), gpu.DeviceArray(np.zeros(s.n))
    for prec in ("icc", "jacobi", "none"):
        d, dinv = oracle.prec_form(A, prec)
        z_ref = oracle.prec_solve(A, d, dinv, r, prec)
        nat.prec_apply(prec, dv, dr, dz)
        z = dz.download()
        assert np.array_equal(z, z_ref), (prec, np.abs(z - z_ref).max())


CASES = [
    ("poisson24", lambda c: c.hex_pressure(24), None, 1000, 1e-8),
    ("twophase20_norm1", lambda c: c.hex_pressure(20, rho_out=1000.0), 1.0, 2000, 1e-8),
    ("knn10", lambda c: c.knn(10, 10, 10), None, 1000, 1e-8),
]


@pytest.mark.parametrize("name,make,norm,miter,tol", CASES, ids=[c[0] for c in CASES])
def test_icc_cg_parity(cases, oracle, gpu, name, make, norm, miter, tol):
    s = make(cases)
    print(f"icc-cg case {name} n={s.n} nnz={s.nnz}", flush=True)

    def solve(so):
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        x = np.zeros(s.n)
        it, fr = oracle.solve("cg", "icc", A, x, s.b.copy(), miter, tol, norm=norm, sum_order=so)
        return it, fr, x

    it_o, res_o, x_o = solve(0)
    env, spread = parity_bar.envelope(lambda so: solve(so)[::2], it_o, x_o)
    nat = gpu.Native(s.row, s.col, s.dia)
    x = np.zeros(s.n)
    it_g, res_g = nat.solve_native("cg", "icc", s.val.copy(), x, s.b.copy(), miter, tol, norm=norm)
    print(f"  oracle niter={it_o} res={res_o:.3e}; gpu niter={it_g} res={res_g:.3e}", flush=True)
    assert it_o < miter
    parity_bar.check(f"icc-cg {name}", it_g, x, it_o, x_o, env, spread)
    assert res_g < tol


def test_icc_bicg_equals_cg_and_caps(cases, oracle, gpu):
    s = cases.hex_pressure(14, rho_out=50.0)
    nat = gpu.Native(s.row, s.col, s.dia)
    xa, xb = np.zeros(s.n), np.zeros(s.n)
    ra = nat.solve_native("cg", "icc", s.val.copy(), xa, s.b.copy(), 500, 1e-9)
    rb = nat.solve_native("bicg", "icc", s.val.copy(), xb, s.b.copy(), 500, 1e-9)
    assert ra == rb and np.array_equal(xa, xb)
    for miter in (0, 1, 3):
        A = oracle.Matrix(s.row, s.col, s.dia, s.val.copy())
        xo = np.zeros(s.n)
        it_o, res_o = oracle.solve("cg", "icc", A, xo, s.b.copy(), miter, 1e-30)
        x = np.zeros(s.n)
        it_g, res_g = nat.solve_native("cg", "icc", s.val.copy(), x, s.b.copy(), miter, 1e-30)
        assert it_g == it_o == miter and abs(res_g - res_o) <= 1e-12 * res_o
        assert _rel(x, xo) < 1e-13 if miter else np.array_equal(x, xo)


def test_icc_fewer_iterations_than_jacobi(cases, gpu):
    s = cases.hex_pressure(32, rho_out=1000.0)
    nat = gpu.Native(s.row, s.col, s.dia)
    o = gpu.Options()
    o.criterion = gpu.CRIT_RELRES
    out = {}
    for prec in ("jacobi", "icc"):
        x = np.zeros(s.n)
        it, res, rep = nat.solve_native("cg", prec, s.val.copy(), x, s.b.copy(), 5000, 1e-8,
                                        options=o, report=True)
        out[prec] = (it, rep["solve_ms"])
        print(f"  {prec}: {it} iterations, {rep['solve_ms']:.1f} ms", flush=True)
    assert out["icc"][0] < out["jacobi"][0]


@pytest.mark.skip(reason="the single-launch box sweeps (a persistent grid taking tiles from a "
                         "queue, waiting on tile flags) hung the first prec_apply on the GPU twice, "
                         "the second time after the flag write-back fix; no further cause was "
                         "found, so the code path was taken out and the sweeps run one launch "
                         "per tile level")
@pytest.mark.parametrize("which", ["twophase40x33x18", "twophase17x19x21", "plane1x50x37",
                                   "twophase64"])
def test_single_launch_sweeps_equal_level_launches(cases, gpu, monkeypatch, which):
    """The box sweeps in one launch (a persistent grid taking tiles from a
    queue in level order, each waiting on its predecessors' flags) give the
    bits of one launch per tile level, for z = M^-1 r and for an ICC-CG
    solve; the 64^3 grid has 8x8x4 tiles, more than the grid holds at once."""
    s = {"twophase40x33x18": lambda: cases.hex_pressure(40, 33, 18, rho_out=1000.0),
         "twophase17x19x21": lambda: cases.hex_pressure(17, 19, 21, rho_out=1000.0),
         "plane1x50x37": lambda: cases.hex_pressure(1, 50, 37, rho_out=100.0),
         "twophase64": lambda: cases.hex_pressure(64, rho_out=1000.0)}[which]()
    r = np.random.default_rng(9).standard_normal(s.n)
    out = {}
    for mode in ("levels", "single"):
        monkeypatch.setenv("TFG_ICC_SWEEP", mode)
        nat = gpu.Native(s.row, s.col, s.dia)
        dv, dr, dz = gpu.DeviceArray(s.val), gpu.DeviceArray(r), gpu.DeviceArray(np.zeros(s.n))
        zs = []
        for _ in range(3):  # repeated sweeps: new epochs over the same flags
            nat.prec_apply("icc", dv, dr, dz)
            zs.append(dz.download())
        assert all(np.array_equal(z, zs[0]) for z in zs)
        x = np.zeros(s.n)
        it, res = nat.solve_native("cg", "icc", s.val.copy(), x, s.b.copy(), 2000, 1e-9)
        print(f"ICC {which} {mode}: {it} iterations, res {res:.3e}", flush=True)
        out[mode] = (zs[0], it, res, x)
        nat.close()
    a, b = out["levels"], out["single"]
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1] and a[2] == b[2] and np.array_equal(a[3], b[3])
