"""GPU SpMV / residual kernels against the oracle's product loops.

The product loop of Cg.f90:177-185 and Residual_Vector.f90:34-42 sum each row
in CSR order; every GPU row engine keeps that order and rounds every product
and sum separately, so all five engines must agree BIT FOR BIT with the
oracle.  Rows longer than the LDS stage (none in T-Flows grids) are summed
by a workgroup tree in the stage engines and are checked to 1e-14 relative
there instead.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _View:
    """A device pointer offset into a DeviceArray (to test misaligned val)."""

    def __init__(self, base, offset_bytes):
        self.base = base
        self.ptr = ctypes.c_void_p(base.ptr.value + offset_bytes)


def _systems(cases):
    return [cases.hex_pressure(40), cases.hex_pressure(24, 20, 16, rho_out=1000.0),
            cases.upwind(20), cases.knn(14, 12, 10), cases.hex_pressure(1, 1, 7)]


# stage2048, stage4096, wave4, wave8, wave16, pipe2048, pipe2048 with offset codes


@pytest.mark.parametrize("idx", range(5))
def test_spmv_bitexact_every_engine(cases, oracle, gpu, idx):
    s = _systems(cases)[idx]
    print(f"spmv case {s.name} n={s.n} nnz={s.nnz}", flush=True)
    rng = np.random.default_rng(idx)
    p = rng.standard_normal(s.n)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    q_ref = oracle.spmv(A, p)
    r_ref = oracle.residual(A, s.b, p)
    nat = gpu.Native(s.row, s.col, s.dia)
    dv, dp, dq = gpu.DeviceArray(s.val), gpu.DeviceArray(p), gpu.DeviceArray(np.zeros(s.n))
    db, dr = gpu.DeviceArray(s.b), gpu.DeviceArray(np.zeros(s.n))
    print(f"  engines {nat.kinds()} (offsets {nat.info()['offsets']})", flush=True)
    for kind in nat.kinds():
        nat.spmv_select(kind)
        nat.spmv(dv, dp, dq)
        q = dq.download()
        assert np.array_equal(q, q_ref), (kind, np.abs(q - q_ref).max())
        nat.residual(dv, db, dp, dr)
        assert np.array_equal(dr.download(), r_ref), kind
    kind, us = nat.spmv_tune(dv)
    print(f"  autotuner picked {nat.SPMV_KINDS[kind]}: " +
          ", ".join(f"{k} {v:.1f} us" for k, v in us.items()), flush=True)
    assert kind in nat.kinds() and us[nat.SPMV_KINDS[kind]] == min(us.values())


def test_symmetric_engine(cases, oracle, gpu):
    """Engine 8 stores the diagonal and upper diagonals once and reads each
    lower entry from its mirror: bit-identical products on a symmetric val.
    A val that is not symmetric bit for bit (a lower entry one ulp off, or
    -0.0 mirroring +0.0) makes that call use the row-pattern engine, with
    the same exactness; the next symmetric val uses engine 8 again."""
    s = cases.hex_pressure(24, 20, 16, rho_out=1000.0)
    nat = gpu.Native(s.row, s.col, s.dia)
    assert nat.info()["sym_diagonals"] == 3 and 8 in nat.kinds()
    nat.spmv_select(8)
    p = np.random.default_rng(5).standard_normal(s.n)
    dp, dq = gpu.DeviceArray(p), gpu.DeviceArray(np.zeros(s.n))
    row0 = s.row - 1
    i = s.n // 3
    lower = [t for t in range(row0[i], row0[i + 1]) if s.col[t] - 1 < i]
    upper = [t for t in range(row0[i], row0[i + 1]) if s.col[t] - 1 > i]
    ulp = s.val.copy()
    ulp[lower[0]] = np.nextafter(ulp[lower[0]], 0.0)
    # +0.0 above the diagonal, -0.0 at its mirror
    zeros = s.val.copy()
    jj = s.col[upper[0]] - 1
    mirror = next(t for t in range(row0[jj], row0[jj + 1]) if s.col[t] - 1 == i)
    zeros[upper[0]], zeros[mirror] = 0.0, -0.0
    for val, state in ((s.val, 1), (ulp, 0), (zeros, 0), (s.val, 1)):
        A = oracle.Matrix(s.row, s.col, s.dia, val)
        dv = gpu.DeviceArray(val)
        nat.spmv(dv, dp, dq)
        assert np.array_equal(dq.download(), oracle.spmv(A, p)), state
        assert nat.info()["sym_state"] == state
    up = cases.upwind(12)   # symmetric structure, non-symmetric values
    nat_u = gpu.Native(up.row, up.col, up.dia)
    nat_u.spmv_select(8)
    pu = np.random.default_rng(6).standard_normal(up.n)
    dq_u = gpu.DeviceArray(np.zeros(up.n))
    nat_u.spmv(gpu.DeviceArray(up.val), gpu.DeviceArray(pu), dq_u)
    assert nat_u.info()["sym_state"] == 0
    assert np.array_equal(dq_u.download(), oracle.spmv(oracle.Matrix(up.row, up.col, up.dia, up.val), pu))


# grid, whether two rows per thread apply (offsets 1 and even ones, n even)
SYM_GRIDS = [((24, 20, 16), True), ((1, 30, 20), True), ((10, 8, 5), False), ((9, 7, 5), False)]


@pytest.mark.parametrize("dims,pair", SYM_GRIDS, ids=[f"{d[0]}x{d[1]}x{d[2]}" for d, _ in SYM_GRIDS])
def test_symmetric_engine_modes(cases, oracle, gpu, dims, pair):
    """Engine 8 takes two rows per thread when the column offsets are 1 and
    even ones and n is even (24x20x16: 1, 16, 320; 1x30x20: 1, 20), and one
    row per thread otherwise (10x8x5: offset 5; 9x7x5: n odd); with x or y
    not 16-byte aligned a pair matrix runs one row per thread over the same
    blocks.  Every mode gives the CSR product bit for bit, and a CG solve
    on it meets the parity bar against the oracle."""
    s = cases.hex_pressure(*dims, rho_out=100.0)
    offs = sorted({int(c) - 1 - i for i in range(s.n)
                   for c in s.col[s.row[i] - 1:s.row[i + 1] - 1] if int(c) - 1 > i})
    assert pair == (offs[0] == 1 and all(o % 2 == 0 for o in offs[1:]) and s.n % 2 == 0)
    nat = gpu.Native(s.row, s.col, s.dia)
    assert nat.info()["sym_diagonals"] == len(offs) and 8 in nat.kinds()
    nat.spmv_select(8)
    p = np.random.default_rng(11).standard_normal(s.n)
    ref = oracle.spmv(oracle.Matrix(s.row, s.col, s.dia, s.val), p)
    dv = gpu.DeviceArray(s.val)
    dq = gpu.DeviceArray(np.zeros(s.n))
    nat.spmv(dv, gpu.DeviceArray(p), dq)
    assert nat.info()["sym_state"] == 1
    assert np.array_equal(dq.download(), ref)
    bp = gpu.DeviceArray(np.concatenate([[0.0], p]))
    bq = gpu.DeviceArray(np.zeros(s.n + 1))
    nat.spmv(dv, _View(bp, 8), _View(bq, 8))
    assert np.array_equal(bq.download()[1:], ref)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    x_o = np.zeros(s.n)
    it_o, _ = oracle.solve("cg", "jacobi", A, x_o, s.b.copy(), 2000, 1e-9)
    x = np.zeros(s.n)
    it, _ = nat.solve_native("cg", "jacobi", s.val.copy(), x, s.b.copy(), 2000, 1e-9)
    assert nat.info()["sym_state"] == 1
    assert abs(it - it_o) <= 2
    assert np.linalg.norm(x - x_o) <= 1e-10 * np.linalg.norm(x_o)


def test_spmv_misaligned_values(cases, oracle, gpu):
    """val not 16-byte aligned: the scalar staging path must give the same bits."""
    s = cases.hex_pressure(30, 20, 10)
    p = np.random.default_rng(7).standard_normal(s.n)
    A = oracle.Matrix(s.row, s.col, s.dia, s.val)
    nat = gpu.Native(s.row, s.col, s.dia)
    buf = gpu.DeviceArray(np.concatenate([[0.0], s.val]))
    dv = _View(buf, 8)
    dp, dq = gpu.DeviceArray(p), gpu.DeviceArray(np.zeros(s.n))
    for kind in nat.kinds():
        nat.spmv_select(kind)
        nat.spmv(dv, dp, dq)
        assert np.array_equal(dq.download(), oracle.spmv(A, p)), kind


def test_spmv_long_rows(oracle, gpu):
    """Rows longer than the LDS stage (arrow matrix: dense first row/column)."""
    n = 6000
    rows, cols, vals = [], [], []
    rng = np.random.default_rng(3)
    for i in range(n):
        c = [i] + ([j for j in range(1, n)] if i == 0 else ([0] + ([i - 1] if i > 1 else []) +
                                                            ([i + 1] if i + 1 < n else [])))
        c = [c[0]] + sorted(set(c[1:]) - {i})
        rows.append(len(c))
        cols.extend(ci + 1 for ci in c)
        vals.extend([4.0 + n] + list(rng.uniform(-1, 0, len(c) - 1)))
    row = np.concatenate([[1], 1 + np.cumsum(rows)]).astype(np.int32)
    col = np.array(cols, dtype=np.int32)
    val = np.array(vals)
    dia = row[:-1].copy()
    p = rng.standard_normal(n)
    A = oracle.Matrix(row, col, dia, val)
    q_ref = oracle.spmv(A, p)
    nat = gpu.Native(row, col, dia)
    dv, dp, dq = gpu.DeviceArray(val), gpu.DeviceArray(p), gpu.DeviceArray(np.zeros(n))
    for kind in nat.kinds():
        nat.spmv_select(kind)
        nat.spmv(dv, dp, dq)
        q = dq.download()
        assert np.array_equal(q[1:], q_ref[1:]), kind
        if kind in (2, 3, 4):   # the wave engines keep CSR order on long rows too
            assert q[0] == q_ref[0]
        else:           # the stage engines sum a row longer than the stage as a tree
            assert abs(q[0] - q_ref[0]) <= 1e-14 * np.abs(val[:n] * p[col[:n] - 1]).sum()


def _nine_point(nx, ny, seed=3):
    """Symmetric 2-D 9-point system (1-based CSR, diagonal first, then
    ascending columns): upper offsets 1, ny-1, ny, ny+1."""
    rng = np.random.default_rng(seed)
    n = nx * ny
    w = {}
    for x in range(nx):
        for y in range(ny):
            i = x * ny + y
            for dx, dy in ((0, 1), (1, -1), (1, 0), (1, 1)):
                if 0 <= x + dx < nx and 0 <= y + dy < ny:
                    j = (x + dx) * ny + y + dy
                    w[(i, j)] = w[(j, i)] = -rng.uniform(0.5, 1.5)
    rows = [[] for _ in range(n)]
    for (i, j), v in w.items():
        rows[i].append((j, v))
    row, col, val = [1], [], []
    for i in range(n):
        ent = sorted(rows[i])
        col += [i + 1] + [j + 1 for j, _ in ent]
        val += [-sum(v for _, v in ent) + 0.1] + [v for _, v in ent]
        row.append(len(col) + 1)
    row = np.array(row, np.int32)
    return row, np.array(col, np.int32), row[:-1].copy(), np.array(val)


def test_symmetric_engine_nine_point(oracle, gpu):
    """A 2-D 9-point stencil: four upper diagonals (1, ny-1, ny, ny+1; odd
    offsets, so one row per thread) in engine 8, bit-exact against the
    oracle's CSR product; a CG solve on it meets the parity bar; and its ICC
    (not a box grid: diagonal neighbours) runs the row-level sweeps, also
    bit-exact."""
    row, col, dia, val = _nine_point(36, 30)
    n = len(dia)
    nat = gpu.Native(row, col, dia)
    assert nat.info()["sym_diagonals"] == 4 and nat.info()["icc_tiles"] == 0
    nat.spmv_select(8)
    A = oracle.Matrix(row, col, dia, val)
    p = np.random.default_rng(12).standard_normal(n)
    dv, dq = gpu.DeviceArray(val), gpu.DeviceArray(np.zeros(n))
    nat.spmv(dv, gpu.DeviceArray(p), dq)
    assert nat.info()["sym_state"] == 1
    assert np.array_equal(dq.download(), oracle.spmv(A, p))
    b = np.random.default_rng(13).standard_normal(n)
    x_o = np.zeros(n)
    # the oracle's solve scales its matrix in place and back (as Cg.f90
    # does), which can move a value by an ulp: a copy of its own
    it_o, _ = oracle.solve("cg", "jacobi", oracle.Matrix(row, col, dia, val.copy()), x_o,
                           b.copy(), 2000, 1e-10)
    x = np.zeros(n)
    it, _ = nat.solve_native("cg", "jacobi", val.copy(), x, b.copy(), 2000, 1e-10)
    assert abs(it - it_o) <= 2
    assert np.linalg.norm(x - x_o) <= 1e-10 * np.linalg.norm(x_o)
    d, dinv = oracle.prec_form(A, "icc")
    z_ref = oracle.prec_solve(A, d, dinv, b, "icc")
    dz = gpu.DeviceArray(np.zeros(n))
    nat.prec_apply("icc", dv, gpu.DeviceArray(b), dz)
    assert np.array_equal(dz.download(), z_ref)
