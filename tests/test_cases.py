"""Properties of the synthetic systems (layout of Create_Matrix.f90:87-119,
physics of each configuration, reproducibility)."""
import numpy as np
import pytest


def _layout_ok(s):
    n = s.n
    assert s.row[0] == 1 and s.row[-1] == s.nnz + 1
    assert np.all(np.diff(s.row) >= 1)
    assert np.array_equal(s.dia, s.row[:-1])                      # diagonal first
    assert np.array_equal(s.col[s.dia - 1], np.arange(1, n + 1))
    for i in range(0, n, max(1, n // 97)):                         # rest ascending
        c = s.col[s.row[i]:s.row[i + 1] - 1]
        assert np.all(np.diff(c) > 0) and i + 1 not in c
    assert s.col.min() >= 1 and s.col.max() <= n


@pytest.mark.parametrize("g", [(5, 4, 3), (1, 1, 7), (12, 12, 12), (2, 9, 4)])
def test_hex_layout_and_laplacian(cases, g):
    s = cases.hex_pressure(*g)
    nx, ny, nz = g
    assert s.n == nx * ny * nz
    assert s.nnz == 7 * s.n - 2 * (ny * nz + nx * nz + nx * ny) == cases.hex_nnz(*g)
    _layout_ok(s)
    A = s.scipy()
    assert abs(A - A.T).max() == 0.0
    rs = np.asarray(A.sum(axis=1)).ravel()
    pin = 2.0 * (1.0 / ny) * (1.0 / nz) / (1.0 / nx)   # A_f / (h_x/2) of cell 1
    assert rs[0] == pytest.approx(pin, rel=1e-12)
    assert np.allclose(rs[1:], 0.0, atol=1e-13)                    # Neumann walls
    assert np.all(A.diagonal() > 0) and np.all(A.data[A.data != A.diagonal().max()] <= A.diagonal().max())
    assert abs(s.b.sum()) < 1e-9 * s.n and np.all(np.abs(s.b) <= 2.0)


def test_two_phase_harmonic_faces(cases):
    N = 16
    s = cases.hex_pressure(N, rho_out=1000.0)
    A = s.scipy().tocoo()
    off = -A.data[A.row != A.col]
    h = 1.0 / N
    # a_f = h * (1/rho)_f with (1/rho)_f in {1, (1+1e-3)/2, 1e-3}
    vals = np.unique(np.round(off / h, 12))
    assert set(vals) <= {1.0, round(0.5 * (1 + 1e-3), 12), 1e-3}
    assert len(vals) == 3


def test_upwind_properties(cases):
    s = cases.upwind(12)
    _layout_ok(s)
    A = s.scipy()
    assert abs(A - A.T).max() > 1e-4                                # non-symmetric
    offd = A - __import__("scipy.sparse", fromlist=["diags"]).diags(A.diagonal())
    assert offd.max() <= 0.0                                        # M-matrix signs
    dom = A.diagonal() - np.asarray(abs(offd).sum(axis=1)).ravel()
    assert dom.min() > 0.0                                          # time term
    # time term vol/dt with dt |u|max / h = 0.5 and |u| <= 1: vol/dt >= 2 h^2
    h = 1.0 / 12
    assert dom.mean() > 1.0 * h * h


def test_knn_properties(cases):
    s = cases.knn(10, 9, 8)
    _layout_ok(s)
    A = s.scipy()
    assert abs(A - A.T).max() == 0.0
    deg = np.diff(s.row) - 1
    assert deg.min() >= 10 and deg.max() <= 64
    assert 11 <= np.mean(deg + 1) <= 22
    # random permutation: bandwidth of the order of n, not of a lattice plane
    coo = A.tocoo()
    assert np.median(np.abs(coo.row - coo.col)) > s.n / 10
    assert np.all(A.diagonal() > 0)


def test_reproducible_and_seeded(cases):
    a, b = cases.hex_pressure(6, rho_out=100.0), cases.hex_pressure(6, rho_out=100.0)
    assert np.array_equal(a.val, b.val) and np.array_equal(a.b, b.b)
    c = cases.hex_pressure(6, rho_out=100.0, seed=99)
    assert not np.array_equal(a.b, c.b)
    k1, k2 = cases.knn(6, 6, 6), cases.knn(6, 6, 6)
    assert np.array_equal(k1.col, k2.col) and np.array_equal(k1.val, k2.val)


def test_batch_params(cases):
    sides, ratios = cases.batch_params(256)
    assert sides.min() >= 20 and sides.max() <= 80
    assert ratios.min() >= 1.0 and ratios.max() <= 1000.0
    assert len(set(sides.tolist())) > 30
