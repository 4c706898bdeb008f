"""Pins of the oracle's pressure-system assembly (Rhie_And_Chow.f90:184-205,
Compute_Pressure.f90:121-178, Set_Singular.f90:28-31) by properties the
face loop fixes, not by restating it:
  * the energy identity x^T A x = sum_internal a12 (x_c1 - x_c2)^2
    + sum_bc fc v_m(c1) x_c1^2 on random x (positions, signs, diagonal
    accumulation and boundary terms of every face);
  * exact symmetry (both off-diagonal slots of a face get the same -a12);
  * the right-hand side: internal fluxes cancel, sum(b) = -sum of the
    boundary fluxes (without pressure regions), b = 0 in pressure cells,
    b = fc v_m pp in ambient cells, the later region winning;
  * vel_max (Compute_Momentum.f90:201-204) against numpy;
  * Set_Singular multiplies exactly the diagonal by (1 + 1e-6);
  * the face order changes the diagonal by rounding only.
"""
import numpy as np
import scipy.sparse as sp


def _matrix(f, val):
    return sp.csr_matrix((val, f.col - 1, f.row - 1), shape=(f.n, f.n))


def _a12(f):
    i = f.c2 > 0
    c1, c2 = f.c1[i] - 1, f.c2[i] - 1
    return i, c1, c2, f.fc[i] * (f.fw[i] * f.vm[c1] + (1.0 - f.fw[i]) * f.vm[c2])


def test_energy_identity_and_symmetry(cases, oracle):
    f = cases.hex_faces(9, 7, 5)
    val, b = oracle.pressure_assemble(f)
    A = _matrix(f, val)
    assert (A != A.T).nnz == 0
    i, c1, c2, a12 = _a12(f)
    bnd = (f.c2 <= 0) & (f.diag_bc == 1)
    cb = f.c1[bnd] - 1
    for seed in range(3):
        x = np.random.default_rng(seed).standard_normal(f.n)
        lhs = x @ (A @ x)
        rhs = np.sum(a12 * (x[c1] - x[c2]) ** 2) + np.sum(f.fc[bnd] * f.vm[cb] * x[cb] ** 2)
        assert abs(lhs - rhs) <= 1e-12 * abs(rhs)
    # an off-diagonal entry holds -a12 of its face
    assert np.array_equal(val[f.pos[0::2][i] - 1], -a12)


def test_rhs(cases, oracle):
    f = cases.hex_faces(8, 6, 5)
    val, b = oracle.pressure_assemble(f)
    zero_cells = np.unique(f.c1[f.b_region == 1] - 1)
    assert np.all(b[zero_cells] == 0.0)
    f.b_region[:] = 0
    val2, b2 = oracle.pressure_assemble(f)
    assert np.array_equal(val, val2)
    flux_bnd = f.vflux[f.c2 <= 0].sum()
    assert abs(b2.sum() + flux_bnd) <= 1e-12 * np.abs(f.vflux).sum()
    # one cell by hand: -flux where it is c1, +flux where it is c2
    c = 17
    ref = -f.vflux[f.c1 - 1 == c].sum() + f.vflux[f.c2 - 1 == c].sum()
    assert abs(b2[c] - ref) <= 1e-14 * np.abs(f.vflux).sum()


def test_ambient_region(cases, oracle):
    """Compute_Pressure.f90:171-189 by hand: an ambient side's cells get
    b = fc v_m(c1) pp(c2) of their face, a pressure side's cells 0, and a
    cell on both sides the value of the later region (the y = 0 side,
    region 3, after the x = 0 side, region 1); the matrix does not change."""
    f = cases.hex_faces(6, 5, 4, ambient=True)
    val, b = oracle.pressure_assemble(f)
    amb = f.b_region == 2
    press = f.b_region == 1
    assert amb.sum() == 6 * 4 and press.sum() == 5 * 4
    nb = f.n_bnd
    expect = {}
    for s in np.flatnonzero(press):
        expect[f.c1[s] - 1] = 0.0
    for s in np.flatnonzero(amb):      # region 3 comes after region 1
        expect[f.c1[s] - 1] = (f.fc[s] * f.vm[f.c1[s] - 1]) * f.pp[f.c2[s] + nb]
    for c, v in expect.items():
        assert b[c] == v, (c, b[c], v)
    corner = [c for c in expect if c < 4]          # x = 0 and y = 0
    assert corner and all(b[c] != 0.0 for c in corner)
    # every other cell keeps the flux balance, and the matrix is the one with
    # the same boundary diagonal terms
    f0 = cases.hex_faces(6, 5, 4, ambient=True)
    f0.b_region[:] = 0
    val0, b0 = oracle.pressure_assemble(f0)
    rest = np.setdiff1d(np.arange(f.n), list(expect))
    assert np.array_equal(b[rest], b0[rest]) and np.array_equal(val, val0)


def test_vel_max(oracle):
    """Compute_Momentum.f90:201-204: the largest |u| over the cells, at least
    MICRO."""
    rng = np.random.default_rng(4)
    u, v, w = (rng.uniform(-1, 1, 1000) for _ in range(3))
    assert oracle.vel_max(u, v, w) == np.max(np.sqrt(u * u + v * v + w * w))
    z = np.zeros(50)
    assert oracle.vel_max(z, z, z) == 1e-6
    assert oracle.vel_max(z + 1e-7, z, z) == 1e-6


def test_singular_and_face_order(cases, oracle):
    f = cases.hex_faces(7, 6, 5)
    val, b = oracle.pressure_assemble(f)
    vs, bs = oracle.pressure_assemble(f, singular=True)
    d = f.dia - 1
    off = np.ones(len(val), bool)
    off[d] = False
    assert np.array_equal(vs[d], val[d] * (1.0 + 1e-6))
    assert np.array_equal(vs[off], val[off]) and np.array_equal(bs, b)
    # the same faces in another order
    import copy
    g = copy.copy(f)
    o = np.lexsort((f.c2, f.c1))
    for name in ("c1", "c2", "diag_bc", "b_region", "region", "fc", "fw", "vflux"):
        setattr(g, name, np.ascontiguousarray(getattr(f, name)[o]))
    g.pos = np.ascontiguousarray(f.pos.reshape(-1, 2)[o].ravel())
    vg, bg = oracle.pressure_assemble(g)
    assert np.array_equal(vg[off], val[off])          # off-diagonals: one face each
    assert np.allclose(vg[d], val[d], rtol=1e-15, atol=0)
    assert np.allclose(bg, b, rtol=0, atol=1e-15)
