"""The Rhie-Chow flux oracle (Rhie_And_Chow.f90:86-259) against what the
interpolation is built to do, on the uniform hex grid's own geometry: a
linear pressure field with its exact gradient leaves no smoothing term; a
checkerboard pressure gets the full face coefficient times the pressure
jump; Choi's correction leaves a steady uniform flow unchanged; Gu's
correction gives no flux in hydrostatic balance; a balanced surface-tension
pressure gives none either.  Each fails for a flipped sign of the term it
tests."""
import dataclasses

import numpy as np
import pytest

DIMS = (6, 5, 4)


def _setup(cases, **kw):
    f = cases.hex_faces(*DIMS, shuffle=True)
    rf = cases.rhie_chow_fields(f, *DIMS)
    z = np.zeros(f.n)
    base = dict(u=z, v=z, w=z, v_m=np.full(f.n, 2e-2))
    base.update(kw)
    return f, dataclasses.replace(rf, **base)


def _linear(cases, f, g):
    xc = cases.cell_centres(*DIMS)
    return xc @ np.asarray(g), [np.full(f.n, gi) for gi in g]


def _scale(f, rf):
    return np.abs(rf.fc).max() * rf.v_m.max()


def test_linear_pressure_leaves_no_smoothing(cases, oracle):
    f, rf = _setup(cases)
    p, (gx, gy, gz) = _linear(cases, f, (0.7, -1.3, 2.1))
    rf = dataclasses.replace(rf, p=p, px=gx, py=gy, pz=gz)
    out = oracle.rhie_chow_fluxes(f, rf, np.zeros(f.n_faces))
    assert np.abs(out[f.c2 > 0]).max() <= 1e-14 * _scale(f, rf)


def test_checkerboard_pressure_gets_the_full_jump(cases, oracle):
    f, rf = _setup(cases)
    xc = cases.cell_centres(*DIMS)
    ijk = np.floor(xc * np.array(DIMS)).astype(int)
    p = np.where(ijk.sum(1) % 2 == 0, 1.0, -1.0)
    z = np.zeros(f.n)
    rf = dataclasses.replace(rf, p=p, px=z, py=z, pz=z)
    out = oracle.rhie_chow_fluxes(f, rf, np.zeros(f.n_faces))
    i = f.c2 > 0
    jump = p[f.c1[i] - 1] - p[f.c2[i] - 1]
    assert np.allclose(out[i], rf.fc[i] * 2e-2 * jump, rtol=1e-15, atol=0.0)
    assert np.all(np.abs(jump) == 2.0)


def test_choi_leaves_a_steady_uniform_flow(cases, oracle):
    U = np.array([0.4, -0.2, 0.9])
    f, rf = _setup(cases)
    n = f.n
    ones = np.ones(n)
    s = np.stack([rf.sx, rf.sy, rf.sz], 1)
    z = np.zeros(n)
    rf = dataclasses.replace(rf, u=U[0] * ones, v=U[1] * ones, w=U[2] * ones,
                             u_o=U[0] * ones, v_o=U[1] * ones, w_o_=U[2] * ones,
                             u_oo=U[0] * ones, v_oo=U[1] * ones, w_oo_=U[2] * ones,
                             density=np.full(n, 3.0), p=z, px=z, py=z, pz=z,
                             flux_o=s @ U, flux_oo=s @ U)
    plain = oracle.rhie_chow_fluxes(f, rf, np.zeros(f.n_faces))
    for w_o, w_oo in ((1.0, 0.0), (2.0, -0.5)):
        choi = oracle.rhie_chow_fluxes(f, rf, np.zeros(f.n_faces), choi=True, w_o=w_o, w_oo=w_oo,
                                       dt=1e-3)
        i = f.c2 > 0
        assert np.allclose(choi[i], plain[i], rtol=1e-12, atol=1e-14 * np.abs(plain).max())
        assert np.allclose(plain[i], (s @ U)[i], rtol=1e-15, atol=1e-17)


def _face_gradient(f, rf, phi):
    """Face vectors g with g . d = phi(c2) - phi(c1) on the domain faces
    (along the face's axis; the other components random)."""
    rng = np.random.default_rng(11)
    g = rng.uniform(-1.0, 1.0, (f.n_faces, 3))
    d = np.stack([rf.dx, rf.dy, rf.dz], 1)
    i = np.flatnonzero(f.c2 > 0)
    axis = np.argmax(np.abs(d[i]), 1)
    g[i, axis] = (phi[f.c2[i] - 1] - phi[f.c1[i] - 1]) / d[i, axis]
    return g


def test_gu_gives_no_flux_in_hydrostatic_balance(cases, oracle):
    """Discrete hydrostatic balance with forces that vary from cell to cell:
    the face forces are the pressure's face differences and the cell
    gradients equal the cell forces.  Gu's correction removes the spurious
    flux that interpolating the cell forces leaves."""
    f, rf = _setup(cases)
    rng = np.random.default_rng(12)
    p = rng.uniform(-1.0, 1.0, f.n)
    g = _face_gradient(f, rf, p)
    cf = rng.uniform(-1.0, 1.0, (f.n, 3))
    rf = dataclasses.replace(rf, p=p, px=cf[:, 0].copy(), py=cf[:, 1].copy(),
                             pz=cf[:, 2].copy(), cell_fx=cf[:, 0].copy(),
                             cell_fy=cf[:, 1].copy(), cell_fz=cf[:, 2].copy(),
                             face_fx=g[:, 0].copy(), face_fy=g[:, 1].copy(),
                             face_fz=g[:, 2].copy())
    out = oracle.rhie_chow_fluxes(f, rf, np.zeros(f.n_faces), gu=True)
    assert np.abs(out[f.c2 > 0]).max() <= 1e-14 * _scale(f, rf)
    bad = oracle.rhie_chow_fluxes(f, rf, np.zeros(f.n_faces), gu=False)
    assert np.abs(bad[f.c2 > 0]).max() > 1e3 * 1e-14 * _scale(f, rf)


def test_balanced_surface_tension_gives_no_flux(cases, oracle):
    """p = sigma kappa col with the cell gradients in balance: the surface
    tension term cancels the smoothing of a non-linear colour function."""
    sigma, kappa = 0.07, 25.0
    f, rf = _setup(cases)
    rng = np.random.default_rng(13)
    col = rng.uniform(0.0, 1.0, f.n)
    cg = rng.uniform(-1.0, 1.0, (f.n, 3))
    p = sigma * kappa * col
    rf = dataclasses.replace(rf, col=col, col_x=cg[:, 0].copy(), col_y=cg[:, 1].copy(),
                             col_z=cg[:, 2].copy(), curv=np.full(f.n, kappa), p=p,
                             px=sigma * kappa * cg[:, 0], py=sigma * kappa * cg[:, 1],
                             pz=sigma * kappa * cg[:, 2])
    out = oracle.rhie_chow_fluxes(f, rf, np.zeros(f.n_faces), vof=True, sigma=sigma)
    assert np.abs(out[f.c2 > 0]).max() <= 1e-14 * _scale(f, rf)
    off = oracle.rhie_chow_fluxes(f, rf, np.zeros(f.n_faces), vof=True, sigma=1e-16)  # <= TINY
    assert np.abs(off[f.c2 > 0]).max() > 1e3 * 1e-14 * _scale(f, rf)


def test_boundary_faces_untouched(cases, oracle):
    f, rf = _setup(cases)
    base = np.full(f.n_faces, 123.0)
    out = oracle.rhie_chow_fluxes(f, rf, base, choi=True, gu=True, vof=True)
    assert np.all(out[f.c2 <= 0] == 123.0) and np.all(out[f.c2 > 0] != 123.0)
