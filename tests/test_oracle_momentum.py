"""The momentum-assembly oracle (Compute_Momentum.f90 with Numerics_Mod
Advection_Term, Min_Max, Advection_Scheme, Inertial_Term, Under_Relax)
against properties the discretisation fixes, not against its own formulas:
conservation and row sums of the face loops, the mass-flux antisymmetry of
implicit upwinding, textbook TVD limiter values psi(r), the backward Euler
and BDF2 steady states, invariance of the residual under under-relaxation.
SMART and AVL_SMART are pinned only in their upwind (r <= 0) range."""
import dataclasses

import numpy as np
import pytest


def _off(mf, **kw):
    """Fields with every term switched off except those given."""
    z = dict(vis_eff=np.zeros_like(mf.vis_eff), f_stress=np.zeros_like(mf.f_stress),
             v_flux=np.zeros_like(mf.v_flux), p_i=np.zeros_like(mf.p_i),
             wall=np.zeros_like(mf.wall))
    z.update(kw)
    return dataclasses.replace(mf, **z)


QUIET = dict(td="none", urf=1.0, p_drop=0.0, cell_force=False, stress=False)


def _rowsums(f, val):
    rows = np.repeat(np.arange(len(f.dia)), np.diff(f.row))
    return np.bincount(rows, weights=val, minlength=len(f.dia))


@pytest.mark.parametrize("blend", [False, True])
def test_matrix_rows_sum_to_zero_without_walls(cases, oracle, blend):
    f = cases.hex_faces(6, 5, 4)
    mf = cases.momentum_fields(f)
    mf = dataclasses.replace(mf, wall=np.zeros_like(mf.wall))
    val, _, _ = oracle.momentum_assemble(f, mf, blend_matrix=blend, **QUIET)
    d = val[f.dia - 1]
    assert np.all(np.abs(_rowsums(f, val)) <= 1e-13 * np.abs(d))
    if not blend:   # pure diffusion: a(c1, c2) = a(c2, c1) bit for bit
        internal = f.c2 > 0
        p = f.pos.reshape(-1, 2)[internal] - 1
        assert np.array_equal(val[p[:, 0]], val[p[:, 1]])


def test_implicit_upwind_antisymmetry_is_the_mass_flux(cases, oracle):
    """First-order upwinding: a(c2, c1) - a(c1, c2) = -F rho (uniform rho)."""
    f = cases.hex_faces(5, 4, 3)
    mf = cases.momentum_fields(f)
    mf = dataclasses.replace(mf, density=np.full_like(mf.density, 2.5))
    val, _, _ = oracle.momentum_assemble(f, mf, blend_matrix=True, **QUIET)
    internal = f.c2 > 0
    p = f.pos.reshape(-1, 2)[internal] - 1
    flux = mf.v_flux[internal] * 2.5
    assert np.allclose(val[p[:, 1]] - val[p[:, 0]], -flux, rtol=1e-14, atol=1e-30)


def _x_flux(f, nx, ny, nz, U=0.3):
    """Uniform +x volume flux through every x face (boundary faces too), so
    every cell's net flux is zero."""
    idx = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    xi = np.zeros(nx * ny * nz, int)
    xi[idx.ravel()] = np.repeat(np.arange(nx), ny * nz)
    flux = np.zeros(f.n_faces)
    c1 = f.c1 - 1
    internal = f.c2 > 0
    c2 = np.where(internal, f.c2 - 1, 0)
    along_x = internal & (xi[c2] == xi[c1] + 1)
    flux[along_x] = U
    # boundary x faces: a cell on x = 0 or x = nx-1 has one such face; which
    # boundary face is the x face is recovered from the face generator's order
    # (faces of the x = 0 side, then x = nx-1, then y, z sides)
    return flux, along_x


@pytest.mark.parametrize("adv", ["central", "upwind", "luds", "quick", "smart", "gamma",
                                 "minmod", "blended", "superbee", "avl_smart"])
def test_advection_of_a_constant_field_is_zero(cases, oracle, adv):
    """Divergence-free flux (zero on every boundary face, a closed
    circulation inside) and constant phi: every scheme's face value is the
    constant, so the advection source vanishes (central: to rounding)."""
    nx = ny = nz = 5
    f = cases.hex_faces(nx, ny, nz)
    mf = cases.momentum_fields(f)
    # circulation: flux = psi(c2) - psi(c1)... is a gradient, not divergence
    # free; use a discrete curl instead: loops of four faces around each
    # x-parallel edge carry +q, +q, -q, -q
    flux = np.zeros(f.n_faces)
    key = {}
    for s in range(f.n_faces):
        if f.c2[s] > 0:
            key[(f.c1[s] - 1, f.c2[s] - 1)] = s
    idx = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    rng = np.random.default_rng(3)
    for x in range(nx):
        for y in range(ny - 1):
            for z in range(nz - 1):
                q = rng.uniform(-1, 1)
                a, b_, c, d = idx[x, y, z], idx[x, y + 1, z], idx[x, y + 1, z + 1], idx[x, y, z + 1]
                for (u, v), sgn in (((a, b_), 1), ((b_, c), 1), ((d, c), -1), ((a, d), -1)):
                    flux[key[(u, v)]] += sgn * q
    const = np.full_like(mf.phi_n, 0.75)
    g0 = np.zeros_like(mf.gx)
    mf = _off(mf, v_flux=flux, phi_n=const, gx=g0, gy=g0, gz=g0)
    _, b, _ = oracle.momentum_assemble(f, mf, adv=adv, blend_matrix=True, **QUIET)
    tol = 1e-15 * np.abs(flux).max() * 1000.0 * 8
    assert np.abs(b).max() <= (tol if adv == "central" else 0.0)


def test_upwind_with_implicit_blend_leaves_no_explicit_part(cases, oracle):
    """UPWIND face values equal the implicit part's upwind values, so the
    deferred-correction source advect - upwind is exactly zero for any phi."""
    f = cases.hex_faces(6, 5, 4)
    mf = cases.momentum_fields(f)
    mf = _off(mf, v_flux=cases.momentum_fields(f).v_flux)
    _, b, _ = oracle.momentum_assemble(f, mf, adv="upwind", blend_matrix=True, **QUIET)
    assert np.array_equal(b, np.zeros_like(b))


def _two_cells(cases, oracle, r, adv, flux=1.0, blend=0.8):
    """Cells 1, 2 along z (one internal face, f = 0.5, d = (1, 0, 0));
    phi_1 = 1 (upwind), phi_2 = 2, grad phi_1 = ((1 + r)/2, 0, 0) so that
    the far-upwind estimate phi_2 - 2 grad.d is 1 - r and the TVD ratio
    (phi_C - phi_U)/(phi_D - phi_C) is r; boundary values wide apart so that
    the min/max clamp is inactive.  Returns the face value phi_f = -b_1."""
    f = cases.hex_faces(1, 1, 2, shuffle=False)
    mf = cases.momentum_fields(f)
    nb = mf.n_bnd
    internal = np.flatnonzero(f.c2 > 0)[0]
    one = np.zeros_like
    phi = np.zeros(nb + 1 + 2)
    # boundary values (cell c < 0 is element nb + c): cell 1's boundary
    # neighbours span [-100, 100], so its min/max clamp is inactive
    b1 = [s for s in range(f.n_faces) if f.c2[s] < 0 and f.c1[s] == 1]
    phi[:nb] = 100.0
    phi[nb + f.c2[b1[0]]] = -100.0
    phi[nb + 1], phi[nb + 2] = 1.0, 2.0
    gx = np.zeros_like(phi)
    gx[nb + 1] = (1.0 + r) / 2.0
    v = np.zeros(f.n_faces)
    v[internal] = flux
    fa = np.full(f.n_faces, 0.5)
    d = np.zeros(f.n_faces)
    d[internal] = 1.0
    if f.c1[internal] != 1:
        raise AssertionError("face orientation")
    mf = _off(mf, v_flux=v, phi_n=phi, gx=gx, gy=one(gx), gz=one(gx), f=fa, dx=d,
              dy=np.zeros_like(d), dz=np.zeros_like(d), density=np.ones_like(phi))
    _, b, _ = oracle.momentum_assemble(f, mf, adv=adv, blend_matrix=False, blend=blend, **QUIET)
    return -b[0]


PSI = {"central": lambda r: 1.0, "upwind": lambda r: 0.0, "luds": lambda r: r,
       "quick": lambda r: (3.0 + r) / 4.0,
       "minmod": lambda r: max(0.0, min(r, 1.0)),
       "superbee": lambda r: max(0.0, min(2.0 * r, 1.0), min(r, 2.0))}


@pytest.mark.parametrize("adv", sorted(PSI))
@pytest.mark.parametrize("r", [-1.0, 0.0, 0.5, 1.0, 1.5, 2.0, 3.0])
def test_limiters_are_the_textbook_psi(cases, oracle, adv, r):
    """phi_f = phi_C + psi(r)/2 (phi_D - phi_C) on a uniform face (fj = 1/2)."""
    got = _two_cells(cases, oracle, r, adv)
    assert got == pytest.approx(1.0 + 0.5 * PSI[adv](r), rel=1e-15, abs=1e-15)


@pytest.mark.parametrize("adv", ["smart", "avl_smart"])
@pytest.mark.parametrize("r", [-2.0, -0.5, 0.0])
def test_bounded_limiters_upwind_for_nonpositive_r(cases, oracle, adv, r):
    assert _two_cells(cases, oracle, r, adv) == 1.0


def test_gamma_regimes(cases, oracle):
    """Jasak's GAMMA: normalised phi_C = 1 - (phi_D - phi_C)/(2 grad.d);
    upwind outside (0, 1), central for beta <= phi_C < 1.  With phi_C = 1,
    phi_D = 2: grad.d = (1+r)/2, normalised value 1 - 1/(1+r)."""
    assert _two_cells(cases, oracle, -0.5, "gamma") == 1.0            # normalised < 0
    assert _two_cells(cases, oracle, 3.0, "gamma") == pytest.approx(1.5, rel=1e-15)  # 0.75: central


@pytest.mark.parametrize("w,expect", [(1.0, 1.5), (0.0, 1.0)])
def test_blended_is_a_mix_of_central_and_upwind(cases, oracle, w, expect):
    assert _two_cells(cases, oracle, 0.5, "blended", blend=w) == pytest.approx(expect, rel=1e-15)


@pytest.mark.parametrize("td", ["linear", "parabolic"])
def test_time_schemes_steady_state(cases, oracle, td):
    """Inertia alone: diag = c rho V / dt, and b / diag is the backward Euler
    (phi_o) or BDF2 ((4 phi_o - phi_oo) / 3) predictor."""
    f = cases.hex_faces(4, 3, 3)
    mf = _off(cases.momentum_fields(f))
    val, b, _ = oracle.momentum_assemble(f, mf, td=td, urf=1.0, p_drop=0.0, cell_force=False,
                                         stress=False)
    d = val[f.dia - 1]
    expect = mf.phi_o if td == "linear" else (4.0 * mf.phi_o - mf.phi_oo) / 3.0
    assert np.allclose(b / d, expect, rtol=1e-14, atol=1e-14)


def test_under_relaxation_keeps_the_residual_of_phi_n(cases, oracle):
    f = cases.hex_faces(5, 4, 3)
    mf = cases.momentum_fields(f)
    phi = mf.phi_n[mf.n_bnd + 1:]

    def residual(urf):
        val, b, _ = oracle.momentum_assemble(f, mf, adv="minmod", urf=urf, td="parabolic")
        A = oracle.Matrix(f.row, f.col, f.dia, val)
        return oracle.residual(A, b, phi)
    r1, r7 = residual(1.0), residual(0.7)
    assert np.allclose(r7, r1, rtol=1e-9, atol=1e-9 * np.abs(r1).max())


def test_cross_diffusion_internal_faces_cancel(cases, oracle):
    """Each internal face adds +/- the same cross term to its two cells: the
    sum over cells does not depend on the internal faces' viscosity."""
    f = cases.hex_faces(5, 4, 3)
    mf = cases.momentum_fields(f)
    mf = _off(mf, vis_eff=mf.vis_eff, f_stress=mf.f_stress)
    _, b1, _ = oracle.momentum_assemble(f, mf, blend_matrix=False, td="none", urf=1.0,
                                        p_drop=0.0, cell_force=False, stress=True)
    vis2 = np.where(f.c2 > 0, 3.0 * mf.vis_eff, mf.vis_eff)
    _, b2, _ = oracle.momentum_assemble(f, dataclasses.replace(mf, vis_eff=vis2),
                                        blend_matrix=False, td="none", urf=1.0, p_drop=0.0,
                                        cell_force=False, stress=True)
    scale = np.abs(b1).max()
    assert abs(b1.sum() - b2.sum()) <= 1e-12 * scale
    assert not np.allclose(b1, b2)
