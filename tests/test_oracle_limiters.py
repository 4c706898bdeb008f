"""Pins of the oracle's advection-scheme face values (the momentum assembly's
limiters, Numerics_Mod/Advection_Scheme.f90:22-183) by properties the
schemes are defined by, not by their formulas:

* linear data: every scheme that is second order on a uniform grid (QUICK,
  LUDS, MINMOD, SMART, AVL_SMART) gives the linear interpolant at the face,
  for any face position (SUPERBEE only at the mid-point);
* quadratic data in the schemes' smooth range: QUICK, SMART and AVL_SMART
  are the quadratic upwind interpolation, exact for quadratics at the face
  position (the upstream value is extrapolated with the exact gradient);
* a local extremum at the upwind cell (r <= 0): the upwind value;
* uniform grids: the published normalised-variable (NVD) forms of SMART
  (Gaskell and Lau, 1988), MINMOD and SUPERBEE over a sweep of the
  normalised upwind value;
* AVL_SMART on non-uniform faces: bounded between the upwind and
  downwind values (and continuous) over the same sweep.

A mistake such as a swapped g_d / g_u, a wrong beta, a dropped sign for
negative fluxes or a missing clip fails one of them.
"""
import numpy as np
import pytest

TOL = 1e-13


def _face(oracle, adv, f, flux, x1, x2, phi, dphi):
    """Cells at x1, x2 on the x axis with values phi(x1), phi(x2) and exact
    gradients dphi; returns (phi_f, face position)."""
    ph = [phi(x1), phi(x2)]
    g = [dphi(x1), 0.0, 0.0, dphi(x2), 0.0, 0.0]
    v = oracle.adv_face(adv, f, flux, ph, g, [x2 - x1, 0.0, 0.0])
    return v, x1 + (1.0 - f) * (x2 - x1)


@pytest.mark.parametrize("adv", ["quick", "luds", "minmod", "smart", "avl_smart"])
@pytest.mark.parametrize("f", [0.5, 0.3, 0.72])
@pytest.mark.parametrize("flux", [1.0, -1.0])
def test_linear_data_exact(oracle, adv, f, flux):
    for a, b in ((0.3, 2.0), (-1.0, -0.7)):
        v, xf = _face(oracle, adv, f, flux, 1.0, 1.8, lambda x: a + b * x, lambda x: b)
        assert abs(v - (a + b * xf)) <= TOL * max(1.0, abs(a + b * xf)), (adv, f, flux, v)


@pytest.mark.parametrize("flux", [1.0, -1.0])
def test_superbee_linear_at_midpoint(oracle, flux):
    v, xf = _face(oracle, "superbee", 0.5, flux, 0.0, 1.0, lambda x: 1.0 + 3.0 * x, lambda x: 3.0)
    assert abs(v - (1.0 + 3.0 * xf)) <= TOL * 4


@pytest.mark.parametrize("adv", ["quick", "smart", "avl_smart"])
@pytest.mark.parametrize("f", [0.5, 0.7, 0.6])
def test_quadratic_data_exact(oracle, adv, f):
    """phi = x^2 + x with the upwind cell at 1 and the downwind cell at 2
    (flux > 0) or mirrored (flux < 0): r = 1/2, inside the QUICK range of
    SMART and AVL_SMART for these face positions."""
    phi, dphi = (lambda x: x * x + x), (lambda x: 2.0 * x + 1.0)
    v, xf = _face(oracle, adv, f, 1.0, 1.0, 2.0, phi, dphi)
    assert abs(v - phi(xf)) <= TOL * phi(xf), (adv, f, v, phi(xf))
    # mirrored: the flux runs from cell 2 (at x = 1) to cell 1 (at x = 2)
    psi, dpsi = (lambda x: phi(3.0 - x)), (lambda x: -dphi(3.0 - x))
    v, xf = _face(oracle, adv, 1.0 - f, -1.0, 1.0, 2.0, psi, dpsi)
    assert abs(v - psi(xf)) <= TOL * psi(xf), (adv, f, v, psi(xf))


@pytest.mark.parametrize("adv", ["minmod", "smart", "avl_smart", "superbee"])
def test_extremum_is_upwind_for_limiters(oracle, adv):
    """Upwind value 0 at x=0 (extrapolated), cell value 2 at x=1, downwind 1
    at x=2: r < 0.  The limited schemes return the upwind cell's value."""
    # gradient at the upwind cell chosen so that the extrapolated upstream
    # value phi_d - 2 g d is 0
    v = oracle.adv_face(adv, 0.4, 1.0, [2.0, 1.0], [0.5, 0, 0, 0, 0, 0], [1.0, 0, 0])
    assert v == 2.0


def _nvd(oracle, adv, t, f=0.5, flux=1.0):
    """phi~_f of the normalised variable phi~_C = t: phi_U = 0, phi_D = 1,
    phi_C = t, with the gradient that extrapolates phi_U exactly."""
    d = 1.0
    g = 0.5 / d  # phi_star = phi_D - 2 g d = 0
    if flux > 0:
        return oracle.adv_face(adv, f, flux, [t, 1.0], [g, 0, 0, 0, 0, 0], [d, 0, 0])
    return oracle.adv_face(adv, f, flux, [1.0, t], [0, 0, 0, -g, 0, 0], [d, 0, 0])


SWEEP = [-0.4, -0.1, 0.0, 0.05, 0.1, 0.15, 0.2, 0.3, 0.45, 0.55, 0.6, 0.7, 0.8, 0.85, 0.9,
         0.97, 1.2, 1.5]


def _smart_nvd(t):      # Gaskell and Lau (1988)
    if 0.0 < t < 1.0 / 6.0:
        return 3.0 * t
    if 1.0 / 6.0 <= t <= 5.0 / 6.0:
        return 3.0 / 8.0 + 3.0 / 4.0 * t
    if 5.0 / 6.0 < t < 1.0:
        return 1.0
    return t


def _minmod_nvd(t):
    if 0.0 < t <= 0.5:
        return 1.5 * t
    if 0.5 < t < 1.0:
        return 0.5 + 0.5 * t
    return t


def _superbee_nvd(t):
    if 0.0 < t <= 1.0 / 3.0:
        return 2.0 * t
    if 1.0 / 3.0 < t <= 0.5:
        return 0.5 + 0.5 * t
    if 0.5 < t <= 2.0 / 3.0:
        return 1.5 * t
    if 2.0 / 3.0 < t < 1.0:
        return 1.0
    return t


@pytest.mark.parametrize("adv,ref", [("smart", _smart_nvd), ("minmod", _minmod_nvd),
                                     ("superbee", _superbee_nvd)])
@pytest.mark.parametrize("flux", [1.0, -1.0])
def test_uniform_grid_nvd(oracle, adv, ref, flux):
    for t in SWEEP:
        v = _nvd(oracle, adv, t, flux=flux)
        assert abs(v - ref(t)) <= 1e-12, (adv, flux, t, v, ref(t))


@pytest.mark.parametrize("f", [0.5, 0.35, 0.65])
def test_avl_smart_bounded_and_continuous(oracle, f):
    ts = np.linspace(-0.5, 1.5, 401)
    ts = ts[np.abs(ts - 1.0) > 1e-9]
    v = np.array([_nvd(oracle, "avl_smart", t, f=f) for t in ts])
    inside = (ts >= 0.0) & (ts <= 1.0)
    assert np.all(v[inside] >= ts[inside] - 1e-14) and np.all(v[inside] <= 1.0 + 1e-14)
    assert np.all(v[~inside] == ts[~inside])
    # piecewise linear in t with no jumps inside (0, 1)
    jumps = np.abs(np.diff(v[inside]))
    assert jumps.max() < 0.05
