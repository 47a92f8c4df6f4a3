"""Pins the oracle's traced states, wave speeds, HLLC fluxes, geometric
pressure source and dust flux by closed forms:

* set_state on linear profiles (limiter and flattening inactive) is the
  textbook half-step characteristic predictor:
    right state (upwind from outside the cell): rho* = rho_L (1 - dt/2 du/dx),
      p* = p_L (1 - gam dt/2 du/dx), u* = u_L - dt/2 (dp/dx)/rho_L;
    left state: the same applied to the mean over the swept region;
* get_pm_simple and get_sm reduce to Toro's PVRS estimates for states of
  equal impedance;
* Einfeldt speeds of identical states are u -/+ c;
* HLLC is consistent (identical states give the physical flux), upwinds
  supersonic states, and a stationary contact carries no mass;
* net_source's geometric pressure term gives -dp/dR per unit volume for a
  linear pressure profile at rest in cylindrical geometry;
* dust_flux of a uniform moving dust is the upwind flux rho*u.
"""
import numpy as np
import pytest

import oracle


def linear_row(n, x0, dx, a, g):
    """cell means of a + g*(x - x0) on a uniform grid (equal to centre values)."""
    xa = x0 + dx * np.arange(n + 1)
    xc = 0.5 * (xa[1:] + xa[:-1])
    return xa, a + g * (xc - x0)


@pytest.mark.parametrize("geom", [0, 3])
def test_set_state_linear_predictor(geom):
    n, dx, dt, gam = 11, 0.01, 2e-3, 1.4
    i = 5
    xa, r = linear_row(n, 0.0, dx, 1.0, 0.5)
    _, p = linear_row(n, 0.0, dx, 1.0, 0.8)
    _, u = linear_row(n, 0.0, dx, 0.2, 1.5)   # u > 0 and increasing: no flattening
    _, v = linear_row(n, 0.0, dx, 0.1, -0.7)
    w = np.zeros(n)
    flat = np.array([oracle.flatten(gam, r[k - 2:k + 3], p[k - 2:k + 3], u[k - 2:k + 3]) if 2 <= k < n - 2 else 0.0
                     for k in range(n)])
    assert not flat.any()
    S = oracle.set_state_row(geom, gam, xa, r, p, u, v, w, flat, i, dt)
    rl, pl, ul, vl, wl, rr, pr, ur, vr, wr = S
    gr, gp, gu, gv = 0.5, 0.8, 1.5, -0.7
    xe = xa[i]
    f = lambda a, g_: a + g_ * xe                      # edge value
    # right state: u_edge > 0, so the swept region is empty (dis = 0)
    rL, pL, uL = f(1.0, gr), f(1.0, gp), f(0.2, gu)
    cs = np.sqrt(gam * pL / rL)
    assert rr == pytest.approx(rL * (1 - 0.5 * dt * gu), rel=1e-12)
    assert pr == pytest.approx(pL * (1 - gam * 0.5 * dt * gu), rel=1e-12)
    assert ur == pytest.approx(uL - 0.5 * dt * gp / rL, rel=1e-12)
    assert vr == pytest.approx(f(0.1, gv) + gv * 0.5 * cs * dt, rel=1e-12)
    # left state: mean over the region swept by the flow in dt, then the
    # same acoustic correction
    c_u = lambda a, g_, L: a + g_ * (xe - 0.5 * L)     # mean over [xe - L, xe]
    cs_l = np.sqrt(gam * p[i - 1] / r[i - 1])
    ul_dd = c_u(0.2, gu, (u[i - 1] + cs_l) * dt)      # u averaged over the fast-wave domain
    L = ul_dd * dt
    r0, p0, u0 = c_u(1.0, gr, L), c_u(1.0, gp, L), c_u(0.2, gu, L)
    assert rl == pytest.approx(r0 * (1 - 0.5 * dt * gu), rel=1e-12)
    assert pl == pytest.approx(p0 * (1 - gam * 0.5 * dt * gu), rel=1e-12)
    assert ul == pytest.approx(u0 - 0.5 * dt * gp / r0, rel=1e-12)


def test_set_state_uniform():
    n, i = 11, 5
    xa = np.linspace(0.0, 0.11, n + 1)
    S = oracle.set_state_row(0, 1.4, xa, np.full(n, 2.0), np.full(n, 3.0), np.full(n, 0.3),
                             np.full(n, 0.1), np.zeros(n), np.zeros(n), i, 1e-3)
    assert np.allclose(S, [2.0, 3.0, 0.3, 0.1, 0.0, 2.0, 3.0, 0.3, 0.1, 0.0], rtol=1e-14)


def S10(rl, pl, ul, rr, pr, ur, vl=0.0, vr=0.0):
    return [rl, pl, ul, vl, 0.0, rr, pr, ur, vr, 0.0]


def test_pvrs_equal_impedance():
    gam = 1.4
    rho, p = 1.3, 0.9
    c = np.sqrt(gam * p / rho)
    for ul, ur in [(0.1, -0.05), (-0.2, 0.3), (0.0, 0.0)]:
        out = oracle.riemann_hllc(gam, S10(rho, p, ul, rho, p, ur))
        pm, sl, sm, sr = out[:4]
        # Toro (2009) eq. 9.20: p* = (pl+pr)/2 - (ur-ul) rho_bar c_bar / 2
        assert pm == pytest.approx(max(p - 0.5 * (ur - ul) * rho * c, 1e-6 * p), rel=1e-12)
        assert sm == pytest.approx(0.5 * (ul + ur), rel=1e-12, abs=1e-15)
    # pressure jump at rest: u* = (pl - pr) / (2 rho_bar c_bar) (PVRS)
    pl, pr = 1.1, 0.9
    out = oracle.riemann_hllc(1.0, S10(1.0, pl, 0.0, 1.0, pr, 0.0))
    cl, cr = np.sqrt(pl), np.sqrt(pr)
    assert out[2] == pytest.approx((pl - pr) / (2.0 * 1.0 * 0.5 * (cl + cr)), rel=1e-12)


def test_einfeldt_identical_states():
    gam = 1.4
    out = oracle.riemann_hllc(gam, S10(1.0, 1.0, 0.3, 1.0, 1.0, 0.3))
    c = np.sqrt(gam)
    assert out[1] == pytest.approx(0.3 - c, rel=1e-14)
    assert out[3] == pytest.approx(0.3 + c, rel=1e-14)


def test_hllc_consistency_and_upwinding():
    gam = 1.0
    # identical states: physical flux rho u, rho u^2, rho u v, pressure p
    for u in (-0.3, 0.0, 0.4):
        out = oracle.riemann_hllc(gam, S10(2.0, 1.0, u, 2.0, 1.0, u, 0.7, 0.7))
        Fr, Fp, Fu, Fv, Fw, pres = out[4:]
        assert Fr == pytest.approx(2.0 * u, abs=1e-15)
        assert Fu == pytest.approx(2.0 * u * u, abs=1e-15)
        assert Fv == pytest.approx(2.0 * u * 0.7, abs=1e-15)
        assert pres == pytest.approx(1.0, rel=1e-14)
        assert Fp == 0.0
    # supersonic to the right: left state's flux
    out = oracle.riemann_hllc(gam, S10(1.0, 1.0, 5.0, 0.5, 0.3, 4.0))
    assert out[1] >= 0 and out[4] == pytest.approx(5.0) and out[9] == 1.0
    # supersonic to the left: right state's flux
    out = oracle.riemann_hllc(gam, S10(1.0, 1.0, -5.0, 0.5, 0.3, -4.0))
    assert out[3] < 0 and out[4] == pytest.approx(0.5 * -4.0) and out[9] == 0.3
    # stationary contact: no mass flux, pressure continuous
    out = oracle.riemann_hllc(gam, S10(3.0, 1.0, 0.0, 1.0, 1.0, 0.0))
    assert out[2] == 0.0 and out[4] == 0.0 and out[9] == pytest.approx(1.0)


def test_pressure_source_cylindrical():
    """A linear pressure p = p0 + g R at rest with uniform density, in
    cylindrical geometry: the momentum change per unit volume and time tends
    to -g as dt -> 0 (the half-sum of edge radii times the pressure jump
    divided by the annulus volume (R+^2 - R-^2)/2)."""
    n = 13
    xa = 1.0 + 0.01 * np.arange(n + 1)
    xc = 0.5 * (xa[1:] + xa[:-1])
    xv = 0.5 * (xa[1:] ** 2 - xa[:-1] ** 2)
    g = -0.3
    p = 1.0 + g * (xc - 1.0)   # cell means of a linear p are not centre values in
    # cylindrical geometry; use the exact weighted means
    p = np.array([(1.0 - g) + g * (2.0 / 3.0) * (xa[k + 1] ** 3 - xa[k] ** 3) / (xa[k + 1] ** 2 - xa[k] ** 2)
                  for k in range(n)])
    r = np.ones(n)
    u = np.zeros(n)
    for dt in (1e-6, 1e-8):
        D = oracle.riemann_row(1, 1.0, xa, r, p, u, np.zeros(n), np.zeros(n), dt)
        acc = D[2][3:n - 3] / xv[3:n - 3]
        assert np.allclose(acc, -g, rtol=1e-4), acc


def test_flux_difference_telescopes():
    """Row sum of the mass changes equals the net flux through the two
    outermost interfaces (conservative differencing)."""
    rng = np.random.default_rng(5)
    n = 24
    xa = np.linspace(0.0, 0.24, n + 1)
    r = 1.0 + 0.1 * rng.random(n)
    p = 1.0 + 0.1 * rng.random(n)
    u = 0.05 * rng.standard_normal(n)
    # make the first and last 6 cells identical so their interface fluxes match
    for a in (r, p, u):
        a[:7] = a[7]
        a[-7:] = a[-8]
    a_edge = oracle.riemann_row(0, 1.0, xa, r, p, u, np.zeros(n), np.zeros(n), 1e-3)
    assert abs(a_edge[0][3:n - 3].sum() - (r[7] * u[7] - r[-8] * u[-8])) < 1e-12


def test_dust_flux_uniform():
    n = 11
    xa = np.linspace(0.0, 0.11, n + 1)
    dt = 1e-2                       # dt/dx = 1 > |u|: the donor cap is inactive
    for u in (0.3, -0.3):
        F = oracle.dust_flux(0, xa, np.full(n, 0.02), np.full(n, u), np.full(n, 0.5), np.zeros(n), 5, dt)
        assert F[0] == pytest.approx(0.02 * u, rel=1e-14)
        assert F[1] == pytest.approx(0.02 * u * u, rel=1e-14)
        assert F[2] == pytest.approx(0.02 * u * 0.5, rel=1e-14)
    # dust at rest: no flux
    F = oracle.dust_flux(0, xa, np.full(n, 0.02), np.zeros(n), np.full(n, 0.5), np.zeros(n), 5, dt)
    assert np.all(F == 0.0)
    # the reference caps the donor flux at rho*dt/dx (dust.cu:96-97), which
    # binds once |u| exceeds dt/dx
    dt = 1e-3
    for u, sgn in ((0.3, 1.0), (-0.3, -1.0)):
        F = oracle.dust_flux(0, xa, np.full(n, 0.02), np.full(n, u), np.full(n, 0.5), np.zeros(n), 5, dt)
        assert F[0] == pytest.approx(sgn * 0.02 * dt / 0.01, rel=1e-12)
