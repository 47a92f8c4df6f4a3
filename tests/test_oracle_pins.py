"""Pins of oracle functions by what the mathematics fixes, independent of
the reading the oracle itself was written from.

* dust drag (sweeps.cu:665-711, get_t_stop dust.cu:1-4): with the gas state
  frozen over the kick and a force that does not depend on the dust
  velocity, du/dt = g + (u_g - u)/t_s is linear with constant coefficients,
  and the reference's exponential update must be its exact solution
  u_g + g t_s + (u - u_g - g t_s) exp(-dt/t_s); at the unperturbed disk's
  density the stopping time t_s is St / Omega_K(R) (the Stokes number's
  definition), and it scales as 1/rho (Epstein drag).
* dust initial drift (init.cu:365-380): the radial drift u_d solves the
  steady equations of a drifting dust fluid with Keplerian scaling, the
  radial u - (St/2) u^2 = St (v^2 - 1) and the azimuthal v (1 + St u/2) = f
  (in units of v_K, f = v_gas/v_K), solved here by Newton's method.
* the 3D spline-softened planet (forces.cu:14-39): the direct force is
  -grad Phi with Phi(d) = -m/d outside the softening length r_s and
  m (2 d^2/r_s^3 - d^3/r_s^4) - 2m/r_s inside (continuous with its
  derivative at d = r_s); the indirect term is -grad of m r.r_p/a^3 in its
  radial component, while its azimuthal component carries an extra
  sin(theta) and its polar one an extra 1/sin(theta) (forces.cu:35-37),
  which the oracle keeps (within the +-3h polar range sin(theta) > 0.988).
* the internal-energy equation (riemann.cu:65-73, HLLC.cu:58-68): for a
  uniform gas with the radial velocity u = a r, one radial sweep gives
  drho/dt = -3 a rho and dp/dt = -3 gamma a p (the p dV work included) to
  first order in dt; without the p dV term the pressure would follow
  -3 a p.  The interface velocity is the contact speed of a moving contact
  and 0 in a symmetric collision.
* flatten (PPM.cu:44-54): its switch is Colella & Woodward's (1984)
  flattening with omega_1 = 0.75, omega_2 = 10: zero for
  dp1/dp2 <= 0.75, 10 (dp1/dp2 - 0.75) above, times the jump strength
  max(|dp1f|, |dp1b|)/p and capped at 1/2; the reference's |ratio - 1/2|
  extends the dead band symmetrically to ratios below 0.25.
"""
import numpy as np
import pytest

import oracle
from penguin_amd import config_c2, disk2d, disk3d


def _radii(o):
    xa, _, _, _ = o.coords()
    return 0.5 * (xa[3:-4] + xa[4:-3])


# --------------------------------------------------------------- dust drag
def _drag_setup(dens_factor=1.0, vd_factor=1.0, seed=4):
    p = config_c2(precision=8, xres=48, yres=16).replace(planet_mass=0.0, visc_flag=0, kill_flag=0,
                                                         noise_rho=0.0, noise_vr=0.0)
    o = oracle.Oracle(p)
    g, dd = o.get_state()
    R = _radii(o)[None, :]
    rng = np.random.default_rng(seed)
    cs2 = g[1] / g[0]
    g[0] = g[0] * dens_factor
    g[1] = cs2 * g[0]
    g[2] = 1e-3 * rng.uniform(-1.0, 1.0, g[2].shape)       # gas radial velocity
    vk = np.sqrt(1.0 / R) * np.ones_like(g[3])
    g[3] = vk * (1.0 + 1e-3 * rng.uniform(-1.0, 1.0, g[3].shape))
    dd[1] = 1e-2 * rng.uniform(-1.0, 1.0, dd[1].shape)
    dd[2] = vk * vd_factor
    o.set_state(g, dd)
    return p, o, g, dd, R


@pytest.mark.parametrize("dens", [1.0, 2.5])
def test_dust_drag_is_the_exact_relaxation(dens):
    p, o, g, dd, R = _drag_setup(dens_factor=dens)
    ts = p.Stokes * R ** 1.5 / dens            # St / Omega_K at the disk's density, / rho ratio
    dt = 0.3 * float(ts.min())
    o.sources(0.0, dt, inplace=True)
    _, d1 = o.get_state()
    e = np.exp(-dt / ts)
    # Keplerian dust: no net radial force, so u relaxes to u_gas
    u_exp = g[2] + (dd[1] - g[2]) * e
    v_exp = g[3] + (dd[2] - g[3]) * e
    assert np.abs(d1[1] - u_exp).max() <= 1e-12 * np.abs(dd[1]).max()
    assert np.abs(d1[2] - v_exp).max() <= 1e-12 * np.abs(dd[2]).max()
    o.close()


def test_dust_drag_with_a_constant_force_tends_to_terminal_drift():
    p, o, g, dd, R = _drag_setup(vd_factor=0.97)
    ts = p.Stokes * R ** 1.5
    gacc = dd[2] ** 2 / R - 1.0 / R ** 2     # the dust's net radial acceleration, fixed over the kick
    u_term = g[2] + gacc * ts
    for dt in (0.05 * float(ts.min()), 2.0 * float(ts.max())):
        o.set_state(g, dd)
        o.sources(0.0, dt, inplace=True)
        _, d1 = o.get_state()
        u_exp = u_term + (dd[1] - u_term) * np.exp(-dt / ts)
        assert np.abs(d1[1] - u_exp).max() <= 1e-11 * np.abs(u_term).max()
    # started at the terminal drift, it stays there
    d2 = dd.copy()
    d2[1] = u_term
    o.set_state(g, d2)
    o.sources(0.0, float(ts.min()), inplace=True)
    _, d1 = o.get_state()
    assert np.abs(d1[1] - u_term).max() <= 1e-12 * np.abs(u_term).max()
    o.close()


# ------------------------------------------------------ dust initial drift
@pytest.mark.parametrize("stokes, tol", [(0.01, 1e-8), (0.3, 1e-8), (1.0, 2e-6)])
def test_initial_dust_drift_solves_the_steady_equations(stokes, tol):
    """init.cu:365-380 runs two passes of a fixed-point iteration, so the
    drift solves the equations to the accuracy of two passes: 1e-8 relative
    up to St 0.3, 8e-7 at St 1 (dropping the advection term u^2 St/2, the
    smallest one, would move it by 2.5e-7 at St 0.01 and 6e-4 at St 1)."""
    p = config_c2(precision=8, xres=64, yres=8).replace(Stokes=stokes)
    o = oracle.Oracle(p)
    g, dd = o.get_state()
    R = _radii(o)
    vk = np.sqrt(1.0 / R)
    f = g[3][0] / vk
    # Newton on F(u) = u - St u^2/2 - St (v(u)^2 - 1), v(u) = f / (1 + St u/2)
    St = stokes
    u = St * (f * f - 1.0)
    for _ in range(60):
        v = f / (1.0 + 0.5 * St * u)
        F = u - 0.5 * St * u * u - St * (v * v - 1.0)
        dv = -v * 0.5 * St / (1.0 + 0.5 * St * u)
        dF = 1.0 - St * u - 2.0 * St * v * dv
        u = u - F / dF
    got = dd[1][0] / vk
    print(f"St {stokes}: drift/v_K {got.min():.6e} .. {got.max():.6e}", flush=True)
    assert np.all(got < 0.0)
    assert np.abs(got - u).max() <= tol * np.abs(u).max()
    o.close()


# ------------------------------------------------------ 3D planet potential
def _phi_planet(m, a, phip, rs, r, phi, th):
    x, y, z = r * np.sin(th) * np.cos(phi), r * np.sin(th) * np.sin(phi), r * np.cos(th)
    px, py = a * np.cos(phip), a * np.sin(phip)
    d = np.sqrt((x - px) ** 2 + (y - py) ** 2 + z ** 2)
    inside = m * (2.0 * d * d / rs ** 3 - d ** 3 / rs ** 4) - 2.0 * m / rs
    direct = np.where(d < rs, inside, -m / d)
    indirect = m * r * np.sin(th) * np.cos(phi - phip) / (a * a)
    return direct, indirect


def _grad_sph(fn, r, phi, th, h=1e-6):
    gr = (fn(r + h, phi, th) - fn(r - h, phi, th)) / (2 * h)
    gp = (fn(r, phi + h, th) - fn(r, phi - h, th)) / (2 * h) / (r * np.sin(th))
    gt = (fn(r, phi, th + h) - fn(r, phi, th - h)) / (2 * h) / r
    return -np.array([gr, gp, gt])


def test_spherical_planet_force_is_minus_the_gradient_of_its_potential():
    m, a, phip, rs = 1e-3, 1.0, np.pi, 0.08
    body = np.array([m, a, 0.0, a, phip, 0.5 * np.pi, 0.0, 0.0, 0.0, rs])
    rng = np.random.default_rng(2)
    pts = [(1.0 + 0.03, np.pi + 0.02, 0.5 * np.pi + 0.01),      # inside the softening sphere
           (1.0 - 0.05, np.pi - 0.03, 0.5 * np.pi - 0.02)]
    pts += [(rng.uniform(0.5, 2.0), rng.uniform(0, 2 * np.pi), 0.5 * np.pi + rng.uniform(-0.15, 0.15))
            for _ in range(20)]
    n_inside = 0
    for r, phi, th in pts:
        f = oracle.planet_force(body, r, phi, th, sph=True)
        gd = _grad_sph(lambda *x: _phi_planet(m, a, phip, rs, *x)[0], r, phi, th)
        gi = _grad_sph(lambda *x: _phi_planet(m, a, phip, rs, *x)[1], r, phi, th)
        expect = gd + gi * np.array([1.0, np.sin(th), 1.0 / np.sin(th)])
        scale = np.abs(gd).max() + np.abs(gi).max()
        d = np.sqrt(r * r + a * a - 2 * a * r * np.sin(th) * np.cos(phi - phip))
        n_inside += d < rs
        assert np.abs(f - expect).max() <= 1e-7 * scale, (r, phi, th, f, expect)
    assert n_inside == 2
    # the softened potential and its derivative are continuous at d = r_s
    for eps in (1e-7,):
        r0 = a + rs
        fin = oracle.planet_force(body, r0 - eps, phip, 0.5 * np.pi, sph=True)
        fout = oracle.planet_force(body, r0 + eps, phip, 0.5 * np.pi, sph=True)
        assert abs(fin[0] - fout[0]) <= 1e-5 * abs(fout[0])


def test_cylindrical_planet_force_is_minus_the_gradient_of_its_potential():
    m, a, phip, rs = 3e-4, 1.0, np.pi, 0.03
    body = np.array([m, a, 0.0, a, phip, 0.5 * np.pi, 0.0, 0.0, 0.0, rs])
    rng = np.random.default_rng(5)
    for _ in range(20):
        R, phi = rng.uniform(0.5, 2.0), rng.uniform(0, 2 * np.pi)

        def pot(R_, phi_):
            d2 = R_ * R_ + a * a - 2 * a * R_ * np.cos(phi_ - phip) + rs * rs
            return -m / np.sqrt(d2) + m * R_ * np.cos(phi_ - phip) / (a * a)
        h = 1e-6
        gR = -(pot(R + h, phi) - pot(R - h, phi)) / (2 * h)
        gp = -(pot(R, phi + h) - pot(R, phi - h)) / (2 * h) / R
        f = oracle.planet_force(body, R, phi, 0.0, sph=False)
        scale = abs(gR) + abs(gp)
        assert abs(f[0] - gR) <= 1e-7 * scale and abs(f[1] - gp) <= 1e-7 * scale and f[2] == 0.0


# ------------------------------------------------- internal-energy equation
@pytest.mark.parametrize("gam", [1.4, 5.0 / 3.0])
def test_radial_sweep_of_a_linear_flow_does_p_dV_work(gam):
    p = disk3d(40, 8, 4, precision=8, gam=gam, noise_rho=0.0)
    o = oracle.Oracle(p)
    g, _ = o.get_state()
    xa, _, _, _ = o.coords()
    r1, r2 = xa[3:-4], xa[4:-3]
    a = 0.01
    rho0, p0 = 1.3, 2e-3
    g[0] = rho0
    g[1] = p0
    # cell means of a r under the spherical volume weight r^2 dr, which the
    # fourth-order reconstruction reproduces exactly at the cell edges
    ubar = a * 0.75 * (r2 ** 4 - r1 ** 4) / (r2 ** 3 - r1 ** 3)
    g[2] = ubar[None, None, :] * np.ones_like(g[2])
    g[3] = 0.0
    g[4] = 0.0
    o.set_state(g)
    out = {}
    for dt in (1e-4, 5e-5):
        o.set_state(g)
        o.sweepx(dt)
        s, _ = o.get_state()
        inner = (slice(None), slice(None), slice(5, -5))
        out[dt] = ((s[0][inner] / rho0 - 1.0) / (-3 * a * dt), (s[1][inner] / p0 - 1.0) / (-3 * gam * a * dt))
    for dt, (kr, kp) in out.items():
        print(f"gamma {gam} dt {dt}: drho ratio {kr.min():.9f}..{kr.max():.9f}, dp ratio {kp.min():.9f}..{kp.max():.9f}")
        assert np.abs(kr - 1.0).max() <= 2e-3 and np.abs(kp - 1.0).max() <= 2e-3
    # the deviation is first order in dt (halving dt halves it)
    e1 = np.abs(out[1e-4][1] - 1.0).max()
    e2 = np.abs(out[5e-5][1] - 1.0).max()
    assert e2 <= 0.6 * e1
    o.close()


def test_interface_velocity_of_the_internal_energy_flux():
    gam = 1.4
    # a contact moving at U: equal pressures and velocities, unequal densities
    for U in (0.3, -0.2, 0.0):
        S = [1.0, 1.0, U, 0.0, 0.0, 0.125, 1.0, U, 0.0, 0.0]   # rl,pl,ul,vl,wl,rr,pr,ur,vr,wr
        out = oracle.riemann_hllc_e(gam, S)
        assert out[2] == pytest.approx(U, abs=1e-14)       # the contact speed
        assert out[10] == pytest.approx(U, abs=1e-14)      # uprs
        assert out[9] == pytest.approx(1.0, abs=1e-14)     # the interface pressure
    # a symmetric collision: the interface is at rest
    S = [1.0, 1.0, 0.4, 0.0, 0.0, 1.0, 1.0, -0.4, 0.0, 0.0]
    out = oracle.riemann_hllc_e(gam, S)
    assert abs(out[2]) <= 1e-15 and abs(out[10]) <= 1e-15
    assert out[9] > 1.0


# ------------------------------------------------------------------ flatten
def _flat(ratio_f, jump=1e-3, ratio_b=0.5):
    """flatten of cell 2 with dp1f/dp2f = ratio_f, dp1b/dp2b = ratio_b and
    converging velocities; dp1f = jump."""
    p2 = 1.0
    dp1f = jump
    dp2f = dp1f / ratio_f                 # p[4] - p[1]
    p3 = p2 + dp1f
    dp1b = jump                           # p[2] - p[1]
    p1 = p2 - dp1b
    p4 = p1 + dp2f
    p0 = p3 - dp1b / ratio_b              # dp2b = p[3] - p[0]
    return oracle.flatten(1.0, np.ones(5), [p0, p1, p2, p3, p4], [0.0, 1.0, 0.5, 0.0, 0.0])


def test_flatten_is_colella_woodward_switch():
    jump = 1e-3
    # the dead band of CW84's omega_1 = 0.75 (and its mirror at 0.25)
    assert _flat(0.74, jump) == 0.0
    assert _flat(0.76, jump) > 0.0
    assert _flat(0.26, jump) == 0.0
    assert _flat(0.24, jump) > 0.0
    # slope omega_2 = 10 above the threshold, times the jump strength
    for ratio in (0.76, 0.8, 0.85, 0.9):
        f = _flat(ratio, jump)
        assert f == pytest.approx(10.0 * (ratio - 0.75) * jump / 1.0, rel=1e-9), ratio
    # the cap at 1/2
    assert _flat(0.9, 0.5) == 0.5
