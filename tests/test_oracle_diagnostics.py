"""The oracle's diagnostics: the planet torque (global_sum + cal_torque +
output_gy, output.cu:336-457, output/forces.cu:41-58, 106-127) and the time
average of the gas (init_average / averaging / add_grid, output.cu:250-329,
applied around each step as main.cu:344-350 does).

Pins, independent of the oracle's arithmetic:
* the torque of one cell is -dPhi/dphi rho dV, with Phi the softened planet
  potential plus the indirect term, differentiated numerically in phi
  (R gy = -dPhi/dphi); this fixes the softening, the indirect term and the
  sign;
* an axisymmetric disk feels no torque from a planet on the grid's mirror
  axis (phi_p = pi: the cells pair up at phi and 2 pi - phi);
* the average over k steps equals the trapezoid sum over the states the
  oracle passes through, each step weighted dt/2 / interval at both ends,
  recorded here one step at a time.
"""
import numpy as np

import oracle
from penguin_amd import disk2d


def _params(**kw):
    base = dict(ramp_time=0.0, noise_rho=0.0, planet_mass=1e-3)
    base.update(kw)
    return disk2d(24, 48, precision=8, **base)


def _centres(o):
    xa, _, ya, _ = o.coords()
    R = 0.5 * (xa[3:-4] + xa[4:-3])
    phi = 0.5 * (ya[3:-4] + ya[4:-3])
    dv = 0.5 * (xa[4:-3] ** 2 - xa[3:-4] ** 2)[None, :] * (ya[4:-3] - ya[3:-4])[:, None]
    return R, phi, dv


def _potential(R, phi, b):
    """Softened point mass plus the indirect term of the star's reflex
    (the potential whose -grad gives star_planet_grav_cyl)."""
    m, x, y, rs = b[0], b[3], b[4], b[9]
    d2 = R * R + x * x - 2.0 * x * R * np.cos(phi - y) + rs * rs
    return -m / np.sqrt(d2) + m * R * np.cos(phi - y) / (x * x)


def test_torque_of_one_cell_is_minus_dphi():
    p = _params()
    o = oracle.Oracle(p)
    R, phi, dv = _centres(o)
    b = o.planet()
    gas, _ = o.get_state()
    for (i, j) in [(5, 30), (12, 20), (20, 26), (9, 2)]:
        g = gas.copy()
        g[0][:] = 0.0
        g[0][j, i] = 2.5
        o.set_state(g)
        tq = o.torque()
        h = 1e-5
        dphi = (_potential(R[i], phi[j] + h, b) - _potential(R[i], phi[j] - h, b)) / (2 * h)
        want = -dphi * 2.5 * dv[j, i]
        assert abs(tq - want) <= 1e-7 * abs(want), (i, j, tq, want)


def test_axisymmetric_disk_feels_no_torque():
    p = _params()
    o = oracle.Oracle(p)
    b = o.planet()
    assert abs(b[4] - np.pi) < 1e-15          # circular orbit: the planet sits at phi = pi
    gas, _ = o.get_state()                    # noise off: Sigma depends on R only
    R, phi, dv = _centres(o)
    scale = np.sum(np.abs(gas[0] * dv)) * p.planet_mass / (R[0] ** 2)
    assert abs(o.torque()) <= 1e-12 * scale


def test_torque_is_linear_in_density():
    p = _params(noise_rho=1e-2)
    o = oracle.Oracle(p)
    gas, _ = o.get_state()
    t1 = o.torque()
    g = gas.copy()
    g[0] *= 3.0
    o.set_state(g)
    assert abs(o.torque() - 3.0 * t1) <= 1e-12 * abs(3.0 * t1) + 1e-300


def test_average_is_the_trapezoid_over_the_steps():
    p = _params(noise_rho=1e-3, planet_mass=3e-4)
    interval = 0.05
    a = oracle.Oracle(p)
    a.average_begin(interval)
    a.run(6)
    ave = a.average()
    # the same steps one at a time, states recorded
    b = oracle.Oracle(p)
    states = [b.get_state()[0]]
    dts = []
    for _ in range(6):
        dts.append(b.run(1)[0])
        states.append(b.get_state()[0])
    want = np.zeros_like(ave)
    for n, dt in enumerate(dts):
        want += 0.5 * dt / interval * (states[n] + states[n + 1])
    assert np.allclose(ave, want, rtol=1e-13, atol=1e-16 * np.abs(want).max())
    # the weights sum to the elapsed time over the interval (Sigma changes by
    # a fraction of a percent over these steps)
    assert abs(ave[0].sum() / states[0][0].sum() / (sum(dts) / interval) - 1.0) < 1e-2
    # stopping keeps the average
    a.average_end()
    a.run(2)
    assert np.array_equal(a.average(), ave)
