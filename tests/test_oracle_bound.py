"""Pins for the oracle's ghost fills (boundx/boundy/boundz, boundary.cu:13-374,
462-718), the reference of pg_bound_x/y/z: the azimuthal ghosts are the
periodic images of the interior rows, the polar ghosts the mirror images
with w negated, and the radial ghosts (boundary type 0) the analytic disk
at the ghost radii: Sigma = (R / R_p)^-p_alpha, p = c_s^2(R) Sigma with
c_s^2 = h_p^2 / R_p (R / R_p)^-p_beta, the interior's power laws continued."""
import numpy as np

import oracle
from penguin_amd import config_c1, disk3d


def _pair2d():
    p = config_c1(precision=8, xres=32, yres=48)
    o = oracle.Oracle(p)
    return p, o


def test_azimuthal_ghosts_are_periodic_images():
    p, o = _pair2d()
    g0, d0 = o.get_state()
    g, d = o.bound(1, dust=True)
    assert g.shape == (5, 54, 32)
    assert np.array_equal(g[:, 3:-3], g0)
    assert np.array_equal(g[:, :3], g0[:, -3:]) and np.array_equal(g[:, -3:], g0[:, :3])


def test_radial_ghosts_continue_the_analytic_disk():
    p, o = _pair2d()
    g0, _ = o.get_state()
    g, _ = o.bound(0)
    assert np.array_equal(g[:, :, 3:-3], g0)
    xa, _, _, _ = o.coords()
    xc = 0.5 * (xa[1:] + xa[:-1])          # centres of the padded cells
    ghosts = np.r_[0:3, 35:38]
    sigma = (xc[ghosts] / p.planet_radius) ** (-p.p_alpha)
    assert np.allclose(g[0][:, ghosts], sigma[None, :], rtol=1e-14, atol=0)
    cs2 = p.sc_h ** 2 / p.planet_radius * (xc[ghosts] / p.planet_radius) ** (-p.p_beta)
    assert np.allclose(g[1][:, ghosts] / g[0][:, ghosts], cs2[None, :], rtol=1e-13)
    # every ghost row is the same analytic state (no noise outside the disk)
    assert np.array_equal(g[:, :, ghosts], np.broadcast_to(g[:, :1, ghosts], g[:, :, ghosts].shape))


def test_polar_ghosts_mirror_with_w_negated():
    o = oracle.Oracle(disk3d(16, 32, 8, precision=8, noise_rho=1e-2))
    g0, _ = o.get_state()
    g, _ = o.bound(2)
    assert g.shape == (5, 14, 32, 16)
    assert np.array_equal(g[:, 3:-3], g0)
    assert np.array_equal(g[:4, :3], g0[:4, 2::-1]) and np.array_equal(g[:4, -3:], g0[:4, :-4:-1])
    assert np.array_equal(g[4, :3], -g0[4, 2::-1]) and np.array_equal(g[4, -3:], -g0[4, :-4:-1])
