"""Pins of the oracle's 3D spherical path (ndim 3: spherical radius, azimuth,
polar angle; PEM4; adiabatic gas with the internal-energy equation).  The
reference ships no expected outputs for its 3D disk, so these checks come
from what the mathematics fixes:

* PEM sub-cell means (PEM.cu:51-93): the means over the first fraction q and
  over the rest of the cell recombine to the cell mean for every q and both
  power-law branches; at slope ratio 2 the power law meets the parabola; the
  end values are the interface values and the cell mean; the power laws of
  get_PEM_0/1 (series branch included) against double-precision closed
  forms;
* adiabatic HLLC (HLLC.cu:10-77, internal_e_flag 1): F(U, U) is the physical
  flux (mass, internal energy e rho u, momenta) with p* = p and u* = u;
* the polar sweep (sweeps.cu:458-553) with reflecting walls conserves the
  mass of every (r, phi) column;
* the full step keeps a disk that is symmetric about the midplane symmetric
  (rho even, w odd in theta - pi/2): any shifted polar index, wrong ghost
  parity or sign of a polar force breaks it;
* the axisymmetric disk without planet stays axisymmetric;
* the tangent-linear build agrees with central differences in 3D (PPM4).
"""
import numpy as np
import pytest

import oracle
from penguin_amd import disk3d


def _p(**kw):
    base = dict(precision=8, noise_rho=0.0)
    base.update(kw)
    return disk3d(24, 32, 12, **base)


@pytest.mark.parametrize("par", [[0.1, 1.0, 0.5], [0.3, 1.0, 0.05], [-0.2, 1.0, -0.9], [0.1, 1.0, 0.15],
                                 [0.1, 1.0, 3.0], [-0.4, 2.0, -0.1]])
def test_pem_means_recombine_to_the_cell_mean(par):
    for q in (1e-3, 0.1, 0.37, 0.5, 0.8, 0.999):
        L, R = oracle.pem_aveL(q, par), oracle.pem_aveR(q, par)
        assert abs(q * L + (1 - q) * R - par[1]) < 1e-14
    # ends: the whole cell's mean, and the interface values a - sL, a + sR
    assert abs(oracle.pem_aveL(1.0, par) - par[1]) < 1e-14
    assert abs(oracle.pem_aveR(0.0, par) - par[1]) < 1e-14
    assert abs(oracle.pem_aveL(0.0, par) - (par[1] - par[0])) < 1e-14
    assert abs(oracle.pem_aveR(1.0, par) - (par[1] + par[2])) < 1e-14


@pytest.mark.parametrize("sL", [0.1, -0.3])
def test_pem_meets_ppm_at_slope_ratio_two(sL):
    # at ratio 2 the power law x^eta of get_PEM_0/1 reduces to the parabola;
    # the single-precision logarithm leaves up to ~1e-7 of the slope (its
    # rounding, amplified by the 1/(eta x) of get_PEM_1)
    for q in (0.0, 0.05, 0.2, 0.7, 1.0):
        for a, b in (([sL, 1.0, 2 * sL * (1 + 1e-12)], [sL, 1.0, 2 * sL]),
                     ([2 * sL * (1 + 1e-12), 1.0, sL], [2 * sL, 1.0, sL])):
            assert abs(oracle.pem_aveL(q, a) - oracle.ppm_aveL(q, b)) < 1e-6 * abs(sL)
            assert abs(oracle.pem_aveR(q, a) - oracle.ppm_aveR(q, b)) < 1e-6 * abs(sL)


def test_pem_power_law_against_double_precision():
    # get_PEM_1 (PEM.cu:17-48): v(x) = (1-x)((1-x)^eta - 1)/(eta x) + 1, by a
    # 4-term series where eta x <= 0.001 and in closed form (with the
    # single-precision logarithm of a float argument) above; aveL = a - sL + sL v
    sL, sR = 0.1, 0.1 / 5.0            # eta = sL / sR = 5 selects get_PEM_1
    eta = sL / sR
    exact = lambda x: (1 - x) * np.expm1(eta * np.log1p(-x)) / (eta * x) + 1
    for x in (1e-6, 1e-5, 1e-4, 1.9e-4):                         # series branch
        v = (oracle.pem_aveL(x, [sL, 1.0, sR]) - 1.0 + sL) / sL
        assert abs(v - exact(x)) < 1e-12, x
    for x in (0.01, 0.1, 0.3, 0.6):                               # closed form, float log
        v = (oracle.pem_aveL(x, [sL, 1.0, sR]) - 1.0 + sL) / sL
        assert abs(v - exact(x)) < 1e-5, x
    # get_PEM_0: the mean over the first fraction is a - sL + sL x^eta
    for x in (0.01, 0.2, 0.9):
        v = oracle.pem_aveL(x, [sR, 1.0, sL])                    # sR/sL = 5 in the argument order
        assert abs(v - (1.0 - sR + sR * x ** eta)) < 1e-6 * sR, x


def test_adiabatic_hllc_is_consistent():
    gam = 1.4
    for r, p, u, v, w in ((1.0, 0.7, 0.3, 0.1, -0.2), (0.5, 2.0, -0.4, 0.0, 0.3), (2.0, 0.1, 1.5, 0.2, 0.1)):
        S = np.array([r, p, u, v, w, r, p, u, v, w], dtype=float)
        out = np.zeros(10)
        oracle.lib().orc_riemann_hllc(gam, 2, oracle.ptr(S), oracle.ptr(out))
        pm, sl, sm, sr, Fr, Fp, Fu, Fv, Fw, pres = out
        e = p / ((gam - 1.0) * r)
        np.testing.assert_allclose([Fr, Fp, Fu, Fv, Fw], [r * u, e * r * u, u * r * u, v * r * u, w * r * u],
                                   rtol=1e-13, atol=1e-15)
        assert abs(pres - p) < 1e-13 * p and abs(sm - u) < 1e-12


def test_polar_sweep_conserves_column_mass():
    p = _p(noise_rho=1e-2, noise_vr=1e-2)
    o = oracle.Oracle(p)
    za, zv = o.zcoords()
    zv_i = zv[3:-3][:, None, None]
    g0, _ = o.get_state()
    dt = 0.5 * o.global_dt(1.0)
    o.sweepz(dt)
    g1, _ = o.get_state()
    m0, m1 = (g0[0] * zv_i).sum(axis=0), (g1[0] * zv_i).sum(axis=0)
    assert np.abs(g1[0] - g0[0]).max() > 1e-6 * g0[0].max()     # the sweep moved mass
    assert np.abs(m1 - m0).max() <= 1e-13 * np.abs(m0).max()


def test_step_keeps_midplane_symmetry():
    # one full step (viscosity, kicks, the three sweeps, orbital advection,
    # predictor-corrector with the planet); over more steps rounding-level
    # asymmetries flip discrete choices (the integer orbital shift of one of
    # two mirrored layers, limiter extrema), which then differ at the level
    # of the reconstruction error
    p = _p(ramp_time=0.0)
    o = oracle.Oracle(p)
    o.run(1)
    g, _ = o.get_state()
    scale = np.abs(g).max(axis=(1, 2, 3))
    for k in (0, 1, 2, 3):                        # rho, p, u, v even
        assert np.abs(g[k] - g[k][::-1]).max() <= 1e-10 * scale[k], k
    assert np.abs(g[4] + g[4][::-1]).max() <= 1e-10 * scale[3]          # w odd
    assert np.abs(g[4]).max() > 1e-8                                    # and not trivially zero


def test_axisymmetric_disk_stays_axisymmetric():
    p = _p(n_planet=0)
    o = oracle.Oracle(p)
    o.run(3)
    g, _ = o.get_state()
    for k in range(5):
        dev = np.abs(g[k] - g[k][:, :1, :]).max()
        assert dev <= 1e-12 * max(np.abs(g[k]).max(), 1e-30), k


def test_tangent_linear_matches_central_differences_3d():
    # the 3D geometry, polar sweep, spherical forces and per-layer orbital
    # advection of the dual-number build, on a state with structure in every
    # direction (so that 1e-7 perturbations cross no limiter decision); PPM4,
    # since PEM's closed form takes the logarithm of a float argument, a
    # staircase at 6e-8 relative spacing that central differences cannot
    # resolve (the adjoint, which this build checks, is two-dimensional)
    p = _p(ramp_time=0.0, recon_flag=6)
    o = oracle.Oracle(p)
    g, _ = o.get_state()
    K, J, I = np.meshgrid(*(np.arange(n) for n in g.shape[1:]), indexing="ij")
    phi = 2 * np.pi * (J + 0.5) / g.shape[2]
    lr = I / g.shape[3] * np.log(2.5 / 0.4)
    th = (K + 0.5) / g.shape[1]
    cs = np.sqrt(g[1] / g[0])
    g[0] *= 1 + 0.1 * np.cos(2 * phi - 6 * lr + 2 * th)
    g[1] *= (1 + 0.1 * np.cos(2 * phi - 6 * lr + 2 * th)) * (1 + 0.05 * np.sin(phi + 3 * th))
    g[2] += 0.05 * cs * np.sin(phi + 2 * lr - th)
    g[3] += 0.05 * cs * np.cos(3 * phi - lr + th)
    g[4] += 0.05 * cs * np.sin(2 * phi + lr) * np.sin(np.pi * th)
    o.set_state(g)
    dt = 0.5 * o.global_dt(1.0)
    o.close()
    d = np.zeros_like(g)
    d[0] = g[0] * 0.3 * np.sin(phi + 3 * lr) * np.cos(np.pi * th)
    d[1] = g[1] / g[0] * d[0]
    tl = oracle.TangentOracle(p)
    tl.set_state(g)
    tl.set_tangent(d)
    tl.solve(0.0, dt)
    jd, _ = tl.get_tangent()
    eps = 1e-7
    out = []
    for s in (1.0, -1.0):
        o = oracle.Oracle(p)
        o.set_state(g + s * eps * d)
        o.solve(0.0, dt)
        out.append(o.get_state()[0])
    fd = (out[0] - out[1]) / (2 * eps)
    # v ~ 1 carries a tangent ~1e-3 of itself: its central difference keeps
    # ~1e-16 / eps of rounding per cell
    for k, tol in ((0, 1e-5), (1, 1e-5), (2, 1e-5), (3, 2e-4), (4, 1e-5)):
        err = np.linalg.norm(jd[k] - fd[k]) / np.linalg.norm(fd[k])
        assert err < tol, (k, err)
