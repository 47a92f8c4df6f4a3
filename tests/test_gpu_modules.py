"""Module-by-module parity of the HIP kernels against the FP64 oracle.

Each test puts the same state (the oracle's state after a few steps, so that
every field has structure) on the GPU and in the oracle, applies one
reference module on both, and compares.  FP64 device storage must agree with
the oracle to rounding; FP32 storage to single-precision rounding.
"""
import numpy as np
import pytest

import oracle
from penguin_amd import Disk, config_c1, config_c2
from tests.helpers import parity_report, rel_l1

pytestmark = pytest.mark.gpu

# relative-L1 bounds per storage precision for one module application
TOL = {8: 1e-12, 4: 2e-6}


def _pair(p, warm_steps=6):
    o = oracle.Oracle(p)
    o.run(warm_steps)
    gas, dust = o.get_state()
    d = Disk(p)
    d.set_state(gas, dust if p.dust_flag else None)
    return o, d


def _small(prec, dust=False):
    if dust:
        return config_c2(precision=prec, xres=96, yres=160)
    return config_c1(precision=prec, xres=96, yres=160)


def _cmp(o, d, tol, label):
    g_ref, d_ref = o.get_state()
    g, dd = d.get_state()
    rep = parity_report(g, g_ref, dd, d_ref if dd is not None else None)
    print(label, {k: f"{v:.2e}" for k, v in rep.items()})
    for k in ("rho", "rho_v", "momentum"):
        assert rep[k] <= tol, (label, k, rep[k])
    return rep


@pytest.mark.parametrize("prec", [8, 4])
def test_initial_state(prec):
    p = _small(prec, dust=True)
    o = oracle.Oracle(p)
    d = Disk(p)
    g_ref, d_ref = o.get_state()
    g, dd = d.get_state()
    tol = 1e-15 if prec == 8 else 1e-7
    for f in (0, 1, 2, 3):
        assert rel_l1(g[f], g_ref[f]) <= tol, f
    for f in (0, 1, 2):
        assert rel_l1(dd[f], d_ref[f]) <= tol, f


@pytest.mark.parametrize("prec", [8, 4])
def test_global_dt(prec):
    p = _small(prec, dust=True)
    o, d = _pair(p)
    dt_ref = o.global_dt(1.0)
    dt = d.global_dt(1.0)
    assert abs(dt / dt_ref - 1.0) <= (1e-13 if prec == 8 else 1e-6)
    # the 1.1x growth cap (timestep.cu:133-137)
    assert d.global_dt(0.5 * dt_ref) == pytest.approx(0.55 * dt_ref, rel=1e-15)


@pytest.mark.parametrize("prec", [8, 4])
def test_sweep_x_gas(prec):
    p = _small(prec)
    o, d = _pair(p)
    dt = o.global_dt(1.0)
    o.sweepx(dt)
    d.sweep_x(dt)
    _cmp(o, d, TOL[prec], "sweep_x")


@pytest.mark.parametrize("prec", [8, 4])
def test_sweep_y_gas(prec):
    p = _small(prec)
    o, d = _pair(p)
    dt = o.global_dt(1.0)
    o.sweepy(0.0, dt)
    d.sweep_y(dt)
    _cmp(o, d, TOL[prec], "sweep_y")


@pytest.mark.parametrize("prec", [8, 4])
def test_sweeps_dust(prec):
    p = _small(prec, dust=True)
    o, d = _pair(p)
    dt = o.global_dt(1.0)
    o.sweepx_dust(dt)
    d.sweep_x_dust(dt)
    rep = _cmp(o, d, TOL[prec], "sweep_x_dust")
    assert rep["rho_d"] <= TOL[prec] and rep["rho_d_v"] <= TOL[prec]
    o.sweepy_dust(0.0, dt)
    d.sweep_y_dust(dt)
    rep = _cmp(o, d, TOL[prec], "sweep_y_dust")
    assert rep["rho_d"] <= TOL[prec] and rep["rho_d_v"] <= TOL[prec]


@pytest.mark.parametrize("prec", [8, 4])
def test_orbital_advection(prec):
    p = _small(prec, dust=True)
    o, d = _pair(p)
    dt = o.global_dt(1.0)
    o.orbital_advection(dt)
    inc, res = d.orbital_advection(dt)
    inc_ref, res_ref = o.orbadv()
    assert np.array_equal(inc, inc_ref)          # integer shift: exact
    assert np.allclose(res, res_ref, rtol=1e-12 if prec == 8 else 1e-6, atol=1e-14)
    rep = _cmp(o, d, TOL[prec], "orbital_advection")
    assert rep["rho_d"] <= TOL[prec]


@pytest.mark.parametrize("prec", [8, 4])
def test_viscosity(prec):
    p = _small(prec)
    o, d = _pair(p)
    v_ref = o.viscosity()
    v = d.viscosity()
    for f in (0, 1):
        e = rel_l1(v[f], v_ref[f])
        print("viscosity", f, e)
        assert e <= (1e-10 if prec == 8 else 2e-3), (f, e)


@pytest.mark.parametrize("prec", [8, 4])
def test_sources_kick(prec):
    p = _small(prec, dust=True)
    o, d = _pair(p)
    t = 0.37
    dt = o.global_dt(1.0)
    o.viscosity()
    d.viscosity()
    o.evolve_planet(t)
    o.sources(0.0, 0.5 * dt, inplace=True)
    d.sources(t, 0.0, 0.5 * dt, inplace=True)
    rep = _cmp(o, d, TOL[prec], "kick")
    assert rep["rho_d_u"] <= (1e-12 if prec == 8 else 1e-5)


@pytest.mark.parametrize("prec", [8, 4])
def test_sources_predictor_corrector(prec):
    p = _small(prec, dust=True)
    o, d = _pair(p)
    t = 0.41
    dt = o.global_dt(1.0)
    o.viscosity()
    d.viscosity()
    o.evolve_planet(t)
    # predictor into the scratch buffer, viscosity of it, corrector in place
    o.sources(0.5 * dt, 0.5 * dt, inplace=False)
    d.sources(t, 0.5 * dt, 0.5 * dt, inplace=False)
    vT_ref = o.viscosity(on_T=True)
    vT = d.viscosity(on_T=True)
    for f in (0, 1):
        e = rel_l1(vT[f], vT_ref[f])
        print("viscosity(T)", f, e)
        assert e <= (1e-10 if prec == 8 else 2e-3)
    o.sources(0.5 * dt, 0.5 * dt, inplace=True)
    d.sources(t, 0.5 * dt, 0.5 * dt, inplace=True)
    rep = _cmp(o, d, TOL[prec], "corrector")
    assert rep["rho_d_u"] <= (1e-12 if prec == 8 else 1e-5)


@pytest.mark.parametrize("prec", [8, 4])
def test_killwave(prec):
    p = _small(prec, dust=True)
    o, d = _pair(p)
    dt = o.global_dt(1.0)
    o.killwave(dt)
    d.killwave(dt)
    _cmp(o, d, TOL[prec], "killwave")


@pytest.mark.parametrize("prec", [8, 4])
def test_planet(prec):
    p = _small(prec)
    d = Disk(p)
    for t in (0.0, 0.3, 2.9, 7.0):
        b = d.planet(t)
        b_ref = oracle.orbit_solution(p, t)
        assert np.allclose(b, b_ref, rtol=1e-14, atol=1e-15), (t, b, b_ref)
