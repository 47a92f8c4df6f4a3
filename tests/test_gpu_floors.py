"""The density floor of update_Cell (sweeps.cu:238-246) on the GPU.

A supersonic cell (8 c_s outward) next to a near-empty region, swept with
four times its CFL step, empties: the reference sets rho = max(1e-10 rho,
smallr), restores the old velocities and sets p = max(1e-10 p, smallp).
The GPU keeps the density and the velocities as the oracle does; its
pressure is not stored but derived, c_s^2 rho, so a floored cell reads
p = c_s^2 r_min (2.9e-13 here) where the reference stores smallp = 1e-10
until the cell's next update resets it to c_s^2 rho.  In FP32 the stored
deviation from the background resolves the density to ~6e-8 rho_b, so the
floor is held at ~5e-7 rho_b there (never at or below zero).
"""
import numpy as np
import pytest

import oracle
from penguin_amd import Disk, config_c1

pytestmark = pytest.mark.gpu


def _state(prec):
    p = config_c1(precision=prec, xres=48, yres=16).replace(noise_rho=0.0, visc_flag=0, planet_mass=0.0)
    o = oracle.Oracle(p.replace(precision=8))
    g, _ = o.get_state()
    cs = np.sqrt(g[1] / g[0])
    i0 = 20
    g[0][:, i0 - 3:i0] *= 1e-4
    g[1][:, i0 - 3:i0] *= 1e-4
    g[2][:, i0] = 8.0 * cs[:, i0]
    o.set_state(g)
    dt = 4.0 * o.global_dt(1e9)
    o.sweepx(dt)
    ref, _ = o.get_state()
    o.close()
    return p, g, dt, ref


def test_floor_fp64_matches_oracle():
    p, g, dt, ref = _state(8)
    d = Disk(p)
    d.set_state(g)
    d.sweep_x(dt)
    got, _ = d.get_state()
    d.close()
    rmin = np.maximum(1e-10 * g[0], 1e-10)
    floored = ref[0] == rmin
    print("floored cells per row:", floored.sum(axis=1), flush=True)
    assert floored.sum() == g.shape[1] and np.all(floored[:, 20])
    # density: the floor (the stored deviation rmin - rho_b carries rho_b's rounding)
    assert np.allclose(got[0][floored], rmin[floored], rtol=1e-5, atol=0)
    # velocities: the old ones, exactly
    assert np.array_equal(got[2][floored], g[2][floored])
    assert np.array_equal(got[3][floored], g[3][floored])
    # pressure: smallp in the reference, c_s^2 rho here
    cs2 = g[1][~floored] / g[0][~floored]
    assert np.all(ref[1][floored] == 1e-10)
    assert np.allclose(got[1][floored], (g[1] / g[0])[floored] * got[0][floored], rtol=1e-12)
    # every other cell as the oracle
    for k in (0, 2, 3):
        a, b = got[k][~floored], ref[k][~floored]
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max(), k
    assert cs2.size > 0


def test_floor_fp32_stays_positive_and_keeps_velocities():
    p, g, dt, ref = _state(4)
    d = Disk(p)
    d.set_state(g)
    d.sweep_x(dt)
    got, _ = d.get_state()
    d.close()
    rmin = np.maximum(1e-10 * g[0], 1e-10)
    floored = ref[0] == rmin
    assert np.all(np.isfinite(got))
    rb = g[0][:, 25:26]   # the background density around the floored column (within 5 %)
    assert np.all(got[0][floored] > 0.0) and np.all(got[0][floored] <= 1e-6 * rb.max())
    assert np.allclose(got[2][floored], g[2][floored], rtol=1e-6)
