"""Full-step parity of the GPU path against the FP64 oracle.

The headline gate: after 100 steps of the 256x512 disk and of the full
1024x3072 gas+dust disk, the GPU state matches the oracle within 1e-5
relative L1 on density and momentum.  Both sides take the oracle's dt sequence (the reference loop,
main.cu:320-359, with dt from global_dt), so only the step arithmetic
differs; the GPU's own dt is compared separately.
"""
import numpy as np
import pytest

import oracle
from penguin_amd import Disk, config_c1, config_c2
from tests.helpers import parity_report

pytestmark = pytest.mark.gpu


def _lockstep(p, nsteps):
    o = oracle.Oracle(p)
    d = Disk(p)
    t, dt_prev = 0.0, p.end_time
    worst_dt = 0.0
    for n in range(nsteps):
        dt = o.global_dt(dt_prev)
        dt_gpu = d.global_dt(dt_prev)
        worst_dt = max(worst_dt, abs(dt_gpu / dt - 1.0))
        o.solve(t, dt)
        d.solve(t, dt)
        t += dt
        dt_prev = dt
    g_ref, d_ref = o.get_state()
    g, dd = d.get_state()
    rep = parity_report(g, g_ref, dd, d_ref if p.dust_flag else None)
    rep["dt"] = worst_dt
    return rep


@pytest.mark.parametrize("prec", [8, 4])
def test_small_20_steps(prec):
    p = config_c2(precision=prec, xres=96, yres=160)
    rep = _lockstep(p, 20)
    print("small gas+dust 20 steps", prec, {k: f"{v:.2e}" for k, v in rep.items()})
    tol = 1e-11 if prec == 8 else 1e-5
    for k in ("rho", "rho_v", "momentum", "rho_d", "rho_d_v"):
        assert rep[k] <= tol, (k, rep[k])


@pytest.mark.parametrize("prec", [8, 4])
def test_c1_100_steps(prec):
    p = config_c1(precision=prec)
    rep = _lockstep(p, 100)
    print("C1 100 steps", prec, {k: f"{v:.2e}" for k, v in rep.items()})
    tol = 1e-10 if prec == 8 else 1e-5
    for k in ("rho", "rho_v", "momentum"):
        assert rep[k] <= tol, (k, rep[k])
    if prec == 8:
        assert rep["rho_u"] <= 1e-8


@pytest.mark.parametrize("prec", [8, 4])
def test_device_loop_matches_host_loop(prec):
    """pg_run (dt on the device, optionally replayed from a hipGraph) gives
    the same trajectory as the host-driven loop."""
    p = config_c2(precision=prec, xres=96, yres=160)
    a = Disk(p)
    dts_a = [a.step_host() for _ in range(12)]
    b = Disk(p)
    dts_b = b.run(12, use_graph=False)
    c = Disk(p)
    dts_c = np.concatenate([c.run(5, use_graph=True), c.run(7, use_graph=True)])
    assert np.allclose(dts_b, dts_a, rtol=1e-12, atol=0)
    assert np.array_equal(dts_b, dts_c)
    ga, _ = a.get_state()
    gb, _ = b.get_state()
    gc, _ = c.get_state()
    assert np.array_equal(gb, gc)
    assert np.allclose(ga, gb, rtol=1e-12 if prec == 8 else 1e-6, atol=1e-14)
    assert abs(b.t - sum(dts_b)) <= 1e-12 * b.t


def test_c2_full_size_100_steps():
    """The headline gate at full size: 100 steps of the 1024x3072 gas+dust
    disk (planet, viscosity, orbital advection, wave killing, dust drag) in
    FP32 within 1e-5 relative L1 of the FP64 oracle on density and momenta,
    gas and dust, with the FP64 GPU error stated beside it.  Both GPU runs
    take the oracle's dt sequence; the GPU's own first dt is checked against
    the oracle's."""
    p = config_c2(precision=4)
    o = oracle.Oracle(p)
    t, dt_prev, dts = 0.0, p.end_time, []
    for n in range(100):
        dt = o.global_dt(dt_prev)
        o.solve(t, dt)
        dts.append(dt)
        t += dt
        dt_prev = dt
    g_ref, d_ref = o.get_state()
    del o
    reps = {}
    for prec in (4, 8):
        d = Disk(config_c2(precision=prec))
        assert abs(d.global_dt(p.end_time) / dts[0] - 1.0) <= (1e-6 if prec == 4 else 1e-12)
        t = 0.0
        for dt in dts:
            d.solve(t, dt)
            t += dt
        g, dd = d.get_state()
        d.close()
        reps[prec] = parity_report(g, g_ref, dd, d_ref)
        print(f"C2 1024x3072 100 steps, precision {prec}:", {k: f"{v:.2e}" for k, v in reps[prec].items()},
              flush=True)
    # the gate: density and the momentum vector, gas and dust
    for k in ("rho", "rho_v", "momentum", "rho_d", "rho_d_v"):
        assert reps[4][k] <= 1e-5, (k, reps[4][k])
        assert reps[8][k] <= 1e-8, (k, reps[8][k])
    # the radial momenta alone are bounded, not gated: rho*u (viscous drift,
    # the seeded noise, the planet's wake) is orders of magnitude below rho*v,
    # so the same absolute rounding is far larger relative to it.  Measured:
    # FP32 5e-4, FP64 2e-6, while the momentum vector's relative L1 exceeds
    # rho*v's by a twentieth: the radial absolute error is ~1/20 of the
    # azimuthal one.
    assert reps[4]["rho_u"] <= 2e-3 and reps[4]["rho_d_u"] <= 5e-4
    assert reps[8]["rho_u"] <= 2e-5 and reps[8]["rho_d_u"] <= 5e-6
