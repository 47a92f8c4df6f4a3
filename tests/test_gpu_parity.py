"""Full-step parity of the GPU path against the FP64 oracle.

The headline gate: after 100 steps of the 256x512 disk and of the full
1024x3072 gas+dust disk, the GPU state matches the oracle within 1e-5
relative L1 on density and on each momentum component, radial and
azimuthal, gas and dust.  The lockstep tests give both sides the oracle's dt
sequence (the reference loop, main.cu:320-359, with dt from global_dt), so
only the step arithmetic differs; the full-size test runs the timed path
itself (pg_run, the captured step, dt chosen on the device) and gives the
oracle the GPU's step sizes.
"""
import numpy as np
import pytest

import oracle
from penguin_amd import Disk, config_c1, config_c2
from tests.helpers import parity_report

pytestmark = pytest.mark.gpu

GAS = ("rho", "rho_u", "rho_v", "momentum")
DUST = ("rho_d", "rho_d_u", "rho_d_v")


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
    print("small gas+dust 20 steps", prec, {k: f"{v:.2e}" for k, v in rep.items()}, flush=True)
    tol = 1e-11 if prec == 8 else 1e-5
    for k in GAS + DUST:
        assert rep[k] <= (tol if k not in ("rho_u", "rho_d_u") or prec == 4 else 1e-9), (k, rep[k])


@pytest.mark.parametrize("prec", [8, 4])
def test_c1_100_steps(prec):
    p = config_c1(precision=prec)
    rep = _lockstep(p, 100)
    print("C1 100 steps", prec, {k: f"{v:.2e}" for k, v in rep.items()}, flush=True)
    tol = 1e-10 if prec == 8 else 1e-5
    for k in ("rho", "rho_v", "momentum"):
        assert rep[k] <= tol, (k, rep[k])
    assert rep["rho_u"] <= (1e-8 if prec == 8 else 1e-5), rep["rho_u"]


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
    disk (planet, viscosity, orbital advection, wave killing, dust drag),
    run in FP32 through the benchmarked path (pg_run replaying its captured
    step, the CFL step chosen on the device), within 1e-5 relative L1 of the
    FP64 oracle given the same step sizes: density and each momentum
    component, gas and dust.  The FP64 GPU error is stated beside it (the
    host-driven pg_solve loop with the same step sizes), and the GPU's
    first step size is checked against the oracle's own."""
    p = config_c2(precision=4)
    d = Disk(p)
    dt0 = d.global_dt(p.end_time)
    dts = d.run(100, use_graph=True)
    g32, d32 = d.get_state()
    d.close()
    o = oracle.Oracle(p)
    assert abs(dt0 / o.global_dt(p.end_time) - 1.0) <= 1e-6
    t = 0.0
    for dt in dts:
        o.solve(t, float(dt))
        t += float(dt)
    g_ref, d_ref = o.get_state()
    del o
    d = Disk(config_c2(precision=8))
    t = 0.0
    for dt in dts:
        d.solve(t, float(dt))
        t += float(dt)
    g64, d64 = d.get_state()
    d.close()
    reps = {4: parity_report(g32, g_ref, d32, d_ref), 8: parity_report(g64, g_ref, d64, d_ref)}
    for prec, rep in reps.items():
        print(f"C2 1024x3072 100 steps, precision {prec}:", {k: f"{v:.2e}" for k, v in rep.items()}, flush=True)
    for k in GAS + DUST:
        assert reps[4][k] <= 1e-5, (k, reps[4][k])
    for k in ("rho", "rho_v", "momentum", "rho_d", "rho_d_v"):
        assert reps[8][k] <= 1e-8, (k, reps[8][k])
    assert reps[8]["rho_u"] <= 2e-5 and reps[8]["rho_d_u"] <= 5e-6
