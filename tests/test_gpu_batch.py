"""Batched disks (pg_batch_*) on the GPU.

A batch advances disks of uneven sizes and parameters together, one launch
per pass over all of them, each disk with its own CFL step.  The gate: every
disk of a batch ends bitwise equal to the same disk run alone through
pg_run (same arithmetic, per-disk reductions), whatever the launch schedule,
with or without the captured graph, and after the autotuner has run; and a
sample of disks matches the FP64 oracle at the single-disk tolerance.
"""
import numpy as np
import pytest

import oracle
from penguin_amd import Batch, Disk, batch_configs, config_c2
from tests.helpers import parity_report

pytestmark = pytest.mark.gpu

# uneven grids, including radii that are not a multiple of any tile width
SIZES = [(128, 256), (64, 96), (96, 200), (256, 128), (40, 64), (128, 512)]


def _small_batch(n=6, prec=4, seed=7):
    ps = batch_configs(n, seed=seed, precision=prec)
    return [p.replace(xres=SIZES[k % len(SIZES)][0], yres=SIZES[k % len(SIZES)][1])
            for k, p in enumerate(ps)]


def _dust_batch(prec=4):
    return [config_c2(precision=prec, xres=nr, yres=nphi, planet_mass=q, seed=11 + k)
            for k, (nr, nphi, q) in enumerate([(96, 160, 3e-4), (64, 256, 1e-4), (128, 96, 1e-3)])]


def _alone(p, nsteps, use_graph=True):
    d = Disk(p)
    dts = d.run(nsteps, use_graph=use_graph)
    g, dd = d.get_state()
    t = d.t
    d.close()
    return g, dd, dts, t


def _assert_same_as_alone(ps, nsteps, b=None, use_graph=True):
    own = b is None
    if own:
        b = Batch(ps)
    dts = b.run(nsteps, use_graph=use_graph)
    for k, p in enumerate(ps):
        g1, d1, dts1, t1 = _alone(p, nsteps)
        g, dd = b.get_state(k)
        print(f"disk {k} {p.xres}x{p.yres}: dt {dts[k, -1]:.6e}", flush=True)
        assert np.array_equal(dts[k], dts1), k
        assert b.t[k] == t1, k
        assert np.array_equal(g, g1), (k, float(np.abs(g - g1).max()))
        if p.dust_flag:
            assert np.array_equal(dd, d1), k
    if own:
        b.close()


@pytest.mark.parametrize("prec", [4, 8])
def test_batch_equals_each_disk_alone(prec):
    _assert_same_as_alone(_small_batch(prec=prec), 12)


def test_batch_gas_dust_equals_each_disk_alone():
    _assert_same_as_alone(_dust_batch(), 10)


def test_batch_without_graph_equals_each_disk_alone():
    _assert_same_as_alone(_small_batch(n=4), 6, use_graph=False)


@pytest.mark.parametrize("sched", [(128, 256, 0, 16), (512, 512, 1, 4), (256, 1024, 2, 8), (384, 512, 0, 16),
                                   (122, 1024, 2, 16), (250, 512, 0, 8), (506, 256, 1, 4)])
def test_every_schedule_gives_the_same_bits(sched):
    ps = _small_batch()
    ref = Batch(ps)
    ref.run(8)
    b = Batch(ps)
    b.set_schedule(*sched)
    assert b.schedule() == sched
    b.run(8)
    for k in range(len(ps)):
        assert np.array_equal(b.get_state(k)[0], ref.get_state(k)[0]), (sched, k)
        assert b.t[k] == ref.t[k]


@pytest.mark.parametrize("tile", [122, 250, 506])
def test_wave_filling_x_tiles_gas_dust(tile):
    # x tiles whose 6 halo cells fill whole waves (one LDS slot per thread),
    # gas and dust sweeps, against the default schedule
    ps = _dust_batch()
    ref = Batch(ps)
    ref.run(6)
    b = Batch(ps)
    tx, rb, order, rows = b.schedule()
    b.set_schedule(tile, rb, order, rows)
    b.run(6)
    for k in range(len(ps)):
        g0, d0 = ref.get_state(k)
        g1, d1 = b.get_state(k)
        assert np.array_equal(g1, g0) and np.array_equal(d1, d0), (tile, k)


def test_autotune_restores_state_and_keeps_bits():
    ps = _small_batch()
    a = Batch(ps)
    a.run(3)
    ms = a.autotune(steps=2)
    assert ms > 0.0
    tile, ring, order, rows = a.schedule()
    print("autotuned schedule", tile, ring, order, rows, f"{ms:.3f} ms/step", flush=True)
    a.run(7)
    b = Batch(ps)
    b.run(3)
    b.run(7)
    for k in range(len(ps)):
        assert np.array_equal(a.get_state(k)[0], b.get_state(k)[0]), k
        assert a.t[k] == b.t[k] and a.dt[k] == b.dt[k]


def test_batch_sample_matches_oracle():
    ps = _small_batch()
    b = Batch(ps)
    b.run(20)
    for k in (0, 2, 5):
        o = oracle.Oracle(ps[k])
        dts_o = o.run(20)
        g_ref, _ = o.get_state()
        rep = parity_report(b.get_state(k)[0], g_ref)
        rep["t"] = abs(b.t[k] / o.t - 1.0)
        print(f"disk {k} vs oracle", {n: f"{v:.2e}" for n, v in rep.items()}, flush=True)
        for n in ("rho", "rho_v", "momentum", "t"):
            assert rep[n] <= 1e-5, (k, n, rep[n])
        assert dts_o[-1] > 0.0


def test_batch_profile_and_introspection():
    ps = _small_batch(n=3)
    b = Batch(ps)
    assert b.device_bytes() > 0
    b.profile(True)
    b.run(3, use_graph=False)
    for name in ("sweep_x_gas", "ring_gas", "corrector_viscosity", "prologue"):
        us, cnt = b.kernel_time_us(name)
        assert cnt == 3 and us > 0.0, name
