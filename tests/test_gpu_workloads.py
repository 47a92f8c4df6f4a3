"""The batched and adjoint workloads at their own sizes.

C4: 64 uneven gas disks (nr in {128..512}, nphi in {256..1536}, seed 1)
advanced together with the autotuned schedule, each with its own CFL step:
a sample of eight disks is bitwise equal to each disk run alone, and two
disks match the FP64 oracle (given the GPU's step sizes) within 1e-5.

C5: the 512x1024 gas disk over 200 steps with a checkpoint every 20 steps
and the gap-profile loss.  The FP64 adjoint passes the dot-product test
against the FP64 host tangent-linear model (the oracle over forward-mode
duals) with the forward run's step sizes frozen in both; the FP32 adjoint's
dL/dq and dL/dalpha match central differences of FP64 forward runs with the
same step sizes within 1e-3.
"""
import numpy as np
import pytest

import oracle
from penguin_amd import Batch, Disk, batch_configs, config_c5
from penguin_amd.adjoint import gap_profile
from tests.helpers import parity_report

pytestmark = pytest.mark.gpu

C5_STEPS, C5_CKPT = 200, 20


def test_c4_batch_of_64_disks():
    ps = batch_configs(64, seed=1, precision=4)
    b = Batch(ps)
    ms = b.autotune(steps=3)
    print("autotuned schedule", b.schedule(), f"{ms:.3f} ms/step", flush=True)
    nsteps = 20
    dts = b.run(nsteps)
    assert np.all(dts > 0.0)
    sample = [0, 9, 18, 27, 36, 45, 54, 63]
    for k in sample:
        d = Disk(ps[k])
        dk = d.run(nsteps, use_graph=True)
        g1, _ = d.get_state()
        d.close()
        g, _ = b.get_state(k)
        print(f"disk {k} {ps[k].xres}x{ps[k].yres}: bitwise {np.array_equal(g, g1)}", flush=True)
        assert np.array_equal(dts[k], dk), k
        assert np.array_equal(g, g1), k
    # two disks against the oracle with the batch's own step sizes
    small = sorted(range(len(ps)), key=lambda k: int(ps[k].xres) * int(ps[k].yres))[:2]
    for k in small:
        o = oracle.Oracle(ps[k])
        t = 0.0
        for h in dts[k]:
            o.solve(t, float(h))
            t += float(h)
        g_ref, _ = o.get_state()
        o.close()
        g, _ = b.get_state(k)
        rep = parity_report(g, g_ref)
        print(f"disk {k} {ps[k].xres}x{ps[k].yres} vs oracle:", {x: f"{v:.2e}" for x, v in rep.items()}, flush=True)
        for x in ("rho", "rho_u", "rho_v", "momentum"):
            assert rep[x] <= 1e-5, (k, x, rep[x])
    b.close()


def _gap_loss_grad(p, g):
    """L = sum_i (<Sigma>_phi(r_i) - Sigma_t(r_i))^2 and dL/dSigma."""
    target = gap_profile(p)
    prof = g[0].mean(axis=0)
    L = float(np.sum((prof - target) ** 2))
    dL = np.zeros_like(g)
    dL[0] = np.broadcast_to(2.0 * (prof - target) / g.shape[1], g[0].shape)
    return L, dL


def _frozen_loss(p, dts):
    d = Disk(p)
    t = 0.0
    for h in dts:
        d.solve(t, float(h))
        t += float(h)
    g, _ = d.get_state()
    d.close()
    return _gap_loss_grad(p, g)[0]


def test_c5_fp32_adjoint_against_finite_differences():
    p = config_c5(precision=4)
    d = Disk(p)
    dts = d.adjoint_forward(C5_STEPS, C5_CKPT)
    g, _ = d.get_state()
    L, dL = _gap_loss_grad(p, g)
    dq, dal, _ = d.adjoint_backward(dL)
    d.close()
    p64 = p.replace(precision=8)
    assert abs(_frozen_loss(p64, dts) / L - 1.0) <= 1e-5
    # central differences of FP64 runs with the FP32 run's step sizes; the
    # alpha step is larger than q's because the loss moves ~1e-6 of itself
    # per unit relative change of alpha (FP64 rounding of the loss would
    # swamp a 1e-6 step)
    for name, val, grad, rel in (("planet_mass", p.planet_mass, dq, 1e-6), ("ss_alpha", p.ss_alpha, dal, 1e-4)):
        h = rel * val
        fd = (_frozen_loss(p64.replace(**{name: val + h}), dts) -
              _frozen_loss(p64.replace(**{name: val - h}), dts)) / (2.0 * h)
        print(f"C5 dL/d{name}: adjoint (FP32) {grad:.9e}  central FD (FP64) {fd:.9e}  "
              f"rel {abs(grad - fd) / abs(fd):.2e}", flush=True)
        assert abs(grad - fd) <= 1e-3 * abs(fd), (name, grad, fd)


def test_c5_fp64_adjoint_dot_product_against_tangent_linear():
    p = config_c5(precision=8)
    d = Disk(p)
    g0, _ = d.get_state()
    dts = d.adjoint_forward(C5_STEPS, C5_CKPT)
    g, _ = d.get_state()
    _, dL = _gap_loss_grad(p, g)
    rng = np.random.default_rng(21)
    cs = np.sqrt(g0[1] / g0[0])
    dL[2] = rng.uniform(-1.0, 1.0, g0[0].shape) * 1e-3
    dL[3] = rng.uniform(-1.0, 1.0, g0[0].shape) * 1e-3
    _, _, d0 = d.adjoint_backward(dL)
    d.close()
    # a smooth direction in the start state (rho, u, v; p = cs2 rho)
    ny, nx = g0[0].shape
    R = np.linspace(0.0, 1.0, nx)[None, :]
    phi = np.linspace(0.0, 2.0 * np.pi, ny, endpoint=False)[:, None]
    dv = np.zeros_like(g0)
    dv[0] = g0[0] * 1e-3 * np.sin(2.0 * phi + 5.0 * R)
    dv[1] = (g0[1] / g0[0]) * dv[0]
    dv[2] = 1e-3 * cs * np.cos(phi - 3.0 * R)
    dv[3] = 1e-3 * cs * np.sin(3.0 * phi + R)
    adj = float(sum(np.sum(d0[k] * dv[k]) for k in (0, 2, 3)))
    tl = oracle.TangentOracle(p)
    tl.set_state(g0)
    tl.set_tangent(dv)
    t = 0.0
    for h in dts:
        tl.solve(t, float(h))
        t += float(h)
    jd, _ = tl.get_tangent()
    tl.close()
    tlm = float(sum(np.sum(dL[k] * jd[k]) for k in (0, 2, 3)))
    rel = abs(adj - tlm) / abs(tlm)
    print(f"C5 dot product, FP64: <J^T lam, d> {adj:.12e}  <lam, J d> {tlm:.12e}  rel {rel:.2e}", flush=True)
    assert rel <= 1e-6
