"""Dot-product test of the GPU adjoint against the FP64 host tangent-linear
model (the oracle over forward-mode duals, oracle.TangentOracle, itself
checked against central differences in tests/test_oracle_tangent.py):

    <lambda, J d>  (tangent-linear, host, FP64)
  = <J^T lambda, d>  (adjoint, GPU)

over several steps with checkpoint segments, the step sizes the GPU forward
run recorded frozen in both.  FP64 adjoint within 1e-9, FP32 within 1e-3.
"""
import numpy as np
import pytest

import oracle
from penguin_amd import Disk, config_c5

pytestmark = pytest.mark.gpu


def _setup(prec):
    p = config_c5(precision=prec, xres=64, yres=96).replace(noise_rho=0.0, p_alpha=0.5)
    o = oracle.Oracle(p.replace(precision=8))
    g, _ = o.get_state()
    xa, _, ya, _ = o.coords()
    o.close()
    R = (0.5 * (xa[3:-4] + xa[4:-3]))[None, :]
    phi = (0.5 * (ya[3:-4] + ya[4:-3]))[:, None]
    cs2 = g[1] / g[0]
    cs = np.sqrt(cs2)
    g[0] = g[0] * (1.0 + 0.05 * np.cos(2 * phi - 4 * np.log(R)))
    g[1] = cs2 * g[0]
    g[2] = g[2] + 0.02 * cs * np.sin(phi + 2 * np.log(R))
    g[3] = g[3] + 0.02 * cs * np.cos(3 * phi - np.log(R))
    d = np.zeros_like(g)
    d[0] = g[0] * 0.3 * np.sin(phi + 3 * R)
    d[1] = cs2 * d[0]                      # p = cs2 rho: a consistent direction
    d[2] = 0.05 * cs * np.cos(2 * phi - R)
    d[3] = 0.05 * cs * np.sin(phi - 2 * R)
    lam = np.random.default_rng(9).uniform(-1.0, 1.0, (3,) + g.shape[1:])
    return p, g, d, lam


@pytest.mark.parametrize("prec, tol", [(8, 1e-9), (4, 1e-3)])
def test_adjoint_dot_product_against_tangent_linear(prec, tol):
    p, g, d, lam = _setup(prec)
    nsteps, ckpt = 7, 3
    dev = Disk(p)
    dev.set_state(g)
    dts = dev.adjoint_forward(nsteps, ckpt)
    dL = np.zeros((5,) + g.shape[1:])
    dL[0], dL[2], dL[3] = lam
    buf = np.full((5,) + g.shape[1:], np.nan)     # a caller's buffer: every field must be written
    _, _, d0 = dev.adjoint_backward(dL, out=buf)
    dev.close()
    assert d0 is buf and np.all(np.isfinite(buf)) and np.all(buf[1] == 0.0) and np.all(buf[4] == 0.0)
    adj = float(np.sum(d0[0] * d[0]) + np.sum(d0[2] * d[2]) + np.sum(d0[3] * d[3]))

    tl = oracle.TangentOracle(p.replace(precision=8))
    tl.set_state(g)
    tl.set_tangent(d)
    t = 0.0
    for h in dts:
        tl.solve(t, float(h))
        t += float(h)
    jd, _ = tl.get_tangent()
    tlm = float(np.sum(lam[0] * jd[0]) + np.sum(lam[1] * jd[2]) + np.sum(lam[2] * jd[3]))
    rel = abs(adj - tlm) / abs(tlm)
    print(f"prec {prec}: <J^T lam, d> {adj:.12e}  <lam, J d> {tlm:.12e}  rel {rel:.2e}", flush=True)
    assert rel <= tol
