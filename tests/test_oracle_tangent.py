"""The oracle's tangent-linear build (liboracle_tlm.so: the same source over
forward-mode dual numbers) against central differences of the FP64 oracle
with the same frozen step sizes.  This pins the tangent-linear model that
the GPU adjoint is dot-product tested against (tests/test_gpu_adjoint_tlm.py)."""
import numpy as np
import pytest

import oracle
from penguin_amd import config_c5


def _setup(kw):
    p = config_c5(precision=8, xres=32, yres=48).replace(noise_rho=0.0, p_alpha=0.5, **kw)
    o = oracle.Oracle(p)
    g, _ = o.get_state()
    xa, _, ya, _ = o.coords()
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
    d[1] = cs2 * d[0]
    d[2] = 0.05 * cs * np.cos(2 * phi - R)
    d[3] = 0.05 * cs * np.sin(phi - 2 * R)
    o.set_state(g)
    dt = 0.5 * o.global_dt(1.0)
    o.close()
    return p, g, d, dt


@pytest.mark.parametrize("case", ["sweeps", "full"])
def test_tangent_matches_central_differences(case):
    kw = {"sweeps": dict(n_planet=0, visc_flag=0), "full": dict(ramp_time=0.0)}[case]
    p, g, d, dt = _setup(kw)
    # two steps of unequal size; over more steps some cell's limiter comes
    # within 1e-6 of a switch, where a central difference straddles the kink
    dts = [dt, 1.05 * dt]
    tl = oracle.TangentOracle(p)
    tl.set_state(g)
    tl.set_tangent(d)
    t = 0.0
    for h in dts:
        tl.solve(t, h)
        t += h
    jd, _ = tl.get_tangent()
    base, _ = tl.get_state()
    eps = 1e-6
    out = []
    for s in (1.0, -1.0):
        o = oracle.Oracle(p)
        o.set_state(g + s * eps * d)
        t = 0.0
        for h in dts:
            o.solve(t, h)
            t += h
        out.append(o.get_state()[0])
        if s == 1.0:
            # the dual build's values are the FP64 oracle's, bit for bit
            o0 = oracle.Oracle(p)
            o0.set_state(g)
            t = 0.0
            for h in dts:
                o0.solve(t, h)
                t += h
            assert np.array_equal(o0.get_state()[0], base)
    fd = (out[0] - out[1]) / (2 * eps)
    for k, name in ((0, "rho"), (2, "u"), (3, "v")):
        err = np.linalg.norm(jd[k] - fd[k]) / np.linalg.norm(fd[k])
        print(f"{case} {name}: tangent vs FD rel L2 {err:.2e}")
        assert err < 1e-6, (name, err)
