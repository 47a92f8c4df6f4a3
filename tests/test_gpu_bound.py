"""pg_bound_x/y/z (the ghost cells the kernels read: analytic radial
ghosts, periodic azimuthal ones, mirrored polar ones) against the oracle's
boundx/boundy/boundz (boundary.cu:13-374, 462-718) on the same state: a
structured 2D gas+dust state and a 3D one, in FP64 storage.  The ghosts that
copy interior cells are exact copies of the handle's own state; against the
oracle every cell agrees to the rounding of the state's round trip through
the device layout (v stored as its deviation from the background rotation)
and of the same closed forms for the analytic radial ghosts."""
import numpy as np
import pytest

import oracle
from penguin_amd import Disk, PenguinError, config_c2, disk3d

pytestmark = pytest.mark.gpu


def _structured(g, amp=0.1):
    sh = g.shape[1:]
    idx = np.meshgrid(*(np.arange(n) for n in sh), indexing="ij")
    phase = sum((m + 1) * 2 * np.pi * (x + 0.5) / n for m, (x, n) in enumerate(zip(idx, sh)))
    g = g.copy()
    g[0] *= 1 + amp * np.cos(phase)
    g[2] += 1e-3 * np.sin(phase)
    if g.shape[0] == 5:
        g[4] += 1e-3 * np.cos(2 * phase)
    return g


def test_bound_2d_gas_and_dust_match_oracle():
    p = config_c2(precision=8, xres=40, yres=64)
    o = oracle.Oracle(p)
    g, d = o.get_state()
    g = _structured(g)
    g[4] = 0.0
    d = _structured(d)
    d[3] = 0.0
    g[1] = o.get_state()[0][1] / o.get_state()[0][0] * g[0]   # isothermal: p = c_s^2 rho
    o.set_state(g, d)
    with Disk(p) as h:
        h.set_state(g, d)
        own, own_d = h.get_state()
        for axis in (0, 1):
            a, ad = h.bound(axis, dust=True)
            b, bd = o.bound(axis, dust=True)
            assert a.shape == b.shape
            if axis == 0:
                assert np.array_equal(a[:, :, 3:-3], own) and np.array_equal(ad[:, :, 3:-3], own_d)
            else:
                assert np.array_equal(a, np.concatenate([own[:, -3:], own, own[:, :3]], axis=1))
                assert np.array_equal(ad, np.concatenate([own_d[:, -3:], own_d, own_d[:, :3]], axis=1))
            assert np.allclose(a, b, rtol=1e-13, atol=1e-18), (axis, np.abs(a - b).max())
            assert np.allclose(ad, bd, rtol=1e-13, atol=1e-18), (axis, np.abs(ad - bd).max())
        with pytest.raises(PenguinError):
            h.bound(2)


def test_bound_3d_match_oracle():
    p = disk3d(24, 32, 12, precision=8, noise_rho=1e-2)
    o = oracle.Oracle(p)
    g, _ = o.get_state()
    g = _structured(g)
    o.set_state(g)
    with Disk(p) as h:
        h.set_state(g)
        own, _ = h.get_state()
        mirror = own[:, ::-1].copy()
        mirror[4] *= -1.0
        for axis in (0, 1, 2):
            a, _ = h.bound(axis)
            b, _ = o.bound(axis)
            assert a.shape == b.shape
            if axis == 0:
                assert np.array_equal(a[..., 3:-3], own)
            elif axis == 1:
                assert np.array_equal(a, np.concatenate([own[:, :, -3:], own, own[:, :, :3]], axis=2))
            else:
                assert np.array_equal(a, np.concatenate([mirror[:, -3:], own, mirror[:, :3]], axis=1))
            assert np.allclose(a, b, rtol=1e-13, atol=1e-18), (axis, np.abs(a - b).max())
        with pytest.raises(PenguinError):
            h.bound(0, dust=True)
