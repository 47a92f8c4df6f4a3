"""Shared helpers of the test suite."""
from __future__ import annotations

import numpy as np


def rel_l1(a, b) -> float:
    """sum |a - b| / sum |b| (b is the reference)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / (den if den > 0 else 1.0))


def conserved(gas):
    """rho, rho*u (radial), rho*v (azimuthal) of a (5, ny, nx) gas state."""
    return gas[0], gas[0] * gas[2], gas[0] * gas[3]


def momentum_l1(a, b) -> float:
    """Relative L1 of the momentum vector field rho*(u, v)."""
    da = np.abs(a[0] * a[2] - b[0] * b[2]) + np.abs(a[0] * a[3] - b[0] * b[3])
    den = np.abs(b[0] * b[2]).sum() + np.abs(b[0] * b[3]).sum()
    return float(da.sum() / den)


def parity_report(gpu_gas, ref_gas, gpu_dust=None, ref_dust=None) -> dict:
    out = {
        "rho": rel_l1(gpu_gas[0], ref_gas[0]),
        "rho_u": rel_l1(gpu_gas[0] * gpu_gas[2], ref_gas[0] * ref_gas[2]),
        "rho_v": rel_l1(gpu_gas[0] * gpu_gas[3], ref_gas[0] * ref_gas[3]),
        "momentum": momentum_l1(gpu_gas, ref_gas),
    }
    if gpu_dust is not None:
        out["rho_d"] = rel_l1(gpu_dust[0], ref_dust[0])
        out["rho_d_u"] = rel_l1(gpu_dust[0] * gpu_dust[1], ref_dust[0] * ref_dust[1])
        out["rho_d_v"] = rel_l1(gpu_dust[0] * gpu_dust[2], ref_dust[0] * ref_dust[2])
    return out
