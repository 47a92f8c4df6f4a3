"""Run-time disk configuration (mirror of include/pg_params.h) and the
standard workloads.

The reference fixes every parameter at compile time in include/parameters.h
(reference include/parameters.h:4-207).  ``Params`` carries the same names as
run-time fields so one library build serves many disks (the batched path
holds one ``Params`` per disk).
"""
from __future__ import annotations

import ctypes
import math

TWOPI = 2.0 * 3.1415926535897932384626434
HPI = 0.5 * 3.1415926535897932384626434


class Params(ctypes.Structure):
    """ctypes mirror of ``pg_params`` (include/pg_params.h), field for field."""

    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("xres", ctypes.c_int32), ("yres", ctypes.c_int32), ("zres", ctypes.c_int32),
        ("geomx", ctypes.c_int32), ("geomy", ctypes.c_int32), ("geomz", ctypes.c_int32),
        ("gridx", ctypes.c_int32), ("gridy", ctypes.c_int32), ("gridz", ctypes.c_int32),
        ("recon_flag", ctypes.c_int32),
        ("EOS_flag", ctypes.c_int32),
        ("visc_flag", ctypes.c_int32),
        ("OrbAdv_flag", ctypes.c_int32),
        ("dust_flag", ctypes.c_int32),
        ("kill_flag", ctypes.c_int32),
        ("n_planet", ctypes.c_int32),
        ("bound_lft", ctypes.c_int32), ("bound_rgh", ctypes.c_int32),
        ("bound_bak", ctypes.c_int32), ("bound_frn", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("bound_bom", ctypes.c_int32), ("bound_top", ctypes.c_int32),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("gam", ctypes.c_double),
        ("CFL", ctypes.c_double),
        ("frame_omega", ctypes.c_double),
        ("sc_h", ctypes.c_double),
        ("p_alpha", ctypes.c_double),
        ("p_beta", ctypes.c_double),
        ("ss_alpha", ctypes.c_double),
        ("kill_width", ctypes.c_double),
        ("D_G_ratio", ctypes.c_double),
        ("Stokes", ctypes.c_double),
        ("planet_mass", ctypes.c_double),
        ("planet_radius", ctypes.c_double),
        ("planet_ecc", ctypes.c_double),
        ("rs_fac", ctypes.c_double),
        ("ramp_time", ctypes.c_double),
        ("sav_interval", ctypes.c_double),
        ("end_time", ctypes.c_double),
        ("noise_rho", ctypes.c_double),
        ("noise_vr", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]

    def copy(self) -> "Params":
        p = Params()
        ctypes.pointer(p)[0] = self
        return p

    def replace(self, **kw) -> "Params":
        p = self.copy()
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return p

    @property
    def ncells(self) -> int:
        return int(self.xres) * int(self.yres) * int(self.zres)

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def disk2d(nr: int, nphi: int, *, precision: int = 4, **kw) -> Params:
    """2D polar locally-isothermal disk: r in [0.4, 2.5] log-spaced, phi in
    [0, 2pi) uniform, Sigma ~ r^-1, h/r = 0.05 without flaring (p_beta = 1),
    PPM4 + HLLC, orbital advection and alpha viscosity on, one planet at r=1
    on a circular orbit with Plummer smoothing 0.6 h, CFL 0.3, frame rotating
    at Omega_f = 1."""
    p = Params()
    p.ndim = 2
    p.xres, p.yres, p.zres = nr, nphi, 1
    p.geomx, p.geomy, p.geomz = 1, 3, 0
    p.gridx, p.gridy, p.gridz = 1, 0, 0
    p.recon_flag = 6
    p.EOS_flag = 0
    p.visc_flag = 1
    p.OrbAdv_flag = 1
    p.dust_flag = 0
    p.kill_flag = 0
    p.n_planet = 1
    p.bound_lft, p.bound_rgh, p.bound_bak, p.bound_frn = 0, 0, 3, 3
    p.precision = precision
    p.xmin, p.xmax = 0.4, 2.5
    p.ymin, p.ymax = 0.0, TWOPI
    p.zmin, p.zmax = HPI - 4.0 * 0.05, HPI
    p.gam = 1.0
    p.CFL = 0.3
    p.frame_omega = 1.0
    p.sc_h = 0.05
    p.p_alpha = 1.0
    p.p_beta = 1.0
    p.ss_alpha = 1.0e-3
    p.kill_width = 1.0
    p.D_G_ratio = 0.01
    p.Stokes = 0.01
    p.planet_mass = 1.0e-4
    p.planet_radius = 1.0
    p.planet_ecc = 0.0
    p.rs_fac = 0.6
    p.ramp_time = 0.5 * TWOPI
    p.sav_interval = 10.0 * TWOPI
    p.end_time = 1000.0 * TWOPI
    p.noise_rho = 1.0e-4
    p.noise_vr = 0.0
    p.seed = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def disk3d(nr: int, nphi: int, ntheta: int, *, precision: int = 4, **kw) -> Params:
    """3D adiabatic disk in spherical coordinates (init_flag 3): r in
    [0.4, 2.5] log-spaced, phi in [0, 2pi) uniform, theta uniform in
    [pi/2 - 3h, pi/2 + 3h] with reflecting polar boundaries; PEM4 + HLLC,
    EOS_flag 2 (gamma 1.4, internal energy), temperature h/r = 0.05 without
    flaring (p_beta = 1); volume density slope p_alpha = 1 - p_beta/2 + 3/2
    = 2, i.e. Sigma ~ r^-1 (the reference's 3D formula, parameters.h:155-159);
    alpha = 1e-4; one planet at r = 1 on a circular orbit with spline
    softening rs_fac (q/3)^(1/3), rs_fac = 0.5; orbital advection; CFL 0.3."""
    p = disk2d(nr, nphi, precision=precision)
    p.ndim = 3
    p.zres = ntheta
    p.geomx, p.geomy, p.geomz = 2, 4, 5
    p.gridx, p.gridy, p.gridz = 1, 0, 0
    p.recon_flag = 5
    p.EOS_flag = 2
    p.gam = 1.4
    p.bound_bom, p.bound_top = 2, 2
    p.zmin, p.zmax = HPI - 3.0 * 0.05, HPI + 3.0 * 0.05
    p.p_alpha = 1.0 - 0.5 * 1.0 + 1.5
    p.ss_alpha = 1.0e-4
    p.rs_fac = 0.5
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def config_c3(precision: int = 4, **kw) -> Params:
    """3D adiabatic disk, 384 x 1536 x 64 (r, phi, theta), q = 1e-4."""
    return disk3d(384, 1536, 64, precision=precision, **kw)


def config_c1(precision: int = 4, **kw) -> Params:
    """Plain method, small: 256 x 512 gas disk, q = 1e-4, alpha = 1e-3,
    100 steps, 1e-4 seeded noise on Sigma."""
    return disk2d(256, 512, precision=precision, **kw)


def config_c2(precision: int = 4, **kw) -> Params:
    """Gas + pressureless dust (St 0.01, dust/gas 0.01), 1024 x 3072, q = 3e-4,
    alpha = 1e-3, wave-killing zones, 1e-4 noise on Sigma and v_r."""
    base = dict(dust_flag=1, kill_flag=1, planet_mass=3.0e-4, noise_vr=1.0e-4)
    base.update(kw)
    return disk2d(1024, 3072, precision=precision, **base)


def config_c5(precision: int = 4, **kw) -> Params:
    """Adjoint workload's forward disk: 512 x 1024 gas, q = 1e-4, alpha = 1e-3,
    the unperturbed disk (the workload names no noise; seeded noise would
    put the limiters of many cells at their switching points, where the
    loss has kinks and finite differences stop converging)."""
    base = dict(noise_rho=0.0, noise_vr=0.0)
    base.update(kw)
    return disk2d(512, 1024, precision=precision, **base)


def batch_configs(n: int = 64, seed: int = 1, precision: int = 4) -> list:
    """The batched workload: n disks with nr from {128,256,384,512}, nphi from
    {256,512,1024,1536}, q log-uniform in [1e-5,1e-3], alpha log-uniform in
    [1e-4,1e-2], h/r uniform in [0.03,0.08], drawn from a seeded generator."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for d in range(n):
        nr = int(rng.choice([128, 256, 384, 512]))
        nphi = int(rng.choice([256, 512, 1024, 1536]))
        q = float(10.0 ** rng.uniform(-5.0, -3.0))
        alpha = float(10.0 ** rng.uniform(-4.0, -2.0))
        h = float(rng.uniform(0.03, 0.08))
        out.append(disk2d(nr, nphi, precision=precision, planet_mass=q, ss_alpha=alpha,
                          sc_h=h, seed=seed * 1000 + d))
    return out


def soft_length(p: Params) -> float:
    """Planet smoothing length in 2D (planet.cu:308-317): rs_fac * h0 * a."""
    return p.rs_fac * p.sc_h * p.planet_radius


def interior_cell_updates(p: Params, steps: int) -> float:
    return float(p.xres) * float(p.yres) * float(p.zres) * steps


__all__ = ["Params", "disk2d", "disk3d", "config_c1", "config_c2", "config_c3", "config_c5", "batch_configs",
           "soft_length", "interior_cell_updates", "TWOPI", "HPI", "math"]
