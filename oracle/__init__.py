"""ctypes binding of the FP64 host oracle (oracle/penguin_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the checker.  Imported by tests/, by the smoke
check in __graft_entry__.py and by bench.py's CPU-baseline leg; the product
package penguin_amd never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from penguin_amd.params import Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
TLM_PATH = os.path.join(_HERE, "_build", "liboracle_tlm.so")
_lib = None
_tlm = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile the oracle and its tangent-linear twin with g++/OpenMP
    (seconds)."""
    if force or not (os.path.exists(LIB_PATH) and os.path.exists(TLM_PATH)):
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["all"], cwd=_HERE, check=True)
    return LIB_PATH


def _bind(L):
    vp = ctypes.c_void_p
    pp = ctypes.POINTER(Params)
    sig = {
        "orc_create": (vp, [pp]),
        "orc_destroy": (None, [vp]),
        "orc_num_threads": (ctypes.c_int, []),
        "orc_get_state": (None, [vp, _dp, _dp]),
        "orc_set_state": (None, [vp, _dp, _dp]),
        "orc_get_coords": (None, [vp, _dp, _dp, _dp, _dp]),
        "orc_get_zcoords": (None, [vp, _dp, _dp]),
        "orc_get_planet": (None, [vp, _dp]),
        "orc_get_rot": (None, [vp, _dp]),
        "orc_global_dt": (ctypes.c_double, [vp, ctypes.c_double]),
        "orc_solve": (None, [vp, ctypes.c_double, ctypes.c_double]),
        "orc_run": (None, [vp, ctypes.c_int, _dp, _dp, _dp]),
        "orc_viscosity": (None, [vp, ctypes.c_int, _dp]),
        "orc_sources": (None, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
        "orc_sweepx": (None, [vp, ctypes.c_double]),
        "orc_sweepz": (None, [vp, ctypes.c_double]),
        "orc_sweepy": (None, [vp, ctypes.c_double, ctypes.c_double]),
        "orc_sweepx_dust": (None, [vp, ctypes.c_double]),
        "orc_sweepy_dust": (None, [vp, ctypes.c_double, ctypes.c_double]),
        "orc_orbital_advection": (None, [vp, ctypes.c_double]),
        "orc_get_orbadv": (None, [vp, _ip, _dp]),
        "orc_killwave": (None, [vp, ctypes.c_double]),
        "orc_write_snapshot": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_double]),
        "orc_bound": (None, [vp, ctypes.c_int, _dp, _dp]),
        "orc_evolve_planet": (None, [vp, ctypes.c_double]),
        "orc_fourth_parameters": (None, [ctypes.c_int, _dp, _dp, _dp]),
        "orc_ppm_adjust": (None, [ctypes.c_double, _dp]),
        "orc_ppm_aveL": (ctypes.c_double, [ctypes.c_double, _dp]),
        "orc_ppm_aveR": (ctypes.c_double, [ctypes.c_double, _dp]),
        "orc_pem_aveL": (ctypes.c_double, [ctypes.c_double, _dp]),
        "orc_pem_aveR": (ctypes.c_double, [ctypes.c_double, _dp]),
        "orc_flatten": (ctypes.c_double, [ctypes.c_double, _dp, _dp, _dp]),
        "orc_set_state_row": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                     _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                     ctypes.c_double, _dp]),
        "orc_riemann_hllc": (None, [ctypes.c_double, ctypes.c_int, _dp, _dp]),
        "orc_riemann_row": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                   _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp]),
        "orc_dust_flux": (None, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                 ctypes.c_int, ctypes.c_double, _dp]),
        "orc_orbit_solution": (None, [pp, ctypes.c_double, _dp]),
        "orc_riemann_hllc_e": (None, [ctypes.c_double, ctypes.c_int, _dp, _dp]),
        "orc_planet_force": (None, [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]),
        "orc_average_begin": (None, [vp, ctypes.c_double]),
        "orc_average_end": (None, [vp]),
        "orc_average_get": (None, [vp, _dp]),
        "orc_torque": (ctypes.c_double, [vp]),
        "orc_set_tangent": (None, [vp, _dp, _dp]),
        "orc_get_tangent": (None, [vp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        if not hasattr(L, name):
            continue   # the tangent build leaves out the unit-level entries, the FP64 one the tangent ones
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def tlm_lib():
    """The tangent-linear build: the same oracle over forward-mode duals."""
    global _tlm
    if _tlm is None:
        if not os.path.exists(TLM_PATH):
            build()
        _tlm = _bind(ctypes.CDLL(TLM_PATH))
    return _tlm


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def arr(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


class Oracle:
    """One disk advanced by the FP64 host oracle."""

    def __init__(self, p: Params, _L=None):
        self._L = lib() if _L is None else _L
        self.p = p.copy()
        self._h = self._L.orc_create(ctypes.byref(self.p))
        self.nx, self.ny, self.nz = int(p.xres), int(p.yres), int(p.zres)
        # interior arrays: (ny, nx) in 2D, (nz, ny, nx) in 3D, x fastest
        self.shape = (self.nz, self.ny, self.nx) if int(p.ndim) == 3 else (self.ny, self.nx)
        self.zarr = self.nz + 6 if int(p.ndim) == 3 else 1
        self.t = 0.0
        self.dt = float(p.end_time)   # main.cu:144: the first cap is end_time

    def close(self):
        if self._h:
            self._L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state: gas (5,) + shape: rho, p, u, v, w; dust (4,) + shape: rho, u, v, w
    def get_state(self):
        gas = np.zeros((5,) + self.shape)
        dust = np.zeros((4,) + self.shape)
        self._L.orc_get_state(self._h, ptr(gas), ptr(dust))
        return gas, dust

    def set_state(self, gas=None, dust=None):
        g = None if gas is None else arr(gas)
        d = None if dust is None else arr(dust)
        self._L.orc_set_state(self._h, ptr(g) if g is not None else None,
                            ptr(d) if d is not None else None)

    def coords(self):
        xa = np.zeros(self.nx + 7); xv = np.zeros(self.nx + 6)
        ya = np.zeros(self.ny + 7); yv = np.zeros(self.ny + 6)
        self._L.orc_get_coords(self._h, ptr(xa), ptr(xv), ptr(ya), ptr(yv))
        return xa, xv, ya, yv

    def zcoords(self):
        """polar-angle edges and volumes with ghosts (3D)"""
        za = np.zeros(self.zarr + 1); zv = np.zeros(self.zarr)
        self._L.orc_get_zcoords(self._h, ptr(za), ptr(zv))
        return za, zv

    def planet(self):
        b = np.zeros(10)
        self._L.orc_get_planet(self._h, ptr(b))
        return b

    def rot(self):
        """orbital-advection speed per radius (per radius and polar layer in 3D), ghosts included"""
        r = np.zeros((self.zarr, self.nx + 6))
        self._L.orc_get_rot(self._h, ptr(r))
        return r[0] if self.zarr == 1 else r

    def global_dt(self, dt_prev: float) -> float:
        return self._L.orc_global_dt(self._h, dt_prev)

    def solve(self, t: float, dt: float):
        self._L.orc_solve(self._h, t, dt)

    def run(self, nsteps: int) -> np.ndarray:
        dts = np.zeros(nsteps)
        t = ctypes.c_double(self.t)
        dt = ctypes.c_double(self.dt)
        self._L.orc_run(self._h, nsteps, ctypes.byref(t), ctypes.byref(dt), ptr(dts))
        self.t, self.dt = t.value, dt.value
        return dts

    def viscosity(self, on_T: bool = False) -> np.ndarray:
        v = np.zeros((2,) + self.shape)
        self._L.orc_viscosity(self._h, int(on_T), ptr(v))
        return v

    def sources(self, mdt: float, dt: float, inplace: bool = True):
        self._L.orc_sources(self._h, mdt, dt, int(inplace))

    def sweepx(self, dt):
        self._L.orc_sweepx(self._h, dt)

    def sweepz(self, dt):
        self._L.orc_sweepz(self._h, dt)

    def sweepy(self, t, dt):
        self._L.orc_sweepy(self._h, t, dt)

    def sweepx_dust(self, dt):
        self._L.orc_sweepx_dust(self._h, dt)

    def sweepy_dust(self, t, dt):
        self._L.orc_sweepy_dust(self._h, t, dt)

    def orbital_advection(self, dt):
        self._L.orc_orbital_advection(self._h, dt)

    def orbadv(self):
        n = self.shape[0] if len(self.shape) == 3 else 1
        inc = np.zeros((n, self.nx), dtype=np.int32)
        res = np.zeros((n, self.nx))
        self._L.orc_get_orbadv(self._h, inc.ctypes.data_as(_ip), ptr(res))
        return (inc[0], res[0]) if n == 1 and len(self.shape) == 2 else (inc, res)

    def killwave(self, dt):
        self._L.orc_killwave(self._h, dt)

    def evolve_planet(self, t):
        self._L.orc_evolve_planet(self._h, t)

    def bound(self, direction: int, dust: bool = False):
        """The state with the ghost cells of one direction filled (boundx/y/z)."""
        shape = list(self.shape)
        shape[len(shape) - 1 - direction] += 6
        g = np.zeros((5,) + tuple(shape))
        d = np.zeros((4,) + tuple(shape)) if dust else None
        self._L.orc_bound(self._h, direction, ptr(g), ptr(d) if dust else None)
        return g, d

    def write_snapshot(self, filename: str, t: float):
        if self._L.orc_write_snapshot(self._h, filename.encode(), t) != 0:
            raise OSError(filename)

    def average_begin(self, interval: float):
        self._L.orc_average_begin(self._h, interval)

    def average_end(self):
        self._L.orc_average_end(self._h)

    def average(self) -> np.ndarray:
        g = np.zeros((5,) + self.shape)
        self._L.orc_average_get(self._h, ptr(g))
        return g

    def torque(self) -> float:
        return self._L.orc_torque(self._h)



class TangentOracle(Oracle):
    """The oracle compiled over forward-mode dual numbers: its FP64
    tangent-linear model.  set_tangent seeds the derivative parts of the
    state with a direction d; solve(t, dt) with the recorded (frozen) step
    sizes then carries J d, read back by get_tangent."""

    def __init__(self, p: Params):
        super().__init__(p, tlm_lib())

    def set_tangent(self, gas=None, dust=None):
        g = None if gas is None else arr(gas)
        d = None if dust is None else arr(dust)
        self._L.orc_set_tangent(self._h, ptr(g) if g is not None else None,
                                ptr(d) if d is not None else None)

    def get_tangent(self):
        gas = np.zeros((5,) + self.shape)
        dust = np.zeros((4,) + self.shape)
        self._L.orc_get_tangent(self._h, ptr(gas), ptr(dust))
        return gas, dust

def num_threads() -> int:
    return lib().orc_num_threads()


# ----------------------------------------------------------- unit functions
def fourth_parameters(geom, x6, a5):
    par = np.zeros(3)
    lib().orc_fourth_parameters(geom, ptr(arr(x6)), ptr(arr(a5)), ptr(par))
    return par


def ppm_adjust(flat, par3):
    par = arr(par3).copy()
    lib().orc_ppm_adjust(flat, ptr(par))
    return par


def ppm_aveL(q, par3):
    return lib().orc_ppm_aveL(q, ptr(arr(par3)))


def ppm_aveR(q, par3):
    return lib().orc_ppm_aveR(q, ptr(arr(par3)))


def pem_aveL(q, par3):
    return lib().orc_pem_aveL(q, ptr(arr(par3)))


def pem_aveR(q, par3):
    return lib().orc_pem_aveR(q, ptr(arr(par3)))


def flatten(gam, r5, p5, u5):
    return lib().orc_flatten(gam, ptr(arr(r5)), ptr(arr(p5)), ptr(arr(u5)))


def set_state_row(geom, gam, xa, r, p, u, v, w, flat, i, dt):
    n = len(r)
    S = np.zeros(10)
    xa, r, p, u, v, w, flat = map(arr, (xa, r, p, u, v, w, flat))
    lib().orc_set_state_row(geom, gam, n, ptr(xa), None, ptr(r), ptr(p), ptr(u), ptr(v), ptr(w),
                            ptr(flat), i, dt, ptr(S))
    return S


def riemann_hllc(gam, S10, eos=0):
    out = np.zeros(10)
    lib().orc_riemann_hllc(gam, eos, ptr(arr(S10)), ptr(out))
    return out


def riemann_hllc_e(gam, S10, eos=2):
    """riemann_hllc plus the interface velocity uprs (HLLC.cu:58-68)."""
    out = np.zeros(11)
    lib().orc_riemann_hllc_e(gam, eos, ptr(arr(S10)), ptr(out))
    return out


def planet_force(body10, rad, azi, pol, sph=True):
    """Planet acceleration (star_planet_grav, or star_planet_grav_cyl with
    sph False and pol = z) at one point, dt = 0."""
    f = np.zeros(3)
    lib().orc_planet_force(int(sph), ptr(arr(body10)), rad, azi, pol, ptr(f))
    return f


def riemann_row(geom, gam, xa, r, p, u, v, w, dt):
    n = len(r)
    D = np.zeros((5, n))
    xa, r, p, u, v, w = map(arr, (xa, r, p, u, v, w))
    lib().orc_riemann_row(geom, gam, n, ptr(xa), None, ptr(r), ptr(p), ptr(u), ptr(v), ptr(w),
                          dt, ptr(D))
    return D


def dust_flux(geom, xa, r, u, v, w, i, dt):
    n = len(r)
    F = np.zeros(4)
    xa, r, u, v, w = map(arr, (xa, r, u, v, w))
    lib().orc_dust_flux(geom, n, ptr(xa), None, ptr(r), ptr(u), ptr(v), ptr(w), i, dt, ptr(F))
    return F


def orbit_solution(p: Params, t: float) -> np.ndarray:
    b = np.zeros(10)
    lib().orc_orbit_solution(ctypes.byref(p), t, ptr(b))
    return b
