"""ctypes binding of libpenguin_amd.so (include/penguin_amd.h) and the host
driver that mirrors the reference's main loop.

``Disk`` plays the role of the reference's ``Grid* dev`` plus the loop in
main.cu:320-359: ``global_dt`` -> ``solve`` -> ``t += dt``.  The numerics run
in the HIP library; this module only moves arguments across the C ABI.
The library must have been built (``python -m penguin_amd.build``); there is
no fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import Params

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpenguin_amd.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

# name -> (restype, argtypes) of every exported entry point of penguin_amd.h
SIGNATURES = {
    "pg_last_error": (ctypes.c_char_p, []),
    "pg_version": (ctypes.c_char_p, []),
    "pg_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(_vp)]),
    "pg_destroy": (ctypes.c_int, [_vp]),
    "pg_get_state": (ctypes.c_int, [_vp, _dp, _dp]),
    "pg_set_state": (ctypes.c_int, [_vp, _dp, _dp]),
    "pg_global_dt": (ctypes.c_int, [_vp, ctypes.c_double, _dp]),
    "pg_solve": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
    "pg_run": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
    "pg_viscosity_eval": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pg_sources": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "pg_sweep_x": (ctypes.c_int, [_vp, ctypes.c_double]),
    "pg_sweep_y": (ctypes.c_int, [_vp, ctypes.c_double]),
    "pg_sweep_z": (ctypes.c_int, [_vp, ctypes.c_double]),
    "pg_sweep_x_dust": (ctypes.c_int, [_vp, ctypes.c_double]),
    "pg_sweep_y_dust": (ctypes.c_int, [_vp, ctypes.c_double]),
    "pg_orbital_advection": (ctypes.c_int, [_vp, ctypes.c_double, _ip, _dp]),
    "pg_killwave": (ctypes.c_int, [_vp, ctypes.c_double]),
    "pg_planet": (ctypes.c_int, [_vp, ctypes.c_double, _dp]),
    "pg_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "pg_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pg_profile_get": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_int)]),
    "pg_device_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "pg_write_snapshot": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_double]),
    "pg_read_snapshot": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
    "pg_snapshot_label": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_char_p, ctypes.c_size_t]),
    "pg_get_energy": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, _dp]),
    "pg_adjoint_forward": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]),
    "pg_adjoint_backward": (ctypes.c_int, [_vp, _dp, _vp]),
    "pg_adjoint_pass": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]),
    "pg_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Params), ctypes.POINTER(_vp)]),
    "pg_batch_destroy": (ctypes.c_int, [_vp]),
    "pg_batch_size": (ctypes.c_int, [_vp, _ip]),
    "pg_batch_get_state": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "pg_batch_set_state": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "pg_batch_run": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
    "pg_average_begin": (ctypes.c_int, [_vp, ctypes.c_double]),
    "pg_average_end": (ctypes.c_int, [_vp]),
    "pg_average_get": (ctypes.c_int, [_vp, _dp]),
    "pg_write_average_snapshot": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_double]),
    "pg_torque": (ctypes.c_int, [_vp, _dp]),
    "pg_torque_noalloc": (ctypes.c_int, [_vp, _dp]),
    "pg_probe_grid_barrier": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "pg_bound_x": (ctypes.c_int, [_vp, _dp, _dp]),
    "pg_bound_y": (ctypes.c_int, [_vp, _dp, _dp]),
    "pg_bound_z": (ctypes.c_int, [_vp, _dp, _dp]),
    "pg_batch_set_schedule": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pg_batch_get_schedule": (ctypes.c_int, [_vp, _ip, _ip, _ip, _ip]),
    "pg_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "pg_batch_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pg_batch_profile_get": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_int)]),
    "pg_batch_device_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
}


class Grads(ctypes.Structure):
    """pg_grads (include/penguin_amd.h)."""
    _fields_ = [("dq", ctypes.c_double), ("dalpha", ctypes.c_double), ("dstate0", _dp)]


class PenguinError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"penguin_amd error {code}: {msg}")
        self.code = code


def load_library(path: str = LIB_PATH):
    """Load the HIP library (built in-tree); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} not found: build it with `python -m penguin_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def probe_grid_barrier(blocks: int, npass: int = 200):
    """Microseconds per grid-wide barrier of a cooperative launch and per
    kernel node of a captured graph, for the same pass over `blocks`
    workgroups of 256 threads (pg_probe_grid_barrier)."""
    ub, un = ctypes.c_double(), ctypes.c_double()
    _check(load_library().pg_probe_grid_barrier(blocks, npass, ctypes.byref(ub), ctypes.byref(un)))
    return ub.value, un.value


def _check(rc: int):
    if rc != 0:
        raise PenguinError(rc, load_library().pg_last_error().decode())


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


KERNELS = ("prologue", "viscosity", "sweep_x_gas", "ring_gas", "sweep_x_dust", "ring_dust",
           "predictor_viscosity", "corrector_viscosity", "sources_corrector", "cfl")


class Disk:
    """One disk on the GPU: the reference's Grid plus its time loop."""

    def __init__(self, p: Params):
        L = load_library()
        self.p = p.copy()
        h = _vp()
        _check(L.pg_create(ctypes.byref(self.p), ctypes.byref(h)))
        self._h = h
        self.nx, self.ny, self.nz = int(p.xres), int(p.yres), int(p.zres)
        # interior arrays, x fastest: (ny, nx) in 2D, (nz, ny, nx) for a 3D disk
        self.shape = (self.nz, self.ny, self.nx) if int(p.ndim) == 3 else (self.ny, self.nx)
        self.t = 0.0
        self.dt = float(p.end_time)   # main.cu:144: the first cap is end_time

    def close(self):
        if getattr(self, "_h", None):
            load_library().pg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state -----------------------------------------------------------
    def get_state(self):
        """gas (5, ny, nx): rho, p, u, v, w; dust (4, ny, nx) or None."""
        gas = np.zeros((5,) + self.shape)
        dust = np.zeros((4,) + self.shape) if self.p.dust_flag else None
        _check(load_library().pg_get_state(self._h, _ptr(gas), _ptr(dust) if dust is not None else None))
        return gas, dust

    def set_state(self, gas=None, dust=None):
        g = None if gas is None else np.ascontiguousarray(gas, dtype=np.float64)
        d = None if dust is None else np.ascontiguousarray(dust, dtype=np.float64)
        _check(load_library().pg_set_state(self._h, _ptr(g) if g is not None else None,
                                           _ptr(d) if d is not None else None))

    # ---- stepping (timestep.cu:113-140, solve.cu:279-292, main.cu:320-359)
    def bound(self, direction: int, dust: bool = False):
        """The state with the ghost cells of one direction filled (pg_bound_x/y/z):
        arrays shaped as get_state's with that axis 6 longer."""
        shape = list(self.shape)
        shape[len(shape) - 1 - direction] += 6
        g = np.zeros((5,) + tuple(shape))
        dd = np.zeros((4,) + tuple(shape)) if dust else None
        f = ("pg_bound_x", "pg_bound_y", "pg_bound_z")[direction]
        _check(getattr(load_library(), f)(self._h, _ptr(g), _ptr(dd) if dust else None))
        return g, dd

    def global_dt(self, dt_prev: float) -> float:
        out = ctypes.c_double()
        _check(load_library().pg_global_dt(self._h, dt_prev, ctypes.byref(out)))
        return out.value

    def solve(self, t: float, dt: float):
        _check(load_library().pg_solve(self._h, t, dt))

    def step_host(self) -> float:
        """One iteration of the reference loop driven from the host."""
        dt = self.global_dt(self.dt)
        if self.t + dt > self.p.end_time:
            dt = self.p.end_time - self.t
        self.solve(self.t, dt)
        self.t += dt
        self.dt = dt
        return dt

    def run(self, nsteps: int, use_graph: bool = False) -> np.ndarray:
        """nsteps of the loop on the device; returns the step sizes."""
        dts = np.zeros(max(nsteps, 1))
        t = ctypes.c_double(self.t)
        dt = ctypes.c_double(self.dt)
        _check(load_library().pg_run(self._h, nsteps, ctypes.byref(t), ctypes.byref(dt), _ptr(dts),
                                     int(use_graph)))
        self.t, self.dt = t.value, dt.value
        return dts[:nsteps]

    # ---- module-level entry points --------------------------------------
    def viscosity(self, on_T: bool = False) -> np.ndarray:
        v = np.zeros((2,) + self.shape)
        _check(load_library().pg_viscosity_eval(self._h, int(on_T), _ptr(v)))
        return v

    def sources(self, t_planet: float, mdt: float, dt: float, inplace: bool = True):
        _check(load_library().pg_sources(self._h, t_planet, mdt, dt, int(inplace)))

    def sweep_x(self, dt: float):
        _check(load_library().pg_sweep_x(self._h, dt))

    def sweep_y(self, dt: float):
        _check(load_library().pg_sweep_y(self._h, dt))

    def sweep_x_dust(self, dt: float):
        _check(load_library().pg_sweep_x_dust(self._h, dt))

    def sweep_y_dust(self, dt: float):
        _check(load_library().pg_sweep_y_dust(self._h, dt))

    def sweep_z(self, dt: float):
        """polar-angle sweep of a 3D disk"""
        _check(load_library().pg_sweep_z(self._h, dt))

    def orbital_advection(self, dt: float):
        """integer shift and residual speed per radius ((nz, nx) in 3D)"""
        n = self.nz if len(self.shape) == 3 else 1
        inc = np.zeros((n, self.nx), dtype=np.int32)
        res = np.zeros((n, self.nx))
        _check(load_library().pg_orbital_advection(self._h, dt, inc.ctypes.data_as(_ip), _ptr(res)))
        return (inc, res) if n > 1 or len(self.shape) == 3 else (inc[0], res[0])

    def killwave(self, dt: float):
        _check(load_library().pg_killwave(self._h, dt))

    def planet(self, t: float) -> np.ndarray:
        b = np.zeros(10)
        _check(load_library().pg_planet(self._h, t, _ptr(b)))
        return b

    # ---- adjoint (pg_adjoint_forward / pg_adjoint_backward) ---------------
    def adjoint_forward(self, nsteps: int, ckpt_every: int = 20) -> np.ndarray:
        """nsteps of the device loop, keeping checkpoints for the adjoint;
        returns the step sizes (frozen for the backward sweep)."""
        dts = np.zeros(nsteps)
        _check(load_library().pg_adjoint_forward(self._h, nsteps, ckpt_every, self.t, self.dt, _ptr(dts)))
        self.t = self.t + float(np.sum(dts))
        self.dt = float(dts[-1])
        return dts

    def adjoint_backward(self, dL_dgas: np.ndarray, out: np.ndarray = None):
        """dL/d(final gas state) (5, ny, nx) -> (dL/dq, dL/dalpha, dL/d(start
        state) (5, ny, nx)).  `out`, a C-contiguous float64 (5, ny, nx) array,
        receives the last (every field is written), so repeated calls need no
        fresh host pages."""
        dL = np.ascontiguousarray(dL_dgas, dtype=np.float64)
        assert dL.shape == (5, self.ny, self.nx)
        if out is None:
            out = np.empty((5, self.ny, self.nx))
        assert out.shape == (5, self.ny, self.nx) and out.dtype == np.float64 and out.flags.c_contiguous
        d0 = out
        g = Grads(0.0, 0.0, _ptr(d0))
        _check(load_library().pg_adjoint_backward(self._h, _ptr(dL), ctypes.byref(g)))
        return g.dq, g.dalpha, d0

    ADJOINT_PASSES = ("sweep_x", "sweep_y", "orbital_advection", "viscosity", "kick", "predictor", "corrector")

    def adjoint_pass(self, name: str, t: float, dt: float, lam_out: np.ndarray):
        """Transpose of one module call at the current state: lam_out
        (3 or 2, ny, nx) -> (lam_in (3, ny, nx) over rho, u, v; (dL/dq, dL/dalpha))."""
        k = self.ADJOINT_PASSES.index(name)
        lo = np.ascontiguousarray(lam_out, dtype=np.float64)
        li = np.zeros((3, self.ny, self.nx))
        dp = np.zeros(2)
        _check(load_library().pg_adjoint_pass(self._h, k, t, dt, _ptr(lo), _ptr(li), _ptr(dp)))
        return li, dp

    # ---- diagnostics (output.cu:250-457) --------------------------------
    def average_begin(self, interval: float):
        """Start a time average of the gas over `interval` (init_average +
        averaging around every device-loop step, main.cu:344-350)."""
        _check(load_library().pg_average_begin(self._h, float(interval)))

    def average_end(self):
        _check(load_library().pg_average_end(self._h))

    def average(self) -> np.ndarray:
        """The gas average so far, (5, ny, nx): rho, p, u, v, w."""
        g = np.empty((5, self.ny, self.nx))
        _check(load_library().pg_average_get(self._h, _ptr(g)))
        return g

    def write_average_snapshot(self, filename: str, t: float = None):
        _check(load_library().pg_write_average_snapshot(self._h, filename.encode(), self.t if t is None else t))

    def torque(self) -> float:
        """Planet torque on the gas (global_sum, output.cu:336-457)."""
        x = ctypes.c_double()
        _check(load_library().pg_torque(self._h, ctypes.byref(x)))
        return x.value

    # ---- snapshots (output.cu:125-248) ----------------------------------
    def write_snapshot(self, filename: str, t: float = None):
        _check(load_library().pg_write_snapshot(self._h, filename.encode(), self.t if t is None else t))

    def read_snapshot(self, filename: str) -> float:
        t = ctypes.c_double()
        _check(load_library().pg_read_snapshot(self._h, filename.encode(), ctypes.byref(t)))
        self.t = t.value
        return t.value

    # ---- introspection ---------------------------------------------------
    def profile(self, on: bool = True):
        _check(load_library().pg_profile_enable(self._h, int(on)))

    def kernel_time_us(self, name: str):
        avg = ctypes.c_double()
        cnt = ctypes.c_int()
        _check(load_library().pg_profile_get(self._h, name.encode(), ctypes.byref(avg), ctypes.byref(cnt)))
        return avg.value, cnt.value

    def device_bytes(self) -> int:
        b = ctypes.c_size_t()
        _check(load_library().pg_device_bytes(self._h, ctypes.byref(b)))
        return b.value


class Batch:
    """Many independent disks advanced together (pg_batch_* entry points):
    one launch per pass covers every disk, each disk keeps its own CFL step,
    and a hipGraph replays a whole step of the batch."""

    def __init__(self, params):
        L = load_library()
        self.params = [p.copy() for p in params]
        arr = (Params * len(self.params))(*self.params)
        h = _vp()
        _check(L.pg_batch_create(len(self.params), arr, ctypes.byref(h)))
        self._h = h
        self.n = len(self.params)
        self.t = np.zeros(self.n)
        self.dt = np.array([float(p.end_time) for p in self.params])  # main.cu:144

    def close(self):
        if getattr(self, "_h", None):
            load_library().pg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return self.n

    def cells(self) -> int:
        return int(sum(int(p.xres) * int(p.yres) for p in self.params))

    def get_state(self, k: int):
        p = self.params[k]
        gas = np.zeros((5, int(p.yres), int(p.xres)))
        dust = np.zeros((4, int(p.yres), int(p.xres))) if p.dust_flag else None
        _check(load_library().pg_batch_get_state(self._h, k, _ptr(gas), _ptr(dust) if dust is not None else None))
        return gas, dust

    def set_state(self, k: int, gas=None, dust=None):
        g = None if gas is None else np.ascontiguousarray(gas, dtype=np.float64)
        d = None if dust is None else np.ascontiguousarray(dust, dtype=np.float64)
        _check(load_library().pg_batch_set_state(self._h, k, _ptr(g) if g is not None else None,
                                                 _ptr(d) if d is not None else None))

    def run(self, nsteps: int, use_graph: bool = True) -> np.ndarray:
        """nsteps of every disk; returns the step sizes, shape (n, nsteps)."""
        dts = np.zeros((self.n, max(nsteps, 1)))
        t = np.ascontiguousarray(self.t, dtype=np.float64)
        dt = np.ascontiguousarray(self.dt, dtype=np.float64)
        _check(load_library().pg_batch_run(self._h, nsteps, _ptr(t), _ptr(dt), _ptr(dts), int(use_graph)))
        self.t, self.dt = t, dt
        return dts[:, :nsteps]

    def set_schedule(self, tile_x: int, ring_b: int, order: int = 0, pt_rows: int = 0):
        _check(load_library().pg_batch_set_schedule(self._h, tile_x, ring_b, order, pt_rows))

    def schedule(self):
        """(x-sweep tile width, azimuthal-tile threads, tile order, pointwise tile rows)"""
        v = [ctypes.c_int32() for _ in range(4)]
        _check(load_library().pg_batch_get_schedule(self._h, *(ctypes.byref(x) for x in v)))
        return tuple(x.value for x in v)

    def autotune(self, steps: int = 5) -> float:
        """Pick the fastest schedule on this batch; returns its ms per step."""
        ms = ctypes.c_double()
        _check(load_library().pg_batch_autotune(self._h, steps, ctypes.byref(ms)))
        return ms.value

    def profile(self, on: bool = True):
        _check(load_library().pg_batch_profile_enable(self._h, int(on)))

    def kernel_time_us(self, name: str):
        avg = ctypes.c_double()
        cnt = ctypes.c_int()
        _check(load_library().pg_batch_profile_get(self._h, name.encode(), ctypes.byref(avg), ctypes.byref(cnt)))
        return avg.value, cnt.value

    def device_bytes(self) -> int:
        b = ctypes.c_size_t()
        _check(load_library().pg_batch_device_bytes(self._h, ctypes.byref(b)))
        return b.value
