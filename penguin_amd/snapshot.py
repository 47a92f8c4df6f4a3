"""Snapshot files of the reference (checkpoint = snapshot).

Layout of write_check_point (reference src/output/output.cu:125-180), all
float64, fields x-fastest: t, xa[xres+1], ya[yres+1] (2D), rho, p, u, v,
then with dust rho_d, u_d, v_d and -- since the reference guards the last
dust field with ndim>1 (output.cu:171) -- w_d in 2D as well.  3D: t, xa, ya,
za[zres+1], rho, p, u, v, w.  File name:
path + "binary_" + label + "_" + 5-digit frame (main.cu:403, util.cu:25-36).
This module reads that layout with its own code.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .params import Params


def frame_num(n: int) -> str:
    return f"{n:05d}"


def label(p: Params) -> str:
    """create_label (output.cu:14-93), computed by the library."""
    from .disk import _check, load_library
    buf = ctypes.create_string_buffer(256)
    _check(load_library().pg_snapshot_label(ctypes.byref(p), buf, 256))
    return buf.value.decode()


def snapshot_name(p: Params, frame: int, path: str = "") -> str:
    return f"{path}binary_{label(p)}_{frame_num(frame)}"


def read_snapshot(filename: str, nx: int, ny: int, dust: bool = False, nz: int = 0) -> dict:
    """Fields of a snapshot; nz > 0 reads the 3D layout (za[zres+1] after ya,
    and w among the gas fields, output.cu:140-170), arrays (nz, ny, nx)."""
    dat = np.fromfile(filename, dtype=np.float64)
    n = nx * ny * max(nz, 1)
    gas = ["rho", "p", "u", "v"] + (["w"] if nz else [])
    names = gas + (["rho_d", "u_d", "v_d", "w_d"] if dust else [])
    expect = 1 + (nx + 1) + (ny + 1) + ((nz + 1) if nz else 0) + len(names) * n
    if dat.size != expect:
        raise ValueError(f"{filename}: {dat.size} doubles, expected {expect}")
    out = {"t": float(dat[0])}
    k = 1
    out["xa"] = dat[k:k + nx + 1]; k += nx + 1
    out["ya"] = dat[k:k + ny + 1]; k += ny + 1
    if nz:
        out["za"] = dat[k:k + nz + 1]; k += nz + 1
    shape = (nz, ny, nx) if nz else (ny, nx)
    for name in names:
        out[name] = dat[k:k + n].reshape(shape)
        k += n
    return out
