"""CPU checks of the C ABI (no GPU needed): the built library exports every
entry point declared in include/penguin_amd.h, the ctypes mirror of
pg_params has the C layout, and invalid configurations fail with an error
code and message instead of crashing."""
import ctypes
import os
import re
import subprocess

import pytest

from penguin_amd.disk import LIB_PATH, SIGNATURES, PenguinError, load_library
from penguin_amd.params import Params, config_c2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "penguin_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(pg_\w+)\s*\(", text, flags=re.M)))


def test_library_is_built():
    assert os.path.exists(LIB_PATH), "run python -m penguin_amd.build"


def test_every_declared_entry_point_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 20
    L = load_library()
    for n in names:
        assert hasattr(L, n), n
        assert n in SIGNATURES, f"{n} has no ctypes signature"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (pg_\w+)", out))
    assert set(names) <= exported
    # nothing else with the pg_ prefix leaks from the library
    assert exported <= set(names)


def test_params_layout_matches_c(tmp_path):
    fields = [f for f, _ in Params._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{os.path.join(ROOT, "include", "pg_params.h")}"',
           'int main(void){', 'printf("%zu\\n", sizeof(pg_params));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof(pg_params, {f}));')
    src.append("return 0;}")
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(Params)
    for f, off in zip(fields, vals[1:]):
        assert getattr(Params, f).offset == off, f


def test_invalid_configuration_is_rejected():
    L = load_library()
    h = ctypes.c_void_p()
    for bad in (dict(ndim=3), dict(EOS_flag=2), dict(recon_flag=5), dict(precision=2),
                dict(bound_bak=2), dict(gam=1.4), dict(xres=4)):
        p = config_c2(xres=64, yres=64).replace(**bad)
        rc = L.pg_create(ctypes.byref(p), ctypes.byref(h))
        assert rc == -1, bad
        assert L.pg_last_error().decode(), bad
    with pytest.raises(PenguinError):
        from penguin_amd import Disk
        Disk(config_c2(xres=64, yres=64).replace(ndim=3))
    # 3D: what the spherical path does not cover is refused before any device work
    from penguin_amd import disk3d
    for bad in (dict(dust_flag=1), dict(kill_flag=1), dict(recon_flag=4), dict(EOS_flag=0), dict(gam=1.0),
                dict(OrbAdv_flag=0), dict(bound_bom=3), dict(bound_top=1), dict(zres=200), dict(geomz=3),
                dict(gridz=1), dict(zmin=-0.1)):
        p = disk3d(32, 32, 8).replace(**bad)
        rc = L.pg_create(ctypes.byref(p), ctypes.byref(h))
        assert rc == -1, bad
        assert "3D" in L.pg_last_error().decode() or "polar" in L.pg_last_error().decode(), bad


def test_null_handle_is_an_error_not_a_crash():
    L = load_library()
    dt = ctypes.c_double()
    assert L.pg_global_dt(None, 1.0, ctypes.byref(dt)) == -1
    assert L.pg_solve(None, 0.0, 1e-3) == -1
    assert L.pg_destroy(None) == 0
    assert L.pg_version().decode().startswith("penguin_amd")


def test_batch_configuration_is_checked_before_any_device_work():
    from penguin_amd import Batch, batch_configs
    L = load_library()
    h = ctypes.c_void_p()
    ps = batch_configs(3)
    # a member that fails the single-disk checks is named by its index
    bad = [ps[0], ps[1].replace(EOS_flag=2), ps[2]]
    arr = (Params * 3)(*bad)
    assert L.pg_batch_create(3, arr, ctypes.byref(h)) == -1
    assert "disk 1" in L.pg_last_error().decode()
    # members must share precision and the dust / viscosity / killing switches
    for change in (dict(precision=8), dict(dust_flag=1), dict(visc_flag=0), dict(kill_flag=1)):
        mixed = [ps[0], ps[1].replace(**change), ps[2]]
        arr = (Params * 3)(*mixed)
        assert L.pg_batch_create(3, arr, ctypes.byref(h)) == -1, change
        assert "must agree" in L.pg_last_error().decode(), change
    assert L.pg_batch_create(0, arr, ctypes.byref(h)) == -1
    with pytest.raises(PenguinError):
        Batch([ps[0], ps[1].replace(precision=8)])
    assert L.pg_batch_destroy(None) == 0
    t = (ctypes.c_double * 1)()
    assert L.pg_batch_run(None, 1, t, t, None, 1) == -1
    assert L.pg_batch_set_schedule(None, 256, 512, 0, 0) == -1


def test_eos_energy_matches_the_oracle_hllc_energy_flux():
    """pg_get_energy (EOS.cu:4-18) against the energy the oracle's HLLC carries
    (F_p / F_rho of a uniform state is the upwinded specific energy)."""
    import numpy as np
    import oracle
    from penguin_amd import disk3d
    L = load_library()
    e = ctypes.c_double()
    p3 = disk3d(16, 16, 8, gam=1.4)
    for r, p, u in ((1.0, 0.7, 0.3), (0.2, 0.05, -0.1), (3.0, 2.0, 1.5)):
        assert L.pg_get_energy(ctypes.byref(p3), r, p, u, 0.1, -0.2, ctypes.byref(e)) == 0
        S = np.array([r, p, u, 0.1, -0.2] * 2, dtype=float)
        out = np.zeros(10)
        oracle.lib().orc_riemann_hllc(1.4, 2, oracle.ptr(S), oracle.ptr(out))
        assert abs(e.value - out[5] / out[4]) <= 1e-14 * e.value
        assert abs(e.value - p / (0.4 * r)) <= 1e-15 * e.value
    p2 = config_c2(xres=64, yres=64)                     # EOS_flag 0: p / rho
    assert L.pg_get_energy(ctypes.byref(p2), 2.0, 3.0, 0.0, 0.0, 0.0, ctypes.byref(e)) == 0 and e.value == 1.5
    assert L.pg_get_energy(ctypes.byref(p3), -1.0, 1.0, 0.0, 0.0, 0.0, ctypes.byref(e)) == -1
