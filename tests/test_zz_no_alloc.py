"""pg_torque_noalloc, the allocation-free variant of pg_torque, through the
C ABI: the same bits as pg_torque on two calls in a row, the second on a
smaller disk (the shared device array keeps the first call's partials, of
which the second reads only its own), then again on the first disk; and a
disk with more tiles than the fixed array holds is refused with an error.
"""
import ctypes

import pytest

from penguin_amd import Disk, config_c2
from penguin_amd.disk import load_library

pytestmark = pytest.mark.gpu

PG_E_INVALID = -1


def _noalloc(d):
    x = ctypes.c_double()
    rc = load_library().pg_torque_noalloc(d._h, ctypes.byref(x))
    assert rc == 0, load_library().pg_last_error().decode()
    return x.value


def test_noalloc_torque_equals_torque():
    big = Disk(config_c2(precision=4, xres=512, yres=1536))    # 8 x 384 tiles of 64 x 4
    small = Disk(config_c2(precision=8, xres=96, yres=160))    # 2 x 40 tiles of 64 x 4
    big.run(3)
    small.run(3)
    a_big = _noalloc(big)
    a_small = _noalloc(small)
    b_big, b_small = big.torque(), small.torque()
    print(f"torque 512x1536 FP32: {a_big!r} {b_big!r}; 96x160 FP64: {a_small!r} {b_small!r}", flush=True)
    assert a_big == b_big and a_small == b_small
    assert a_big != 0.0 and a_small != 0.0
    big.run(2)
    assert _noalloc(big) == big.torque()
    big.close()
    small.close()


def test_noalloc_torque_refuses_a_disk_beyond_its_array():
    d = Disk(config_c2(precision=4, xres=1024, yres=16400).replace(dust_flag=0))   # 16 x 1025 tiles of 64 x 16
    x = ctypes.c_double()
    rc = load_library().pg_torque_noalloc(d._h, ctypes.byref(x))
    msg = load_library().pg_last_error().decode()
    print(rc, msg, flush=True)
    assert rc == PG_E_INVALID and "16384" in msg
    d.close()
