"""Snapshot files: the reference's checkpoint layout and naming
(write_check_point output.cu:125-180, create_label output.cu:14-93,
frame_num util.cu:25-36), written by the library and read back."""
import os

import numpy as np
import pytest

import oracle
from penguin_amd import Disk, config_c1, config_c2
from penguin_amd.snapshot import frame_num, label, read_snapshot, snapshot_name


def test_labels_and_names():
    # int_to_string truncates: 1e-4/3e-6 = 33.3 -> 33, 3e-4/3e-6 = 99.999... -> 99
    assert label(config_c1()) == "256x512_h50_1p33E_e0_a-30_OA_PPM4"
    assert label(config_c2()) == "1024x3072_h50_1p99E_e0_a-30_OA_St-20_PPM4"
    assert label(config_c1(planet_mass=2e-3)) == "256x512_h50_1p2J_e0_a-30_OA_PPM4"
    assert frame_num(7) == "00007" and frame_num(12345) == "12345"
    assert snapshot_name(config_c1(), 3, "/tmp/") == "/tmp/binary_256x512_h50_1p33E_e0_a-30_OA_PPM4_00003"


def test_oracle_snapshot_layout(tmp_path):
    p = config_c2(precision=8, xres=24, yres=32)
    o = oracle.Oracle(p)
    o.run(2)
    f = str(tmp_path / "snap")
    o.write_snapshot(f, 0.125)
    s = read_snapshot(f, 24, 32, dust=True)
    gas, dust = o.get_state()
    xa, _, ya, _ = o.coords()
    assert s["t"] == 0.125
    assert np.array_equal(s["xa"], xa[3:28]) and np.array_equal(s["ya"], ya[3:36])
    for k, name in enumerate(("rho", "p", "u", "v")):
        assert np.array_equal(s[name], gas[k])
    for k, name in enumerate(("rho_d", "u_d", "v_d", "w_d")):
        assert np.array_equal(s[name], dust[k])


@pytest.mark.gpu
def test_gpu_snapshot_is_byte_identical_and_restarts(tmp_path):
    p = config_c2(precision=8, xres=64, yres=96)
    o = oracle.Oracle(p)
    d = Disk(p)
    fo, fd = str(tmp_path / "oracle"), str(tmp_path / "gpu")
    o.write_snapshot(fo, 0.0)
    d.write_snapshot(fd, 0.0)
    assert open(fo, "rb").read() == open(fd, "rb").read()
    # restart: the state (gas and dust) comes back
    d.run(5)
    g1, d1 = d.get_state()
    fr = str(tmp_path / "restart")
    d.write_snapshot(fr)
    e = Disk(p)
    t = e.read_snapshot(fr)
    assert t == d.t
    g2, d2 = e.get_state()
    assert np.allclose(g2, g1, rtol=1e-15, atol=0) and np.allclose(d2, d1, rtol=1e-15, atol=0)
    # a snapshot of another resolution is rejected
    f = Disk(config_c2(precision=8, xres=32, yres=96))
    with pytest.raises(Exception):
        f.read_snapshot(fr)
