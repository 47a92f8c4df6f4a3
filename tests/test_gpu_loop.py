"""The device loop's end and its failure signal (pg_run, pg_batch_run).

The reference loop stops at end_time (main.cu:323-327): the last step is
clipped to reach it exactly, and no step follows.  pg_run takes a step
count; the steps it has left once t reaches end_time have size 0 and leave
the state as it is.  A state whose CFL rate is not finite (NaN or infinite
velocities) stops the loop where it is found and pg_run returns PG_E_STATE
naming the step, instead of running on with the last finite step size.
"""
import numpy as np
import pytest

from penguin_amd import Batch, Disk, PenguinError, batch_configs, config_c1, config_c2

pytestmark = pytest.mark.gpu

PG_E_STATE = -4


@pytest.mark.parametrize("use_graph", [False, True])
def test_run_stops_at_end_time(use_graph):
    p = config_c2(precision=4, xres=96, yres=160)
    probe = Disk(p)
    dts = probe.run(6, use_graph=False)
    probe.close()
    end = float(dts[:3].sum() + 0.5 * dts[3])   # inside the fourth step
    d = Disk(p.replace(end_time=end))
    got = d.run(10, use_graph=use_graph)
    print("step sizes", got, flush=True)
    assert np.array_equal(got[:3], dts[:3])
    assert 0.0 < got[3] < dts[3]
    assert np.all(got[4:] == 0.0)
    assert d.t == end
    g, dd = d.get_state()
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(dd))
    # more steps past the end change nothing
    again = d.run(4, use_graph=use_graph)
    assert np.all(again == 0.0) and d.t == end
    g2, dd2 = d.get_state()
    assert np.array_equal(g, g2) and np.array_equal(dd, dd2)
    d.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_run_reports_a_state_that_is_not_finite(use_graph):
    p = config_c1(precision=4, xres=64, yres=96)
    d = Disk(p)
    d.run(2, use_graph=use_graph)
    g, _ = d.get_state()
    t0 = d.t
    g[2, 40, 17] = np.nan
    d.set_state(g)
    with pytest.raises(PenguinError) as e:
        d.run(5, use_graph=use_graph)
    print(e.value, flush=True)
    assert e.value.code == PG_E_STATE
    assert "step 0" in str(e.value)
    d.close()
    assert t0 > 0.0


def test_batch_reports_the_disk_whose_state_is_not_finite():
    ps = [p.replace(xres=64, yres=96) for p in batch_configs(3, seed=5, precision=4)]
    b = Batch(ps)
    b.run(2)
    g, _ = b.get_state(1)
    g[3, 10, 5] = np.inf
    b.set_state(1, g)
    with pytest.raises(PenguinError) as e:
        b.run(3)
    print(e.value, flush=True)
    assert e.value.code == PG_E_STATE and "disk 1" in str(e.value)
    b.close()
