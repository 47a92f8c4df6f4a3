"""The grid-barrier probe behind the batch's launch design: a cooperative
launch with a grid-wide barrier after each pass and a captured graph of one
kernel node per pass carry the same data-dependent words (the entry point
checks them), and both report a positive time per pass boundary."""
import pytest

from penguin_amd import PenguinError
from penguin_amd.disk import probe_grid_barrier

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("blocks", [1, 64, 1024])
def test_probe_runs_and_checks_its_words(blocks):
    ub, un = probe_grid_barrier(blocks, 20)
    assert ub > 0.0 and un > 0.0


def test_probe_refuses_a_grid_that_cannot_be_resident():
    with pytest.raises(PenguinError):
        probe_grid_barrier(1 << 20, 4)
    with pytest.raises(PenguinError):
        probe_grid_barrier(64, 1)
