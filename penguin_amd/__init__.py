"""penguin_amd -- MI355X-native Godunov sweep of PEnGUIn.

Hand-written HIP kernels for gfx950 behind a C ABI (include/penguin_amd.h);
this package holds the Python host side: the run-time configuration
(``Params``, mirror of include/pg_params.h), the ctypes binding and the
driver loop (``Disk``) and batches of disks advanced together (``Batch``).
"""
from .params import (Params, batch_configs, config_c1, config_c2, config_c3, config_c5, disk2d, disk3d,
                     interior_cell_updates)
from .disk import KERNELS, Batch, Disk, PenguinError, load_library

__all__ = ["Params", "Disk", "Batch", "PenguinError", "load_library", "KERNELS", "disk2d", "disk3d",
           "config_c1", "config_c2", "config_c3", "config_c5", "batch_configs", "interior_cell_updates"]
