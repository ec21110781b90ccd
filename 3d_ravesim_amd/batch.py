"""Batched fused multislice over many (source, energy) instances.

The reference runs every instance as its own simulation process
(fast-wave/Readme.md:27).  Here the instances share one captured hipGraph:
``rsim_batch_create`` groups them by field shape into launches of up to 16
fields and records the whole slice loop; ``run`` replays it with one graph
launch.  Each instance is what one ``Sample.apply`` call with the fused
complex64 engine would do (big-wave/optical_element.py:502-596).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib
from .optical_element import DeviceVolume
from .vector import DeviceVector


@dataclass
class BatchItem:
    """One instance: fields, grid geometry, wavelength and its voxel volume."""
    u: DeviceVector
    U: DeviceVector
    nx: int
    ny: int
    dx: float
    wl: float
    cutoff_freq: float
    volume: DeviceVolume
    pixel_size_x: float
    pixel_size_z: float
    dy: float = 0.0
    pixel_size_y: float = 0.0
    x_pos: float = 0.0
    y_pos: float = 0.0


class MultisliceBatch:
    """A captured batched run of ``n_slices`` slices from ``first_slice``."""

    def __init__(self, items: list[BatchItem], first_slice: int, n_slices: int, flags: int = 0):
        if not items:
            raise ValueError("a batch needs at least one instance")
        arr = (_lib.BatchInstance * len(items))()
        for k, it in enumerate(items):
            if it.u.dtype.itemsize != 8 or it.U.dtype.itemsize != 8:
                raise ValueError("batched instances are complex64")
            arr[k] = _lib.BatchInstance(
                it.u.ptr, it.U.ptr, it.nx, it.ny, it.dx, it.dy or it.dx, it.wl, it.cutoff_freq,
                it.volume.handle, it.pixel_size_x, it.pixel_size_y or it.pixel_size_x,
                it.pixel_size_z, it.x_pos, it.y_pos)
        h = C.c_void_p()
        _lib.call("rsim_batch_create", C.byref(h), arr, len(items), first_slice, n_slices, flags,
                  None)
        self._h = h.value
        self._items = list(items)      # the graph points into these buffers
        g, k = C.c_int(), C.c_int()
        _lib.call("rsim_batch_info", self._h, C.byref(g), C.byref(k))
        self.groups, self.kernels = g.value, k.value

    def run(self, stream=None, sync: bool = True) -> None:
        if self._h is None:
            raise ValueError("batch closed")
        _lib.call("rsim_batch_run", self._h, stream)
        if sync:
            _lib.call("rsim_synchronize", stream)

    def close(self) -> None:
        if self._h is not None:
            _lib.call("rsim_batch_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and _lib._lib is not None:
                self.close()
        except Exception:
            pass
