"""Vector protocol and the HBM-resident DeviceVector.

Mirrors big-wave's ``Vector`` protocol (big-wave/vector.py:13-90).  The field
lives in device memory allocated through the library; ``fft2``/``ifft2``
run the HIP FFT.  The chunked callbacks keep the reference's contract
(row-aligned chunks, mutate in place) by staging each chunk through host
memory -- they are for setup and inspection, not the hot loop, which runs
entirely on the device (``Sample.apply``, ``propagate_2d``).
"""

from __future__ import annotations

import ctypes as C
from typing import Callable, Protocol

import numpy as np

from . import _lib


class Vector(Protocol):
    """big-wave/vector.py:13-90."""

    def fft2(self, destination: "Vector", nx: int, ny: int) -> None: ...
    def ifft2(self, destination: "Vector", nx: int, ny: int) -> None: ...
    def __len__(self) -> int: ...
    def write_chunked(self, chunk_size: int, fn: Callable[[int, np.ndarray], None]) -> None: ...
    def modify_chunked(self, chunk_size: int, fn: Callable[[int, np.ndarray], None]) -> None: ...
    def read_chunked(self, chunk_size: int, fn: Callable[[int, np.ndarray], None]) -> None: ...
    def copy(self) -> "Vector": ...
    def copy_to(self, destination: "Vector") -> None: ...
    def drop(self) -> None: ...

    @property
    def dtype(self) -> np.dtype: ...


class DeviceVector:
    """A complex64/complex128 vector resident in HBM (one allocation).

    ``ptr`` is the device address (an int), usable by the C ABI and by
    ``torch`` via ``from_blob``-style wrappers.  ``progress_cb(stage,
    completed, total)`` and ``cancel_token`` (any token form big-wave
    accepts) serve the slice loops run on this field, as DiskVector's
    fft2_progress_cb / fft2_cancel_token serve its out-of-core FFT
    (big-wave/vector.py:115-143): a report every ``progress_every`` slices,
    and a cancel stops at the next slice boundary with InterruptedError.
    """

    progress_cb = None
    cancel_token = None
    progress_every = 1

    def __init__(self, n: int, dtype=np.complex128, zero: bool = True):
        self._dtype = np.dtype(dtype)
        self._code = _lib.dtype_code(self._dtype)
        if n <= 0:
            raise ValueError(f"vector length must be positive, got {n}")
        self._n = int(n)
        p = C.c_void_p()
        _lib.call("rsim_malloc", C.byref(p), self.nbytes)
        self._ptr = p.value
        if zero:
            _lib.call("rsim_memset", self._ptr, 0, self.nbytes, None)
            _lib.call("rsim_synchronize", None)

    # ---- construction / transfer -------------------------------------
    @classmethod
    def from_numpy(cls, vec: np.ndarray) -> "DeviceVector":
        vec = np.ascontiguousarray(vec).reshape(-1)
        assert vec.dtype in (np.complex64, np.complex128)
        v = cls(len(vec), vec.dtype, zero=False)
        _lib.call("rsim_memcpy_h2d", v._ptr, vec.ctypes.data, v.nbytes, None)
        return v

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self._n, dtype=self._dtype)
        _lib.call("rsim_memcpy_d2h", out.ctypes.data, self._ptr, self.nbytes, None)
        return out

    def upload(self, vec: np.ndarray) -> None:
        vec = np.ascontiguousarray(vec, dtype=self._dtype).reshape(-1)
        if len(vec) != self._n:
            raise ValueError("length mismatch")
        _lib.call("rsim_memcpy_h2d", self._ptr, vec.ctypes.data, self.nbytes, None)

    # ---- protocol ------------------------------------------------------
    @property
    def ptr(self) -> int:
        if self._ptr is None:
            raise ValueError("vector has been dropped")
        return self._ptr

    @property
    def dtype(self) -> np.dtype:
        return self._dtype

    @property
    def code(self) -> int:
        return self._code

    @property
    def nbytes(self) -> int:
        return self._n * self._dtype.itemsize

    def __len__(self) -> int:
        return self._n

    def _check_pair(self, dst: "DeviceVector", nx: int, ny: int) -> None:
        if not isinstance(dst, DeviceVector):
            raise TypeError("destination must be a DeviceVector")
        if nx * ny != self._n or len(dst) != self._n:
            raise ValueError("2D vector lengths must equal nx*ny")
        if dst.dtype != self._dtype:
            raise ValueError("source and destination dtypes differ")

    def fft2(self, destination: "DeviceVector", nx: int, ny: int) -> None:
        """Forward 2D FFT of the (ny, nx) view (vector.py:341-343)."""
        self._check_pair(destination, nx, ny)
        _lib.call("rsim_fft2", destination.ptr, self.ptr, nx, ny, self._code, _lib.RSIM_FORWARD, None)

    def ifft2(self, destination: "DeviceVector", nx: int, ny: int) -> None:
        """Inverse 2D FFT with 1/(nx*ny) (vector.py:345-347)."""
        self._check_pair(destination, nx, ny)
        _lib.call("rsim_fft2", destination.ptr, self.ptr, nx, ny, self._code, _lib.RSIM_INVERSE, None)

    def write_chunked(self, chunk_size: int, fn: Callable[[int, np.ndarray], None]) -> None:
        for i in range(0, self._n, chunk_size):
            m = min(chunk_size, self._n - i)
            chunk = np.zeros(m, dtype=self._dtype)
            fn(i, chunk)
            _lib.call("rsim_memcpy_h2d", self.ptr + i * self._dtype.itemsize, chunk.ctypes.data,
                      chunk.nbytes, None)

    def modify_chunked(self, chunk_size: int, fn: Callable[[int, np.ndarray], None]) -> None:
        for i in range(0, self._n, chunk_size):
            m = min(chunk_size, self._n - i)
            chunk = np.empty(m, dtype=self._dtype)
            off = self.ptr + i * self._dtype.itemsize
            _lib.call("rsim_memcpy_d2h", chunk.ctypes.data, off, chunk.nbytes, None)
            fn(i, chunk)
            _lib.call("rsim_memcpy_h2d", off, chunk.ctypes.data, chunk.nbytes, None)

    def read_chunked(self, chunk_size: int, fn: Callable[[int, np.ndarray], None]) -> None:
        for i in range(0, self._n, chunk_size):
            m = min(chunk_size, self._n - i)
            chunk = np.empty(m, dtype=self._dtype)
            _lib.call("rsim_memcpy_d2h", chunk.ctypes.data, self.ptr + i * self._dtype.itemsize,
                      chunk.nbytes, None)
            chunk.flags.writeable = False
            fn(i, chunk)

    def copy(self) -> "DeviceVector":
        v = DeviceVector(self._n, self._dtype, zero=False)
        self.copy_to(v)
        return v

    def copy_to(self, destination: "DeviceVector") -> None:
        if len(destination) != self._n or destination.dtype != self._dtype:
            raise ValueError("copy_to: shape/dtype mismatch")
        _lib.call("rsim_memcpy_d2d", destination.ptr, self.ptr, self.nbytes, None)

    def drop(self) -> None:
        if self._ptr is not None:
            _lib.call("rsim_free", self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            if getattr(self, "_ptr", None) is not None and _lib._lib is not None:
                _lib._lib.rsim_free(self._ptr)
                self._ptr = None
        except Exception:
            pass
