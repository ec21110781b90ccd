"""Histories of the 2D path: detector frames recorded along z.

What a history is follows big-wave (big-wave/history.py:52-435): frames of
the area_v1 detector image at planes along the beam, kept in canonical
(z, y, x) order, cut to a region of interest and strided before they are
counted, under an optional frame limit and cancel token; free-space legs get
a frame every ``history_dz``, each propagated from the leg's start plane;
the leg from the point source gets analytic 1/r^2 frames.

How it is kept is this package's own:
- ``FrameWindow`` crops and strides an incoming frame;
- frames go to a ``_MemoryFrames`` store, an array that grows by doubling;
- once the memory budget would be exceeded they move to a ``_SpillFrames``
  store: an append-only raw file next to ``storage_path`` with a small JSON
  index, read back through a memory map (big-wave streams to HDF5 at this
  point; h5py is not part of this image, and the raw store needs nothing
  beyond numpy).  A resumed run reopens an existing spill store and appends.
- inside a sample the frames are computed on the GPU by the slice loop
  itself (``rsim_sample_apply_2d_history``) into a device buffer that is
  copied out once per chunk of slices (``slice_loop_with_history``).
"""

from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Callable, Optional, Tuple

import numpy as np

from . import _lib
from ._control import token_is_set
from .propagation import SimParams, detector_shape, propagate_2d, square_and_downsample_2d
from .vector import DeviceVector

HISTORY_FORMAT_VERSION = 2
HISTORY_AXIS_ORDER_2D = ("z", "y", "x")
HISTORY_AXIS_ORDER_1D = ("x", "z")


class FrameWindow:
    """Region of interest (y0, y1, x0, x1) and (sy, sx) stride of 2D frames."""

    def __init__(self, roi, downsample):
        if isinstance(downsample, int):
            downsample = (downsample, downsample)
        stride = tuple(downsample)
        if len(stride) != 2 or not all(isinstance(s, int) and s > 0 for s in stride):
            raise ValueError("history downsample must contain two positive integers")
        if roi is not None:
            roi = tuple(roi)
            if len(roi) != 4 or not all(isinstance(v, int) for v in roi):
                raise ValueError("history roi must be (y_start, y_stop, x_start, x_stop)")
            y0, y1, x0, x1 = roi
            if min(y0, x0) < 0 or y1 <= y0 or x1 <= x0:
                raise ValueError("history roi bounds must be non-negative and increasing")
        self.roi = roi
        self.stride = (int(stride[0]), int(stride[1]))

    def apply(self, frame) -> np.ndarray:
        a = np.asanyarray(frame)
        if a.ndim == 1:
            return a
        if a.ndim != 2:
            raise ValueError(f"history frame must be 1D or 2D, got shape {a.shape}")
        if self.roi is not None:
            y0, y1, x0, x1 = self.roi
            if y1 > a.shape[0] or x1 > a.shape[1]:
                raise ValueError(f"history roi {self.roi} exceeds frame shape {a.shape}")
            a = a[y0:y1, x0:x1]
        return a[::self.stride[0], ::self.stride[1]]


class _MemoryFrames:
    storage = "npy"

    def __init__(self):
        self._buf: Optional[np.ndarray] = None
        self.count = 0

    def nbytes(self) -> int:
        return 0 if self._buf is None else self.count * self._buf[0].nbytes

    def append(self, frame: np.ndarray) -> None:
        if self._buf is None:
            self._buf = np.empty((4, *frame.shape), dtype=frame.dtype)
        elif self.count == len(self._buf):
            grown = np.empty((2 * len(self._buf), *frame.shape), dtype=self._buf.dtype)
            grown[:self.count] = self._buf
            self._buf = grown
        self._buf[self.count] = frame
        self.count += 1

    def frames(self) -> np.ndarray:
        return self._buf[:self.count].copy()

    def close(self) -> None:
        pass


class _SpillFrames:
    """Frames appended to ``<path>.frames`` (raw, C order) with an index
    ``<path>.frames.json`` holding dtype, frame shape, count and the z of
    every frame."""

    storage = "spill"

    def __init__(self, path: Path, shape: tuple, dtype):
        self.data_path = path.with_name(path.name + ".frames")
        self.index_path = path.with_name(path.name + ".frames.json")
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.count = 0
        self.zs: list[float] = []
        self.data_path.parent.mkdir(parents=True, exist_ok=True)
        self._fh = self.data_path.open("wb")
        self._write_index()

    @classmethod
    def reopen(cls, path: Path) -> "_SpillFrames":
        index_path = path.with_name(path.name + ".frames.json")
        info = json.loads(index_path.read_text())
        if int(info.get("format_version", 0)) != HISTORY_FORMAT_VERSION:
            raise ValueError("cannot resume a history spill of another format")
        self = cls.__new__(cls)
        self.data_path = path.with_name(path.name + ".frames")
        self.index_path = index_path
        self.shape, self.dtype = tuple(info["frame_shape"]), np.dtype(info["dtype"])
        self.count = int(info["frames"])
        self.zs = [float(z) for z in info["z"]]
        self._fh = self.data_path.open("r+b")
        self._fh.truncate(self.count * self._frame_bytes())
        self._fh.seek(0, os.SEEK_END)
        return self

    def _frame_bytes(self) -> int:
        return int(np.prod(self.shape)) * self.dtype.itemsize

    def _write_index(self) -> None:
        tmp = self.index_path.with_name(self.index_path.name + ".tmp")
        tmp.write_text(json.dumps({"format_version": HISTORY_FORMAT_VERSION,
                                   "dtype": self.dtype.str, "frame_shape": list(self.shape),
                                   "frames": self.count, "z": self.zs}))
        os.replace(tmp, self.index_path)

    def nbytes(self) -> int:
        return 0

    def append(self, frame: np.ndarray, z: float) -> None:
        self._fh.write(np.ascontiguousarray(frame, dtype=self.dtype).tobytes())
        self._fh.flush()
        self.count += 1
        self.zs.append(float(z))
        self._write_index()

    def frames(self) -> np.ndarray:
        if self.count == 0:
            return np.empty((0, *self.shape), dtype=self.dtype)
        self._fh.flush()
        return np.memmap(self.data_path, dtype=self.dtype, mode="r",
                         shape=(self.count, *self.shape))

    def close(self) -> None:
        if self._fh is not None and not self._fh.closed:
            self._fh.close()


class History:
    """Detector frames along z (big-wave/history.py:52-253 interface).

    ``memory_budget_bytes`` bounds the frames held in memory once a
    ``storage_path`` is given: 0 spills from the first 2D frame, a positive
    budget spills when the next frame would exceed it.  ``max_frames`` is a
    hard limit (RuntimeError); ``cancel_token`` stops recording
    (InterruptedError); ``progress_cb("history_frame", n, max_frames or 0)``
    follows every frame.
    """

    def __init__(self, storage_path: Path | str | None = None,
                 memory_budget_bytes: int | None = None, max_frames: int | None = None,
                 downsample=(1, 1), roi: tuple[int, int, int, int] | None = None,
                 progress_cb: Callable[[str, int, int], Any] | None = None,
                 cancel_token: Any | None = None, resume: bool = False) -> None:
        if memory_budget_bytes is not None and memory_budget_bytes < 0:
            raise ValueError("history memory_budget_bytes must be non-negative")
        if max_frames is not None and max_frames <= 0:
            raise ValueError("history max_frames must be positive")
        self.window = FrameWindow(roi, downsample)
        self.storage_path = Path(storage_path) if storage_path is not None else None
        self.memory_budget_bytes = memory_budget_bytes
        self.max_frames = max_frames
        self.progress_cb = progress_cb
        self.cancel_token = cancel_token
        self._store: Any = _MemoryFrames()
        self._zs: list[float] = []
        self._shape: Optional[tuple] = None
        if resume and self.storage_path is not None and \
                self.storage_path.with_name(self.storage_path.name + ".frames.json").exists():
            self._store = _SpillFrames.reopen(self.storage_path)
            self._zs = list(self._store.zs)
            self._shape = self._store.shape

    # -- configuration as big-wave exposes it --
    @property
    def roi(self):
        return self.window.roi

    @property
    def downsample(self) -> tuple[int, int]:
        return self.window.stride

    @property
    def frame_ndim(self) -> Optional[int]:
        return None if self._shape is None else len(self._shape)

    @property
    def frame_shape(self) -> Optional[tuple]:
        return self._shape

    @property
    def is_streaming(self) -> bool:
        return isinstance(self._store, _SpillFrames)

    # -- recording --
    def _spill_needed(self, frame: np.ndarray) -> bool:
        if frame.ndim != 2 or self.storage_path is None:
            return False
        if self.is_streaming or self.memory_budget_bytes == 0:
            return True
        return (self.memory_budget_bytes is not None
                and self._store.nbytes() + frame.nbytes > self.memory_budget_bytes)

    def _start_spill(self, frame: np.ndarray) -> None:
        spill = _SpillFrames(self.storage_path, frame.shape, frame.dtype)
        if self._store.count:
            for old, z in zip(self._store.frames(), self._zs):
                spill.append(old, z)
        self._store = spill

    def push(self, downscaled: np.ndarray, z: float) -> None:
        """Record the frame ``downscaled`` taken at ``z``."""
        if token_is_set(self.cancel_token):
            raise InterruptedError("history recording cancelled")
        if self.max_frames is not None and len(self) >= self.max_frames:
            raise RuntimeError(f"history maximum frame count {self.max_frames} exceeded")
        frame = self.window.apply(downscaled)
        if self._shape is None:
            self._shape = tuple(frame.shape)
        elif tuple(frame.shape) != self._shape:
            raise ValueError(f"history frame shape changed from {self._shape} to {frame.shape}")
        if self._spill_needed(frame):
            if not self.is_streaming:
                self._start_spill(frame)
            self._store.append(frame, z)
        else:
            self._store.append(frame)
        self._zs.append(float(z))
        if self.progress_cb is not None:
            self.progress_cb("history_frame", len(self), self.max_frames or 0)

    # -- reading --
    def get_history(self) -> np.ndarray:
        """(z, y, x) for 2D frames; 1D frames as (x, z)."""
        if len(self) == 0:
            return np.empty((0, 0))
        stack = self._store.frames()
        return np.moveaxis(stack, 0, -1) if self.frame_ndim == 1 else stack

    def get_z(self) -> list[float]:
        return self._zs

    def metadata(self) -> dict[str, Any]:
        two_d = self.frame_ndim == 2
        shape = [len(self), *self._shape] if self._shape is not None else [0, 0]
        if self.frame_ndim == 1:
            shape = [self._shape[0], len(self)]
        return {
            "format_version": HISTORY_FORMAT_VERSION if two_d else 1,
            "axis_order": list(HISTORY_AXIS_ORDER_2D if two_d else HISTORY_AXIS_ORDER_1D),
            "storage": self._store.storage,
            "shape": shape,
            "frames": len(self),
            "downsample_yx": list(self.downsample),
            "roi_yxyx": list(self.roi) if self.roi is not None else None,
        }

    def finalize(self) -> None:
        if self.is_streaming:
            self._store._fh.flush()

    def close(self) -> None:
        self._store.close()

    def remove_spill(self) -> None:
        """Delete the spill files once the frames have been saved elsewhere."""
        if self.is_streaming:
            self._store.close()
            self._store.data_path.unlink(missing_ok=True)
            self._store.index_path.unlink(missing_ok=True)

    def __len__(self) -> int:
        return len(self._zs)


@dataclass
class HistoryData:
    values: np.ndarray
    z: np.ndarray
    x: np.ndarray | None
    y: np.ndarray | None = None
    metadata: dict | None = None


def _read_yaml(path: Path) -> dict:
    import yaml
    with open(path) as handle:
        return yaml.safe_load(handle) or {}


def load_history(path: Path | str) -> HistoryData:
    """A saved history (big-wave/history.py:265-314 behaviour): ``history.npy``
    in canonical (z, y, x) order when its metadata says format 2, otherwise
    a legacy (y, x, z) array whose last axis matches the z list, returned
    transposed.  HDF5 files need h5py and are refused."""
    root = Path(path)
    folder = root if root.is_dir() else root.parent
    if root.suffix in (".h5", ".hdf5") or (root.is_dir() and (folder / "history.h5").exists()):
        raise NotImplementedError("reading an HDF5 history needs h5py")

    def optional(name):
        p = folder / name
        return np.load(p, allow_pickle=False) if p.exists() else None

    x, y, z = optional("history_x.npy"), optional("history_y.npy"), optional("history_z.npy")
    array_path = root if root.suffix == ".npy" else folder / "history.npy"
    values = np.load(array_path, mmap_mode="r", allow_pickle=False)
    if z is None:
        z = np.arange(values.shape[-1])
    if root.is_file() and root.name != "history.npy":
        meta_path = root.with_name(root.stem + "_metadata.yaml")
    else:
        meta_path = folder / "history_metadata.yaml"
    meta = _read_yaml(meta_path) if meta_path.exists() else {}
    version = int(meta.get("format_version", 0))
    order = list(meta.get("axis_order", []))
    if values.ndim != 3:
        return HistoryData(np.asarray(values), z, x, y,
                           {"format_version": version, "axis_order": order, "storage": "npy"})
    canonical = version == HISTORY_FORMAT_VERSION and order == list(HISTORY_AXIS_ORDER_2D)
    if not canonical:
        if values.shape[-1] != len(z):
            raise ValueError("cannot infer legacy 2D history axis order")
        values = values.transpose(2, 0, 1)
    return HistoryData(np.asarray(values), z, x, y,
                       {"format_version": HISTORY_FORMAT_VERSION,
                        "axis_order": list(HISTORY_AXIS_ORDER_2D), "storage": "npy"})


def migrate_legacy_history(path: Path | str, output_path: Path | str | None = None) -> Path:
    """Save a legacy (y, x, z) history as a canonical one beside it
    (``history_v2.npy`` by default); the source file is left as it is."""
    from . import config
    root = Path(path)
    data = load_history(root)
    if data.values.ndim != 3:
        raise ValueError("only legacy 2D histories require migration")
    folder = root if root.is_dir() else root.parent
    target = Path(output_path) if output_path is not None else folder / "history_v2.npy"
    np.save(target, data.values)
    config.save(target.with_name(target.stem + "_metadata.yaml"),
                {"format_version": HISTORY_FORMAT_VERSION,
                 "axis_order": list(HISTORY_AXIS_ORDER_2D), "storage": "npy",
                 "shape": list(data.values.shape)})
    return target


def generate_analytic_history_2d(history: History, zs: np.ndarray, x_source: float,
                                 y_source: float, params: SimParams, cutoff_freq: float) -> None:
    """Frames of the unobstructed point source: pixel area x z / r^3 at each
    plane z, zero outside the cone |x|, |y| <= cutoff_freq * wl * z that the
    frequency cutoff lets through (big-wave/history.py:348-367)."""
    dsx, dsy = params.detector_pixel_size_x, params.detector_pixel_size_y
    ox = int(params.get_detector_size_x() // dsx)
    oy = int(params.get_detector_size_y() // dsy)
    xs = (np.arange(ox) - ox // 2) * dsx - x_source
    ys = (np.arange(oy) - oy // 2) * dsy - y_source
    x_sq, y_sq = xs * xs, ys * ys
    slope = cutoff_freq * params.wl
    for z in zs:
        r_sq = x_sq[None, :] + y_sq[:, None] + z * z
        image = dsx * dsy * z / np.power(r_sq, 1.5)
        reach = slope * z
        image[:, np.abs(xs) > reach] = 0.0
        image[np.abs(ys) > reach, :] = 0.0
        history.push(image, z)


def _history_z_steps(dz: float, hist_dz: float) -> np.ndarray:
    """Frame planes of a free-space leg of length dz, relative to its start:
    0, hist_dz, 2 hist_dz, ... below dz, then dz itself; a last regular
    plane closer than hist_dz / 2 to dz (other than the start) moves to dz
    (big-wave/history.py:370-378)."""
    grid = np.arange(0.0, dz, hist_dz)
    close = grid.size > 1 and dz - grid[-1] < 0.5 * hist_dz
    head = grid[:-1] if close else (grid if grid.size else np.zeros(1))
    return np.concatenate([head, [dz]])


def propagate_with_history_2d(u: DeviceVector, U: DeviceVector, sim_params: SimParams, dz: float,
                              cutoff_freq: float, current_z: float, skip_fft: bool = False,
                              history: Optional[Tuple[History, float]] = None) -> None:
    """Free space over dz with a frame at each plane of ``_history_z_steps``.
    Every plane is reached from the leg's start spectrum, as big-wave does
    (big-wave/history.py:406-435), so the last frame's field is exactly the
    field of one propagation over dz.  U = FFT2(u) afterwards."""
    if dz == 0:
        return
    nx, ny, dx, dy, wl = sim_params.nx, sim_params.ny, sim_params.dx, sim_params.get_dy(), \
        sim_params.wl
    if history is None:
        propagate_2d(u, U, dx, dy, wl, dz, sim_params.chunk_size, cutoff_freq, nx, ny, skip_fft)
        return
    recorder, step = history
    planes = _history_z_steps(dz, step)
    if not skip_fft:
        u.fft2(U, nx, ny)
    start_spectrum = U.copy()
    try:
        for k, plane in enumerate(planes[1:], start=1):
            if k > 1:
                start_spectrum.copy_to(U)
            propagate_2d(u, U, dx, dy, wl, float(plane - planes[0]), sim_params.chunk_size,
                         cutoff_freq, nx, ny, True)
            z = float(plane + current_z)
            recorder.push(square_and_downsample_2d(u, sim_params, z), z)
    finally:
        start_spectrum.drop()


class DeviceFrames:
    """Device buffer of n detector frames (float64) for a slice loop's history."""

    def __init__(self, n: int, out_y: int, out_x: int):
        self.n, self.oy, self.ox = n, out_y, out_x
        self.nbytes = max(1, n * out_y * out_x * 8)
        p = C.c_void_p()
        _lib.call("rsim_malloc", C.byref(p), self.nbytes)
        self.ptr = p.value

    def spec(self, sim_params: SimParams, z_start: float) -> "_lib.HistoryFrames":
        return _lib.HistoryFrames(self.ox, self.oy, float(sim_params.detector_pixel_size_x),
                                  float(sim_params.detector_pixel_size_y), float(z_start),
                                  self.ptr)

    def to_numpy(self, n: int) -> np.ndarray:
        out = np.empty((n, self.oy, self.ox), dtype=np.float64)
        _lib.call("rsim_memcpy_d2h", out.ctypes.data, self.ptr, out.nbytes, None)
        return out

    def drop(self) -> None:
        if self.ptr is not None:
            _lib.call("rsim_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None) is not None and _lib._lib is not None:
                _lib._lib.rsim_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


# device memory for one chunk of slice-loop frames
FRAME_CHUNK_BYTES = 256 * 1024 ** 2


def slice_loop_with_history(run, history: History, sim_params: SimParams, z_start: float,
                            pz: float, first: int, count: int) -> None:
    """Run slices [first, first + count) in chunks through ``run(first, n,
    frames_spec)`` and push frame k at slice_z = z_start + k * pz
    (optical_element.py:548-551)."""
    oy, ox = detector_shape(sim_params)
    if ox <= 0 or oy <= 0:
        raise ValueError(f"2D detector output must be non-empty, got ({oy}, {ox})")
    per = max(1, FRAME_CHUNK_BYTES // (oy * ox * 8))
    frames = DeviceFrames(min(per, max(count, 1)), oy, ox)
    try:
        k = first
        while k < first + count:
            n = min(per, first + count - k)
            spec = frames.spec(sim_params, z_start)
            run(k, n, spec)
            imgs = frames.to_numpy(n)
            for j in range(n):
                slice_z = z_start + (k + j) * pz
                history.push(imgs[j], slice_z)
            k += n
    finally:
        frames.drop()


def drive_slices(run, element, sim_params: SimParams, history: Optional[History], start: int,
                 nz: int) -> None:
    """The slice loop of Sample / PlasmaSample.apply around ``run(first, n,
    frames_spec)``: history frames at every slice (optical_element.py:548-551)
    and the runtime-checkpoint callback after slices (:590-593).  With a
    ``_checkpoint_cadence`` set by the driver, the slices run in chunks that
    end on multiples of the cadence and the callback is called at chunk ends
    only (the driver's callback acts on those slices alone)."""
    callback = getattr(element, "_checkpoint_callback", None)
    cadence = int(getattr(element, "_checkpoint_cadence", 0) or 0)
    pz = element.pixel_size_z
    z_start = element.z_start

    def segment(a: int, n: int) -> None:
        if history is None:
            run(a, n, None)
        else:
            slice_loop_with_history(run, history, sim_params, z_start, pz, a, n)

    if callback is None:
        if nz > start:
            segment(start, nz - start)
        return
    a = start
    while a < nz:
        b = a + 1 if cadence <= 0 else min(nz, (a // cadence + 1) * cadence)
        segment(a, b - a)
        callback(b, z_start + b * pz)
        a = b
