"""Simulation histories of the 2D path (big-wave/history.py:52-435).

``History`` collects detector frames with the reference's ROI, downsampling,
frame limit, cancellation and metadata; frames are kept in memory (HDF5
streaming needs h5py, which this image does not ship: a history that must
stream raises NotImplementedError).  The frames themselves are area_v1
images computed on the GPU: inside a sample's slice loop by the library
(rsim_sample_apply_2d_history / rsim_plasma_apply_2d_history), in free space
by ``propagate_with_history_2d``.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Callable, Optional, Tuple

import numpy as np

from . import _lib
from .propagation import SimParams, detector_shape, propagate_2d, square_and_downsample_2d
from .vector import DeviceVector

HISTORY_FORMAT_VERSION = 2
HISTORY_AXIS_ORDER_2D = ("z", "y", "x")


def _cancelled(token: Any | None) -> bool:
    if token is None:
        return False
    if isinstance(token, bool):
        return token
    for name in ("is_cancelled", "is_set", "cancelled"):
        value = getattr(token, name, None)
        if value is not None:
            return bool(value() if callable(value) else value)
    return False


def _pair(value, name: str) -> tuple[int, int]:
    pair = (value, value) if isinstance(value, int) else tuple(value)
    if len(pair) != 2 or any(not isinstance(item, int) or item <= 0 for item in pair):
        raise ValueError(f"{name} must contain two positive integers")
    return int(pair[0]), int(pair[1])


class History:
    """Frames (z, y, x) with ROI / downsampling applied before accounting
    (history.py:52-253)."""

    def __init__(self, storage_path: Path | str | None = None,
                 memory_budget_bytes: int | None = None, max_frames: int | None = None,
                 downsample=(1, 1), roi: tuple[int, int, int, int] | None = None,
                 progress_cb: Callable[[str, int, int], Any] | None = None,
                 cancel_token: Any | None = None, resume: bool = False) -> None:
        if memory_budget_bytes is not None and memory_budget_bytes < 0:
            raise ValueError("history memory_budget_bytes must be non-negative")
        if max_frames is not None and max_frames <= 0:
            raise ValueError("history max_frames must be positive")
        if roi is not None:
            if len(roi) != 4 or any(not isinstance(item, int) for item in roi):
                raise ValueError("history roi must be (y_start, y_stop, x_start, x_stop)")
            if roi[0] < 0 or roi[2] < 0 or roi[1] <= roi[0] or roi[3] <= roi[2]:
                raise ValueError("history roi bounds must be non-negative and increasing")
        if resume and storage_path is not None and Path(storage_path).exists():
            raise NotImplementedError("resuming an HDF5 history needs h5py")
        self.entries: list[np.ndarray] = []
        self.zs: list[float] = []
        self.storage_path = Path(storage_path) if storage_path is not None else None
        self.memory_budget_bytes = memory_budget_bytes
        self.max_frames = max_frames
        self.downsample = _pair(downsample, "history downsample")
        self.roi = roi
        self.progress_cb = progress_cb
        self.cancel_token = cancel_token
        self.frame_ndim: int | None = None
        self.frame_shape: tuple[int, ...] | None = None

    @property
    def is_streaming(self) -> bool:
        return False

    def _prepare_frame(self, frame: np.ndarray) -> np.ndarray:
        array = np.asanyarray(frame)
        if array.ndim not in (1, 2):
            raise ValueError(f"history frame must be 1D or 2D, got shape {array.shape}")
        if array.ndim == 2:
            if self.roi is not None:
                y0, y1, x0, x1 = self.roi
                if y1 > array.shape[0] or x1 > array.shape[1]:
                    raise ValueError(f"history roi {self.roi} exceeds frame shape {array.shape}")
                array = array[y0:y1, x0:x1]
            sy, sx = self.downsample
            array = array[::sy, ::sx]
        return array

    def push(self, downscaled: np.ndarray, z: float) -> None:
        """Add a frame (history.py:186-226)."""
        if _cancelled(self.cancel_token):
            raise InterruptedError("history recording cancelled")
        if self.max_frames is not None and len(self) >= self.max_frames:
            raise RuntimeError(f"history maximum frame count {self.max_frames} exceeded")
        frame = self._prepare_frame(downscaled)
        if self.frame_ndim is None:
            self.frame_ndim = frame.ndim
            self.frame_shape = tuple(frame.shape)
        elif frame.ndim != self.frame_ndim or tuple(frame.shape) != self.frame_shape:
            raise ValueError(f"history frame shape changed from {self.frame_shape} to {frame.shape}")
        should_stream = frame.ndim == 2 and self.storage_path is not None and (
            self.memory_budget_bytes == 0
            or (self.memory_budget_bytes is not None
                and sum(item.nbytes for item in self.entries) + frame.nbytes
                > self.memory_budget_bytes))
        if should_stream:
            raise NotImplementedError(
                "streaming a history to HDF5 needs h5py, which is not installed; "
                "use history storage 'memory' or a larger memory budget")
        self.entries.append(np.array(frame, copy=True))
        self.zs.append(float(z))
        if self.progress_cb is not None:
            self.progress_cb("history_frame", len(self), self.max_frames or 0)

    def get_history(self) -> np.ndarray:
        if not self.entries:
            return np.empty((0, 0))
        if self.frame_ndim == 1:
            return np.stack(self.entries, axis=-1)
        return np.stack(self.entries, axis=0)

    def get_z(self) -> list[float]:
        return self.zs

    def metadata(self) -> dict[str, Any]:
        shape = tuple(int(item) for item in self.get_history().shape)
        return {
            "format_version": HISTORY_FORMAT_VERSION if self.frame_ndim == 2 else 1,
            "axis_order": list(HISTORY_AXIS_ORDER_2D if self.frame_ndim == 2 else ("x", "z")),
            "storage": "npy",
            "shape": list(shape),
            "frames": len(self),
            "downsample_yx": list(self.downsample),
            "roi_yxyx": list(self.roi) if self.roi is not None else None,
        }

    def finalize(self) -> None:
        pass

    def close(self) -> None:
        pass

    def __len__(self) -> int:
        return len(self.entries)


@dataclass
class HistoryData:
    values: np.ndarray
    z: np.ndarray
    x: np.ndarray | None
    y: np.ndarray | None = None
    metadata: dict | None = None


def load_history(path: Path | str) -> HistoryData:
    """Load a history written by a run (history.py:265-314): canonical
    (z, y, x) .npy with its metadata, or a legacy (y, x, z) array, which is
    transposed.  HDF5 histories need h5py."""
    import yaml
    root = Path(path)
    directory = root if root.is_dir() else root.parent
    h5_path = root if root.suffix in (".h5", ".hdf5") else directory / "history.h5"
    if h5_path.exists():
        raise NotImplementedError("reading an HDF5 history needs h5py")
    x_path, y_path = directory / "history_x.npy", directory / "history_y.npy"
    z_path = directory / "history_z.npy"
    x = np.load(x_path, allow_pickle=False) if x_path.exists() else None
    y = np.load(y_path, allow_pickle=False) if y_path.exists() else None
    npy_path = root if root.suffix == ".npy" else directory / "history.npy"
    values = np.load(npy_path, mmap_mode="r", allow_pickle=False)
    z = np.load(z_path, allow_pickle=False) if z_path.exists() else np.arange(values.shape[-1])
    metadata_path = (root.with_name(f"{root.stem}_metadata.yaml")
                     if root.is_file() and root.name != "history.npy"
                     else directory / "history_metadata.yaml")
    version, axis_order = 0, None
    if metadata_path.exists():
        with open(metadata_path) as f:
            loaded = yaml.safe_load(f) or {}
        version, axis_order = int(loaded.get("format_version", 0)), list(loaded.get("axis_order", []))
    if values.ndim == 3:
        if version == HISTORY_FORMAT_VERSION and axis_order == list(HISTORY_AXIS_ORDER_2D):
            canonical = np.asarray(values)
        elif values.shape[-1] == len(z):
            canonical = np.transpose(values, (2, 0, 1))
        else:
            raise ValueError("cannot infer legacy 2D history axis order")
        metadata = {"format_version": HISTORY_FORMAT_VERSION,
                    "axis_order": list(HISTORY_AXIS_ORDER_2D), "storage": "npy"}
        return HistoryData(canonical, z, x, y, metadata)
    return HistoryData(np.asarray(values), z, x, y,
                       {"format_version": version, "axis_order": axis_order, "storage": "npy"})


def migrate_legacy_history(path: Path | str, output_path: Path | str | None = None) -> Path:
    """Write a legacy 2D history as canonical v2 without overwriting it
    (history.py:317-333)."""
    from . import config
    root = Path(path)
    directory = root if root.is_dir() else root.parent
    loaded = load_history(root)
    if loaded.values.ndim != 3:
        raise ValueError("only legacy 2D histories require migration")
    target = Path(output_path) if output_path is not None else directory / "history_v2.npy"
    np.save(target, loaded.values)
    config.save(target.with_name(f"{target.stem}_metadata.yaml"),
                {"format_version": HISTORY_FORMAT_VERSION,
                 "axis_order": list(HISTORY_AXIS_ORDER_2D), "storage": "npy",
                 "shape": list(loaded.values.shape)})
    return target


def generate_analytic_history_2d(history: History, zs: np.ndarray, x_source: float,
                                 y_source: float, params: SimParams, cutoff_freq: float) -> None:
    """Point-source history frames from the analytic 1/r^2 intensity
    (history.py:348-367)."""
    gradient = cutoff_freq * params.wl
    nx_pix = int(params.get_detector_size_x() // params.detector_pixel_size_x)
    ny_pix = int(params.get_detector_size_y() // params.detector_pixel_size_y)
    x = (np.arange(nx_pix) - nx_pix // 2) * params.detector_pixel_size_x - x_source
    y = (np.arange(ny_pix) - ny_pix // 2) * params.detector_pixel_size_y - y_source
    x2 = x * x
    for z in zs:
        intensity = np.empty((ny_pix, nx_pix), dtype=np.float64)
        for row, y_value in enumerate(y):
            r2 = x2 + y_value * y_value + z * z
            intensity[row] = (params.detector_pixel_size_x * params.detector_pixel_size_y * z
                              / np.power(r2, 1.5))
            intensity[row, np.abs(x) > gradient * z] = 0
            if abs(y_value) > gradient * z:
                intensity[row] = 0
        history.push(intensity, z)


def _history_z_steps(dz: float, hist_dz: float) -> np.ndarray:
    """history.py:370-378."""
    zs = np.arange(0, dz, hist_dz)
    if len(zs) == 0:
        return np.array([0.0, dz])
    if dz - zs[-1] < hist_dz / 2.0 and len(zs) > 1:
        zs[-1] = dz
    else:
        zs = np.append(zs, [dz])
    return zs


def propagate_with_history_2d(u: DeviceVector, U: DeviceVector, sim_params: SimParams, dz: float,
                              cutoff_freq: float, current_z: float, skip_fft: bool = False,
                              history: Optional[Tuple[History, float]] = None) -> None:
    """Free space with a frame every hist_dz, each leg from the start plane
    (history.py:406-435).  U = FFT2(u) afterwards."""
    if dz == 0:
        return
    nx, ny = sim_params.nx, sim_params.ny
    if history is None:
        propagate_2d(u, U, sim_params.dx, sim_params.get_dy(), sim_params.wl, dz,
                     sim_params.chunk_size, cutoff_freq, nx, ny, skip_fft)
        return
    hist_obj, hist_dz = history
    zs = _history_z_steps(dz, hist_dz)
    if not skip_fft:
        u.fft2(U, nx, ny)
    U_start = U.copy()
    try:
        for i in range(1, len(zs)):
            dz_step = float(zs[i] - zs[0])
            propagate_2d(u, U, sim_params.dx, sim_params.get_dy(), sim_params.wl, dz_step,
                         sim_params.chunk_size, cutoff_freq, nx, ny, True)
            z = float(zs[i] + current_z)
            hist_obj.push(square_and_downsample_2d(u, sim_params, z), z)
            if i < len(zs) - 1:
                U_start.copy_to(U)
    finally:
        U_start.drop()


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

