"""Optical elements of the 2D path: Material, material_factor, Sample.

Mirrors big-wave/optical_element.py:24-116 and the 2D Sample
(:383-596).  The voxel volume is uploaded once to HBM as uint8 material ids
(``DeviceVolume``); ``Sample.apply`` runs the whole slice loop on the GPU.
"""

from __future__ import annotations

import ctypes as C
import logging
from dataclasses import dataclass
from typing import Optional, Protocol

import numpy as np

from . import _lib
from .propagation import SimParams, require_row_aligned_chunk_size
from .vector import DeviceVector

logger = logging.getLogger("big-wave")

ENGINES = {"plain": _lib.RSIM_PLAIN, "fused": _lib.RSIM_FUSED}


@dataclass
class Material:
    """big-wave/optical_element.py:24-29."""

    mat: str
    density: float


DeltabetaTable = list  # list[tuple[Material, complex]] (optical_element.py:35)


def material_factor(deltabeta, thickness: float, wl: float):
    """exp(-2*pi*i*t/wl*conj(db)) (big-wave/optical_element.py:39-42).

    Host helper for scalar/tabular values (e.g. per-material factors); the
    per-voxel factors of a Sample are evaluated on the GPU.
    """
    return np.exp(-2j * np.pi * thickness / wl * np.conj(deltabeta))


def extract_from_deltabeta_table(material: Material, deltabeta_table) -> complex:
    """big-wave/optical_element.py:45-66."""
    candidates = sorted(
        [(abs(tm[0].density - material.density), tm[1]) for tm in deltabeta_table
         if tm[0].mat == material.mat], key=lambda c: c[0])
    assert len(candidates) > 0, f"Material {material.mat} not found in deltabeta table"
    assert candidates[0][0] < 1e-5, f"Material {material.mat} found with density error of {candidates[0][0]}"
    return complex(candidates[0][1])


class DeviceVolume:
    """HBM-resident uint8 material-id volume plus its delta-beta table."""

    def __init__(self, handle: int, shape: tuple[int, int, int], db_list: np.ndarray):
        self._h = handle
        self.shape = shape
        self.db_list = db_list

    @staticmethod
    def _db_arg(db_list) -> tuple[np.ndarray, int]:
        db = np.ascontiguousarray(np.asarray(db_list, dtype=np.complex128))
        flat = np.ascontiguousarray(db.view(np.float64))
        return flat, len(db)

    @classmethod
    def from_numpy(cls, grid: np.ndarray, db_list) -> "DeviceVolume":
        if grid.ndim != 3:
            raise ValueError(f"Sample grid must be 3D for the 2D path, got shape {grid.shape}")
        if grid.dtype not in (np.uint8, np.uint16, np.uint32, np.int32):
            raise ValueError(f"grid dtype must be uint8/uint16/uint32/int32, got {grid.dtype}")
        g = np.ascontiguousarray(grid)
        if g.dtype == np.int32:
            if g.size and g.min() < 0:
                raise ValueError("negative material id in grid")
            g = g.view(np.uint32)
        flat, n = cls._db_arg(db_list)
        h = C.c_void_p()
        _lib.call("rsim_volume_create", C.byref(h), g.ctypes.data, g.dtype.itemsize,
                  g.shape[0], g.shape[1], g.shape[2], flat.ctypes.data_as(_lib._dp), n)
        return cls(h.value, tuple(g.shape), np.asarray(db_list, dtype=np.complex128))

    @classmethod
    def empty(cls, shape: tuple[int, int, int], db_list) -> "DeviceVolume":
        flat, n = cls._db_arg(db_list)
        h = C.c_void_p()
        _lib.call("rsim_volume_create_empty", C.byref(h), shape[0], shape[1], shape[2],
                  flat.ctypes.data_as(_lib._dp), n)
        return cls(h.value, tuple(shape), np.asarray(db_list, dtype=np.complex128))

    def paint_spheres(self, centres_radii: np.ndarray, materials: np.ndarray) -> None:
        """Spheres (cz, cy, cx, r) in voxel units, painted in order."""
        s = np.ascontiguousarray(centres_radii, dtype=np.float64).reshape(-1, 4)
        m = np.ascontiguousarray(materials, dtype=np.int32).reshape(-1)
        if len(s) != len(m):
            raise ValueError("one material per sphere")
        _lib.call("rsim_volume_paint_spheres", self.handle, s.ctypes.data_as(_lib._dp),
                  m.ctypes.data_as(_lib._ip), len(m), None)
        _lib.call("rsim_synchronize", None)

    def ids_to_numpy(self) -> np.ndarray:
        p = C.c_void_p()
        _lib.call("rsim_volume_info", self.handle, C.byref(p), None, None, None)
        out = np.empty(self.shape, dtype=np.uint8)
        _lib.call("rsim_memcpy_d2h", out.ctypes.data, p.value, out.nbytes, None)
        return out

    @property
    def handle(self) -> int:
        if self._h is None:
            raise ValueError("volume destroyed")
        return self._h

    def destroy(self) -> None:
        if self._h is not None:
            _lib.call("rsim_volume_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and _lib._lib is not None:
                _lib._lib.rsim_volume_destroy(self._h)
                self._h = None
        except Exception:
            pass


class OpticalElement(Protocol):
    """big-wave/optical_element.py:69-116."""

    z_start: float
    x_positions: np.ndarray

    def get_thickness(self) -> float: ...
    def apply(self, u, U, sim_params: SimParams, cutoff_freq: float, stepping_iteration: int,
              history) -> None: ...
    def check_valid(self) -> None: ...
    def store_deltabetas(self, deltabeta_table) -> None: ...


@dataclass
class Sample:
    """Voxel sample grid[z, y, x] of material ids, 0 = vacuum.

    big-wave/optical_element.py:383-596.  ``grid`` may be a numpy array
    (uploaded on first use) or a ``DeviceVolume``.  ``engine`` selects the
    multislice kernels: "fused" (complex64 fused row/column passes), "plain"
    (unfused kernels in the field precision) or "auto" (fused for complex64).
    """

    z_start: float
    pixel_size_x: float
    pixel_size_z: float
    grid: object
    materials: list
    x_positions: np.ndarray
    pixel_size_y: float = 0.0
    y_positions: Optional[np.ndarray] = None
    engine: str = "auto"

    def check_valid(self) -> None:
        shape = self.grid.shape
        if len(shape) != 3:
            raise ValueError(f"Sample grid must be 3D for the 2D path, got shape {shape}")
        assert shape[0] > 0 and shape[1] >= 2 and shape[2] >= 2
        if isinstance(self.grid, np.ndarray):
            assert self.grid.dtype in (np.uint8, np.uint16, np.uint32, np.int32), \
                f"grid dtype must be uint8/uint16/uint32/int32, got {self.grid.dtype}"
            assert len(self.materials) >= np.max(self.grid)
        assert self.z_start > 0
        assert self.pixel_size_x > 0
        assert self.pixel_size_z > 0
        if self.y_positions is not None:
            assert len(self.y_positions) == len(self.x_positions)

    def get_thickness(self) -> float:
        return self.pixel_size_z * self.grid.shape[0]

    def store_deltabetas(self, deltabeta_table) -> None:
        """big-wave/optical_element.py:433-440 (db_list[0] = 0 is vacuum)."""
        self.db_list = np.zeros(len(self.materials) + 1, dtype=np.complex128)
        for idx, material in enumerate(self.materials):
            self.db_list[idx + 1] = extract_from_deltabeta_table(material, deltabeta_table)
        self._device_volume = None

    def device_volume(self) -> DeviceVolume:
        if isinstance(self.grid, DeviceVolume):
            return self.grid
        vol = getattr(self, "_device_volume", None)
        if vol is None:
            if not hasattr(self, "db_list"):
                raise ValueError("store_deltabetas must be called before apply")
            vol = DeviceVolume.from_numpy(np.asarray(self.grid), self.db_list)
            self._device_volume = vol
        return vol

    def engine_code(self, dtype) -> int:
        e = self.engine
        if e == "auto":
            e = "fused" if np.dtype(dtype) == np.complex64 else "plain"
        if e not in ENGINES:
            raise ValueError(f"unknown engine {self.engine!r}; expected auto, fused or plain")
        return ENGINES[e]

    def apply(self, u: DeviceVector, U: DeviceVector, sim_params: SimParams, cutoff_freq: float,
              stepping_iteration: int, history) -> None:
        """Sample._apply_2d (big-wave/optical_element.py:502-596) on the GPU.

        After return U == FFT2(u) (OpticalElement.apply contract, :97).
        """
        if not isinstance(u, DeviceVector) or not isinstance(U, DeviceVector):
            raise TypeError("Sample.apply needs DeviceVector fields")
        nx, ny = sim_params.nx, sim_params.ny
        require_row_aligned_chunk_size(sim_params.chunk_size, nx, ny)
        vol = self.device_volume()
        z_len = vol.shape[0]
        py = self.pixel_size_y if self.pixel_size_y != 0.0 else self.pixel_size_x
        x_pos = float(self.x_positions[stepping_iteration])
        y_pos = float(self.y_positions[stepping_iteration]) if self.y_positions is not None else 0.0
        start = int(getattr(self, "_checkpoint_start_slice", 0))
        if start < 0 or start > z_len:
            raise ValueError(f"invalid Sample checkpoint slice {start}/{z_len}")
        engine = self.engine_code(u.dtype)
        args = (u.ptr, U.ptr, nx, ny, u.code, float(sim_params.dx), float(sim_params.get_dy()),
                float(sim_params.wl), float(cutoff_freq), vol.handle, float(self.pixel_size_x),
                float(py), float(self.pixel_size_z), x_pos, y_pos)
        pz = self.pixel_size_z
        scales = dzs = None
        if sim_params.use_cone_beam_bpm:
            # moving coordinates: slice k samples the voxels at m(z_k + pz/2) times
            # the field coordinates and propagates effective_slice_dz (:549-592)
            zs = self.z_start + np.arange(z_len, dtype=np.float64) * pz
            scales = np.array([sim_params.fresnel_transverse_scale(z + 0.5 * pz) for z in zs])
            dzs = np.array([sim_params.effective_slice_dz(z, pz) for z in zs])

        def run(first: int, n: int, frames) -> None:
            if frames is not None:
                _lib.call("rsim_sample_apply_2d_history", *args, first, n,
                          scales[first:].ctypes.data_as(_lib._dp) if scales is not None else None,
                          dzs[first:].ctypes.data_as(_lib._dp) if dzs is not None else None,
                          engine, 0, C.byref(frames), None)
            elif scales is not None:
                _lib.call("rsim_sample_apply_2d_scaled", *args, first, n,
                          scales[first:].ctypes.data_as(_lib._dp),
                          dzs[first:].ctypes.data_as(_lib._dp), engine, 0, None)
            else:
                _lib.call("rsim_sample_apply_2d", *args, first, n, engine, 0, None)

        from ._control import run_scope
        from .history import drive_slices
        with run_scope(u):
            drive_slices(run, self, sim_params, history, start, z_len)
