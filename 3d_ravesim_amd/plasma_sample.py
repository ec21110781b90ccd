"""PlasmaSample of the 2D path (big-wave/plasma_sample.py:40-233).

A plasma region given by electron/ion density, temperature and ionisation
grids (nz, ny, nx).  ``apply`` runs the whole slice loop on the GPU through
rsim_plasma_apply_2d: per slice the plasma optics per voxel, bilinear
interpolation onto the field (vacuum outside), the transmission and
propagate_2d -- plain or fused engine, cone-beam scaling as for Sample.
"""

from __future__ import annotations

import ctypes as C
import logging
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .optical_element import ENGINES
from .plasma import DevicePlasma
from .propagation import SimParams, convert_energy_wavelength, require_row_aligned_chunk_size
from .vector import DeviceVector

logger = logging.getLogger("big-wave")


@dataclass
class PlasmaSample:
    """Fields of big-wave's PlasmaSample, plus ``optics`` = (f1bar, f2eff_bar,
    zbar, z2bar) -- the Chantler quadruple the reference looks up with xraydb
    (plasma_sample.py:68-84) or reads from a prepared subconfig's
    plasma_optics_table -- and ``engine`` as in Sample."""

    z_start: float
    pixel_size_x: float
    pixel_size_z: float
    ne_grid: np.ndarray
    ni_grid: np.ndarray
    te_grid: np.ndarray
    zstar_grid: np.ndarray
    Z: int
    x_positions: np.ndarray
    formula: Optional[str] = None
    pixel_size_y: float = 0.0
    y_positions: Optional[np.ndarray] = None
    optics: Optional[tuple] = None
    engine: str = "auto"

    def _scattering_factors(self, energy: float) -> tuple:
        if self.optics is not None:
            return tuple(float(v) for v in self.optics)
        raise NotImplementedError(
            "Chantler scattering factors need xraydb (nist_lookup), which the HIP engine does "
            "not ship; pass optics=(f1bar, f2eff_bar, zbar, z2bar) or a plasma_optics_table")

    def check_valid(self) -> None:
        """plasma_sample.py:86-106."""
        shape = self.ne_grid.shape
        assert len(shape) in (2, 3), f"PlasmaSample grids must be 2D (z, x) or 3D (z, y, x), got shape {shape}"
        assert shape[0] > 0
        for name, g in (("ni", self.ni_grid), ("te", self.te_grid), ("zstar", self.zstar_grid)):
            assert g.shape == shape, f"{name}_grid shape mismatch: {g.shape} != {shape}"
        assert self.z_start > 0
        assert self.pixel_size_x > 0
        assert self.pixel_size_z > 0
        if len(shape) == 3:
            assert self.pixel_size_y > 0, "3D grids require pixel_size_y > 0"
            assert shape[1] >= 2 and shape[2] >= 2, "3D grids require at least 2x2 transverse pixels"

    def get_thickness(self) -> float:
        return self.pixel_size_z * self.ne_grid.shape[0]

    def store_deltabetas(self, deltabeta_table) -> None:
        """delta/beta come from the plasma state; nothing to store (:111-113)."""

    def device_plasma(self) -> DevicePlasma:
        dev = getattr(self, "_device_plasma", None)
        if dev is None:
            dev = DevicePlasma(self.ne_grid, self.ni_grid, self.te_grid, self.zstar_grid)
            self._device_plasma = dev
        return dev

    def engine_code(self, dtype) -> int:
        e = self.engine
        if e == "auto":
            e = "fused" if np.dtype(dtype) == np.complex64 else "plain"
        if e not in ENGINES:
            raise ValueError(f"unknown engine {self.engine!r}; expected auto, fused or plain")
        return ENGINES[e]

    def apply(self, u: DeviceVector, U: DeviceVector, sim_params: SimParams, cutoff_freq: float,
              stepping_iteration: int, history) -> None:
        """PlasmaSample._apply_2d (plasma_sample.py:126-233); U = FFT2(u) after."""
        if not (sim_params.is_2d and self.ne_grid.ndim == 3):
            raise NotImplementedError("the HIP engine implements the 2D plasma path (3D grids) only")
        if not isinstance(u, DeviceVector) or not isinstance(U, DeviceVector):
            raise TypeError("PlasmaSample.apply needs DeviceVector fields")
        nx, ny = sim_params.nx, sim_params.ny
        require_row_aligned_chunk_size(sim_params.chunk_size, nx, ny)
        energy = convert_energy_wavelength(sim_params.wl)
        optics = np.ascontiguousarray(self._scattering_factors(energy), dtype=np.float64)
        dev = self.device_plasma()
        nz = dev.shape[0]
        py = self.pixel_size_y if self.pixel_size_y != 0.0 else self.pixel_size_x
        x_pos = float(self.x_positions[stepping_iteration])
        y_pos = float(self.y_positions[stepping_iteration]) if self.y_positions is not None else 0.0
        start = int(getattr(self, "_checkpoint_start_slice", 0))
        if start < 0 or start > nz:
            raise ValueError(f"invalid PlasmaSample checkpoint slice {start}/{nz}")
        pz = self.pixel_size_z
        engine = self.engine_code(u.dtype)
        scales = dzs = None
        if sim_params.use_cone_beam_bpm:
            zs = self.z_start + np.arange(nz, dtype=np.float64) * pz
            scales = np.array([sim_params.fresnel_transverse_scale(z + 0.5 * pz) for z in zs])
            dzs = np.array([sim_params.effective_slice_dz(z, pz) for z in zs])

        def run(first: int, n: int, frames) -> None:
            sp = scales[first:].ctypes.data_as(_lib._dp) if scales is not None else None
            dp = dzs[first:].ctypes.data_as(_lib._dp) if dzs is not None else None
            args = (u.ptr, U.ptr, nx, ny, u.code, float(sim_params.dx), float(sim_params.get_dy()),
                    float(sim_params.wl), float(cutoff_freq), dev.handle, float(energy),
                    optics.ctypes.data_as(_lib._dp), float(self.pixel_size_x), float(py),
                    float(pz), x_pos, y_pos, first, n, sp, dp, engine, 0)
            if frames is None:
                _lib.call("rsim_plasma_apply_2d", *args, None)
            else:
                _lib.call("rsim_plasma_apply_2d_history", *args, C.byref(frames), None)

        from ._control import run_scope
        from .history import drive_slices
        with run_scope(u):
            drive_slices(run, self, sim_params, history, start, nz)
