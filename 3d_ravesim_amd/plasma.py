"""Plasma optics of big-wave/plasma.py (:380-434) on the GPU.

``plasma_delta_beta_grid_multi`` takes the same arguments as the reference
and returns (delta, beta, attenuation_length_cm); the per-voxel arithmetic
runs in the HIP library (FP64, the reference's operation order).  The
Chantler scattering factors (f1bar, f2eff_bar, zbar, z2bar) come from the
caller: the reference looks them up with xraydb (nist_lookup), which this
engine does not ship.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib

# nist_lookup/nist_lookup/physical_constants.py:20-27 (CODATA 2018, scipy~=1.10)
PLANCK_HC = 1.e10 * 6.62607015e-34 * 299792458.0 / 1.602176634e-19
R_ELECTRON_CM = 100.0 * 2.8179403262e-15


class DevicePlasma:
    """HBM-resident (n_e, n_i, T_e, Z*) grids of shape (nz, ny, nx)."""

    def __init__(self, ne, ni, te, zstar):
        grids = [np.asarray(g) for g in (ne, ni, te, zstar)]
        shape = grids[0].shape
        if any(g.shape != shape for g in grids):
            raise ValueError("plasma grid tiles must have identical shapes")
        if len(shape) != 3:
            raise ValueError(f"plasma grids must be 3D (z, y, x), got shape {shape}")
        dt = np.float32 if all(g.dtype == np.float32 for g in grids) else np.float64
        self._grids = [np.ascontiguousarray(g, dtype=dt) for g in grids]
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dt)
        h = C.c_void_p()
        _lib.call("rsim_plasma_create", C.byref(h), *[g.ctypes.data for g in self._grids],
                  self.dtype.itemsize, *self.shape)
        self._h = h.value
        self._grids = None

    @property
    def handle(self) -> int:
        if self._h is None:
            raise ValueError("plasma destroyed")
        return self._h

    def deltabeta(self, energy: float, optics, first_slice: int = 0,
                  n_slices: int | None = None) -> np.ndarray:
        """complex128 delta + i*beta of slices [first, first + n) (host copy)."""
        from .vector import DeviceVector
        nz, ny, nx = self.shape
        n = nz - first_slice if n_slices is None else n_slices
        out = DeviceVector(max(n, 1) * ny * nx, np.complex128, zero=False)
        opt = np.ascontiguousarray(optics, dtype=np.float64)
        _lib.call("rsim_plasma_deltabeta", self.handle, first_slice, n, float(energy),
                  opt.ctypes.data_as(_lib._dp), out.ptr, None)
        res = out.to_numpy()[: n * ny * nx].reshape(n, ny, nx)
        out.drop()
        return res

    def destroy(self) -> None:
        if self._h is not None:
            _lib.call("rsim_plasma_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and _lib._lib is not None:
                _lib._lib.rsim_plasma_destroy(self._h)
                self._h = None
        except Exception:
            pass


def plasma_delta_beta_grid_multi(ne_grid, ni_grid, te_grid, zstar_grid, f1bar: float,
                                 f2eff_bar: float, zbar: float, z2bar: float, energy: float):
    """(delta, beta, attenuation_length_cm) of equally shaped tiles
    (big-wave/plasma.py:380-434)."""
    ne = np.asarray(ne_grid, dtype=np.float64)
    ni = np.asarray(ni_grid, dtype=np.float64)
    te = np.asarray(te_grid, dtype=np.float64)
    zs = np.asarray(zstar_grid, dtype=np.float64)
    if not (ne.shape == ni.shape == te.shape == zs.shape):
        raise ValueError("plasma grid tiles must have identical shapes")
    if energy <= 0.0:
        raise ValueError(f"energy must be positive, got {energy}")
    if ne.size == 0:
        z = np.zeros(ne.shape)
        return z, z.copy(), np.full(ne.shape, np.inf)
    dev = DevicePlasma(*(a.reshape(1, 1, -1) for a in (ne, ni, te, zs)))
    db = dev.deltabeta(energy, (f1bar, f2eff_bar, zbar, z2bar)).reshape(ne.shape)
    dev.destroy()
    delta, beta = db.real.copy(), db.imag.copy()
    lamb_cm = 1.0e-8 * PLANCK_HC / energy
    atlen = np.full(ne.shape, np.inf, dtype=np.float64)
    pos = beta > 0.0
    atlen[pos] = lamb_cm / (4.0 * math.pi * beta[pos])
    return delta, beta, atlen
