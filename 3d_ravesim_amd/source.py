"""Point source of the 2D path (big-wave/source.py:44-113)."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .propagation import (
    SimParams,
    apply_frequency_cutoff_2d,
    fresnel_plane_2d,
    grid_density_check_2d,
    propagate_analytically_2d,
)
from .vector import DeviceVector


@dataclass
class PointSource:
    x: float
    z: float
    y: float = 0.0

    def propagate_to(self, z_out: float, sim_params: SimParams, cutoff_freq: float,
                     u: DeviceVector, U: DeviceVector, history) -> None:
        """Analytic spherical wave to z_out, then the circular cutoff.

        big-wave/source.py:97-107; under Fresnel scaling the tilted plane
        wave of source.py:62-90 instead.  U is scratch afterwards (not
        FFT2(u)), as in the reference (source.py:38-40).
        """
        dz = z_out - self.z
        if sim_params.use_fresnel_scaling:
            if sim_params.fresnel_effective_z <= 0 or sim_params.fresnel_magnification <= 0:
                raise ValueError("Fresnel detector geometry has not been configured")
            if dz <= 0:
                raise ValueError("Fresnel source-to-sample distance must be positive")
            if history is not None:
                raise ValueError("physical-z source history is not supported with Fresnel scaling")
            fresnel_plane_2d(u, dz, self.x, self.y, sim_params)
        else:
            grid_density_check_2d(dz, self.x, sim_params.nx, sim_params.dx, self.y,
                                  sim_params.ny, sim_params.get_dy(), sim_params.wl)
            propagate_analytically_2d(u, dz, self.x, self.y, sim_params)
        u.fft2(U, sim_params.nx, sim_params.ny)
        apply_frequency_cutoff_2d(U, cutoff_freq, sim_params.dx, sim_params.get_dy(),
                                  sim_params.nx, sim_params.ny, sim_params.chunk_size)
        U.ifft2(u, sim_params.nx, sim_params.ny)
        if history is not None:
            # analytic frames between the source and z_out (source.py:109-113)
            from .history import generate_analytic_history_2d
            zs = np.arange(0, z_out, history[1])[1:]
            generate_analytic_history_2d(history[0], zs, self.x, self.y, sim_params, cutoff_freq)
