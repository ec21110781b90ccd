"""MI355X-native RAVE-SIM 2D Fresnel multislice engine.

The public names mirror big-wave's 2D API (SimParams, propagate_2d,
Sample, PointSource, run_simulation, ...); the arithmetic runs in hand-written
HIP kernels for gfx950 reached through the C ABI in include/ravesim.h.
The package directory name starts with a digit, so import it with
``importlib.import_module("3d_ravesim_amd")``.
"""

from . import _lib
from .vector import Vector, DeviceVector
from .propagation import (
    SimParams,
    apply_frequency_cutoff_2d,
    convert_energy_wavelength,
    detector_shape,
    fresnel_plane_2d,
    grid_density_check,
    grid_density_check_2d,
    propagate_2d,
    propagate_analytically_2d,
    require_row_aligned_chunk_size,
    square_and_downsample_2d,
)
from .optical_element import (
    DeviceVolume,
    Material,
    OpticalElement,
    Sample,
    extract_from_deltabeta_table,
    material_factor,
)
from .plasma import DevicePlasma, plasma_delta_beta_grid_multi
from .plasma_sample import PlasmaSample
from .source import PointSource
from .wavesim import run_simulation
from .multisim import compute_fresnel_cutoff_geometry, fresnel_mode_name, validate_fresnel_mode
from . import config
from . import graph
from . import spectral

__all__ = [
    "Vector", "DeviceVector", "SimParams", "apply_frequency_cutoff_2d", "convert_energy_wavelength",
    "detector_shape", "fresnel_plane_2d", "grid_density_check", "grid_density_check_2d", "propagate_2d",
    "propagate_analytically_2d", "require_row_aligned_chunk_size", "square_and_downsample_2d",
    "DeviceVolume", "Material", "OpticalElement", "Sample", "extract_from_deltabeta_table",
    "material_factor", "DevicePlasma", "plasma_delta_beta_grid_multi", "PlasmaSample", "PointSource", "run_simulation", "compute_fresnel_cutoff_geometry",
    "fresnel_mode_name", "validate_fresnel_mode", "config", "graph", "spectral",
]


def library():
    """Load libravesim_hip.so (raises if it is not built)."""
    return _lib.load()
