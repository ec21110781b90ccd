"""Per-source simulation driver for the 2D path (big-wave/wavesim.py:89-409).

Same sequencing as the reference: source -> gaps -> elements -> final leg
(skipping the forward FFT when the last element left U = FFT2(u)) ->
area_v1 detector, with phase stepping replayed from a snapshot.  Runtime
checkpoints and history frames are not produced by this engine.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .propagation import SimParams, propagate_2d, square_and_downsample_2d
from .vector import DeviceVector

logger = logging.getLogger("big-wave")


@dataclass
class Snapshot:
    """big-wave/wavesim.py:24-37."""

    z: float
    element_idx: int
    u: DeviceVector
    nr_phase_steps: int


def run_simulation(params: SimParams, source, elements: list, cutoff_angles: list[float],
                   u: DeviceVector, U: DeviceVector, deltabeta_table, sub_dir=None,
                   vectors_dir=None, save_final_u_vectors: bool = False, history=None,
                   save_keypoints_path=None, checkpoint=None,
                   resume_checkpoint: bool = False) -> list[np.ndarray]:
    """Run one source through the setup; returns one detector image per phase step."""
    if history is not None or checkpoint is not None or resume_checkpoint:
        raise NotImplementedError("history and runtime checkpoints are not supported by the HIP engine")
    if save_final_u_vectors or save_keypoints_path is not None:
        raise NotImplementedError("saving wave vectors is not supported by the HIP engine")
    z_tolerance = 1e-8
    cutoff_frequencies = [np.sin(a) / params.wl for a in cutoff_angles]
    for el in elements:
        el.store_deltabetas(deltabeta_table)

    u_fourier_valid = False
    cutoff_freq = cutoff_frequencies[0]
    current_z = elements[0].z_start if len(elements) > 0 else params.z_detector
    source.propagate_to(current_z, params, cutoff_freq, u, U, None)

    snapshot: Optional[Snapshot] = None

    def gap(dz_from: float, dz_to: float, freq: float) -> None:
        gap_dz = params.effective_slice_dz(dz_from, dz_to - dz_from)
        propagate_2d(u, U, params.dx, params.get_dy(), params.wl, gap_dz, params.chunk_size,
                     freq, params.nx, params.ny)

    for el_idx, el in enumerate(elements):
        if current_z < el.z_start:
            logger.info(f"Propagating from z {current_z}m to {el.z_start}m")
            gap(current_z, el.z_start, cutoff_freq)
            u_fourier_valid = True
        current_z = el.z_start
        if len(el.x_positions) > 1 and snapshot is None:
            snapshot = Snapshot(z=current_z, element_idx=el_idx, u=u.copy(),
                                nr_phase_steps=len(el.x_positions))
        cutoff_freq = cutoff_frequencies[el_idx + 1]
        el.apply(u, U, params, cutoff_freq, 0, None)
        current_z = el.z_start + el.get_thickness()
        u_fourier_valid = True

    if current_z < params.z_detector - z_tolerance:
        propagate_2d(u, U, params.dx, params.get_dy(), params.wl,
                     params.effective_final_dz(current_z), params.chunk_size, cutoff_freq,
                     params.nx, params.ny, skip_fft=u_fourier_valid)
        current_z = params.z_detector

    detector_outputs = [square_and_downsample_2d(u, params, params.z_detector)]

    if snapshot is not None:
        el_start_idx = snapshot.element_idx
        for step in range(1, snapshot.nr_phase_steps):
            current_z = snapshot.z
            snapshot.u.copy_to(u)
            cutoff_freq = cutoff_frequencies[el_start_idx]
            for el_idx, el in enumerate(elements[el_start_idx:], start=el_start_idx):
                if current_z < el.z_start:
                    gap(current_z, el.z_start, cutoff_freq)
                current_z = el.z_start
                cutoff_freq = cutoff_frequencies[el_idx + 1]
                el_step = step if len(el.x_positions) > 1 else 0
                el.apply(u, U, params, cutoff_freq, el_step, None)
                current_z = el.z_start + el.get_thickness()
            if current_z < params.z_detector - z_tolerance:
                propagate_2d(u, U, params.dx, params.get_dy(), params.wl,
                             params.effective_final_dz(current_z), params.chunk_size,
                             cutoff_freq, params.nx, params.ny)
            detector_outputs.append(square_and_downsample_2d(u, params, params.z_detector))
        snapshot.u.drop()

    return detector_outputs
