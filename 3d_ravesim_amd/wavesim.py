"""Per-source simulation driver for the 2D path (big-wave/wavesim.py:89-409).

Same sequencing as the reference: source -> gaps -> elements -> final leg
(skipping the forward FFT when the last element left U = FFT2(u)) ->
area_v1 detector, with phase stepping replayed from a snapshot, history
frames (analytic for the source leg, GPU detector frames in free space and
at every sample slice) and transactional runtime checkpoints with resume.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

from .checkpoint import RunCheckpoint
from .history import History, propagate_with_history_2d
from .propagation import SimParams, propagate_2d, square_and_downsample_2d
from .vector import DeviceVector

# Runtime checkpoint cadence: every N-th sample slice (wavesim.py:17-19).
CHECKPOINT_SLICE_CADENCE = 5

logger = logging.getLogger("big-wave")


@dataclass
class Snapshot:
    """big-wave/wavesim.py:24-37."""

    z: float
    element_idx: int
    u: DeviceVector
    nr_phase_steps: int


def save_keypoint(save_keypoints_path: Optional[Path], element_idx: int, sub_idx: int,
                  u: DeviceVector) -> None:
    """keypoint_<element>_<0 before | 1 after>.npy (wavesim.py:40-62)."""
    if save_keypoints_path is None:
        return
    path = Path(save_keypoints_path) / f"keypoint_{element_idx:02}_{sub_idx}.npy"
    np.save(path, u.to_numpy())


def save_u(u: DeviceVector, phase_step: int, sub_dir: Path, vectors_dir: Path,
           save_final_u_vectors: bool) -> None:
    """u_<phase>.npy in the source's directory (wavesim.py:65-86)."""
    if save_final_u_vectors:
        np.save(Path(sub_dir) / f"u_{phase_step:04}.npy", u.to_numpy())


def run_simulation(params: SimParams, source, elements: list, cutoff_angles: list[float],
                   u: DeviceVector, U: DeviceVector, deltabeta_table, sub_dir=None,
                   vectors_dir=None, save_final_u_vectors: bool = False,
                   history: Optional[Tuple[History, float]] = None, save_keypoints_path=None,
                   checkpoint: Optional[RunCheckpoint] = None,
                   resume_checkpoint: bool = False) -> list[np.ndarray]:
    """Run one source through the setup; one detector image per phase step."""
    z_tolerance = 1e-8
    if save_keypoints_path is not None and not isinstance(save_keypoints_path, Path):
        save_keypoints_path = Path(save_keypoints_path)
    if save_final_u_vectors and sub_dir is None:
        raise ValueError("save_final_u_vectors needs sub_dir")
    cutoff_frequencies = [np.sin(a) / params.wl for a in cutoff_angles]
    for el in elements:
        el.store_deltabetas(deltabeta_table)
    if checkpoint is not None and any(len(el.x_positions) > 1 for el in elements):
        raise ValueError(
            "runtime checkpointing is not supported together with phase stepping; "
            "phase Snapshot state is intentionally separate")

    resume_slice = 0
    u_fourier_valid = False
    if resume_checkpoint:
        if checkpoint is None or not checkpoint.exists:
            raise ValueError("resume_checkpoint requested but no ready checkpoint exists")
        state = checkpoint.load(u, U)
        if state.phase_step != 0:
            raise ValueError("only phase_step=0 runtime checkpoints are supported")
        current_z = state.current_z
        start_element_idx = state.element_index
        resume_slice = state.slice_index
        u_fourier_valid = state.u_fourier_valid
        cutoff_index = start_element_idx + (1 if resume_slice > 0 else 0)
        cutoff_freq = cutoff_frequencies[min(cutoff_index, len(cutoff_frequencies) - 1)]
        logger.info("Resuming checkpoint at z=%sm, element=%s, slice=%s", current_z,
                    start_element_idx, resume_slice)
    else:
        cutoff_freq = cutoff_frequencies[0]
        current_z = elements[0].z_start if len(elements) > 0 else params.z_detector
        source.propagate_to(current_z, params, cutoff_freq, u, U, history)
        start_element_idx = 0
        if checkpoint is not None:
            checkpoint.save(u, U, current_z=current_z, element_index=0, slice_index=0,
                            u_fourier_valid=False, stage="source_propagated")

    snapshot: Optional[Snapshot] = None
    for el_idx, el in enumerate(elements[start_element_idx:], start=start_element_idx):
        element_start_slice = resume_slice if el_idx == start_element_idx else 0
        if element_start_slice == 0 and current_z < el.z_start:
            logger.info(f"Propagating from z {current_z}m to {el.z_start}m")
            gap_dz = params.effective_slice_dz(current_z, el.z_start - current_z)
            propagate_with_history_2d(u, U, params, gap_dz, cutoff_freq, current_z, False, history)
            u_fourier_valid = True
        if element_start_slice == 0:
            current_z = el.z_start
        if len(el.x_positions) > 1 and snapshot is None:
            snapshot = Snapshot(z=current_z, element_idx=el_idx, u=u.copy(),
                                nr_phase_steps=len(el.x_positions))
        cutoff_freq = cutoff_frequencies[el_idx + 1]
        save_keypoint(save_keypoints_path, el_idx, 0, u)
        if element_start_slice > 0:
            setattr(el, "_checkpoint_start_slice", element_start_slice)
        if checkpoint is not None:
            def checkpoint_slice(next_slice: int, slice_z: float, el_idx=el_idx) -> None:
                if next_slice % CHECKPOINT_SLICE_CADENCE != 0:
                    return
                checkpoint.save(u, U, current_z=slice_z, element_index=el_idx,
                                slice_index=next_slice, u_fourier_valid=True,
                                stage="element_slice")
            setattr(el, "_checkpoint_callback", checkpoint_slice)
            setattr(el, "_checkpoint_cadence", CHECKPOINT_SLICE_CADENCE)
        try:
            el.apply(u, U, params, cutoff_freq, 0, history[0] if history is not None else None)
        finally:
            for attr in ("_checkpoint_start_slice", "_checkpoint_callback", "_checkpoint_cadence"):
                if hasattr(el, attr):
                    delattr(el, attr)
        save_keypoint(save_keypoints_path, el_idx, 1, u)
        current_z = el.z_start + el.get_thickness()
        u_fourier_valid = True
        resume_slice = 0
        if checkpoint is not None:
            checkpoint.save(u, U, current_z=current_z, element_index=el_idx + 1, slice_index=0,
                            u_fourier_valid=True, stage="element_complete")

    if current_z < params.z_detector - z_tolerance:
        logger.info(f"Propagating from z {current_z}m to {params.z_detector}m")
        propagate_with_history_2d(
            u, U, params, params.effective_final_dz(current_z), cutoff_freq, current_z,
            skip_fft=u_fourier_valid,
            history=None if params.use_fresnel_scaling else history)
        current_z = params.z_detector
        u_fourier_valid = True
    if checkpoint is not None:
        checkpoint.save(u, U, current_z=current_z, element_index=len(elements), slice_index=0,
                        u_fourier_valid=u_fourier_valid, stage="detector_ready")

    detector_outputs = [square_and_downsample_2d(u, params, params.z_detector)]
    save_u(u, 0, sub_dir, vectors_dir, save_final_u_vectors)

    # remaining phase steps replay from the snapshot, without history
    if snapshot is not None:
        el_start_idx = snapshot.element_idx
        for step in range(1, snapshot.nr_phase_steps):
            current_z = snapshot.z
            snapshot.u.copy_to(u)
            cutoff_freq = cutoff_frequencies[el_start_idx]
            for el_idx, el in enumerate(elements[el_start_idx:], start=el_start_idx):
                if current_z < el.z_start:
                    gap_dz = params.effective_slice_dz(current_z, el.z_start - current_z)
                    propagate_2d(u, U, params.dx, params.get_dy(), params.wl, gap_dz,
                                 params.chunk_size, cutoff_freq, params.nx, params.ny)
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
            save_u(u, step, sub_dir, vectors_dir, save_final_u_vectors)
        snapshot.u.drop()

    return detector_outputs
