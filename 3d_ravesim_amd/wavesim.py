"""One source through the setup on HBM-resident fields.

``run_simulation`` keeps the calling contract of big-wave's driver
(big-wave/wavesim.py:89-409), because the runner, tests and callers of the
reference use it: the same arguments and return value (one area_v1 image per
phase step), the same order of side effects (source, gap, snapshot, keypoint,
element, keypoint, checkpoint, ..., final leg, checkpoint, detector, saved u,
replayed phase steps) and the same ValueErrors for runs that cannot be
resumed.  Its structure is this package's own: a ``_Beam`` records where the
field is (z, the element it has reached, how many of that element's slices
are done, and whether U holds FFT2(u)), and each stage of the run is one
method that advances it.
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

# A runtime checkpoint is taken after every CHECKPOINT_SLICE_CADENCE-th slice
# of a sample (big-wave/wavesim.py:17-19): u and U are GiB-sized at the large
# grids, so saving at every slice would cost more than the slices.
CHECKPOINT_SLICE_CADENCE = 5

# the final leg is skipped when the field is already this close to the detector
Z_TOLERANCE = 1e-8

logger = logging.getLogger("big-wave")

# attributes the driver lends an element for the duration of one apply()
_LENT = ("_checkpoint_start_slice", "_checkpoint_callback", "_checkpoint_cadence")


@dataclass
class Snapshot:
    """The field in front of the first phase-stepped element, kept so that
    later phase steps restart there (big-wave/wavesim.py:24-37)."""

    z: float
    element_idx: int
    u: DeviceVector
    nr_phase_steps: int


def save_keypoint(save_keypoints_path: Optional[Path], element_idx: int, sub_idx: int,
                  u: DeviceVector) -> None:
    """``keypoint_<element:02>_<0 before, 1 after>.npy`` of u
    (big-wave/wavesim.py:40-62 naming)."""
    if save_keypoints_path is not None:
        np.save(Path(save_keypoints_path) / f"keypoint_{element_idx:02}_{sub_idx}.npy",
                u.to_numpy())


def save_u(u: DeviceVector, phase_step: int, sub_dir: Path, vectors_dir: Path,
           save_final_u_vectors: bool) -> None:
    """``u_<phase:04>.npy`` in the source's directory when asked for
    (big-wave/wavesim.py:65-86 naming)."""
    if save_final_u_vectors:
        np.save(Path(sub_dir) / f"u_{phase_step:04}.npy", u.to_numpy())


@dataclass
class _Beam:
    z: float
    element: int = 0
    slices_done: int = 0        # of ``element``, for a resumed sample
    spectrum_valid: bool = False


class _SourceRun:
    def __init__(self, params: SimParams, source, elements: list, cutoff_angles: list,
                 u, U, history, keypoints: Optional[Path], checkpoint: Optional[RunCheckpoint]):
        self.p = params
        self.source = source
        self.elements = elements
        self.freqs = [float(np.sin(a)) / params.wl for a in cutoff_angles]
        self.u, self.U = u, U
        self.history = history
        self.keypoints = keypoints
        self.checkpoint = checkpoint
        self.snapshot: Optional[Snapshot] = None

    # -- the cutoff in force: index 0 before the first element, i + 1 from
    # the start of element i on --
    def freq_before(self, element: int) -> float:
        return self.freqs[element]

    def freq_inside(self, element: int) -> float:
        return self.freqs[element + 1]

    def _save(self, beam: _Beam, stage: str) -> None:
        if self.checkpoint is not None:
            self.checkpoint.save(self.u, self.U, current_z=beam.z, element_index=beam.element,
                                 slice_index=beam.slices_done,
                                 u_fourier_valid=beam.spectrum_valid, stage=stage)

    # -- stages --
    def from_source(self) -> _Beam:
        first_z = self.elements[0].z_start if self.elements else self.p.z_detector
        self.source.propagate_to(first_z, self.p, self.freq_before(0), self.u, self.U,
                                 self.history)
        beam = _Beam(z=first_z)
        self._save(beam, "source_propagated")
        return beam

    def from_checkpoint(self) -> _Beam:
        if self.checkpoint is None or not self.checkpoint.exists:
            raise ValueError("resume_checkpoint requested but no ready checkpoint exists")
        state = self.checkpoint.load(self.u, self.U)
        if state.phase_step != 0:
            raise ValueError("only phase_step=0 runtime checkpoints are supported")
        logger.info("Resuming checkpoint at z=%sm, element=%s, slice=%s", state.current_z,
                    state.element_index, state.slice_index)
        return _Beam(z=state.current_z, element=state.element_index,
                     slices_done=state.slice_index, spectrum_valid=state.u_fourier_valid)

    def cutoff_after_resume(self, beam: _Beam) -> float:
        # the field is inside an element once one of its slices is done
        idx = beam.element + (1 if beam.slices_done > 0 else 0)
        return self.freqs[min(idx, len(self.freqs) - 1)]

    def through_element(self, beam: _Beam, idx: int, cutoff: float) -> float:
        """Gap up to element idx, then the element; returns the cutoff in force."""
        el = self.elements[idx]
        resuming_inside = beam.slices_done > 0
        if not resuming_inside:
            if beam.z < el.z_start:
                logger.info(f"Propagating from z {beam.z}m to {el.z_start}m")
                gap = self.p.effective_slice_dz(beam.z, el.z_start - beam.z)
                propagate_with_history_2d(self.u, self.U, self.p, gap, cutoff, beam.z, False,
                                          self.history)
                beam.spectrum_valid = True
            beam.z = el.z_start
        if self.snapshot is None and len(el.x_positions) > 1:
            self.snapshot = Snapshot(z=beam.z, element_idx=idx, u=self.u.copy(),
                                     nr_phase_steps=len(el.x_positions))
        cutoff = self.freq_inside(idx)
        save_keypoint(self.keypoints, idx, 0, self.u)
        self._lend(el, idx, beam.slices_done)
        try:
            el.apply(self.u, self.U, self.p, cutoff, 0,
                     self.history[0] if self.history is not None else None)
        finally:
            for name in _LENT:
                if hasattr(el, name):
                    delattr(el, name)
        save_keypoint(self.keypoints, idx, 1, self.u)
        beam.z = el.z_start + el.get_thickness()
        beam.element, beam.slices_done, beam.spectrum_valid = idx + 1, 0, True
        self._save(beam, "element_complete")
        return cutoff

    def _lend(self, el, idx: int, first_slice: int) -> None:
        if first_slice > 0:
            el._checkpoint_start_slice = first_slice
        if self.checkpoint is None:
            return
        u, U, ck = self.u, self.U, self.checkpoint

        def after_slices(done: int, z: float) -> None:
            if done % CHECKPOINT_SLICE_CADENCE == 0:
                ck.save(u, U, current_z=z, element_index=idx, slice_index=done,
                        u_fourier_valid=True, stage="element_slice")

        el._checkpoint_callback = after_slices
        el._checkpoint_cadence = CHECKPOINT_SLICE_CADENCE

    def to_detector(self, beam: _Beam, cutoff: float) -> None:
        if beam.z < self.p.z_detector - Z_TOLERANCE:
            logger.info(f"Propagating from z {beam.z}m to {self.p.z_detector}m")
            # a source leaves U undefined; an element leaves U = FFT2(u)
            propagate_with_history_2d(
                self.u, self.U, self.p, self.p.effective_final_dz(beam.z), cutoff, beam.z,
                skip_fft=beam.spectrum_valid,
                history=None if self.p.use_fresnel_scaling else self.history)
            beam.z, beam.spectrum_valid = self.p.z_detector, True
        beam.element, beam.slices_done = len(self.elements), 0
        self._save(beam, "detector_ready")

    def image(self) -> np.ndarray:
        return square_and_downsample_2d(self.u, self.p, self.p.z_detector)

    def replay_step(self, step: int) -> np.ndarray:
        """Phase step ``step`` > 0: from the snapshot to the detector, without
        history or checkpoints."""
        snap = self.snapshot
        snap.u.copy_to(self.u)
        z, cutoff = snap.z, self.freq_before(snap.element_idx)
        nx, ny, dx, dy, wl = self.p.nx, self.p.ny, self.p.dx, self.p.get_dy(), self.p.wl
        for idx in range(snap.element_idx, len(self.elements)):
            el = self.elements[idx]
            if z < el.z_start:
                propagate_2d(self.u, self.U, dx, dy, wl,
                             self.p.effective_slice_dz(z, el.z_start - z),
                             self.p.chunk_size, cutoff, nx, ny)
            cutoff = self.freq_inside(idx)
            el.apply(self.u, self.U, self.p, cutoff, step if len(el.x_positions) > 1 else 0, None)
            z = el.z_start + el.get_thickness()
        if z < self.p.z_detector - Z_TOLERANCE:
            propagate_2d(self.u, self.U, dx, dy, wl, self.p.effective_final_dz(z),
                         self.p.chunk_size, cutoff, nx, ny)
        return self.image()


def run_simulation(params: SimParams, source, elements: list, cutoff_angles: list[float],
                   u: DeviceVector, U: DeviceVector, deltabeta_table, sub_dir=None,
                   vectors_dir=None, save_final_u_vectors: bool = False,
                   history: Optional[Tuple[History, float]] = None, save_keypoints_path=None,
                   checkpoint: Optional[RunCheckpoint] = None,
                   resume_checkpoint: bool = False) -> list[np.ndarray]:
    """Run one source through ``elements`` to the detector.

    ``cutoff_angles`` has one entry more than ``elements``: entry 0 holds
    before the first element, entry i + 1 from the start of element i on.
    Returns one detector image per phase step of the phase-stepped element
    (one image without phase stepping).  ``checkpoint`` saves u and U after
    the source leg, every CHECKPOINT_SLICE_CADENCE slices, after each element
    and before the detector; ``resume_checkpoint`` continues from the last
    save instead of starting at the source.
    """
    if save_final_u_vectors and sub_dir is None:
        raise ValueError("save_final_u_vectors needs sub_dir")
    keypoints = Path(save_keypoints_path) if save_keypoints_path is not None else None
    for el in elements:
        el.store_deltabetas(deltabeta_table)
    if checkpoint is not None and any(len(el.x_positions) > 1 for el in elements):
        raise ValueError(
            "runtime checkpointing is not supported together with phase stepping; "
            "the phase-step snapshot is kept in memory only")

    run = _SourceRun(params, source, elements, cutoff_angles, u, U, history, keypoints,
                     checkpoint)
    if resume_checkpoint:
        beam = run.from_checkpoint()
        cutoff = run.cutoff_after_resume(beam)
    else:
        beam = run.from_source()
        cutoff = run.freq_before(0)
    for idx in range(beam.element, len(elements)):
        cutoff = run.through_element(beam, idx, cutoff)
    run.to_detector(beam, cutoff)

    images = [run.image()]
    save_u(u, 0, sub_dir, vectors_dir, save_final_u_vectors)
    if run.snapshot is not None:
        try:
            for step in range(1, run.snapshot.nr_phase_steps):
                images.append(run.replay_step(step))
                save_u(u, step, sub_dir, vectors_dir, save_final_u_vectors)
        finally:
            run.snapshot.u.drop()
    return images
