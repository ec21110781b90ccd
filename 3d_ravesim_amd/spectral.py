"""Many (source, energy) instances through one setup into one weighted image.

A polychromatic or extended-source simulation is a set of independent
instances -- one point source at one energy each -- whose detector images
are summed with spectral weights.  The reference runs every instance as its
own simulation (fast-wave/Readme.md:25-27; multisim's per-source
sub-directories, big-wave/multisim.py:258-519) and sums afterwards.  Here
all instances live in HBM at once, each on its own grid (sizes may differ
per energy), and one run is:

    per instance:  point source (big-wave/source.py:97-107: analytic
                   spherical wave, FFT2, circular cutoff, IFFT2)
                   -> for each stage: a voxel sample (Sample._apply_2d,
                   big-wave/optical_element.py:502-596) or free space
                   (propagate_2d, big-wave/propagation.py:396-442)
                   -> the last leg to the detector
                   -> area_v1 (big-wave/propagation.py:995-1130), accumulated
                   with the instance's weight into one float64 image.

Instances of one grid and stage structure run their chains together
(``rsim_chain_apply_2d_batch``: each pass one launch for up to 16 of them);
the detector sum follows grid by grid.  ``run()`` issues it call by call;
``capture()`` records the same calls once into a graph, so that ``launch()``
submits the whole run -- every instance, source to detector sum -- as a
single graph launch.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .graph import CapturedGraph
from .optical_element import DeviceVolume
from .vector import DeviceVector


@dataclass
class SampleStage:
    """n_slices slices of ``volume`` (voxel pitch px, py, slice thickness pz)."""
    volume: DeviceVolume
    pixel_size_z: float
    n_slices: int
    pixel_size_x: float = 0.0     # 0: the instance's dx
    pixel_size_y: float = 0.0     # 0: the x pitch
    x_pos: float = 0.0
    y_pos: float = 0.0


@dataclass
class GapStage:
    """Free space over dz."""
    dz: float


@dataclass
class Instance:
    nx: int
    ny: int
    dx: float
    wl: float
    cutoff_freq: float
    source_x: float
    source_y: float
    source_distance: float        # from the source to the first stage
    stages: list
    weight: float = 1.0
    dy: float = 0.0


@dataclass
class Detector:
    out_x: int
    out_y: int
    pixel_size_x: float
    pixel_size_y: float
    z: float                      # source-to-detector distance (obliquity factor)
    final_dz: float               # free space after the last stage


class SpectralRun:
    """All instances on device fields of ``dtype`` (complex64: fused engine)."""

    # grids below this many points share launches by default: one such field
    # leaves most of the GPU idle per launch (64 x 512^2 Talbot-Lau
    # instances: 5.1 ms batched vs 21.3 ms one chain each), while at 2048^2
    # and above every field fills the GPU on its own and separate launches
    # keep each instance's working set in cache (239 vs 244 ms at 2048^2-6144^2)
    BATCH_BELOW_POINTS = 2048 * 2048

    def __init__(self, instances: list[Instance], detector: Detector, dtype=np.complex64,
                 engine: Optional[int] = None, stream=None, batched: Optional[bool] = None):
        if not instances:
            raise ValueError("a spectral run needs at least one instance")
        self.instances = list(instances)
        self.det = detector
        self.dtype = np.dtype(dtype)
        self.code = _lib.dtype_code(self.dtype)
        if engine is None:
            engine = _lib.RSIM_FUSED if self.dtype == np.complex64 else _lib.RSIM_PLAIN
        self.engine = engine
        self.stream = stream
        # instances of one grid and stage structure share launches
        # (rsim_chain_apply_2d_batch): True for every grid, False for none,
        # None for grids below BATCH_BELOW_POINTS
        self.batched = batched
        self.fields = [(DeviceVector(q.nx * q.ny, self.dtype), DeviceVector(q.nx * q.ny, self.dtype))
                       for q in self.instances]
        self.image_bytes = detector.out_x * detector.out_y * 8
        p = C.c_void_p()
        _lib.call("rsim_malloc", C.byref(p), self.image_bytes)
        self.image_ptr = p.value
        self.graph: Optional[CapturedGraph] = None

    # -- the run as library calls --
    def _steps(self, q: Instance):
        """The instance's stages and last leg as rsim_chain_step records."""
        rows = []
        for st in q.stages:
            if isinstance(st, GapStage):
                rows.append(_lib.ChainStep(None, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0, st.dz))
            else:
                px = st.pixel_size_x or q.dx
                rows.append(_lib.ChainStep(st.volume.handle, px, st.pixel_size_y or px,
                                           st.pixel_size_z, st.x_pos, st.y_pos, 0, st.n_slices,
                                           0.0))
        if self.det.final_dz > 0:
            rows.append(_lib.ChainStep(None, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0, self.det.final_dz))
        return (_lib.ChainStep * max(len(rows), 1))(*rows), len(rows)

    def _instance(self, q: Instance, u: DeviceVector, U: DeviceVector) -> None:
        s, code = self.stream, self.code
        dy = q.dy or q.dx
        _lib.call("rsim_propagate_analytically_2d", u.ptr, q.nx, q.ny, code, q.dx, dy,
                  q.source_distance, q.source_x, q.source_y, q.wl, s)
        # the source's band limit, the stages and the last leg: one fused
        # chain on the complex64 engine, call by call otherwise
        steps, n = self._steps(q)
        _lib.call("rsim_chain_apply_2d", u.ptr, U.ptr, q.nx, q.ny, code, q.dx, dy, q.wl,
                  q.cutoff_freq, 1, steps, n, self.engine, _lib.RSIM_FLAG_NO_SPECTRUM, s)
        d = self.det
        _lib.call("rsim_detector_area_v1_accumulate", u.ptr, q.nx, q.ny, code, q.dx, dy, d.out_x,
                  d.out_y, d.pixel_size_x, d.pixel_size_y, d.z, float(q.weight),
                  self.image_ptr, s)

    def _groups(self):
        """Instances that can share launches: one grid, the same stage kinds
        and slice counts (their own volumes, distances, energies, cutoffs)."""
        groups = {}
        for k, q in enumerate(self.instances):
            sig = (q.nx, q.ny, tuple(("gap",) if isinstance(st, GapStage) else ("vol", st.n_slices)
                                     for st in q.stages))
            groups.setdefault(sig, []).append(k)
        return list(groups.values())

    def _shares(self, q: Instance) -> bool:
        if self.batched is None:
            return q.nx * q.ny < self.BATCH_BELOW_POINTS
        return bool(self.batched)

    def _issue(self) -> None:
        _lib.call("rsim_fill_zero", self.image_ptr, self.image_bytes, self.stream)
        s, code, d = self.stream, self.code, self.det
        for idx in self._groups():
            if len(idx) == 1 or not self._shares(self.instances[idx[0]]):
                for k in idx:
                    self._instance(self.instances[k], *self.fields[k])
                continue
            keep = []
            rows = []
            for k in idx:
                q = self.instances[k]
                u, U = self.fields[k]
                dy = q.dy or q.dx
                _lib.call("rsim_propagate_analytically_2d", u.ptr, q.nx, q.ny, code, q.dx, dy,
                          q.source_distance, q.source_x, q.source_y, q.wl, s)
                steps, n = self._steps(q)
                keep.append(steps)
                rows.append(_lib.ChainInstance(u.ptr, U.ptr, q.dx, dy, q.wl, q.cutoff_freq, steps))
            q0 = self.instances[idx[0]]
            arr = (_lib.ChainInstance * len(rows))(*rows)
            _lib.call("rsim_chain_apply_2d_batch", arr, len(rows), q0.nx, q0.ny, code, 1, n,
                      self.engine, _lib.RSIM_FLAG_NO_SPECTRUM, s)
            for k in idx:
                q = self.instances[k]
                u, _ = self.fields[k]
                _lib.call("rsim_detector_area_v1_accumulate", u.ptr, q.nx, q.ny, code, q.dx,
                          q.dy or q.dx, d.out_x, d.out_y, d.pixel_size_x, d.pixel_size_y, d.z,
                          float(q.weight), self.image_ptr, s)

    def run(self) -> None:
        """Issue the whole run call by call (stream-ordered, asynchronous)."""
        self._issue()

    def capture(self) -> CapturedGraph:
        """Record the run as one graph (runs it once first, eagerly, so that
        every table and launch shape it needs exists)."""
        self._issue()
        _lib.call("rsim_synchronize", self.stream)
        with CapturedGraph.record(self.stream) as rec:
            self._issue()
        if self.graph is not None:
            self.graph.close()
        self.graph = rec.graph
        return self.graph

    def launch(self) -> None:
        if self.graph is None:
            raise ValueError("capture() the run first")
        self.graph.launch(self.stream)

    def synchronize(self) -> None:
        _lib.call("rsim_synchronize", self.stream)

    def image(self) -> np.ndarray:
        out = np.empty((self.det.out_y, self.det.out_x), dtype=np.float64)
        _lib.call("rsim_memcpy_d2h", out.ctypes.data, self.image_ptr, out.nbytes, self.stream)
        return out

    def close(self) -> None:
        if self.graph is not None:
            self.graph.close()
            self.graph = None
        for u, U in self.fields:
            u.drop()
            U.drop()
        self.fields = []
        if self.image_ptr is not None:
            _lib.call("rsim_free", self.image_ptr)
            self.image_ptr = None
