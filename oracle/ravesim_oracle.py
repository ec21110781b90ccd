"""CPU FP64 oracle for RAVE-SIM's 2D Fresnel multislice path.

TEST INFRASTRUCTURE ONLY.  This module restates, in numpy/scipy complex128,
the arithmetic of big-wave's in-memory (``NumpyVector``) 2D path so that the
HIP engine can be checked against it.  It is imported only by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``; the
product package never imports it and has no CPU fallback.

Every function cites the reference file:line whose semantics it follows
(paths relative to the reference checkout).  The restatement is written from
the reference's documented behaviour, not copied: operation order is kept
where it changes rounding (phase constants, interpolation order), because the
parity tests compare at 1e-12..2e-14 relative.

Pinning (what checks each function against something other than itself):

* ``fft2`` / ``ifft2``          -- numpy-derived vectors from
  fast-wave/test/fwtest.cu:89-334 (tests/golden/fwtest_fft_vectors.json) and
  brute-force DFT on tiny sizes.
* ``frequency_squared`` / ``apply_frequency_cutoff_2d`` -- fftfreq N=10 vector
  (fwtest.cu:59-69) and the axis/diagonal boundary cells of
  tests/big_wave_2d/test_p5.py:126-140.
* ``propagate_2d``              -- plane and tilted plane wave closed form
  (test_p5.py:84-100), Parseval power conservation, rectangular aperture vs
  Fresnel integrals (test_p5.py:142-172).
* ``propagate_analytically_2d`` -- 1/r amplitude, circular phase, peak position
  (test_p5.py:64-82).
* ``material_factor`` / ``sample_deltabeta_slice`` / ``sample_apply_2d`` --
  uniform slab phase + Beer-Lambert (test_p5.py:102-124), offset slab vs
  explicit transmission (test_p5.py:174-213), per-pixel brute-force bilinear
  map (test_p1.py:283-317), hollow-capsule x/y symmetry (test_p5.py:215-256).
* ``area_v1``                   -- brute-force overlap loops
  (tests/big_wave_2d/test_p3.py:85-123) and uniform-field area preservation
  (test_p3.py:222-243).
* ``area_v1_matrices`` / ``area_v1_adjoint`` -- the dense form reproduces
  ``area_v1`` (itself pinned by brute force); the adjoint passes the
  dot-product test <gI, A x> = <A^T gI, x>.
* ``multislice_db_2d`` / ``multislice_db_intensity`` -- equals
  ``sample_apply_2d`` on an integral-pitch id volume carrying the same
  delta/beta per voxel.
* ``multislice_db_gradient``    -- central finite differences of
  ``multislice_db_intensity`` (FP64) and complex128 torch autograd of an
  independent torch forward, on small volumes.
* ``FresnelGeometry``           -- the reference's geometry known answers
  (tests/fresnel_scaling/test_cone_beam_geometry.py:32-66: magnification,
  local scale, effective slice and final distances, partition sum = z_eff).
* ``fresnel_plane_source_2d``   -- centred source = constant 1/dz, off-axis
  linear phase (tests/fresnel_scaling/test_big_wave_source.py:46-73).
* ``sample_apply_2d`` (CB-BPM)  -- sampled voxel coordinates at m = 1.1 and
  the slice distance 0.2/1.2 (test_big_wave_cone_beam_sample.py:21-76).
* ``plasma_delta_beta_grid_multi`` -- delta_free = 1.077221e-5 at n_e = 1e24
  cm^-3, 8 keV, and the Kramers beta_ff of fully / partially ionised Al
  (7.61e-12 / 4.68e-12; tests/plasma_sample/TEST_REPORT.md:76, 100-106);
  scaling laws of beta_ff (n_e n_i Z*^2 T_e^-1/2 E^-4), q clipping, vacuum.
* ``plasma_sample_apply_2d``     -- uniform-slab phase pi and 2 pi and
  Beer-Lambert modulus by the reference's gold-standard ratio method
  (tests/plasma_sample/test_plasma_phase_absorption.py:60-101, 187-250);
  equal to ``sample_apply_2d`` on a material-id volume with the same
  per-voxel (delta, beta).
* ``run_point_source_sample``   -- the reference's own end-to-end golden
  fingerprints tests/big_wave_2d/p5_results/{vacuum,sample}_convergence_4096.json
  (inputs in tests/big_wave_2d/p5_benchmark.py:69-185).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy import fft as spfft

# big-wave/propagation.py:199-201
PLANCK_H = 6.62607004 * 10 ** (-34)
SPEED_OF_LIGHT = 299792458
EV_TO_JOULE = 1.602176634 * 10 ** (-19)


def convert_energy_wavelength(energy_or_wavelength: float) -> float:
    """big-wave/propagation.py:260-262."""
    return PLANCK_H * SPEED_OF_LIGHT / (energy_or_wavelength * EV_TO_JOULE)


# ---------------------------------------------------------------------------
# FFT conventions: big-wave/vector.py:341-347 (scipy.fft.fft2/ifft2 on the
# (ny, nx) view; forward unnormalised, inverse 1/(nx*ny)).
# ---------------------------------------------------------------------------

def fft2(field2d: np.ndarray, workers: int = 1) -> np.ndarray:
    return spfft.fft2(field2d, workers=workers)


def ifft2(field2d: np.ndarray, workers: int = 1) -> np.ndarray:
    return spfft.ifft2(field2d, workers=workers)


def fftfreq(n: int) -> np.ndarray:
    """numpy.fft.fftfreq(n) with d=1: integer bins times 1/n."""
    val = 1.0 / n
    bins = np.empty(n, dtype=np.int64)
    half = (n - 1) // 2 + 1
    bins[:half] = np.arange(0, half)
    bins[half:] = np.arange(-(n // 2), 0)
    return bins * val


def frequency_squared(n: int, spacing: float) -> np.ndarray:
    """big-wave/propagation.py:293-295: (fftfreq(n)/spacing)**2."""
    f = fftfreq(n) / spacing
    return f * f


def apply_frequency_cutoff_2d(U2d: np.ndarray, freq: float, dx: float, dy: float) -> np.ndarray:
    """big-wave/propagation.py:298-326: zero where kx^2+ky^2 > 2 f^2 (boundary kept)."""
    ny, nx = U2d.shape
    freq_sq = 2.0 * freq ** 2
    kx2 = frequency_squared(nx, dx)
    ky2 = frequency_squared(ny, dy)
    mask = ky2[:, None] + kx2[None, :] > freq_sq
    out = U2d.copy()
    out[mask] = 0.0
    return out


def chirp_2d(nx: int, ny: int, dx: float, dy: float, wl: float, dz: float) -> np.ndarray:
    """Transfer function of big-wave/propagation.py:427-435 (no cutoff).

    H = exp(2*pi*i*dz/wl) * exp(-i*pi*wl*dz*(kx^2+ky^2)), with the same
    expression order as the reference so the FP64 phase rounds identically.
    """
    kx2 = frequency_squared(nx, dx)
    ky2 = frequency_squared(ny, dy)
    global_phase = np.exp(1j * 2 * np.pi / wl * dz)
    phase_scale = -1j * np.pi * wl * dz
    k_sq = ky2[:, None] + kx2[None, :]
    return global_phase * np.exp(phase_scale * k_sq)


def propagate_2d(u2d: np.ndarray, dx: float, dy: float, wl: float, dz: float,
                 cutoff_frequency: float, U2d: np.ndarray | None = None,
                 workers: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """big-wave/propagation.py:396-442.

    Returns (u_new, U_new) where U_new = FFT2(u_new).  When ``U2d`` is given it
    is taken as FFT2(u2d) and the forward transform is skipped (``skip_fft``).
    The result keeps the input dtype, as NumpyVector does (vector.py:341-347):
    complex64 inputs are transformed in single precision by scipy.
    """
    ny, nx = u2d.shape
    U = fft2(u2d, workers) if U2d is None else U2d.copy()
    H = chirp_2d(nx, ny, dx, dy, wl, dz)
    # In-place `tile *= global_phase * exp(...)` on a complex64 tile computes in
    # complex128 and casts back (propagation.py:433-435).
    U = (U * H).astype(U.dtype, copy=False)
    U = apply_frequency_cutoff_2d(U, cutoff_frequency, dx, dy)
    return ifft2(U, workers), U


# ---------------------------------------------------------------------------
# Point source: big-wave/propagation.py:458-484 and source.py:50-113
# ---------------------------------------------------------------------------

def grid_density_check(dz: float, x_source: float, n: int, dx: float, wl: float) -> bool:
    """big-wave/propagation.py:204-225 (returns the predicate instead of raising)."""
    def dist_x(idx):
        return np.sqrt(((n / 2 - idx) * dx + np.abs(x_source)) ** 2 + dz ** 2)
    return bool(np.abs(dist_x(0) - dist_x(1)) <= wl / 2)


def propagate_analytically_2d(nx: int, ny: int, dx: float, dy: float, z: float,
                              x_source: float, y_source: float, wl: float) -> np.ndarray:
    """u = exp(2*pi*i*r/wl)/r, r = sqrt(y^2 + x^2 + z^2) (propagation.py:474-482)."""
    x = (np.arange(nx, dtype=np.float64) - nx / 2.0) * dx - x_source
    y = (np.arange(ny, dtype=np.float64) - ny / 2.0) * dy - y_source
    x2 = np.square(x)
    y2 = np.square(y)
    wave_number = 2j * np.pi / wl
    r = np.sqrt(y2[:, None] + x2[None, :] + z ** 2)
    return np.exp(r * wave_number) / r


def point_source_propagate_to(nx, ny, dx, dy, dz, xs, ys, wl, cutoff_freq,
                              dtype=np.complex128, check_density=True, workers=1):
    """PointSource.propagate_to, 2D non-Fresnel branch (source.py:97-107).

    Analytic init, FFT2, circular cutoff, IFFT2.  The spectrum is NOT returned:
    the reference does not guarantee U = FFT(u) after a source (source.py:38-40).
    """
    if check_density:
        if not (grid_density_check(dz, xs, nx, dx, wl) and grid_density_check(dz, ys, ny, dy, wl)):
            raise ValueError("Nyquist not satisfied, dx too large")
    u = propagate_analytically_2d(nx, ny, dx, dy, dz, xs, ys, wl).astype(dtype)
    U = apply_frequency_cutoff_2d(fft2(u, workers), cutoff_freq, dx, dy)
    return ifft2(U, workers)


# ---------------------------------------------------------------------------
# Sample transmission: big-wave/optical_element.py:39-42, 383-596
# ---------------------------------------------------------------------------

def material_factor(deltabeta, thickness: float, wl: float):
    """exp(-2*pi*i*t/wl*conj(db)) (optical_element.py:39-42)."""
    return np.exp(-2j * np.pi * thickness / wl * np.conj(deltabeta))


def axis_map(n: int, spacing: float, scale: float, pos: float, grid_len: int, pixel: float):
    """Field axis -> voxel axis mapping (optical_element.py:538-546, 555-563).

    Returns (floor_clipped, frac, inside) for each field index.
    """
    sim = (np.arange(n, dtype=np.float64) - n / 2.0) * spacing
    idx = (sim * scale + pos + grid_len * pixel * 0.5) / pixel
    floor_raw = np.floor(idx).astype(np.int64)
    frac = idx - floor_raw
    inside = (idx >= 0.0) & (idx < float(grid_len) - 1.0)
    floor_c = np.clip(floor_raw, 0, grid_len - 2)
    return floor_c, frac, inside


def sample_deltabeta_slice(grid: np.ndarray, db_list: np.ndarray, rowidx: int,
                           nx: int, ny: int, dx: float, dy: float,
                           px: float, py: float, x_pos: float, y_pos: float,
                           scale: float = 1.0) -> np.ndarray:
    """Bilinear delta+i*beta of one z-slice on the field grid, 0 outside.

    Follows optical_element.py:566-579: gather the four corner ids, interpolate
    in delta-beta space along x then y, zero outside the voxel box.
    """
    _, y_len, x_len = grid.shape
    x0, xf, inside_x = axis_map(nx, dx, scale, x_pos, x_len, px)
    y0, yf, inside_y = axis_map(ny, dy, scale, y_pos, y_len, py)
    g = grid[rowidx]
    X0 = x0[None, :]
    Y0 = y0[:, None]
    db00 = db_list[g[Y0, X0]]
    db01 = db_list[g[Y0, X0 + 1]]
    db10 = db_list[g[Y0 + 1, X0]]
    db11 = db_list[g[Y0 + 1, X0 + 1]]
    db_y0 = db00 * (1.0 - xf[None, :]) + db01 * xf[None, :]
    db_y1 = db10 * (1.0 - xf[None, :]) + db11 * xf[None, :]
    interp = db_y0 * (1.0 - yf[:, None]) + db_y1 * yf[:, None]
    interp[~(inside_y[:, None] & inside_x[None, :])] = 0.0
    return interp


@dataclass
class SampleSpec:
    """The fields of big-wave's Sample (optical_element.py:383-425) the 2D path reads."""
    grid: np.ndarray            # (nz, ny_g, nx_g) integer material ids, 0 = vacuum
    db_list: np.ndarray         # complex128, db_list[0] = 0 (store_deltabetas, :433-440)
    pixel_size_x: float
    pixel_size_z: float
    pixel_size_y: float = 0.0
    x_pos: float = 0.0
    y_pos: float = 0.0
    z_start: float = 0.0

    @property
    def py(self) -> float:
        return self.pixel_size_y if self.pixel_size_y != 0.0 else self.pixel_size_x


@dataclass
class FresnelGeometry:
    """Fresnel-scaling / cone-beam BPM geometry of SimParams
    (big-wave/propagation.py:114-196): thin mode (cone_beam False) or CB-BPM."""
    z_detector: float
    cone_beam: bool = False
    source_z: float = 0.0
    reference_z: float = 0.0
    magnification: float = 1.0
    effective_z: float = 0.0

    def configure(self, z_source: float, z_sample: float) -> "FresnelGeometry":
        """configure_fresnel_detector (propagation.py:114-135)."""
        a = z_sample - z_source
        b = self.z_detector - z_sample
        if a <= 0 or b <= 0:
            raise ValueError("Fresnel scaling requires z_source < z_sample < z_detector")
        self.source_z, self.reference_z = z_source, z_sample
        self.effective_z = a * b / (a + b)
        self.magnification = (a + b) / a
        return self

    def transverse_scale(self, z: float) -> float:
        """m(z) = (z - z_source) / (z_reference - z_source) under CB-BPM, else 1."""
        if not self.cone_beam:
            return 1.0
        return (z - self.source_z) / (self.reference_z - self.source_z)

    def slice_dz(self, z: float, dz: float) -> float:
        """effective_slice_dz: dz / (m(z) m(z + dz)) under CB-BPM."""
        if not self.cone_beam:
            return dz
        return dz / (self.transverse_scale(z) * self.transverse_scale(z + dz))

    def final_dz(self, current_z: float) -> float:
        """effective_final_dz (propagation.py:158-170)."""
        physical = self.z_detector - current_z
        if self.cone_beam:
            return physical / (self.transverse_scale(current_z) * self.magnification)
        return self.effective_z

    def detector(self, ds_x: float, ds_y: float) -> tuple[float, float, float, float]:
        """effective_detector_geometry at the detector: (ds_x/M, ds_y/M, z_eff, M)."""
        M = self.magnification
        return ds_x / M, ds_y / M, self.effective_z, M


def fresnel_plane_source_2d(nx: int, ny: int, dx: float, dy: float, wl: float, dz: float,
                            xs: float, ys: float, cutoff_freq: float,
                            dtype=np.complex128) -> np.ndarray:
    """PointSource.propagate_to under Fresnel scaling (source.py:62-90): the
    equivalent tilted plane wave exp(-i k (x xs + y ys) / dz) / dz, then FFT,
    cutoff, IFFT."""
    k = 2.0 * np.pi / wl
    x = (np.arange(nx, dtype=np.float64) - nx / 2.0) * dx
    y = (np.arange(ny, dtype=np.float64) - ny / 2.0) * dy
    phase = -k * (x[None, :] * xs + y[:, None] * ys) / dz
    u = (np.exp(1j * phase) / dz).astype(dtype)
    U = apply_frequency_cutoff_2d(fft2(u), cutoff_freq, dx, dy)
    return ifft2(U)


def sample_apply_2d(u2d: np.ndarray, sample: SampleSpec, dx: float, dy: float, wl: float,
                    cutoff_freq: float, first_slice: int = 0, n_slices: int | None = None,
                    workers: int = 1, geometry: FresnelGeometry | None = None,
                    history: list | None = None, frame=None
                    ) -> tuple[np.ndarray, np.ndarray]:
    """Sample._apply_2d (optical_element.py:502-596).

    Per slice: u *= material_factor(interpolated db, pz, wl), the voxel map
    scaled by m(slice_z + pz/2) under CB-BPM; then propagate_2d over
    effective_slice_dz(slice_z, pz).  Returns (u, U) with U = FFT2(u) (the
    apply contract, optical_element.py:97).  With ``history`` (a list) and
    ``frame`` = (size_x, size_y, ds_x, ds_y), the area_v1 image of the field
    entering each slice at z = slice_z is appended with its z first
    (optical_element.py:548-551).
    """
    ny, nx = u2d.shape
    nz = sample.grid.shape[0]
    last = nz if n_slices is None else min(nz, first_slice + n_slices)
    u = u2d.copy()
    U = None
    pz = sample.pixel_size_z
    for rowidx in range(first_slice, last):
        slice_z = sample.z_start + rowidx * pz
        scale = geometry.transverse_scale(slice_z + 0.5 * pz) if geometry else 1.0
        dz = geometry.slice_dz(slice_z, pz) if geometry else pz
        if history is not None:
            history.append((area_v1(u, dx, dy, *frame, slice_z), slice_z))
        db = sample_deltabeta_slice(sample.grid, sample.db_list, rowidx, nx, ny, dx, dy,
                                    sample.pixel_size_x, sample.py, sample.x_pos, sample.y_pos,
                                    scale)
        t = material_factor(db, pz, wl)
        u = (u * t).astype(u.dtype, copy=False)
        u, U = propagate_2d(u, dx, dy, wl, dz, cutoff_freq, workers=workers)
    return u, U


# ---------------------------------------------------------------------------
# PlasmaSample: big-wave/plasma.py:380-434, big-wave/plasma_sample.py:126-233
# ---------------------------------------------------------------------------

# nist_lookup/nist_lookup/physical_constants.py:20-27 under the reference's
# pinned scipy~=1.10 (requirements.txt:4), i.e. CODATA 2018.
PLANCK_HC = 1.e10 * 6.62607015e-34 * 299792458.0 / 1.602176634e-19
R_ELECTRON_CM = 100.0 * 2.8179403262e-15


def plasma_delta_beta_grid_multi(ne, ni, te, zstar, f1bar, f2eff_bar, zbar, z2bar, energy):
    """(delta, beta) per voxel (plasma.py:380-434): free electrons, the bound
    fraction (1 - q) of the Chantler terms, q = clip(Z*/zbar, 0, 1), and
    Kramers inverse bremsstrahlung with mean-square charge q^2 z2bar."""
    ne = np.asarray(ne, dtype=np.float64)
    ni = np.asarray(ni, dtype=np.float64)
    te = np.asarray(te, dtype=np.float64)
    zstar = np.asarray(zstar, dtype=np.float64)
    if energy <= 0.0:
        raise ValueError(f"energy must be positive, got {energy}")
    lamb_cm = 1.0e-8 * PLANCK_HC / energy
    prefactor = R_ELECTRON_CM * lamb_cm * lamb_cm / (2.0 * np.pi)
    delta = ne * prefactor
    beta = np.zeros(ne.shape, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(zbar > 0.0, zstar / zbar, 0.0)
    q = np.clip(q, 0.0, 1.0)
    fraction_bound = 1.0 - q
    bound = (fraction_bound > 0.0) & (ni > 0.0)
    bound_scale = ni * prefactor * fraction_bound
    delta = delta + np.where(bound, bound_scale * f1bar, 0.0)
    beta = beta + np.where(bound, bound_scale * f2eff_bar, 0.0)
    ff = (ne > 0.0) & (ni > 0.0) & (q > 0.0) & (z2bar > 0.0) & (te > 0.0)
    nu = energy / 4.135667e-15
    alpha = np.zeros(ne.shape, dtype=np.float64)
    alpha[ff] = (3.7e8 * ne[ff] * ni[ff] * (np.square(q[ff]) * z2bar)
                 / (np.sqrt(te[ff] * 11604.5) * nu ** 3))
    beta = beta + alpha * lamb_cm / (4.0 * np.pi)
    return delta, beta


@dataclass
class PlasmaSpec:
    """The fields of big-wave's PlasmaSample (plasma_sample.py:40-66) the 2D path reads."""
    ne: np.ndarray              # (nz, ny_g, nx_g) [cm^-3]
    ni: np.ndarray
    te: np.ndarray              # [eV]
    zstar: np.ndarray
    optics: tuple               # (f1bar, f2eff_bar, zbar, z2bar) at the run's energy
    pixel_size_x: float
    pixel_size_z: float
    pixel_size_y: float = 0.0
    x_pos: float = 0.0
    y_pos: float = 0.0
    z_start: float = 0.0

    @property
    def py(self) -> float:
        return self.pixel_size_y if self.pixel_size_y != 0.0 else self.pixel_size_x


def plasma_deltabeta_slice(spec: PlasmaSpec, rowidx: int, nx: int, ny: int, dx: float, dy: float,
                           energy: float, scale: float = 1.0) -> np.ndarray:
    """Bilinear delta+i*beta of one plasma slice on the field grid, 0 outside
    (plasma_sample.py:160-217): the four corners' plasma optics, then the
    Sample interpolation."""
    _, y_len, x_len = spec.ne.shape
    x0, xf, inside_x = axis_map(nx, dx, scale, spec.x_pos, x_len, spec.pixel_size_x)
    y0, yf, inside_y = axis_map(ny, dy, scale, spec.y_pos, y_len, spec.py)
    X0 = x0[None, :]
    Y0 = y0[:, None]

    def corner(oy, ox):
        idx = (rowidx, Y0 + oy, X0 + ox)
        d, b = plasma_delta_beta_grid_multi(spec.ne[idx], spec.ni[idx], spec.te[idx],
                                            spec.zstar[idx], *spec.optics, energy)
        return d + 1j * b

    db00, db01, db10, db11 = corner(0, 0), corner(0, 1), corner(1, 0), corner(1, 1)
    db_y0 = db00 * (1.0 - xf[None, :]) + db01 * xf[None, :]
    db_y1 = db10 * (1.0 - xf[None, :]) + db11 * xf[None, :]
    interp = db_y0 * (1.0 - yf[:, None]) + db_y1 * yf[:, None]
    interp[~(inside_y[:, None] & inside_x[None, :])] = 0.0
    return interp


def plasma_sample_apply_2d(u2d: np.ndarray, spec: PlasmaSpec, dx: float, dy: float, wl: float,
                           cutoff_freq: float, first_slice: int = 0, n_slices: int | None = None,
                           workers: int = 1, geometry: FresnelGeometry | None = None,
                           history: list | None = None, frame=None
                           ) -> tuple[np.ndarray, np.ndarray]:
    """PlasmaSample._apply_2d (plasma_sample.py:126-233): per slice u *=
    exp(-2 pi i pz/wl conj(db)) with db from the plasma optics at
    energy = convert_energy_wavelength(wl), then propagate_2d."""
    energy = convert_energy_wavelength(wl)
    ny, nx = u2d.shape
    nz = spec.ne.shape[0]
    last = nz if n_slices is None else min(nz, first_slice + n_slices)
    u = u2d.copy()
    U = None
    pz = spec.pixel_size_z
    for rowidx in range(first_slice, last):
        slice_z = spec.z_start + rowidx * pz
        scale = geometry.transverse_scale(slice_z + 0.5 * pz) if geometry else 1.0
        dz = geometry.slice_dz(slice_z, pz) if geometry else pz
        if history is not None:   # plasma_sample.py:172-176
            history.append((area_v1(u, dx, dy, *frame, slice_z), slice_z))
        db = plasma_deltabeta_slice(spec, rowidx, nx, ny, dx, dy, energy, scale)
        u = (u * material_factor(db, pz, wl)).astype(u.dtype, copy=False)
        u, U = propagate_2d(u, dx, dy, wl, dz, cutoff_freq, workers=workers)
    return u, U


# ---------------------------------------------------------------------------
# Detector: big-wave/propagation.py:764-816 (_detector_geometry),
# 949-977 (_area_axis_plan), 980-992 (_prefix_integral_at), 995-1110 (_area_v1_stream)
# ---------------------------------------------------------------------------

def detector_shape(size_x: float, size_y: float, ds_x: float, ds_y: float) -> tuple[int, int]:
    """(out_y, out_x) = (int(size_y // ds_y), int(size_x // ds_x)) (propagation.py:793-794)."""
    return int(size_y // ds_y), int(size_x // ds_x)


def area_axis_plan(positions: np.ndarray, ds: float, n: int, spacing: float):
    """Clipped overlap plan of one detector axis (propagation.py:949-977)."""
    grid_left = (-(n // 2) - 0.5) * spacing
    grid_right = grid_left + n * spacing
    lo = positions - ds * 0.5
    hi = positions + ds * 0.5
    clo = np.clip(lo, grid_left, grid_right)
    chi = np.maximum(np.clip(hi, grid_left, grid_right), clo)
    slo = (clo - grid_left) / spacing
    shi = (chi - grid_left) / spacing
    ilo = np.floor(slo).astype(np.int64)
    ihi = np.floor(shi).astype(np.int64)
    flo = slo - ilo
    fhi = shi - ihi
    rlo = ilo >= n
    rhi = ihi >= n
    ilo[rlo] = n
    ihi[rhi] = n
    flo[rlo] = 0.0
    fhi[rhi] = 0.0
    return ilo, flo, ihi, fhi, lo, hi


def area_v1(u2d: np.ndarray, dx: float, dy: float, size_x: float, size_y: float,
            ds_x: float, ds_y: float, z: float, out_shape=None) -> np.ndarray:
    """Area-weighted |u|^2 detector, FP64 (propagation.py:995-1110).

    I[q,p] = sum_cells |u|^2 * overlap_x * overlap_y * z/sqrt(x_p^2+y_q^2+z^2).
    The x overlap is integrated with a row prefix sum; the y overlap with the
    explicit min/max interval formula, as the reference does.  ``out_shape``
    (out_y, out_x) overrides the shape derived from size / ds: under Fresnel
    scaling the shape comes from the physical pixels while ds and z are the
    effective ones (propagation.py:791-794).
    """
    ny, nx = u2d.shape
    out_y, out_x = (out_shape if out_shape is not None
                    else detector_shape(size_x, size_y, ds_x, ds_y))
    if out_x <= 0 or out_y <= 0:
        raise ValueError("2D detector output must be non-empty")
    x_det = (np.arange(out_x, dtype=np.float64) - out_x // 2) * ds_x
    y_det = (np.arange(out_y, dtype=np.float64) - out_y // 2) * ds_y
    xi0, xf0, xi1, xf1, _, _ = area_axis_plan(x_det, ds_x, nx, dx)
    _, _, _, _, y_lo, y_hi = area_axis_plan(y_det, ds_y, ny, dy)
    grid_y_left = (-(ny // 2) - 0.5) * dy
    if u2d.dtype == np.complex64:
        inten = (u2d.real * u2d.real + u2d.imag * u2d.imag).astype(np.float64)
    else:
        inten = u2d.real * u2d.real + u2d.imag * u2d.imag
    prefix = np.zeros((ny, nx + 1), dtype=np.float64)
    np.cumsum(inten, axis=1, out=prefix[:, 1:])

    def at(idx, frac):
        res = prefix[:, idx] * dx
        interior = idx < nx
        res[:, interior] += frac[interior] * dx * inten[:, idx[interior]]
        return res

    x_int = at(xi1, xf1) - at(xi0, xf0)          # (ny, out_x)
    cell_lo = grid_y_left + np.arange(ny) * dy
    cell_hi = cell_lo + dy
    overlap = np.minimum(cell_hi[None, :], y_hi[:, None]) - np.maximum(cell_lo[None, :], y_lo[:, None])
    overlap = np.where(overlap > 0, overlap, 0.0)  # (out_y, ny)
    cosine = z / np.sqrt(np.square(x_det)[None, :] + np.square(y_det)[:, None] + z * z)
    return (overlap @ x_int) * cosine


# ---------------------------------------------------------------------------
# History: big-wave/history.py:348-435
# ---------------------------------------------------------------------------

def history_z_steps(dz: float, hist_dz: float) -> np.ndarray:
    """_history_z_steps (history.py:370-378): 0, hist_dz, 2 hist_dz, ... with
    the last step stretched to dz when less than half a step remains."""
    zs = np.arange(0, dz, hist_dz)
    if len(zs) == 0:
        return np.array([0.0, dz])
    if dz - zs[-1] < hist_dz / 2.0 and len(zs) > 1:
        zs[-1] = dz
    else:
        zs = np.append(zs, [dz])
    return zs


def propagate_with_history_2d(u2d, dx, dy, wl, dz, cutoff_freq, current_z, hist_dz, frame,
                              U2d=None, workers=1):
    """propagate_with_history_2d (history.py:406-435): every frame is a fresh
    propagation of the start spectrum over zs[i] - zs[0]; returns
    (u, U, [(frame, z), ...])."""
    zs = history_z_steps(dz, hist_dz)
    U0 = fft2(u2d, workers=workers) if U2d is None else U2d
    out = []
    u = U = None
    for i in range(1, len(zs)):
        step = float(zs[i] - zs[0])
        u, U = propagate_2d(u2d, dx, dy, wl, step, cutoff_freq, U2d=U0, workers=workers)
        z = float(zs[i] + current_z)
        out.append((area_v1(u, dx, dy, *frame, z), z))
    return u, U, out


def analytic_history_2d(zs, x_source, y_source, size_x, size_y, ds_x, ds_y, wl, cutoff_freq):
    """generate_analytic_history_2d (history.py:348-367): ds_x ds_y z / r^3
    inside the cutoff cone, 0 outside."""
    gradient = cutoff_freq * wl
    nxp, nyp = int(size_x // ds_x), int(size_y // ds_y)
    x = (np.arange(nxp) - nxp // 2) * ds_x - x_source
    y = (np.arange(nyp) - nyp // 2) * ds_y - y_source
    out = []
    for z in zs:
        r2 = x[None, :] ** 2 + y[:, None] ** 2 + z * z
        img = ds_x * ds_y * z / np.power(r2, 1.5)
        img[:, np.abs(x) > gradient * z] = 0
        img[np.abs(y) > gradient * z, :] = 0
        out.append((img, z))
    return out


# ---------------------------------------------------------------------------
# Driver ordering: big-wave/wavesim.py:89-317 (single source, no phase stepping)
# ---------------------------------------------------------------------------

def run_point_source_sample(nx, ny, dx, dy, wl, source_xyz, cutoff_freq, z_detector,
                            det_size_x, det_size_y, ds_x, ds_y, sample: SampleSpec | None = None,
                            dtype=np.complex128, workers=1, return_field=False,
                            geometry: FresnelGeometry | None = None):
    """Point source -> [Sample] -> final leg (skip_fft) -> area_v1 detector.

    Mirrors wavesim.run_simulation for one 2D Sample and one cutoff frequency
    (wavesim.py:175-186, 257-260, 279-303, 314-315).  With a configured
    ``geometry`` (Fresnel scaling, thin or CB-BPM): the tilted-plane source
    (source.py:62-90), scaled slices, effective final distance and the
    effective detector geometry at the detector plane (propagation.py:172-196).
    """
    xs, ys, zs = source_xyz
    current_z = sample.z_start if sample is not None else z_detector
    if geometry is not None:
        u = fresnel_plane_source_2d(nx, ny, dx, dy, wl, current_z - zs, xs, ys, cutoff_freq,
                                    dtype=dtype)
    else:
        u = point_source_propagate_to(nx, ny, dx, dy, current_z - zs, xs, ys, wl, cutoff_freq,
                                      dtype=dtype, workers=workers)
    U = None
    if sample is not None:
        u, U = sample_apply_2d(u, sample, dx, dy, wl, cutoff_freq, workers=workers,
                               geometry=geometry)
        current_z = sample.z_start + sample.pixel_size_z * sample.grid.shape[0]
    if current_z < z_detector - 1e-8:
        leg = geometry.final_dz(current_z) if geometry is not None else z_detector - current_z
        u, U = propagate_2d(u, dx, dy, wl, leg, cutoff_freq, U2d=U, workers=workers)
    if geometry is not None:
        shape = detector_shape(det_size_x, det_size_y, ds_x, ds_y)
        eds_x, eds_y, ez, _ = geometry.detector(ds_x, ds_y)
        det = area_v1(u, dx, dy, det_size_x, det_size_y, eds_x, eds_y, ez, out_shape=shape)
    else:
        det = area_v1(u, dx, dy, det_size_x, det_size_y, ds_x, ds_y, z_detector)
    return (det, u) if return_field else det


# ---------------------------------------------------------------------------
# Per-voxel delta/beta multislice and its adjoint (the gradient path; the
# reference has no adjoint).  The forward restates Sample._apply_2d
# (optical_element.py:533-596) with the voxel grid equal to the field grid and
# delta/beta given per voxel, as fast-wave's per-voxel complex grid
# (fast-wave/fwcuda/kernels.cu:411-491).
# ---------------------------------------------------------------------------

def area_v1_matrices(nx, ny, dx, dy, size_x, size_y, ds_x, ds_y, z):
    """(Wy, Wx, cosine) with area_v1(u) == cosine * (Wy @ |u|^2 @ Wx.T).

    Wx[p, i] is the x overlap weight of cell i in pixel p as area_v1's prefix
    difference applies it (propagation.py:980-992); Wy[q, j] the y overlap.
    """
    out_y, out_x = detector_shape(size_x, size_y, ds_x, ds_y)
    x_det = (np.arange(out_x, dtype=np.float64) - out_x // 2) * ds_x
    y_det = (np.arange(out_y, dtype=np.float64) - out_y // 2) * ds_y
    xi0, xf0, xi1, xf1, _, _ = area_axis_plan(x_det, ds_x, nx, dx)
    _, _, _, _, y_lo, y_hi = area_axis_plan(y_det, ds_y, ny, dy)
    i = np.arange(nx)
    Wx = dx * ((i[None, :] < xi1[:, None]).astype(np.float64) - (i[None, :] < xi0[:, None]))
    Wx += np.where(i[None, :] == xi1[:, None], xf1[:, None] * dx, 0.0)
    Wx -= np.where(i[None, :] == xi0[:, None], xf0[:, None] * dx, 0.0)
    grid_y_left = (-(ny // 2) - 0.5) * dy
    cell_lo = grid_y_left + np.arange(ny) * dy
    cell_hi = cell_lo + dy
    Wy = np.minimum(cell_hi[None, :], y_hi[:, None]) - np.maximum(cell_lo[None, :], y_lo[:, None])
    Wy = np.where(Wy > 0, Wy, 0.0)
    cosine = z / np.sqrt(np.square(x_det)[None, :] + np.square(y_det)[:, None] + z * z)
    return Wy, Wx, cosine


def area_v1_adjoint(v2d, grad_img, dx, dy, size_x, size_y, ds_x, ds_y, z):
    """dL/dv* = v * A^T dL/dI for I = area_v1(v) = A |v|^2."""
    ny, nx = v2d.shape
    Wy, Wx, cosine = area_v1_matrices(nx, ny, dx, dy, size_x, size_y, ds_x, ds_y, z)
    return v2d * (Wy.T @ (cosine * grad_img) @ Wx)


def multislice_db_2d(u2d, delta, beta, dx, dy, wl, pz, cutoff_freq, workers=1, keep=False):
    """Per slice s: u *= material_factor(delta_s + i beta_s, pz, wl); propagate_2d(pz).

    With keep=True also returns the field before each slice.
    """
    u = u2d.astype(np.complex128)
    before = []
    for s in range(delta.shape[0]):
        if keep:
            before.append(u)
        t = material_factor(delta[s].astype(np.float64) + 1j * beta[s].astype(np.float64), pz, wl)
        u, _ = propagate_2d(u * t, dx, dy, wl, pz, cutoff_freq, workers=workers)
    return (u, before) if keep else u


def multislice_db_intensity(u2d, delta, beta, dx, dy, wl, pz, cutoff_freq, z_leg,
                            size_x, size_y, ds_x, ds_y, z_det, workers=1):
    """Slices, a free-space leg over z_leg, then area_v1 at z_det."""
    u = multislice_db_2d(u2d, delta, beta, dx, dy, wl, pz, cutoff_freq, workers)
    v, _ = propagate_2d(u, dx, dy, wl, z_leg, cutoff_freq, workers=workers)
    return area_v1(v, dx, dy, size_x, size_y, ds_x, ds_y, z_det)


def multislice_db_gradient(u2d, delta, beta, dx, dy, wl, pz, cutoff_freq, z_leg,
                           size_x, size_y, ds_x, ds_y, z_det, grad_img, workers=1):
    """(dL/d delta, dL/d beta) for a loss L with dL/dI = grad_img.

    Reverse-mode Wirtinger calculus for real L: with g = dL/du* after a
    slice, h = P^H g (propagation over -pz), w = t_s u_s,
    dL/d delta_s = 2 K Im(conj(h) w), dL/d beta_s = -2 K Re(conj(h) w),
    K = 2 pi pz / wl, and g <- conj(t_s) h.
    """
    u, before = multislice_db_2d(u2d, delta, beta, dx, dy, wl, pz, cutoff_freq, workers, keep=True)
    v, _ = propagate_2d(u, dx, dy, wl, z_leg, cutoff_freq, workers=workers)
    g = area_v1_adjoint(v, grad_img, dx, dy, size_x, size_y, ds_x, ds_y, z_det)
    g, _ = propagate_2d(g, dx, dy, wl, -z_leg, cutoff_freq, workers=workers)
    K = 2 * np.pi * pz / wl
    gd = np.zeros(delta.shape, np.float64)
    gb = np.zeros(beta.shape, np.float64)
    for s in range(delta.shape[0] - 1, -1, -1):
        h, _ = propagate_2d(g, dx, dy, wl, -pz, cutoff_freq, workers=workers)
        t = material_factor(delta[s].astype(np.float64) + 1j * beta[s].astype(np.float64), pz, wl)
        c = np.conj(h) * (t * before[s])
        gd[s] = 2 * K * c.imag
        gb[s] = -2 * K * c.real
        g = np.conj(t) * h
    return gd, gb


def relative_l2(actual, expected) -> float:
    """tests/big_wave_2d/test_p5.py:44-50."""
    a = np.asarray(actual)
    e = np.asarray(expected)
    return float(np.linalg.norm((a - e).reshape(-1)) / max(np.linalg.norm(e.reshape(-1)), np.finfo(float).eps))
