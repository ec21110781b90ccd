"""Pins of the CPU oracle against reference-held vectors, closed forms and invariants.

Each test names the reference test or file it mirrors.  None of them restates
the oracle's own formula: they compare against golden numbers produced by the
reference, against analytic physics (Fresnel integrals, plane-wave transfer,
Beer-Lambert, 1/r spherical amplitude), against brute-force loops, or against
symmetries a sign/index mistake would break.
"""

import json
import math
import os

import numpy as np
import pytest
from scipy.special import fresnel

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _cplx(pairs):
    return np.array([a + 1j * b for a, b in pairs], dtype=np.complex128)


# --------------------------------------------------------------------------
# FFT conventions: fast-wave/test/fwtest.cu:59-334 known answers
# --------------------------------------------------------------------------

def test_fftfreq_matches_fwtest_vector():
    g = json.load(open(os.path.join(GOLDEN, "fwtest_fft_vectors.json")))
    np.testing.assert_allclose(O.fftfreq(10), g["fftfreq_10"], atol=1e-7)


def test_fft_known_answers_from_fwtest():
    g = json.load(open(os.path.join(GOLDEN, "fwtest_fft_vectors.json")))
    c = g["fft_8"]
    out = O.fft2(_cplx(c["input"]).reshape(1, 8))[0]
    np.testing.assert_allclose(out, _cplx(c["expected"]), atol=c["tol"])
    for key in ("fft2_ones_3x3", "fft2_arange_4x4", "fft2_random_3x4"):
        c = g[key]
        x = _cplx(c["input"]).reshape(c["ny"], c["nx"])
        out = O.fft2(x)
        np.testing.assert_allclose(out.reshape(-1), _cplx(c["expected"]), atol=c.get("tol", 1e-9))
        np.testing.assert_allclose(O.ifft2(out), x, atol=1e-12)
    ones = np.ones((1, 16), dtype=np.complex128)
    f = O.fft2(ones)
    assert f[0, 0] == 16.0 and np.all(f[0, 1:] == 0.0)


def test_fft2_against_brute_force_dft():
    rng = np.random.default_rng(5)
    for ny, nx in ((3, 5), (4, 6), (8, 3)):
        x = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
        wy = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
        wx = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
        np.testing.assert_allclose(O.fft2(x), wy @ x @ wx.T, atol=1e-12)


# --------------------------------------------------------------------------
# Closed forms: tests/big_wave_2d/test_p5.py
# --------------------------------------------------------------------------

def test_point_source_amplitude_phase_and_offset():
    """test_p5.py:64-82."""
    nx, ny, dx, dy, wl = 40, 28, 0.7e-6, 1.1e-6, 1.3e-10
    sx, sy, dist = 1.2e-6, -2.4e-6, 0.3
    u = O.propagate_analytically_2d(nx, ny, dx, dy, dist, sx, sy, wl)
    # independent geometry: per-pixel loop in pure Python
    amp = np.empty((ny, nx))
    for j in range(ny):
        for i in range(nx):
            xx = (i - nx / 2) * dx - sx
            yy = (j - ny / 2) * dy - sy
            amp[j, i] = 1.0 / math.sqrt(xx * xx + yy * yy + dist * dist)
    assert O.relative_l2(np.abs(u), amp) < 2e-15
    expected = np.exp(2j * np.pi * (1.0 / amp) / wl) * amp
    assert np.max(np.abs(np.angle(u * np.conj(expected)))) < 3e-6
    peak = np.unravel_index(np.argmax(np.abs(u)), u.shape)
    assert abs(peak[1] - (nx / 2 + sx / dx)) <= 1.0
    assert abs(peak[0] - (ny / 2 + sy / dy)) <= 1.0


def test_plane_and_tilted_plane_wave_transfer():
    """test_p5.py:84-100: exact H for a pure Fourier mode, 2e-14."""
    nx, ny, dx, dy, wl, z = 64, 32, 1.0, 1.0, 0.02, 0.7
    x = np.arange(nx)[None, :]
    y = np.arange(ny)[:, None]
    for mx, my in ((0, 0), (5, -3)):
        field = np.exp(2j * np.pi * (mx * x / nx + my * y / ny))
        u, U = O.propagate_2d(field, dx, dy, wl, z, 20.0)
        fx, fy = mx / (nx * dx), my / (ny * dy)
        transfer = np.exp(2j * np.pi * z / wl) * np.exp(-1j * np.pi * wl * z * (fx * fx + fy * fy))
        assert O.relative_l2(u, field * transfer) < 2e-14
        # the returned spectrum is the FFT of the returned field (apply contract)
        assert O.relative_l2(U, O.fft2(u)) < 1e-13


def test_uniform_slab_phase_and_beer_lambert():
    """test_p5.py:102-124."""
    nx, ny, dx, wl = 32, 16, 1e-6, 1e-10
    thickness = 2e-7
    db = 1.2e-7 + 2.4e-8j
    s = O.SampleSpec(grid=np.ones((1, ny + 2, nx + 2), dtype=np.uint32),
                     db_list=np.array([0, db]), pixel_size_x=dx, pixel_size_y=dx,
                     pixel_size_z=thickness)
    u, _ = O.sample_apply_2d(np.ones((ny, nx), dtype=np.complex128), s, dx, dx, wl, 1.0 / dx)
    # closed form written from the physics, not from material_factor
    phase = -2 * np.pi * thickness / wl * db.real + 2 * np.pi * thickness / wl
    expected = np.exp(-2 * np.pi * thickness / wl * db.imag) * np.exp(1j * phase)
    assert np.max(np.abs(np.angle(u * np.conj(expected)))) < 2e-12
    assert abs(float(np.mean(np.abs(u) ** 2)) - math.exp(-4 * math.pi * thickness * db.imag / wl)) < 2e-13


def test_circular_cutoff_axis_and_diagonal_cells():
    """test_p5.py:126-140."""
    nx = ny = 64
    cutoff = 10 / nx
    spec = np.zeros((ny, nx), dtype=np.complex128)
    spec[0, 14] = 1.0
    spec[0, 15] = 2.0
    spec[10, 10] = 3.0
    spec[10, 11] = 4.0
    out = O.apply_frequency_cutoff_2d(spec, cutoff, 1.0, 1.0)
    assert out[0, 14] == 1.0 and out[0, 15] == 0.0
    assert out[10, 10] == 3.0 and out[10, 11] == 0.0


def test_rectangular_aperture_matches_fresnel_integrals():
    """test_p5.py:142-172."""
    nx = ny = 256
    dx, wl, z = 1e-6, 1e-10, 0.4
    ap = np.zeros((ny, nx), dtype=np.complex128)
    y0, x0 = ny // 2, nx // 2
    ap[y0 - 8:y0 + 8, x0 - 8:x0 + 8] = 1.0
    u, _ = O.propagate_2d(ap, dx, dx, wl, z, 1.0 / dx)
    coord = (np.arange(nx) - nx // 2) * dx
    width = 16 * dx
    scale = np.sqrt(2.0 / (wl * z))
    center = -0.5 * dx
    s0, c0 = fresnel(scale * (center - width / 2 - coord))
    s1, c1 = fresnel(scale * (center + width / 2 - coord))
    one_d = (c1 - c0) + 1j * (s1 - s0)
    expected = np.abs(one_d[:, None] * one_d[None, :]) ** 2
    actual = np.abs(u) ** 2
    expected /= expected.max()
    actual /= actual.max()
    c = np.s_[48:-48, 48:-48]
    assert O.relative_l2(actual[c], expected[c]) < 0.12


def test_parseval_power_conserved_without_clipping():
    rng = np.random.default_rng(11)
    u0 = rng.normal(size=(32, 48)) + 1j * rng.normal(size=(32, 48))
    u, _ = O.propagate_2d(u0, 1e-6, 1.3e-6, 1e-10, 0.3, 1e7)
    assert abs(np.sum(np.abs(u) ** 2) / np.sum(np.abs(u0) ** 2) - 1) < 1e-13


def test_offset_slab_matches_explicit_transmission():
    """test_p5.py:174-213 with a fixed delta-beta (the logic does not depend on nist_lookup)."""
    nx = ny = 64
    wl = O.convert_energy_wavelength(8000.0)
    dx = 1e-6
    thickness = 2e-7
    margin = 12
    grid = np.zeros((1, ny + margin, nx + margin), dtype=np.uint32)
    grid[0, 20:56, 27:45] = 1
    db = 3.1e-6 + 2.2e-7j
    s = O.SampleSpec(grid=grid, db_list=np.array([0, db]), pixel_size_x=dx, pixel_size_y=dx,
                     pixel_size_z=thickness, x_pos=2 * dx, y_pos=-3 * dx)
    cutoff = 0.4 / dx
    actual, _ = O.sample_apply_2d(np.ones((ny, nx), dtype=np.complex128), s, dx, dx, wl, cutoff)
    mask = grid[0, 3:3 + ny, 8:8 + nx]
    t = np.exp(-2 * np.pi * thickness / wl * db.imag * (mask == 1)) * np.exp(
        -2j * np.pi * thickness / wl * db.real * (mask == 1))
    expected, _ = O.propagate_2d(t.astype(np.complex128), dx, dx, wl, thickness, cutoff)
    assert O.relative_l2(actual, expected) < 2e-13


def test_bilinear_map_matches_per_pixel_brute_force():
    """test_p1.py:267-317: per-pixel reference of the bilinear/vacuum map, non-integer ratio."""
    nx, ny = 24, 16
    grid = np.zeros((1, 13, 19), dtype=np.uint32)
    grid[:, 2:11, 3:16] = 1
    grid[:, 5:7, 6:9] = 2
    db_list = np.array([0, 1e-6 + 1e-9j, 3e-6 + 4e-8j])
    dx, dy, px, py = 1e-6, 1.2e-6, 1.7e-6, 1.9e-6
    xpos, ypos = 0.3e-6, -0.45e-6
    got = O.sample_deltabeta_slice(grid, db_list, 0, nx, ny, dx, dy, px, py, xpos, ypos)
    for j in range(ny):
        for i in range(nx):
            xi = ((i - nx / 2) * dx + xpos + 19 * px * 0.5) / px
            yi = ((j - ny / 2) * dy + ypos + 13 * py * 0.5) / py
            if not (0 <= xi < 18 and 0 <= yi < 12):
                assert got[j, i] == 0
                continue
            x0 = min(max(math.floor(xi), 0), 17)
            y0 = min(max(math.floor(yi), 0), 11)
            fx, fy = xi - math.floor(xi), yi - math.floor(yi)
            v = ((db_list[grid[0, y0, x0]] * (1 - fx) + db_list[grid[0, y0, x0 + 1]] * fx) * (1 - fy)
                 + (db_list[grid[0, y0 + 1, x0]] * (1 - fx) + db_list[grid[0, y0 + 1, x0 + 1]] * fx) * fy)
            assert abs(got[j, i] - v) <= 1e-15 * abs(v) + 1e-30


def test_hollow_capsule_is_xy_symmetric():
    """test_p5.py:215-256."""
    nx = ny = 72
    wl = O.convert_energy_wavelength(8000.0)
    dx = 0.5e-6
    nz, margin = 7, 8
    shape = (nz, ny + margin, nx + margin)
    z = (np.arange(nz) - (nz - 1) / 2.0)[:, None, None]
    y = (np.arange(shape[1]) - shape[1] / 2.0)[None, :, None]
    x = (np.arange(shape[2]) - shape[2] / 2.0)[None, None, :]
    r = np.sqrt(x * x + y * y + z * z)
    db = np.array([0, 1.4e-5 + 3e-8j])

    def sim(grid):
        s = O.SampleSpec(grid=grid, db_list=db, pixel_size_x=dx, pixel_size_y=dx, pixel_size_z=0.25e-6)
        u, _ = O.sample_apply_2d(np.ones((ny, nx), dtype=np.complex128), s, dx, dx, wl, 0.4 / dx)
        return u

    shell = sim(((r <= 18.0) & (r >= 11.0)).astype(np.uint32))
    solid = sim((r <= 18.0).astype(np.uint32))
    inten = np.abs(shell) ** 2
    assert O.relative_l2(inten[ny // 2, :], inten[:, nx // 2]) < 2e-13
    assert O.relative_l2(shell, solid) > 1e-3


# --------------------------------------------------------------------------
# Detector: tests/big_wave_2d/test_p3.py
# --------------------------------------------------------------------------

def _area_brute(field, dx, dy, sx, sy, dsx, dsy, z):
    ny, nx = field.shape
    oy, ox = int(sy // dsy), int(sx // dsx)
    glx = (-(nx // 2) - 0.5) * dx
    gly = (-(ny // 2) - 0.5) * dy
    res = np.zeros((oy, ox))
    for q in range(oy):
        yq = (q - oy // 2) * dsy
        for p in range(ox):
            xp = (p - ox // 2) * dsx
            v = 0.0
            for j in range(ny):
                cy = gly + j * dy
                ovy = max(0.0, min(yq + dsy / 2, cy + dy) - max(yq - dsy / 2, cy))
                for i in range(nx):
                    cx = glx + i * dx
                    ovx = max(0.0, min(xp + dsx / 2, cx + dx) - max(xp - dsx / 2, cx))
                    s = field[j, i]
                    v += (s.real * s.real + s.imag * s.imag) * ovx * ovy
            res[q, p] = v * z / math.sqrt(xp * xp + yq * yq + z * z)
    return res


def test_area_v1_matches_brute_force_non_integer_ratio():
    """test_p3.py:208-220 (params test_p3.py:29-55)."""
    rng = np.random.default_rng(23)
    nx, ny, dx, dy = 11, 7, 0.7, 1.1
    f = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    got = O.area_v1(f, dx, dy, 5.2, 5.1, 1.3, 1.7, 8.0)
    exp = _area_brute(f, dx, dy, 5.2, 5.1, 1.3, 1.7, 8.0)
    np.testing.assert_allclose(got, exp, rtol=2e-14, atol=2e-14)


def test_area_v1_binning_4x4_matches_brute_force():
    rng = np.random.default_rng(3)
    nx = ny = 16
    f = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    got = O.area_v1(f, 1.0, 1.0, 16.0, 16.0, 4.0, 4.0, 50.0)
    exp = _area_brute(f, 1.0, 1.0, 16.0, 16.0, 4.0, 4.0, 50.0)
    np.testing.assert_allclose(got, exp, rtol=2e-14, atol=2e-14)


def test_area_v1_uniform_field_preserves_area():
    """test_p3.py:222-243."""
    nx, ny = 11, 7
    f = np.ones((ny, nx), dtype=np.complex64)
    got = O.area_v1(f, 0.7, 1.1, 3.9, 3.4, 1.3, 1.7, 8.0)
    oy, ox = got.shape
    x = (np.arange(ox) - ox // 2) * 1.3
    y = (np.arange(oy) - oy // 2) * 1.7
    cos = 8.0 / np.sqrt(x[None, :] ** 2 + y[:, None] ** 2 + 64.0)
    np.testing.assert_allclose(got / cos, 1.3 * 1.7, rtol=2e-7, atol=2e-7)


# --------------------------------------------------------------------------
# End to end: the reference's own golden fingerprints (P5, 4096^2)
# --------------------------------------------------------------------------

def _p5_case(mode):
    g = json.load(open(os.path.join(GOLDEN, "p5_convergence_4096.json")))[mode]
    sample = None
    if mode == "sample":
        px = 4e-4 / 256
        sample = O.SampleSpec(grid=np.ones((1, 258, 258), dtype=np.uint32),
                              db_list=np.array([0, 1.2e-7 + 2.4e-9j]), pixel_size_x=px,
                              pixel_size_y=px, pixel_size_z=1e-6, z_start=5.0)
    return g, sample


@pytest.mark.parametrize("mode", ["vacuum", "sample"])
def test_end_to_end_matches_reference_p5_golden_4096(mode):
    """p5_results/*_convergence_4096.json, inputs per p5_benchmark.py:69-185.

    The goldens come from the reference's complex64 out-of-core path; the
    oracle (complex128) agrees to ~5e-7 relative.
    """
    g, sample = _p5_case(mode)
    nx = ny = 4096
    dx = 3.2768e-3 / nx
    det = O.run_point_source_sample(nx, ny, dx, dx, 1e-9, (1e-6, -1.4e-6, 0.0), 2.5e5, 10.0,
                                    4e-4, 4e-4, 8e-7, 8e-7, sample, workers=os.cpu_count() or 1)
    assert list(det.shape) == g["detected_shape"][1:]
    s = g["centerline_start_pixel"]
    line = det[det.shape[0] // 2, s:s + 256]
    assert O.relative_l2(line, g["centerline_256"]) < 5e-6
    assert abs(det.sum() - g["detected_energy"]) / g["detected_energy"] < 5e-6
    assert abs(det.max() - g["detected_peak"]) / g["detected_peak"] < 5e-6
