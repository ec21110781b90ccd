"""GPU parity: the HIP engine through the C ABI versus the FP64 oracle.

Tolerances: complex128 paths are compared at round-off (1e-12 relative L2
unless a reference test states a tighter bound); complex64 paths at the
single-precision levels stated per test.  Every case prints its name and
shape before touching the GPU.
"""

import json
import math
import os
import sys

import numpy as np
import pytest

import oracle as O
from _rs import lib, rs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return O.relative_l2(a, b)


def dev(x):
    return rs.DeviceVector.from_numpy(np.ascontiguousarray(x).reshape(-1))


def host(v, ny, nx):
    return v.to_numpy().reshape(ny, nx)


def gpu_fft2(x, direction):
    ny, nx = x.shape
    a = dev(x)
    b = rs.DeviceVector(nx * ny, x.dtype)
    (a.fft2 if direction < 0 else a.ifft2)(b, nx, ny)
    return host(b, ny, nx)


def test_device_visible():
    assert lib.device_count() >= 1, "no HIP device visible"


# ---------------------------------------------------------------- FFT ----

def test_fft_known_answers_from_fwtest():
    g = json.load(open(os.path.join(GOLDEN, "fwtest_fft_vectors.json")))
    for dt, tol in ((np.complex128, 1e-5), (np.complex64, 2e-5)):
        for key in ("fft2_ones_3x3", "fft2_arange_4x4", "fft2_random_3x4"):
            c = g[key]
            say("fwtest", key, dt.__name__)
            x = np.array([a + 1j * b for a, b in c["input"]], dtype=dt).reshape(c["ny"], c["nx"])
            e = np.array([a + 1j * b for a, b in c["expected"]]).reshape(c["ny"], c["nx"])
            out = gpu_fft2(x, -1)
            np.testing.assert_allclose(out, e, atol=c.get("tol", tol) if dt == np.complex128 else tol)
            back = gpu_fft2(out.astype(dt), 1)
            np.testing.assert_allclose(back, x, atol=1e-12 if dt == np.complex128 else 2e-6)
        c = g["fft_8"]
        x = np.array([a + 1j * b for a, b in c["input"]], dtype=dt).reshape(1, 8)
        # a 1 x 8 field: rows of length 8
        say("fwtest fft_8", dt.__name__)
        a = dev(x)
        b = rs.DeviceVector(8, dt)
        lib.call("rsim_fft1d_rows", b.ptr, a.ptr, 8, 1, a.code, lib.RSIM_FORWARD, None)
        np.testing.assert_allclose(b.to_numpy(), [p + 1j * q for p, q in c["expected"]], atol=1e-5)


FFT_SIZES = [(1, 2), (2, 3), (5, 6), (8, 8), (16, 12), (30, 20), (64, 32), (72, 72), (128, 256),
             (256, 256), (512, 128), (1024, 1024), (2048, 512), (4096, 256), (256, 4096),
             (8192, 64), (64, 8192), (3072, 96), (96, 3072), (1000, 10),
             # beyond one LDS pass (c128 > 5120, c64 > 10240): four-step schedule
             (8, 6144), (6144, 8), (6, 16384), (16384, 6), (5, 12000), (24, 10240)]


@pytest.mark.parametrize("ny,nx", FFT_SIZES)
def test_fft2_matches_oracle(ny, nx):
    rng = np.random.default_rng(ny * 7919 + nx)
    x = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    ref_f = O.fft2(x)
    ref_i = O.ifft2(x)
    for dt, tol in ((np.complex128, 1e-13), (np.complex64, 2e-6)):
        say("fft2", ny, nx, dt.__name__)
        xf = x.astype(dt)
        assert rel(gpu_fft2(xf, -1), ref_f) < tol
        assert rel(gpu_fft2(xf, 1), ref_i) < tol


@pytest.mark.parametrize("ny,nx", [(256, 512), (16, 6144), (6144, 16)])
def test_fft_in_place(ny, nx):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    a = dev(x)
    say("fft2 in place", ny, nx)
    a.fft2(a, nx, ny)
    assert rel(host(a, ny, nx), O.fft2(x)) < 1e-13


# ------------------------------------------------- cutoff / propagation ----

def test_circular_cutoff_axis_and_diagonal_cells():
    """tests/big_wave_2d/test_p5.py:126-140 (exact)."""
    nx = ny = 64
    spec = np.zeros((ny, nx), dtype=np.complex128)
    spec[0, 14], spec[0, 15], spec[10, 10], spec[10, 11] = 1, 2, 3, 4
    v = dev(spec)
    say("cutoff cells")
    rs.apply_frequency_cutoff_2d(v, 10 / nx, 1.0, 1.0, nx, ny, 4 * nx)
    out = host(v, ny, nx)
    assert out[0, 14] == 1.0 and out[0, 15] == 0.0 and out[10, 10] == 3.0 and out[10, 11] == 0.0


def test_cutoff_mask_matches_oracle_exactly():
    rng = np.random.default_rng(2)
    ny, nx = 96, 128
    x = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    for dt in (np.complex128, np.complex64):
        v = dev(x.astype(dt))
        say("cutoff mask", dt.__name__)
        rs.apply_frequency_cutoff_2d(v, 0.21 / 1e-6, 1e-6, 1.3e-6, nx, ny, nx)
        exp = O.apply_frequency_cutoff_2d(x.astype(dt), 0.21 / 1e-6, 1e-6, 1.3e-6)
        np.testing.assert_array_equal(host(v, ny, nx), exp)


def test_plane_and_tilted_plane_wave_transfer():
    """test_p5.py:84-100: exact H on a Fourier mode, 2e-14 (c128)."""
    nx, ny, wl, z = 64, 32, 0.02, 0.7
    x = np.arange(nx)[None, :]
    y = np.arange(ny)[:, None]
    for mx, my in ((0, 0), (5, -3)):
        field = np.exp(2j * np.pi * (mx * x / nx + my * y / ny))
        u, U = dev(field), rs.DeviceVector(nx * ny, np.complex128)
        say("plane wave", mx, my)
        rs.propagate_2d(u, U, 1.0, 1.0, wl, z, 4 * nx, 20.0, nx, ny)
        fx, fy = mx / nx, my / ny
        t = np.exp(2j * np.pi * z / wl) * np.exp(-1j * np.pi * wl * z * (fx * fx + fy * fy))
        assert rel(host(u, ny, nx), field * t) < 2e-14
        assert rel(host(U, ny, nx), O.fft2(host(u, ny, nx))) < 1e-13


@pytest.mark.parametrize("ny,nx", [(32, 64), (256, 256), (72, 72), (512, 1024)])
def test_propagate_2d_matches_oracle(ny, nx):
    rng = np.random.default_rng(ny + nx)
    x = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    dx, dy, wl, dz = 1e-6, 0.8e-6, 1e-10, 0.05
    cut = 0.35 / dx
    ue, Ue = O.propagate_2d(x, dx, dy, wl, dz, cut)
    for dt, tol in ((np.complex128, 1e-12), (np.complex64, 3e-6)):
        u, U = dev(x.astype(dt)), rs.DeviceVector(nx * ny, dt)
        say("propagate_2d", ny, nx, dt.__name__)
        rs.propagate_2d(u, U, dx, dy, wl, dz, nx, cut, nx, ny)
        assert rel(host(u, ny, nx), ue) < tol
        assert rel(host(U, ny, nx), Ue) < tol
        # skip_fft: reuse U as the spectrum of u
        rs.propagate_2d(u, U, dx, dy, wl, 0.02, nx, cut, nx, ny, skip_fft=True)
        ue2, _ = O.propagate_2d(ue, dx, dy, wl, 0.02, cut, U2d=Ue)
        assert rel(host(u, ny, nx), ue2) < 2 * tol


def test_point_source_amplitude_phase_offset():
    """test_p5.py:64-82 against the oracle and the closed form."""
    nx, ny, dx, dy, wl = 40, 28, 0.7e-6, 1.1e-6, 1.3e-10
    sx, sy, dist = 1.2e-6, -2.4e-6, 0.3
    p = rs.SimParams(N=nx * ny, nx=nx, ny=ny, dx=dx, dy=dy, z_detector=dist, detector_size=nx * dx,
                     detector_pixel_size_x=dx, detector_pixel_size_y=dy, wl=wl, chunk_size=nx * 4)
    exp = O.propagate_analytically_2d(nx, ny, dx, dy, dist, sx, sy, wl)
    for dt in (np.complex128, np.complex64):
        u = rs.DeviceVector(nx * ny, dt)
        say("point source", dt.__name__)
        rs.propagate_analytically_2d(u, dist, sx, sy, p)
        got = host(u, ny, nx)
        tol = 2e-15 if dt == np.complex128 else 1e-7
        assert rel(np.abs(got), np.abs(exp)) < tol
        assert np.max(np.abs(np.angle(got * np.conj(exp)))) < (3e-6 if dt == np.complex128 else 2e-7 + 3e-6)


# ------------------------------------------------------- transmission ----

def _spheres_grid(rng, nz, ny, nx, nsph, nmat):
    g = np.zeros((nz, ny, nx), dtype=np.uint8)
    zz, yy, xx = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    for _ in range(nsph):
        c = rng.uniform([0, 0, 0], [nz, ny, nx])
        r = rng.uniform(2, max(3, min(ny, nx) / 6))
        m = rng.integers(1, nmat + 1)
        g[(zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2 <= r * r] = m
    return g


DB3 = np.array([0, 3.5e-7 + 2e-10j, 1.1e-6 + 3e-9j, 2.6e-7 + 1.5e-10j])


@pytest.mark.parametrize("aligned", [True, False])
def test_transmission_matches_oracle(aligned):
    rng = np.random.default_rng(9)
    nx, ny = 96, 64
    dx, dy = 1e-6, 1e-6
    if aligned:
        grid = _spheres_grid(rng, 2, ny + 6, nx + 4, 12, 3)
        px, py, xpos, ypos = dx, dy, 1 * dx, -2 * dy
    else:
        grid = _spheres_grid(rng, 2, 50, 70, 12, 3)
        px, py, xpos, ypos = 1.37e-6, 1.21e-6, 0.33e-6, -0.71e-6
    wl, pz = 0.5e-10, 0.5e-6
    vol = rs.DeviceVolume.from_numpy(grid, DB3 * 40)
    u0 = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    for sl in (0, 1):
        db = O.sample_deltabeta_slice(grid, DB3 * 40, sl, nx, ny, dx, dy, px, py, xpos, ypos)
        exp = u0 * O.material_factor(db, pz, wl)
        for dt, tol in ((np.complex128, 1e-14), (np.complex64, 1e-7)):
            u = dev(u0.astype(dt))
            say("transmission", aligned, sl, dt.__name__)
            lib.call("rsim_transmission_2d", u.ptr, nx, ny, u.code, dx, dy, wl, vol.handle, sl, px, py,
                     pz, xpos, ypos, None)
            assert rel(host(u, ny, nx), exp) < tol


# ---------------------------------------------------------- multislice ----

def _sample_case(n=256, nz=8, seed=4, aligned="unit"):
    """aligned: "unit" (voxel pitch == field pitch 0.5 um: the voxel
    coordinates carry ~1e-13 rounding residues), "exact" (pitch 2**-21 m:
    exactly integral coordinates), "half" (voxel pitch dx/2: affine map
    x0 = 2x + c) or "bilinear" (non-integral map)."""
    rng = np.random.default_rng(seed)
    dx = 2.0 ** -21 if aligned == "exact" else 0.5e-6
    if aligned in ("unit", "exact"):
        grid = _spheres_grid(rng, nz, n, n, 20, 3)
        px = dx
    elif aligned == "half":
        grid = _spheres_grid(rng, nz, 2 * n + 6, 2 * n + 2, 20, 3)
        px = dx / 2
    else:
        grid = _spheres_grid(rng, nz, n // 2 + 3, n // 2 + 5, 10, 3)
        px = 2.13 * dx
    u0 = np.exp(1j * rng.uniform(0, 0.1, size=(n, n))) * (1 + 0.01 * rng.normal(size=(n, n)))
    spec = O.SampleSpec(grid=grid, db_list=DB3 * 30, pixel_size_x=px, pixel_size_y=px,
                        pixel_size_z=0.5e-6, x_pos=0.0, y_pos=0.0)
    return grid, spec, u0, dx


@pytest.mark.parametrize("engine", ["plain", "fused"])
@pytest.mark.parametrize("aligned", ["unit", "exact", "half", "bilinear"])
def test_sample_apply_matches_oracle(engine, aligned):
    n = 256
    grid, spec, u0, dx = _sample_case(n, 8, aligned=aligned)
    wl, cut = 0.5e-10, 0.4 / dx
    ue, Ue = O.sample_apply_2d(u0, spec, dx, dx, wl, cut)
    p = rs.SimParams(N=n * n, nx=n, ny=n, dx=dx, z_detector=1.0, detector_size=n * dx,
                     detector_pixel_size_x=dx, detector_pixel_size_y=dx, wl=wl, chunk_size=n)
    cases = [(np.complex64, 2e-6)] if engine == "fused" else [(np.complex128, 1e-12), (np.complex64, 3e-6)]
    for dt, tol in cases:
        s = rs.Sample(z_start=0.1, pixel_size_x=spec.pixel_size_x, pixel_size_z=spec.pixel_size_z,
                      grid=grid, materials=[rs.Material(str(i), 1.0) for i in range(3)],
                      x_positions=np.array([0.0]), pixel_size_y=spec.pixel_size_y,
                      y_positions=np.array([0.0]), engine=engine)
        s.db_list = DB3 * 30
        u, U = dev(u0.astype(dt)), rs.DeviceVector(n * n, dt)
        say("sample_apply", engine, aligned, dt.__name__)
        s.apply(u, U, p, cut, 0, None)
        assert rel(host(u, n, n), ue) < tol, (engine, dt)
        assert rel(host(U, n, n), Ue) < tol, (engine, dt)


@pytest.mark.parametrize("case", ["smaller", "larger_half"])
def test_fused_lookup_rows_at_volume_edges(case):
    """Integral voxel maps whose volume does not cover the field row for row:
    a volume smaller than the field, shifted (columns and rows outside it get
    no factor), and a half-pitch volume wider than the field, shifted so far
    that the field's last columns fall beyond it (affine map x0 = 2x + c).  The MID rows load
    their voxel ids before testing which columns lie inside, so offsets off
    either end of a volume row are part of this case; fused c64 vs the FP64
    oracle."""
    n, nz = 256, 4
    rng = np.random.default_rng(17)
    dx = 0.5e-6
    if case == "smaller":
        grid = _spheres_grid(rng, nz, 180, 200, 12, 3)
        px, xpos, ypos = dx, 9 * dx, -5 * dx
    else:
        grid = _spheres_grid(rng, nz, 2 * n + 10, 2 * n + 40, 20, 3)
        px, xpos, ypos = dx / 2, 70 * dx, 4 * dx
    u0 = np.exp(1j * rng.uniform(0, 0.1, size=(n, n))) * (1 + 0.01 * rng.normal(size=(n, n)))
    spec = O.SampleSpec(grid=grid, db_list=DB3 * 30, pixel_size_x=px, pixel_size_y=px,
                        pixel_size_z=0.5e-6, x_pos=xpos, y_pos=ypos)
    wl, cut = 0.5e-10, 0.4 / dx
    ue, _ = O.sample_apply_2d(u0, spec, dx, dx, wl, cut)
    vol = rs.DeviceVolume.from_numpy(grid, DB3 * 30)
    u, U = dev(u0.astype(np.complex64)), rs.DeviceVector(n * n, np.complex64)
    say("fused lookup rows at volume edges", case)
    lib.call("rsim_sample_apply_2d", u.ptr, U.ptr, n, n, u.code, dx, dx, wl, cut, vol.handle,
             px, px, 0.5e-6, xpos, ypos, 0, nz, lib.RSIM_FUSED, 0, None)
    err = rel(host(u, n, n), ue)
    say("  vs oracle", err)
    assert err < 2e-6


@pytest.mark.parametrize("n", [512, 1024, 2048, 3072, 6144])
def test_fused_sizes_match_plain_c128(n):
    """Fused complex64 against the plain complex128 GPU path (itself pinned to the oracle)."""
    rng = np.random.default_rng(n)
    nz = 4
    grid = _spheres_grid(rng, nz, n, n, 6, 3) if n <= 1024 else np.zeros((nz, n, n), np.uint8)
    if n > 1024:
        grid[:, n // 4: n // 2, n // 3: n // 2] = 2
    dx, wl = 0.5e-6, 0.5e-10
    vol = rs.DeviceVolume.from_numpy(grid, DB3 * 30)
    u0 = np.ones((n, n))
    res = {}
    for dt, eng in ((np.complex128, lib.RSIM_PLAIN), (np.complex64, lib.RSIM_FUSED)):
        u, U = dev(u0.astype(dt)), rs.DeviceVector(n * n, dt)
        say("fused size", n, dt.__name__)
        lib.call("rsim_sample_apply_2d", u.ptr, U.ptr, n, n, u.code, dx, dx, wl, 0.4 / dx, vol.handle,
                 dx, dx, 0.5e-6, 0.0, 0.0, 0, nz, eng, 0, None)
        res[dt] = (host(u, n, n), host(U, n, n))
    assert rel(res[np.complex64][0], res[np.complex128][0]) < 2e-6
    assert rel(res[np.complex64][1], res[np.complex128][1]) < 2e-6


@pytest.mark.parametrize("ny,nx", [(1024, 16384), (16384, 1024), (16384, 16384), (3072, 2048),
                                   (2048, 6144), (6144, 3072)])
def test_fused_16384_axes_match_plain_c128(ny, nx):
    """Mixed and large shapes: 16384-point rows (1024-thread register pass),
    16384-point columns (three-pass 128 x 128 schedule), 3072/6144-point axes
    (prime-factor radix-6/12 stages) against the plain complex128 path (whose
    6144- and 16384-point axes use the four-step out-of-LDS transform)."""
    nz = 3
    rng = np.random.default_rng(ny + 3 * nx)
    vol = rs.DeviceVolume.empty((nz, ny, nx), DB3 * 30)
    sph = np.stack([rng.uniform(0, nz, 40), rng.uniform(0, ny, 40), rng.uniform(0, nx, 40),
                    rng.uniform(20, 300, 40)], axis=1)
    vol.paint_spheres(sph, rng.integers(1, 3, 40))
    dx, wl = 0.5e-6, 0.5e-10
    res = {}
    for dt, eng in ((np.complex128, lib.RSIM_PLAIN), (np.complex64, lib.RSIM_FUSED)):
        u = rs.DeviceVector.from_numpy(np.ones(nx * ny, dtype=dt))
        U = rs.DeviceVector(nx * ny, dt)
        say("fused 16384 axes", ny, nx, dt.__name__)
        lib.call("rsim_sample_apply_2d", u.ptr, U.ptr, nx, ny, u.code, dx, dx, wl, 0.4 / dx,
                 vol.handle, dx, dx, 0.5e-6, 0.0, 0.0, 0, nz, eng, 0, None)
        res[dt] = (host(u, ny, nx), host(U, ny, nx))
        u.drop()
        U.drop()
    vol.destroy()
    assert rel(res[np.complex64][0], res[np.complex128][0]) < 2e-6
    assert rel(res[np.complex64][1], res[np.complex128][1]) < 2e-6


# ------------------------------------------------------------ detector ----

def test_area_v1_matches_oracle_non_integer_ratio():
    rng = np.random.default_rng(23)
    nx, ny = 11, 7
    f = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    p = rs.SimParams(N=nx * ny, nx=nx, ny=ny, dx=0.7, dy=1.1, z_detector=8.0, detector_size=5.2,
                     detector_size_x=5.2, detector_size_y=5.1, detector_pixel_size_x=1.3,
                     detector_pixel_size_y=1.7, wl=1e-10, chunk_size=2 * nx)
    exp = O.area_v1(f, 0.7, 1.1, 5.2, 5.1, 1.3, 1.7, 8.0)
    say("area_v1 11x7")
    got = rs.square_and_downsample_2d(dev(f), p, 8.0)
    np.testing.assert_allclose(got, exp, rtol=2e-14, atol=2e-14)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_area_v1_binning_matches_oracle(dt):
    rng = np.random.default_rng(5)
    n = 512
    f = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))).astype(dt)
    dx = 1e-6
    p = rs.SimParams(N=n * n, nx=n, ny=n, dx=dx, z_detector=0.5, detector_size=n * dx,
                     detector_pixel_size_x=4 * dx, detector_pixel_size_y=4 * dx, wl=1e-10, chunk_size=n)
    exp = O.area_v1(f, dx, dx, n * dx, n * dx, 4 * dx, 4 * dx, 0.5)
    say("area_v1 512 4x4", np.dtype(dt).name)
    got = rs.square_and_downsample_2d(dev(f), p, 0.5)
    assert got.shape == (128, 128)
    np.testing.assert_allclose(got, exp, rtol=1e-12 if dt == np.complex128 else 1e-6, atol=0)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("ny,nx,pix,span", [
    (96, 512, (37.3, 5.1), (1.5, 1.3)),   # detector overhangs the field on all sides
    (64, 8192, (20.0, 3.0), (1.0, 0.9)),  # 256 detector columns span > 4096 field cells
])
def test_area_v1_overhang_and_wide_spans(dt, ny, nx, pix, span):
    """Detector pixels outside the field, partial edge pixels and x spans wider
    than the row stage's LDS span (read straight from the field)."""
    rng = np.random.default_rng(29)
    f = (rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))).astype(dt)
    dx, dy = 0.8e-6, 1.3e-6
    sx, sy = span[0] * nx * dx, span[1] * ny * dy
    psx, psy = pix[0] * dx, pix[1] * dy
    p = rs.SimParams(N=nx * ny, nx=nx, ny=ny, dx=dx, dy=dy, z_detector=0.4, detector_size=sx,
                     detector_size_x=sx, detector_size_y=sy, detector_pixel_size_x=psx,
                     detector_pixel_size_y=psy, wl=1e-10, chunk_size=nx)
    exp = O.area_v1(f, dx, dy, sx, sy, psx, psy, 0.4)
    say("area_v1 overhang/wide", ny, nx, np.dtype(dt).name)
    got = rs.square_and_downsample_2d(dev(f), p, 0.4)
    assert got.shape == exp.shape
    np.testing.assert_allclose(got, exp, rtol=1e-12 if dt == np.complex128 else 1e-6,
                               atol=1e-12 * np.abs(exp).max())


# ----------------------------------------------------- end to end (P5) ----

@pytest.mark.parametrize("n", [4096, 8192, 16384])
@pytest.mark.parametrize("mode", ["vacuum", "sample"])
@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_end_to_end_reference_p5_golden(n, mode, dt):
    """run_simulation vs the reference's own P5 fingerprints at 4096^2,
    8192^2 and 16384^2 (the same physical field sampled ever finer; at
    16384^2 complex64 takes the three-pass column schedule)."""
    g = json.load(open(os.path.join(GOLDEN, f"p5_convergence_{n}.json")))[mode]
    nx = ny = n
    dx = 3.2768e-3 / nx
    wl = 1e-9
    p = rs.SimParams(N=nx * ny, nx=nx, ny=ny, dx=dx, dy=dx, z_detector=10.0, detector_size=4e-4,
                     detector_size_x=4e-4, detector_size_y=4e-4, detector_pixel_size_x=8e-7,
                     detector_pixel_size_y=8e-7, wl=wl, chunk_size=nx * 32)
    elements, table = [], []
    if mode == "sample":
        mat = rs.Material("P5-benchmark", 1.0)
        elements = [rs.Sample(z_start=5.0, pixel_size_x=4e-4 / 256, pixel_size_y=4e-4 / 256,
                              pixel_size_z=1e-6, grid=np.ones((1, 258, 258), dtype=np.uint32),
                              materials=[mat], x_positions=np.array([0.0]),
                              y_positions=np.array([0.0]))]
        table = [(mat, 1.2e-7 + 2.4e-9j)]
    ang = math.asin(2.5e5 * wl)
    u, U = rs.DeviceVector(nx * ny, dt), rs.DeviceVector(nx * ny, dt)
    say("p5 e2e", n, mode, np.dtype(dt).name)
    det = rs.run_simulation(p, rs.PointSource(x=1e-6, y=-1.4e-6, z=0.0), elements,
                            [ang] * (len(elements) + 1), u, U, table)[0]
    s = g["centerline_start_pixel"]
    line = det[det.shape[0] // 2, s:s + 256]
    assert list(det.shape) == g["detected_shape"][1:]
    assert rel(line, g["centerline_256"]) < 5e-6
    assert abs(det.sum() - g["detected_energy"]) / g["detected_energy"] < 5e-6
    assert abs(det.max() - g["detected_peak"]) / g["detected_peak"] < 5e-6


def test_phase_stepping_replays_from_snapshot():
    n = 128
    dx, wl = 1e-6, 1e-9
    p = rs.SimParams(N=n * n, nx=n, ny=n, dx=dx, z_detector=1.3, detector_size=n * dx,
                     detector_pixel_size_x=2 * dx, detector_pixel_size_y=2 * dx, wl=wl, chunk_size=n)
    grid = np.zeros((2, n + 2, n + 2), dtype=np.uint32)
    grid[:, 40:80, 30:70] = 1
    mat = rs.Material("X", 1.0)
    xs = np.array([0.0, 3 * dx, 7 * dx])
    el = rs.Sample(z_start=1.0, pixel_size_x=dx, pixel_size_z=1e-6, grid=grid, materials=[mat],
                   x_positions=xs, pixel_size_y=dx, y_positions=np.zeros(3), engine="plain")
    u, U = rs.DeviceVector(n * n, np.complex128), rs.DeviceVector(n * n, np.complex128)
    say("phase stepping")
    outs = rs.run_simulation(p, rs.PointSource(x=0.0, z=0.0), [el], [1e-4, 1e-4], u, U,
                             [(mat, 2e-6 + 1e-8j)])
    assert len(outs) == 3
    cut = math.sin(1e-4) / wl
    for k, xpos in enumerate(xs):
        spec = O.SampleSpec(grid=grid, db_list=np.array([0, 2e-6 + 1e-8j]), pixel_size_x=dx,
                            pixel_size_y=dx, pixel_size_z=1e-6, x_pos=xpos, z_start=1.0)
        exp = O.run_point_source_sample(n, n, dx, dx, wl, (0.0, 0.0, 0.0), cut, 1.3, n * dx, n * dx,
                                        2 * dx, 2 * dx, spec)
        assert rel(outs[k], exp) < 1e-11


N_ROW_VARIANTS, N_COL_VARIANTS = 10, 9
VARIANT_CASES = ([(1024, 1024, (r, 0)) for r in range(N_ROW_VARIANTS)]
                 # the four-waves row budget at every fused length
                 + [(256, n, (9, 0)) for n in (256, 512, 2048, 3072, 4096, 6144, 8192)]
                 + [(1024, 1024, (r, c)) for c in range(1, N_COL_VARIANTS)
                    for r in (0, N_ROW_VARIANTS - 1)]
                 + [(4096, 4096, (i % N_ROW_VARIANTS, i)) for i in range(N_COL_VARIANTS)]
                 + [(256, 16384, (r, 0)) for r in range(N_ROW_VARIANTS)]
                 + [(3072, 6144, (i % N_ROW_VARIANTS, i)) for i in range(N_COL_VARIANTS)]
                 + [(6144, 3072, (i % N_ROW_VARIANTS, i)) for i in range(N_COL_VARIANTS)])


@pytest.mark.parametrize("ny,nx,variant", VARIANT_CASES)
def test_every_fused_launch_shape_matches_plain_c128(ny, nx, variant, monkeypatch):
    """Each row/column launch-shape variant the autotuner may pick, forced in turn."""
    nz = 2
    grid = np.zeros((nz, ny, nx), np.uint8)
    grid[:, ny // 5: ny // 2, nx // 3: (2 * nx) // 3] = 1
    grid[1, ny // 2: (3 * ny) // 4, nx // 4: nx // 2] = 2
    dx, wl = 0.5e-6, 0.5e-10
    vol = rs.DeviceVolume.from_numpy(grid, DB3 * 30)
    rng = np.random.default_rng(7)
    u0 = np.exp(1j * rng.uniform(0, 0.2, size=(ny, nx)))
    out = {}
    monkeypatch.setenv("RAVESIM_FUSED_VARIANT", f"{variant[0]},{variant[1]}")
    for dt, eng in ((np.complex128, lib.RSIM_PLAIN), (np.complex64, lib.RSIM_FUSED)):
        u, U = dev(u0.astype(dt)), rs.DeviceVector(nx * ny, dt)
        say("variant", ny, nx, variant, np.dtype(dt).name)
        lib.call("rsim_sample_apply_2d", u.ptr, U.ptr, nx, ny, u.code, dx, dx, wl, 0.4 / dx, vol.handle,
                 dx, dx, 0.5e-6, 0.0, 0.0, 0, nz, eng, 0, None)
        out[dt] = (host(u, ny, nx), host(U, ny, nx))
    assert rel(out[np.complex64][0], out[np.complex128][0]) < 1e-6
    assert rel(out[np.complex64][1], out[np.complex128][1]) < 1e-6
