"""GPU parity of the Fresnel-scaled / cone-beam BPM path versus the oracle.

Tolerances: complex128 at round-off (1e-12 relative L2; the reference's own
source tests state absolute 1e-13 / 2e-12); complex64 at the single-precision
levels stated per case.  Every case prints its name before touching the GPU.
"""

import sys

import numpy as np
import pytest

import oracle as O
from _rs import lib, rs

pytestmark = pytest.mark.gpu

DB3 = np.array([0, 3.5e-7 + 2e-10j, 1.1e-6 + 3e-9j, 2.6e-7 + 1.5e-10j])


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def dev(x):
    return rs.DeviceVector.from_numpy(np.ascontiguousarray(x).reshape(-1))


def host(v, ny, nx):
    return v.to_numpy().reshape(ny, nx)


def source_params(**kw):
    p = rs.SimParams(N=16 * 8, nx=16, ny=8, dx=2.0e-6, dy=3.0e-6, z_detector=5.0,
                     detector_size=1.0e-3, detector_size_x=1.0e-3, detector_size_y=1.0e-3,
                     detector_pixel_size_x=10.0e-6, detector_pixel_size_y=10.0e-6, wl=1.0e-10,
                     chunk_size=16 * 4, use_fresnel_scaling=True, **kw)
    p.configure_fresnel_detector(z_source=0.0, z_sample=1.0)
    return p


@pytest.mark.parametrize("dt,atol", [(np.complex128, 2e-12), (np.complex64, 2e-6)])
def test_fresnel_point_source_known_answers(dt, atol):
    """Centred source -> constant 1/dz; off-axis -> the linear phase ramp
    (tests/fresnel_scaling/test_big_wave_source.py:46-73)."""
    p = source_params()
    u, U = rs.DeviceVector(p.N, dt), rs.DeviceVector(p.N, dt)
    say("fresnel source centred", dt.__name__)
    rs.PointSource(x=0.0, y=0.0, z=0.0).propagate_to(1.0, p, 1.0 / p.dx, u, U, None)
    np.testing.assert_allclose(u.to_numpy(), 1.0 + 0.0j, rtol=0.0,
                               atol=1e-13 if dt == np.complex128 else 1e-6)
    src = rs.PointSource(x=25.0e-6, y=-30.0e-6, z=0.0)
    say("fresnel source off-axis", dt.__name__)
    src.propagate_to(1.0, p, 1.0 / p.dx, u, U, None)
    iy, ix = np.indices((8, 16))
    x = (ix - 8.0) * p.dx
    y = (iy - 4.0) * p.dy
    exp = np.exp(-2j * np.pi * (x * src.x + y * src.y) / p.wl)
    np.testing.assert_allclose(host(u, 8, 16), exp, rtol=0.0, atol=atol)


def test_tilted_plane_matches_oracle_with_cutoff():
    """A 256x128 grid, a source off-axis by non-bin multiples (spectral peak at
    bins (30.4, -14.6), inside the cutoff), dz = 0.37 and a cutoff that
    removes the outer part of the leakage: field after FFT/cutoff/IFFT."""
    nx, ny, dx, dy, wl = 256, 128, 0.8e-6, 1.1e-6, 0.6e-10
    dz, xs, ys, cut = 0.37, 3.3e-6, -2.3e-6, 0.2 / 0.8e-6
    p = rs.SimParams(N=nx * ny, ny=ny, dx=dx, dy=dy, z_detector=2.0, detector_size=nx * dx,
                     detector_pixel_size_x=dx, detector_pixel_size_y=dy, wl=wl, chunk_size=nx,
                     use_fresnel_scaling=True)
    p.configure_fresnel_detector(z_source=0.0, z_sample=dz)
    exp = O.fresnel_plane_source_2d(nx, ny, dx, dy, wl, dz, xs, ys, cut)
    for dt, tol in ((np.complex128, 1e-12), (np.complex64, 2e-6)):
        u, U = rs.DeviceVector(nx * ny, dt), rs.DeviceVector(nx * ny, dt)
        say("tilted plane + cutoff", dt.__name__)
        rs.PointSource(x=xs, y=ys, z=0.0).propagate_to(dz, p, cut, u, U, None)
        assert O.relative_l2(host(u, ny, nx), exp) < tol, dt


def _spheres_grid(rng, nz, ny, nx, nsph, nmat):
    g = np.zeros((nz, ny, nx), dtype=np.uint8)
    zz, yy, xx = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5,
                             indexing="ij")
    for _ in range(nsph):
        c = rng.uniform([0, 0, 0], [nz, ny, nx])
        r = rng.uniform(2, max(3, min(ny, nx) / 6))
        g[(zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2 <= r * r] = rng.integers(1, nmat + 1)
    return g


@pytest.mark.parametrize("scale", [1.0, 1.0375, 0.91])
def test_scaled_transmission_matches_oracle(scale):
    rng = np.random.default_rng(19)
    nx, ny, dx, dy = 128, 96, 1e-6, 1e-6
    grid = _spheres_grid(rng, 2, ny + 6, nx + 4, 12, 3)
    wl, pz, xpos, ypos = 0.5e-10, 0.5e-6, 1.0 * dx, -2.0 * dy
    vol = rs.DeviceVolume.from_numpy(grid, DB3 * 40)
    u0 = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    db = O.sample_deltabeta_slice(grid, DB3 * 40, 1, nx, ny, dx, dy, dx, dy, xpos, ypos, scale)
    exp = u0 * O.material_factor(db, pz, wl)
    for dt, tol in ((np.complex128, 1e-14), (np.complex64, 1e-7)):
        u = dev(u0.astype(dt))
        say("scaled transmission", scale, dt.__name__)
        lib.call("rsim_transmission_2d_scaled", u.ptr, nx, ny, u.code, dx, dy, wl, vol.handle, 1,
                 dx, dy, pz, xpos, ypos, scale, None)
        assert O.relative_l2(host(u, ny, nx), exp) < tol


def _cb_case(n, nz, seed):
    """Source 100 um upstream of the sample: the local magnification grows
    from 1.0025 to 1.0375 across the 8 slices, so edge coordinates move by
    several voxels and the slice distances shrink by 7 %."""
    rng = np.random.default_rng(seed)
    dx = 0.5e-6
    grid = _spheres_grid(rng, nz, n + 8, n + 8, 20, 3)
    spec = O.SampleSpec(grid=grid, db_list=DB3 * 30, pixel_size_x=dx, pixel_size_y=dx,
                        pixel_size_z=0.5e-6, x_pos=0.0, y_pos=0.0, z_start=1e-4)
    u0 = np.exp(1j * rng.uniform(0, 0.1, size=(n, n))) * (1 + 0.01 * rng.normal(size=(n, n)))
    return grid, spec, u0, dx


@pytest.mark.parametrize("engine", ["plain", "fused"])
def test_cb_bpm_sample_apply_matches_oracle(engine):
    n, nz = 256, 8
    grid, spec, u0, dx = _cb_case(n, nz, 41)
    wl, cut = 0.5e-10, 0.4 / dx
    geom = O.FresnelGeometry(z_detector=0.05, cone_beam=True).configure(0.0, 1e-4)
    ue, Ue = O.sample_apply_2d(u0, spec, dx, dx, wl, cut, geometry=geom)
    u_flat, _ = O.sample_apply_2d(u0, spec, dx, dx, wl, cut)
    assert O.relative_l2(ue, u_flat) > 1e-3        # the scaling matters at this geometry
    p = rs.SimParams(N=n * n, ny=n, dx=dx, z_detector=0.05, detector_size=n * dx,
                     detector_pixel_size_x=dx, detector_pixel_size_y=dx, wl=wl, chunk_size=n,
                     use_cone_beam_bpm=True)
    p.configure_fresnel_detector(z_source=0.0, z_sample=1e-4)
    cases = [(np.complex64, 3e-6)] if engine == "fused" else [(np.complex128, 1e-12),
                                                              (np.complex64, 3e-6)]
    for dt, tol in cases:
        s = rs.Sample(z_start=spec.z_start, pixel_size_x=dx, pixel_size_z=spec.pixel_size_z,
                      grid=grid, materials=[rs.Material(str(i), 1.0) for i in range(3)],
                      x_positions=np.array([0.0]), engine=engine)
        s.db_list = DB3 * 30
        u, U = dev(u0.astype(dt)), rs.DeviceVector(n * n, dt)
        say("cb-bpm sample_apply", engine, dt.__name__)
        s.apply(u, U, p, cut, 0, None)
        assert O.relative_l2(host(u, n, n), ue) < tol, (engine, dt)
        assert O.relative_l2(host(U, n, n), Ue) < tol, (engine, dt)


def test_cb_bpm_fused_deep_sample_matches_plain_c128():
    """4096-class depth policy on a 512^2 field: 40 scaled slices through the
    fused engine (per-slice tables) versus the plain complex128 engine."""
    n, nz = 512, 40
    grid, spec, u0, dx = _cb_case(n, nz, 43)
    wl, cut = 0.5e-10, 0.4 / dx
    p = rs.SimParams(N=n * n, ny=n, dx=dx, z_detector=0.05, detector_size=n * dx,
                     detector_pixel_size_x=dx, detector_pixel_size_y=dx, wl=wl, chunk_size=n,
                     use_cone_beam_bpm=True)
    p.configure_fresnel_detector(z_source=0.0, z_sample=1e-4)
    out = {}
    for engine, dt in (("plain", np.complex128), ("fused", np.complex64)):
        s = rs.Sample(z_start=spec.z_start, pixel_size_x=dx, pixel_size_z=spec.pixel_size_z,
                      grid=grid, materials=[rs.Material(str(i), 1.0) for i in range(3)],
                      x_positions=np.array([0.0]), engine=engine)
        s.db_list = DB3 * 30
        u, U = dev(u0.astype(dt)), rs.DeviceVector(n * n, dt)
        say("cb-bpm deep", engine, dt.__name__)
        s.apply(u, U, p, cut, 0, None)
        out[engine] = host(u, n, n)
    assert O.relative_l2(out["fused"], out["plain"]) < 1e-5


def _end_to_end(mode, dt):
    n, nz = 256, 6
    grid, spec, u0, dx = _cb_case(n, nz, 47)
    wl, cut = 0.5e-10, 0.4 / dx
    zs, z_det = -0.02, 0.35
    spec.z_start = 0.03
    flags = {"use_cone_beam_bpm": True} if mode == "cb" else {"use_fresnel_scaling": True}
    p = rs.SimParams(N=n * n, ny=n, dx=dx, z_detector=z_det, detector_size=n * dx * 4.0,
                     detector_size_y=n * dx * 3.5, detector_pixel_size_x=3.1 * dx,
                     detector_pixel_size_y=2.7 * dx, wl=wl, chunk_size=n, **flags)
    p.configure_fresnel_detector(z_source=zs, z_sample=spec.z_start)
    geom = O.FresnelGeometry(z_detector=z_det, cone_beam=(mode == "cb")).configure(zs, spec.z_start)
    xs, ys = 3.1e-6, -1.7e-6
    exp = O.run_point_source_sample(n, n, dx, dx, wl, (xs, ys, zs), cut, z_det,
                                    p.get_detector_size_x(), p.get_detector_size_y(),
                                    p.detector_pixel_size_x, p.detector_pixel_size_y,
                                    sample=spec, geometry=geom)
    s = rs.Sample(z_start=spec.z_start, pixel_size_x=dx, pixel_size_z=spec.pixel_size_z,
                  grid=grid, materials=[rs.Material(str(i), 1.0) for i in range(3)],
                  x_positions=np.array([0.0]))
    table = [(rs.Material(str(i), 1.0), DB3[i + 1] * 30) for i in range(3)]
    u, U = rs.DeviceVector(n * n, dt), rs.DeviceVector(n * n, dt)
    angle = float(np.arcsin(cut * wl))
    say("fresnel end-to-end", mode, dt.__name__)
    det = rs.run_simulation(p, rs.PointSource(x=xs, y=ys, z=zs), [s], [angle, angle], u, U,
                            table)[0]
    meta = p.detector_last_metadata["effective_geometry"]
    assert meta["fresnel_scaled"] and meta["magnification"] == pytest.approx(geom.magnification)
    return det, exp


@pytest.mark.parametrize("mode", ["thin", "cb"])
@pytest.mark.parametrize("dt,tol", [(np.complex128, 1e-11), (np.complex64, 2e-5)])
def test_fresnel_end_to_end_matches_oracle(mode, dt, tol):
    det, exp = _end_to_end(mode, dt)
    assert det.shape == exp.shape
    assert O.relative_l2(det, exp) < tol, (mode, dt)
