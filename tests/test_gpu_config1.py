"""The reference's own CPU case (1024^2 grating) on the GPU against the oracle.

Field 1024^2, pixel 1 um, lambda 0.5 A.  A point source 1 m upstream gives
u0 = exp(2 pi i r / lambda) / r (big-wave/propagation.py:474-482;
PointSource.propagate_to would reject this grid as under-sampled, so the
analytic field is used directly).  A binary phase grating along x (period
4.2 um, duty 0.5, mask = mod(x / p + dc / 2, 1) < dc with x = i dx,
big-wave/optical_element.py:354-360; delta 2.3e-6, beta 1e-9, 5 um) is one
Sample slice on the integral voxel map (grid (1, ny + 2, nx + 2), voxel pitch
dx), then 0.1 m of free space (cutoff 0.4 / dx) and an area_v1 detector with
4 um pixels (256^2).

Relative L2 bounds, from the oracle's u0 (field and image) and end to end
from the GPU's own source (image): complex128 on the plain engine 1e-12,
complex64 on the fused engine 3e-6.
"""

import sys

import numpy as np
import pytest

import oracle as O
from _rs import rs

pytestmark = pytest.mark.gpu
N, DX, WL = 1024, 1e-6, 0.5e-10
Z_SRC, THICK, LEG = 1.0, 5e-6, 0.1
DB = np.array([0, 2.3e-6 + 1e-9j])
CUT = 0.4 / DX


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def grating_grid(n, dx, period=4.2e-6, dc=0.5):
    """Field column i is grid column i + 1 (integral map, margin 2)."""
    x = (np.arange(n + 2) - 1) * dx
    mask = np.mod(x / period + dc / 2, 1) < dc
    g = np.zeros((1, n + 2, n + 2), np.uint8)
    g[0][:, mask] = 1
    return g


def params():
    return rs.SimParams(N=N * N, nx=N, ny=N, dx=DX, dy=DX, z_detector=Z_SRC + THICK + LEG,
                        detector_size=N * DX, detector_pixel_size_x=4 * DX,
                        detector_pixel_size_y=4 * DX, wl=WL, chunk_size=N)


@pytest.fixture(scope="module")
def expected():
    grid = grating_grid(N, DX)
    u0 = O.propagate_analytically_2d(N, N, DX, DX, Z_SRC, 0.0, 0.0, WL)
    spec = O.SampleSpec(grid=grid, db_list=DB, pixel_size_x=DX, pixel_size_z=THICK, z_start=Z_SRC)
    us, Us = O.sample_apply_2d(u0, spec, DX, DX, WL, CUT)
    u1, _ = O.propagate_2d(us, DX, DX, WL, LEG, CUT, U2d=Us)
    img = O.area_v1(u1, DX, DX, N * DX, N * DX, 4 * DX, 4 * DX, Z_SRC + THICK + LEG)
    return grid, u0, u1, img


def gpu_pipeline(grid, dt, engine, u0=None):
    p = params()
    if u0 is None:
        u = rs.DeviceVector(N * N, dt)
        rs.propagate_analytically_2d(u, Z_SRC, 0.0, 0.0, p)
    else:
        u = rs.DeviceVector.from_numpy(np.ascontiguousarray(u0.astype(dt)).reshape(-1))
    U = rs.DeviceVector(N * N, dt)
    s = rs.Sample(z_start=Z_SRC, pixel_size_x=DX, pixel_size_z=THICK, grid=grid,
                  materials=[rs.Material("grating", 1.0)], x_positions=np.array([0.0]),
                  engine=engine)
    s.db_list = DB
    s.apply(u, U, p, CUT, 0, None)
    rs.propagate_2d(u, U, DX, DX, WL, LEG, N, CUT, N, N, skip_fft=True)
    img = rs.square_and_downsample_2d(u, p, Z_SRC + THICK + LEG)
    return u.to_numpy().reshape(N, N), img


@pytest.mark.parametrize("dt,engine,tol", [(np.complex128, "plain", 1e-12),
                                           (np.complex64, "fused", 3e-6)])
def test_grating_pipeline_from_oracle_source(expected, dt, engine, tol):
    grid, u0, u1, img = expected
    say("config-1 pipeline from the oracle source", np.dtype(dt).name, engine)
    got_u, got_img = gpu_pipeline(grid, dt, engine, u0)
    assert got_img.shape == (256, 256)
    ef, ei = O.relative_l2(got_u, u1), O.relative_l2(got_img, img)
    say(f"  field {ef:.3e} image {ei:.3e}")
    assert ef < tol and ei < tol


@pytest.mark.parametrize("dt,engine,tol", [(np.complex128, "plain", 1e-12),
                                           (np.complex64, "fused", 3e-6)])
def test_grating_pipeline_end_to_end(expected, dt, engine, tol):
    grid, u0, u1, img = expected
    say("config-1 pipeline end to end", np.dtype(dt).name, engine)
    _, got_img = gpu_pipeline(grid, dt, engine)
    ei = O.relative_l2(got_img, img)
    say(f"  image {ei:.3e}")
    assert ei < tol
