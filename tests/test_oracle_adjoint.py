"""CPU pins of the oracle's per-voxel multislice, area_v1 adjoint and gradient.

The gradient path has no counterpart in the reference, so its oracle is
pinned by mathematics: the dense detector form against area_v1 (itself
pinned by brute force), the dot-product identity of the adjoint, central
finite differences in FP64, and complex128 torch autograd of an independent
torch forward.
"""

import numpy as np
import pytest
import oracle as O
from _adjoint_ref import case, loss_and_grad_img, oracle_gradient, torch_gradient


@pytest.mark.parametrize("shape,ds", [((7, 11), (1.3, 1.7)), ((16, 16), (2.0, 2.0)),
                                      ((12, 10), (0.6, 0.45))])
def test_dense_area_matrices_reproduce_area_v1(shape, ds):
    rng = np.random.default_rng(3)
    ny, nx = shape
    f = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    size_x, size_y = 0.8 * nx, 0.9 * ny
    Wy, Wx, cos = O.area_v1_matrices(nx, ny, 1.0, 1.1, size_x, size_y, ds[0], ds[1], 5.0)
    dense = cos * (Wy @ np.abs(f) ** 2 @ Wx.T)
    ref = O.area_v1(f, 1.0, 1.1, size_x, size_y, ds[0], ds[1], 5.0)
    assert O.relative_l2(dense, ref) < 1e-13


def test_area_adjoint_dot_product_identity():
    rng = np.random.default_rng(4)
    ny, nx = 9, 13
    x = rng.uniform(0.1, 2.0, (ny, nx))          # an intensity |v|^2
    v = np.sqrt(x).astype(np.complex128)
    args = (1.0, 1.0, 0.85 * nx, 0.9 * ny, 1.7, 1.4, 3.0)
    Ax = O.area_v1(v, *args)
    gI = rng.normal(size=Ax.shape)
    ATg = O.area_v1_adjoint(np.ones((ny, nx), np.complex128), gI, *args).real
    assert abs(np.sum(gI * Ax) - np.sum(ATg * x)) < 1e-12 * np.sum(np.abs(gI * Ax))


def test_per_voxel_multislice_equals_sample_apply_on_id_volume():
    rng = np.random.default_rng(5)
    n, nz = 32, 3
    dx = 2.0 ** -21                      # exactly integral voxel map
    grid = rng.integers(0, 3, (nz, n, n)).astype(np.uint8)
    f32 = lambda v: float(np.float32(v))  # noqa: E731  (values exact in float32)
    db = np.array([0, f32(3e-6) + 1j * f32(2e-9), f32(1e-6) + 1j * f32(5e-10)])
    spec = O.SampleSpec(grid=grid, db_list=db, pixel_size_x=dx, pixel_size_y=dx, pixel_size_z=1e-6)
    u0 = np.exp(1j * rng.uniform(0, 0.2, (n, n)))
    ue, _ = O.sample_apply_2d(u0, spec, dx, dx, 0.5e-10, 0.4 / dx)
    # the id map reads voxel (y0, x0) = (y, x); the last row/column is outside
    vol = db[grid]
    vol[:, -1, :] = 0
    vol[:, :, -1] = 0
    delta = vol.real.astype(np.float32)
    beta = vol.imag.astype(np.float32)
    got = O.multislice_db_2d(u0, delta, beta, dx, dx, 0.5e-10, 1e-6, 0.4 / dx)
    assert O.relative_l2(got, ue) < 1e-14


def test_oracle_gradient_matches_central_finite_differences():
    delta, beta, u0, geo = case()
    rng = np.random.default_rng(7)
    target = O.multislice_db_intensity(u0, delta * 1.05, beta * 0.95, geo["dx"], geo["dy"],
                                       geo["wl"], geo["pz"], geo["cutoff_freq"], geo["z_leg"],
                                       geo["size_x"], geo["size_y"], geo["ds_x"], geo["ds_y"],
                                       geo["z_det"])
    _, gI = loss_and_grad_img(delta, beta, u0, geo, target)
    gd, gb = oracle_gradient(delta, beta, u0, geo, gI)
    d64, b64 = delta.astype(np.float64), beta.astype(np.float64)
    for _ in range(6):
        s, y, x = (int(rng.integers(0, k)) for k in delta.shape)
        for arr, grad, h in ((d64, gd, 1e-9), (b64, gb, 1e-12)):
            old = arr[s, y, x]
            arr[s, y, x] = old + h
            lp, _ = loss_and_grad_img(d64, b64, u0, geo, target)
            arr[s, y, x] = old - h
            lm, _ = loss_and_grad_img(d64, b64, u0, geo, target)
            arr[s, y, x] = old
            fd = (lp - lm) / (2 * h)
            scale = np.abs(grad).max()
            assert abs(fd - grad[s, y, x]) < 1e-5 * scale, (s, y, x, fd, grad[s, y, x])


def test_oracle_gradient_matches_torch_complex128_autograd():
    delta, beta, u0, geo = case(n=24, nz=4, seed=8)
    target = O.multislice_db_intensity(u0, delta * 0.97, beta * 1.04, geo["dx"], geo["dy"],
                                       geo["wl"], geo["pz"], geo["cutoff_freq"], geo["z_leg"],
                                       geo["size_x"], geo["size_y"], geo["ds_x"], geo["ds_y"],
                                       geo["z_det"])
    td, tb, tI = torch_gradient(delta, beta, u0, geo, target)
    _, gI = loss_and_grad_img(delta, beta, u0, geo, target)
    I = target + gI / 2
    assert O.relative_l2(tI, I) < 1e-12
    gd, gb = oracle_gradient(delta, beta, u0, geo, gI)
    assert O.relative_l2(gd, td) < 1e-10
    assert O.relative_l2(gb, tb) < 1e-10
