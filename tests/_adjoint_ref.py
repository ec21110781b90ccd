"""Shared reference helpers of the differentiable-path tests.

``torch_intensity`` is an independent complex128 torch forward of the
per-voxel multislice + leg + area_v1 detector; its autograd gradient is the
CPU reference the HIP adjoint is checked against.
"""

import numpy as np
import torch

import oracle as O


def case(n=16, nz=3, seed=6, ny=None, dy=0.5e-6):
    """Seeded (nz, ny, n) volumes and an (ny, n) field; ny defaults to n."""
    ny = n if ny is None else ny
    rng = np.random.default_rng(seed)
    delta = rng.uniform(0, 1e-6, (nz, ny, n)).astype(np.float32)
    beta = rng.uniform(0, 1e-9, (nz, ny, n)).astype(np.float32)
    u0 = np.exp(1j * rng.uniform(0, 0.3, (ny, n))) * (1 + 0.05 * rng.normal(size=(ny, n)))
    geo = dict(dx=0.5e-6, dy=dy, wl=0.5e-10, pz=2e-6, cutoff_freq=0.4 / 0.5e-6, z_leg=2e-4,
               size_x=n * 0.5e-6, size_y=ny * dy, ds_x=2 * 0.5e-6, ds_y=2 * dy, z_det=0.2)
    return delta, beta, u0, geo


def loss_and_grad_img(delta, beta, u0, geo, target):
    I = O.multislice_db_intensity(u0, delta, beta, geo["dx"], geo["dy"], geo["wl"], geo["pz"],
                                  geo["cutoff_freq"], geo["z_leg"], geo["size_x"], geo["size_y"],
                                  geo["ds_x"], geo["ds_y"], geo["z_det"])
    return float(np.sum((I - target) ** 2)), 2 * (I - target)


def oracle_gradient(delta, beta, u0, geo, gI):
    return O.multislice_db_gradient(u0, delta, beta, geo["dx"], geo["dy"], geo["wl"], geo["pz"],
                                    geo["cutoff_freq"], geo["z_leg"], geo["size_x"],
                                    geo["size_y"], geo["ds_x"], geo["ds_y"], geo["z_det"], gI)


def torch_intensity(delta, beta, u0, geo):
    """Independent complex128 torch forward (autograd reference)."""
    nz, ny, nx = delta.shape
    K = 2 * np.pi * geo["pz"] / geo["wl"]

    def H(dz):
        h = O.chirp_2d(nx, ny, geo["dx"], geo["dy"], geo["wl"], dz)
        k2 = O.frequency_squared(ny, geo["dy"])[:, None] + O.frequency_squared(nx, geo["dx"])[None, :]
        h = np.where(k2 > 2 * geo["cutoff_freq"] ** 2, 0, h)
        return torch.from_numpy(h)

    Hs, Hl = H(geo["pz"]), H(geo["z_leg"])
    u = torch.from_numpy(u0.astype(np.complex128))
    for s in range(nz):
        u = u * torch.exp(-K * (beta[s] + 1j * delta[s]))
        u = torch.fft.ifft2(torch.fft.fft2(u) * Hs)
    v = torch.fft.ifft2(torch.fft.fft2(u) * Hl)
    Wy, Wx, cos = O.area_v1_matrices(nx, ny, geo["dx"], geo["dy"], geo["size_x"], geo["size_y"],
                                     geo["ds_x"], geo["ds_y"], geo["z_det"])
    inten = v.real ** 2 + v.imag ** 2
    return torch.from_numpy(cos) * (torch.from_numpy(Wy) @ inten @ torch.from_numpy(Wx).T)


def torch_gradient(delta, beta, u0, geo, target):
    d = torch.tensor(delta.astype(np.float64), requires_grad=True)
    b = torch.tensor(beta.astype(np.float64), requires_grad=True)
    I = torch_intensity(d, b, u0, geo)
    loss = ((I - torch.from_numpy(target)) ** 2).sum()
    loss.backward()
    return d.grad.numpy(), b.grad.numpy(), I.detach().numpy()
