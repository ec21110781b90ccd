"""Batched fused multislice (one captured hipGraph over many instances).

Each instance of a batch must equal, bit for bit, the same instance run on
its own through rsim_sample_apply_2d with the fused engine: the batched
launches run the same kernels with per-instance descriptors.
"""

import ctypes as C
import importlib
import sys

import numpy as np
import pytest

from _rs import rs

pytestmark = pytest.mark.gpu
lib = rs._lib
B = importlib.import_module("3d_ravesim_amd.batch")
DB3 = np.array([0, 1.1e-6 + 3e-9j, 3.5e-7 + 2e-10j])


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def spheres(rng, nz, ny, nx, k):
    g = np.zeros((nz, ny, nx), np.uint8)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for _ in range(k):
        c = rng.uniform(0, 1, 3) * [nz, ny, nx]
        r = rng.uniform(0.1, 0.3) * min(ny, nx)
        g[(zz - c[0]) ** 2 * 16 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2 < r * r] = rng.integers(1, 3)
    return g


def make_items(rng, nz):
    """Instances: 18 at 256^2 (two launches of <= 16), 3 at 512^2, 2 at 256x512,
    2 at 1024^2 on a non-integral voxel map (bilinear transmission)."""
    shapes = [(256, 256)] * 18 + [(512, 512)] * 3 + [(256, 512)] * 2 + [(1024, 1024)] * 2
    vols = {}
    items = []
    for k, (ny, nx) in enumerate(shapes):
        bilinear = nx == 1024
        key = (ny, nx)
        if key not in vols:
            gny, gnx = ((ny // 2 + 3, nx // 2 + 5) if bilinear else (ny, nx))
            vols[key] = rs.DeviceVolume.from_numpy(spheres(rng, nz, gny, gnx, 5), DB3 * 20)
        dx = 0.5e-6 * (1 + 0.1 * (k % 3))
        wl = (0.4 + 0.05 * (k % 5)) * 1e-10
        u0 = (np.exp(1j * rng.uniform(0, 0.3, size=(ny, nx))) *
              (1 + 0.02 * rng.normal(size=(ny, nx)))).astype(np.complex64).reshape(-1)
        it = B.BatchItem(u=rs.DeviceVector.from_numpy(u0), U=rs.DeviceVector(nx * ny, np.complex64),
                         nx=nx, ny=ny, dx=dx, wl=wl, cutoff_freq=0.4 / dx, volume=vols[key],
                         pixel_size_x=(2.13 * dx if bilinear else dx), pixel_size_z=0.5e-6,
                         x_pos=(k % 4) * dx if not bilinear else 0.3 * dx)
        items.append((it, u0))
    return items, vols


def individual(it, u0, first, n, flags=0):
    u = rs.DeviceVector.from_numpy(u0)
    U = rs.DeviceVector(it.nx * it.ny, np.complex64)
    lib.call("rsim_sample_apply_2d", u.ptr, U.ptr, it.nx, it.ny, lib.RSIM_C64, it.dx, it.dx, it.wl,
             it.cutoff_freq, it.volume.handle, it.pixel_size_x, it.pixel_size_x, it.pixel_size_z,
             it.x_pos, it.y_pos, first, n, lib.RSIM_FUSED, flags, None)
    return u.to_numpy(), U.to_numpy()


def test_batch_equals_individual_runs_bitwise():
    rng = np.random.default_rng(31)
    nz, first, n = 5, 1, 4
    items, vols = make_items(rng, nz)
    say("batch create", len(items), "instances")
    b = B.MultisliceBatch([it for it, _ in items], first, n)
    assert b.groups == 4                   # 256^2, 512^2, 256x512, 1024^2 (bilinear)
    assert b.kernels == (n + 1) * 5 + n * 5  # 5 launches per pass (256^2 needs two)
    say("batch run")
    b.run()
    for k, (it, u0) in enumerate(items):
        say("instance", k, it.ny, it.nx)
        eu, eU = individual(it, u0, first, n)
        assert np.array_equal(it.u.to_numpy(), eu), k
        assert np.array_equal(it.U.to_numpy(), eU), k
    b.close()


def test_batch_graph_replays_match_eager_results():
    rng = np.random.default_rng(32)
    nz = 3
    items, vols = make_items(rng, nz)
    items = items[16:]                     # 2 x 256^2, 3 x 512^2, 2 x 256x512, 2 x 1024^2
    b = B.MultisliceBatch([it for it, _ in items], 0, nz)
    expect = [individual(it, u0, 0, nz) for it, u0 in items]
    for rep in range(3):
        say("replay", rep)
        for it, u0 in items:
            it.u.upload(u0)
        # eager work of the program between replays
        scratch = rs.DeviceVector.from_numpy(np.ones(256 * 256, np.complex64))
        scratch.fft2(scratch, 256, 256)
        lib.call("rsim_synchronize", None)
        b.run()
        for k, (it, _) in enumerate(items):
            assert np.array_equal(it.u.to_numpy(), expect[k][0]), (rep, k)
            assert np.array_equal(it.U.to_numpy(), expect[k][1]), (rep, k)
    b.close()


def test_batch_without_spectrum_leaves_u_identical():
    rng = np.random.default_rng(33)
    items, vols = make_items(rng, 2)
    items = items[18:21]
    b = B.MultisliceBatch([it for it, _ in items], 0, 2, flags=lib.RSIM_FLAG_NO_SPECTRUM)
    b.run()
    for it, u0 in items:
        eu, _ = individual(it, u0, 0, 2)
        assert np.array_equal(it.u.to_numpy(), eu)
    b.close()
