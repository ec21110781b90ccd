"""Tiled intermediate layout of the fused slice chain.

Between the passes of a fused slice the column pass may store its output as
column tiles (element (y, x) at ((x / w) * ny + y) * w + x % w, w = the
column pass's columns per workgroup) that the next row pass reads back in
rows, so the column pass writes whole 128-byte lines.  The layout moves the
same values through the same arithmetic, so a tiled chain equals the in-place
chain bit for bit: for each length, both launch-shape sets, the spectrum
output, history frames and a cancel at a slice boundary.  The schedule is
forced with RAVESIM_FUSED_VARIANT="row,col,tiled".
"""

import ctypes as C
import importlib
import sys

import numpy as np
import pytest

from _rs import rs

pytestmark = pytest.mark.gpu
lib = rs._lib
CTL = importlib.import_module("3d_ravesim_amd._control")
H = importlib.import_module("3d_ravesim_amd.history")
DB = np.array([0, 3.5e-7 + 2e-10j, 1.1e-6 + 3e-9j, 2.6e-7 + 1.5e-10j]) * 30
DX, WL, PZ = 0.5e-6, 0.5e-10, 0.5e-6
ROW_LD2 = 7      # the LD list's three-workgroup row shape
COL_BASE1 = 1    # the column shape without forced split exchange


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def sphere_volume(nz, ny, nx, seed):
    rng = np.random.default_rng(seed)
    n = 12
    cz, cy, cx = rng.uniform(0, nz, n), rng.uniform(0, ny, n), rng.uniform(0, nx, n)
    r = rng.uniform(0.05, 0.2, n) * min(ny, nx)
    vol = rs.DeviceVolume.empty((nz, ny, nx), DB)
    vol.paint_spheres(np.stack([cz, cy, cx, r], axis=1), rng.integers(1, 4, n))
    return vol


def tile_width(nx, ny):
    w = C.c_int()
    lib.call("rsim_fused_tile", nx, ny, C.byref(w))
    return w.value


def run(u0, vol, nz, flags=0):
    ny, nx = u0.shape
    u = rs.DeviceVector.from_numpy(u0.reshape(-1))
    U = rs.DeviceVector(nx * ny, np.complex64)
    lib.call("rsim_sample_apply_2d", u.ptr, U.ptr, nx, ny, u.code, DX, DX, WL, 0.4 / DX,
             vol.handle, DX, DX, PZ, 0.0, 0.0, 0, nz, lib.RSIM_FUSED, flags, None)
    out = u.to_numpy(), U.to_numpy()
    u.drop()
    U.drop()
    return out


SHAPES = [(256, 256), (512, 512), (1024, 1024), (2048, 2048), (3072, 3072), (4096, 4096),
          (6144, 6144), (1024, 2048), (6144, 3072), (3072, 6144)]


@pytest.mark.parametrize("variant", [(0, 0), (ROW_LD2, COL_BASE1)])
@pytest.mark.parametrize("ny,nx", SHAPES)
def test_tiled_chain_is_bit_identical_to_in_place(ny, nx, variant, monkeypatch):
    nz = 3
    vol = sphere_volume(nz, ny, nx, seed=ny + nx)
    rng = np.random.default_rng(5)
    u0 = np.exp(1j * rng.uniform(0, 0.3, (ny, nx))).astype(np.complex64)
    r, c = variant
    monkeypatch.setenv("RAVESIM_FUSED_VARIANT", f"{r},{c},1")
    w = tile_width(nx, ny)
    say("tiled", ny, nx, variant, "tile", w)
    if w == 0:
        vol.destroy()
        pytest.skip(f"column variant {c} at {ny} rows has fewer than 4 columns per workgroup")
    for flags in (0, lib.RSIM_FLAG_MATH_COMP32):
        monkeypatch.setenv("RAVESIM_FUSED_VARIANT", f"{r},{c},1")
        ft, st = run(u0, vol, nz, flags)
        monkeypatch.setenv("RAVESIM_FUSED_VARIANT", f"{r},{c},0")
        assert tile_width(nx, ny) == 0
        fp, sp = run(u0, vol, nz, flags)
        assert np.array_equal(ft, fp), flags
        assert np.array_equal(st, sp), flags        # the spectrum of the result
        assert np.isfinite(ft).all() and np.abs(ft).max() > 0.5
    vol.destroy()


def test_tiled_cancel_leaves_the_field_after_k_slices(monkeypatch):
    n, nz, k = 1024, 6, 2
    vol = sphere_volume(nz, n, n, seed=9)
    rng = np.random.default_rng(6)
    u0 = np.exp(1j * rng.uniform(0, 0.3, (n, n))).astype(np.complex64)
    monkeypatch.setenv("RAVESIM_FUSED_VARIANT", "0,0,1")
    assert tile_width(n, n) > 0
    say("tiled cancel after", k)

    def progress(stage, done, total):
        if done == k:
            ctl.cancel()

    ctl = CTL.RunControl(progress)
    u = rs.DeviceVector.from_numpy(u0.reshape(-1))
    U = rs.DeviceVector(n * n, np.complex64)
    with ctl:
        with pytest.raises(InterruptedError):
            lib.call("rsim_sample_apply_2d", u.ptr, U.ptr, n, n, u.code, DX, DX, WL, 0.4 / DX,
                     vol.handle, DX, DX, PZ, 0.0, 0.0, 0, nz, lib.RSIM_FUSED, 0, None)
    ctl.close()
    got = u.to_numpy()
    ref, _ = run(u0, vol, k)
    assert np.array_equal(got, ref)
    vol.destroy()


def test_tiled_history_frames_match_in_place(monkeypatch):
    rng = np.random.default_rng(41)
    n, nz, dx = 256, 7, DX
    g = np.zeros((nz, n, n), np.uint8)
    yy, xx = np.mgrid[0:n, 0:n]
    for z in range(nz):
        cy, cx, r = rng.uniform(60, 196), rng.uniform(60, 196), rng.uniform(20, 50)
        g[z][(yy - cy) ** 2 + (xx - cx) ** 2 < r * r] = rng.integers(1, 4)
    u0 = np.exp(1j * rng.uniform(0, 0.2, (n, n))).astype(np.complex64)
    p = rs.SimParams(N=n * n, ny=n, dx=dx, z_detector=0.3, detector_size=n * dx * 0.9,
                     detector_size_y=n * dx * 0.8, detector_pixel_size_x=2.3 * dx,
                     detector_pixel_size_y=1.7 * dx, wl=WL, chunk_size=n)
    frames = {}
    for t in (1, 0):
        monkeypatch.setenv("RAVESIM_FUSED_VARIANT", f"0,0,{t}")
        assert (tile_width(n, n) > 0) == bool(t)
        s = rs.Sample(z_start=0.1, pixel_size_x=dx, pixel_size_z=PZ, grid=g,
                      materials=[rs.Material(str(i), 1.0) for i in range(3)],
                      x_positions=np.array([0.0]), engine="fused")
        s.db_list = DB
        u, U = rs.DeviceVector.from_numpy(u0.reshape(-1)), rs.DeviceVector(n * n, np.complex64)
        hist = H.History()
        say("tiled history", t)
        s.apply(u, U, p, 0.4 / dx, 0, hist)
        frames[t] = (hist.get_history(), u.to_numpy())
    assert np.array_equal(frames[1][0], frames[0][0])
    assert np.array_equal(frames[1][1], frames[0][1])
