"""Chunked Vector access on device fields, and tile-row invariance.

The reference's Vector protocol walks a field in row tiles
(big-wave/vector.py:349-371; tiles of complete rows,
big-wave/propagation.py:265-290): write_chunked fills fresh chunks,
modify_chunked hands out chunks that are mutated in place, read_chunked
hands out read-only chunks.  DeviceVector keeps that contract over HBM.  Its
operators never iterate tiles on the device, so their results are the same,
bit for bit, whatever tile size the parameters carry (the reference's
tests/big_wave_2d/test_p1.py:90-140 asserts equality for the source and the
cutoff and 2e-7 for propagation).
"""

import sys

import numpy as np
import pytest

import oracle as O
from _rs import rs

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def params(nx, ny, rows):
    return rs.SimParams(N=nx * ny, nx=nx, ny=ny, dx=0.5e-6, dy=0.7e-6, z_detector=1.0,
                        detector_size=nx * 0.5e-6, detector_pixel_size_x=1e-6,
                        detector_pixel_size_y=1e-6, wl=1e-10, chunk_size=rows * nx)


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("chunk", [24, 72, 100, 768])
def test_chunked_write_modify_read_round_trip(dt, chunk):
    nx, ny = 32, 24
    n = nx * ny
    say("chunked access", np.dtype(dt).name, chunk)
    v = rs.DeviceVector(n, dt)
    starts = []

    def fill(i, c):
        starts.append(i)
        assert c.dtype == dt and not c.any()          # fresh chunks start at zero
        c[:] = np.arange(i, i + len(c)) * (1 + 0.5j)

    v.write_chunked(chunk, fill)
    assert starts == list(range(0, n, chunk))
    expected = (np.arange(n) * (1 + 0.5j)).astype(dt)
    np.testing.assert_array_equal(v.to_numpy(), expected)

    def bump(i, c):
        c *= 2j
        c += i

    v.modify_chunked(chunk, bump)
    for i in range(0, n, chunk):
        expected[i:i + chunk] = expected[i:i + chunk] * 2j + i
    np.testing.assert_array_equal(v.to_numpy(), expected)

    seen = []

    def read(i, c):
        with pytest.raises(ValueError):
            c[0] = 0
        seen.append((i, c.copy()))

    v.read_chunked(chunk, read)
    got = np.concatenate([c for _, c in seen])
    np.testing.assert_array_equal(got, expected)
    assert [i for i, _ in seen] == list(range(0, n, chunk))


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_source_cutoff_and_propagation_are_tile_rows_invariant(dt):
    nx, ny = 32, 24
    rng = np.random.default_rng(42)
    initial = (rng.normal(size=nx * ny) + 1j * rng.normal(size=nx * ny)).astype(dt)
    say("tile-row invariance", np.dtype(dt).name)
    src, cut, prop = [], [], []
    for rows in (1, 3, 4, 8, 24):
        p = params(nx, ny, rows)
        v = rs.DeviceVector(nx * ny, dt)
        rs.propagate_analytically_2d(v, 0.02, 0.3e-6, -0.2e-6, p)
        src.append(v.to_numpy())
        c = rs.DeviceVector.from_numpy(initial)
        rs.apply_frequency_cutoff_2d(c, 2.0e5, p.dx, p.get_dy(), nx, ny, p.chunk_size)
        cut.append(c.to_numpy())
        u = rs.DeviceVector.from_numpy(initial)
        U = rs.DeviceVector(nx * ny, dt)
        rs.propagate_2d(u, U, p.dx, p.get_dy(), p.wl, 0.003, p.chunk_size, 2.0e5, nx, ny)
        prop.append(u.to_numpy())
    for r in src[1:]:
        np.testing.assert_array_equal(r, src[0])
    for r in cut[1:]:
        np.testing.assert_array_equal(r, cut[0])
    for r in prop[1:]:
        np.testing.assert_array_equal(r, prop[0])
    # and each equals the oracle
    tol = 1e-12 if dt == np.complex128 else 2e-6
    p = params(nx, ny, 1)
    ref_src = O.propagate_analytically_2d(nx, ny, p.dx, p.get_dy(), 0.02, 0.3e-6, -0.2e-6, p.wl)
    assert O.relative_l2(src[0].reshape(ny, nx), ref_src) < tol
    ref_prop, _ = O.propagate_2d(initial.reshape(ny, nx).astype(np.complex128), p.dx, p.get_dy(),
                                 p.wl, 0.003, 2.0e5)
    assert O.relative_l2(prop[0].reshape(ny, nx), ref_prop) < tol


def test_chunk_sizes_must_hold_complete_rows():
    nx, ny = 16, 8
    v = rs.DeviceVector(nx * ny, np.complex64)
    with pytest.raises(ValueError, match="complete rows"):
        rs.apply_frequency_cutoff_2d(v, 1e5, 1e-6, 1e-6, nx, ny, 17)
    with pytest.raises(ValueError, match="must not exceed"):
        rs.apply_frequency_cutoff_2d(v, 1e5, 1e-6, 1e-6, nx, ny, 16 * 9)
