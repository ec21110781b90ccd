"""The detector entry point that allocates nothing, against the one that does.

rsim_detector_area_v1_accumulate_noalloc keeps its plan and x-integrals in a
static device array instead of allocating.  Two calls in a row -- a large
detector, then a smaller field and detector, so the second runs over what the
first left in the array -- each equal rsim_detector_area_v1_accumulate bit for
bit, for complex64 and complex128 fields; a detector beyond the array's size
is refused with ValueError.
"""

import ctypes as C
import sys

import numpy as np
import pytest

from _rs import rs

pytestmark = pytest.mark.gpu
lib = rs._lib


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def detector(name, f, nx, ny, dx, ox, oy, ds, z, weight, base):
    buf = C.c_void_p()
    lib.call("rsim_malloc", C.byref(buf), ox * oy * 8)
    lib.call("rsim_memcpy_h2d", buf, base.ctypes.data, base.nbytes, None)
    lib.call(name, f.ptr, nx, ny, f.code, dx, dx, ox, oy, ds, ds, z, weight, buf, None)
    got = np.empty((oy, ox))
    lib.call("rsim_memcpy_d2h", got.ctypes.data, buf, got.nbytes, None)
    lib.call("rsim_free", buf)
    return got


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_noalloc_detector_matches_allocating_one_twice(dt):
    rng = np.random.default_rng(5)
    cases = [(4096, 4096, 0.5e-6, 1024, 1024, 2e-6, 0.7),   # 32 MiB of x-integrals
             (768, 512, 0.8e-6, 100, 64, 3.1e-6, 0.3)]      # smaller, after the large one
    for nx, ny, dx, ox, oy, ds, w in cases:
        say("no-alloc detector", np.dtype(dt).name, nx, ny, ox, oy)
        f = rs.DeviceVector.from_numpy((rng.normal(size=nx * ny)
                                        + 1j * rng.normal(size=nx * ny)).astype(dt))
        base = rng.uniform(0, 1, (oy, ox))
        ref = detector("rsim_detector_area_v1_accumulate", f, nx, ny, dx, ox, oy, ds, 0.5, w, base)
        got = detector("rsim_detector_area_v1_accumulate_noalloc", f, nx, ny, dx, ox, oy, ds, 0.5,
                       w, base)
        assert np.array_equal(got, ref)
        assert not np.array_equal(ref, base)
        f.drop()


def test_noalloc_detector_refuses_what_does_not_fit():
    cap = C.c_size_t()
    lib.call("rsim_detector_noalloc_capacity", C.byref(cap))
    assert cap.value == 64 << 20
    nx, ny = 256, 16384
    f = rs.DeviceVector(nx * ny, np.complex64)
    buf = C.c_void_p()
    ox, oy = 1024, 4096                   # 16384 rows x 1024 doubles = 128 MiB > 64 MiB
    lib.call("rsim_malloc", C.byref(buf), ox * oy * 8)
    with pytest.raises(ValueError, match="scratch"):
        lib.call("rsim_detector_area_v1_accumulate_noalloc", f.ptr, nx, ny, f.code, 1e-6, 1e-6,
                 ox, oy, 0.25e-6, 4e-6, 0.5, 1.0, buf, None)
    lib.call("rsim_free", buf)
