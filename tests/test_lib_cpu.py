"""CPU-side checks of the C ABI: the library loads, exports every declared
symbol, and rejects bad arguments before touching a device."""

import ctypes as C

import numpy as np
import pytest

from _rs import header_functions, lib, rs


def test_library_loads_and_exports_every_declared_symbol():
    L = lib.load()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
        assert n in lib.SIGNATURES, f"{n} missing from the ctypes signature table"
    assert set(lib.SIGNATURES) == set(names)
    assert b"gfx950" in L.rsim_version()


def test_invalid_arguments_map_to_value_error_without_a_device():
    L = lib.load()
    assert L.rsim_fft2(None, None, 4, 4, lib.RSIM_C64, lib.RSIM_FORWARD, None) == lib.RSIM_E_INVALID_ARG
    assert b"NULL" in L.rsim_last_error()
    with pytest.raises(ValueError):
        lib.call("rsim_fft2", C.c_void_p(16), C.c_void_p(16), 7, 4, lib.RSIM_C64, lib.RSIM_FORWARD, None)
    with pytest.raises(ValueError):
        lib.call("rsim_fft2", C.c_void_p(16), C.c_void_p(16), 4, 4, 9, lib.RSIM_FORWARD, None)
    with pytest.raises(ValueError):
        lib.call("rsim_propagate_2d", C.c_void_p(16), C.c_void_p(16), 4, 4, lib.RSIM_C128,
                 1.0, 1.0, -1.0, 1.0, 1.0, 0, None)
    vol = C.c_void_p()
    db = np.array([1.0, 0.0], dtype=np.float64)
    ids = np.zeros((1, 2, 2), dtype=np.uint8)
    rc = L.rsim_volume_create(C.byref(vol), ids.ctypes.data, 1, 1, 2, 2, db.ctypes.data_as(lib._dp), 1)
    assert rc == lib.RSIM_E_INVALID_ARG and b"vacuum" in L.rsim_last_error()
    rc = L.rsim_volume_create(C.byref(vol), ids.ctypes.data, 3, 1, 2, 2, db.ctypes.data_as(lib._dp), 1)
    assert rc == lib.RSIM_E_INVALID_ARG
    # material ids are checked on the host before any device allocation
    db2 = np.zeros(4, dtype=np.float64)
    ids2 = np.array([[[0, 1], [2, 3]]], dtype=np.uint16)
    rc = L.rsim_volume_create(C.byref(vol), ids2.ctypes.data, 2, 1, 2, 2, db2.ctypes.data_as(lib._dp), 2)
    assert rc == lib.RSIM_E_INVALID_ARG and b"material id 2 at voxel 2" in L.rsim_last_error()
    assert not vol.value
    with pytest.raises(ValueError):
        lib.call("rsim_kernel_timing", -1)
    b = C.c_void_p()
    with pytest.raises(ValueError):
        lib.call("rsim_batch_create", C.byref(b), None, 0, 0, 1, 0, None)
    assert L.rsim_batch_destroy(None) == 0
    assert L.rsim_kernel_timing(16) == 0 and L.rsim_kernel_timing(0) == 0


def test_cancel_flag_roundtrip_and_progress_registration():
    L = lib.load()
    assert L.rsim_request_cancel() == 0
    assert L.rsim_clear_cancel() == 0
    seen = []
    cb = lib.PROGRESS_FN(lambda s, a, b, u: seen.append((s, a, b)))
    assert L.rsim_set_progress(cb, None) == 0
    assert L.rsim_set_progress(lib.PROGRESS_FN(), None) == 0


def test_simparams_mirrors_reference_validation():
    with pytest.raises(ValueError):
        rs.SimParams(N=100, nx=10, ny=11, dx=1e-6, z_detector=1, detector_size=1e-4,
                     detector_pixel_size_x=1e-6, detector_pixel_size_y=1e-6, wl=1e-10, chunk_size=10)
    p = rs.SimParams(N=64 * 32, ny=32, dx=1e-6, z_detector=1, detector_size=6.4e-5,
                     detector_pixel_size_x=1e-6, detector_pixel_size_y=1e-6, wl=1e-10, chunk_size=64)
    assert p.nx == 64 and p.get_dy() == 1e-6 and p.get_detector_size_y() == 32e-6
    with pytest.raises(ValueError):
        rs.require_row_aligned_chunk_size(65, 64, 32)
    assert rs.require_row_aligned_chunk_size(128, 64, 32) == 2


def test_energy_wavelength_matches_reference_constants():
    # big-wave/propagation.py:199-201,260-262; 24.7968 keV <-> 0.5 Angstrom
    wl = rs.convert_energy_wavelength(24796.8)
    assert abs(wl - 0.5e-10) / 0.5e-10 < 1e-4
    assert rs.convert_energy_wavelength(rs.convert_energy_wavelength(8000.0)) == pytest.approx(8000.0, rel=1e-14)


def test_new_entry_points_reject_bad_arguments_without_a_device():
    """Argument checks of the Fresnel, plasma and history entry points run
    before any device call (no GPU here)."""
    L = lib.load()
    f = C.c_void_p(16)
    # tilted plane: non-positive distance
    assert L.rsim_tilted_plane_2d(f, 4, 4, lib.RSIM_C64, 1e-6, 1e-6, 1e-10, 0.0, 0.0, 0.0,
                                  None) == lib.RSIM_E_INVALID_ARG
    assert b"distance" in L.rsim_last_error()
    # plasma: grid pointers, element size, dimensions
    p = C.c_void_p()
    g = np.zeros(8)
    ptr = g.ctypes.data
    assert L.rsim_plasma_create(C.byref(p), None, ptr, ptr, ptr, 8, 1, 2, 4) == lib.RSIM_E_INVALID_ARG
    assert L.rsim_plasma_create(C.byref(p), ptr, ptr, ptr, ptr, 2, 1, 2, 4) == lib.RSIM_E_INVALID_ARG
    assert b"elem_size" in L.rsim_last_error()
    assert L.rsim_plasma_create(C.byref(p), ptr, ptr, ptr, ptr, 8, 0, 2, 4) == lib.RSIM_E_INVALID_ARG
    assert L.rsim_plasma_destroy(None) == 0
    opt = np.array([1.0, 0.0, 1.0, 1.0])
    assert L.rsim_plasma_deltabeta(None, 0, 1, 8000.0, opt.ctypes.data_as(lib._dp), f,
                                   None) == lib.RSIM_E_INVALID_ARG
    assert L.rsim_plasma_apply_2d(f, f, 4, 4, lib.RSIM_C64, 1e-6, 1e-6, 1e-10, 1e5, None, 8000.0,
                                  opt.ctypes.data_as(lib._dp), 1e-6, 1e-6, 1e-6, 0.0, 0.0, 0, 1,
                                  None, None, lib.RSIM_FUSED, 0, None) == lib.RSIM_E_INVALID_ARG
    # history: a NULL spec is refused
    assert L.rsim_sample_apply_2d_history(f, f, 4, 4, lib.RSIM_C64, 1e-6, 1e-6, 1e-10, 1e5, None,
                                          1e-6, 1e-6, 1e-6, 0.0, 0.0, 0, 1, None, None,
                                          lib.RSIM_FUSED, 0, None, None) == lib.RSIM_E_INVALID_ARG
    assert b"history" in L.rsim_last_error()
