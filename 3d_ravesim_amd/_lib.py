"""ctypes binding of libravesim_hip.so (the C ABI declared in include/ravesim.h).

The product has no CPU fallback: if the library is not built, or no HIP
device is visible, calls raise instead of computing anything on the host.
Error codes map to the exceptions big-wave raises for the same conditions
(big-wave/propagation.py:265-290 ValueError, :685-686 InterruptedError).
"""

from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RAVESIM_HIP_LIB", os.path.join(HERE, "libravesim_hip.so"))

RSIM_OK = 0
RSIM_E_INVALID_ARG = -1
RSIM_E_HIP = -2
RSIM_E_CANCELLED = -3
RSIM_E_NOMEM = -4
RSIM_E_UNSUPPORTED = -5

RSIM_C64 = 0
RSIM_C128 = 1
RSIM_FORWARD = -1
RSIM_INVERSE = 1
RSIM_PLAIN = 0
RSIM_FUSED = 1
RSIM_FLAG_NO_SPECTRUM = 1
RSIM_FLAG_MATH_FP32 = 2
RSIM_FLAG_MATH_FP64 = 4
RSIM_FLAG_MATH_COMP32 = 8

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_sz = C.c_size_t
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

PROGRESS_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_int64, C.c_int64, C.c_void_p)


class BatchInstance(C.Structure):
    """rsim_batch_instance of include/ravesim.h."""
    _fields_ = [("u", _vp), ("U", _vp), ("nx", _i), ("ny", _i), ("dx", _d), ("dy", _d),
                ("wl", _d), ("cutoff_freq", _d), ("vol", _vp), ("pixel_size_x", _d),
                ("pixel_size_y", _d), ("pixel_size_z", _d), ("x_pos", _d), ("y_pos", _d)]


class ChainStep(C.Structure):
    """rsim_chain_step of include/ravesim.h."""
    _fields_ = [("vol", _vp), ("pixel_size_x", _d), ("pixel_size_y", _d), ("pixel_size_z", _d),
                ("x_pos", _d), ("y_pos", _d), ("first_slice", _i), ("n_slices", _i), ("dz", _d)]


class ChainInstance(C.Structure):
    """rsim_chain_instance of include/ravesim.h."""
    _fields_ = [("u", _vp), ("U", _vp), ("dx", _d), ("dy", _d), ("wl", _d), ("cutoff_freq", _d),
                ("steps", C.POINTER(ChainStep))]


class HistoryFrames(C.Structure):
    """rsim_history_frames of include/ravesim.h."""
    _fields_ = [("out_x", _i), ("out_y", _i), ("pixel_size_x", _d), ("pixel_size_y", _d),
                ("z_start", _d), ("frames", _vp)]


# name -> argtypes (restype is int unless listed in _RESTYPES)
SIGNATURES = {
    "rsim_last_error": [],
    "rsim_version": [],
    "rsim_device_count": [_ip],
    "rsim_set_device": [_i],
    "rsim_malloc": [C.POINTER(_vp), _sz],
    "rsim_free": [_vp],
    "rsim_memcpy_h2d": [_vp, _vp, _sz, _vp],
    "rsim_memcpy_d2h": [_vp, _vp, _sz, _vp],
    "rsim_memcpy_d2d": [_vp, _vp, _sz, _vp],
    "rsim_memset": [_vp, _i, _sz, _vp],
    "rsim_synchronize": [_vp],
    "rsim_stream_create": [C.POINTER(_vp)],
    "rsim_stream_destroy": [_vp],
    "rsim_release_caches": [],
    "rsim_cache_info": [C.POINTER(_sz), _ip],
    "rsim_graph_begin": [_vp],
    "rsim_graph_end": [_vp, C.POINTER(_vp)],
    "rsim_graph_launch": [_vp, _vp],
    "rsim_graph_info": [_vp, C.POINTER(_sz)],
    "rsim_graph_destroy": [_vp],
    "rsim_fill_zero": [_vp, _sz, _vp],
    "rsim_stream_copy": [_vp, _vp, _sz, _vp],
    "rsim_kernel_timing": [_i],
    "rsim_kernel_timing_read": [_dp, _ip],
    "rsim_set_progress": [PROGRESS_FN, _vp],
    "rsim_request_cancel": [],
    "rsim_clear_cancel": [],
    "rsim_control_create": [C.POINTER(_vp)],
    "rsim_control_destroy": [_vp],
    "rsim_control_set_progress": [_vp, PROGRESS_FN, _vp, _i],
    "rsim_control_request_cancel": [_vp],
    "rsim_control_clear_cancel": [_vp],
    "rsim_control_bind": [_vp],
    "rsim_control_completed": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "rsim_fft2": [_vp, _vp, _i, _i, _i, _i, _vp],
    "rsim_fft1d_rows": [_vp, _vp, _i, _i, _i, _i, _vp],
    "rsim_apply_frequency_cutoff_2d": [_vp, _i, _i, _i, _d, _d, _d, _vp],
    "rsim_propagate_2d": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _i, _vp],
    "rsim_propagate_analytically_2d": [_vp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _vp],
    "rsim_volume_create": [C.POINTER(_vp), _vp, _i, _i, _i, _i, _dp, _i],
    "rsim_volume_create_empty": [C.POINTER(_vp), _i, _i, _i, _dp, _i],
    "rsim_volume_destroy": [_vp],
    "rsim_volume_info": [_vp, C.POINTER(_vp), _ip, _ip, _ip],
    "rsim_volume_paint_spheres": [_vp, _dp, _ip, _i, _vp],
    "rsim_transmission_2d": [_vp, _i, _i, _i, _d, _d, _d, _vp, _i, _d, _d, _d, _d, _d, _vp],
    "rsim_sample_apply_2d": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _vp, _d, _d, _d, _d, _d, _i,
                             _i, _i, _i, _vp],
    "rsim_chain_apply_2d": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _i, C.POINTER(ChainStep), _i,
                            _i, _i, _vp],
    "rsim_chain_apply_2d_batch": [C.POINTER(ChainInstance), _i, _i, _i, _i, _i, _i, _i, _i, _vp],
    "rsim_fused_schedule": [_i, _i, _ip, _ip],
    "rsim_fused_tile": [_i, _i, _ip],
    "rsim_detector_area_v1": [_vp, _i, _i, _i, _d, _d, _i, _i, _d, _d, _d, _dp, _vp],
    "rsim_detector_area_v1_accumulate": [_vp, _i, _i, _i, _d, _d, _i, _i, _d, _d, _d, _d, _vp,
                                         _vp],
    "rsim_detector_area_v1_accumulate_noalloc": [_vp, _i, _i, _i, _d, _d, _i, _i, _d, _d, _d, _d,
                                                 _vp, _vp],
    "rsim_detector_noalloc_capacity": [C.POINTER(_sz)],
    "rsim_transmission_db_2d": [_vp, _i, _i, _i, _d, _d, _vp, _vp, _vp],
    "rsim_multislice_db_2d": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _vp, _vp, _d, _i, _vp, _i,
                              _i, _i, _vp],
    "rsim_multislice_adjoint_workspace": [_i, _i, _i, _i, _i, C.POINTER(_sz), C.POINTER(_sz)],
    "rsim_multislice_db_adjoint_2d": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _vp, _vp, _d, _i,
                                      _vp, _i, _vp, _vp, _vp, _i, _i, _vp],
    "rsim_detector_area_v1_adjoint": [_vp, _vp, _i, _i, _i, _d, _d, _i, _i, _d, _d, _d, _vp,
                                      _vp],
    "rsim_sample_apply_2d_scaled": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _vp, _d, _d, _d, _d,
                                    _d, _i, _i, _dp, _dp, _i, _i, _vp],
    "rsim_transmission_2d_scaled": [_vp, _i, _i, _i, _d, _d, _d, _vp, _i, _d, _d, _d, _d, _d,
                                    _d, _vp],
    "rsim_tilted_plane_2d": [_vp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _vp],
    "rsim_sample_apply_2d_history": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _vp, _d, _d, _d, _d,
                                     _d, _i, _i, _dp, _dp, _i, _i, C.POINTER(HistoryFrames), _vp],
    "rsim_plasma_apply_2d_history": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _vp, _d, _dp, _d, _d,
                                     _d, _d, _d, _i, _i, _dp, _dp, _i, _i,
                                     C.POINTER(HistoryFrames), _vp],
    "rsim_plasma_create": [C.POINTER(_vp), _vp, _vp, _vp, _vp, _i, _i, _i, _i],
    "rsim_plasma_destroy": [_vp],
    "rsim_plasma_deltabeta": [_vp, _i, _i, _d, _dp, _vp, _vp],
    "rsim_plasma_transmission_2d": [_vp, _i, _i, _i, _d, _d, _d, _vp, _i, _d, _dp, _d, _d, _d,
                                    _d, _d, _d, _vp],
    "rsim_plasma_apply_2d": [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _vp, _d, _dp, _d, _d, _d, _d,
                             _d, _i, _i, _dp, _dp, _i, _i, _vp],
    "rsim_batch_create": [C.POINTER(_vp), C.POINTER(BatchInstance), _i, _i, _i, _i, _vp],
    "rsim_batch_run": [_vp, _vp],
    "rsim_batch_info": [_vp, _ip, _ip],
    "rsim_batch_destroy": [_vp],
}
_RESTYPES = {"rsim_last_error": C.c_char_p, "rsim_version": C.c_char_p}

_lib = None
_lock = threading.Lock()


class RavesimError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libravesim_hip.so not found at {LIB_PATH}; build it with "
                "`python 3d_ravesim_amd/build.py` (the engine has no CPU fallback)")
        # PyTorch-ROCm wheels carry their own HIP/HSA runtime.  Loading torch
        # first lets the library bind to that runtime, so one runtime serves
        # the process and torch's streams and allocations are valid here; the
        # reverse order leaves torch unable to find the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, C.c_int)
        _lib = lib
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc == RSIM_OK:
        return
    msg = load().rsim_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == RSIM_E_INVALID_ARG:
        raise ValueError(text)
    if rc == RSIM_E_CANCELLED:
        raise InterruptedError(text)
    if rc == RSIM_E_NOMEM:
        raise MemoryError(text)
    if rc == RSIM_E_UNSUPPORTED:
        raise NotImplementedError(text)
    raise RavesimError(text)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def device_count() -> int:
    n = C.c_int(0)
    rc = load().rsim_device_count(C.byref(n))
    if rc != RSIM_OK:
        return 0
    return int(n.value)


def dtype_code(dtype) -> int:
    import numpy as np
    dt = np.dtype(dtype)
    if dt == np.complex64:
        return RSIM_C64
    if dt == np.complex128:
        return RSIM_C128
    raise ValueError(f"field dtype must be complex64 or complex128, got {dt}")
