"""Differentiable multislice: a torch.autograd.Function over the HIP adjoint.

Forward: a plane (or given) incident field passes ``n`` slices of per-voxel
float32 delta/beta (one voxel per field pixel), each slice applying
big-wave's material_factor (big-wave/optical_element.py:39-42) and
propagate_2d (big-wave/propagation.py:396-442); then a free-space leg over
``z_leg`` and the area_v1 detector (big-wave/propagation.py:995-1130).
Backward: the hand-written adjoint (rsim_multislice_db_adjoint_2d) with
sqrt(n) slice checkpointing returns dL/d(delta, beta) per voxel.

PyTorch is used for device memory, streams and the autograd graph only; all
arithmetic runs in the HIP library on torch's current stream.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _lib as lib


def _code(dtype: torch.dtype) -> int:
    if dtype == torch.complex64:
        return lib.RSIM_C64
    if dtype == torch.complex128:
        return lib.RSIM_C128
    raise ValueError(f"field dtype must be complex64 or complex128, got {dtype}")


_SIDE: dict = {}


class _OnStream:
    """Stream for library calls, ordered with torch's current stream.

    A non-default current stream is used directly.  torch's default stream has
    handle 0, which the C ABI reads as "the library's own stream", so then the
    calls run on a side stream that first waits for the current stream and
    that the current stream waits for afterwards.
    """

    def __enter__(self) -> C.c_void_p:
        cur = torch.cuda.current_stream()
        self.cur, self.side = cur, None
        if cur.cuda_stream != 0:
            return C.c_void_p(cur.cuda_stream)
        side = _SIDE.get(cur.device)
        if side is None:
            side = _SIDE[cur.device] = torch.cuda.Stream(device=cur.device)
        side.wait_stream(cur)
        self.side = side
        return C.c_void_p(side.cuda_stream)

    def __exit__(self, *exc) -> None:
        if self.side is not None:
            self.cur.wait_stream(self.side)


def _p(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


@dataclass(frozen=True)
class MultisliceGeometry:
    """Geometry of the differentiable path (SI units)."""
    dx: float
    wl: float
    pz: float                 # slice thickness
    cutoff: float             # circular cutoff frequency f (kx^2 + ky^2 <= 2 f^2)
    z_leg: float              # free space after the last slice
    out_x: int                # detector pixels
    out_y: int
    ds_x: float               # detector pixel size
    ds_y: float
    z_det: float              # current_z of area_v1's obliquity factor
    dy: float = 0.0           # 0: dy = dx
    ckpt_every: int = 0       # 0: ceil(sqrt(n_slices))
    engine: str = "fused"     # "fused" (complex64 fused chain where the shape allows) or "plain"
    math: str = "fp32"        # fused butterflies: "fp32", "comp32" (compensated) or "fp64"
    # Precision of the forward pass that forms the image, the loss gradient
    # dL/dI and the checkpoints: "complex128" (plain FP64 engine), "field" (the
    # field dtype), or "auto" = complex128 for complex64 fields.  A loss such
    # as ||I - I_t||^2 divides the image error by |I - I_t| / |I| (4e-3 for a
    # 5 % perturbation), so a complex64 image error of 5e-6 becomes a 1e-3
    # gradient error; with the image in complex128 only the adjoint chain runs
    # in complex64, and the gradient stays within 1e-4.
    forward: str = "auto"

    @property
    def dyy(self) -> float:
        return self.dy if self.dy != 0.0 else self.dx

    @property
    def engine_code(self) -> int:
        if self.engine not in ("fused", "plain"):
            raise ValueError(f"engine must be 'fused' or 'plain', got {self.engine!r}")
        return lib.RSIM_FUSED if self.engine == "fused" else lib.RSIM_PLAIN

    def forward_c128(self, dtype: torch.dtype) -> bool:
        if self.forward not in ("auto", "field", "complex128"):
            raise ValueError(f"forward must be 'auto', 'field' or 'complex128', got {self.forward!r}")
        if self.forward == "complex128":
            return True
        return self.forward == "auto" and dtype == torch.complex64

    @property
    def math_flags(self) -> int:
        flags = {"fp32": lib.RSIM_FLAG_MATH_FP32, "comp32": lib.RSIM_FLAG_MATH_COMP32,
                 "fp64": lib.RSIM_FLAG_MATH_FP64}
        if self.math not in flags:
            raise ValueError(f"math must be one of {sorted(flags)}, got {self.math!r}")
        return flags[self.math]


def _check_volume(name: str, t: torch.Tensor) -> None:
    if t.dtype != torch.float32 or t.dim() != 3 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float32 CUDA tensor (n_slices, ny, nx)")


class MultisliceIntensity(torch.autograd.Function):
    """I = area_v1(|P_leg prod_s P T_s u0|^2) as a function of (delta, beta)."""

    @staticmethod
    def forward(ctx, delta, beta, u0, geom: MultisliceGeometry, dtype: torch.dtype):
        _check_volume("delta", delta)
        _check_volume("beta", beta)
        if delta.shape != beta.shape:
            raise ValueError("delta and beta must have the same shape")
        n, ny, nx = delta.shape
        code = _code(dtype)
        every = geom.ckpt_every or max(1, math.ceil(math.sqrt(n)))
        dev = delta.device
        if geom.forward_c128(dtype):
            return MultisliceIntensity._forward_c128(ctx, delta, beta, u0, geom, dtype, every)
        u = u0.to(device=dev, dtype=dtype).reshape(ny, nx).contiguous().clone()
        U = torch.empty_like(u)
        ck_b, wk_b = C.c_size_t(), C.c_size_t()
        lib.call("rsim_multislice_adjoint_workspace", nx, ny, code, n, every, C.byref(ck_b),
                 C.byref(wk_b))
        ckpt = torch.empty(ck_b.value // u.element_size(), dtype=dtype, device=dev)
        img = torch.zeros((geom.out_y, geom.out_x), dtype=torch.float64, device=dev)
        with _OnStream() as s:
            lib.call("rsim_multislice_db_2d", _p(u), _p(U), nx, ny, code, geom.dx, geom.dyy,
                     geom.wl, geom.cutoff, _p(delta), _p(beta), geom.pz, n, _p(ckpt), every,
                     geom.engine_code, geom.math_flags, s)
            lib.call("rsim_propagate_2d", _p(u), _p(U), nx, ny, code, geom.dx, geom.dyy, geom.wl,
                     geom.z_leg, geom.cutoff, 0, s)
            lib.call("rsim_detector_area_v1_accumulate", _p(u), nx, ny, code, geom.dx, geom.dyy,
                     geom.out_x, geom.out_y, geom.ds_x, geom.ds_y, geom.z_det, 1.0, _p(img), s)
        ctx.save_for_backward(delta, beta, ckpt, u)
        ctx.geom, ctx.every, ctx.code, ctx.work_bytes = geom, every, code, wk_b.value
        ctx.c128_front = False
        return img

    @staticmethod
    def _forward_c128(ctx, delta, beta, u0, geom, dtype, every):
        """Image, final field and checkpoints from the complex128 plain engine;
        the checkpoints are handed to the adjoint chain in ``dtype``."""
        n, ny, nx = delta.shape
        dev = delta.device
        u = u0.to(device=dev, dtype=torch.complex128).reshape(ny, nx).contiguous().clone()
        U = torch.empty_like(u)
        ck_b, wk_b = C.c_size_t(), C.c_size_t()
        lib.call("rsim_multislice_adjoint_workspace", nx, ny, lib.RSIM_C128, n, every,
                 C.byref(ck_b), C.byref(wk_b))
        ckpt128 = torch.empty(ck_b.value // u.element_size(), dtype=torch.complex128, device=dev)
        img = torch.zeros((geom.out_y, geom.out_x), dtype=torch.float64, device=dev)
        with _OnStream() as s:
            lib.call("rsim_multislice_db_2d", _p(u), _p(U), nx, ny, lib.RSIM_C128, geom.dx,
                     geom.dyy, geom.wl, geom.cutoff, _p(delta), _p(beta), geom.pz, n,
                     _p(ckpt128), every, lib.RSIM_PLAIN, 0, s)
            lib.call("rsim_propagate_2d", _p(u), _p(U), nx, ny, lib.RSIM_C128, geom.dx, geom.dyy,
                     geom.wl, geom.z_leg, geom.cutoff, 0, s)
            lib.call("rsim_detector_area_v1_accumulate", _p(u), nx, ny, lib.RSIM_C128, geom.dx,
                     geom.dyy, geom.out_x, geom.out_y, geom.ds_x, geom.ds_y, geom.z_det, 1.0,
                     _p(img), s)
        ckpt = ckpt128.to(dtype)
        del ckpt128, U
        code = _code(dtype)
        lib.call("rsim_multislice_adjoint_workspace", nx, ny, code, n, every, C.byref(ck_b),
                 C.byref(wk_b))
        ctx.save_for_backward(delta, beta, ckpt, u)
        ctx.geom, ctx.every, ctx.code, ctx.work_bytes = geom, every, code, wk_b.value
        ctx.c128_front = True
        return img

    @staticmethod
    def backward(ctx, grad_img):
        delta, beta, ckpt, v = ctx.saved_tensors
        geom, every, code = ctx.geom, ctx.every, ctx.code
        n, ny, nx = delta.shape
        gimg = grad_img.to(dtype=torch.float64).contiguous()
        vcode = lib.RSIM_C128 if v.dtype == torch.complex128 else code
        g = torch.empty_like(v)
        U = torch.empty_like(v)
        with _OnStream() as s:
            lib.call("rsim_detector_area_v1_adjoint", _p(v), _p(g), nx, ny, vcode, geom.dx,
                     geom.dyy, geom.out_x, geom.out_y, geom.ds_x, geom.ds_y, geom.z_det,
                     _p(gimg), s)
            # adjoint of the leg: propagation over -z_leg (conjugate chirp, same mask)
            lib.call("rsim_propagate_2d", _p(g), _p(U), nx, ny, vcode, geom.dx, geom.dyy,
                     geom.wl, -geom.z_leg, geom.cutoff, 0, s)
        if ctx.c128_front:
            # the chain runs in the checkpoints' dtype
            g = g.to(ckpt.dtype)
            U = torch.empty_like(g)
        work = torch.empty(max(ctx.work_bytes // g.element_size(), 1), dtype=g.dtype,
                           device=g.device)
        gd = torch.empty_like(delta)
        gb = torch.empty_like(beta)
        with _OnStream() as s:
            lib.call("rsim_multislice_db_adjoint_2d", _p(g), _p(U), nx, ny, code, geom.dx,
                     geom.dyy, geom.wl, geom.cutoff, _p(delta), _p(beta), geom.pz, n, _p(ckpt),
                     every, _p(work), _p(gd), _p(gb), geom.engine_code, geom.math_flags, s)
        return gd, gb, None, None, None


def multislice_intensity(delta: torch.Tensor, beta: torch.Tensor, geom: MultisliceGeometry,
                         u0: torch.Tensor | None = None,
                         dtype: torch.dtype = torch.complex64) -> torch.Tensor:
    """Detector image (out_y, out_x) float64, differentiable in delta and beta.

    ``u0`` is the incident field (default: unit plane wave).  ``dtype`` is the
    field precision (complex64: FP32 FFTs; complex128: FP64 throughout).
    """
    if u0 is None:
        u0 = torch.ones(delta.shape[1:], dtype=dtype, device=delta.device)
    return MultisliceIntensity.apply(delta, beta, u0, geom, dtype)
