"""GPU parity of the differentiable multislice (per-voxel delta/beta) and its adjoint.

Every case runs the HIP library through the C ABI (directly, or through the
torch.autograd.Function wrapper on torch-owned device memory) and compares
with the FP64 oracle or the complex128 CPU torch autograd reference.
"""

import ctypes as C
import importlib
import sys

import numpy as np
import pytest
import torch

import oracle as O
from _adjoint_ref import case, loss_and_grad_img, oracle_gradient, torch_gradient
from _rs import rs

pytestmark = pytest.mark.gpu
lib = rs._lib
AG = importlib.import_module("3d_ravesim_amd.autograd")
TORCH_C = {np.complex64: torch.complex64, np.complex128: torch.complex128}


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def P(t):
    return C.c_void_p(t.data_ptr())


def S():
    h = torch.cuda.current_stream().cuda_stream
    assert h != 0, "tests run on a non-default torch stream (0 selects the library's own)"
    return C.c_void_p(h)


@pytest.fixture(autouse=True)
def _torch_stream():
    """Run each test's torch work and library calls on one non-default stream."""
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        yield
    torch.cuda.synchronize()


def code(dt):
    return lib.RSIM_C64 if dt == np.complex64 else lib.RSIM_C128


def field(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a.astype(dt))).cuda()


def vol(a):
    return torch.from_numpy(np.ascontiguousarray(a.astype(np.float32))).cuda()


def rel(a, e):
    return O.relative_l2(a, e)


@pytest.mark.parametrize("dt,tol", [(np.complex128, 1e-14), (np.complex64, 2e-7)])
def test_transmission_db_matches_oracle(dt, tol):
    delta, beta, u0, geo = case(n=64, nz=1, seed=11)
    say("transmission_db", dt.__name__)
    u = field(u0, dt)
    d, b = vol(delta), vol(beta)   # held: a temporary's memory could be reused at once
    lib.call("rsim_transmission_db_2d", P(u), 64, 64, code(dt), geo["wl"], geo["pz"], P(d), P(b),
             S())
    t = O.material_factor(delta[0].astype(np.float64) + 1j * beta[0].astype(np.float64),
                          geo["pz"], geo["wl"])
    assert rel(u.cpu().numpy(), u0 * t) < tol


@pytest.mark.parametrize("shape,ds,size", [((7, 11), (1.3, 1.7), (9.0, 6.5)),
                                           ((64, 64), (4.0, 4.0), (64.0, 64.0)),
                                           ((12, 10), (0.6, 0.45), (8.0, 10.5))])
@pytest.mark.parametrize("dt,tol", [(np.complex128, 1e-13), (np.complex64, 1e-6)])
def test_area_v1_adjoint_matches_oracle(shape, ds, size, dt, tol):
    rng = np.random.default_rng(12)
    ny, nx = shape
    v = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    oy, ox = O.detector_shape(size[0], size[1], ds[0], ds[1])
    gI = rng.normal(size=(oy, ox))
    exp = O.area_v1_adjoint(v.astype(dt).astype(np.complex128), gI, 1.0, 1.1, size[0], size[1],
                            ds[0], ds[1], 5.0)
    say("area_v1_adjoint", shape, ds, dt.__name__)
    vd = field(v, dt)
    g = torch.empty_like(vd)
    gi = torch.from_numpy(gI).cuda()
    lib.call("rsim_detector_area_v1_adjoint", P(vd), P(g), nx, ny, code(dt), 1.0, 1.1, ox, oy,
             ds[0], ds[1], 5.0, P(gi), S())
    assert rel(g.cpu().numpy(), exp) < tol


@pytest.mark.parametrize("dt,tol", [(np.complex128, 1e-12), (np.complex64, 3e-6)])
def test_multislice_db_forward_and_checkpoints_match_oracle(dt, tol):
    delta, beta, u0, geo = case(n=128, nz=6, seed=13)
    exp, before = O.multislice_db_2d(u0, delta, beta, geo["dx"], geo["dy"], geo["wl"], geo["pz"],
                                     geo["cutoff_freq"], keep=True)
    say("multislice_db forward", dt.__name__)
    u = field(u0, dt)
    U = torch.empty_like(u)
    every = 4
    ck = torch.empty((2, 128, 128), dtype=u.dtype, device=u.device)
    d, b = vol(delta), vol(beta)
    lib.call("rsim_multislice_db_2d", P(u), P(U), 128, 128, code(dt), geo["dx"], geo["dy"],
             geo["wl"], geo["cutoff_freq"], P(d), P(b), geo["pz"], 6, P(ck), every, lib.RSIM_PLAIN,
             0, S())
    assert rel(u.cpu().numpy(), exp) < tol
    assert rel(ck[0].cpu().numpy(), before[0]) < tol
    assert rel(ck[1].cpu().numpy(), before[4]) < tol
    assert rel(U.cpu().numpy(), O.fft2(exp)) < tol


def _run_adjoint(delta, beta, u0, geo, gI, dt, every, engine=None, flags=0):
    """HIP forward (with checkpoints), leg, detector adjoint, leg adjoint, slice adjoint."""
    engine = lib.RSIM_PLAIN if engine is None else engine
    nz, ny, nx = delta.shape
    c = code(dt)
    d, b = vol(delta), vol(beta)
    u = field(u0, dt)
    U = torch.empty_like(u)
    ckb, wkb = C.c_size_t(), C.c_size_t()
    lib.call("rsim_multislice_adjoint_workspace", nx, ny, c, nz, every, C.byref(ckb), C.byref(wkb))
    es = u.element_size()
    ck = torch.empty(ckb.value // es, dtype=u.dtype, device=u.device)
    work = torch.empty(max(wkb.value // es, 1), dtype=u.dtype, device=u.device)
    s = S()
    lib.call("rsim_multislice_db_2d", P(u), P(U), nx, ny, c, geo["dx"], geo["dy"], geo["wl"],
             geo["cutoff_freq"], P(d), P(b), geo["pz"], nz, P(ck), every, engine, flags, s)
    lib.call("rsim_propagate_2d", P(u), P(U), nx, ny, c, geo["dx"], geo["dy"], geo["wl"],
             geo["z_leg"], geo["cutoff_freq"], 0, s)
    oy, ox = gI.shape
    g = torch.empty_like(u)
    gi = torch.from_numpy(gI).cuda()
    lib.call("rsim_detector_area_v1_adjoint", P(u), P(g), nx, ny, c, geo["dx"], geo["dy"], ox, oy,
             geo["ds_x"], geo["ds_y"], geo["z_det"], P(gi), s)
    lib.call("rsim_propagate_2d", P(g), P(U), nx, ny, c, geo["dx"], geo["dy"], geo["wl"],
             -geo["z_leg"], geo["cutoff_freq"], 0, s)
    gd, gb = torch.empty_like(d), torch.empty_like(b)
    lib.call("rsim_multislice_db_adjoint_2d", P(g), P(U), nx, ny, c, geo["dx"], geo["dy"],
             geo["wl"], geo["cutoff_freq"], P(d), P(b), geo["pz"], nz, P(ck), every, P(work),
             P(gd), P(gb), engine, flags, s)
    torch.cuda.synchronize()
    return gd.cpu().numpy(), gb.cpu().numpy()


def _target(delta, beta, u0, geo):
    return O.multislice_db_intensity(u0, delta * 1.05, beta * 0.96, geo["dx"], geo["dy"],
                                     geo["wl"], geo["pz"], geo["cutoff_freq"], geo["z_leg"],
                                     geo["size_x"], geo["size_y"], geo["ds_x"], geo["ds_y"],
                                     geo["z_det"])


@pytest.mark.parametrize("every", [1, 3, 7])
@pytest.mark.parametrize("dt,tol", [(np.complex128, 1e-9), (np.complex64, 1e-4)])
def test_adjoint_gradient_matches_oracle(every, dt, tol):
    delta, beta, u0, geo = case(n=96, nz=7, seed=14)
    target = _target(delta, beta, u0, geo)
    _, gI = loss_and_grad_img(delta, beta, u0, geo, target)
    ed, eb = oracle_gradient(delta, beta, u0, geo, gI)
    say("adjoint gradient", dt.__name__, "every", every)
    gd, gb = _run_adjoint(delta, beta, u0, geo, gI, dt, every)
    assert rel(gd, ed) < tol
    assert rel(gb, eb) < tol


# ---- fused complex64 chain (FP32 butterflies; 256^2 fields) ----

@pytest.mark.parametrize("nz,every", [(6, 1), (6, 4), (6, 6), (1, 1)])
def test_fused_multislice_db_forward_and_checkpoints_match_oracle(nz, every):
    n = 256
    delta, beta, u0, geo = case(n=n, nz=nz, seed=21)
    exp, before = O.multislice_db_2d(u0, delta, beta, geo["dx"], geo["dy"], geo["wl"], geo["pz"],
                                     geo["cutoff_freq"], keep=True)
    say("fused multislice_db forward", nz, every)
    u = field(u0, np.complex64)
    U = torch.empty_like(u)
    nck = -(-nz // every)
    ck = torch.empty((nck, n, n), dtype=u.dtype, device=u.device)
    d, b = vol(delta), vol(beta)
    lib.call("rsim_multislice_db_2d", P(u), P(U), n, n, lib.RSIM_C64, geo["dx"], geo["dy"],
             geo["wl"], geo["cutoff_freq"], P(d), P(b), geo["pz"], nz, P(ck), every,
             lib.RSIM_FUSED, 0, S())
    assert rel(u.cpu().numpy(), exp) < 3e-6
    for j in range(nck):
        assert rel(ck[j].cpu().numpy(), before[j * every]) < 3e-6, j
    assert rel(U.cpu().numpy(), O.fft2(exp)) < 3e-6


@pytest.mark.parametrize("nz,every,math", [(7, 1, "fp32"), (7, 3, "fp32"), (7, 7, "fp32"),
                                           (1, 1, "fp32"), (7, 3, "comp32"), (7, 3, "fp64")])
def test_fused_adjoint_gradient_matches_oracle(nz, every, math):
    delta, beta, u0, geo = case(n=256, nz=nz, seed=22)
    target = _target(delta, beta, u0, geo)
    _, gI = loss_and_grad_img(delta, beta, u0, geo, target)
    ed, eb = oracle_gradient(delta, beta, u0, geo, gI)
    say("fused adjoint gradient", nz, every, math)
    flags = {"fp32": lib.RSIM_FLAG_MATH_FP32, "comp32": lib.RSIM_FLAG_MATH_COMP32,
             "fp64": lib.RSIM_FLAG_MATH_FP64}[math]
    gd, gb = _run_adjoint(delta, beta, u0, geo, gI, np.complex64, every, lib.RSIM_FUSED, flags)
    assert rel(gd, ed) < 1e-4
    assert rel(gb, eb) < 1e-4


@pytest.mark.parametrize("nx,ny", [(512, 256), (256, 1024)])
def test_fused_adjoint_non_square_matches_oracle(nx, ny):
    """Rectangular fields with dy != dx: the row (x) and column (y) passes of the
    fused chain and its adjoint use their own lengths and pitches."""
    delta, beta, u0, geo = case(n=nx, ny=ny, dy=0.6e-6, nz=5, seed=25)
    exp = O.multislice_db_2d(u0, delta, beta, geo["dx"], geo["dy"], geo["wl"], geo["pz"],
                             geo["cutoff_freq"])
    say("fused non-square", nx, ny)
    u = field(u0, np.complex64)
    U = torch.empty_like(u)
    d, b = vol(delta), vol(beta)
    lib.call("rsim_multislice_db_2d", P(u), P(U), nx, ny, lib.RSIM_C64, geo["dx"], geo["dy"],
             geo["wl"], geo["cutoff_freq"], P(d), P(b), geo["pz"], 5, None, 1, lib.RSIM_FUSED, 0,
             S())
    assert rel(u.cpu().numpy(), exp) < 3e-6
    target = _target(delta, beta, u0, geo)
    _, gI = loss_and_grad_img(delta, beta, u0, geo, target)
    ed, eb = oracle_gradient(delta, beta, u0, geo, gI)
    gd, gb = _run_adjoint(delta, beta, u0, geo, gI, np.complex64, 2, lib.RSIM_FUSED)
    assert rel(gd, ed) < 1e-4
    assert rel(gb, eb) < 1e-4


def test_db_entry_points_reject_bad_engine_and_flags():
    delta, beta, u0, geo = case(n=256, nz=2, seed=24)
    say("db argument checks")
    u = field(u0, np.complex64)
    U = torch.empty_like(u)
    d, b = vol(delta), vol(beta)
    for engine, flags in ((2, 0), (lib.RSIM_FUSED, 1),
                          (lib.RSIM_FUSED, lib.RSIM_FLAG_MATH_FP32 | lib.RSIM_FLAG_MATH_FP64)):
        with pytest.raises(ValueError, match="engine|flags|RSIM_FLAG"):
            lib.call("rsim_multislice_db_2d", P(u), P(U), 256, 256, lib.RSIM_C64, geo["dx"],
                     geo["dy"], geo["wl"], geo["cutoff_freq"], P(d), P(b), geo["pz"], 2, None, 1,
                     engine, flags, S())


def test_fused_adjoint_checkpoint_interval_and_plain_engine_agree():
    """Rebuilt fields equal the stored ones up to FP32 rounding: the gradients
    hardly move with the checkpoint interval, and the fused chain agrees with
    the plain complex64 engine."""
    delta, beta, u0, geo = case(n=256, nz=9, seed=23)
    target = _target(delta, beta, u0, geo)
    _, gI = loss_and_grad_img(delta, beta, u0, geo, target)
    say("fused adjoint: checkpoint intervals and plain engine")
    ref = _run_adjoint(delta, beta, u0, geo, gI, np.complex64, 1, lib.RSIM_FUSED)
    for every in (2, 3, 9, 20):
        got = _run_adjoint(delta, beta, u0, geo, gI, np.complex64, every, lib.RSIM_FUSED)
        assert rel(got[0], ref[0]) < 1e-6 and rel(got[1], ref[1]) < 1e-6, every
    plain = _run_adjoint(delta, beta, u0, geo, gI, np.complex64, 3, lib.RSIM_PLAIN)
    assert rel(plain[0], ref[0]) < 2e-5 and rel(plain[1], ref[1]) < 2e-5


def test_checkpoint_interval_does_not_change_gradients():
    delta, beta, u0, geo = case(n=64, nz=9, seed=15)
    target = _target(delta, beta, u0, geo)
    _, gI = loss_and_grad_img(delta, beta, u0, geo, target)
    say("checkpoint invariance")
    ref = _run_adjoint(delta, beta, u0, geo, gI, np.complex128, 1)
    for every in (2, 3, 9, 20):
        got = _run_adjoint(delta, beta, u0, geo, gI, np.complex128, every)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), every


def _geometry(geo, n, every=0, engine="fused"):
    oy, ox = O.detector_shape(geo["size_x"], geo["size_y"], geo["ds_x"], geo["ds_y"])
    return AG.MultisliceGeometry(dx=geo["dx"], dy=geo["dy"], wl=geo["wl"], pz=geo["pz"],
                                 cutoff=geo["cutoff_freq"], z_leg=geo["z_leg"], out_x=ox, out_y=oy,
                                 ds_x=geo["ds_x"], ds_y=geo["ds_y"], z_det=geo["z_det"],
                                 ckpt_every=every, engine=engine)


# n = 64 runs the plain engine (no fused chain at that length); n = 256 the
# fused complex64 chain
@pytest.mark.parametrize("dt,tol,n", [(np.complex128, 1e-9, 64), (np.complex64, 1e-4, 64),
                                      (np.complex64, 1e-4, 256)])
def test_autograd_function_matches_cpu_complex128_autograd(dt, tol, n):
    delta, beta, u0, geo = case(n=n, nz=6, seed=16)
    target = _target(delta, beta, u0, geo)
    td, tb, tI = torch_gradient(delta, beta, u0, geo, target)
    say("autograd.Function", dt.__name__, n)
    d = vol(delta).requires_grad_(True)
    b = vol(beta).requires_grad_(True)
    I = AG.multislice_intensity(d, b, _geometry(geo, n), u0=field(u0, dt), dtype=TORCH_C[dt])
    loss = ((I - torch.from_numpy(target).cuda()) ** 2).sum()
    loss.backward()
    assert rel(I.detach().cpu().numpy(), tI) < tol
    assert rel(d.grad.cpu().numpy(), td) < tol
    assert rel(b.grad.cpu().numpy(), tb) < tol


def test_gradient_matches_central_differences_on_gpu():
    """dL/d(delta, beta) of the complex128 HIP adjoint vs central differences of
    the complex128 HIP forward at 8 random voxels."""
    delta, beta, u0, geo = case(n=64, nz=5, seed=17)
    target = torch.from_numpy(_target(delta, beta, u0, geo)).cuda()
    G = _geometry(geo, 64, every=2)
    u = field(u0, np.complex128)
    d = vol(delta).requires_grad_(True)
    b = vol(beta).requires_grad_(True)
    say("gpu finite differences (autograd on torch's default stream)")
    with torch.cuda.stream(torch.cuda.default_stream()):
        loss = ((AG.multislice_intensity(d, b, G, u0=u, dtype=torch.complex128) - target) ** 2).sum()
        loss.backward()
    torch.cuda.synchronize()   # the gradients were produced on the default stream
    gd, gb = d.grad.cpu().numpy(), b.grad.cpu().numpy()
    rng = np.random.default_rng(18)

    def L(dd, bb):
        with torch.no_grad():
            I = AG.multislice_intensity(dd, bb, G, u0=u, dtype=torch.complex128)
            return float(((I - target) ** 2).sum())

    for _ in range(8):
        s, y, x = (int(rng.integers(0, k)) for k in delta.shape)
        for which, grad, h in ((0, gd, 2e-8), (1, gb, 2e-11)):
            dp, bp = d.detach().clone(), b.detach().clone()
            dm, bm = d.detach().clone(), b.detach().clone()
            tgt_p, tgt_m = (dp, dm) if which == 0 else (bp, bm)
            # float32 volumes: use the exactly representable perturbed values
            v0 = float(tgt_p[s, y, x])
            vp = float(np.float32(v0 + h))
            vm = float(np.float32(v0 - h))
            tgt_p[s, y, x] = vp
            tgt_m[s, y, x] = vm
            fd = (L(dp, bp) - L(dm, bm)) / (vp - vm)
            scale = np.abs(grad).max()
            assert abs(fd - grad[s, y, x]) < 2e-4 * scale, (which, s, y, x, fd, grad[s, y, x])


def _smooth_case(n, nz, seed):
    """Gaussian-filtered (sigma 4 voxels) seeded noise, delta in [0, 1e-6] and
    beta in [0, 1e-9] from independent fields; target from a +5 % smooth
    perturbation of delta (the differentiable configuration at small size)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)

    def field01(shape):
        f = gaussian_filter(rng.normal(size=shape), 4.0, mode="wrap")
        return (f - f.min()) / (f.max() - f.min())

    delta = (1e-6 * field01((nz, n, n))).astype(np.float32)
    beta = (1e-9 * field01((nz, n, n))).astype(np.float32)
    pert = field01((nz, n, n))
    u0 = np.ones((n, n), np.complex128)
    geo = dict(dx=0.5e-6, dy=0.5e-6, wl=0.5e-10, pz=0.5e-6, cutoff_freq=0.4 / 0.5e-6, z_leg=0.2,
               size_x=n * 0.5e-6, size_y=n * 0.5e-6, ds_x=2e-6, ds_y=2e-6, z_det=0.2 + nz * 0.5e-6)
    target = O.multislice_db_intensity(u0, (delta * (1 + 0.05 * pert)).astype(np.float32), beta,
                                       geo["dx"], geo["dy"], geo["wl"], geo["pz"],
                                       geo["cutoff_freq"], geo["z_leg"], geo["size_x"],
                                       geo["size_y"], geo["ds_x"], geo["ds_y"], geo["z_det"])
    return delta, beta, u0, geo, target


def test_complex64_gradient_end_to_end_from_the_gpu_forward():
    """The gradient a complex64 user gets: the GPU's own forward, loss and
    backward (no oracle dL/dI), with |I - I_t| / |I| of a few 1e-3 that
    amplifies image errors.  Default mode (image in complex128, adjoint chain
    in fused complex64): within 1e-4 of the FP64 oracle gradient."""
    n, nz = 256, 32
    delta, beta, u0, geo, target = _smooth_case(n, nz, seed=41)
    _, gI = loss_and_grad_img(delta, beta, u0, geo, target)
    I_ref = target + gI / 2
    ratio = np.linalg.norm(I_ref - target) / np.linalg.norm(I_ref)
    ref_d, ref_b = oracle_gradient(delta, beta, u0, geo, gI)
    say(f"end-to-end c64 gradient, |I-I_t|/|I| = {ratio:.2e}")
    assert 1e-3 < ratio < 2e-2
    errs = {}
    oy, ox = O.detector_shape(geo["size_x"], geo["size_y"], geo["ds_x"], geo["ds_y"])
    for forward in ("auto", "field"):
        G = AG.MultisliceGeometry(dx=geo["dx"], wl=geo["wl"], pz=geo["pz"],
                                  cutoff=geo["cutoff_freq"], z_leg=geo["z_leg"], out_x=ox,
                                  out_y=oy, ds_x=geo["ds_x"], ds_y=geo["ds_y"],
                                  z_det=geo["z_det"], forward=forward)
        d = vol(delta).requires_grad_(True)
        b = vol(beta).requires_grad_(True)
        I = AG.multislice_intensity(d, b, G, dtype=torch.complex64)
        ((I - torch.from_numpy(target).cuda()) ** 2).sum().backward()
        torch.cuda.synchronize()
        errs[forward] = (rel(d.grad.cpu().numpy(), ref_d), rel(b.grad.cpu().numpy(), ref_b))
        say(f"  forward={forward}: delta {errs[forward][0]:.2e} beta {errs[forward][1]:.2e}")
    assert errs["auto"][0] < 1e-4 and errs["auto"][1] < 1e-4
