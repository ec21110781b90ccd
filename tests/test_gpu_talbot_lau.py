"""Batched Talbot-Lau setup (sources x energies, uneven grids) versus the oracle.

The workload of the batched configuration at reduced size: point sources
20 m upstream with seeded offsets, energies 15 / 27.5 / 40 keV on 256^2,
384^2 and 512^2 grids over the same 128 um field, a binary phase grating
(4.2 um, pi shift at 27.5 keV, 4 slices), 0.3 m of free space, an absorption
analyzer (2.1 um, delta 5e-6, beta 4e-7 at 27.5 keV, 50 um in 10 slices), a
5 cm leg and the area_v1 detector on 8 um pixels, summed with seeded
spectral weights.  Every instance is checked against the FP64 oracle
(big-wave's order: source.py:97-107, optical_element.py:502-596,
propagation.py:396-442 and 995-1130): complex128 plain 1e-12, complex64
(fused where the grid has a fused schedule, plain kernels at 384^2) 3e-6,
per field and for the weighted image.  The whole run captured as one graph
replays bit for bit like the call-by-call run, also at 2048^2-6144^2, and the
batched slice loop matches single calls bit for bit at those sizes.
"""

import ctypes as C
import importlib
import sys

import numpy as np
import pytest

import oracle as O
from _rs import rs

pytestmark = pytest.mark.gpu
lib = rs._lib
SP = importlib.import_module("3d_ravesim_amd.spectral")
B = importlib.import_module("3d_ravesim_amd.batch")

FOV = 128e-6
E_DESIGN = 27.5
DELTA_G = 6.4e-7
Z_SRC = 20.0
GAP = 0.3
LEG = 0.05
DS = 8e-6


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def grating_grid(n, dx, nz, period, duty=0.5):
    """Binary grating along x, extruded in y, on an (nz, n+2, n+2) grid of
    pitch dx centred on the field (mask mod(x/p + duty/2, 1) < duty,
    big-wave/optical_element.py:354-360)."""
    x = (np.arange(n + 2) - (n + 2) / 2.0 + 0.5) * dx
    row = (np.mod(x / period + duty / 2, 1.0) < duty).astype(np.uint8)
    return np.ascontiguousarray(np.broadcast_to(row[None, None, :], (nz, n + 2, n + 2)))


def setup(sizes=(256, 384, 512), energies=(15.0, 27.5, 40.0), n_src=2, seed=3,
          analyzer_slices=10):
    rng = np.random.default_rng(seed)
    src = rng.uniform(-20e-6, 20e-6, size=(n_src, 2))
    weights = rng.uniform(0.5, 1.5, len(energies))
    weights /= weights.sum()
    wl_d = rs.convert_energy_wavelength(E_DESIGN * 1e3)
    t_g = wl_d / (2 * DELTA_G)
    cases = []
    for e, n, w in zip(energies, sizes, weights):
        wl = rs.convert_energy_wavelength(e * 1e3)
        dx = FOV / n
        s2 = (E_DESIGN / e) ** 2
        db_g = np.array([0, DELTA_G * s2 + 1e-9j * s2 ** 1.5])
        db_a = np.array([0, 5e-6 * s2 + 4e-7j * s2 ** 2])
        g = grating_grid(n, dx, 4, 4.2e-6)
        a = grating_grid(n, dx, analyzer_slices, 2.1e-6)
        for xs, ys in src:
            cases.append(dict(n=n, dx=dx, wl=wl, xs=float(xs), ys=float(ys), w=float(w) / n_src,
                              g=g, a=a, db_g=db_g, db_a=db_a, pz_g=t_g / 4,
                              pz_a=50e-6 / analyzer_slices))
    z_end = Z_SRC + t_g + GAP + 50e-6
    return cases, z_end + LEG


def oracle_instance(q):
    n, dx, wl = q["n"], q["dx"], q["wl"]
    cut = 0.4 / dx
    u = O.point_source_propagate_to(n, n, dx, dx, Z_SRC, q["xs"], q["ys"], wl, cut)
    spec_g = O.SampleSpec(grid=q["g"], db_list=q["db_g"], pixel_size_x=dx, pixel_size_y=dx,
                          pixel_size_z=q["pz_g"])
    u, U = O.sample_apply_2d(u, spec_g, dx, dx, wl, cut)
    u, U = O.propagate_2d(u, dx, dx, wl, GAP, cut, U2d=U)
    spec_a = O.SampleSpec(grid=q["a"], db_list=q["db_a"], pixel_size_x=dx, pixel_size_y=dx,
                          pixel_size_z=q["pz_a"])
    u, U = O.sample_apply_2d(u, spec_a, dx, dx, wl, cut)
    u, U = O.propagate_2d(u, dx, dx, wl, LEG, cut, U2d=U)
    return u


def spectral_run(cases, z_det, dtype, keep):
    inst = []
    for q in cases:
        vg = rs.DeviceVolume.from_numpy(q["g"], q["db_g"])
        va = rs.DeviceVolume.from_numpy(q["a"], q["db_a"])
        keep += [vg, va]
        inst.append(SP.Instance(
            nx=q["n"], ny=q["n"], dx=q["dx"], wl=q["wl"], cutoff_freq=0.4 / q["dx"],
            source_x=q["xs"], source_y=q["ys"], source_distance=Z_SRC, weight=q["w"],
            stages=[SP.SampleStage(vg, q["pz_g"], q["g"].shape[0]), SP.GapStage(GAP),
                    SP.SampleStage(va, q["pz_a"], q["a"].shape[0])]))
    nd = int(FOV // DS)
    return SP.SpectralRun(inst, SP.Detector(nd, nd, DS, DS, z_det, LEG), dtype=dtype)


@pytest.mark.parametrize("dt,tol", [(np.complex128, 1e-12), (np.complex64, 3e-6)])
def test_talbot_lau_instances_and_weighted_image_match_oracle(dt, tol):
    cases, z_det = setup()
    keep = []
    run = spectral_run(cases, z_det, dt, keep)
    say("talbot-lau", np.dtype(dt).name, len(cases), "instances")
    run.run()
    run.synchronize()
    nd = int(FOV // DS)
    img_ref = np.zeros((nd, nd))
    for k, q in enumerate(cases):
        u_ref = oracle_instance(q)
        got = run.fields[k][0].to_numpy().reshape(q["n"], q["n"])
        err = O.relative_l2(got, u_ref)
        say(f"  instance {k} n={q['n']} field err {err:.2e}")
        assert err < tol, (k, err)
        img_ref += q["w"] * O.area_v1(u_ref, q["dx"], q["dx"], FOV, FOV, DS, DS, z_det)
    err_img = O.relative_l2(run.image(), img_ref)
    say(f"  weighted image err {err_img:.2e}")
    assert err_img < tol
    run.close()


def test_captured_run_replays_bit_for_bit():
    """One graph launch for the whole run (sources, both gratings, the gap,
    the leg and the weighted detector sum); replays interleaved with eager
    runs and host synchronisations all equal the eager image exactly."""
    cases, z_det = setup()
    keep = []
    run = spectral_run(cases, z_det, np.complex64, keep)
    say("captured talbot-lau")
    run.run()
    run.synchronize()
    eager = run.image()
    g = run.capture()
    assert g.nodes >= 8 * len(cases)
    for rep in range(3):
        run.launch()
        run.synchronize()
        replay = run.image()
        assert np.array_equal(replay, eager), rep
        run.run()                      # eager work between replays
        run.synchronize()
        assert np.array_equal(run.image(), eager)
    lib.call("rsim_release_caches")
    with pytest.raises(ValueError, match="release"):
        run.launch()
    run.close()


def test_capture_refuses_calls_that_would_build_tables():
    """A call whose tables do not exist yet cannot be recorded: it raises
    and the capture yields no graph."""
    say("capture miss")
    lib.call("rsim_release_caches")
    n, dx = 256, 0.5e-6
    u = rs.DeviceVector(n * n, np.complex64)
    U = rs.DeviceVector(n * n, np.complex64)
    with pytest.raises(NotImplementedError, match="before"):
        with rs.graph.CapturedGraph.record() as rec:
            lib.call("rsim_propagate_2d", u.ptr, U.ptr, n, n, lib.RSIM_C64, dx, dx, 0.5e-10,
                     0.01, 0.4 / dx, 0, None)
    with pytest.raises(NotImplementedError, match="fill_zero"):
        with rs.graph.CapturedGraph.record():
            lib.call("rsim_memset", u.ptr, 0, 8, None)
    # the library works normally afterwards
    lib.call("rsim_propagate_2d", u.ptr, U.ptr, n, n, lib.RSIM_C64, dx, dx, 0.5e-10, 0.01,
             0.4 / dx, 0, None)
    lib.call("rsim_synchronize", None)


def test_large_grids_graph_and_batch_are_bit_identical():
    """2 instances per grid size 2048^2, 3072^2, 4096^2 and 6144^2, 2 slices:
    the captured spectral run equals its eager run, and the batched slice
    loop (rsim_batch) equals one call per instance, bit for bit."""
    sizes = (2048, 3072, 4096, 6144)
    fov = 1.024e-3
    rng = np.random.default_rng(9)
    keep, inst, raw = [], [], []
    wl = rs.convert_energy_wavelength(27.5e3)
    for n in sizes:
        dx = fov / n
        g = grating_grid(n, dx, 2, 4.2e-6)
        vol = rs.DeviceVolume.from_numpy(g, np.array([0, DELTA_G + 1e-9j]))
        keep.append(vol)
        for _ in range(2):
            xs, ys = rng.uniform(-20e-6, 20e-6, 2)
            inst.append(SP.Instance(nx=n, ny=n, dx=dx, wl=wl, cutoff_freq=0.4 / dx, source_x=xs,
                                    source_y=ys, source_distance=Z_SRC, weight=0.125,
                                    stages=[SP.SampleStage(vol, 0.5e-6, 2), SP.GapStage(0.1)]))
            raw.append((n, dx, vol, xs, ys))
    say("large graph", sizes)
    run = SP.SpectralRun(inst, SP.Detector(128, 128, 8e-6, 8e-6, Z_SRC + 0.2, 0.1))
    run.run()
    run.synchronize()
    eager = run.image()
    run.capture()
    for _ in range(2):
        run.launch()
        run.synchronize()
        assert np.array_equal(run.image(), eager)
    run.close()

    say("large batch", sizes)
    fields = []
    for n, dx, vol, xs, ys in raw:
        u = rs.DeviceVector(n * n, np.complex64)
        lib.call("rsim_propagate_analytically_2d", u.ptr, n, n, lib.RSIM_C64, dx, dx, Z_SRC, xs,
                 ys, wl, None)
        fields.append((u, rs.DeviceVector(n * n, np.complex64), u.to_numpy()))
    items = [B.BatchItem(u=u, U=U, nx=n, ny=n, dx=dx, wl=wl, cutoff_freq=0.4 / dx, volume=vol,
                         pixel_size_x=dx, pixel_size_z=0.5e-6)
             for (u, U, _), (n, dx, vol, _, _) in zip(fields, raw)]
    batch = B.MultisliceBatch(items, 0, 2)
    batch.run()
    outs = [(u.to_numpy(), U.to_numpy()) for u, U, _ in fields]
    batch.close()
    for (u, U, u0), (n, dx, vol, _, _), (ub, Ub) in zip(fields, raw, outs):
        u.upload(u0)
        lib.call("rsim_sample_apply_2d", u.ptr, U.ptr, n, n, lib.RSIM_C64, dx, dx, wl, 0.4 / dx,
                 vol.handle, dx, dx, 0.5e-6, 0.0, 0.0, 0, 2, lib.RSIM_FUSED, 0, None)
        assert np.array_equal(u.to_numpy(), ub), n
        assert np.array_equal(U.to_numpy(), Ub), n


def test_batched_and_per_instance_runs_agree():
    """The spectral run with instances of one grid sharing launches equals the
    run with one chain call per instance: every field bit for bit, the
    weighted image to its summation order (1e-13)."""
    cases, z_det = setup()
    keep = []
    say("batched vs per-instance spectral run")
    a = spectral_run(cases, z_det, np.complex64, keep)
    a.run()
    a.synchronize()
    b = spectral_run(cases, z_det, np.complex64, keep)
    b.batched = False
    b.run()
    b.synchronize()
    for k in range(len(cases)):
        assert np.array_equal(a.fields[k][0].to_numpy(), b.fields[k][0].to_numpy()), k
    ia, ib = a.image(), b.image()
    assert np.linalg.norm(ia - ib) <= 1e-13 * np.linalg.norm(ib)
    a.close()
    b.close()
