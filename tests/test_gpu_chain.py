"""Propagation chains (rsim_chain_apply_2d) versus the FP64 oracle.

A chain is the element loop of one source in big-wave's run_simulation
(big-wave/wavesim.py:199-300): the source's band limit (FFT2, circular
cutoff, IFFT2; big-wave/source.py:97-107), then samples
(Sample._apply_2d, big-wave/optical_element.py:502-596) and free-space legs
(propagate_2d, big-wave/propagation.py:396-442) in order.  The oracle runs
every step as the reference does; the fused complex64 chain folds the
propagations with no transmission between them into one column pass over the
summed distance.  Tolerances: complex128 plain 1e-12, complex64 fused 3e-6 on
the field, and the spectrum U = FFT2(u) when it is asked for.  Chains of every
shape are covered: free space only, a sample first (no band limit), a gap
between two samples, a leg after the last sample, zero-slice steps, and sizes
without a fused schedule (step by step).
"""

import sys

import numpy as np
import pytest

import oracle as O
from _rs import rs

pytestmark = pytest.mark.gpu
lib = rs._lib
DB = np.array([0, 3.5e-7 + 2e-10j, 1.1e-6 + 3e-9j, 2.6e-7 + 1.5e-10j]) * 30
DX, WL, PZ = 0.5e-6, 0.5e-10, 0.5e-6


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def volume(n, nz, seed):
    rng = np.random.default_rng(seed)
    g = np.zeros((nz, n, n), np.uint8)
    yy, xx = np.mgrid[0:n, 0:n]
    for k in range(nz):
        for _ in range(3):
            cy, cx, r = rng.uniform(0.2 * n, 0.8 * n), rng.uniform(0.2 * n, 0.8 * n), rng.uniform(4, n / 6)
            g[k][(yy - cy) ** 2 + (xx - cx) ** 2 < r * r] = rng.integers(1, 4)
    return g


def oracle_chain(u0, band_limit, steps, cut):
    u = u0.copy()
    if band_limit:
        u = O.ifft2(O.apply_frequency_cutoff_2d(O.fft2(u), cut, DX, DX))
    U = None
    for st in steps:
        if st[0] == "gap":
            u, U = O.propagate_2d(u, DX, DX, WL, st[1], cut)
        else:
            _, grid, first, n = st
            spec = O.SampleSpec(grid=grid, db_list=DB, pixel_size_x=DX, pixel_size_y=DX,
                                pixel_size_z=PZ)
            u, U = O.sample_apply_2d(u, spec, DX, DX, WL, cut, first_slice=first, n_slices=n)
    return u, (O.fft2(u) if U is None else U)


CHAINS = {
    "free_space": (True, [("gap", 0.02), ("gap", 0.01)]),
    "sample_first": (False, [("vol", 0, 3)]),
    "gap_between": (True, [("vol", 0, 2), ("gap", 0.05), ("vol", 2, 3), ("gap", 0.02)]),
    "gap_first": (True, [("gap", 0.1), ("vol", 1, 4)]),
    "zero_slices": (True, [("vol", 0, 0), ("gap", 0.03), ("vol", 0, 2)]),
    "nothing": (True, []),
}


@pytest.mark.parametrize("n", [256, 384])
@pytest.mark.parametrize("name", sorted(CHAINS))
@pytest.mark.parametrize("dt,engine,tol", [(np.complex128, lib.RSIM_PLAIN, 1e-12),
                                           (np.complex64, lib.RSIM_FUSED, 3e-6)])
def test_chain_matches_oracle(name, n, dt, engine, tol):
    band_limit, spec = CHAINS[name]
    grid = volume(n, 6, 11)
    vol = rs.DeviceVolume.from_numpy(grid, DB)
    cut = 0.4 / DX
    rng = np.random.default_rng(5)
    u0 = np.exp(1j * rng.uniform(0, 0.3, size=(n, n))) * (1 + 0.1 * rng.standard_normal((n, n)))
    steps = []
    ostep = []
    for st in spec:
        if st[0] == "gap":
            steps.append(lib.ChainStep(None, 0, 0, 0, 0, 0, 0, 0, st[1]))
            ostep.append(st)
        else:
            steps.append(lib.ChainStep(vol.handle, DX, DX, PZ, 0.0, 0.0, st[1], st[2], 0.0))
            ostep.append(("vol", grid, st[1], st[2]))
    arr = (lib.ChainStep * max(len(steps), 1))(*steps)
    u = rs.DeviceVector.from_numpy(u0.astype(dt).reshape(-1))
    U = rs.DeviceVector(n * n, dt)
    say("chain", name, n, np.dtype(dt).name)
    lib.call("rsim_chain_apply_2d", u.ptr, U.ptr, n, n, u.code, DX, DX, WL, cut, int(band_limit),
             arr, len(steps), engine, 0, None)
    ue, Ue = oracle_chain(u0, band_limit, ostep, cut)
    eu = O.relative_l2(u.to_numpy().reshape(n, n), ue)
    eU = O.relative_l2(U.to_numpy().reshape(n, n), Ue)
    say(f"  field {eu:.2e} spectrum {eU:.2e}")
    assert eu < tol and eU < tol


def test_fused_chain_and_its_steps_against_the_oracle():
    """The fused chain and the same steps issued one call at a time on the
    complex64 engine (band limit, 4 slices, 0.3 m, 4 slices, 5 cm), both
    against the FP64 oracle: folding the free space into the column passes
    costs no accuracy."""
    n = 512
    grid = volume(n, 8, 3)
    vol = rs.DeviceVolume.from_numpy(grid, DB)
    cut = 0.4 / DX
    rng = np.random.default_rng(8)
    u0 = np.exp(1j * rng.uniform(0, 0.3, size=(n, n))).astype(np.complex64)
    steps = [lib.ChainStep(vol.handle, DX, DX, PZ, 0, 0, 0, 4, 0),
             lib.ChainStep(None, 0, 0, 0, 0, 0, 0, 0, 0.3),
             lib.ChainStep(vol.handle, DX, DX, PZ, 0, 0, 4, 4, 0),
             lib.ChainStep(None, 0, 0, 0, 0, 0, 0, 0, 0.05)]
    arr = (lib.ChainStep * 4)(*steps)
    u = rs.DeviceVector.from_numpy(u0.reshape(-1))
    U = rs.DeviceVector(n * n, np.complex64)
    lib.call("rsim_chain_apply_2d", u.ptr, U.ptr, n, n, u.code, DX, DX, WL, cut, 1, arr, 4,
             lib.RSIM_FUSED, 0, None)
    chained = u.to_numpy()
    v = rs.DeviceVector.from_numpy(u0.reshape(-1))
    V = rs.DeviceVector(n * n, np.complex64)
    c64 = lib.RSIM_C64
    lib.call("rsim_fft2", V.ptr, v.ptr, n, n, c64, lib.RSIM_FORWARD, None)
    lib.call("rsim_apply_frequency_cutoff_2d", V.ptr, n, n, c64, cut, DX, DX, None)
    lib.call("rsim_fft2", v.ptr, V.ptr, n, n, c64, lib.RSIM_INVERSE, None)
    lib.call("rsim_sample_apply_2d", v.ptr, V.ptr, n, n, c64, DX, DX, WL, cut, vol.handle, DX, DX,
             PZ, 0.0, 0.0, 0, 4, lib.RSIM_FUSED, 0, None)
    lib.call("rsim_propagate_2d", v.ptr, V.ptr, n, n, c64, DX, DX, WL, 0.3, cut, 1, None)
    lib.call("rsim_sample_apply_2d", v.ptr, V.ptr, n, n, c64, DX, DX, WL, cut, vol.handle, DX, DX,
             PZ, 0.0, 0.0, 4, 4, lib.RSIM_FUSED, 0, None)
    lib.call("rsim_propagate_2d", v.ptr, V.ptr, n, n, c64, DX, DX, WL, 0.05, cut, 1, None)
    ue, _ = oracle_chain(u0.astype(np.complex128), True,
                         [("vol", grid, 0, 4), ("gap", 0.3), ("vol", grid, 4, 4), ("gap", 0.05)],
                         cut)
    e_chain = O.relative_l2(chained.reshape(n, n), ue)
    e_steps = O.relative_l2(v.to_numpy().reshape(n, n), ue)
    say(f"vs oracle: fused chain {e_chain:.2e}, step by step {e_steps:.2e}")
    assert e_chain < 1.5 * e_steps + 1e-7
    assert e_chain < 1e-5


def test_chain_argument_checks():
    n = 256
    u = rs.DeviceVector(n * n, np.complex64)
    U = rs.DeviceVector(n * n, np.complex64)
    bad = (lib.ChainStep * 1)(lib.ChainStep(None, 0, 0, 0, 0, 0, 0, 0, float("nan")))
    with pytest.raises(ValueError, match="finite"):
        lib.call("rsim_chain_apply_2d", u.ptr, U.ptr, n, n, u.code, DX, DX, WL, 0.4 / DX, 1, bad,
                 1, lib.RSIM_FUSED, 0, None)
    grid = volume(n, 2, 1)
    vol = rs.DeviceVolume.from_numpy(grid, DB)
    over = (lib.ChainStep * 1)(lib.ChainStep(vol.handle, DX, DX, PZ, 0, 0, 1, 5, 0))
    with pytest.raises(ValueError, match="chain step 0"):
        lib.call("rsim_chain_apply_2d", u.ptr, U.ptr, n, n, u.code, DX, DX, WL, 0.4 / DX, 1, over,
                 1, lib.RSIM_FUSED, 0, None)
    with pytest.raises(ValueError):
        lib.call("rsim_chain_apply_2d", u.ptr, U.ptr, n, n, u.code, DX, DX, WL, 0.4 / DX, 1, None,
                 2, lib.RSIM_FUSED, 0, None)


def _instances(n, nz, dt, band_rng=5):
    """Five instances of one grid with their own wavelengths, cutoffs,
    distances and volumes; one volume is sampled at 1.06 dx (bilinear
    transmission), the others on the field grid (material lookup)."""
    rng = np.random.default_rng(band_rng)
    out = []
    for i in range(5):
        grid = volume(n, nz, 20 + i)
        vol = rs.DeviceVolume.from_numpy(grid, DB)
        wl = WL * (0.8 + 0.1 * i)
        px = DX * (1.06 if i == 3 else 1.0)
        steps = (lib.ChainStep * 4)(
            lib.ChainStep(vol.handle, px, px, PZ, 0.0, 0.0, 0, 2, 0.0),
            lib.ChainStep(None, 0, 0, 0, 0, 0, 0, 0, 0.01 * (i + 1)),
            lib.ChainStep(vol.handle, px, px, PZ, 0.0, 0.0, 2, 3, 0.0),
            lib.ChainStep(None, 0, 0, 0, 0, 0, 0, 0, 0.02))
        u0 = np.exp(1j * rng.uniform(0, 0.3, size=(n, n))).astype(dt)
        out.append(dict(vol=vol, steps=steps, wl=wl, cut=(0.35 + 0.02 * i) / DX, u0=u0))
    return out


@pytest.mark.parametrize("dt,engine", [(np.complex64, lib.RSIM_FUSED), (np.complex128, lib.RSIM_PLAIN)])
def test_batched_chain_equals_single_chains(dt, engine):
    """rsim_chain_apply_2d_batch over five instances (different wavelengths,
    cutoffs, distances, volumes and transmission modes) equals five calls of
    rsim_chain_apply_2d, bit for bit, field and spectrum."""
    n = 256
    inst = _instances(n, 5, dt)
    say("batched chain", np.dtype(dt).name)
    single = []
    for q in inst:
        u = rs.DeviceVector.from_numpy(q["u0"].reshape(-1))
        U = rs.DeviceVector(n * n, dt)
        lib.call("rsim_chain_apply_2d", u.ptr, U.ptr, n, n, u.code, DX, DX, q["wl"], q["cut"], 1,
                 q["steps"], 4, engine, 0, None)
        single.append((u.to_numpy(), U.to_numpy()))
    fields = [(rs.DeviceVector.from_numpy(q["u0"].reshape(-1)), rs.DeviceVector(n * n, dt))
              for q in inst]
    rows = [lib.ChainInstance(u.ptr, U.ptr, DX, DX, q["wl"], q["cut"], q["steps"])
            for (u, U), q in zip(fields, inst)]
    arr = (lib.ChainInstance * len(rows))(*rows)
    lib.call("rsim_chain_apply_2d_batch", arr, len(rows), n, n, fields[0][0].code, 1, 4, engine, 0,
             None)
    for k, ((u, U), (su, sU)) in enumerate(zip(fields, single)):
        assert np.array_equal(u.to_numpy(), su), k
        assert np.array_equal(U.to_numpy(), sU), k


def test_batched_chain_instance_against_the_oracle():
    """One instance of a fused batch (bilinear map, 1.06 dx) against the FP64
    oracle run step by step."""
    n = 256
    inst = _instances(n, 5, np.complex64)
    fields = [(rs.DeviceVector.from_numpy(q["u0"].reshape(-1)), rs.DeviceVector(n * n, np.complex64))
              for q in inst]
    rows = [lib.ChainInstance(u.ptr, U.ptr, DX, DX, q["wl"], q["cut"], q["steps"])
            for (u, U), q in zip(fields, inst)]
    arr = (lib.ChainInstance * len(rows))(*rows)
    lib.call("rsim_chain_apply_2d_batch", arr, len(rows), n, n, lib.RSIM_C64, 1, 4, lib.RSIM_FUSED,
             lib.RSIM_FLAG_NO_SPECTRUM, None)
    for k in (0, 3):
        q = inst[k]
        grid = volume(n, 5, 20 + k)
        px = DX * (1.06 if k == 3 else 1.0)
        u = q["u0"].astype(np.complex128)
        u = O.ifft2(O.apply_frequency_cutoff_2d(O.fft2(u), q["cut"], DX, DX))
        spec = O.SampleSpec(grid=grid, db_list=DB, pixel_size_x=px, pixel_size_y=px, pixel_size_z=PZ)
        u, _ = O.sample_apply_2d(u, spec, DX, DX, q["wl"], q["cut"], first_slice=0, n_slices=2)
        u, _ = O.propagate_2d(u, DX, DX, q["wl"], 0.01 * (k + 1), q["cut"])
        u, _ = O.sample_apply_2d(u, spec, DX, DX, q["wl"], q["cut"], first_slice=2, n_slices=3)
        u, _ = O.propagate_2d(u, DX, DX, q["wl"], 0.02, q["cut"])
        err = O.relative_l2(fields[k][0].to_numpy().reshape(n, n), u)
        say(f"  batched instance {k} vs oracle {err:.2e}")
        assert err < 3e-6


def test_batched_chain_argument_checks():
    n = 256
    inst = _instances(n, 5, np.complex64)
    fields = [(rs.DeviceVector(n * n, np.complex64), rs.DeviceVector(n * n, np.complex64)) for _ in range(2)]
    other = (lib.ChainStep * 4)(*[inst[1]["steps"][j] for j in (1, 0, 2, 3)])   # gap first
    rows = [lib.ChainInstance(fields[0][0].ptr, fields[0][1].ptr, DX, DX, WL, 0.4 / DX, inst[0]["steps"]),
            lib.ChainInstance(fields[1][0].ptr, fields[1][1].ptr, DX, DX, WL, 0.4 / DX, other)]
    with pytest.raises(ValueError, match="same step kinds"):
        lib.call("rsim_chain_apply_2d_batch", (lib.ChainInstance * 2)(*rows), 2, n, n, lib.RSIM_C64, 1,
                 4, lib.RSIM_FUSED, 0, None)
    rows[1] = lib.ChainInstance(fields[0][0].ptr, fields[1][1].ptr, DX, DX, WL, 0.4 / DX,
                                inst[1]["steps"])
    with pytest.raises(ValueError, match="share a field"):
        lib.call("rsim_chain_apply_2d_batch", (lib.ChainInstance * 2)(*rows), 2, n, n, lib.RSIM_C64, 1,
                 4, lib.RSIM_FUSED, 0, None)
