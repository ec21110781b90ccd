"""History frames and runtime checkpoints: host logic on the CPU.

Mirrors the reference's own P4 tests (tests/big_wave_2d/test_p4.py:102-290):
canonical (z, y, x) order, ROI and downsampling, frame limit and
cancellation, the disk spill of frames, the checksummed checkpoint file, and a slice-level resume
that replays neither the source nor the finished slices.  The field vectors
are host stand-ins with the DeviceVector transfer methods, so no GPU work
runs here.  Also pins the oracle's history helpers to hand-derived values.
"""

import importlib
import json

import numpy as np
import pytest

import oracle as O
from _rs import rs

H = importlib.import_module("3d_ravesim_amd.history")
CK = importlib.import_module("3d_ravesim_amd.checkpoint")
WS = importlib.import_module("3d_ravesim_amd.wavesim")


class HostVec:
    """Host vector with the transfer methods the driver and checkpoints use."""

    def __init__(self, a):
        self.vec = np.array(a)

    def to_numpy(self):
        return self.vec.copy()

    def upload(self, a):
        self.vec[:] = a

    def copy(self):
        return HostVec(self.vec.copy())

    def copy_to(self, d):
        d.vec[:] = self.vec

    def drop(self):
        pass


@pytest.mark.parametrize("dz,hist,expected", [
    (1.0, 0.3, [0.0, 0.3, 0.6, 1.0]),      # 0.1 left < 0.15: last step stretched
    (1.0, 0.25, [0.0, 0.25, 0.5, 0.75, 1.0]),  # a whole step left: one more step
    (0.1, 1.0, [0.0, 0.1]),                # shorter than one step
    (0.0, 0.5, [0.0, 0.0]),
])
def test_history_z_steps_known_answers(dz, hist, expected):
    np.testing.assert_allclose(H._history_z_steps(dz, hist), expected, rtol=0, atol=1e-15)
    np.testing.assert_allclose(O.history_z_steps(dz, hist), expected, rtol=0, atol=1e-15)


def test_oracle_history_leg_ends_at_the_plain_propagation():
    rng = np.random.default_rng(2)
    u = np.exp(1j * rng.uniform(0, 1, (32, 32)))
    dx, wl, cut = 1e-6, 1e-10, 0.4e6
    frame = (32e-6, 32e-6, 2e-6, 2e-6)
    uh, Uh, frames = O.propagate_with_history_2d(u, dx, dx, wl, 0.05, cut, 0.3, 0.02, frame)
    up, Up = O.propagate_2d(u, dx, dx, wl, 0.05, cut)
    assert [z for _, z in frames] == pytest.approx([0.32, 0.34, 0.35])
    assert np.array_equal(uh, up) and np.array_equal(Uh, Up)
    assert np.allclose(frames[-1][0], O.area_v1(up, dx, dx, *frame, 0.35), rtol=1e-15, atol=0)


def test_oracle_analytic_history_centre_and_cone():
    frames = O.analytic_history_2d([0.5, 1.0], 0.0, 0.0, 8e-6, 8e-6, 1e-6, 1e-6, 1e-10, 1e4)
    img, z = frames[1]
    assert img.shape == (8, 8) and z == 1.0
    assert img[4, 4] == pytest.approx(1e-12 / 1.0, rel=1e-12)   # ds^2 z / z^3
    # cone |x| <= cutoff * wl * z = 1e-6 m: only the 3 x 3 centre survives
    assert np.count_nonzero(img) == 9


def test_canonical_axis_order_for_small_in_memory_history():
    h = H.History()
    h.push(np.full((2, 3), 1.0), 0.1)
    h.push(np.full((2, 3), 2.0), 0.2)
    v = h.get_history()
    assert v.shape == (2, 2, 3)
    np.testing.assert_array_equal(v[:, 0, 0], [1.0, 2.0])
    assert h.metadata()["axis_order"] == ["z", "y", "x"]


def test_roi_and_downsample_before_accounting():
    h = H.History(downsample=(2, 2), roi=(1, 5, 1, 7))
    frame = np.arange(48, dtype=np.float64).reshape(6, 8)
    h.push(frame, 0.1)
    h.push(frame + 100, 0.2)
    assert h.metadata()["shape"] == [2, 2, 3]
    np.testing.assert_array_equal(h.get_history(), np.stack([frame[1:5:2, 1:7:2],
                                                            (frame + 100)[1:5:2, 1:7:2]]))
    with pytest.raises(ValueError, match="shape changed"):
        h.push(np.ones(4), 0.3)


def test_max_frames_and_cancel_are_hard_gates():
    h = H.History(max_frames=1)
    h.push(np.ones((2, 2)), 0.1)
    with pytest.raises(RuntimeError, match="maximum frame"):
        h.push(np.ones((2, 2)), 0.2)
    with pytest.raises(InterruptedError):
        H.History(cancel_token=True).push(np.ones((2, 2)), 0.1)


def test_budget_overflow_spills_to_disk_and_resumes(tmp_path):
    """A zero budget spills from the first frame (big-wave streams at that
    point); frames read back equal, and a resumed History appends to them."""
    path = tmp_path / "history"
    h = H.History(path, memory_budget_bytes=0, downsample=2)
    frames = [np.arange(36, dtype=np.float64).reshape(6, 6) + 100 * k for k in range(3)]
    for k, f in enumerate(frames):
        h.push(f, 0.1 * (k + 1))
    assert h.is_streaming and h.metadata()["storage"] == "spill"
    np.testing.assert_array_equal(h.get_history(), np.stack([f[::2, ::2] for f in frames]))
    h.close()
    again = H.History(path, memory_budget_bytes=0, downsample=2, resume=True)
    assert len(again) == 3 and again.get_z() == pytest.approx([0.1, 0.2, 0.3])
    again.push(frames[0] - 1, 0.4)
    got = again.get_history()
    assert got.shape == (4, 3, 3) and got[3, 0, 0] == -1
    again.remove_spill()
    assert not any(tmp_path.iterdir())


def test_positive_budget_spills_only_when_exceeded(tmp_path):
    h = H.History(tmp_path / "history", memory_budget_bytes=2 * 4 * 8)
    h.push(np.ones((2, 2)), 0.1)
    h.push(np.ones((2, 2)) * 2, 0.2)
    assert not h.is_streaming
    h.push(np.ones((2, 2)) * 3, 0.3)
    assert h.is_streaming
    np.testing.assert_array_equal(h.get_history()[:, 0, 0], [1, 2, 3])
    # without a storage path everything stays in memory
    m = H.History(memory_budget_bytes=0)
    m.push(np.ones((2, 2)), 0.1)
    assert not m.is_streaming


def test_legacy_axis_migration_does_not_overwrite_source(tmp_path):
    legacy = np.arange(24).reshape(2, 3, 4)
    np.save(tmp_path / "history.npy", legacy)
    np.save(tmp_path / "history_z.npy", [0.1, 0.2, 0.3, 0.4])
    loaded = H.load_history(tmp_path)
    assert loaded.values.shape == (4, 2, 3)
    target = H.migrate_legacy_history(tmp_path)
    assert target.name == "history_v2.npy"
    np.testing.assert_array_equal(np.load(tmp_path / "history.npy"), legacy)
    migrated = H.load_history(target)
    np.testing.assert_array_equal(migrated.values, np.transpose(legacy, (2, 0, 1)))


def test_checkpoint_restores_both_vectors_and_rejects_damage(tmp_path):
    ck = CK.RunCheckpoint(tmp_path, {"N": 8, "dtype": np.dtype(np.complex64).str})
    u = HostVec(np.arange(8, dtype=np.complex64))
    U = HostVec((np.arange(8) * 3j).astype(np.complex64))
    stages = []
    ck.progress_cb = lambda *a: stages.append(a)
    ck.save(u, U, current_z=2.5, element_index=1, slice_index=4, u_fourier_valid=True,
            stage="element_slice")
    assert stages[-1] == ("checkpoint_commit", 3, 3)
    assert sorted(p.name for p in tmp_path.iterdir()) == ["run.ckpt"]
    u.vec[:] = 0
    U.vec[:] = 0
    st = ck.load(u, U)
    np.testing.assert_array_equal(u.vec, np.arange(8, dtype=np.complex64))
    np.testing.assert_array_equal(U.vec, (np.arange(8) * 3j).astype(np.complex64))
    assert (st.current_z, st.element_index, st.slice_index, st.u_fourier_valid,
            st.stage) == (2.5, 1, 4, True, "element_slice")
    # a flipped payload byte (inside U) fails its digest
    raw = bytearray((tmp_path / "run.ckpt").read_bytes())
    raw[8 * 8 + 5] ^= 0xFF
    (tmp_path / "run.ckpt").write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        ck.load(u, U)
    other = CK.RunCheckpoint(tmp_path, {"N": 8, "dtype": np.dtype(np.complex128).str})
    with pytest.raises(ValueError, match="fingerprint"):
        other.load(u, U)
    # a truncated file is not mistaken for a checkpoint
    (tmp_path / "run.ckpt").write_bytes(bytes(raw[:40]))
    with pytest.raises(ValueError):
        ck.load(u, U)


def test_checkpoint_cancel_complete_and_atomic_replace(tmp_path):
    ck = CK.RunCheckpoint(tmp_path, {"N": 4, "dtype": "<c16"})
    a = HostVec(np.ones(4, np.complex128))
    ck.save(a, a, current_z=1.0, element_index=0, slice_index=0, u_fourier_valid=False,
            stage="source_propagated")
    before = (tmp_path / "run.ckpt").read_bytes()
    ck.cancel_token = lambda: True
    with pytest.raises(InterruptedError):
        ck.save(a, a, current_z=2.0, element_index=1, slice_index=0, u_fourier_valid=True,
                stage="element_complete")
    # the interrupted save left the previous checkpoint untouched, no partial file
    assert (tmp_path / "run.ckpt").read_bytes() == before
    assert sorted(p.name for p in tmp_path.iterdir()) == ["run.ckpt"]
    ck.cancel_token = None
    ck.mark_complete()
    assert not ck.exists
    with pytest.raises(ValueError, match="no ready checkpoint"):
        ck.load(a, a)


def test_cancel_token_forms():
    from importlib import import_module
    ctl = import_module("3d_ravesim_amd._control")
    import threading
    ev = threading.Event()
    assert not ctl.token_is_set(None) and ctl.token_is_set(True) and not ctl.token_is_set(False)
    assert not ctl.token_is_set(ev)
    ev.set()
    assert ctl.token_is_set(ev) and ctl.token_is_set(lambda: 1)

    class Flag:
        cancelled = True
    assert ctl.token_is_set(Flag())


class DummySource:
    z = 0.0

    def __init__(self, forbid=False):
        self.forbid = forbid

    def propagate_to(self, z_out, params, cutoff_freq, u, U, history):
        if self.forbid:
            raise AssertionError("source must not run while resuming")
        u.vec[:] = 0.0
        U.vec[:] = 0.0


class DummySliceElement:
    z_start = 1.0
    x_positions = np.array([0.0])

    def __init__(self, fail_after_first=False):
        self.fail_after_first = fail_after_first

    def store_deltabetas(self, table):
        pass

    def get_thickness(self):
        return 3.0

    def apply(self, u, U, params, cutoff_freq, stepping_iteration, history):
        start = int(getattr(self, "_checkpoint_start_slice", 0))
        callback = getattr(self, "_checkpoint_callback", None)
        for index in range(start, 3):
            u.vec += 1.0
            U.vec[:] = 2.0 * u.vec
            if callback is not None:
                callback(index + 1, self.z_start + index + 1)
            if self.fail_after_first:
                raise RuntimeError("simulated process interruption")


def _params():
    return rs.SimParams(N=16, nx=4, ny=4, dx=1.0, dy=1.0, z_detector=4.0, detector_size=4.0,
                        detector_size_x=4.0, detector_size_y=4.0, detector_pixel_size_x=1.0,
                        detector_pixel_size_y=1.0, wl=1e-10, chunk_size=8)


def test_slice_checkpoint_resumes_without_replaying_source_or_slice(tmp_path, monkeypatch):
    monkeypatch.setattr(WS, "square_and_downsample_2d", lambda u, p, z: np.zeros((1, 1)))
    monkeypatch.setattr(WS, "CHECKPOINT_SLICE_CADENCE", 1)
    ck = CK.RunCheckpoint(tmp_path / "checkpoint", {"N": 16, "dtype": np.dtype(np.complex64).str})
    u = HostVec(np.zeros(16, np.complex64))
    U = HostVec(np.zeros(16, np.complex64))
    with pytest.raises(RuntimeError, match="simulated"):
        WS.run_simulation(_params(), DummySource(), [DummySliceElement(True)], [0.0, 0.0], u, U,
                          [], checkpoint=ck)
    ru = HostVec(np.zeros(16, np.complex64))
    rU = HostVec(np.zeros(16, np.complex64))
    out = WS.run_simulation(_params(), DummySource(forbid=True), [DummySliceElement(False)],
                            [0.0, 0.0], ru, rU, [], checkpoint=ck, resume_checkpoint=True)
    np.testing.assert_array_equal(ru.vec, np.full(16, 3, dtype=np.complex64))
    assert len(out) == 1


def test_runtime_checkpoint_rejects_phase_snapshot_configuration(tmp_path):
    el = DummySliceElement()
    el.x_positions = np.array([0.0, 1.0])
    ck = CK.RunCheckpoint(tmp_path / "checkpoint", {"N": 16, "dtype": "<c8"})
    with pytest.raises(ValueError, match="phase stepping"):
        WS.run_simulation(_params(), DummySource(), [el], [0.0, 0.0],
                          HostVec(np.zeros(16, np.complex64)), HostVec(np.zeros(16, np.complex64)),
                          [], checkpoint=ck)


def test_drive_slices_chunks_on_the_cadence():
    calls = []

    class El:
        pixel_size_z = 0.5
        z_start = 2.0

    el = El()
    el._checkpoint_callback = lambda s, z: calls.append(("cb", s, z))
    el._checkpoint_cadence = 5
    H.drive_slices(lambda a, n, f: calls.append(("run", a, n)), el, None, None, 3, 12)
    assert calls == [("run", 3, 2), ("cb", 5, 4.5), ("run", 5, 5), ("cb", 10, 7.0),
                     ("run", 10, 2), ("cb", 12, 8.0)]
    calls.clear()
    del el._checkpoint_cadence
    H.drive_slices(lambda a, n, f: calls.append(("run", a, n)), el, None, None, 0, 2)
    assert calls == [("run", 0, 1), ("cb", 1, 2.5), ("run", 1, 1), ("cb", 2, 3.0)]


def test_history_axis_coordinates_follow_roi_and_stride():
    """history_x/y.npy: pixel i of a detector axis of ``count`` pixels sits at
    (i - floor(count/2)) * pitch, then the ROI span and the stride select
    (big-wave/multisim.py:475-484).  Hand-derived values."""
    MS = importlib.import_module("3d_ravesim_amd.multisim")
    np.testing.assert_array_equal(MS._frame_axis(5, 1.0, None, 1), [-2, -1, 0, 1, 2])
    np.testing.assert_array_equal(MS._frame_axis(6, 2.0, None, 1), [-6, -4, -2, 0, 2, 4])
    np.testing.assert_array_equal(MS._frame_axis(5, 1.0, (1, 4), 2), [-1, 1])
    np.testing.assert_array_equal(MS._frame_axis(8, 0.5, (2, 8), 3), [-1.0, 0.5])
