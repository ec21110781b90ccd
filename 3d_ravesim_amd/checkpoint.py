"""Runtime checkpoints of one source's run: the fields u and U plus where to
continue, in one self-describing file that is replaced atomically.

The interface is the one ``wavesim.run_simulation`` and the runner use
(big-wave/checkpoint.py:83-208: ``RunCheckpoint(directory, fingerprint,
progress_cb, cancel_token)`` with ``exists``, ``save``, ``load`` and
``mark_complete``; ``CheckpointState``; ValueError for anything that cannot
be resumed, InterruptedError on cancel).  The storage is this package's own:

    run.ckpt  =  payload  |  trailer (UTF-8 JSON)  |  trailer length (u64 LE)  |  MAGIC

The payload is u's bytes followed by U's bytes, copied from HBM one field at
a time and hashed (BLAKE2b-256) while they are written, so a save reads
nothing back.  The trailer records the format, the run fingerprint, the
state to resume from, each field's dtype, length, offset and digest.  A save
writes ``run.ckpt.partial``, flushes it to stable storage and renames it
over ``run.ckpt`` -- a reader sees the previous checkpoint or the new one,
never a mixture.  ``mark_complete`` renames the file to
``run.ckpt.complete``: the run finished, and there is nothing to resume.
"""

from __future__ import annotations

import hashlib
import json
import os
import struct
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Any, Callable

import numpy as np

from ._control import token_is_set

FORMAT = "ravesim-run-checkpoint"
FORMAT_VERSION = 2
MAGIC = b"RSCKPT02"
_TAIL = struct.Struct("<Q8s")           # trailer length, magic
_COPY_BLOCK = 64 * 1024 ** 2            # bytes hashed / written per step
_FIELDS = ("u", "U")


@dataclass(frozen=True)
class CheckpointState:
    """Where the run continues: after ``slice_index`` slices of element
    ``element_index`` at ``current_z``; ``u_fourier_valid`` tells whether U
    holds FFT2(u)."""

    current_z: float
    element_index: int
    slice_index: int
    phase_step: int
    u_fourier_valid: bool
    stage: str


def _host_copy(vector) -> np.ndarray:
    if not hasattr(vector, "to_numpy"):
        raise TypeError(f"cannot checkpoint a {type(vector).__name__}: it has no to_numpy()")
    return np.ascontiguousarray(vector.to_numpy()).reshape(-1)


def _stream(handle, array: np.ndarray) -> str:
    """Write ``array``'s bytes to ``handle`` in blocks; its BLAKE2b-256 digest."""
    digest = hashlib.blake2b(digest_size=32)
    raw = memoryview(array.view(np.uint8))
    for start in range(0, len(raw), _COPY_BLOCK):
        block = raw[start:start + _COPY_BLOCK]
        digest.update(block)
        handle.write(block)
    return digest.hexdigest()


def _digest_of(view: memoryview) -> str:
    digest = hashlib.blake2b(digest_size=32)
    for start in range(0, len(view), _COPY_BLOCK):
        digest.update(view[start:start + _COPY_BLOCK])
    return digest.hexdigest()


def _fsync_dir(directory: Path) -> None:
    try:
        fd = os.open(directory, os.O_RDONLY)
    except OSError:
        return
    try:
        os.fsync(fd)
    except OSError:
        pass
    finally:
        os.close(fd)


class RunCheckpoint:
    """The checkpoint of one source under ``directory``; ``fingerprint`` (a
    JSON-able dict holding at least the field length ``N`` and numpy dtype
    string ``dtype``) must match for a checkpoint to be loaded."""

    def __init__(self, directory: Path | str, fingerprint: dict[str, Any],
                 progress_cb: Callable[[str, int, int], Any] | None = None,
                 cancel_token: Any | None = None) -> None:
        self.directory = Path(directory)
        self.fingerprint = fingerprint
        self.progress_cb = progress_cb
        self.cancel_token = cancel_token
        self.path = self.directory / "run.ckpt"
        self.completed_path = self.directory / "run.ckpt.complete"

    # -- state --
    @property
    def exists(self) -> bool:
        """A checkpoint is ready to resume from."""
        return self.path.exists()

    def _step(self, stage: str, completed: int, total: int) -> None:
        if token_is_set(self.cancel_token):
            raise InterruptedError(f"checkpoint cancelled during {stage}")
        if self.progress_cb is not None:
            self.progress_cb(stage, completed, total)

    def _canonical_fingerprint(self) -> Any:
        # what the trailer will hold after a JSON round trip
        return json.loads(json.dumps(self.fingerprint, sort_keys=True))

    # -- write --
    def save(self, u, U, *, current_z: float, element_index: int, slice_index: int,
             phase_step: int = 0, u_fourier_valid: bool, stage: str) -> None:
        state = CheckpointState(float(current_z), int(element_index), int(slice_index),
                                int(phase_step), bool(u_fourier_valid), str(stage))
        self.directory.mkdir(parents=True, exist_ok=True)
        partial = self.path.with_name(self.path.name + ".partial")
        layout = []
        offset = 0
        try:
            with partial.open("wb") as handle:
                for k, (name, vector) in enumerate(zip(_FIELDS, (u, U))):
                    self._step("checkpoint_u" if name == "u" else "checkpoint_U", k, 3)
                    array = _host_copy(vector)
                    digest = _stream(handle, array)
                    layout.append({"name": name, "dtype": array.dtype.str, "count": int(array.size),
                                   "offset": offset, "nbytes": int(array.nbytes),
                                   "blake2b": digest})
                    offset += int(array.nbytes)
                    del array
                trailer = json.dumps({
                    "format": FORMAT, "version": FORMAT_VERSION,
                    "fingerprint": self.fingerprint, "state": asdict(state),
                    "fields": layout,
                }, sort_keys=True).encode()
                handle.write(trailer)
                handle.write(_TAIL.pack(len(trailer), MAGIC))
                handle.flush()
                os.fsync(handle.fileno())
            os.replace(partial, self.path)
        except BaseException:
            partial.unlink(missing_ok=True)
            raise
        self.completed_path.unlink(missing_ok=True)
        _fsync_dir(self.directory)
        self._step("checkpoint_commit", 3, 3)

    # -- read --
    def _trailer(self, size: int, handle) -> tuple[dict, int]:
        if size < _TAIL.size:
            raise ValueError(f"checkpoint {self.path} is truncated")
        handle.seek(size - _TAIL.size)
        length, magic = _TAIL.unpack(handle.read(_TAIL.size))
        if magic != MAGIC or length > size - _TAIL.size:
            raise ValueError(f"{self.path} is not a checkpoint of this format")
        handle.seek(size - _TAIL.size - length)
        try:
            trailer = json.loads(handle.read(length).decode())
        except ValueError as exc:
            raise ValueError(f"checkpoint trailer of {self.path} is unreadable: {exc}") from None
        if trailer.get("format") != FORMAT or int(trailer.get("version", 0)) != FORMAT_VERSION:
            raise ValueError("unsupported checkpoint format")
        return trailer, size - _TAIL.size - length

    def load(self, u, U) -> CheckpointState:
        """Verify the checkpoint and restore u and U from it."""
        if not self.path.exists():
            raise ValueError(f"no ready checkpoint at {self.path}")
        size = self.path.stat().st_size
        with self.path.open("rb") as handle:
            trailer, payload_end = self._trailer(size, handle)
        if trailer.get("fingerprint") != self._canonical_fingerprint():
            raise ValueError("checkpoint fingerprint differs from this simulation/source")
        want_count = int(self.fingerprint["N"])
        want_dtype = np.dtype(self.fingerprint["dtype"])
        fields = {f["name"]: f for f in trailer.get("fields", [])}
        if set(fields) != set(_FIELDS):
            raise ValueError("checkpoint does not hold both u and U")
        raw = np.memmap(self.path, dtype=np.uint8, mode="r", shape=(size,))
        try:
            views = {}
            for name in _FIELDS:
                f = fields[name]
                if int(f["count"]) != want_count or np.dtype(f["dtype"]) != want_dtype:
                    raise ValueError(
                        f"checkpoint field {name} is {f['count']} x {f['dtype']}, "
                        f"expected {want_count} x {want_dtype.str}")
                lo, hi = int(f["offset"]), int(f["offset"]) + int(f["nbytes"])
                if hi > payload_end or hi - lo != want_count * want_dtype.itemsize:
                    raise ValueError(f"checkpoint field {name} lies outside the payload")
                view = memoryview(raw[lo:hi])
                if _digest_of(view) != f["blake2b"]:
                    raise ValueError(f"checkpoint checksum mismatch in field {name}")
                views[name] = raw[lo:hi].view(want_dtype)
            self._step("checkpoint_restore", 0, 2)
            for vector, name in ((u, "u"), (U, "U")):
                if not hasattr(vector, "upload"):
                    raise TypeError(f"cannot restore into a {type(vector).__name__}")
                vector.upload(np.array(views[name]))
            self._step("checkpoint_restore", 2, 2)
        finally:
            del raw
        return CheckpointState(**trailer["state"])

    def mark_complete(self) -> None:
        """The run finished: keep the last checkpoint as ``run.ckpt.complete``
        so that it is no longer offered for resume."""
        if self.path.exists():
            os.replace(self.path, self.completed_path)
            _fsync_dir(self.directory)
