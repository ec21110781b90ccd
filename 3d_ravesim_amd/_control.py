"""Cancellation tokens and per-run control of the HIP slice loops.

Token protocol.  big-wave accepts the same cancellation tokens everywhere a
long operation can stop early (detector: big-wave/propagation.py:669-688;
history and checkpoints: big-wave/history.py, big-wave/checkpoint.py;
bfpy's out-of-core FFT: big-fourier/src/lib.rs:132-146): nothing, a bool, a
zero-argument callable, or an object with an ``is_cancelled``, ``is_set`` or
``cancelled`` attribute or method (``threading.Event`` is the usual one).
``token_is_set`` evaluates any of them.

Run control.  ``RunControl`` owns one ``rsim_control`` of the library and
binds it to the calling thread for the duration of a ``with`` block: the
slice loops started on that thread report their progress to it and stop at
the next slice boundary once it is cancelled.  Runs on other threads have
controls of their own, so two runs in one process never cancel each other.
"""

from __future__ import annotations

import ctypes as C
from typing import Any, Callable, Optional

from . import _lib

_TOKEN_FIELDS = ("is_cancelled", "is_set", "cancelled")


def token_is_set(token: Any) -> bool:
    """True when ``token`` asks to stop (see the module docstring)."""
    if token is None:
        return False
    if isinstance(token, bool):
        return token
    for field in _TOKEN_FIELDS:
        probe = getattr(token, field, None)
        if probe is not None:
            return bool(probe() if callable(probe) else probe)
    if callable(token):
        return bool(token())
    return False


class RunControl:
    """Progress and cooperative cancel of the library calls made on this thread.

    ``progress_cb(stage, completed, total)`` is called after every
    ``every``-th slice; the slice loop waits for the GPU before each report,
    so a report means the work is done.  ``cancel_token`` is polled at each
    report; once set, the loop stops at the next slice boundary with the
    field valid there and the call raises ``InterruptedError``.
    ``cancel()`` may be called from any thread.  ``completed`` is the number
    of slices the last slice loop on this control finished.
    """

    def __init__(self, progress_cb: Optional[Callable[[str, int, int], Any]] = None,
                 cancel_token: Any = None, every: int = 1):
        if every < 1:
            raise ValueError("progress interval must be at least one slice")
        self.progress_cb = progress_cb
        self.cancel_token = cancel_token
        self.every = int(every)
        self._handle = None
        self._trampoline = None
        self._bound = False

    # -- library handle --
    def _ensure(self):
        if self._handle is not None:
            return self._handle
        h = C.c_void_p()
        _lib.call("rsim_control_create", C.byref(h))
        self._handle = h.value
        if self.progress_cb is not None or self.cancel_token is not None:
            def relay(stage, done, total, _user):
                try:
                    if self.progress_cb is not None:
                        self.progress_cb(stage.decode(), int(done), int(total))
                finally:
                    if token_is_set(self.cancel_token):
                        _lib.load().rsim_control_request_cancel(self._handle)
            self._trampoline = _lib.PROGRESS_FN(relay)
            _lib.call("rsim_control_set_progress", self._handle, self._trampoline, None,
                      self.every)
        return self._handle

    def cancel(self) -> None:
        _lib.call("rsim_control_request_cancel", self._ensure())

    def clear(self) -> None:
        _lib.call("rsim_control_clear_cancel", self._ensure())

    @property
    def completed(self) -> int:
        done, total = C.c_int64(0), C.c_int64(0)
        _lib.call("rsim_control_completed", self._ensure(), C.byref(done), C.byref(total))
        return int(done.value)

    def __enter__(self) -> "RunControl":
        if token_is_set(self.cancel_token):
            raise InterruptedError("run cancelled before it started")
        _lib.call("rsim_control_bind", self._ensure())
        self._bound = True
        return self

    def __exit__(self, *exc) -> None:
        if self._bound:
            _lib.call("rsim_control_bind", None)
            self._bound = False

    def close(self) -> None:
        if self._handle is not None:
            if self._bound:
                self.__exit__(None, None, None)
            _lib.call("rsim_control_destroy", self._handle)
            self._handle = None

    def __del__(self):
        try:
            if getattr(self, "_handle", None) is not None and _lib._lib is not None:
                _lib._lib.rsim_control_destroy(self._handle)
                self._handle = None
        except Exception:
            pass


class run_scope:
    """``with run_scope(u):`` around the library calls of one operation on
    the field ``u``: when ``u`` carries a ``progress_cb`` or ``cancel_token``
    (and optionally ``progress_every``), a RunControl serves those calls and
    is released afterwards; otherwise nothing is bound."""

    def __init__(self, vector):
        cb = getattr(vector, "progress_cb", None)
        token = getattr(vector, "cancel_token", None)
        self.control = None
        if cb is not None or token is not None:
            self.control = RunControl(cb, token, int(getattr(vector, "progress_every", 1) or 1))

    def __enter__(self):
        if self.control is not None:
            self.control.__enter__()
        return self.control

    def __exit__(self, *exc) -> None:
        if self.control is not None:
            self.control.close()
