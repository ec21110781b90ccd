"""Captured pipelines: a sequence of library calls replayed as one graph launch.

``with CapturedGraph.record(stream) as g: <library calls>`` records every
call the block makes on ``stream`` (or with no stream) into one hipGraph;
``g.launch()`` replays it with a single launch.  The calls must have been
made once before with the same arguments (the tables, detector plans and
launch shapes they use are built then); a call that would build something
inside the capture raises NotImplementedError.  The reference launches every
kernel from the host with a synchronisation per slice
(fast-wave/fwcuda/kernels.cu:365); here a whole multi-instance run is one
submission.
"""

from __future__ import annotations

import ctypes as C

from . import _lib


class CapturedGraph:
    def __init__(self, handle: int):
        self._h = handle
        n = C.c_size_t(0)
        _lib.call("rsim_graph_info", handle, C.byref(n))
        self.nodes = int(n.value)

    @classmethod
    def record(cls, stream=None) -> "_Recording":
        return _Recording(stream)

    def launch(self, stream=None, sync: bool = False) -> None:
        if self._h is None:
            raise ValueError("graph destroyed")
        _lib.call("rsim_graph_launch", self._h, stream)
        if sync:
            _lib.call("rsim_synchronize", stream)

    def close(self) -> None:
        if self._h is not None:
            _lib.call("rsim_graph_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and _lib._lib is not None:
                _lib._lib.rsim_graph_destroy(self._h)
                self._h = None
        except Exception:
            pass


class _Recording:
    def __init__(self, stream):
        self.stream = stream
        self.graph: CapturedGraph | None = None

    def __enter__(self) -> "_Recording":
        _lib.call("rsim_graph_begin", self.stream)
        return self

    def __exit__(self, exc_type, exc, tb) -> bool:
        h = C.c_void_p()
        rc = _lib.load().rsim_graph_end(self.stream, C.byref(h))
        if exc_type is not None:   # the block's own error wins; the capture is closed
            if rc == _lib.RSIM_OK and h.value:
                _lib.call("rsim_graph_destroy", h.value)
            return False
        _lib.check(rc, "rsim_graph_end")
        self.graph = CapturedGraph(h.value)
        return False
