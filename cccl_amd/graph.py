"""hipGraph capture and replay of sort calls (cccl_amd_capture_begin/_end,
cccl_amd_graph_launch).  Record any sequence of Sorter / SegmentedSorter calls issued
on a non-default stream once, then replay it with a single graph launch."""

from __future__ import annotations

import contextlib
import ctypes

from . import _bindings as B
from .radix_sort import SortError, _stream_handle


class SortGraph:
    def __init__(self):
        self._exec = ctypes.c_void_p()

    @contextlib.contextmanager
    def capture(self, stream):
        """Calls issued on `stream` (a torch.cuda.Stream, not the default one) inside the
        block are recorded instead of run."""
        L = B.lib()
        h = _stream_handle(stream)
        rc = L.cccl_amd_capture_begin(h)
        if rc != 0:
            raise SortError("cccl_amd_capture_begin", rc)
        try:
            yield self
        finally:
            rc = L.cccl_amd_capture_end(h, ctypes.byref(self._exec))
            if rc != 0:
                raise SortError("cccl_amd_capture_end", rc)

    def launch(self, stream=None):
        rc = B.lib().cccl_amd_graph_launch(self._exec, _stream_handle(stream))
        if rc != 0:
            raise SortError("cccl_amd_graph_launch", rc)

    def __del__(self):
        try:
            if self._exec:
                B.lib().cccl_amd_graph_destroy(self._exec)
        except Exception:
            pass
