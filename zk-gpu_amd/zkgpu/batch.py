"""Batches of MSMs and NTTs (no counterpart in the reference's API; see include/zkgpu.h)."""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import check


class Batch:
    """Queue many NTTs and MSMs of uneven size and run them as one unit on the device."""

    def __init__(self, gpu):
        h = ctypes.c_void_p()
        check(_lib.load().zkg_batch_create(gpu.handle, ctypes.byref(h)))
        self._h = h
        self._gpu = gpu
        self._keep = []  # buffers and plans referenced by queued jobs

    def add_ntt(self, data, log_n: int, inverse: bool = False) -> None:
        check(_lib.load().zkg_batch_add_ntt(self._h, data.handle, int(log_n), int(bool(inverse))))
        self._keep.append(data)

    def add_msm(self, plan, bases, scalars) -> None:
        check(_lib.load().zkg_batch_add_msm(self._h, plan._h, bases.handle, scalars.handle))
        self._keep.extend((plan, bases, scalars))

    def run(self) -> None:
        check(_lib.load().zkg_batch_run(self._h))

    def capture(self) -> None:
        """Capture the batch into one hipGraph; run() then replays it with one launch."""
        check(_lib.load().zkg_batch_capture(self._h))

    def captured(self) -> bool:
        return bool(_lib.load().zkg_batch_captured(self._h))

    def clear(self) -> None:
        check(_lib.load().zkg_batch_clear(self._h))
        self._keep = []

    def __del__(self):
        try:
            if self._h:
                _lib.load().zkg_batch_destroy(self._h)
                self._h = ctypes.c_void_p(0)
        except Exception:
            pass
