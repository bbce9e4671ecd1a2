"""NTT — mirror of the reference's `ntt` module (src/ntt.rs:16-68)."""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import check


class Ntt:
    """Plan for in-place, natural-order transforms of 2^log_n Fr elements (10 <= log_n <= 27)."""

    def __init__(self, gpu, log_n: int):
        h = ctypes.c_void_p()
        check(_lib.load().zkg_ntt_create(gpu.handle, int(log_n), ctypes.byref(h)))
        self._h = h
        self._gpu = gpu
        self.log_n = int(_lib.load().zkg_ntt_log_n(h))

    @staticmethod
    def new(gpu, log_n: int) -> "Ntt":
        return Ntt(gpu, log_n)

    def passes(self) -> int:
        """Kernel launches (LDS-tiled passes) per transform."""
        return int(_lib.load().zkg_ntt_passes(self._h))

    def forward(self, gpu, data) -> None:
        """In-place forward transform: natural order in, natural order out (asynchronous)."""
        check(_lib.load().zkg_ntt_forward(self._h, data.handle))

    def inverse(self, gpu, data) -> None:
        """In-place inverse transform, including the 1/n scaling (asynchronous)."""
        check(_lib.load().zkg_ntt_inverse(self._h, data.handle))

    def __del__(self):
        try:
            if self._h:
                _lib.load().zkg_ntt_destroy(self._h)
                self._h = ctypes.c_void_p(0)
        except Exception:
            pass
