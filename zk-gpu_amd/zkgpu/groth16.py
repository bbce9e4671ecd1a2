"""Groth16 quotient polynomial H = (A B - C) / (x^n - 1) from the evaluations of A, B, C on the
domain (no counterpart in the reference; include/zkgpu.h "Groth16 quotient polynomial")."""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import check


COEFFICIENTS, ODD_COSET = 0, 1


class Groth16H:
    """form = COEFFICIENTS: a <- H's coefficients (bases tau^i Z(tau)/delta); ODD_COSET: a <- A B - C
    on the odd coset w_2n H (snarkjs keys' H bases; no division by Z)."""

    def __init__(self, gpu, log_n: int, form: int = COEFFICIENTS):
        h = ctypes.c_void_p()
        check(_lib.load().zkg_groth16_h_create_form(gpu.handle, int(log_n), int(form), ctypes.byref(h)))
        self._h = h
        self._gpu = gpu
        self.log_n = int(log_n)

    def run(self, gpu, a, b, c) -> None:
        """a <- H's coefficients (or the odd coset values); b, c are clobbered. All hold 2^log_n elements."""
        check(_lib.load().zkg_groth16_h_run(self._h, a.handle, b.handle, c.handle))

    def __del__(self):
        try:
            if self._h:
                _lib.load().zkg_groth16_h_destroy(self._h)
                self._h = ctypes.c_void_p(0)
        except Exception:
            pass
