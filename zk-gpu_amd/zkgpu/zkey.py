"""snarkjs Groth16 proving keys (.zkey, version 1) as MSM bases (no counterpart in the reference).

The reference accepts bases "already in Montgomery form with R = 2^256 (the form snarkjs/ffjavascript
keep points in)" (src/msm.rs:190-197) and leaves the snarkjs front end to another project
(README.md:224-225). `Zkey` maps a key file through the library (zkg_zkey_*), checks it (magic,
version, Groth16, BN254 q and r, section sizes) and uploads its point sections unchanged: the
stored words are already the library's device form.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

IC, A, B1, B2, C, H = 3, 5, 6, 7, 8, 9


class ZkeyInfo(ctypes.Structure):
    _fields_ = [
        ("n_vars", ctypes.c_uint32),
        ("n_public", ctypes.c_uint32),
        ("domain_size", ctypes.c_uint32),
        ("alpha1", ctypes.c_uint32 * 16),
        ("beta1", ctypes.c_uint32 * 16),
        ("beta2", ctypes.c_uint32 * 32),
        ("gamma2", ctypes.c_uint32 * 32),
        ("delta1", ctypes.c_uint32 * 16),
        ("delta2", ctypes.c_uint32 * 32),
    ]


class Zkey:
    def __init__(self, path: str):
        h = ctypes.c_void_p()
        check(_lib.load().zkg_zkey_open(str(path).encode(), ctypes.byref(h)))
        self._h = h
        info = ZkeyInfo()
        check(_lib.load().zkg_zkey_get_info(self._h, ctypes.byref(info)))
        self.n_vars, self.n_public, self.domain_size = info.n_vars, info.n_public, info.domain_size
        # verification key: plain affine words (infinity = zeros)
        self.vk = {name: np.array(getattr(info, name), dtype=np.uint32)
                   for name in ("alpha1", "beta1", "beta2", "gamma2", "delta1", "delta2")}

    @property
    def handle(self):
        return self._h

    def points(self, section: int) -> int:
        out = ctypes.c_size_t()
        check(_lib.load().zkg_zkey_points(self._h, int(section), ctypes.byref(out)))
        return out.value

    def upload(self, gpu, section: int):
        """The section's points as an MSM bases buffer (a copy: already the library's form)."""
        from .gpu import Buffer

        h = ctypes.c_void_p()
        check(_lib.load().zkg_zkey_upload(gpu.handle, self._h, int(section), ctypes.byref(h)))
        return Buffer(gpu, h.value)

    def close(self) -> None:
        if self._h:
            _lib.load().zkg_zkey_close(self._h)
            self._h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
