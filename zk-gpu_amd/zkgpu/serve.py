"""Proof-job server: resident proving keys, continuous batching (no counterpart in the reference).

A job = seven in-place NTTs of the key's size + the five Groth16 MSMs (A, B1 and B2 over the
witness, C over the witness from the key's C offset, H over the first outputs of the seventh
transform). Jobs are admitted in arrival order while the HBM budget and the in-flight limit allow,
and admitted jobs overlap on the device. Keys come from device buffers (`add_key`,
`add_key_shaped`) or from a snarkjs proving key (`add_zkey`).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

RESULT_WORDS = 4 * 32 + 64


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("jobs_per_s", ctypes.c_double),
        ("bytes_in_use", ctypes.c_size_t),
        ("peak_bytes", ctypes.c_size_t),
        ("inflight", ctypes.c_uint32),
        ("peak_inflight", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class KeyShape(ctypes.Structure):
    """MSM sizes of a key: points = (A, B1, C, H, B2); C's scalars start at witness[c_offset]."""

    _fields_ = [("log_n", ctypes.c_uint32), ("points", ctypes.c_size_t * 5), ("c_offset", ctypes.c_size_t),
                ("h_form", ctypes.c_int)]

    @classmethod
    def make(cls, log_n: int, points, c_offset: int = 0, h_form: int = 0) -> "KeyShape":
        """h_form: the H bases' convention for quotient jobs (groth16.COEFFICIENTS or ODD_COSET)."""
        return cls(int(log_n), (ctypes.c_size_t * 5)(*[int(p) for p in points]), int(c_offset), int(h_form))

    def as_dict(self):
        return {"log_n": self.log_n, "points": tuple(self.points), "c_offset": self.c_offset, "h_form": self.h_form}


class Server:
    def __init__(self, gpu, mem_budget_bytes: int, max_inflight: int = 4):
        h = ctypes.c_void_p()
        check(_lib.load().zkg_server_create(gpu.handle, int(mem_budget_bytes), int(max_inflight), ctypes.byref(h)))
        self._h = h
        self._gpu = gpu
        self._keep = {}  # key id -> bases buffers; job id -> its buffers

    def add_key(self, log_n: int, g1, g2) -> int:
        """g1 = (A, B1, C, H) bases buffers (H holds n - 1 points), g2 = B2 bases buffer."""
        arr = (ctypes.c_void_p * 4)(*[b.handle for b in g1])
        kid = ctypes.c_int()
        check(_lib.load().zkg_server_add_key(self._h, int(log_n), arr, g2.handle, ctypes.byref(kid)))
        self._keep[("key", kid.value)] = (tuple(g1), g2)
        return kid.value

    def add_key_shaped(self, shape: KeyShape, g1, g2) -> int:
        """As add_key with explicit MSM sizes (`KeyShape.make(log_n, points, c_offset)`)."""
        arr = (ctypes.c_void_p * 4)(*[b.handle for b in g1])
        kid = ctypes.c_int()
        check(_lib.load().zkg_server_add_key_shaped(self._h, ctypes.byref(shape), arr, g2.handle, ctypes.byref(kid)))
        self._keep[("key", kid.value)] = (tuple(g1), g2)
        return kid.value

    def add_zkey(self, zkey) -> int:
        """Upload a snarkjs proving key (`zkgpu.Zkey`) into server-owned HBM; returns the key id."""
        kid = ctypes.c_int()
        check(_lib.load().zkg_server_add_zkey(self._h, zkey.handle, ctypes.byref(kid)))
        return kid.value

    def key_shape(self, key: int) -> dict:
        out = KeyShape()
        check(_lib.load().zkg_server_key_shape(self._h, int(key), ctypes.byref(out)))
        return out.as_dict()

    def key_slot_bytes(self, key: int) -> int:
        out = ctypes.c_size_t()
        check(_lib.load().zkg_server_key_slot_bytes(self._h, int(key), ctypes.byref(out)))
        return out.value

    def slot_bytes(self, log_n: int) -> int:
        out = ctypes.c_size_t()
        check(_lib.load().zkg_server_slot_bytes(self._h, int(log_n), ctypes.byref(out)))
        return out.value

    def submit(self, key: int, polys, inverse_mask: int, witness) -> int:
        assert len(polys) == 7
        arr = (ctypes.c_void_p * 7)(*[p.handle for p in polys])
        jid = ctypes.c_uint64()
        check(_lib.load().zkg_server_submit(self._h, int(key), arr, int(inverse_mask), witness.handle,
                                            ctypes.byref(jid)))
        self._keep[("job", jid.value)] = (tuple(polys), witness)
        return jid.value

    def submit_groth16(self, key: int, a, b, c, witness) -> int:
        """A quotient job: a, b, c hold A, B, C's domain evaluations (clobbered; a receives the H
        scalars in the key's H form); then the five MSMs."""
        jid = ctypes.c_uint64()
        check(_lib.load().zkg_server_submit_groth16(self._h, int(key), a.handle, b.handle, c.handle, witness.handle,
                                                    ctypes.byref(jid)))
        self._keep[("job", jid.value)] = ((a, b, c), witness)
        return jid.value

    def done(self, job: int) -> bool:
        d = ctypes.c_int()
        check(_lib.load().zkg_server_done(self._h, int(job), ctypes.byref(d)))
        return bool(d.value)

    def wait(self, job: int):
        """Block until `job` finishes: (results as 192 u32 words, latency in ms)."""
        out = np.zeros(RESULT_WORDS, dtype=np.uint32)
        lat = ctypes.c_double()
        try:
            check(_lib.load().zkg_server_wait(self._h, int(job), _lib.u32_ptr(out), ctypes.byref(lat)))
        finally:
            self._keep.pop(("job", int(job)), None)
        return out, lat.value

    def stats(self) -> dict:
        st = ServerStats()
        check(_lib.load().zkg_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def reset_stats(self) -> None:
        check(_lib.load().zkg_server_reset_stats(self._h))

    def close(self) -> None:
        if self._h:
            _lib.load().zkg_server_destroy(self._h)
            self._h = ctypes.c_void_p(0)
            self._keep = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
