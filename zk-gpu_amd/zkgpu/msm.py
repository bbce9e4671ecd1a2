"""MSM — mirror of the reference's `msm` module (src/msm.rs:25-387)."""
from __future__ import annotations

import ctypes
import weakref
from typing import Optional, Tuple

import numpy as np

from . import _lib, host
from ._lib import check, u32_ptr
from .gpu import Buffer
from .group import Group

Fq2Words = Tuple[host.U256, host.U256]


class Msm:
    """Plan for MSMs of n points in one group with window c (src/msm.rs:95-180). Not re-entrant:
    the plan owns its workspace and the result buffer returned by `run`."""

    def __init__(self, gpu, group: Group, n: int, c: int, item_points: Optional[int] = None, _tuned: bool = False,
                 _precomputed: bool = False):
        h = ctypes.c_void_p()
        if _tuned:
            check(_lib.load().zkg_msm_create_tuned(gpu.handle, group.value, int(n), ctypes.byref(h)))
        elif _precomputed:
            check(_lib.load().zkg_msm_create_precomputed(gpu.handle, group.value, int(n), int(c),
                                                         int(item_points or 0), ctypes.byref(h)))
        else:
            check(_lib.load().zkg_msm_create(gpu.handle, group.value, int(n), int(c), int(item_points or 0),
                                             ctypes.byref(h)))
        self._h = h
        self._gpu = gpu
        self._group = group
        self._result = None

    @staticmethod
    def default_window(n: int) -> int:
        return int(_lib.load().zkg_msm_default_window(int(n)))

    @staticmethod
    def new(gpu, n: int, c: int) -> "Msm":
        return Msm(gpu, Group.G1, n, c, None)

    @staticmethod
    def with_item_points(gpu, n: int, c: int, item_points: int) -> "Msm":
        return Msm(gpu, Group.G1, n, c, item_points)

    @staticmethod
    def for_group(gpu, group: Group, n: int, c: int, item_points: Optional[int] = None) -> "Msm":
        return Msm(gpu, group, n, c, item_points)

    @staticmethod
    def plan_check(group: Group, n: int, c: int, item_points: Optional[int] = None,
                   precomputed: bool = False) -> int:
        """Device workspace bytes of the plan `for_group` (or `precomputed`) would make, checked
        without a device; raises ZkgError with the same message as the constructor for a plan it
        would reject (the reference's argument checks, src/msm.rs:96-98, and this library's 32-bit
        index limits). No counterpart in the reference."""
        out = ctypes.c_size_t()
        check(_lib.load().zkg_msm_plan_check(group.value, int(n), int(c), int(item_points or 0), int(precomputed),
                                             ctypes.byref(out)))
        return int(out.value)

    @staticmethod
    def tuned(gpu, group: Group, n: int) -> "Msm":
        """A plan with the tuning table's window, item and sort sizes for this GPU and size
        (default_window(n) when the size is untuned). No counterpart in the reference."""
        return Msm(gpu, group, n, 0, None, _tuned=True)

    @staticmethod
    def precomputed(gpu, group: Group, n: int, c: int, item_points: Optional[int] = None) -> "Msm":
        """A plan whose `run` takes `precompute_bases`'s table (2n points) as its bases: 4n virtual
        points over ceil(64 / c) windows, same result. c in 9..=16. No counterpart in the reference."""
        return Msm(gpu, group, n, c, item_points, _precomputed=True)

    @staticmethod
    def precompute_bases(gpu, group: Group, bases: Buffer, n: int) -> Buffer:
        """New buffer: the n points of `bases` (library form) followed by 2^64 times each, computed
        on the device. No counterpart in the reference."""
        h = ctypes.c_void_p()
        check(_lib.load().zkg_msm_precompute_bases(gpu.handle, group.value, bases.handle, int(n), ctypes.byref(h)))
        return Buffer(gpu, h.value)

    def base_points(self) -> int:
        """Points the bases buffer must hold: n, or 2n for a plan over precomputed bases."""
        return int(_lib.load().zkg_msm_base_points(self._h))

    def item_points(self) -> int:
        return int(_lib.load().zkg_msm_item_points(self._h))

    @staticmethod
    def upload_bases(gpu, plain_xy) -> Buffer:
        arr = np.ascontiguousarray(np.asarray(plain_xy, dtype=np.uint32).reshape(-1))
        h = ctypes.c_void_p()
        check(_lib.load().zkg_msm_upload_bases(gpu.handle, u32_ptr(arr), arr.size, ctypes.byref(h)))
        return Buffer(gpu, h.value)

    @staticmethod
    def upload_bases_montgomery_256(gpu, xy) -> Buffer:
        arr = np.ascontiguousarray(np.asarray(xy, dtype=np.uint32).reshape(-1))
        h = ctypes.c_void_p()
        check(_lib.load().zkg_msm_upload_bases_montgomery_256(gpu.handle, u32_ptr(arr), arr.size, ctypes.byref(h)))
        return Buffer(gpu, h.value)

    @staticmethod
    def convert_bases(buf: Buffer, words: int) -> None:
        """Convert plain coordinates already on the device to the library's form, in place."""
        check(_lib.load().zkg_msm_convert_bases(buf.handle, int(words)))

    @staticmethod
    def check_bases(gpu, group: Group, bases: Buffer, n: int):
        """(count, first index) of the first n bases that are neither infinity nor on the curve (G2:
        the twist and the order-r subgroup). No counterpart in the reference: plans assume G2 bases
        in the subgroup (the GLV endomorphism acts as lambda only there)."""
        bad, first = ctypes.c_uint64(), ctypes.c_uint64()
        check(_lib.load().zkg_msm_check_bases(gpu.handle, group.value, bases.handle, int(n), ctypes.byref(bad),
                                              ctypes.byref(first)))
        return int(bad.value), int(first.value)

    def group(self) -> Group:
        return self._group

    def window(self) -> int:
        return int(_lib.load().zkg_msm_window(self._h))

    def _result_for(self, gpu, handle: int) -> Buffer:
        # The returned buffer refers to plan-owned memory, so it holds the plan (which would free
        # that memory) alive; the plan keeps only a weak reference back, so dropping both frees them
        # at once (the reference's borrow of the plan, which Python cannot express).
        buf = self._result() if self._result is not None else None
        if buf is None or buf.handle.value != handle:
            buf = Buffer(gpu, handle, owned=False, keep=self)
            self._result = weakref.ref(buf)
        return buf

    def result_buffer(self, gpu) -> Buffer:
        """The plan-owned result buffer (written by `run` or by a batch that ran this plan)."""
        h = ctypes.c_void_p()
        check(_lib.load().zkg_msm_result(self._h, ctypes.byref(h)))
        return self._result_for(gpu, h.value)

    def run(self, gpu, bases: Buffer, scalars: Buffer) -> Buffer:
        """Queue sum(scalars[i] * bases[i]); returns the plan-owned XYZZ result buffer."""
        h = ctypes.c_void_p()
        check(_lib.load().zkg_msm_run(self._h, bases.handle, scalars.handle, ctypes.byref(h)))
        return self._result_for(gpu, h.value)

    async def run_yielding(self, gpu, bases: Buffer, scalars: Buffer) -> Buffer:
        """Msm::run_yielding (src/msm.rs:300-318): waits for the GPU after the heavy stages so a
        display sharing the GPU gets it in between; same result as run."""
        import asyncio

        h = ctypes.c_void_p()
        lib = _lib.load()
        await asyncio.to_thread(lambda: check(lib.zkg_msm_run_yielding(self._h, bases.handle, scalars.handle,
                                                                       ctypes.byref(h))))
        return self._result_for(gpu, h.value)

    def profile(self, gpu, bases: Buffer, scalars: Buffer):
        """Run once with a wait around every stage; [(stage, milliseconds)] (src/msm.rs:321-342)."""
        cap = 64
        names = ctypes.create_string_buffer(4096)
        ms = (ctypes.c_double * cap)()
        count = ctypes.c_size_t()
        check(_lib.load().zkg_msm_profile(self._h, bases.handle, scalars.handle, names, 4096, ms, cap,
                                          ctypes.byref(count)))
        labels = names.raw.split(b"\0")[: count.value]
        return [(labels[i].decode(), ms[i]) for i in range(count.value)]

    def __del__(self):
        try:
            if self._h:
                _lib.load().zkg_msm_destroy(self._h)
                self._h = ctypes.c_void_p(0)
        except Exception:
            pass


def _to_u256(words) -> host.U256:
    return host.from_words32(words)


def xyzz_to_affine(words) -> Optional[Tuple[host.U256, host.U256]]:
    """The GPU's XYZZ result (32 words) -> plain affine (x, y) as [u64; 4] each; None = infinity."""
    w = np.ascontiguousarray(np.asarray(words, dtype=np.uint32).reshape(-1)[:32])
    out = np.zeros(16, dtype=np.uint32)
    fin = ctypes.c_int()
    check(_lib.load().zkg_xyzz_to_affine(u32_ptr(w), u32_ptr(out), ctypes.byref(fin)))
    if not fin.value:
        return None
    return _to_u256(out[:8]), _to_u256(out[8:16])


def xyzz_to_affine_g2(words) -> Optional[Tuple[Fq2Words, Fq2Words]]:
    """The GPU's G2 XYZZ result (64 words) -> ((x.a, x.b), (y.a, y.b)); None = infinity."""
    w = np.ascontiguousarray(np.asarray(words, dtype=np.uint32).reshape(-1)[:64])
    out = np.zeros(32, dtype=np.uint32)
    fin = ctypes.c_int()
    check(_lib.load().zkg_xyzz_to_affine_g2(u32_ptr(w), u32_ptr(out), ctypes.byref(fin)))
    if not fin.value:
        return None
    return (_to_u256(out[0:8]), _to_u256(out[8:16])), (_to_u256(out[16:24]), _to_u256(out[24:32]))
