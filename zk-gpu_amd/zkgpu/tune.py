"""Autotuning of MSM window / work-item / sort-block sizes and the NTT tile (no counterpart in
the reference; include/zkgpu.h "autotuning"). Tuned choices change only how the work is split."""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import check


def tune_msm(gpu, group: int, n: int) -> dict:
    c, ip, sp = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    ms = ctypes.c_double()
    check(_lib.load().zkg_tune_msm(gpu.handle, int(group), int(n), ctypes.byref(c), ctypes.byref(ip), ctypes.byref(sp),
                                   ctypes.byref(ms)))
    return {"c": c.value, "item_points": ip.value, "sort_points": sp.value, "ms": ms.value}


def tune_ntt(gpu, log_n: int) -> dict:
    tile, ms = ctypes.c_uint32(), ctypes.c_double()
    check(_lib.load().zkg_tune_ntt(gpu.handle, int(log_n), ctypes.byref(tile), ctypes.byref(ms)))
    return {"tile_log": tile.value, "ms": ms.value}


def lookup_msm(arch: str, group: int, n: int):
    found = ctypes.c_int()
    c, ip, sp = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    check(_lib.load().zkg_tune_lookup_msm(arch.encode(), int(group), int(n), ctypes.byref(found), ctypes.byref(c),
                                          ctypes.byref(ip), ctypes.byref(sp)))
    return {"c": c.value, "item_points": ip.value, "sort_points": sp.value} if found.value else None


def lookup_ntt(arch: str, log_n: int):
    found, tile = ctypes.c_int(), ctypes.c_uint32()
    check(_lib.load().zkg_tune_lookup_ntt(arch.encode(), int(log_n), ctypes.byref(found), ctypes.byref(tile)))
    return {"tile_log": tile.value} if found.value else None


def load(path: str) -> None:
    check(_lib.load().zkg_tune_load(str(path).encode()))


def save(path: str) -> None:
    check(_lib.load().zkg_tune_save(str(path).encode()))


def clear() -> None:
    check(_lib.load().zkg_tune_clear())
