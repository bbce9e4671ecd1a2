"""Device and buffers — mirror of the reference's `gpu` module (src/gpu.rs).

`Gpu()` / `Gpu.new()` opens the device (src/gpu.rs:51-55), `buffer`/`empty_buffer` create
device buffers of u32 words (:57-69), `read` is the blocking readback (:192-199) and `wait`
drains the queue (:151-155). Everything goes through libzkgpu.so's C ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import ZkgError, check, u32_ptr


ZKG_ERR_DEVICE = 3


class Buffer:
    """A device buffer (the reference's wgpu::Buffer). Owned buffers are freed with the handle."""

    def __init__(self, gpu: "Gpu", handle: int, owned: bool = True, keep=None):
        self._gpu = gpu  # keeps the device open while the buffer lives
        self._h = ctypes.c_void_p(handle)
        self._owned = owned
        self._keep = keep  # the owner of memory this buffer does not own (e.g. an MSM plan)

    @property
    def handle(self):
        return self._h

    def size(self) -> int:
        """Capacity in bytes (wgpu::Buffer::size)."""
        return int(_lib.load().zkg_buffer_size(self._h))

    def device_ptr(self) -> int:
        return int(_lib.load().zkg_buffer_device_ptr(self._h) or 0)

    def free(self) -> None:
        if self._owned and self._h:
            _lib.load().zkg_buffer_free(self._h)
        self._h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Gpu:
    def __init__(self, ordinal: int = 0):
        lib = _lib.load()
        h = ctypes.c_void_p()
        check(lib.zkg_device_open(ordinal, ctypes.byref(h)))
        self._h = h

    @staticmethod
    def new() -> "Gpu":
        return Gpu()

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        cu, clk, mem = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        name = ctypes.create_string_buffer(64)
        check(_lib.load().zkg_device_info(self._h, ctypes.byref(cu), ctypes.byref(clk), ctypes.byref(mem), name, 64))
        return {"cu_count": cu.value, "clock_khz": clk.value, "total_mem": mem.value, "arch": name.value.decode()}

    def stream(self) -> int:
        """The hipStream_t (as an integer) on which the library queues this device's work."""
        return int(_lib.load().zkg_device_stream(self._h) or 0)

    def report_errors(self, log) -> None:
        """Gpu::report_errors (src/gpu.rs:40-49): ``log(message)`` receives "GPU error: ..." for every
        error the HIP runtime returns on this device, including those no call waits for (a
        server's scheduler thread). ``log=None`` removes the callback. ``log`` may run on another
        thread and must not call back into this device."""
        lib = _lib.load()
        if log is None:
            check(lib.zkg_device_report_errors(self._h, _lib.ERROR_FN(), None))
            self._on_error = None
            return

        def trampoline(message, _user):
            try:
                log(message.decode(errors="replace"))
            except Exception:  # an exception must not unwind into the library
                pass

        fn = _lib.ERROR_FN(trampoline)
        check(lib.zkg_device_report_errors(self._h, fn, None))
        self._on_error = fn  # the library holds a raw pointer: keep the thunk alive

    def buffer(self, data) -> Buffer:
        arr = np.ascontiguousarray(np.asarray(data, dtype=np.uint32).reshape(-1))
        h = ctypes.c_void_p()
        check(_lib.load().zkg_buffer_create(self._h, u32_ptr(arr), arr.size, ctypes.byref(h)))
        return Buffer(self, h.value)

    def empty_buffer(self, words: int) -> Buffer:
        h = ctypes.c_void_p()
        check(_lib.load().zkg_buffer_empty(self._h, int(words), ctypes.byref(h)))
        return Buffer(self, h.value)

    def wrap(self, device_ptr: int, nbytes: int) -> Buffer:
        """View caller-owned device memory (e.g. a torch tensor) as a Buffer, without copying."""
        h = ctypes.c_void_p()
        check(_lib.load().zkg_buffer_wrap(self._h, ctypes.c_void_p(device_ptr), int(nbytes), ctypes.byref(h)))
        return Buffer(self, h.value, owned=True)  # freeing the handle never frees wrapped memory

    def write(self, buf: Buffer, data, offset_words: int = 0) -> None:
        arr = np.ascontiguousarray(np.asarray(data, dtype=np.uint32).reshape(-1))
        check(_lib.load().zkg_buffer_write(buf.handle, int(offset_words), u32_ptr(arr), arr.size))

    def read(self, buf: Buffer, words: int) -> np.ndarray:
        out = np.zeros(int(words), dtype=np.uint32)
        check(_lib.load().zkg_buffer_read(buf.handle, int(words), u32_ptr(out)))
        return out

    def wait(self) -> None:
        check(_lib.load().zkg_wait(self._h))

    # ---- asynchronous forms (src/gpu.rs:22-40, :157-190): awaitable without blocking an event loop

    @staticmethod
    async def new_async() -> "Gpu":
        """Gpu::new_async: raises ZkgError("GPU init failed ...") instead of returning Err."""
        import asyncio

        return await asyncio.to_thread(Gpu)

    async def finished_async(self) -> None:
        """Gpu::finished_async: resolves when all work queued so far has finished."""
        import asyncio

        await asyncio.to_thread(self.wait)

    async def try_read_async(self, buf: Buffer, words: int):
        """Gpu::try_read_async: (True, words) or (False, "GPU read failed: ...")."""
        import asyncio

        out = np.zeros(int(words), dtype=np.uint32)
        ticket = ctypes.c_void_p()
        lib = _lib.load()
        if lib.zkg_buffer_read_async(buf.handle, int(words), u32_ptr(out), ctypes.byref(ticket)) != 0:
            return False, lib.zkg_last_error().decode(errors="replace")
        done = ctypes.c_int()
        while True:
            if lib.zkg_read_done(ticket, ctypes.byref(done)) != 0 or done.value:
                break
            await asyncio.sleep(0)
        if lib.zkg_read_wait(ticket) != 0:
            return False, lib.zkg_last_error().decode(errors="replace")
        return True, out

    async def read_async(self, buf: Buffer, words: int) -> np.ndarray:
        """Gpu::read_async: the words, or ZkgError on failure."""
        ok, val = await self.try_read_async(buf, words)
        if not ok:
            raise ZkgError(ZKG_ERR_DEVICE, val)
        return val

    def close(self) -> None:
        if self._h:
            _lib.load().zkg_device_close(self._h)
            self._on_error = None
            self._h = ctypes.c_void_p(0)
