// Device runtime implementation (see runtime.hpp).
#include "runtime.hpp"

namespace zkg {

Device* device_open(int ordinal) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0 || ordinal < 0 || ordinal >= count)
        throw Error(ZKG_ERR_DEVICE, "GPU init failed");
    ZKG_HIP(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    ZKG_HIP(hipGetDeviceProperties(&prop, ordinal));
    Device* d = new Device();
    d->ordinal = ordinal;
    d->cu_count = prop.multiProcessorCount;
    d->clock_khz = prop.clockRate;
    d->total_mem = prop.totalGlobalMem;
    d->name = prop.gcnArchName;
    // A non-blocking stream at the default priority keeps FIFO order for everything queued here.
    hipError_t se = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (se != hipSuccess) {
        delete d;
        throw Error(ZKG_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(se));
    }
    return d;
}

void device_bind(Device* d) { ZKG_HIP(hipSetDevice(d->ordinal)); }

void device_close(Device* d) {
    if (!d) return;
    (void)hipSetDevice(d->ordinal);
    (void)hipStreamSynchronize(d->stream);
    (void)hipStreamDestroy(d->stream);
    delete d;
}

Buffer* buffer_alloc(Device* d, size_t bytes) {
    device_bind(d);
    Buffer* b = new Buffer();
    b->dev = d;
    b->bytes = bytes;
    if (bytes) {
        hipError_t e = hipMalloc(&b->ptr, bytes);
        if (e != hipSuccess) {
            delete b;
            throw Error(ZKG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
    }
    return b;
}

void buffer_free(Buffer* b) {
    if (!b) return;
    if (b->owned && b->ptr) {
        (void)hipSetDevice(b->dev->ordinal);
        // Work queued on the stream may still read this memory.
        (void)hipStreamSynchronize(b->dev->stream);
        (void)hipFree(b->ptr);
    }
    delete b;
}

}  // namespace zkg
