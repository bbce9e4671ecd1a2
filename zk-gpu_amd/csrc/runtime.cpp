// Device runtime implementation (see runtime.hpp).
#include "runtime.hpp"

#include "msm.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace zkg {

namespace {
// Open devices, so that an error raised by a HIP call can reach the callback of the device the
// failing call was bound to.
std::mutex g_devices_mu;
std::vector<Device*> g_devices;
}  // namespace

void device_set_error_callback(Device* d, ErrorFn fn, void* user) {
    std::lock_guard<std::mutex> lk(g_devices_mu);
    d->on_error = fn;
    d->on_error_user = fn ? user : nullptr;
}

void report_gpu_error(Device* d, const std::string& msg) {
    std::vector<std::pair<ErrorFn, void*>> calls;
    {
        std::lock_guard<std::mutex> lk(g_devices_mu);
        int ordinal = -1;
        if (!d && hipGetDevice(&ordinal) != hipSuccess) return;
        for (Device* x : g_devices)
            if (x->on_error && (d ? x == d : x->ordinal == ordinal)) calls.emplace_back(x->on_error, x->on_error_user);
    }
    const std::string text = "GPU error: " + msg;
    for (auto& c : calls) c.first(text.c_str(), c.second);
}

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
    try {
        d->msm_scratch = msm_scratch_address();  // this device's copy of the static array
    } catch (...) {
        (void)hipStreamDestroy(d->stream);
        delete d;
        throw;
    }
    std::lock_guard<std::mutex> lk(g_devices_mu);
    g_devices.push_back(d);
    return d;
}

void device_bind(Device* d) { ZKG_HIP(hipSetDevice(d->ordinal)); }

void device_close(Device* d) {
    if (!d) return;
    {
        std::lock_guard<std::mutex> lk(g_devices_mu);
        g_devices.erase(std::remove(g_devices.begin(), g_devices.end(), d), g_devices.end());
    }
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
            (void)hipGetLastError();  // leave no stale error for the next launch check to pick up
            throw Error(ZKG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e), true);
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
