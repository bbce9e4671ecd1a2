// Device runtime: one HIP device with one in-order stream, device buffers, and the error channel
// of the C ABI.
//
// Mirrors the reference's device wrapper (src/gpu.rs:51-69 new/buffer/empty_buffer,
// :151-199 read/wait): every call is queued in FIFO order on the device's stream and only
// reading a buffer back or waiting synchronises with the host. Pipelines and bind groups have no
// counterpart: kernels are compiled ahead of time for gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace zkg {

// Raised inside the library and turned into a status code + message at the C ABI. `gpu` marks
// errors returned by the HIP runtime (they also go to the device's error callback).
struct Error : std::runtime_error {
    int code;
    bool gpu;
    Error(int c, const std::string& msg, bool from_gpu = false) : std::runtime_error(msg), code(c), gpu(from_gpu) {}
};

enum Status : int {
    ZKG_OK = 0,
    ZKG_ERR_INVALID = 1,   // argument out of range (same conditions as the reference's asserts)
    ZKG_ERR_SIZE = 2,      // buffer smaller than the plan needs
    ZKG_ERR_DEVICE = 3,    // HIP runtime error
    ZKG_ERR_NOMEM = 4,     // allocation failure / over the admission budget
    ZKG_ERR_INTERNAL = 5,
};

#define ZKG_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::zkg::Error(e_ == hipErrorOutOfMemory ? ::zkg::ZKG_ERR_NOMEM : ::zkg::ZKG_ERR_DEVICE, \
                               std::string(#call) + ": " + hipGetErrorString(e_), true);           \
    } while (0)

// Callback for GPU errors (Gpu::report_errors, src/gpu.rs:40-49): message and the caller's pointer.
using ErrorFn = void (*)(const char* message, void* user);

struct Device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    int clock_khz = 0;
    size_t total_mem = 0;
    std::string name;
    ErrorFn on_error = nullptr;  // guarded by the device registry's lock
    void* on_error_user = nullptr;
    // This device's copy of the no-allocation MSM's static scratch array (msm_scratch_address,
    // taken at device_open with the device current).
    void* msm_scratch = nullptr;
};

struct Buffer {
    Device* dev = nullptr;
    void* ptr = nullptr;
    size_t bytes = 0;
    bool owned = true;  // false for buffers wrapping caller-owned device memory
};

Device* device_open(int ordinal);
void device_close(Device* d);
void device_bind(Device* d);  // make d's ordinal the calling thread's current HIP device
// Install (fn != nullptr) or remove the device's GPU-error callback.
void device_set_error_callback(Device* d, ErrorFn fn, void* user);
// Pass "GPU error: <msg>" to the callback of `d`, or, with d == nullptr, of every open device
// on the calling thread's current HIP device. Called outside any library lock.
void report_gpu_error(Device* d, const std::string& msg);

Buffer* buffer_alloc(Device* d, size_t bytes);
void buffer_free(Buffer* b);

// Device scratch that grows on demand and is released with its owner.
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t need) {
        if (need <= bytes) return;
        if (ptr) ZKG_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        ZKG_HIP(hipMalloc(&ptr, need));
        bytes = need;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// p[0 .. count) = value, queued on s as a kernel (not a memset, so it may be graph-captured).
void fill_words(uint32_t* p, uint32_t value, size_t count, hipStream_t s);
// Debug mode ZKGPU_POISON=1: plans fill their workspace (MSM) or scratch (NTT) with 0xA5A5A5A5
// before every run, so that reading a word the run did not write gives a wrong result.
bool poison_enabled();

}  // namespace zkg
