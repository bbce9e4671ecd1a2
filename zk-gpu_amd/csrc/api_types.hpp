// Opaque handle types of the C ABI and the exception-to-status guard shared by every
// translation unit that implements entry points of include/zkgpu.h.
#pragma once
#include <new>
#include <string>

#include "runtime.hpp"

struct zkg_device {
    zkg::Device* d;
};
struct zkg_buffer {
    zkg::Buffer* b;
};

namespace zkg {
// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);
// As fail(), for an error the HIP runtime returned: the message also goes to the error callback
// of `d` (nullptr: of the devices on the calling thread's current HIP device).
int fail_gpu(Device* d, int code, const std::string& msg);
}  // namespace zkg

#define ZKG_GUARD_BEGIN try {
#define ZKG_GUARD_END                                                   \
    }                                                                   \
    catch (const zkg::Error& e) {                                       \
        if (e.gpu) return zkg::fail_gpu(nullptr, e.code, e.what());     \
        return zkg::fail(e.code, e.what());                             \
    }                                                                   \
    catch (const std::bad_alloc&) {                                     \
        return zkg::fail(zkg::ZKG_ERR_NOMEM, "host allocation failed"); \
    }                                                                   \
    catch (const std::exception& e) {                                   \
        return zkg::fail(zkg::ZKG_ERR_INTERNAL, e.what());              \
    }

#define ZKG_REQUIRE(cond, msg) \
    if (!(cond)) return zkg::fail(zkg::ZKG_ERR_INVALID, msg)
