// C ABI (include/zkgpu.h): device, buffers, NTT, and host-side affine conversion.
// Every entry point catches library exceptions and turns them into a status code plus a
// thread-local message, so no C++ exception crosses the boundary.
#include "../../include/zkgpu.h"

#include <string.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "api_types.hpp"
#include "host_math.hpp"
#include "ntt.hpp"
#include "runtime.hpp"

struct zkg_ntt {
    zkg::NttPlan* plan;
};

namespace zkg {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int fail_gpu(Device* d, int code, const std::string& msg) {
    report_gpu_error(d, msg);
    return fail(code, msg);
}
}  // namespace zkg

extern "C" {

const char* zkg_last_error(void) { return zkg::g_last_error.c_str(); }

int zkg_version(void) { return 100; }

int zkg_device_open(int ordinal, zkg_device** out) {
    ZKG_REQUIRE(out, "null output pointer");
    ZKG_GUARD_BEGIN
    zkg::Device* d = zkg::device_open(ordinal);
    *out = new zkg_device{d};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_device_close(zkg_device* dev) {
    if (!dev) return zkg::ZKG_OK;
    zkg::device_close(dev->d);
    delete dev;
    return zkg::ZKG_OK;
}

int zkg_device_info(zkg_device* dev, int* cu_count, int* clock_khz, size_t* total_mem, char* name,
                    size_t name_len) {
    ZKG_REQUIRE(dev, "null device");
    if (cu_count) *cu_count = dev->d->cu_count;
    if (clock_khz) *clock_khz = dev->d->clock_khz;
    if (total_mem) *total_mem = dev->d->total_mem;
    if (name && name_len) {
        strncpy(name, dev->d->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    return zkg::ZKG_OK;
}

void* zkg_device_stream(zkg_device* dev) { return dev ? (void*)dev->d->stream : nullptr; }

int zkg_device_report_errors(zkg_device* dev, zkg_error_fn fn, void* user) {
    ZKG_REQUIRE(dev, "null device");
    zkg::device_set_error_callback(dev->d, fn, user);
    return zkg::ZKG_OK;
}

int zkg_buffer_create(zkg_device* dev, const uint32_t* words, size_t n_words, zkg_buffer** out) {
    ZKG_REQUIRE(dev && out && (words || n_words == 0), "null argument");
    ZKG_GUARD_BEGIN
    zkg::Buffer* b = zkg::buffer_alloc(dev->d, 4 * n_words);
    if (n_words) {
        hipError_t e = hipMemcpyAsync(b->ptr, words, 4 * n_words, hipMemcpyHostToDevice, dev->d->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(dev->d->stream);  // `words` may be freed on return
        if (e != hipSuccess) {
            zkg::buffer_free(b);
            return zkg::fail_gpu(dev->d, zkg::ZKG_ERR_DEVICE, std::string("upload: ") + hipGetErrorString(e));
        }
    }
    *out = new zkg_buffer{b};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_buffer_empty(zkg_device* dev, size_t n_words, zkg_buffer** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    ZKG_GUARD_BEGIN
    zkg::Buffer* b = zkg::buffer_alloc(dev->d, 4 * n_words);
    if (n_words) {
        hipError_t e = hipMemsetAsync(b->ptr, 0, 4 * n_words, dev->d->stream);
        if (e != hipSuccess) {
            zkg::buffer_free(b);
            return zkg::fail_gpu(dev->d, zkg::ZKG_ERR_DEVICE, std::string("memset: ") + hipGetErrorString(e));
        }
    }
    *out = new zkg_buffer{b};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_buffer_wrap(zkg_device* dev, void* device_ptr, size_t bytes, zkg_buffer** out) {
    ZKG_REQUIRE(dev && out && (device_ptr || bytes == 0), "null argument");
    ZKG_GUARD_BEGIN
    zkg::Buffer* b = new zkg::Buffer();
    b->dev = dev->d;
    b->ptr = device_ptr;
    b->bytes = bytes;
    b->owned = false;
    *out = new zkg_buffer{b};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_buffer_read(zkg_buffer* buf, size_t n_words, uint32_t* out) {
    ZKG_REQUIRE(buf && (out || n_words == 0), "null argument");
    if (4 * n_words > buf->b->bytes)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "read of " + std::to_string(4 * n_words) + " bytes from a buffer of " +
                                                 std::to_string(buf->b->bytes));
    ZKG_GUARD_BEGIN
    zkg::device_bind(buf->b->dev);
    ZKG_HIP(hipMemcpyAsync(out, buf->b->ptr, 4 * n_words, hipMemcpyDeviceToHost, buf->b->dev->stream));
    hipError_t e = hipStreamSynchronize(buf->b->dev->stream);
    if (e != hipSuccess)
        return zkg::fail_gpu(buf->b->dev, zkg::ZKG_ERR_DEVICE, std::string("GPU read failed: ") + hipGetErrorString(e));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

// An asynchronous read: the copy lands in pinned staging; wait() moves it to the caller's array.
struct zkg_read {
    zkg::Device* dev;
    uint32_t* staging;
    uint32_t* out;
    size_t words;
    hipEvent_t done;
};

int zkg_buffer_read_async(zkg_buffer* buf, size_t n_words, uint32_t* out, zkg_read** ticket) {
    ZKG_REQUIRE(buf && ticket && (out || n_words == 0), "null argument");
    if (4 * n_words > buf->b->bytes)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "read of " + std::to_string(4 * n_words) + " bytes from a buffer of " +
                                                 std::to_string(buf->b->bytes));
    ZKG_GUARD_BEGIN
    zkg::device_bind(buf->b->dev);
    std::unique_ptr<zkg_read> t(new zkg_read{buf->b->dev, nullptr, out, n_words, nullptr});
    ZKG_HIP(hipHostMalloc(reinterpret_cast<void**>(&t->staging), 4 * std::max<size_t>(n_words, 1)));
    hipError_t e = hipEventCreateWithFlags(&t->done, hipEventDisableTiming);
    if (e == hipSuccess && n_words)
        e = hipMemcpyAsync(t->staging, buf->b->ptr, 4 * n_words, hipMemcpyDeviceToHost, buf->b->dev->stream);
    if (e == hipSuccess) e = hipEventRecord(t->done, buf->b->dev->stream);
    if (e != hipSuccess) {
        (void)hipHostFree(t->staging);
        if (t->done) (void)hipEventDestroy(t->done);
        return zkg::fail_gpu(buf->b->dev, zkg::ZKG_ERR_DEVICE, std::string("GPU read failed: ") + hipGetErrorString(e));
    }
    *ticket = t.release();
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_read_done(zkg_read* t, int* done) {
    ZKG_REQUIRE(t && done, "null argument");
    const hipError_t e = hipEventQuery(t->done);
    if (e == hipErrorNotReady) {
        *done = 0;
        return zkg::ZKG_OK;
    }
    *done = 1;
    if (e != hipSuccess)
        return zkg::fail_gpu(t->dev, zkg::ZKG_ERR_DEVICE, std::string("GPU read failed: ") + hipGetErrorString(e));
    return zkg::ZKG_OK;
}

int zkg_read_wait(zkg_read* t) {
    ZKG_REQUIRE(t, "null argument");
    const hipError_t e = hipEventSynchronize(t->done);
    if (e == hipSuccess && t->words) memcpy(t->out, t->staging, 4 * t->words);
    zkg::Device* d = t->dev;
    (void)hipHostFree(t->staging);
    (void)hipEventDestroy(t->done);
    delete t;
    if (e != hipSuccess) return zkg::fail_gpu(d, zkg::ZKG_ERR_DEVICE, std::string("GPU read failed: ") + hipGetErrorString(e));
    return zkg::ZKG_OK;
}

int zkg_buffer_write(zkg_buffer* buf, size_t offset_words, const uint32_t* words, size_t n_words) {
    ZKG_REQUIRE(buf && (words || n_words == 0), "null argument");
    if (4 * (offset_words + n_words) > buf->b->bytes) return zkg::fail(zkg::ZKG_ERR_SIZE, "write past the end of a buffer");
    ZKG_GUARD_BEGIN
    zkg::device_bind(buf->b->dev);
    ZKG_HIP(hipMemcpyAsync(static_cast<uint32_t*>(buf->b->ptr) + offset_words, words, 4 * n_words,
                           hipMemcpyHostToDevice, buf->b->dev->stream));
    ZKG_HIP(hipStreamSynchronize(buf->b->dev->stream));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

size_t zkg_buffer_size(const zkg_buffer* buf) { return buf ? buf->b->bytes : 0; }

void* zkg_buffer_device_ptr(const zkg_buffer* buf) { return buf ? buf->b->ptr : nullptr; }

int zkg_buffer_free(zkg_buffer* buf) {
    if (!buf) return zkg::ZKG_OK;
    zkg::buffer_free(buf->b);
    delete buf;
    return zkg::ZKG_OK;
}

int zkg_wait(zkg_device* dev) {
    ZKG_REQUIRE(dev, "null device");
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    ZKG_HIP(hipStreamSynchronize(dev->d->stream));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

// ------------------------------------------------------------------ NTT

int zkg_ntt_create(zkg_device* dev, uint32_t log_n, zkg_ntt** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    ZKG_GUARD_BEGIN
    zkg::NttPlan* p = new zkg::NttPlan(dev->d, log_n);
    *out = new zkg_ntt{p};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_ntt_destroy(zkg_ntt* plan) {
    if (!plan) return zkg::ZKG_OK;
    (void)hipSetDevice(plan->plan->dev->ordinal);
    (void)hipStreamSynchronize(plan->plan->dev->stream);
    delete plan->plan;
    delete plan;
    return zkg::ZKG_OK;
}

uint32_t zkg_ntt_log_n(const zkg_ntt* plan) { return plan ? plan->plan->log_n : 0; }
uint32_t zkg_ntt_passes(const zkg_ntt* plan) { return plan ? (uint32_t)plan->plan->passes.size() : 0; }

static int ntt_run(zkg_ntt* plan, zkg_buffer* data, bool inverse) {
    ZKG_REQUIRE(plan && data, "null argument");
    const size_t need = (size_t)32 << plan->plan->log_n;
    if (data->b->bytes < need)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "data buffer holds " + std::to_string(data->b->bytes) +
                                                " bytes; a 2^" + std::to_string(plan->plan->log_n) +
                                                " transform needs " + std::to_string(need));
    ZKG_GUARD_BEGIN
    zkg::device_bind(plan->plan->dev);
    plan->plan->run(static_cast<uint32_t*>(data->b->ptr), inverse, plan->plan->dev->stream);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_ntt_forward(zkg_ntt* plan, zkg_buffer* data) { return ntt_run(plan, data, false); }
int zkg_ntt_inverse(zkg_ntt* plan, zkg_buffer* data) { return ntt_run(plan, data, true); }

// ------------------------------------------------------------------ affine conversion (host)
// The library's Montgomery radix is 2^256: x = X * 2^-256 mod p. Then (x/zz, y/zzz).

int zkg_xyzz_to_affine(const uint32_t* w, uint32_t* out, int* is_finite) {
    ZKG_REQUIRE(w && out && is_finite, "null argument");
    const zkg::host::Modulus& F = zkg::host::fq();
    zkg::host::U256 c[4];
    for (int i = 0; i < 4; ++i) c[i] = F.from_mont(zkg::host::from_words32(w + 8 * i));
    if (zkg::host::is_zero(c[2])) {
        memset(out, 0, 16 * sizeof(uint32_t));
        *is_finite = 0;
        return zkg::ZKG_OK;
    }
    zkg::host::to_words32(out, F.mul(c[0], F.inverse(c[2])));
    zkg::host::to_words32(out + 8, F.mul(c[1], F.inverse(c[3])));
    *is_finite = 1;
    return zkg::ZKG_OK;
}

// Fq2 = Fq[u]/(u^2+1); 1/(a+bu) = (a-bu)/(a^2+b^2)   (src/msm.rs:376-387)
static void fq2_mul(const zkg::host::Modulus& F, const zkg::host::U256 x[2], const zkg::host::U256 y[2],
                    zkg::host::U256 o[2]) {
    zkg::host::U256 ac = F.mul(x[0], y[0]), bd = F.mul(x[1], y[1]);
    zkg::host::U256 ad = F.mul(x[0], y[1]), bc = F.mul(x[1], y[0]);
    o[0] = F.sub(ac, bd);
    o[1] = F.add(ad, bc);
}

static void fq2_inv(const zkg::host::Modulus& F, const zkg::host::U256 x[2], zkg::host::U256 o[2]) {
    zkg::host::U256 nrm = F.add(F.mul(x[0], x[0]), F.mul(x[1], x[1]));
    zkg::host::U256 ni = F.inverse(nrm);
    o[0] = F.mul(x[0], ni);
    o[1] = F.sub(zkg::host::U256{{0, 0, 0, 0}}, F.mul(x[1], ni));
}

int zkg_xyzz_to_affine_g2(const uint32_t* w, uint32_t* out, int* is_finite) {
    ZKG_REQUIRE(w && out && is_finite, "null argument");
    const zkg::host::Modulus& F = zkg::host::fq();
    zkg::host::U256 c[4][2];
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 2; ++k) c[i][k] = F.from_mont(zkg::host::from_words32(w + 8 * (2 * i + k)));
    if (zkg::host::is_zero(c[2][0]) && zkg::host::is_zero(c[2][1])) {
        memset(out, 0, 32 * sizeof(uint32_t));
        *is_finite = 0;
        return zkg::ZKG_OK;
    }
    zkg::host::U256 zi[2], zzi[2], x[2], y[2];
    fq2_inv(F, c[2], zi);
    fq2_inv(F, c[3], zzi);
    fq2_mul(F, c[0], zi, x);
    fq2_mul(F, c[1], zzi, y);
    zkg::host::to_words32(out, x[0]);
    zkg::host::to_words32(out + 8, x[1]);
    zkg::host::to_words32(out + 16, y[0]);
    zkg::host::to_words32(out + 24, y[1]);
    *is_finite = 1;
    return zkg::ZKG_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ MSM

#include "msm.hpp"
#include "tune.hpp"

struct zkg_msm {
    zkg::MsmPlan* plan;
    zkg_buffer* result;  // plan-owned view of plan->result
};

extern "C" {

uint32_t zkg_msm_default_window(size_t n) { return zkg::msm_default_window(n); }

static zkg_msm* wrap_plan(zkg::Device* d, std::unique_ptr<zkg::MsmPlan> p) {
    std::unique_ptr<zkg::Buffer> rb(new zkg::Buffer());
    rb->dev = d;
    rb->ptr = p->result;
    rb->bytes = 4 * (size_t)p->xw;
    rb->owned = false;
    std::unique_ptr<zkg_buffer> rh(new zkg_buffer{rb.get()});
    zkg_msm* m = new zkg_msm{p.get(), rh.get()};
    p.release();
    rb.release();
    rh.release();
    return m;
}

int zkg_msm_create(zkg_device* dev, int group, size_t n, uint32_t c, uint32_t item_points, zkg_msm** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    ZKG_GUARD_BEGIN
    *out = wrap_plan(dev->d, std::unique_ptr<zkg::MsmPlan>(new zkg::MsmPlan(dev->d, group, n, c, item_points)));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_plan_check(int group, size_t n, uint32_t c, uint32_t item_points, int precomputed,
                       size_t* workspace_bytes) {
    ZKG_GUARD_BEGIN
    const zkg::MsmShape sh = zkg::msm_shape(group, n, c, item_points, 0, precomputed != 0);
    if (workspace_bytes) *workspace_bytes = sh.workspace_bytes(n);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_create_tuned(zkg_device* dev, int group, size_t n, zkg_msm** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    ZKG_GUARD_BEGIN
    if (n < 1) return zkg::fail(zkg::ZKG_ERR_INVALID, "an MSM needs at least one point");
    const zkg::MsmTune t = zkg::msm_config(dev->d, group, n);
    *out = wrap_plan(dev->d, std::unique_ptr<zkg::MsmPlan>(
                                 new zkg::MsmPlan(dev->d, group, n, t.c, t.item_points, t.sort_points)));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_create_precomputed(zkg_device* dev, int group, size_t n, uint32_t c, uint32_t item_points,
                               zkg_msm** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    ZKG_GUARD_BEGIN
    *out = wrap_plan(dev->d, std::unique_ptr<zkg::MsmPlan>(new zkg::MsmPlan(dev->d, group, n, c, item_points, 0, true)));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_precompute_bases(zkg_device* dev, int group, const zkg_buffer* bases, size_t n, zkg_buffer** out) {
    ZKG_REQUIRE(dev && bases && out, "null argument");
    ZKG_REQUIRE(group == ZKG_G1 || group == ZKG_G2, "group must be G1 (1) or G2 (2)");
    ZKG_REQUIRE(n >= 1, "an MSM needs at least one point");
    const uint64_t point_bytes = 64ull * group;
    if (bases->b->bytes < n * point_bytes)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "bases buffer holds " + std::to_string(bases->b->bytes) + " bytes; " +
                                                std::to_string(n) + " points need " + std::to_string(n * point_bytes));
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    zkg::Buffer* b = zkg::buffer_alloc(dev->d, 2 * n * point_bytes);  // every word is written below
    try {
        zkg::msm_precompute_bases(group, static_cast<const uint32_t*>(bases->b->ptr), n,
                                  static_cast<uint32_t*>(b->ptr), dev->d->stream);
    } catch (...) {
        zkg::buffer_free(b);
        throw;
    }
    *out = new zkg_buffer{b};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

size_t zkg_msm_base_points(const zkg_msm* m) { return m ? m->plan->base_points() : 0; }

uint32_t zkg_msm_item_points(const zkg_msm* m) { return m ? m->plan->CH : 0; }

int zkg_msm_destroy(zkg_msm* m) {
    if (!m) return zkg::ZKG_OK;
    (void)hipSetDevice(m->plan->dev->ordinal);
    (void)hipStreamSynchronize(m->plan->dev->stream);
    zkg::buffer_free(m->result->b);
    delete m->result;
    delete m->plan;
    delete m;
    return zkg::ZKG_OK;
}

int zkg_msm_group(const zkg_msm* m) { return m ? m->plan->group : 0; }
uint32_t zkg_msm_window(const zkg_msm* m) { return m ? m->plan->c : 0; }
size_t zkg_msm_points(const zkg_msm* m) { return m ? m->plan->n : 0; }

// convert: plain coordinates -> Montgomery form; else Montgomery-2^256 words, reduced below p
static int upload(zkg_device* dev, const uint32_t* xy, size_t n_words, zkg_buffer** out, bool convert) {
    ZKG_REQUIRE(dev && out && (xy || n_words == 0), "null argument");
    ZKG_REQUIRE(n_words % 8 == 0, "coordinates are 8 words each");
    int rc = zkg_buffer_create(dev, xy, n_words, out);
    if (rc != zkg::ZKG_OK) return rc;
    if (n_words == 0) return zkg::ZKG_OK;
    ZKG_GUARD_BEGIN
    uint32_t* p = static_cast<uint32_t*>((*out)->b->ptr);
    if (convert) zkg::msm_convert_bases(p, n_words / 8, dev->d->stream);
    else zkg::msm_reduce_coords(p, n_words / 8, dev->d->stream);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_upload_bases(zkg_device* dev, const uint32_t* plain_xy, size_t n_words, zkg_buffer** out) {
    return upload(dev, plain_xy, n_words, out, true);
}

int zkg_msm_upload_bases_montgomery_256(zkg_device* dev, const uint32_t* xy, size_t n_words, zkg_buffer** out) {
    return upload(dev, xy, n_words, out, false);
}

int zkg_msm_convert_bases(zkg_buffer* buf, size_t n_words) {
    ZKG_REQUIRE(buf, "null argument");
    ZKG_REQUIRE(n_words % 8 == 0, "coordinates are 8 words each");
    if (4 * n_words > buf->b->bytes) return zkg::fail(zkg::ZKG_ERR_SIZE, "conversion past the end of a buffer");
    ZKG_GUARD_BEGIN
    zkg::device_bind(buf->b->dev);
    zkg::msm_convert_bases(static_cast<uint32_t*>(buf->b->ptr), n_words / 8, buf->b->dev->stream);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_check_bases(zkg_device* dev, int group, const zkg_buffer* bases, size_t n, uint64_t* bad,
                        uint64_t* first_bad) {
    ZKG_REQUIRE(dev && bases && bad && first_bad, "null argument");
    ZKG_REQUIRE(group == ZKG_G1 || group == ZKG_G2, "group must be G1 (1) or G2 (2)");
    const uint64_t point_bytes = group == ZKG_G1 ? 64 : 128;
    if (bases->b->bytes < n * point_bytes)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "bases buffer holds " + std::to_string(bases->b->bytes) + " bytes; " +
                                                std::to_string(n) + " points need " + std::to_string(n * point_bytes));
    ZKG_GUARD_BEGIN
    zkg::msm_check_bases(dev->d, group, static_cast<const uint32_t*>(bases->b->ptr), n, bad, first_bad);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

// Msm::check_buffers (src/msm.rs:271-282)
static int msm_check(const zkg_msm* m, const zkg_buffer* bases, const zkg_buffer* scalars) {
    ZKG_REQUIRE(m && bases && scalars, "null argument");
    const uint64_t n = m->plan->n, point_bytes = 4ull * m->plan->aw, pts = m->plan->base_points();
    if (scalars->b->bytes < n * 32)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "scalars buffer holds " + std::to_string(scalars->b->bytes) + " bytes; " +
                                                std::to_string(n) + " scalars need " + std::to_string(n * 32));
    if (bases->b->bytes < pts * point_bytes)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "bases buffer holds " + std::to_string(bases->b->bytes) + " bytes; " +
                                                std::to_string(pts) + " points need " + std::to_string(pts * point_bytes));
    return zkg::ZKG_OK;
}

int zkg_msm_run(zkg_msm* m, const zkg_buffer* bases, const zkg_buffer* scalars, zkg_buffer** result) {
    int rc = msm_check(m, bases, scalars);
    if (rc != zkg::ZKG_OK) return rc;
    ZKG_GUARD_BEGIN
    zkg::device_bind(m->plan->dev);
    m->plan->run(static_cast<const uint32_t*>(bases->b->ptr), static_cast<const uint32_t*>(scalars->b->ptr),
                 m->plan->dev->stream);
    if (result) *result = m->result;
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_run_noalloc(zkg_device* dev, int group, size_t n, uint32_t c, const zkg_buffer* bases,
                        const zkg_buffer* scalars, zkg_buffer* result) {
    ZKG_REQUIRE(dev && bases && scalars && result, "null argument");
    ZKG_GUARD_BEGIN
    const zkg::MsmShape sh = zkg::msm_shape(group, n, c, 0);
    const size_t need = zkg::msm_carved_bytes(sh, n);
    if (need > zkg::MSM_SCRATCH_BYTES)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "an MSM of " + std::to_string(n) + " points at c = " + std::to_string(c) +
                                                " needs a workspace of " + std::to_string(need) +
                                                " bytes; the static scratch holds " +
                                                std::to_string(zkg::MSM_SCRATCH_BYTES));
    const uint64_t point_bytes = 4ull * sh.aw;
    if (scalars->b->bytes < n * 32)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "scalars buffer holds " + std::to_string(scalars->b->bytes) + " bytes; " +
                                                std::to_string(n) + " scalars need " + std::to_string(n * 32));
    if (bases->b->bytes < n * point_bytes)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "bases buffer holds " + std::to_string(bases->b->bytes) + " bytes; " +
                                                std::to_string(n) + " points need " + std::to_string(n * point_bytes));
    if (result->b->bytes < 4ull * sh.xw)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "result buffer holds " + std::to_string(result->b->bytes) +
                                                " bytes; the point needs " + std::to_string(4ull * sh.xw));
    if (!dev->d->msm_scratch) return zkg::fail(zkg::ZKG_ERR_DEVICE, "no static scratch on this device");
    zkg::device_bind(dev->d);
    const zkg::MsmPlan plan(dev->d, group, n, c, sh, dev->d->msm_scratch, static_cast<uint32_t*>(result->b->ptr));
    plan.run(static_cast<const uint32_t*>(bases->b->ptr), static_cast<const uint32_t*>(scalars->b->ptr),
             dev->d->stream);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_run_yielding(zkg_msm* m, const zkg_buffer* bases, const zkg_buffer* scalars, zkg_buffer** result) {
    int rc = msm_check(m, bases, scalars);
    if (rc != zkg::ZKG_OK) return rc;
    ZKG_GUARD_BEGIN
    zkg::device_bind(m->plan->dev);
    m->plan->run_yielding(static_cast<const uint32_t*>(bases->b->ptr), static_cast<const uint32_t*>(scalars->b->ptr),
                          m->plan->dev->stream);
    if (result) *result = m->result;
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_msm_result(zkg_msm* m, zkg_buffer** result) {
    ZKG_REQUIRE(m && result, "null argument");
    *result = m->result;
    return zkg::ZKG_OK;
}

int zkg_msm_profile(zkg_msm* m, const zkg_buffer* bases, const zkg_buffer* scalars, char* names, size_t names_len,
                    double* ms, size_t cap, size_t* count) {
    int rc = msm_check(m, bases, scalars);
    if (rc != zkg::ZKG_OK) return rc;
    ZKG_REQUIRE(names && ms && count, "null argument");
    ZKG_GUARD_BEGIN
    zkg::device_bind(m->plan->dev);
    std::vector<std::pair<std::string, double>> st;
    m->plan->run(static_cast<const uint32_t*>(bases->b->ptr), static_cast<const uint32_t*>(scalars->b->ptr),
                 m->plan->dev->stream, &st);
    size_t k = 0, off = 0;
    for (auto& e : st) {
        if (k >= cap || off + e.first.size() + 1 > names_len) break;
        memcpy(names + off, e.first.c_str(), e.first.size() + 1);
        off += e.first.size() + 1;
        ms[k++] = e.second;
    }
    *count = k;
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

}  // extern "C"

// ------------------------------------------------------------------ batches

#include "batch.hpp"

struct zkg_batch {
    zkg::Batch* b;
};

extern "C" {

int zkg_batch_create(zkg_device* dev, zkg_batch** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    ZKG_GUARD_BEGIN
    *out = new zkg_batch{new zkg::Batch(dev->d)};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_batch_destroy(zkg_batch* batch) {
    if (!batch) return zkg::ZKG_OK;
    delete batch->b;
    delete batch;
    return zkg::ZKG_OK;
}

int zkg_batch_add_ntt(zkg_batch* batch, zkg_buffer* data, uint32_t log_n, int inverse) {
    ZKG_REQUIRE(batch && data, "null argument");
    if (log_n < 10 || log_n > 27) return zkg::fail(zkg::ZKG_ERR_INVALID, "log_n must be in 10..=27");
    const size_t need = (size_t)32 << log_n;
    if (data->b->bytes < need)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "data buffer holds " + std::to_string(data->b->bytes) + " bytes; a 2^" +
                                                std::to_string(log_n) + " transform needs " + std::to_string(need));
    ZKG_GUARD_BEGIN
    batch->b->add_ntt(log_n, static_cast<uint32_t*>(data->b->ptr), inverse != 0);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_batch_add_msm(zkg_batch* batch, zkg_msm* plan, const zkg_buffer* bases, const zkg_buffer* scalars) {
    ZKG_REQUIRE(batch, "null argument");
    int rc = msm_check(plan, bases, scalars);
    if (rc != zkg::ZKG_OK) return rc;
    ZKG_GUARD_BEGIN
    batch->b->add_msm(plan->plan, static_cast<const uint32_t*>(bases->b->ptr),
                      static_cast<const uint32_t*>(scalars->b->ptr));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_batch_run(zkg_batch* batch) {
    ZKG_REQUIRE(batch, "null argument");
    ZKG_GUARD_BEGIN
    batch->b->run();
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_batch_clear(zkg_batch* batch) {
    ZKG_REQUIRE(batch, "null argument");
    batch->b->clear();
    return zkg::ZKG_OK;
}

int zkg_batch_capture(zkg_batch* batch) {
    ZKG_REQUIRE(batch, "null argument");
    ZKG_GUARD_BEGIN
    batch->b->capture();
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_batch_captured(const zkg_batch* batch) { return batch && batch->b->captured() ? 1 : 0; }

}  // extern "C"

// ------------------------------------------------------------------ proof-job server

#include "serve.hpp"
#include "zkey.hpp"

// A section's uploaded points must be infinity or on the curve (and, for B2, in the order-r
// subgroup): a file whose coordinates are swapped, not in Montgomery form, or cut from another
// curve fails here instead of giving wrong MSMs.
static void zkey_check_section(zkg::Device* d, int section, const uint32_t* words, size_t points) {
    const int group = (int)(zkg::Zkey::point_words(section) / 16);
    uint64_t bad = 0, first = 0;
    zkg::msm_check_bases(d, group, words, points, &bad, &first);
    if (bad)
        throw zkg::Error(zkg::ZKG_ERR_INVALID,
                         "the .zkey section " + std::to_string(section) + " holds " + std::to_string(bad) + " point" +
                             (bad == 1 ? "" : "s") + (group == 1 ? " off the curve" : " off the twist or outside the order-r subgroup") +
                             " (first at index " + std::to_string(first) + ")");
}


struct zkg_server {
    zkg::Server* s;
    zkg::Device* d;
};

struct zkg_zkey {
    zkg::Zkey* z;
};

namespace {

std::string bases_short(const zkg::Buffer* b, size_t pts, size_t words_per_point) {
    return "bases buffer holds " + std::to_string(b->bytes) + " bytes; " + std::to_string(pts) + " points need " +
           std::to_string(4 * words_per_point * pts);
}

int add_key_checked(zkg_server* server, const zkg::KeyShape& shape, zkg_buffer* const g1[4], zkg_buffer* g2,
                    int* key_id) {
    ZKG_REQUIRE(server && g1 && g2 && key_id, "null argument");
    for (int i = 0; i < 4; ++i) ZKG_REQUIRE(g1[i], "null argument");
    ZKG_GUARD_BEGIN
    shape.validate();
    for (int i = 0; i < 4; ++i)
        if (g1[i]->b->bytes < 64 * shape.points[i])
            return zkg::fail(zkg::ZKG_ERR_SIZE, bases_short(g1[i]->b, shape.points[i], 16));
    if (g2->b->bytes < 128 * shape.points[4]) return zkg::fail(zkg::ZKG_ERR_SIZE, bases_short(g2->b, shape.points[4], 32));
    const uint32_t* p[4];
    size_t bytes = g2->b->bytes;
    for (int i = 0; i < 4; ++i) {
        p[i] = static_cast<const uint32_t*>(g1[i]->b->ptr);
        bytes += g1[i]->b->bytes;
    }
    *key_id = server->s->add_key(shape, p, static_cast<const uint32_t*>(g2->b->ptr), bytes);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

}  // namespace

extern "C" {

int zkg_server_create(zkg_device* dev, size_t mem_budget_bytes, uint32_t max_inflight, zkg_server** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    *out = new zkg_server{new zkg::Server(dev->d, mem_budget_bytes, max_inflight), dev->d};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_server_destroy(zkg_server* server) {
    if (!server) return zkg::ZKG_OK;
    delete server->s;
    delete server;
    return zkg::ZKG_OK;
}

int zkg_server_add_key(zkg_server* server, uint32_t log_n, zkg_buffer* const g1[4], zkg_buffer* g2, int* key_id) {
    if (log_n < 10 || log_n > 27) return zkg::fail(zkg::ZKG_ERR_INVALID, "log_n must be in 10..=27");
    return add_key_checked(server, zkg::KeyShape::groth16(log_n), g1, g2, key_id);
}

int zkg_server_add_key_shaped(zkg_server* server, const zkg_key_shape* shape, zkg_buffer* const g1[4], zkg_buffer* g2,
                              int* key_id) {
    ZKG_REQUIRE(shape, "null argument");
    if (shape->h_form != ZKG_H_COEFFICIENTS && shape->h_form != ZKG_H_ODD_COSET)
        return zkg::fail(zkg::ZKG_ERR_INVALID, "unknown H form");
    zkg::KeyShape k;
    k.log_n = shape->log_n;
    for (int m = 0; m < 5; ++m) k.points[m] = shape->points[m];
    k.c_offset = shape->c_offset;
    k.h_form = static_cast<zkg::HForm>(shape->h_form);
    return add_key_checked(server, k, g1, g2, key_id);
}

int zkg_server_add_zkey(zkg_server* server, const zkg_zkey* zkey, int* key_id) {
    ZKG_REQUIRE(server && zkey && key_id, "null argument");
    const zkg::Zkey& z = *zkey->z;
    const uint32_t log_n = (uint32_t)__builtin_ctz(z.domain_size);
    if (log_n < 10 || log_n > 27)
        return zkg::fail(zkg::ZKG_ERR_INVALID, "the .zkey domain is 2^" + std::to_string(log_n) +
                                                   "; a served key needs log_n in 10..=27");
    zkg::KeyShape shape;
    shape.log_n = log_n;
    const int sections[5] = {zkg::ZKEY_A, zkg::ZKEY_B1, zkg::ZKEY_C, zkg::ZKEY_H, zkg::ZKEY_B2};
    for (int m = 0; m < 5; ++m) shape.points[m] = z.points(sections[m]);
    shape.c_offset = (size_t)z.n_public + 1;
    shape.h_form = zkg::HForm::odd_coset;  // snarkjs's H bases
    std::vector<zkg::Buffer*> owned;
    ZKG_GUARD_BEGIN
    shape.validate();
    size_t bytes = 0;
    for (int m = 0; m < 5; ++m) {
        const size_t nbytes = 4 * zkg::Zkey::point_words(sections[m]) * shape.points[m];
        owned.push_back(zkg::buffer_alloc(server->d, nbytes));
        // the stored words are the library's Montgomery form (R = 2^256), reduced below p on the
        // device as upload_bases_montgomery_256 does
        ZKG_HIP(hipMemcpy(owned.back()->ptr, z.words(sections[m]), nbytes, hipMemcpyHostToDevice));
        zkg::msm_reduce_coords(static_cast<uint32_t*>(owned.back()->ptr), nbytes / 32, server->d->stream);
        zkey_check_section(server->d, sections[m], static_cast<const uint32_t*>(owned.back()->ptr), shape.points[m]);
        bytes += nbytes;
    }
    const uint32_t* g1[4];
    for (int m = 0; m < 4; ++m) g1[m] = static_cast<const uint32_t*>(owned[m]->ptr);
    *key_id = server->s->add_key(shape, g1, static_cast<const uint32_t*>(owned[4]->ptr), bytes, owned);
    return zkg::ZKG_OK;
    }
    catch (const zkg::Error& e) {
        for (zkg::Buffer* b : owned) zkg::buffer_free(b);
        if (e.gpu) return zkg::fail_gpu(server->d, e.code, e.what());
        return zkg::fail(e.code, e.what());
    }
    catch (const std::bad_alloc&) {
        for (zkg::Buffer* b : owned) zkg::buffer_free(b);
        return zkg::fail(zkg::ZKG_ERR_NOMEM, "host allocation failed");
    }
    catch (const std::exception& e) {
        for (zkg::Buffer* b : owned) zkg::buffer_free(b);
        return zkg::fail(zkg::ZKG_ERR_INTERNAL, e.what());
    }
}

int zkg_server_key_shape(zkg_server* server, int key_id, zkg_key_shape* out) {
    ZKG_REQUIRE(server && out, "null argument");
    ZKG_GUARD_BEGIN
    const zkg::KeyShape k = server->s->key_shape(key_id);
    out->log_n = k.log_n;
    for (int m = 0; m < 5; ++m) out->points[m] = k.points[m];
    out->c_offset = k.c_offset;
    out->h_form = (int)k.h_form;
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_server_slot_bytes(zkg_server* server, uint32_t log_n, size_t* bytes) {
    ZKG_REQUIRE(server && bytes, "null argument");
    if (log_n < 10 || log_n > 27) return zkg::fail(zkg::ZKG_ERR_INVALID, "log_n must be in 10..=27");
    ZKG_GUARD_BEGIN
    *bytes = server->s->slot_bytes(zkg::KeyShape::groth16(log_n));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_server_key_slot_bytes(zkg_server* server, int key_id, size_t* bytes) {
    ZKG_REQUIRE(server && bytes, "null argument");
    ZKG_GUARD_BEGIN
    *bytes = server->s->slot_bytes(server->s->key_shape(key_id));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_server_submit(zkg_server* server, int key_id, zkg_buffer* const polys[7], uint32_t inverse_mask,
                      const zkg_buffer* witness, uint64_t* job_id) {
    ZKG_REQUIRE(server && polys && witness && job_id, "null argument");
    ZKG_GUARD_BEGIN
    const zkg::KeyShape shape = server->s->key_shape(key_id);
    const uint32_t log_n = shape.log_n;
    const size_t need = (size_t)32 << log_n;
    for (int p = 0; p < 7; ++p) {
        ZKG_REQUIRE(polys[p], "null argument");
        if (polys[p]->b->bytes < need)
            return zkg::fail(zkg::ZKG_ERR_SIZE, "data buffer holds " + std::to_string(polys[p]->b->bytes) +
                                                    " bytes; a 2^" + std::to_string(log_n) + " transform needs " +
                                                    std::to_string(need));
    }
    const size_t wlen = shape.witness_len();
    if (witness->b->bytes < 32 * wlen)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "scalars buffer holds " + std::to_string(witness->b->bytes) + " bytes; " +
                                                std::to_string(wlen) + " scalars need " + std::to_string(32 * wlen));
    uint32_t* p[7];
    for (int k = 0; k < 7; ++k) p[k] = static_cast<uint32_t*>(polys[k]->b->ptr);
    *job_id = server->s->submit(key_id, p, inverse_mask, static_cast<const uint32_t*>(witness->b->ptr));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_server_submit_groth16(zkg_server* server, int key_id, zkg_buffer* a, zkg_buffer* b, zkg_buffer* c,
                              const zkg_buffer* witness, uint64_t* job_id) {
    ZKG_REQUIRE(server && a && b && c && witness && job_id, "null argument");
    if (a->b->ptr == b->b->ptr || a->b->ptr == c->b->ptr || b->b->ptr == c->b->ptr)
        return zkg::fail(zkg::ZKG_ERR_INVALID, "a, b and c must be distinct buffers");
    ZKG_GUARD_BEGIN
    const zkg::KeyShape shape = server->s->key_shape(key_id);
    const size_t need = (size_t)32 << shape.log_n;
    for (zkg_buffer* x : {a, b, c})
        if (x->b->bytes < need)
            return zkg::fail(zkg::ZKG_ERR_SIZE, "data buffer holds " + std::to_string(x->b->bytes) + " bytes; a 2^" +
                                                    std::to_string(shape.log_n) + " transform needs " +
                                                    std::to_string(need));
    const size_t wlen = shape.witness_len();
    if (witness->b->bytes < 32 * wlen)
        return zkg::fail(zkg::ZKG_ERR_SIZE, "scalars buffer holds " + std::to_string(witness->b->bytes) + " bytes; " +
                                                std::to_string(wlen) + " scalars need " + std::to_string(32 * wlen));
    *job_id = server->s->submit_groth16(key_id, static_cast<uint32_t*>(a->b->ptr), static_cast<uint32_t*>(b->b->ptr),
                                        static_cast<uint32_t*>(c->b->ptr),
                                        static_cast<const uint32_t*>(witness->b->ptr));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_server_done(zkg_server* server, uint64_t job_id, int* done) {
    ZKG_REQUIRE(server && done, "null argument");
    ZKG_GUARD_BEGIN
    *done = server->s->done(job_id) ? 1 : 0;
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_server_wait(zkg_server* server, uint64_t job_id, uint32_t* results, double* latency_ms) {
    ZKG_REQUIRE(server, "null argument");
    ZKG_GUARD_BEGIN
    server->s->wait(job_id, results, latency_ms);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_server_get_stats(zkg_server* server, zkg_server_stats* out) {
    ZKG_REQUIRE(server && out, "null argument");
    const zkg::ServerStats st = server->s->stats();
    out->submitted = st.submitted;
    out->completed = st.completed;
    out->p50_ms = st.p50_ms;
    out->p99_ms = st.p99_ms;
    out->mean_ms = st.mean_ms;
    out->jobs_per_s = st.jobs_per_s;
    out->bytes_in_use = st.bytes_in_use;
    out->peak_bytes = st.peak_bytes;
    out->inflight = st.inflight;
    out->peak_inflight = st.peak_inflight;
    return zkg::ZKG_OK;
}

int zkg_server_reset_stats(zkg_server* server) {
    ZKG_REQUIRE(server, "null argument");
    server->s->reset_stats();
    return zkg::ZKG_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ snarkjs .zkey

extern "C" {

int zkg_zkey_open(const char* path, zkg_zkey** out) {
    ZKG_REQUIRE(path && out, "null argument");
    ZKG_GUARD_BEGIN
    std::unique_ptr<zkg::Zkey> z(new zkg::Zkey(path));
    *out = new zkg_zkey{z.release()};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_zkey_close(zkg_zkey* zkey) {
    if (!zkey) return zkg::ZKG_OK;
    delete zkey->z;
    delete zkey;
    return zkg::ZKG_OK;
}

int zkg_zkey_get_info(const zkg_zkey* zkey, zkg_zkey_info* out) {
    ZKG_REQUIRE(zkey && out, "null argument");
    const zkg::Zkey& z = *zkey->z;
    out->n_vars = z.n_vars;
    out->n_public = z.n_public;
    out->domain_size = z.domain_size;
    // verification-key coordinates: Montgomery (R = 2^256) as stored -> plain words
    const zkg::host::Modulus& F = zkg::host::fq();
    auto plain = [&](uint32_t* dst, const uint32_t* src, int coords) {
        for (int i = 0; i < coords; ++i)
            zkg::host::to_words32(dst + 8 * i, F.from_mont(zkg::host::from_words32(src + 8 * i)));
    };
    plain(out->alpha1, z.alpha1, 2);
    plain(out->beta1, z.beta1, 2);
    plain(out->beta2, z.beta2, 4);
    plain(out->gamma2, z.gamma2, 4);
    plain(out->delta1, z.delta1, 2);
    plain(out->delta2, z.delta2, 4);
    return zkg::ZKG_OK;
}

int zkg_zkey_points(const zkg_zkey* zkey, int section, size_t* count) {
    ZKG_REQUIRE(zkey && count, "null argument");
    ZKG_GUARD_BEGIN
    *count = zkey->z->points(section);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_zkey_upload(zkg_device* dev, const zkg_zkey* zkey, int section, zkg_buffer** out) {
    ZKG_REQUIRE(dev && zkey && out, "null argument");
    ZKG_GUARD_BEGIN
    const size_t points = zkey->z->points(section);
    const size_t words = points * zkg::Zkey::point_words(section);
    int rc = zkg_msm_upload_bases_montgomery_256(dev, zkey->z->words(section), words, out);
    if (rc != zkg::ZKG_OK) return rc;
    try {
        zkey_check_section(dev->d, section, static_cast<const uint32_t*>((*out)->b->ptr), points);
    } catch (...) {
        zkg_buffer_free(*out);
        *out = nullptr;
        throw;
    }
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

}  // extern "C"

// ------------------------------------------------------------------ autotuning

#include "tune.hpp"

extern "C" {

int zkg_tune_msm(zkg_device* dev, int group, size_t n, uint32_t* c, uint32_t* item_points, uint32_t* sort_points,
                 double* ms) {
    ZKG_REQUIRE(dev, "null argument");
    ZKG_GUARD_BEGIN
    const zkg::MsmTune t = zkg::tune_msm(dev->d, group, n);
    if (c) *c = t.c;
    if (item_points) *item_points = t.item_points;
    if (sort_points) *sort_points = t.sort_points;
    if (ms) *ms = t.ms;
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_tune_ntt(zkg_device* dev, uint32_t log_n, uint32_t* tile_log, double* ms) {
    ZKG_REQUIRE(dev, "null argument");
    ZKG_GUARD_BEGIN
    const zkg::NttTune t = zkg::tune_ntt(dev->d, log_n);
    if (tile_log) *tile_log = (uint32_t)t.tile_log;
    if (ms) *ms = t.ms;
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_tune_lookup_msm(const char* arch, int group, size_t n, int* found, uint32_t* c, uint32_t* item_points,
                        uint32_t* sort_points) {
    ZKG_REQUIRE(arch && found, "null argument");
    ZKG_GUARD_BEGIN
    zkg::MsmTune t;
    *found = zkg::tune_lookup_msm(arch, group, zkg::tune_size_key(n), &t) ? 1 : 0;
    if (*found) {
        if (c) *c = t.c;
        if (item_points) *item_points = t.item_points;
        if (sort_points) *sort_points = t.sort_points;
    }
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_tune_lookup_ntt(const char* arch, uint32_t log_n, int* found, uint32_t* tile_log) {
    ZKG_REQUIRE(arch && found, "null argument");
    ZKG_GUARD_BEGIN
    zkg::NttTune t;
    *found = zkg::tune_lookup_ntt(arch, log_n, &t) ? 1 : 0;
    if (*found && tile_log) *tile_log = (uint32_t)t.tile_log;
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_tune_load(const char* path) {
    ZKG_REQUIRE(path, "null argument");
    ZKG_GUARD_BEGIN
    zkg::tune_load(path);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_tune_save(const char* path) {
    ZKG_REQUIRE(path, "null argument");
    ZKG_GUARD_BEGIN
    zkg::tune_save(path);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_tune_clear(void) {
    ZKG_GUARD_BEGIN
    zkg::tune_clear();
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

}  // extern "C"

// ------------------------------------------------------------------ Groth16 quotient polynomial

#include "h_poly.hpp"

struct zkg_groth16_h {
    zkg::GrothH* p;
};

extern "C" {

int zkg_groth16_h_create(zkg_device* dev, uint32_t log_n, zkg_groth16_h** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    if (log_n < 10 || log_n > 27) return zkg::fail(zkg::ZKG_ERR_INVALID, "log_n must be in 10..=27");
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    *out = new zkg_groth16_h{new zkg::GrothH(dev->d, log_n)};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_groth16_h_create_form(zkg_device* dev, uint32_t log_n, int form, zkg_groth16_h** out) {
    ZKG_REQUIRE(dev && out, "null argument");
    if (log_n < 10 || log_n > 27) return zkg::fail(zkg::ZKG_ERR_INVALID, "log_n must be in 10..=27");
    if (form != ZKG_H_COEFFICIENTS && form != ZKG_H_ODD_COSET) return zkg::fail(zkg::ZKG_ERR_INVALID, "unknown H form");
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    *out = new zkg_groth16_h{new zkg::GrothH(dev->d, log_n, static_cast<zkg::HForm>(form))};
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_groth16_h_destroy(zkg_groth16_h* plan) {
    if (!plan) return zkg::ZKG_OK;
    (void)hipSetDevice(plan->p->dev->ordinal);
    (void)hipStreamSynchronize(plan->p->dev->stream);
    delete plan->p;
    delete plan;
    return zkg::ZKG_OK;
}

int zkg_groth16_h_run(zkg_groth16_h* plan, zkg_buffer* a, zkg_buffer* b, zkg_buffer* c) {
    ZKG_REQUIRE(plan && a && b && c, "null argument");
    const size_t need = (size_t)32 << plan->p->log_n;
    for (zkg_buffer* x : {a, b, c})
        if (x->b->bytes < need)
            return zkg::fail(zkg::ZKG_ERR_SIZE, "data buffer holds " + std::to_string(x->b->bytes) + " bytes; a 2^" +
                                                    std::to_string(plan->p->log_n) + " transform needs " +
                                                    std::to_string(need));
    if (a->b->ptr == b->b->ptr || a->b->ptr == c->b->ptr || b->b->ptr == c->b->ptr)
        return zkg::fail(zkg::ZKG_ERR_INVALID, "a, b and c must be distinct buffers");
    ZKG_GUARD_BEGIN
    zkg::device_bind(plan->p->dev);
    plan->p->run(static_cast<uint32_t*>(a->b->ptr), static_cast<uint32_t*>(b->b->ptr),
                 static_cast<uint32_t*>(c->b->ptr), plan->p->dev->stream);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

}  // extern "C"
