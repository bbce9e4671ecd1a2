// Test and benchmark hooks of the C ABI (include/zkgpu.h, "test and benchmark hooks"):
// element-wise field ops, point-law known-answer kernels, the field-multiply throughput
// benchmark, and a fixed-base generator-multiples kernel for synthetic inputs.
//
// They stand in for the reference's WGSL point-op harness (tests/curve.rs:8-30) and field
// benchmark (examples/field_bench.rs:10-97), using the same device code as the MSM and NTT.
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/zkgpu.h"
#include "api_types.hpp"
#include "curve.hpp"
#include "device_util.hpp"
#include "field29.hpp"
#include "msm.hpp"
#include "util_point.hpp"
#include "runtime.hpp"

namespace zkg {

// ---------------------------------------------------------------- field ops
template <class Fe>
__global__ void field_op_kernel(int op, const uint32_t* a, const uint32_t* b, uint32_t* out, size_t count) {
    typedef FeTraits<Fe> T;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Fe x = T::to_mont(T::load(a + T::WORDS * i));
    const Fe y = T::to_mont(T::load(b + T::WORDS * i));
    Fe r;
    switch (op) {
        case 0: r = x * y; break;
        case 1: r = x + y; break;
        case 2: r = x - y; break;
        default: r = T::sqr(x); break;
    }
    T::store(out + T::WORDS * i, T::from_mont(r));
}

// ---------------------------------------------------------------- field-multiply benchmark
template <class P, int WAYS>
__global__ __launch_bounds__(256) void field_bench_kernel(uint32_t chain, uint32_t* sink) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    Fp<P> x[WAYS], y;
#pragma unroll
    for (int k = 0; k < 8; ++k) y.w[k] = (t * 2654435761u + k * 40503u) & 0x0fffffffu;
#pragma unroll
    for (int w = 0; w < WAYS; ++w)
#pragma unroll
        for (int k = 0; k < 8; ++k) x[w].w[k] = (t + 97u * w + k) & 0x0fffffffu;
    for (uint32_t i = 0; i < chain; ++i)
#pragma unroll
        for (int w = 0; w < WAYS; ++w) x[w] = x[w] * y;
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < WAYS; ++w)
#pragma unroll
        for (int k = 0; k < 8; ++k) acc ^= x[w].w[k];
    if (acc == 0x9e3779b9u) sink[0] = acc;  // practically never true; keeps the chains live
}

// The same chains in the MSM's nine 29-bit limbs (f29_mul modulo p): operands normalised below
// 2p, so each product (bound 4p^2, far inside the 169p^2 the product allows) is again below 2p.
template <int WAYS>
__global__ __launch_bounds__(256) void field29_bench_kernel(uint32_t chain, uint32_t* sink) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    F29 x[WAYS], y;
#pragma unroll
    for (int k = 0; k < 9; ++k) y.l[k] = (t * 2654435761u + k * 40503u) & (k == 8 ? 0xfffffu : f29::MASK);
#pragma unroll
    for (int w = 0; w < WAYS; ++w)
#pragma unroll
        for (int k = 0; k < 9; ++k) x[w].l[k] = (t + 97u * w + k) & (k == 8 ? 0xfffffu : f29::MASK);
    for (uint32_t i = 0; i < chain; ++i)
#pragma unroll
        for (int w = 0; w < WAYS; ++w) x[w] = f29_mul(x[w], y);
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < WAYS; ++w)
#pragma unroll
        for (int k = 0; k < 9; ++k) acc ^= x[w].l[k];
    if (acc == 0x9e3779b9u) sink[0] = acc;  // practically never true; keeps the chains live
}

// Fq inversions in the 29-bit limbs by Fermat's little theorem, x^-1 = x^(p-2), square-and-multiply
// over the bits of p - 2 from the top: 253 squarings and 109 multiplications, 362 products. Each
// lane inverts its value `chain` times in sequence (x <- x^-1, so every inversion depends on the
// previous one), checks x * x^-1 = 1 (the Montgomery form 2^261 mod p) after each, and counts
// failures in sink[1].
ZKG_DEV constexpr uint32_t p_minus_2_word(int i) {
    return i == 0 ? 0xd87cfd45u : i == 1 ? 0x3c208c16u : i == 2 ? 0x6871ca8du : i == 3 ? 0x97816a91u
         : i == 4 ? 0x8181585du : i == 5 ? 0xb85045b6u : i == 6 ? 0xe131a029u : 0x30644e72u;
}

__global__ __launch_bounds__(256) void field29_inv_bench_kernel(uint32_t chain, uint32_t* sink) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    F29 x;
#pragma unroll
    for (int k = 0; k < 9; ++k) x.l[k] = (t * 2654435761u + k * 40503u + 1u) & (k == 8 ? 0xfffffu : f29::MASK);
    uint32_t bad = 0;
    for (uint32_t i = 0; i < chain; ++i) {
        F29 r = x;  // bit 253, the top bit of p - 2
        for (int bit = 252; bit >= 0; --bit) {
            r = f29_mul(r, r);
            if ((p_minus_2_word(bit >> 5) >> (bit & 31)) & 1u) r = f29_mul(r, x);
        }
        const F29 d = f29_sub<2>(f29_mul(r, x), f29_const_one());  // x x^-1 - 1 + 2p, below 4p
        bad += f29_is_zero_below12p(d) ? 0u : 1u;
        x = r;
    }
    if (bad) atomicAdd(&sink[1], bad);
}

extern template void launch_point_op<Fq2>(int, const uint32_t*, const uint32_t*, uint32_t*, size_t, hipStream_t);
extern template void launch_generator_multiples<Fq2>(const uint32_t*, uint32_t*, size_t, hipStream_t);
extern template void launch_shift64<Fq2>(const uint32_t*, uint32_t*, size_t, hipStream_t);

extern template void launch_check_bases<Fq2>(const uint32_t*, size_t, unsigned long long*, hipStream_t);

void msm_check_bases(Device* d, int group, const uint32_t* bases, size_t n, uint64_t* bad, uint64_t* first_bad) {
    device_bind(d);
    DeviceArray<unsigned long long> flags(2);
    const unsigned long long init[2] = {0ull, ~0ull};
    hipStream_t s = d->stream;
    ZKG_HIP(hipMemcpyAsync(flags.ptr, init, sizeof init, hipMemcpyHostToDevice, s));
    if (n) {
        if (group == 1) launch_check_bases<Fq>(bases, n, flags.ptr, s);
        else launch_check_bases<Fq2>(bases, n, flags.ptr, s);
        ZKG_HIP(hipGetLastError());
    }
    unsigned long long out[2];
    ZKG_HIP(hipMemcpyAsync(out, flags.ptr, sizeof out, hipMemcpyDeviceToHost, s));
    ZKG_HIP(hipStreamSynchronize(s));
    *bad = out[0];
    *first_bad = out[0] ? out[1] : 0;
}

void msm_precompute_bases(int group, const uint32_t* bases, size_t n, uint32_t* out, hipStream_t stream) {
    if (!n) return;
    if (group == 1) launch_shift64<Fq>(bases, out, n, stream);
    else launch_shift64<Fq2>(bases, out, n, stream);
    ZKG_HIP(hipGetLastError());
}

}  // namespace zkg

namespace zkg {

// dst = src, 16 bytes per lane, grid-stride: the reference copy for HBM counter calibration
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace zkg

namespace zkg {

// Synthetic MSM inputs for tuning and benchmarks: bases k_i * G with k_i < 2^253 from a seeded
// xorshift64* stream, converted in place to the library's Montgomery form, and scalars below
// 2^252 from the same stream. `bases` holds 16 * group * n words, `scalars` 8 * n.
void synthetic_msm_inputs(Device* d, int group, size_t n, uint64_t seed, uint32_t* bases, uint32_t* scalars) {
    device_bind(d);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&x]() {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        return (uint32_t)((x * 0x2545F4914F6CDD1Dull) >> 32);
    };
    std::vector<uint32_t> host(8 * n);
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 8; ++k) host[8 * i + k] = next();
        host[8 * i + 7] &= 0x1fffffffu;
    }
    DeviceArray<uint32_t> ks(8 * n);
    hipStream_t s = d->stream;
    ZKG_HIP(hipMemcpyAsync(ks.ptr, host.data(), 32 * n, hipMemcpyHostToDevice, s));
    if (group == 1) launch_generator_multiples<Fq>(ks.ptr, bases, n, s);
    else launch_generator_multiples<Fq2>(ks.ptr, bases, n, s);
    ZKG_HIP(hipGetLastError());
    msm_convert_bases(bases, 2 * (size_t)group * n, s);
    ZKG_HIP(hipStreamSynchronize(s));  // `host` is rewritten below
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 8; ++k) host[8 * i + k] = next();
        host[8 * i + 7] &= 0x0fffffffu;
    }
    ZKG_HIP(hipMemcpyAsync(scalars, host.data(), 32 * n, hipMemcpyHostToDevice, s));
    ZKG_HIP(hipStreamSynchronize(s));
}

}  // namespace zkg

extern "C" {

int zkg_test_field_op(zkg_device* dev, int field, int op, const uint32_t* a, const uint32_t* b, uint32_t* out,
                      size_t count) {
    if (!dev || !a || !b || !out) return zkg::fail(zkg::ZKG_ERR_INVALID, "null argument");
    if (field < 0 || field > 2 || op < 0 || op > 3) return zkg::fail(zkg::ZKG_ERR_INVALID, "bad field or op");
    if (count == 0) return zkg::ZKG_OK;
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    const size_t words = (field == 2 ? 16 : 8) * count;
    zkg::DeviceArray<uint32_t> da(words), db(words), dout(words);
    hipStream_t s = dev->d->stream;
    ZKG_HIP(hipMemcpyAsync(da.ptr, a, 4 * words, hipMemcpyHostToDevice, s));
    ZKG_HIP(hipMemcpyAsync(db.ptr, b, 4 * words, hipMemcpyHostToDevice, s));
    const unsigned grid = zkg::ceil_div(count, 128);
    if (field == 0)
        hipLaunchKernelGGL(zkg::field_op_kernel<zkg::Fq>, dim3(grid), dim3(128), 0, s, op, da.ptr, db.ptr, dout.ptr, count);
    else if (field == 1)
        hipLaunchKernelGGL(zkg::field_op_kernel<zkg::Fr>, dim3(grid), dim3(128), 0, s, op, da.ptr, db.ptr, dout.ptr, count);
    else
        hipLaunchKernelGGL(zkg::field_op_kernel<zkg::Fq2>, dim3(grid), dim3(128), 0, s, op, da.ptr, db.ptr, dout.ptr, count);
    ZKG_HIP(hipGetLastError());
    ZKG_HIP(hipMemcpyAsync(out, dout.ptr, 4 * words, hipMemcpyDeviceToHost, s));
    ZKG_HIP(hipStreamSynchronize(s));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_test_point_op(zkg_device* dev, int group, int op, const uint32_t* p, const uint32_t* q, uint32_t* out,
                      size_t count) {
    if (!dev || !p || !q || !out) return zkg::fail(zkg::ZKG_ERR_INVALID, "null argument");
    if ((group != ZKG_G1 && group != ZKG_G2) || op < 0 || op > 3) return zkg::fail(zkg::ZKG_ERR_INVALID, "bad group or op");
    if (count == 0) return zkg::ZKG_OK;
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    const size_t words = (group == ZKG_G2 ? 32 : 16) * count;
    zkg::DeviceArray<uint32_t> dp(words), dq(words), dout(words);
    hipStream_t s = dev->d->stream;
    ZKG_HIP(hipMemcpyAsync(dp.ptr, p, 4 * words, hipMemcpyHostToDevice, s));
    ZKG_HIP(hipMemcpyAsync(dq.ptr, q, 4 * words, hipMemcpyHostToDevice, s));
    if (group == ZKG_G1) zkg::launch_point_op<zkg::Fq>(op, dp.ptr, dq.ptr, dout.ptr, count, s);
    else zkg::launch_point_op<zkg::Fq2>(op, dp.ptr, dq.ptr, dout.ptr, count, s);
    ZKG_HIP(hipGetLastError());
    ZKG_HIP(hipMemcpyAsync(out, dout.ptr, 4 * words, hipMemcpyDeviceToHost, s));
    ZKG_HIP(hipStreamSynchronize(s));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_bench_field_mul(zkg_device* dev, int field, uint32_t threads, uint32_t chain, uint32_t ways, double* ms) {
    if (!dev || !ms) return zkg::fail(zkg::ZKG_ERR_INVALID, "null argument");
    if (field < 0 || field > 3 || threads == 0 || threads % 256 || !(ways == 1 || ways == 2 || ways == 4))
        return zkg::fail(zkg::ZKG_ERR_INVALID, "field 0/1/2/3, threads a multiple of 256, ways 1/2/4");
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    hipStream_t s = dev->d->stream;
    zkg::DeviceArray<uint32_t> sink(2);
    ZKG_HIP(hipMemsetAsync(sink.ptr, 0, 2 * sizeof(uint32_t), s));
    hipEvent_t e0, e1;
    ZKG_HIP(hipEventCreate(&e0));
    ZKG_HIP(hipEventCreate(&e1));
    const dim3 grid(threads / 256), block(256);
    ZKG_HIP(hipEventRecord(e0, s));
#define ZKG_BENCH_LAUNCH(P)                                                                                     \
    if (ways == 1) hipLaunchKernelGGL((zkg::field_bench_kernel<P, 1>), grid, block, 0, s, chain, sink.ptr);    \
    else if (ways == 2) hipLaunchKernelGGL((zkg::field_bench_kernel<P, 2>), grid, block, 0, s, chain, sink.ptr); \
    else hipLaunchKernelGGL((zkg::field_bench_kernel<P, 4>), grid, block, 0, s, chain, sink.ptr);
    if (field == 3) {
        hipLaunchKernelGGL(zkg::field29_inv_bench_kernel, grid, block, 0, s, chain, sink.ptr);
    } else if (field == 0) {
        ZKG_BENCH_LAUNCH(zkg::FqParams)
    } else if (field == 1) {
        ZKG_BENCH_LAUNCH(zkg::FrParams)
    } else if (ways == 1) {
        hipLaunchKernelGGL(zkg::field29_bench_kernel<1>, grid, block, 0, s, chain, sink.ptr);
    } else if (ways == 2) {
        hipLaunchKernelGGL(zkg::field29_bench_kernel<2>, grid, block, 0, s, chain, sink.ptr);
    } else {
        hipLaunchKernelGGL(zkg::field29_bench_kernel<4>, grid, block, 0, s, chain, sink.ptr);
    }
#undef ZKG_BENCH_LAUNCH
    ZKG_HIP(hipGetLastError());
    ZKG_HIP(hipEventRecord(e1, s));
    ZKG_HIP(hipEventSynchronize(e1));
    float f = 0;
    ZKG_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    uint32_t failures[2] = {0, 0};
    ZKG_HIP(hipMemcpy(failures, sink.ptr, sizeof(failures), hipMemcpyDeviceToHost));
    if (field == 3 && failures[1])
        return zkg::fail(zkg::ZKG_ERR_INTERNAL, std::to_string(failures[1]) + " inversions failed x * x^-1 = 1");
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_bench_copy(zkg_device* dev, size_t bytes, uint32_t reps, double* ms) {
    if (!dev || !ms) return zkg::fail(zkg::ZKG_ERR_INVALID, "null argument");
    if (bytes == 0 || bytes % 16 || reps == 0) return zkg::fail(zkg::ZKG_ERR_INVALID, "bytes a positive multiple of 16, reps >= 1");
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    hipStream_t s = dev->d->stream;
    zkg::DeviceArray<uint4> src(bytes / 16), dst(bytes / 16);
    ZKG_HIP(hipMemsetAsync(src.ptr, 0, bytes, s));
    hipEvent_t e0, e1;
    ZKG_HIP(hipEventCreate(&e0));
    ZKG_HIP(hipEventCreate(&e1));
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    ZKG_HIP(hipEventRecord(e0, s));
    for (uint32_t r = 0; r < reps; ++r) hipLaunchKernelGGL(zkg::copy_kernel, dim3(grid), dim3(256), 0, s, src.ptr, dst.ptr, n);
    ZKG_HIP(hipGetLastError());
    ZKG_HIP(hipEventRecord(e1, s));
    ZKG_HIP(hipEventSynchronize(e1));
    float f = 0;
    ZKG_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_test_glv_split(zkg_device* dev, const uint32_t* scalars, size_t n, uint32_t* out) {
    if (!dev || !scalars || !out) return zkg::fail(zkg::ZKG_ERR_INVALID, "null argument");
    if (n == 0) return zkg::ZKG_OK;
    if (n > 0x3fffffffu) return zkg::fail(zkg::ZKG_ERR_INVALID, "too many scalars");
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    zkg::DeviceArray<uint32_t> ds(8 * n), dout(8 * n);
    hipStream_t s = dev->d->stream;
    ZKG_HIP(hipMemcpyAsync(ds.ptr, scalars, 32 * n, hipMemcpyHostToDevice, s));
    zkg::msm_glv_split(ds.ptr, (uint32_t)n, dout.ptr, s);
    ZKG_HIP(hipMemcpyAsync(out, dout.ptr, 32 * n, hipMemcpyDeviceToHost, s));
    ZKG_HIP(hipStreamSynchronize(s));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_test_gls_split(zkg_device* dev, const uint32_t* scalars, size_t n, uint32_t* out) {
    if (!dev || !scalars || !out) return zkg::fail(zkg::ZKG_ERR_INVALID, "null argument");
    if (n == 0) return zkg::ZKG_OK;
    if (n > 0x3fffffffu) return zkg::fail(zkg::ZKG_ERR_INVALID, "too many scalars");
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    zkg::DeviceArray<uint32_t> ds(8 * n), dout(12 * n);
    hipStream_t s = dev->d->stream;
    ZKG_HIP(hipMemcpyAsync(ds.ptr, scalars, 32 * n, hipMemcpyHostToDevice, s));
    zkg::msm_gls_split(ds.ptr, (uint32_t)n, dout.ptr, s);
    ZKG_HIP(hipMemcpyAsync(out, dout.ptr, 48 * n, hipMemcpyDeviceToHost, s));
    ZKG_HIP(hipStreamSynchronize(s));
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

int zkg_util_generator_multiples(zkg_device* dev, int group, const zkg_buffer* ks, size_t n, zkg_buffer* out) {
    if (!dev || !ks || !out) return zkg::fail(zkg::ZKG_ERR_INVALID, "null argument");
    if (group != ZKG_G1 && group != ZKG_G2) return zkg::fail(zkg::ZKG_ERR_INVALID, "bad group");
    const size_t pw = group == ZKG_G2 ? 32 : 16;
    if (ks->b->bytes < 32 * n || out->b->bytes < 4 * pw * n) return zkg::fail(zkg::ZKG_ERR_SIZE, "buffer too small");
    if (n == 0) return zkg::ZKG_OK;
    ZKG_GUARD_BEGIN
    zkg::device_bind(dev->d);
    hipStream_t s = dev->d->stream;
    const uint32_t* k = static_cast<const uint32_t*>(ks->b->ptr);
    uint32_t* o = static_cast<uint32_t*>(out->b->ptr);
    if (group == ZKG_G1) zkg::launch_generator_multiples<zkg::Fq>(k, o, n, s);
    else zkg::launch_generator_multiples<zkg::Fq2>(k, o, n, s);
    ZKG_HIP(hipGetLastError());
    return zkg::ZKG_OK;
    ZKG_GUARD_END
}

}  // extern "C"
