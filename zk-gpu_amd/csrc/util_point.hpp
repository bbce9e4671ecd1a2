// Point-law test kernels and the fixed-base generator-multiples kernel, templated over the
// coordinate field; G1 instances live in util_kernels.hip, G2 instances in util_g2.hip.
#pragma once
#include "curve.hpp"
#include "device_util.hpp"

namespace zkg {

// ---------------------------------------------------------------- point ops
template <class Fe>
__global__ void point_op_kernel(int op, const uint32_t* p, const uint32_t* q, uint32_t* out, size_t count) {
    typedef FeTraits<Fe> T;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Affine<Fe> a = load_plain_affine<Fe>(p + 2 * T::WORDS * i);
    const Affine<Fe> b = load_plain_affine<Fe>(q + 2 * T::WORDS * i);
    const Xyzz<Fe> pa = xyzz_from_affine(a), pb = xyzz_from_affine(b);
    Xyzz<Fe> r;
    switch (op) {
        case 0: r = xyzz_add(pa, pb); break;
        case 1: r = xyzz_dbl(pa); break;
        case 2: r = xyzz_madd(pa, b); break;
        default: r = xyzz_add(xyzz_add(pa, pb), xyzz_dbl(pa)); break;
    }
    store_plain_affine<Fe>(out + 2 * T::WORDS * i, r);
}

// ---------------------------------------------------------------- generator multiples
// out[i] = k_i * G by left-to-right double-and-add in XYZZ, then one Fermat inversion each.
template <class Fe>
__global__ __launch_bounds__(256) void generator_multiples_kernel(const uint32_t* ks, uint32_t* out, size_t n) {
    typedef FeTraits<Fe> T;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Affine<Fe> g = generator<Fe>();
    Xyzz<Fe> acc = xyzz_inf<Fe>();
    for (int word = 7; word >= 0; --word) {
        const uint32_t k = ks[8 * i + word];
        for (int bit = 31; bit >= 0; --bit) {
            acc = xyzz_dbl(acc);
            if ((k >> bit) & 1u) acc = xyzz_madd_nonzero(acc, g);
        }
    }
    store_plain_affine<Fe>(out + 2 * T::WORDS * i, acc);
}

// ---------------------------------------------------------------- precomputed MSM bases
// out[i] = bases[i] and out[n + i] = 2^64 bases[i], i < n, in the library's form (Montgomery,
// affine; infinity stays (0, 0)): 64 XYZZ doublings and one Fermat inversion per point
// (1/ZZ = t ZZZ and 1/ZZZ = t ZZ with t = 1/(ZZ ZZZ)). No counterpart in the reference.
template <class Fe>
__global__ __launch_bounds__(256) void msm_shift64_kernel(const uint32_t* __restrict__ bases, uint32_t* __restrict__ out,
                                                         size_t n) {
    typedef FeTraits<Fe> T;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Affine<Fe> a = load_affine<Fe>(bases + 2 * T::WORDS * i);
    T::store(out + 2 * T::WORDS * i, a.x);
    T::store(out + 2 * T::WORDS * i + T::WORDS, a.y);
    Xyzz<Fe> q = xyzz_from_affine(a);
    for (int k = 0; k < 64; ++k) q = xyzz_dbl(q);
    uint32_t* o = out + 2 * T::WORDS * (n + i);
    if (xyzz_is_inf(q)) {
        for (int k = 0; k < 2 * T::WORDS; ++k) o[k] = 0;
        return;
    }
    const Fe t = fe_inverse(q.zz * q.zzz);
    T::store(o, q.x * (t * q.zzz));
    T::store(o + T::WORDS, q.y * (t * q.zz));
}

// ---------------------------------------------------------------- base-point checks
// The curve coefficient b in the library's Montgomery form (R = 2^256): G1 y^2 = x^3 + 3; the G2
// twist y^2 = x^3 + 3 / (9 + u), 3 / (9 + u) = (27 - 3u) / 82.
template <class Fe>
__device__ __forceinline__ Fe curve_b();
template <>
__device__ __forceinline__ Fq curve_b<Fq>() {
    const uint32_t w[8] = {0x50ad28d7u, 0x7a17caa9u, 0xe15521b9u, 0x1f6ac17au,
                           0x696bd284u, 0x334bea4eu, 0xce179d8eu, 0x2a1f6744u};
    Fq r;
    for (int i = 0; i < 8; ++i) r.w[i] = w[i];
    return r;
}
template <>
__device__ __forceinline__ Fq2 curve_b<Fq2>() {
    const uint32_t w[16] = {0x77b802a8u, 0x3bf938e3u, 0x3633535du, 0x020b1b27u, 0x49755260u, 0x26b7edf0u,
                            0x4384a86du, 0x2514c632u, 0xd1dcff67u, 0x38e7ecccu, 0x93ce0d3eu, 0x65f0b37du,
                            0x22ac00aau, 0xd749d0ddu, 0x4a688d4du, 0x0141b9ceu};
    Fq2 r;
    for (int i = 0; i < 8; ++i) {
        r.a.w[i] = w[i];
        r.b.w[i] = w[8 + i];
    }
    return r;
}

// Base i (library form) is accepted when it is infinity (0, 0) or lies on the curve and, for G2,
// in the order-r subgroup (r P = O, by double-and-add over the bits of r). G1 has cofactor 1, so
// on the curve means in the subgroup. Rejected points are counted in bad[0] and the smallest
// rejected index is kept in bad[1].
template <class Fe>
__global__ __launch_bounds__(256) void check_bases_kernel(const uint32_t* __restrict__ bases, size_t n,
                                                         unsigned long long* __restrict__ bad) {
    typedef FeTraits<Fe> T;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Affine<Fe> a = load_affine<Fe>(bases + 2 * T::WORDS * i);
    if (affine_is_inf(a)) return;
    const Fe lhs = T::sqr(a.y), rhs = T::sqr(a.x) * a.x + curve_b<Fe>();
    bool ok = T::is_zero(lhs - rhs);
    if (ok && T::WORDS == 16) {
        Xyzz<Fe> acc = xyzz_inf<Fe>();
        for (int word = 7; word >= 0; --word) {
            const uint32_t k = FrParams::mod(word);
            for (int bit = 31; bit >= 0; --bit) {
                acc = xyzz_dbl(acc);
                if ((k >> bit) & 1u) acc = xyzz_madd(acc, a);
            }
        }
        ok = xyzz_is_inf(acc);
    }
    if (!ok) {
        atomicAdd(&bad[0], 1ull);
        atomicMin(&bad[1], (unsigned long long)i);
    }
}

template <class Fe>
void launch_check_bases(const uint32_t* bases, size_t n, unsigned long long* bad, hipStream_t s) {
    hipLaunchKernelGGL(check_bases_kernel<Fe>, dim3(ceil_div(n, 256)), dim3(256), 0, s, bases, n, bad);
}

template <class Fe>
void launch_shift64(const uint32_t* bases, uint32_t* out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(msm_shift64_kernel<Fe>, dim3(ceil_div(n, 256)), dim3(256), 0, s, bases, out, n);
}

template <class Fe>
void launch_point_op(int op, const uint32_t* p, const uint32_t* q, uint32_t* out, size_t count, hipStream_t s) {
    hipLaunchKernelGGL(point_op_kernel<Fe>, dim3(ceil_div(count, 64)), dim3(64), 0, s, op, p, q, out, count);
}

template <class Fe>
void launch_generator_multiples(const uint32_t* ks, uint32_t* out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(generator_multiples_kernel<Fe>, dim3(ceil_div(n, 256)), dim3(256), 0, s, ks, out, n);
}

}  // namespace zkg
