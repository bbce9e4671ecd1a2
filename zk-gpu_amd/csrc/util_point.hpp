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
