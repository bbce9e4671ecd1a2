// Device helpers shared by kernels: RAII device arrays, Fermat inversion, plain <-> Montgomery
// affine conversion, and the group generators in Montgomery form.
//
// Every device function is force-inlined: the library's kernels contain no calls. An out-of-line
// function large enough to need relaxed (s_setpc) branches has had its return address in
// s[30:31] overwritten by those long jumps, so it never returned; the build checks that no code
// object contains a call (check_code_objects.py).
#pragma once
#include "curve.hpp"
#include "runtime.hpp"

namespace zkg {

// Temporary device allocation released on scope exit (after the stream work using it has been
// synchronised by the caller).
template <class T>
struct DeviceArray {
    T* ptr = nullptr;
    explicit DeviceArray(size_t count) {
        if (count) ZKG_HIP(hipMalloc(&ptr, count * sizeof(T)));
    }
    ~DeviceArray() {
        if (ptr) (void)hipFree(ptr);
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
};

template <>
struct FeTraits<Fr> {
    static constexpr int WORDS = 8;
    static ZKG_DEV Fr zero() { return fp_zero<FrParams>(); }
    static ZKG_DEV Fr one() { return fp_one<FrParams>(); }
    static ZKG_DEV bool is_zero(const Fr& x) { return fp_is_zero(x); }
    static ZKG_DEV Fr sqr(const Fr& x) { return fp_sqr(x); }
    static ZKG_DEV Fr load(const uint32_t* p) { return fp_load<FrParams>(p); }
    static ZKG_DEV void store(uint32_t* p, const Fr& x) { fp_store(p, x); }
    static ZKG_DEV Fr to_mont(const Fr& x) { return fp_to_mont(x); }
    static ZKG_DEV Fr from_mont(const Fr& x) { return fp_from_mont(x); }
};

// x^(m-2) (Fermat inverse, Montgomery in and out); 0 maps to 0.
template <class P>
__device__ __forceinline__ Fp<P> fp_inverse(const Fp<P>& x) {
    Fp<P> acc = fp_one<P>();
    for (int word = 7; word >= 0; --word) {
        // m - 2: the low word of either BN254 modulus is >= 2, so there is no borrow
        const uint32_t e = word == 0 ? P::mod(0) - 2u : P::mod(word);
        for (int bit = 31; bit >= 0; --bit) {
            acc = acc * acc;
            if ((e >> bit) & 1u) acc = acc * x;
        }
    }
    return acc;
}

// 1/(a + bu) = (a - bu)/(a^2 + b^2)
__device__ __forceinline__ Fq2 fq2_inverse(const Fq2& x) {
    const Fq ni = fp_inverse(x.a * x.a + x.b * x.b);
    return Fq2{x.a * ni, fp_zero<FqParams>() - x.b * ni};
}

__device__ __forceinline__ Fq fe_inverse(const Fq& x) { return fp_inverse(x); }
__device__ __forceinline__ Fq2 fe_inverse(const Fq2& x) { return fq2_inverse(x); }

// XYZZ (Montgomery) -> plain affine words; infinity -> zeros.
template <class Fe>
__device__ __forceinline__ void store_plain_affine(uint32_t* out, const Xyzz<Fe>& p) {
    typedef FeTraits<Fe> T;
    if (xyzz_is_inf(p)) {
        for (int i = 0; i < 2 * T::WORDS; ++i) out[i] = 0;
        return;
    }
    const Fe x = p.x * fe_inverse(p.zz);
    const Fe y = p.y * fe_inverse(p.zzz);
    T::store(out, T::from_mont(x));
    T::store(out + T::WORDS, T::from_mont(y));
}

template <class Fe>
__device__ __forceinline__ Affine<Fe> load_plain_affine(const uint32_t* p) {
    typedef FeTraits<Fe> T;
    const Affine<Fe> a = load_affine<Fe>(p);
    return Affine<Fe>{T::to_mont(a.x), T::to_mont(a.y)};
}

template <class Fe>
__device__ __forceinline__ Affine<Fe> generator();

template <>
__device__ __forceinline__ Affine<Fq> generator<Fq>() {  // G1 = (1, 2)
    Fq x = fp_zero<FqParams>(), y = fp_zero<FqParams>();
    x.w[0] = 1;
    y.w[0] = 2;
    return Affine<Fq>{fp_to_mont(x), fp_to_mont(y)};
}

template <>
__device__ __forceinline__ Affine<Fq2> generator<Fq2>() {
    // BN254 G2 generator: x.a, x.b, y.a, y.b as plain little-endian words.
    const uint32_t g[32] = {
        0xd992f6edu, 0x46debd5cu, 0xf75edaddu, 0x674322d4u, 0x5e5c4479u, 0x426a0066u, 0x121f1e76u, 0x1800deefu,
        0xaef312c2u, 0x97e485b7u, 0x35a9e712u, 0xf1aa4933u, 0x31fb5d25u, 0x7260bfb7u, 0x920d483au, 0x198e9393u,
        0x66fa7daau, 0x4ce6cc01u, 0x0c43d37bu, 0xe3d1e769u, 0x8dcb408fu, 0x4aab7180u, 0xdb8c6debu, 0x12c85ea5u,
        0xd122975bu, 0x55acdadcu, 0x70b38ef3u, 0xbc4b3133u, 0x690c3395u, 0xec9e99adu, 0x585ff075u, 0x090689d0u};
    Fq c[4];
    for (int k = 0; k < 4; ++k) {
        for (int i = 0; i < 8; ++i) c[k].w[i] = g[8 * k + i];
        c[k] = fp_to_mont(c[k]);
    }
    return Affine<Fq2>{Fq2{c[0], c[1]}, Fq2{c[2], c[3]}};
}

}  // namespace zkg
