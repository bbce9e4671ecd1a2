// BN254 prime-field arithmetic for gfx950: 8 x 32-bit limbs, Montgomery form with R = 2^256.
//
// Semantics restated from the reference's field layer (src/field.rs:17-20 moduli, :106-181
// mul/add/sub). The reference uses 20 x 13-bit limbs and R = 2^260 because WGSL has no wide
// multiply; here every limb product is one native v_mad_u64_u32 (32 x 32 + 64 -> 64 bits), and
// R = 2^256 so that 2^256-Montgomery inputs (snarkjs/ffjavascript form) need no conversion.
//
// Montgomery multiplication is CIOS ("coarsely integrated operand scanning"): for each word b_i,
// t += a * b_i, then t += m * p with m = t_0 * (-p^-1) mod 2^32, then t >>= 32. Each row's eight
// partial products are independent v_mad_u64_u32 (t_j + a_j * b_i cannot overflow 64 bits); their
// high halves are folded in by a single add-with-carry chain, so a row costs 8 mads + 8 carries.
// p < 2^254 leaves two spare bits, so the running value never needs a tenth word and the result
// is below 2p before the one conditional subtraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zkg {

#define ZKG_DEV __device__ __forceinline__
#define ZKG_HD __host__ __device__ __forceinline__

// Field parameters as constexpr word functions, so that unrolled loops fold them into literals.
struct FqParams {  // p = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
    static constexpr uint32_t INV = 0xe4866389u;  // -p^-1 mod 2^32
    static ZKG_HD constexpr uint32_t mod(int i) {
        return i == 0 ? 0xd87cfd47u : i == 1 ? 0x3c208c16u : i == 2 ? 0x6871ca8du : i == 3 ? 0x97816a91u
             : i == 4 ? 0x8181585du : i == 5 ? 0xb85045b6u : i == 6 ? 0xe131a029u : 0x30644e72u;
    }
    static ZKG_HD constexpr uint32_t one(int i) {  // 2^256 mod p
        return i == 0 ? 0xc58f0d9du : i == 1 ? 0xd35d438du : i == 2 ? 0xf5c70b3du : i == 3 ? 0x0a78eb28u
             : i == 4 ? 0x7879462cu : i == 5 ? 0x666ea36fu : i == 6 ? 0x9a07df2fu : 0x0e0a77c1u;
    }
    static ZKG_HD constexpr uint32_t r2(int i) {  // 2^512 mod p
        return i == 0 ? 0x538afa89u : i == 1 ? 0xf32cfc5bu : i == 2 ? 0xd44501fbu : i == 3 ? 0xb5e71911u
             : i == 4 ? 0x0a417ff6u : i == 5 ? 0x47ab1effu : i == 6 ? 0xcab8351fu : 0x06d89f71u;
    }
};

struct FrParams {  // r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001
    static constexpr uint32_t INV = 0xefffffffu;
    static ZKG_HD constexpr uint32_t mod(int i) {
        return i == 0 ? 0xf0000001u : i == 1 ? 0x43e1f593u : i == 2 ? 0x79b97091u : i == 3 ? 0x2833e848u
             : i == 4 ? 0x8181585du : i == 5 ? 0xb85045b6u : i == 6 ? 0xe131a029u : 0x30644e72u;
    }
    static ZKG_HD constexpr uint32_t one(int i) {
        return i == 0 ? 0x4ffffffbu : i == 1 ? 0xac96341cu : i == 2 ? 0x9f60cd29u : i == 3 ? 0x36fc7695u
             : i == 4 ? 0x7879462eu : i == 5 ? 0x666ea36fu : i == 6 ? 0x9a07df2fu : 0x0e0a77c1u;
    }
    static ZKG_HD constexpr uint32_t r2(int i) {
        return i == 0 ? 0xae216da7u : i == 1 ? 0x1bb8e645u : i == 2 ? 0xe35c59e3u : i == 3 ? 0x53fe3ab1u
             : i == 4 ? 0x53bb8085u : i == 5 ? 0x8c49833du : i == 6 ? 0x7f4e44a5u : 0x0216d0b1u;
    }
};

template <class P>
struct Fp {
    uint32_t w[8];
};

using Fq = Fp<FqParams>;
using Fr = Fp<FrParams>;

template <class P>
ZKG_DEV Fp<P> fp_zero() {
    Fp<P> r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.w[i] = 0;
    return r;
}

template <class P>
ZKG_DEV Fp<P> fp_one() {  // Montgomery form of 1
    Fp<P> r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.w[i] = P::one(i);
    return r;
}

template <class P>
ZKG_DEV bool fp_is_zero(const Fp<P>& a) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) x |= a.w[i];
    return x == 0;
}

template <class P>
ZKG_DEV bool fp_eq(const Fp<P>& a, const Fp<P>& b) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) x |= a.w[i] ^ b.w[i];
    return x == 0;
}

// r = a - p if a >= p, else a   (a < 2p)
template <class P>
ZKG_DEV void fp_reduce_once(uint32_t a[8]) {
    uint32_t d[8];
    unsigned borrow = 0, b2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        d[i] = __builtin_subc(a[i], P::mod(i), borrow, &b2);
        borrow = b2;
    }
    // borrow == 0  <=>  a >= p
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = borrow ? a[i] : d[i];
}

template <class P>
ZKG_DEV Fp<P> operator+(const Fp<P>& a, const Fp<P>& b) {
    Fp<P> r;
    unsigned c = 0, c2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        r.w[i] = __builtin_addc(a.w[i], b.w[i], c, &c2);
        c = c2;
    }
    // a + b < 2p < 2^255: no carry out of the top word.
    fp_reduce_once<P>(r.w);
    return r;
}

template <class P>
ZKG_DEV Fp<P> operator-(const Fp<P>& a, const Fp<P>& b) {
    Fp<P> r;
    unsigned borrow = 0, b2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        r.w[i] = __builtin_subc(a.w[i], b.w[i], borrow, &b2);
        borrow = b2;
    }
    // add p back when the subtraction wrapped
    const uint32_t mask = 0u - (uint32_t)borrow;
    unsigned c = 0, c2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        r.w[i] = __builtin_addc(r.w[i], P::mod(i) & mask, c, &c2);
        c = c2;
    }
    return r;
}

template <class P>
ZKG_DEV Fp<P> fp_neg(const Fp<P>& a) {
    return fp_zero<P>() - a;
}

template <class P>
ZKG_DEV Fp<P> fp_dbl(const Fp<P>& a) {
    return a + a;
}

// Montgomery product a * b * 2^-256 mod p (inputs < p, output < p).
template <class P>
ZKG_DEV Fp<P> operator*(const Fp<P>& a, const Fp<P>& b) {
    uint32_t t[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        // t += a * b_i : eight independent mads, then one carry chain over the high halves
        uint32_t lo[8], hi[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint64_t s = (uint64_t)a.w[j] * b.w[i] + t[j];
            lo[j] = (uint32_t)s;
            hi[j] = (uint32_t)(s >> 32);
        }
        unsigned c = 0, c2;
        t[0] = lo[0];
#pragma unroll
        for (int j = 1; j < 8; ++j) {
            t[j] = __builtin_addc(lo[j], hi[j - 1], c, &c2);
            c = c2;
        }
        const uint32_t top = hi[7] + c;  // word 8 of the running value; cannot overflow
        // t += m * p, t >>= 32   (lo[0] becomes 0 by the choice of m)
        const uint32_t m = t[0] * P::INV;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint64_t s = (uint64_t)m * P::mod(j) + t[j];
            lo[j] = (uint32_t)s;
            hi[j] = (uint32_t)(s >> 32);
        }
        c = 0;
#pragma unroll
        for (int j = 1; j < 8; ++j) {
            t[j - 1] = __builtin_addc(lo[j], hi[j - 1], c, &c2);
            c = c2;
        }
        t[7] = top + hi[7] + c;  // the running value is below 2p < 2^255 here
    }
    fp_reduce_once<P>(t);
    Fp<P> r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.w[j] = t[j];
    return r;
}

template <class P>
ZKG_DEV Fp<P> fp_sqr(const Fp<P>& a) {
    return a * a;
}

// plain -> Montgomery (x * R^2 * R^-1 = x R)
template <class P>
ZKG_DEV Fp<P> fp_to_mont(const Fp<P>& a) {
    Fp<P> r2;
#pragma unroll
    for (int i = 0; i < 8; ++i) r2.w[i] = P::r2(i);
    return a * r2;
}

// Montgomery -> plain (x R * 1 * R^-1 = x)
template <class P>
ZKG_DEV Fp<P> fp_from_mont(const Fp<P>& a) {
    Fp<P> one;
#pragma unroll
    for (int i = 0; i < 8; ++i) one.w[i] = i == 0 ? 1u : 0u;
    return a * one;
}

// Global-memory helpers: an element is 8 consecutive u32 (two 16-byte vectors).
template <class P>
ZKG_DEV Fp<P> fp_load(const uint32_t* __restrict__ src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4 a = s[0], b = s[1];
    Fp<P> r;
    r.w[0] = a.x; r.w[1] = a.y; r.w[2] = a.z; r.w[3] = a.w;
    r.w[4] = b.x; r.w[5] = b.y; r.w[6] = b.z; r.w[7] = b.w;
    return r;
}

template <class P>
ZKG_DEV void fp_store(uint32_t* __restrict__ dst, const Fp<P>& v) {
    uint4* d = reinterpret_cast<uint4*>(dst);
    d[0] = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
    d[1] = make_uint4(v.w[4], v.w[5], v.w[6], v.w[7]);
}

}  // namespace zkg
