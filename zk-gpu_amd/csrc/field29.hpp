// BN254 base-field arithmetic in nine 29-bit limbs (Montgomery R = 2^261), for the MSM's
// bucket accumulation.
//
// Why a second representation: with 32-bit limbs every partial product of a Montgomery multiply
// needs an add-with-carry and a register move besides its v_mad_u64_u32 (the hardware adds one
// 64-bit operand and the compiler does not use the multiply-add's carry-out), about 390
// instructions per product. With 29-bit limbs a product of two limbs is below 2^60 and a whole
// column of the schoolbook product plus the reduction terms (at most 18 such products) stays
// below 2^64, so product scanning ("finely integrated product scanning", Koc, Acar and Kaliski,
// IEEE Micro 1996) accumulates every partial product with one v_mad_u64_u32 and no carry
// handling; a column costs one shift and mask. About 235 instructions per product.
//
// Values are kept lazily reduced. A value v represents v * 2^-261 mod p; limbs 0..7 are below
// 2^29 after a multiplication or a subtraction ("normalised") and limb 8 holds the rest. The
// multiplication needs limbs below 2^30 and v(a) * v(b) < p * 2^261 (about 169 p^2), and then
// returns a normalised value below 2p. Callers track value bounds in units of p (the bucket
// addition in msm_kernels29.hpp states its bounds step by step).
#pragma once
#include <stdint.h>

#include "curve.hpp"
#include "field.hpp"

namespace zkg {

struct F29 {
    uint32_t l[9];
};

namespace f29 {

static constexpr uint32_t MASK = (1u << 29) - 1;
static constexpr uint32_t PINV = 0x04866389u;  // -p^-1 mod 2^29

ZKG_DEV constexpr uint32_t P(int i) {  // p
    return i == 0 ? 0x187cfd47u : i == 1 ? 0x010460b6u : i == 2 ? 0x1c72a34fu : i == 3 ? 0x02d522d0u
         : i == 4 ? 0x1585d978u : i == 5 ? 0x02db40c0u : i == 6 ? 0x00a6e141u : i == 7 ? 0x0e5c2634u : 0x0030644eu;
}

// c * p in "borrowed" limbs: limb 0 plus 2^31, limbs 1..7 plus 2^31 - 4, limb 8 minus 4, so that
// subtracting any limb below 2^31 - 4 leaves every limb non-negative (same value c * p).
template <int C>
ZKG_DEV constexpr uint32_t K(int i);
template <>
ZKG_DEV constexpr uint32_t K<2>(int i) {
    return i == 0 ? 0x90f9fa8eu : i == 1 ? 0x8208c169u : i == 2 ? 0x98e5469au : i == 3 ? 0x85aa459du
         : i == 4 ? 0x8b0bb2ecu : i == 5 ? 0x85b6817du : i == 6 ? 0x814dc27eu : i == 7 ? 0x9cb84c64u : 0x0060c898u;
}
template <>
ZKG_DEV constexpr uint32_t K<3>(int i) {
    return i == 0 ? 0x8976f7d5u : i == 1 ? 0x830d2220u : i == 2 ? 0x9557e9e9u : i == 3 ? 0x887f686eu
         : i == 4 ? 0x80918c64u : i == 5 ? 0x8891c23eu : i == 6 ? 0x81f4a3bfu : i == 7 ? 0x8b147298u : 0x00912ce7u;
}
template <>
ZKG_DEV constexpr uint32_t K<4>(int i) {
    return i == 0 ? 0x81f3f51cu : i == 1 ? 0x841182d7u : i == 2 ? 0x91ca8d38u : i == 3 ? 0x8b548b3fu
         : i == 4 ? 0x961765dcu : i == 5 ? 0x8b6d02feu : i == 6 ? 0x829b8500u : i == 7 ? 0x997098ccu : 0x00c19135u;
}
template <>
ZKG_DEV constexpr uint32_t K<6>(int i) {
    return i == 0 ? 0x92edefaau : i == 1 ? 0x861a4444u : i == 2 ? 0x8aafd3d6u : i == 3 ? 0x90fed0e1u
         : i == 4 ? 0x812318ccu : i == 5 ? 0x91238480u : i == 6 ? 0x83e94782u : i == 7 ? 0x9628e534u : 0x012259d2u;
}
template <>
ZKG_DEV constexpr uint32_t K<7>(int i) {
    return i == 0 ? 0x8b6aecf1u : i == 1 ? 0x871ea4fbu : i == 2 ? 0x87227725u : i == 3 ? 0x93d3f3b2u
         : i == 4 ? 0x96a8f244u : i == 5 ? 0x93fec540u : i == 6 ? 0x849028c3u : i == 7 ? 0x84850b68u : 0x0152be21u;
}
template <>
ZKG_DEV constexpr uint32_t K<8>(int i) {
    return i == 0 ? 0x83e7ea38u : i == 1 ? 0x882305b2u : i == 2 ? 0x83951a74u : i == 3 ? 0x96a91683u
         : i == 4 ? 0x8c2ecbbcu : i == 5 ? 0x96da0601u : i == 6 ? 0x85370a04u : i == 7 ? 0x92e1319cu : 0x0183226fu;
}
template <>
ZKG_DEV constexpr uint32_t K<10>(int i) {
    return i == 0 ? 0x94e1e4c6u : i == 1 ? 0x8a2bc71fu : i == 2 ? 0x9c7a6112u : i == 3 ? 0x9c535c24u
         : i == 4 ? 0x973a7eacu : i == 5 ? 0x9c908782u : i == 6 ? 0x8684cc86u : i == 7 ? 0x8f997e04u : 0x01e3eb0cu;
}

template <>
ZKG_DEV constexpr uint32_t K<12>(int i) {
    return i == 0 ? 0x85dbdf54u : i == 1 ? 0x8c34888du : i == 2 ? 0x955fa7b0u : i == 3 ? 0x81fda1c6u
         : i == 4 ? 0x8246319du : i == 5 ? 0x82470904u : i == 6 ? 0x87d28f09u : i == 7 ? 0x8c51ca6cu : 0x0244b3a9u;
}

ZKG_DEV constexpr uint32_t TWO_P(int i) {  // 2p, normalised
    return i == 0 ? 0x10f9fa8eu : i == 1 ? 0x0208c16du : i == 2 ? 0x18e5469eu : i == 3 ? 0x05aa45a1u
         : i == 4 ? 0x0b0bb2f0u : i == 5 ? 0x05b68181u : i == 6 ? 0x014dc282u : i == 7 ? 0x1cb84c68u : 0x0060c89cu;
}

ZKG_DEV constexpr uint32_t ONE(int i) {  // 2^261 mod p: the Montgomery form of 1
    return i == 0 ? 0x157ccc21u : i == 1 ? 0x141c2758u : i == 2 ? 0x185230d3u : i == 3 ? 0x014c0419u
         : i == 4 ? 0x0aa36fb9u : i == 5 ? 0x1d4240ceu : i == 6 ? 0x11d54c07u : i == 7 ? 0x052ac7a8u : 0x000dc836u;
}

ZKG_DEV constexpr uint32_t C256(int i) {  // 2^256 mod p: multiplying by it maps R = 2^261 form to R = 2^256 form
    return i == 0 ? 0x058f0d9du : i == 1 ? 0x1aea1c6eu : i == 2 ? 0x11c2cf74u : i == 3 ? 0x11d651ebu
         : i == 4 ? 0x1462c0a7u : i == 5 ? 0x11b7bc3cu : i == 6 ? 0x1cbd99bau : i == 7 ? 0x183340fbu : 0x000e0a77u;
}

// Limb 0 of k * p for k = 0..11 (a value below 12p is 0 mod p only if its limb 0 is one of these).
ZKG_DEV constexpr uint32_t KP0(int k) {
    return k == 0 ? 0u : k == 1 ? 0x187cfd47u : k == 2 ? 0x10f9fa8eu : k == 3 ? 0x0976f7d5u : k == 4 ? 0x01f3f51cu
         : k == 5 ? 0x1a70f263u : k == 6 ? 0x12edefaau : k == 7 ? 0x0b6aecf1u : k == 8 ? 0x03e7ea38u
         : k == 9 ? 0x1c64e77fu : k == 10 ? 0x14e1e4c6u : 0x0d5ee20du;
}


// p 2^257 in borrowed 64-bit columns (f29_mul2): column k's 29-bit digit of p 2^257, plus 2^62
// for k <= 15, minus 2^33 for k >= 1 (the two cancel between neighbouring columns).
ZKG_DEV constexpr uint64_t M2COL(int k) {  // p 2^257 in borrowed 64-bit columns (f29_mul2)
    return k == 0 ? 0x4000000000000000ull : k == 1 ? 0x3ffffffe00000000ull : k == 2 ? 0x3ffffffe00000000ull
         : k == 3 ? 0x3ffffffe00000000ull : k == 4 ? 0x3ffffffe00000000ull : k == 5 ? 0x3ffffffe00000000ull
         : k == 6 ? 0x3ffffffe00000000ull : k == 7 ? 0x3ffffffe00000000ull : k == 8 ? 0x3ffffffe0e000000ull
         : k == 9 ? 0x3ffffffe0d87cfd4ull : k == 10 ? 0x3ffffffe1e10460bull : k == 11 ? 0x3ffffffe01c72a34ull
         : k == 12 ? 0x3ffffffe102d522dull : k == 13 ? 0x3ffffffe01585d97ull : k == 14 ? 0x3ffffffe022db40cull
         : k == 15 ? 0x3ffffffe080a6e14ull
         : 0x000060c69ce5c263ull;
}

// p 2^261 in borrowed 64-bit columns (f29_mul2<true>, for lazier inputs): p's limbs at columns
// 9..15 and the rest of p in column 16, with the same 2^62 / 2^33 borrows
ZKG_DEV constexpr uint64_t M2COL_WIDE(int k) {
    return (k <= 15 ? (1ull << 62) : 0ull) - (k >= 1 ? (1ull << 33) : 0ull) +
           (k >= 9 && k <= 15 ? (uint64_t)P(k - 9) : 0ull) +
           (k == 16 ? (uint64_t)P(7) + ((uint64_t)P(8) << 29) : 0ull);
}

}  // namespace f29

ZKG_DEV F29 f29_const_one() {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = f29::ONE(i);
    return r;
}

// The moduli the 29-bit product serves: limbs (0..7 below 2^29) and -modulus^-1 mod 2^29.
struct Fq29Mod {  // p: the base field
    static ZKG_DEV constexpr uint32_t P(int i) { return f29::P(i); }
    static constexpr uint32_t PINV = f29::PINV;
};
struct Fr29Mod {  // r: the scalar field (the NTT's); 2^261 / r is about 169.28 as for p
    static ZKG_DEV constexpr uint32_t P(int i) {
        return i == 0 ? 0x10000001u : i == 1 ? 0x1f0fac9fu : i == 2 ? 0x0e5c2450u : i == 3 ? 0x07d090f3u
             : i == 4 ? 0x1585d283u : i == 5 ? 0x02db40c0u : i == 6 ? 0x00a6e141u : i == 7 ? 0x0e5c2634u : 0x0030644eu;
    }
    static constexpr uint32_t PINV = 0x0fffffffu;  // r = 1 + 2^28 mod 2^29
};

// Montgomery product a * b * 2^-261 (FIPS) modulo M's modulus m (p unless stated). Needs limbs
// < 2^30 and v(a) v(b) < m 2^261; the result is normalised and below a b / 2^261 + m < 2m.
// Column sums: at most nine products below 2^60, nine below 2^58 and a carry below 2^35, i.e.
// below 1.3 * 2^63.
template <class M = Fq29Mod>
ZKG_DEV F29 f29_mul(const F29& a, const F29& b) {
    uint64_t acc = 0;
    uint32_t m[9];
    F29 r;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int j = 0; j < k; ++j) acc += (uint64_t)m[j] * M::P(k - j);
        m[k] = ((uint32_t)acc * M::PINV) & f29::MASK;
        acc += (uint64_t)m[k] * M::P(0);  // clears the low 29 bits
        acc >>= 29;
    }
#pragma unroll
    for (int k = 9; k < 17; ++k) {
#pragma unroll
        for (int i = k - 8; i <= 8; ++i) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int j = k - 8; j <= 8; ++j) acc += (uint64_t)m[j] * M::P(k - j);
        r.l[k - 9] = (uint32_t)acc & f29::MASK;
        acc >>= 29;
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// (a b - c d) 2^-261 (sub) or (a b + c d) 2^-261 mod p with the two products summed on the columns
// and one Montgomery reduction (FIPS, as f29_mul): 243 multiply-adds where two products and an
// addition take 324. For a b - c d a multiple of p is added in borrowed columns (every column at
// least 2^62 - 2^33), which keeps each column above c d's: p 2^257 (the last column p's bits from
// 207 on) for inputs normalised below 2p, where c d's last column is below 2^45.2; p 2^261 (WIDE:
// the last column p's bits from 203 on) for normalised inputs with c d below about 12p each,
// whose top limbs' product stays below 2^50.5. Columns stay below 2^63.2. The result is normalised
// and below (a b + c d) / 2^261 + p + p / 16 < 1.1p, or + 2p with WIDE.
template <bool WIDE = false>
ZKG_DEV F29 f29_mul2(const F29& a, const F29& b, const F29& c, const F29& d, bool sub) {
    uint64_t acc = 0;
    uint32_t m[9];
    F29 r;
#pragma unroll
    for (int k = 0; k < 17; ++k) {
        const int lo = k < 9 ? 0 : k - 8, hi = k < 9 ? k : 8;
        uint64_t s = 0;
#pragma unroll
        for (int i = lo; i <= hi; ++i) {
            acc += (uint64_t)a.l[i] * b.l[k - i];
            s += (uint64_t)c.l[i] * d.l[k - i];
        }
        acc += sub ? (WIDE ? f29::M2COL_WIDE(k) : f29::M2COL(k)) - s : s;
        if (k < 9) {
#pragma unroll
            for (int j = 0; j < k; ++j) acc += (uint64_t)m[j] * f29::P(k - j);
            m[k] = ((uint32_t)acc * f29::PINV) & f29::MASK;
            acc += (uint64_t)m[k] * f29::P(0);  // clears the low 29 bits
            acc >>= 29;
        } else {
#pragma unroll
            for (int j = k - 8; j <= 8; ++j) acc += (uint64_t)m[j] * f29::P(k - j);
            r.l[k - 9] = (uint32_t)acc & f29::MASK;
            acc >>= 29;
        }
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// Limb-wise sum, not normalised: limbs add, value adds.
ZKG_DEV F29 f29_add(const F29& a, const F29& b) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = a.l[i] + b.l[i];
    return r;
}

// a + C p - b, normalised. Needs v(b) < (C - 1) p, b's limbs < 2^31 - 4 and a's limbs < 2^30, so
// that no limb leaves [0, 2^32); the value is below v(a) + C p.
template <int C>
ZKG_DEV F29 f29_sub(const F29& a, const F29& b) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = a.l[i] + f29::K<C>(i) - b.l[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        r.l[i + 1] += r.l[i] >> 29;
        r.l[i] &= f29::MASK;
    }
    return r;
}

template <int C>
ZKG_DEV F29 f29_neg(const F29& b) {
    F29 z;
#pragma unroll
    for (int i = 0; i < 9; ++i) z.l[i] = 0;
    return f29_sub<C>(z, b);
}

// For a normalised value below 12p: is it 0 mod p? (k p is normalised-unique, so equality of the
// limbs decides; limb 0 filters first and the full comparison runs only on a match.)
ZKG_DEV bool f29_is_zero_below12p(const F29& a) {
    bool maybe = false;
#pragma unroll
    for (int k = 0; k < 12; ++k) maybe |= a.l[0] == f29::KP0(k);
    if (!maybe) return false;
    // rare: compare every limb with k p, computed here
    bool zero = false;
    for (uint32_t k = 0; k < 12; ++k) {
        uint64_t carry = 0;
        bool eq = true;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            carry += (uint64_t)k * f29::P(i);
            const uint32_t limb = i < 8 ? (uint32_t)carry & f29::MASK : (uint32_t)carry;
            carry >>= 29;
            eq &= a.l[i] == limb;
        }
        zero |= eq;
    }
    return zero;
}

// Eight 32-bit words holding x (< 2^256, e.g. an Fq in R = 2^256 form) -> limbs of 32 x: the same
// field element in R = 2^261 form, not reduced (value below 32p for x < p).
ZKG_DEV F29 f29_from_r256_words(const uint32_t w[8]) {
    F29 r;
    r.l[0] = (w[0] << 5) & f29::MASK;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        const int o = 29 * i - 5, q = o >> 5, s = o & 31;
        // bits [o, o + 29) of w: a funnel shift of two 32-bit words (v_alignbit_b32), which keeps w
        // in registers (a 64-bit read at a 4-byte offset would put it in scratch memory)
        r.l[i] = __builtin_amdgcn_alignbit(w[q + 1], w[q], s) & f29::MASK;
    }
    r.l[8] = w[7] >> 3;
    return r;
}

// A normalised value below 2^256 -> its eight 32-bit words.
ZKG_DEV void f29_to_words(const F29& a, uint32_t w[8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int b = 32 * j, i = b / 29, s = b - 29 * i;
        uint32_t v = (a.l[i] >> s) | (a.l[i + 1] << (29 - s));  // 29 - s <= 29: no undefined shift
        if (i + 2 <= 8 && 58 - s < 32) v |= a.l[i + 2] << (58 - s);
        w[j] = v;
    }
}

// Fq (R = 2^256, canonical) -> F29 below 2p.
ZKG_DEV F29 f29_from_fq(const Fq& x) { return f29_mul(f29_from_r256_words(x.w), f29_const_one()); }

// F29 (value below 12p) -> canonical Fq in R = 2^256 form.
ZKG_DEV Fq f29_to_fq(const F29& a) {
    F29 c;
#pragma unroll
    for (int i = 0; i < 9; ++i) c.l[i] = f29::C256(i);
    const F29 t = f29_mul(a, c);  // a 2^256 2^-261 = the R = 2^256 form, below 2p
    Fq r;
    f29_to_words(t, r.w);
    fp_reduce_once<FqParams>(r.w);
    return r;
}

// Eight 32-bit words of a value below 2^256 -> its nine limbs (no scaling).
ZKG_DEV F29 f29_unpack(const uint32_t w[8]) {
    F29 r;
    r.l[0] = w[0] & f29::MASK;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        const int o = 29 * i, q = o >> 5, s = o & 31;
        r.l[i] = __builtin_amdgcn_alignbit(w[q + 1], w[q], s) & f29::MASK;
    }
    r.l[8] = w[7] >> 8;
    return r;
}

// ---------------------------------------------------------------- Fq29: reduced values
// The same limbs with every operation returning a normalised value below 2p, so generic point
// formulas (msm_kernels.hpp's merge, finalize, reduce, window-sum and combine kernels) can use
// it like Fq. In memory an element is the R = 2^261 value (below 2p) in eight 32-bit words.
struct Fq29 {
    uint32_t l[9];
};

// limbs in [0, 2^32) of a value below 4p -> normalised, below 2p
ZKG_DEV Fq29 fq29_reduce(F29 s) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s.l[i + 1] += s.l[i] >> 29;
        s.l[i] &= f29::MASK;
    }
    Fq29 t;
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const uint32_t d = s.l[i] - f29::TWO_P(i) - br;  // in (-2^30, 2^29): the sign bit is the borrow
        br = d >> 31;
        t.l[i] = i < 8 ? d & f29::MASK : d;
    }
    Fq29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = br ? s.l[i] : t.l[i];  // borrow out: s < 2p already
    return r;
}

ZKG_DEV Fq29 operator*(const Fq29& a, const Fq29& b) {  // 2p * 2p < 169 p^2: below 2p
    F29 x, y;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        x.l[i] = a.l[i];
        y.l[i] = b.l[i];
    }
    const F29 m = f29_mul(x, y);
    Fq29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = m.l[i];
    return r;
}

ZKG_DEV Fq29 operator+(const Fq29& a, const Fq29& b) {
    F29 s;
#pragma unroll
    for (int i = 0; i < 9; ++i) s.l[i] = a.l[i] + b.l[i];
    return fq29_reduce(s);
}

// a + 2p - b in (0, 4p): limbs of 2p borrowed so none goes negative (the top limb may wrap by at
// most 4 and is made whole again by the carries of the borrowed low limbs)
ZKG_DEV Fq29 operator-(const Fq29& a, const Fq29& b) {
    F29 s;
#pragma unroll
    for (int i = 0; i < 9; ++i) s.l[i] = a.l[i] + f29::K<2>(i) - b.l[i];
    return fq29_reduce(s);
}

ZKG_DEV bool fq29_is_zero(const Fq29& a) {  // 0 or p
    uint32_t z = 0, q = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        z |= a.l[i];
        q |= a.l[i] ^ f29::P(i);
    }
    return z == 0 || q == 0;
}

template <>
struct FeTraits<Fq29> {
    static constexpr int WORDS = 8;
    static ZKG_DEV Fq29 zero() {
        Fq29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.l[i] = 0;
        return r;
    }
    static ZKG_DEV Fq29 one() {
        Fq29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.l[i] = f29::ONE(i);
        return r;
    }
    static ZKG_DEV bool is_zero(const Fq29& x) { return fq29_is_zero(x); }
    static ZKG_DEV Fq29 sqr(const Fq29& x) { return x * x; }
    static ZKG_DEV Fq29 load(const uint32_t* p) {
        const Fq w = fp_load<FqParams>(p);
        const F29 u = f29_unpack(w.w);
        Fq29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.l[i] = u.l[i];
        return r;
    }
    static ZKG_DEV void store(uint32_t* p, const Fq29& x) {
        F29 u;
#pragma unroll
        for (int i = 0; i < 9; ++i) u.l[i] = x.l[i];
        Fq w;
        f29_to_words(u, w.w);
        fp_store(p, w);
    }
    // the library's result form: canonical R = 2^256 words
    static ZKG_DEV void store_result(uint32_t* p, const Fq29& x) {
        F29 u;
#pragma unroll
        for (int i = 0; i < 9; ++i) u.l[i] = x.l[i];
        fp_store(p, f29_to_fq(u));
    }
};

// a b - c d with one reduction for both products (reduced inputs below 2p; the result below 1.1p)
ZKG_DEV Fq29 fe_mul_sub(const Fq29& a, const Fq29& b, const Fq29& c, const Fq29& d) {
    F29 fa, fb, fc, fd;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        fa.l[i] = a.l[i];
        fb.l[i] = b.l[i];
        fc.l[i] = c.l[i];
        fd.l[i] = d.l[i];
    }
    const F29 t = f29_mul2(fa, fb, fc, fd, true);
    Fq29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = t.l[i];
    return r;
}

ZKG_DEV Fq29 fe_select(bool c, const Fq29& a, const Fq29& b) {
    Fq29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = c ? a.l[i] : b.l[i];
    return r;
}

// Canonical Fq (R = 2^256) -> Fq29 (R = 2^261, below 2p).
ZKG_DEV Fq29 fq29_from_fq(const Fq& x) {
    const F29 t = f29_from_fq(x);
    Fq29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = t.l[i];
    return r;
}

// ---------------------------------------------------------------- Fq2 over Fq29
// Fq2 = Fq[u]/(u^2 + 1) as in the 32-bit Fq2 (Karatsuba product, (a+b)(a-b) and 2ab square),
// every component a reduced Fq29: the G2 stages after the bucket accumulation.
struct Fq2_29 {
    Fq29 a, b;
};

ZKG_DEV Fq2_29 operator+(const Fq2_29& x, const Fq2_29& y) { return Fq2_29{x.a + y.a, x.b + y.b}; }
ZKG_DEV Fq2_29 operator-(const Fq2_29& x, const Fq2_29& y) { return Fq2_29{x.a - y.a, x.b - y.b}; }
ZKG_DEV Fq2_29 operator*(const Fq2_29& x, const Fq2_29& y) {
    const Fq29 ac = x.a * y.a, bd = x.b * y.b;
    const Fq29 m = (x.a + x.b) * (y.a + y.b);
    return Fq2_29{ac - bd, (m - ac) - bd};
}

template <>
struct FeTraits<Fq2_29> {
    static constexpr int WORDS = 16;
    static ZKG_DEV Fq2_29 zero() { return Fq2_29{FeTraits<Fq29>::zero(), FeTraits<Fq29>::zero()}; }
    static ZKG_DEV Fq2_29 one() { return Fq2_29{FeTraits<Fq29>::one(), FeTraits<Fq29>::zero()}; }
    static ZKG_DEV bool is_zero(const Fq2_29& x) { return fq29_is_zero(x.a) && fq29_is_zero(x.b); }
    static ZKG_DEV Fq2_29 sqr(const Fq2_29& x) {
        const Fq29 ab = x.a * x.b;
        return Fq2_29{(x.a + x.b) * (x.a - x.b), ab + ab};
    }
    static ZKG_DEV Fq2_29 load(const uint32_t* p) {
        return Fq2_29{FeTraits<Fq29>::load(p), FeTraits<Fq29>::load(p + 8)};
    }
    static ZKG_DEV void store(uint32_t* p, const Fq2_29& x) {
        FeTraits<Fq29>::store(p, x.a);
        FeTraits<Fq29>::store(p + 8, x.b);
    }
    static ZKG_DEV void store_result(uint32_t* p, const Fq2_29& x) {
        FeTraits<Fq29>::store_result(p, x.a);
        FeTraits<Fq29>::store_result(p + 8, x.b);
    }
};

ZKG_DEV Fq2_29 fe_select(bool c, const Fq2_29& x, const Fq2_29& y) {
    return Fq2_29{fe_select(c, x.a, y.a), fe_select(c, x.b, y.b)};
}

// ---------------------------------------------------------------- Fq2 over a pair of lanes
// The same Fq2 elements, in the same memory form as Fq2_29, held by two adjacent lanes: the even
// lane holds component a, the odd lane component b. A product costs each lane two Fq29 products
// (after exchanging components with its partner: even lane ac - bd, odd lane ad + bc) where
// Karatsuba's formula on one lane costs three in sequence, a square costs one ((a + b)(a - b) and
// 2ab), and a point takes half the registers. The G2 stages after the bucket accumulation are
// chains of point additions on few lanes, so they run with shorter chains and more waves this
// way. Both lanes of a pair must execute every operation together (branches on pair-uniform
// conditions only: is_zero combines both lanes' answers).
struct Fq2P {
    Fq29 v;
};

ZKG_DEV bool pair_odd() { return (__lane_id() & 1u) != 0; }

// Lane exchanges within each group of four lanes as DPP moves (quad_perm: lane i of a group takes
// the value of lane sel_i, dpp_ctrl = sel0 | sel1 << 2 | sel2 << 4 | sel3 << 6): one VALU move per
// word, where __shfl and __shfl_xor go through the LDS crossbar (ds_bpermute) and wait on its
// latency. Every lane of the group must be active, as for the shuffles they replace.
constexpr int DPP_XOR1 = 0xb1;  // lanes 1, 0, 3, 2
constexpr int DPP_XOR2 = 0x4e;  // lanes 2, 3, 0, 1
template <int CTRL>
ZKG_DEV uint32_t dpp_quad(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
ZKG_DEV Fq29 fq29_dpp(const Fq29& x) {
    Fq29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = dpp_quad<CTRL>(x.l[i]);
    return r;
}

ZKG_DEV Fq29 pair_other(const Fq29& x) { return fq29_dpp<DPP_XOR1>(x); }

ZKG_DEV Fq2P operator+(const Fq2P& x, const Fq2P& y) { return Fq2P{x.v + y.v}; }
ZKG_DEV Fq2P operator-(const Fq2P& x, const Fq2P& y) { return Fq2P{x.v - y.v}; }
ZKG_DEV Fq2P operator*(const Fq2P& x, const Fq2P& y) {
    const bool odd = pair_odd();
    const Fq29 px = pair_other(x.v), py = pair_other(y.v);
    // even (a): x.a y.a - x.b y.b = x y - px py; odd (b): x.a y.b + x.b y.a = px y + x py, each
    // lane's two products summed before one reduction (below 1.1p: a reduced Fq29)
    const Fq29 u = fe_select(odd, px, x.v), w = fe_select(odd, x.v, px);
    F29 fu, fy, fw, fpy;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        fu.l[i] = u.l[i];
        fy.l[i] = y.v.l[i];
        fw.l[i] = w.l[i];
        fpy.l[i] = py.l[i];
    }
    const F29 t = f29_mul2(fu, fy, fw, fpy, !odd);
    Fq29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = t.l[i];
    return Fq2P{r};
}

template <>
struct FeTraits<Fq2P> {
    static constexpr int WORDS = 16;  // memory: the whole element, a then b (each lane reads its half)
    static ZKG_DEV Fq2P zero() { return Fq2P{FeTraits<Fq29>::zero()}; }
    static ZKG_DEV Fq2P one() { return Fq2P{fe_select(pair_odd(), FeTraits<Fq29>::zero(), FeTraits<Fq29>::one())}; }
    static ZKG_DEV bool is_zero(const Fq2P& x) {
        const int z = fq29_is_zero(x.v) ? 1 : 0;
        return (z & (int)dpp_quad<DPP_XOR1>((uint32_t)z)) != 0;
    }
    static ZKG_DEV Fq2P sqr(const Fq2P& x) {  // even: (a + b)(a - b); odd: (b + b) a
        const bool odd = pair_odd();
        const Fq29 p = pair_other(x.v);
        const Fq29 l = fe_select(odd, x.v + x.v, x.v + p);
        const Fq29 r = fe_select(odd, p, x.v - p);
        return Fq2P{l * r};
    }
    static ZKG_DEV Fq2P load(const uint32_t* p) { return Fq2P{FeTraits<Fq29>::load(p + (pair_odd() ? 8 : 0))}; }
    static ZKG_DEV void store(uint32_t* p, const Fq2P& x) { FeTraits<Fq29>::store(p + (pair_odd() ? 8 : 0), x.v); }
    static ZKG_DEV void store_result(uint32_t* p, const Fq2P& x) {
        FeTraits<Fq29>::store_result(p + (pair_odd() ? 8 : 0), x.v);
    }
};

ZKG_DEV Fq2P fe_select(bool c, const Fq2P& x, const Fq2P& y) { return Fq2P{fe_select(c, x.v, y.v)}; }

// ---------------------------------------------------------------- G1 points on a pair of lanes
// An Fq29 value held by both lanes of a pair (the same memory form as Fq29). Field operations act
// on both copies; the point formulas (xyzz_add and xyzz_dbl below) split each level of
// independent products between the two lanes and swap the results, so a point addition costs a
// lane the latency of 7 products instead of 14 and a doubling 5 instead of 9. Used by the G1
// bucket reduction and window sums, whose serial chains of additions on too few lanes set their
// time. Both lanes of a pair must execute every operation together (branches on pair-uniform
// conditions only: both lanes hold the same values).
struct Fq29Pair {
    Fq29 v;
};

ZKG_DEV Fq29Pair operator+(const Fq29Pair& x, const Fq29Pair& y) { return Fq29Pair{x.v + y.v}; }
ZKG_DEV Fq29Pair operator-(const Fq29Pair& x, const Fq29Pair& y) { return Fq29Pair{x.v - y.v}; }
ZKG_DEV Fq29Pair operator*(const Fq29Pair& x, const Fq29Pair& y) { return Fq29Pair{x.v * y.v}; }

template <>
struct FeTraits<Fq29Pair> {
    static constexpr int WORDS = 8;
    static ZKG_DEV Fq29Pair zero() { return Fq29Pair{FeTraits<Fq29>::zero()}; }
    static ZKG_DEV Fq29Pair one() { return Fq29Pair{FeTraits<Fq29>::one()}; }
    static ZKG_DEV bool is_zero(const Fq29Pair& x) { return fq29_is_zero(x.v); }
    static ZKG_DEV Fq29Pair sqr(const Fq29Pair& x) { return Fq29Pair{x.v * x.v}; }
    static ZKG_DEV Fq29Pair load(const uint32_t* p) { return Fq29Pair{FeTraits<Fq29>::load(p)}; }
    // both lanes hold the value: the even lane writes it
    static ZKG_DEV void store(uint32_t* p, const Fq29Pair& x) {
        if (!pair_odd()) FeTraits<Fq29>::store(p, x.v);
    }
    static ZKG_DEV void store_result(uint32_t* p, const Fq29Pair& x) {
        if (!pair_odd()) FeTraits<Fq29>::store_result(p, x.v);
    }
};

ZKG_DEV Fq29Pair fe_select(bool c, const Fq29Pair& x, const Fq29Pair& y) { return Fq29Pair{fe_select(c, x.v, y.v)}; }

// (even lane's m, odd lane's m) of a pair, on both lanes
ZKG_DEV void pair_swap(const Fq29& m, bool odd, Fq29* even_val, Fq29* odd_val) {
    const Fq29 o = pair_other(m);
    *even_val = fe_select(odd, o, m);
    *odd_val = fe_select(odd, m, o);
}

// 2P (dbl-2008-s-1, the formulas of the generic xyzz_dbl) with each level's products split
// between the lanes of a pair. Infinity (ZZ = ZZZ = 0) maps to infinity.
template <>
ZKG_DEV Xyzz<Fq29Pair> xyzz_dbl<Fq29Pair>(const Xyzz<Fq29Pair>& p) {
    const bool odd = pair_odd();
    const Fq29 x = p.x.v, y = p.y.v;
    const Fq29 u = y + y;
    Fq29 v, x2, s, m2, w, msx, wy, zzz3;
    // V = U^2 | X^2
    const Fq29 t0 = fe_select(odd, x, u);
    pair_swap(t0 * t0, odd, &v, &x2);
    const Fq29 mm = (x2 + x2) + x2;
    // S = X V | M^2, then ZZ3 = V ZZ on both lanes (each lane has one product slot per level)
    pair_swap(fe_select(odd, mm, x) * fe_select(odd, mm, v), odd, &s, &m2);
    const Fq29 x3 = m2 - (s + s);
    const Fq29 zz3 = v * p.zz.v;
    // W = U V | M (S - X3)
    pair_swap(fe_select(odd, mm, u) * fe_select(odd, s - x3, v), odd, &w, &msx);
    // W Y | W ZZZ
    pair_swap(w * fe_select(odd, p.zzz.v, y), odd, &wy, &zzz3);
    return Xyzz<Fq29Pair>{{x3}, {msx - wy}, {zz3}, {zzz3}};
}

// P + Q (add-2008-s, the formulas of the generic xyzz_add, every special case) with each level's
// products split between the lanes of a pair.
template <>
ZKG_DEV Xyzz<Fq29Pair> xyzz_add<Fq29Pair>(const Xyzz<Fq29Pair>& p, const Xyzz<Fq29Pair>& q) {
    if (xyzz_is_inf(q)) return p;
    if (xyzz_is_inf(p)) return q;
    const bool odd = pair_odd();
    Fq29 u1, u2, s1, s2, zz12, zzz12, ppsq, r2, ppp, qq, ye, yo, zz3, zzz3;
    // U1 = X1 ZZ2 | U2 = X2 ZZ1;  S1 = Y1 ZZZ2 | S2 = Y2 ZZZ1;  ZZ1 ZZ2 | ZZZ1 ZZZ2
    pair_swap(fe_select(odd, q.x.v, p.x.v) * fe_select(odd, p.zz.v, q.zz.v), odd, &u1, &u2);
    pair_swap(fe_select(odd, q.y.v, p.y.v) * fe_select(odd, p.zzz.v, q.zzz.v), odd, &s1, &s2);
    pair_swap(fe_select(odd, p.zzz.v, p.zz.v) * fe_select(odd, q.zzz.v, q.zz.v), odd, &zz12, &zzz12);
    const Fq29 pp = u2 - u1, r = s2 - s1;
    if (fq29_is_zero(pp)) {
        if (fq29_is_zero(r)) return xyzz_dbl(p);
        return xyzz_inf<Fq29Pair>();
    }
    // PP = P^2 | R^2
    const Fq29 t = fe_select(odd, r, pp);
    pair_swap(t * t, odd, &ppsq, &r2);
    // PPP = P PP | Q = U1 PP
    pair_swap(fe_select(odd, u1, pp) * ppsq, odd, &ppp, &qq);
    const Fq29 x3 = (r2 - ppp) - (qq + qq);
    // R (Q - X3) | S1 PPP;  ZZ3 = ZZ1 ZZ2 PP | ZZZ3 = ZZZ1 ZZZ2 PPP
    pair_swap(fe_select(odd, s1, r) * fe_select(odd, ppp, qq - x3), odd, &ye, &yo);
    pair_swap(fe_select(odd, zzz12, zz12) * fe_select(odd, ppp, ppsq), odd, &zz3, &zzz3);
    return Xyzz<Fq29Pair>{{x3}, {ye - yo}, {zz3}, {zzz3}};
}

// ---------------------------------------------------------------- G2 points on four lanes
// An Fq2P value (one Fq component per lane of a pair) held by both pairs of a group of four lanes;
// the G2 point formulas below split each level of independent Fq2 products between the two pairs,
// as Fq29Pair does for G1, halving the serial chains of the G2 bucket reduction and window sums.
// The pairs of a group execute every operation together (branches on group-uniform conditions).
struct Fq2Q {
    Fq2P v;
};

ZKG_DEV bool quad_odd() { return (__lane_id() & 2u) != 0; }

// the same component from the other pair of the group
ZKG_DEV Fq2P quad_other(const Fq2P& x) { return Fq2P{fq29_dpp<DPP_XOR2>(x.v)}; }

ZKG_DEV Fq2Q operator+(const Fq2Q& x, const Fq2Q& y) { return Fq2Q{x.v + y.v}; }
ZKG_DEV Fq2Q operator-(const Fq2Q& x, const Fq2Q& y) { return Fq2Q{x.v - y.v}; }
ZKG_DEV Fq2Q operator*(const Fq2Q& x, const Fq2Q& y) { return Fq2Q{x.v * y.v}; }

template <>
struct FeTraits<Fq2Q> {
    static constexpr int WORDS = 16;
    static ZKG_DEV Fq2Q zero() { return Fq2Q{FeTraits<Fq2P>::zero()}; }
    static ZKG_DEV Fq2Q one() { return Fq2Q{FeTraits<Fq2P>::one()}; }
    static ZKG_DEV bool is_zero(const Fq2Q& x) { return FeTraits<Fq2P>::is_zero(x.v); }
    static ZKG_DEV Fq2Q sqr(const Fq2Q& x) { return Fq2Q{FeTraits<Fq2P>::sqr(x.v)}; }
    static ZKG_DEV Fq2Q load(const uint32_t* p) { return Fq2Q{FeTraits<Fq2P>::load(p)}; }
    // both pairs hold the value: the first pair writes it
    static ZKG_DEV void store(uint32_t* p, const Fq2Q& x) {
        if (!quad_odd()) FeTraits<Fq2P>::store(p, x.v);
    }
    static ZKG_DEV void store_result(uint32_t* p, const Fq2Q& x) {
        if (!quad_odd()) FeTraits<Fq2P>::store_result(p, x.v);
    }
};

ZKG_DEV Fq2Q fe_select(bool c, const Fq2Q& x, const Fq2Q& y) { return Fq2Q{fe_select(c, x.v, y.v)}; }

// (first pair's m, second pair's m) of a group, on all four lanes
ZKG_DEV void quad_swap(const Fq2P& m, bool odd, Fq2P* even_val, Fq2P* odd_val) {
    const Fq2P o = quad_other(m);
    *even_val = fe_select(odd, o, m);
    *odd_val = fe_select(odd, m, o);
}

// 2P (dbl-2008-s-1, the formulas of the generic xyzz_dbl), each level's products split between
// the pairs of the group, as xyzz_dbl<Fq29Pair>.
template <>
ZKG_DEV Xyzz<Fq2Q> xyzz_dbl<Fq2Q>(const Xyzz<Fq2Q>& p) {
    const bool odd = quad_odd();
    const Fq2P x = p.x.v, y = p.y.v;
    const Fq2P u = y + y;
    Fq2P v, x2, s, m2, w, msx, wy, zzz3;
    const Fq2P t0 = fe_select(odd, x, u);
    quad_swap(FeTraits<Fq2P>::sqr(t0), odd, &v, &x2);  // V = U^2 | X^2
    const Fq2P mm = (x2 + x2) + x2;
    quad_swap(fe_select(odd, mm, x) * fe_select(odd, mm, v), odd, &s, &m2);  // S = X V | M^2
    const Fq2P x3 = m2 - (s + s);
    const Fq2P zz3 = v * p.zz.v;
    quad_swap(fe_select(odd, mm, u) * fe_select(odd, s - x3, v), odd, &w, &msx);  // W = U V | M (S - X3)
    quad_swap(w * fe_select(odd, p.zzz.v, y), odd, &wy, &zzz3);                    // W Y | W ZZZ
    return Xyzz<Fq2Q>{{x3}, {msx - wy}, {zz3}, {zzz3}};
}

// P + Q (add-2008-s, the formulas of the generic xyzz_add, every special case), products split
// between the pairs of the group, as xyzz_add<Fq29Pair>.
template <>
ZKG_DEV Xyzz<Fq2Q> xyzz_add<Fq2Q>(const Xyzz<Fq2Q>& p, const Xyzz<Fq2Q>& q) {
    typedef FeTraits<Fq2P> T;
    if (xyzz_is_inf(q)) return p;
    if (xyzz_is_inf(p)) return q;
    const bool odd = quad_odd();
    Fq2P u1, u2, s1, s2, zz12, zzz12, ppsq, r2, ppp, qq, ye, yo, zz3, zzz3;
    quad_swap(fe_select(odd, q.x.v, p.x.v) * fe_select(odd, p.zz.v, q.zz.v), odd, &u1, &u2);
    quad_swap(fe_select(odd, q.y.v, p.y.v) * fe_select(odd, p.zzz.v, q.zzz.v), odd, &s1, &s2);
    quad_swap(fe_select(odd, p.zzz.v, p.zz.v) * fe_select(odd, q.zzz.v, q.zz.v), odd, &zz12, &zzz12);
    const Fq2P pp = u2 - u1, r = s2 - s1;
    if (T::is_zero(pp)) {
        if (T::is_zero(r)) return xyzz_dbl(p);
        return xyzz_inf<Fq2Q>();
    }
    quad_swap(T::sqr(fe_select(odd, r, pp)), odd, &ppsq, &r2);           // PP = P^2 | R^2
    quad_swap(fe_select(odd, u1, pp) * ppsq, odd, &ppp, &qq);            // PPP = P PP | Q = U1 PP
    const Fq2P x3 = (r2 - ppp) - (qq + qq);
    quad_swap(fe_select(odd, s1, r) * fe_select(odd, ppp, qq - x3), odd, &ye, &yo);  // R (Q - X3) | S1 PPP
    quad_swap(fe_select(odd, zzz12, zz12) * fe_select(odd, ppp, ppsq), odd, &zz3, &zzz3);
    return Xyzz<Fq2Q>{{x3}, {ye - yo}, {zz3}, {zzz3}};
}

// Lanes that hold one field element: 4 for Fq2Q, 2 for Fq2P and Fq29Pair, else 1.
template <class Fe>
struct FeLanes {
    static constexpr uint32_t value = 1;
};
template <>
struct FeLanes<Fq2P> {
    static constexpr uint32_t value = 2;
};
template <>
struct FeLanes<Fq29Pair> {
    static constexpr uint32_t value = 2;
};
template <>
struct FeLanes<Fq2Q> {
    static constexpr uint32_t value = 4;
};

// The storage form of the Fq29 / Fq2_29 stages, from canonical 32-bit coordinates.
ZKG_DEV Fq29 to_storage_field(const Fq& x, const Fq29*) { return fq29_from_fq(x); }
ZKG_DEV Fq2_29 to_storage_field(const Fq2& x, const Fq2_29*) { return Fq2_29{fq29_from_fq(x.a), fq29_from_fq(x.b)}; }

}  // namespace zkg
