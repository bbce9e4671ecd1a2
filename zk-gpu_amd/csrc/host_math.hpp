// 256-bit modular arithmetic on the host, for plan constants (roots of unity, Montgomery
// scale factors, twiddle seeds) and for turning a device XYZZ result into affine coordinates.
//
// Restates the semantics of the reference's host module (src/host.rs:1-115: add_mod, sub_mod,
// mul_mod, pow_mod, root_of_unity with generator 5, Fermat inverse, LE u32 words). The reference
// multiplies bit-serially; here a Montgomery product on 4 x 64-bit limbs (R = 2^256) does it, which
// makes building a 2^12-entry twiddle seed table or an affine conversion take microseconds.
#pragma once
#include <stdint.h>
#include <string.h>

namespace zkg {
namespace host {

struct U256 {
    uint64_t l[4];
};

typedef unsigned __int128 u128;

inline bool geq(const U256& a, const U256& b) {
    for (int i = 3; i >= 0; --i)
        if (a.l[i] != b.l[i]) return a.l[i] > b.l[i];
    return true;
}

inline bool is_zero(const U256& a) { return (a.l[0] | a.l[1] | a.l[2] | a.l[3]) == 0; }

inline uint64_t sub_raw(U256& o, const U256& a, const U256& b) {
    uint64_t br = 0;
    for (int i = 0; i < 4; ++i) {
        u128 d = (u128)a.l[i] - b.l[i] - br;
        o.l[i] = (uint64_t)d;
        br = (uint64_t)(d >> 64) & 1;
    }
    return br;
}

inline uint64_t add_raw(U256& o, const U256& a, const U256& b) {
    uint64_t c = 0;
    for (int i = 0; i < 4; ++i) {
        u128 s = (u128)a.l[i] + b.l[i] + c;
        o.l[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    return c;
}

// A modulus below 2^255 with its Montgomery constants (R = 2^256).
struct Modulus {
    U256 m;
    uint64_t inv;  // -m^-1 mod 2^64
    U256 one;      // R mod m
    U256 r2;       // R^2 mod m

    explicit Modulus(const U256& mod) : m(mod) {
        uint64_t x = 1;
        for (int i = 0; i < 7; ++i) x *= 2 - m.l[0] * x;
        inv = 0 - x;
        U256 acc = {{1, 0, 0, 0}};
        for (int i = 0; i < 512; ++i) {
            U256 t;
            add_raw(t, acc, acc);
            if (geq(t, m)) sub_raw(t, t, m);
            acc = t;
            if (i == 255) one = acc;
        }
        r2 = acc;
    }

    U256 add(const U256& a, const U256& b) const {
        U256 t;
        uint64_t c = add_raw(t, a, b);
        if (c || geq(t, m)) sub_raw(t, t, m);
        return t;
    }
    U256 sub(const U256& a, const U256& b) const {
        U256 t;
        if (sub_raw(t, a, b)) add_raw(t, t, m);
        return t;
    }
    // Montgomery product a*b/R mod m
    U256 mont(const U256& a, const U256& b) const {
        uint64_t t[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 4; ++i) {
            uint64_t c = 0;
            for (int j = 0; j < 4; ++j) {
                u128 s = (u128)a.l[j] * b.l[i] + t[j] + c;
                t[j] = (uint64_t)s;
                c = (uint64_t)(s >> 64);
            }
            u128 s = (u128)t[4] + c;
            t[4] = (uint64_t)s;
            t[5] = (uint64_t)(s >> 64);
            uint64_t k = t[0] * inv;
            s = (u128)k * m.l[0] + t[0];
            c = (uint64_t)(s >> 64);
            for (int j = 1; j < 4; ++j) {
                s = (u128)k * m.l[j] + t[j] + c;
                t[j - 1] = (uint64_t)s;
                c = (uint64_t)(s >> 64);
            }
            s = (u128)t[4] + c;
            t[3] = (uint64_t)s;
            t[4] = t[5] + (uint64_t)(s >> 64);
        }
        U256 r = {{t[0], t[1], t[2], t[3]}};
        if (t[4] || geq(r, m)) sub_raw(r, r, m);
        return r;
    }
    U256 to_mont(const U256& a) const { return mont(a, r2); }
    U256 from_mont(const U256& a) const { return mont(a, U256{{1, 0, 0, 0}}); }
    // plain a*b mod m
    U256 mul(const U256& a, const U256& b) const { return mont(to_mont(a), b); }
    // plain base^e mod m
    U256 pow(const U256& base, const U256& e) const {
        U256 acc = one, b = to_mont(base);
        for (int bit = 255; bit >= 0; --bit) {
            acc = mont(acc, acc);
            if ((e.l[bit / 64] >> (bit % 64)) & 1) acc = mont(acc, b);
        }
        return from_mont(acc);
    }
    U256 inverse(const U256& a) const {  // Fermat: a^(m-2)
        U256 e;
        sub_raw(e, m, U256{{2, 0, 0, 0}});
        return pow(a, e);
    }
};

inline const Modulus& fq() {
    static const Modulus M(U256{{0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL,
                                 0x30644e72e131a029ULL}});
    return M;
}

inline const Modulus& fr() {
    static const Modulus M(U256{{0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                                 0x30644e72e131a029ULL}});
    return M;
}

inline U256 from_words32(const uint32_t* w) {
    U256 r;
    for (int i = 0; i < 4; ++i) r.l[i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
    return r;
}

inline void to_words32(uint32_t* w, const U256& a) {
    for (int i = 0; i < 4; ++i) {
        w[2 * i] = (uint32_t)a.l[i];
        w[2 * i + 1] = (uint32_t)(a.l[i] >> 32);
    }
}

// Primitive 2^log_n-th root of unity of Fr: 5^((r-1) >> log_n)  (src/host.rs:93-98).
inline U256 root_of_unity(unsigned log_n) {
    const Modulus& F = fr();
    U256 e;
    sub_raw(e, F.m, U256{{1, 0, 0, 0}});
    for (unsigned k = 0; k < log_n; ++k)
        for (int i = 0; i < 4; ++i) e.l[i] = (e.l[i] >> 1) | (i < 3 ? (e.l[i + 1] << 63) : 0);
    return F.pow(U256{{5, 0, 0, 0}}, e);
}

}  // namespace host
}  // namespace zkg
