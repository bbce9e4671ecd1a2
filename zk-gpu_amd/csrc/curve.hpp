// Fq2 and the XYZZ point law for BN254 G1 (coordinates in Fq) and G2 (coordinates in Fq2).
//
// Semantics restated from the reference's group layer: Fq2 = Fq[u]/(u^2 + 1) with Karatsuba
// multiplication (3 Fq products) and the (a+b)(a-b), 2ab squaring (2 products)
// (src/group.rs:30-55); XYZZ coordinates x = X/ZZ, y = Y/ZZZ with infinity <=> ZZ = 0 and affine
// infinity (0, 0) (src/shaders/curve.wgsl:1-20). The formulas are the public EFD ones for
// short-Weierstrass curves with a = 0: dbl-2008-s-1 (6M + 3S), madd-2008-s (8M + 2S) and
// add-2008-s (12M + 2S), each with every special case (curve.wgsl:22-75).
//
// Everything is templated over the coordinate field Fe, so one kernel source serves G1 and G2.
#pragma once
#include "field.hpp"

namespace zkg {

struct Fq2 {
    Fq a, b;  // a + b*u
};

ZKG_DEV Fq2 operator+(const Fq2& x, const Fq2& y) { return Fq2{x.a + y.a, x.b + y.b}; }
ZKG_DEV Fq2 operator-(const Fq2& x, const Fq2& y) { return Fq2{x.a - y.a, x.b - y.b}; }
ZKG_DEV Fq2 operator*(const Fq2& x, const Fq2& y) {
    const Fq ac = x.a * y.a, bd = x.b * y.b;
    const Fq cross = (x.a + x.b) * (y.a + y.b);
    return Fq2{ac - bd, (cross - ac) - bd};
}

// Field-generic helpers ---------------------------------------------------------------
template <class Fe>
struct FeTraits;

template <>
struct FeTraits<Fq> {
    static constexpr int WORDS = 8;
    static ZKG_DEV Fq zero() { return fp_zero<FqParams>(); }
    static ZKG_DEV Fq one() { return fp_one<FqParams>(); }
    static ZKG_DEV bool is_zero(const Fq& x) { return fp_is_zero(x); }
    static ZKG_DEV Fq sqr(const Fq& x) { return fp_sqr(x); }
    static ZKG_DEV Fq load(const uint32_t* p) { return fp_load<FqParams>(p); }
    static ZKG_DEV void store(uint32_t* p, const Fq& x) { fp_store(p, x); }
    static ZKG_DEV void store_result(uint32_t* p, const Fq& x) { fp_store(p, x); }
    static ZKG_DEV Fq to_mont(const Fq& x) { return fp_to_mont(x); }
    static ZKG_DEV Fq from_mont(const Fq& x) { return fp_from_mont(x); }
};

template <>
struct FeTraits<Fq2> {
    static constexpr int WORDS = 16;
    static ZKG_DEV Fq2 zero() { return Fq2{fp_zero<FqParams>(), fp_zero<FqParams>()}; }
    static ZKG_DEV Fq2 one() { return Fq2{fp_one<FqParams>(), fp_zero<FqParams>()}; }
    static ZKG_DEV bool is_zero(const Fq2& x) { return fp_is_zero(x.a) && fp_is_zero(x.b); }
    static ZKG_DEV Fq2 sqr(const Fq2& x) {
        const Fq ab = x.a * x.b;
        return Fq2{(x.a + x.b) * (x.a - x.b), ab + ab};
    }
    static ZKG_DEV Fq2 load(const uint32_t* p) { return Fq2{fp_load<FqParams>(p), fp_load<FqParams>(p + 8)}; }
    static ZKG_DEV void store(uint32_t* p, const Fq2& x) {
        fp_store(p, x.a);
        fp_store(p + 8, x.b);
    }
    static ZKG_DEV void store_result(uint32_t* p, const Fq2& x) { store(p, x); }
    static ZKG_DEV Fq2 to_mont(const Fq2& x) { return Fq2{fp_to_mont(x.a), fp_to_mont(x.b)}; }
    static ZKG_DEV Fq2 from_mont(const Fq2& x) { return Fq2{fp_from_mont(x.a), fp_from_mont(x.b)}; }
};

// Branch-free selects (v_cndmask per word), used to share one inlined point-operation body
// between several operand choices.
template <class P>
ZKG_DEV Fp<P> fe_select(bool c, const Fp<P>& a, const Fp<P>& b) {
    Fp<P> r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.w[i] = c ? a.w[i] : b.w[i];
    return r;
}

ZKG_DEV Fq2 fe_select(bool c, const Fq2& a, const Fq2& b) { return Fq2{fe_select(c, a.a, b.a), fe_select(c, a.b, b.b)}; }

template <class Fe>
struct Affine {
    Fe x, y;
};

template <class Fe>
struct Xyzz {
    Fe x, y, zz, zzz;
};

template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_inf() {
    const Fe z = FeTraits<Fe>::zero();
    return Xyzz<Fe>{z, z, z, z};
}

template <class Fe>
ZKG_DEV bool xyzz_is_inf(const Xyzz<Fe>& p) {
    return FeTraits<Fe>::is_zero(p.zz);
}

template <class Fe>
ZKG_DEV bool affine_is_inf(const Affine<Fe>& p) {
    return FeTraits<Fe>::is_zero(p.x) && FeTraits<Fe>::is_zero(p.y);
}

template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_from_affine(const Affine<Fe>& p) {
    if (affine_is_inf(p)) return xyzz_inf<Fe>();
    const Fe one = FeTraits<Fe>::one();
    return Xyzz<Fe>{p.x, p.y, one, one};
}

template <class Fe>
ZKG_DEV Affine<Fe> affine_neg(const Affine<Fe>& p) {
    return Affine<Fe>{p.x, FeTraits<Fe>::zero() - p.y};
}

template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_neg(const Xyzz<Fe>& p) {
    return Xyzz<Fe>{p.x, FeTraits<Fe>::zero() - p.y, p.zz, p.zzz};
}

// 2P, dbl-2008-s-1 (a = 0): U = 2Y, V = U^2, W = UV, S = XV, M = 3X^2,
// a b - c d for the point laws' Y3; a field type with a cheaper form (both products under one
// reduction) overloads it, found by argument-dependent lookup where the laws are instantiated
template <class Fe>
ZKG_DEV Fe fe_mul_sub(const Fe& a, const Fe& b, const Fe& c, const Fe& d) {
    return a * b - c * d;
}

// X3 = M^2 - 2S, Y3 = M(S - X3) - WY, ZZ3 = V ZZ, ZZZ3 = W ZZZ. Infinity maps to infinity.
template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_dbl(const Xyzz<Fe>& p) {
    typedef FeTraits<Fe> T;
    const Fe u = p.y + p.y;
    const Fe v = T::sqr(u);
    const Fe w = u * v;
    const Fe s = p.x * v;
    const Fe x2 = T::sqr(p.x);
    const Fe m = (x2 + x2) + x2;
    const Fe x3 = T::sqr(m) - (s + s);
    const Fe y3 = fe_mul_sub(m, s - x3, w, p.y);
    return Xyzz<Fe>{x3, y3, v * p.zz, w * p.zzz};
}

// P + Q with Q affine (not infinity), madd-2008-s, all special cases.
template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_madd_nonzero(const Xyzz<Fe>& p, const Affine<Fe>& q) {
    typedef FeTraits<Fe> T;
    if (xyzz_is_inf(p)) return Xyzz<Fe>{q.x, q.y, T::one(), T::one()};
    const Fe u2 = q.x * p.zz;
    const Fe s2 = q.y * p.zzz;
    const Fe pp = u2 - p.x;
    const Fe r = s2 - p.y;
    if (T::is_zero(pp)) {
        if (T::is_zero(r)) return xyzz_dbl(p);
        return xyzz_inf<Fe>();
    }
    const Fe ppsq = T::sqr(pp);
    const Fe ppp = pp * ppsq;
    const Fe qq = p.x * ppsq;
    const Fe x3 = (T::sqr(r) - ppp) - (qq + qq);
    const Fe y3 = fe_mul_sub(r, qq - x3, p.y, ppp);
    return Xyzz<Fe>{x3, y3, p.zz * ppsq, p.zzz * ppp};
}

template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_madd(const Xyzz<Fe>& p, const Affine<Fe>& q) {
    if (affine_is_inf(q)) return p;
    return xyzz_madd_nonzero(p, q);
}

// P + Q, both XYZZ, add-2008-s, all special cases.
template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_add(const Xyzz<Fe>& p, const Xyzz<Fe>& q) {
    typedef FeTraits<Fe> T;
    if (xyzz_is_inf(q)) return p;
    if (xyzz_is_inf(p)) return q;
    const Fe u1 = p.x * q.zz;
    const Fe u2 = q.x * p.zz;
    const Fe s1 = p.y * q.zzz;
    const Fe s2 = q.y * p.zzz;
    const Fe pp = u2 - u1;
    const Fe r = s2 - s1;
    if (T::is_zero(pp)) {
        if (T::is_zero(r)) return xyzz_dbl(p);
        return xyzz_inf<Fe>();
    }
    const Fe ppsq = T::sqr(pp);
    const Fe ppp = pp * ppsq;
    const Fe qq = u1 * ppsq;
    const Fe x3 = (T::sqr(r) - ppp) - (qq + qq);
    const Fe y3 = fe_mul_sub(r, qq - x3, s1, ppp);
    return Xyzz<Fe>{x3, y3, (p.zz * q.zz) * ppsq, (p.zzz * q.zzz) * ppp};
}

template <class Fe>
ZKG_DEV Xyzz<Fe> select_xyzz(bool c, const Xyzz<Fe>& a, const Xyzz<Fe>& b) {
    return Xyzz<Fe>{fe_select(c, a.x, b.x), fe_select(c, a.y, b.y), fe_select(c, a.zz, b.zz), fe_select(c, a.zzz, b.zzz)};
}

// Memory format: affine = x, y (WORDS each); XYZZ = x, y, zz, zzz.
template <class Fe>
ZKG_DEV Affine<Fe> load_affine(const uint32_t* p) {
    typedef FeTraits<Fe> T;
    return Affine<Fe>{T::load(p), T::load(p + T::WORDS)};
}

template <class Fe>
ZKG_DEV Xyzz<Fe> load_xyzz(const uint32_t* p) {
    typedef FeTraits<Fe> T;
    return Xyzz<Fe>{T::load(p), T::load(p + T::WORDS), T::load(p + 2 * T::WORDS), T::load(p + 3 * T::WORDS)};
}

template <class Fe>
ZKG_DEV void store_xyzz(uint32_t* p, const Xyzz<Fe>& v) {
    typedef FeTraits<Fe> T;
    T::store(p, v.x);
    T::store(p + T::WORDS, v.y);
    T::store(p + 2 * T::WORDS, v.zz);
    T::store(p + 3 * T::WORDS, v.zzz);
}

// The MSM result in the library's API form (coordinate types with another internal form
// convert here).
template <class Fe>
ZKG_DEV void store_xyzz_result(uint32_t* p, const Xyzz<Fe>& v) {
    typedef FeTraits<Fe> T;
    T::store_result(p, v.x);
    T::store_result(p + T::WORDS, v.y);
    T::store_result(p + 2 * T::WORDS, v.zz);
    T::store_result(p + 3 * T::WORDS, v.zzz);
}

}  // namespace zkg
