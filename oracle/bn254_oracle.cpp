// CPU oracle for the BN254 MSM and Fr NTT — TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU and with independent formulas, what the reference's GPU path
// computes. Only tests/, the smoke check and the benchmark's CPU-baseline leg may load it; the
// product library never links or calls it.
//
// What it follows in the reference (zerotruss/zk-gpu):
//   moduli p, r                       src/field.rs:17-20
//   8 little-endian u32 words/element src/host.rs:104-110, README.md:77-78
//   root of unity 5^((r-1)>>log_n)    src/host.rs:93-98
//   forward = plain DFT, natural order; inverse uses w^-1 and 1/n   src/ntt.rs:26-50
//   MSM = sum s_i * P_i, infinity = (0,0) words                      src/msm.rs:284-298, 345-354
//   G2 over Fq2 = Fq[u]/(u^2+1)                                      src/group.rs:30-55
//   affine result of an MSM, infinity as zeros                       tests/golden.rs:44-49
//
// Independence from the GPU code: the oracle uses 4x64-bit limbs, Jacobian coordinates
// (dbl-2009-l, add-2007-bl, madd-2007-bl) and unsigned-window Pippenger, while the GPU path uses
// 8x32-bit limbs, XYZZ coordinates and signed digits. A formula slip on one side is not
// repeated on the other.
//
// Pinning (see tests/test_oracle.py): ffjavascript golden vectors (ntt.json, msm.json), Python
// big-integer textbook arithmetic on random/boundary values, the O(n^2) DFT, Horner evaluation,
// the G2 generator's curve equation and order, and the closed form sum s_i*(k_i*G) = (sum k_i s_i)*G.

#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>

typedef unsigned __int128 u128;

namespace {

struct Field {
    uint64_t p[4];
    uint64_t inv;     // -p^-1 mod 2^64
    uint64_t r2[4];   // 2^512 mod p
    uint64_t one[4];  // 2^256 mod p (Montgomery one)
};

// p = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47 (src/field.rs:18)
// r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001 (src/field.rs:20)
Field make_field(const uint64_t m[4]);

const uint64_t P_WORDS[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL,
                             0x30644e72e131a029ULL};
const uint64_t R_WORDS[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                             0x30644e72e131a029ULL};

inline bool geq(const uint64_t a[4], const uint64_t b[4]) {
    for (int i = 3; i >= 0; --i) {
        if (a[i] != b[i]) return a[i] > b[i];
    }
    return true;
}

inline uint64_t sub4(uint64_t out[4], const uint64_t a[4], const uint64_t b[4]) {
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        u128 d = (u128)a[i] - b[i] - borrow;
        out[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
    return borrow;
}

inline uint64_t add4(uint64_t out[4], const uint64_t a[4], const uint64_t b[4]) {
    uint64_t carry = 0;
    for (int i = 0; i < 4; ++i) {
        u128 s = (u128)a[i] + b[i] + carry;
        out[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
    return carry;
}

struct Fe {
    uint64_t l[4];
};

inline void f_add(Fe& o, const Fe& a, const Fe& b, const Field& F) {
    uint64_t t[4];
    uint64_t c = add4(t, a.l, b.l);
    if (c || geq(t, F.p)) sub4(t, t, F.p);
    std::memcpy(o.l, t, sizeof t);
}

inline void f_sub(Fe& o, const Fe& a, const Fe& b, const Field& F) {
    uint64_t t[4];
    uint64_t br = sub4(t, a.l, b.l);
    if (br) add4(t, t, F.p);
    std::memcpy(o.l, t, sizeof t);
}

// Montgomery product a*b*2^-256 mod p, CIOS on 4x64-bit limbs.
inline void f_mul(Fe& o, const Fe& a, const Fe& b, const Field& F) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < 4; ++j) {
            u128 s = (u128)a.l[j] * b.l[i] + t[j] + carry;
            t[j] = (uint64_t)s;
            carry = (uint64_t)(s >> 64);
        }
        u128 s = (u128)t[4] + carry;
        t[4] = (uint64_t)s;
        t[5] = (uint64_t)(s >> 64);
        uint64_t m = t[0] * F.inv;
        s = (u128)m * F.p[0] + t[0];
        carry = (uint64_t)(s >> 64);
        for (int j = 1; j < 4; ++j) {
            s = (u128)m * F.p[j] + t[j] + carry;
            t[j - 1] = (uint64_t)s;
            carry = (uint64_t)(s >> 64);
        }
        s = (u128)t[4] + carry;
        t[3] = (uint64_t)s;
        t[4] = t[5] + (uint64_t)(s >> 64);
    }
    uint64_t r[4] = {t[0], t[1], t[2], t[3]};
    if (t[4] || geq(r, F.p)) sub4(r, r, F.p);
    std::memcpy(o.l, r, sizeof r);
}

inline bool f_is_zero(const Fe& a) { return (a.l[0] | a.l[1] | a.l[2] | a.l[3]) == 0; }
inline bool f_eq(const Fe& a, const Fe& b) { return std::memcmp(a.l, b.l, sizeof a.l) == 0; }

Field make_field(const uint64_t m[4]) {
    Field F;
    std::memcpy(F.p, m, sizeof F.p);
    // inv = -m^-1 mod 2^64 by Newton iteration.
    uint64_t x = 1;
    for (int i = 0; i < 7; ++i) x *= 2 - m[0] * x;
    F.inv = (uint64_t)0 - x;
    // one = 2^256 mod m, r2 = 2^512 mod m, by doubling (m < 2^255, so doubling never overflows).
    uint64_t acc[4] = {1, 0, 0, 0};
    for (int i = 0; i < 512; ++i) {
        uint64_t t[4];
        add4(t, acc, acc);
        if (geq(t, m)) sub4(t, t, m);
        std::memcpy(acc, t, sizeof t);
        if (i == 255) std::memcpy(F.one, acc, sizeof acc);
    }
    std::memcpy(F.r2, acc, sizeof acc);
    return F;
}

const Field& FQ() {
    static const Field f = make_field(P_WORDS);
    return f;
}
const Field& FR() {
    static const Field f = make_field(R_WORDS);
    return f;
}

// Plain 8-word LE integer -> Montgomery form (value must be < modulus).
inline Fe f_from_words(const uint32_t* w, const Field& F) {
    Fe plain, r2, o;
    for (int i = 0; i < 4; ++i) plain.l[i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
    std::memcpy(r2.l, F.r2, sizeof r2.l);
    f_mul(o, plain, r2, F);
    return o;
}

inline void f_to_words(uint32_t* w, const Fe& a, const Field& F) {
    Fe one_plain = {{1, 0, 0, 0}}, o;
    f_mul(o, a, one_plain, F);
    for (int i = 0; i < 4; ++i) {
        w[2 * i] = (uint32_t)o.l[i];
        w[2 * i + 1] = (uint32_t)(o.l[i] >> 32);
    }
}

inline Fe f_one(const Field& F) {
    Fe o;
    std::memcpy(o.l, F.one, sizeof o.l);
    return o;
}

Fe f_pow(const Fe& base, const uint64_t e[4], const Field& F) {
    Fe acc = f_one(F);
    for (int bit = 255; bit >= 0; --bit) {
        f_mul(acc, acc, acc, F);
        if ((e[bit / 64] >> (bit % 64)) & 1) f_mul(acc, acc, base, F);
    }
    return acc;
}

Fe f_inv(const Fe& a, const Field& F) {  // Fermat: a^(m-2)
    uint64_t e[4], two[4] = {2, 0, 0, 0};
    sub4(e, F.p, two);
    return f_pow(a, e, F);
}

// ---------------------------------------------------------------- Fq2 = Fq[u]/(u^2+1)
struct Fe2 {
    Fe a, b;
};

struct Fq1Ops {
    typedef Fe T;
    static void add(T& o, const T& x, const T& y) { f_add(o, x, y, FQ()); }
    static void sub(T& o, const T& x, const T& y) { f_sub(o, x, y, FQ()); }
    static void mul(T& o, const T& x, const T& y) { f_mul(o, x, y, FQ()); }
    static bool is_zero(const T& x) { return f_is_zero(x); }
    static T zero() { return Fe{{0, 0, 0, 0}}; }
    static T one() { return f_one(FQ()); }
    static T inv(const T& x) { return f_inv(x, FQ()); }
    static const int WORDS = 8;
    static T load(const uint32_t* w) { return f_from_words(w, FQ()); }
    static void store(uint32_t* w, const T& x) { f_to_words(w, x, FQ()); }
};

struct Fq2Ops {
    typedef Fe2 T;
    static void add(T& o, const T& x, const T& y) {
        f_add(o.a, x.a, y.a, FQ());
        f_add(o.b, x.b, y.b, FQ());
    }
    static void sub(T& o, const T& x, const T& y) {
        f_sub(o.a, x.a, y.a, FQ());
        f_sub(o.b, x.b, y.b, FQ());
    }
    // Schoolbook (4 multiplications), deliberately not Karatsuba, so it shares no shortcut with
    // the GPU's Fq2 code: (a+bu)(c+du) = (ac-bd) + (ad+bc)u.
    static void mul(T& o, const T& x, const T& y) {
        Fe ac, bd, ad, bc;
        f_mul(ac, x.a, y.a, FQ());
        f_mul(bd, x.b, y.b, FQ());
        f_mul(ad, x.a, y.b, FQ());
        f_mul(bc, x.b, y.a, FQ());
        f_sub(o.a, ac, bd, FQ());
        f_add(o.b, ad, bc, FQ());
    }
    static bool is_zero(const T& x) { return f_is_zero(x.a) && f_is_zero(x.b); }
    static T zero() { return Fe2{Fe{{0, 0, 0, 0}}, Fe{{0, 0, 0, 0}}}; }
    static T one() { return Fe2{f_one(FQ()), Fe{{0, 0, 0, 0}}}; }
    // 1/(a+bu) = (a-bu)/(a^2+b^2)
    static T inv(const T& x) {
        Fe aa, bb, nrm, ni;
        f_mul(aa, x.a, x.a, FQ());
        f_mul(bb, x.b, x.b, FQ());
        f_add(nrm, aa, bb, FQ());
        ni = f_inv(nrm, FQ());
        T o;
        f_mul(o.a, x.a, ni, FQ());
        Fe t;
        f_mul(t, x.b, ni, FQ());
        f_sub(o.b, Fe{{0, 0, 0, 0}}, t, FQ());
        return o;
    }
    static const int WORDS = 16;
    static T load(const uint32_t* w) { return Fe2{f_from_words(w, FQ()), f_from_words(w + 8, FQ())}; }
    static void store(uint32_t* w, const T& x) {
        f_to_words(w, x.a, FQ());
        f_to_words(w + 8, x.b, FQ());
    }
};

// ---------------------------------------------------------------- Jacobian points, a = 0
template <class O>
struct Jac {
    typedef typename O::T T;
    T x, y, z;  // (X/Z^2, Y/Z^3); infinity: Z = 0
};

template <class O>
struct Aff {
    typedef typename O::T T;
    T x, y;
    bool inf;
};

template <class O>
Jac<O> jac_inf() {
    return Jac<O>{O::one(), O::one(), O::zero()};
}

// dbl-2009-l (a = 0): A=X^2 B=Y^2 C=B^2 D=2((X+B)^2-A-C) E=3A F=E^2 X3=F-2D Y3=E(D-X3)-8C Z3=2YZ
template <class O>
Jac<O> jac_dbl(const Jac<O>& p) {
    typedef typename O::T T;
    if (O::is_zero(p.z)) return p;
    T A, B, C, D, E, F, t, X3, Y3, Z3;
    O::mul(A, p.x, p.x);
    O::mul(B, p.y, p.y);
    O::mul(C, B, B);
    O::add(t, p.x, B);
    O::mul(t, t, t);
    O::sub(t, t, A);
    O::sub(t, t, C);
    O::add(D, t, t);
    O::add(E, A, A);
    O::add(E, E, A);
    O::mul(F, E, E);
    O::sub(X3, F, D);
    O::sub(X3, X3, D);
    O::sub(t, D, X3);
    O::mul(Y3, E, t);
    T c8;
    O::add(c8, C, C);
    O::add(c8, c8, c8);
    O::add(c8, c8, c8);
    O::sub(Y3, Y3, c8);
    O::mul(Z3, p.y, p.z);
    O::add(Z3, Z3, Z3);
    return Jac<O>{X3, Y3, Z3};
}

// add-2007-bl with all special cases.
template <class O>
Jac<O> jac_add(const Jac<O>& p, const Jac<O>& q) {
    typedef typename O::T T;
    if (O::is_zero(p.z)) return q;
    if (O::is_zero(q.z)) return p;
    T Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, r, V, t, X3, Y3, Z3;
    O::mul(Z1Z1, p.z, p.z);
    O::mul(Z2Z2, q.z, q.z);
    O::mul(U1, p.x, Z2Z2);
    O::mul(U2, q.x, Z1Z1);
    O::mul(t, q.z, Z2Z2);
    O::mul(S1, p.y, t);
    O::mul(t, p.z, Z1Z1);
    O::mul(S2, q.y, t);
    O::sub(H, U2, U1);
    O::sub(r, S2, S1);
    if (O::is_zero(H)) {
        if (O::is_zero(r)) return jac_dbl(p);
        return jac_inf<O>();
    }
    O::add(I, H, H);
    O::mul(I, I, I);
    O::mul(J, H, I);
    O::add(r, r, r);
    O::mul(V, U1, I);
    O::mul(X3, r, r);
    O::sub(X3, X3, J);
    O::sub(X3, X3, V);
    O::sub(X3, X3, V);
    O::sub(t, V, X3);
    O::mul(Y3, r, t);
    O::mul(t, S1, J);
    O::add(t, t, t);
    O::sub(Y3, Y3, t);
    O::add(t, p.z, q.z);
    O::mul(t, t, t);
    O::sub(t, t, Z1Z1);
    O::sub(t, t, Z2Z2);
    O::mul(Z3, t, H);
    return Jac<O>{X3, Y3, Z3};
}

// madd-2007-bl (Q affine) with all special cases.
template <class O>
Jac<O> jac_madd(const Jac<O>& p, const Aff<O>& q) {
    typedef typename O::T T;
    if (q.inf) return p;
    if (O::is_zero(p.z)) return Jac<O>{q.x, q.y, O::one()};
    T Z1Z1, U2, S2, H, HH, I, J, r, V, t, X3, Y3, Z3;
    O::mul(Z1Z1, p.z, p.z);
    O::mul(U2, q.x, Z1Z1);
    O::mul(t, p.z, Z1Z1);
    O::mul(S2, q.y, t);
    O::sub(H, U2, p.x);
    O::sub(r, S2, p.y);
    if (O::is_zero(H)) {
        if (O::is_zero(r)) return jac_dbl(p);
        return jac_inf<O>();
    }
    O::mul(HH, H, H);
    O::add(I, HH, HH);
    O::add(I, I, I);
    O::mul(J, H, I);
    O::add(r, r, r);
    O::mul(V, p.x, I);
    O::mul(X3, r, r);
    O::sub(X3, X3, J);
    O::sub(X3, X3, V);
    O::sub(X3, X3, V);
    O::sub(t, V, X3);
    O::mul(Y3, r, t);
    O::mul(t, p.y, J);
    O::add(t, t, t);
    O::sub(Y3, Y3, t);
    O::add(t, p.z, H);
    O::mul(t, t, t);
    O::sub(t, t, Z1Z1);
    O::sub(Z3, t, HH);
    return Jac<O>{X3, Y3, Z3};
}

template <class O>
Aff<O> jac_to_aff(const Jac<O>& p) {
    typedef typename O::T T;
    if (O::is_zero(p.z)) return Aff<O>{O::zero(), O::zero(), true};
    T zi = O::inv(p.z), zi2, zi3, x, y;
    O::mul(zi2, zi, zi);
    O::mul(zi3, zi2, zi);
    O::mul(x, p.x, zi2);
    O::mul(y, p.y, zi3);
    return Aff<O>{x, y, false};
}

template <class O>
Aff<O> load_aff(const uint32_t* w) {
    bool inf = true;
    for (int i = 0; i < 2 * O::WORDS; ++i) inf = inf && (w[i] == 0);
    if (inf) return Aff<O>{O::zero(), O::zero(), true};
    return Aff<O>{O::load(w), O::load(w + O::WORDS), false};
}

template <class O>
void store_aff(uint32_t* w, const Aff<O>& a) {
    if (a.inf) {
        std::memset(w, 0, sizeof(uint32_t) * 2 * O::WORDS);
        return;
    }
    O::store(w, a.x);
    O::store(w + O::WORDS, a.y);
}

inline int scalar_bit(const uint32_t* s, int bit) { return (s[bit >> 5] >> (bit & 31)) & 1; }

template <class O>
Jac<O> scalar_mul(const Aff<O>& p, const uint32_t* s) {
    Jac<O> acc = jac_inf<O>();
    for (int bit = 255; bit >= 0; --bit) {
        acc = jac_dbl(acc);
        if (scalar_bit(s, bit)) acc = jac_madd(acc, p);
    }
    return acc;
}

inline uint32_t window_value(const uint32_t* s, int pos, int c) {
    uint32_t v = 0;
    for (int k = 0; k < c && pos + k < 256; ++k) v |= (uint32_t)scalar_bit(s, pos + k) << k;
    return v;
}

// Unsigned-window Pippenger over points [lo, hi) with Jacobian buckets.
template <class O>
Jac<O> pippenger_range(const uint32_t* bases, const uint32_t* scalars, size_t lo, size_t hi, int c) {
    const int windows = (256 + c - 1) / c;
    const size_t nb = ((size_t)1 << c) - 1;
    std::vector<Aff<O>> pts(hi - lo);
    for (size_t i = lo; i < hi; ++i) pts[i - lo] = load_aff<O>(bases + i * 2 * O::WORDS);
    std::vector<Jac<O>> buckets(nb);
    Jac<O> total = jac_inf<O>();
    for (int w = windows - 1; w >= 0; --w) {
        for (int k = 0; k < c; ++k) total = jac_dbl(total);
        std::fill(buckets.begin(), buckets.end(), jac_inf<O>());
        for (size_t i = lo; i < hi; ++i) {
            uint32_t v = window_value(scalars + 8 * i, w * c, c);
            if (v) buckets[v - 1] = jac_madd(buckets[v - 1], pts[i - lo]);
        }
        Jac<O> run = jac_inf<O>(), sum = jac_inf<O>();
        for (size_t b = nb; b-- > 0;) {
            run = jac_add(run, buckets[b]);
            sum = jac_add(sum, run);
        }
        total = jac_add(total, sum);
    }
    return total;
}

template <class O>
void msm(const uint32_t* bases, const uint32_t* scalars, size_t n, uint32_t* out, int threads) {
    if (threads < 1) threads = 1;
    if ((size_t)threads > n) threads = (int)std::max<size_t>(n, 1);
    int c = 4;
    while (c < 16 && ((size_t)1 << (c + 3)) < n / (size_t)threads) ++c;
    std::vector<Jac<O>> part(threads, jac_inf<O>());
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        size_t lo = n * t / threads, hi = n * (t + 1) / threads;
        pool.emplace_back([&, t, lo, hi] { part[t] = pippenger_range<O>(bases, scalars, lo, hi, c); });
    }
    for (auto& th : pool) th.join();
    Jac<O> acc = jac_inf<O>();
    for (auto& p : part) acc = jac_add(acc, p);
    store_aff<O>(out, jac_to_aff(acc));
}

template <class O>
void msm_naive(const uint32_t* bases, const uint32_t* scalars, size_t n, uint32_t* out) {
    Jac<O> acc = jac_inf<O>();
    for (size_t i = 0; i < n; ++i) {
        Aff<O> p = load_aff<O>(bases + i * 2 * O::WORDS);
        acc = jac_add(acc, scalar_mul(p, scalars + 8 * i));
    }
    store_aff<O>(out, jac_to_aff(acc));
}

// Fixed-base multiples k_i * P with Montgomery batch normalisation; used to build test inputs.
template <class O>
void multiples(const uint32_t* point, const uint32_t* ks, size_t n, uint32_t* out, int threads) {
    Aff<O> p = load_aff<O>(point);
    // Comb table: T[j][v] = v * 2^(8j) * P for 32 byte positions.
    std::vector<Aff<O>> table(32 * 256);
    {
        Jac<O> base = p.inf ? jac_inf<O>() : Jac<O>{p.x, p.y, O::one()};
        for (int j = 0; j < 32; ++j) {
            Jac<O> acc = jac_inf<O>();
            for (int v = 0; v < 256; ++v) {
                table[j * 256 + v] = jac_to_aff(acc);
                acc = jac_add(acc, base);
            }
            for (int k = 0; k < 8; ++k) base = jac_dbl(base);
        }
    }
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        size_t lo = n * t / threads, hi = n * (t + 1) / threads;
        pool.emplace_back([&, lo, hi] {
            std::vector<Jac<O>> js(hi - lo);
            for (size_t i = lo; i < hi; ++i) {
                Jac<O> acc = jac_inf<O>();
                const uint32_t* k = ks + 8 * i;
                for (int j = 0; j < 32; ++j) {
                    uint32_t byte = (k[j / 4] >> (8 * (j % 4))) & 0xff;
                    if (byte) acc = jac_madd(acc, table[j * 256 + byte]);
                }
                js[i - lo] = acc;
            }
            // Batch inversion of the Z coordinates.
            typedef typename O::T T;
            std::vector<T> prefix(js.size());
            T run = O::one();
            for (size_t i = 0; i < js.size(); ++i) {
                prefix[i] = run;
                if (!O::is_zero(js[i].z)) O::mul(run, run, js[i].z);
            }
            T inv = O::inv(run);
            for (size_t i = js.size(); i-- > 0;) {
                uint32_t* o = out + (lo + i) * 2 * O::WORDS;
                if (O::is_zero(js[i].z)) {
                    std::memset(o, 0, sizeof(uint32_t) * 2 * O::WORDS);
                    continue;
                }
                T zi, zi2, zi3, x, y;
                O::mul(zi, inv, prefix[i]);
                O::mul(inv, inv, js[i].z);
                O::mul(zi2, zi, zi);
                O::mul(zi3, zi2, zi);
                O::mul(x, js[i].x, zi2);
                O::mul(y, js[i].y, zi3);
                O::store(o, x);
                O::store(o + O::WORDS, y);
            }
        });
    }
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------- NTT over Fr
inline uint64_t bitrev(uint64_t x, int bits) {
    uint64_t r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

Fe root_of_unity(int log_n) {  // 5^((r-1) >> log_n), src/host.rs:93-98
    const Field& F = FR();
    uint64_t e[4], one[4] = {1, 0, 0, 0};
    sub4(e, F.p, one);
    for (int k = 0; k < log_n; ++k) {  // shift right by one bit, log_n times
        for (int i = 0; i < 4; ++i) e[i] = (e[i] >> 1) | (i < 3 ? (e[i + 1] << 63) : 0);
    }
    uint32_t five[8] = {5, 0, 0, 0, 0, 0, 0, 0};
    return f_pow(f_from_words(five, F), e, F);
}

template <class Fn>
void parallel_for(size_t n, int threads, Fn fn) {
    if (threads <= 1 || n < 4096) {
        fn((size_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        size_t lo = n * t / threads, hi = n * (t + 1) / threads;
        pool.emplace_back([=] { fn(lo, hi); });
    }
    for (auto& th : pool) th.join();
}

// Iterative radix-2 decimation-in-time: bit-reverse permutation, then log_n butterfly stages.
// X_j = sum_i x_i w^(ij); inverse uses w^-1 and multiplies by n^-1.
void ntt(uint32_t* data, int log_n, bool inverse, int threads) {
    const Field& F = FR();
    const size_t n = (size_t)1 << log_n;
    std::vector<Fe> a(n);
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) a[bitrev(i, log_n)] = f_from_words(data + 8 * i, F);
    });
    Fe w = root_of_unity(log_n);
    if (inverse) w = f_inv(w, F);
    std::vector<Fe> tw(std::max<size_t>(n / 2, 1));
    tw[0] = f_one(F);
    for (size_t i = 1; i < n / 2; ++i) f_mul(tw[i], tw[i - 1], w, F);
    for (int s = 1; s <= log_n; ++s) {
        const size_t half = (size_t)1 << (s - 1);
        const size_t step = n >> s;  // twiddle stride
        parallel_for(n / 2, threads, [&](size_t lo, size_t hi) {
            for (size_t b = lo; b < hi; ++b) {
                size_t grp = b / half, j = b % half;
                size_t i0 = grp * 2 * half + j, i1 = i0 + half;
                Fe t, u = a[i0];
                f_mul(t, a[i1], tw[j * step], F);
                f_add(a[i0], u, t, F);
                f_sub(a[i1], u, t, F);
            }
        });
    }
    Fe scale = f_one(F);
    if (inverse) {
        uint32_t nw[8] = {0};
        nw[log_n / 32] = 1u << (log_n % 32);
        scale = f_inv(f_from_words(nw, F), F);
    }
    parallel_for(n, threads, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            Fe v = a[i];
            if (inverse) f_mul(v, v, scale, F);
            f_to_words(data + 8 * i, v, F);
        }
    });
}

const Field& pick(int which) { return which == 0 ? FQ() : FR(); }

}  // namespace

// ================================================================ exported C ABI (ctypes)
extern "C" {

// which: 0 = Fq (base field), 1 = Fr (scalar field). All values plain 8-word LE, < modulus.
void orc_field_mul(int which, const uint32_t* a, const uint32_t* b, uint32_t* out) {
    const Field& F = pick(which);
    Fe x = f_from_words(a, F), y = f_from_words(b, F), o;
    f_mul(o, x, y, F);
    f_to_words(out, o, F);
}
void orc_field_add(int which, const uint32_t* a, const uint32_t* b, uint32_t* out) {
    const Field& F = pick(which);
    Fe x = f_from_words(a, F), y = f_from_words(b, F), o;
    f_add(o, x, y, F);
    f_to_words(out, o, F);
}
void orc_field_sub(int which, const uint32_t* a, const uint32_t* b, uint32_t* out) {
    const Field& F = pick(which);
    Fe x = f_from_words(a, F), y = f_from_words(b, F), o;
    f_sub(o, x, y, F);
    f_to_words(out, o, F);
}
void orc_field_inv(int which, const uint32_t* a, uint32_t* out) {
    const Field& F = pick(which);
    f_to_words(out, f_inv(f_from_words(a, F), F), F);
}
void orc_field_pow(int which, const uint32_t* a, const uint32_t* e, uint32_t* out) {
    const Field& F = pick(which);
    uint64_t ew[4];
    for (int i = 0; i < 4; ++i) ew[i] = (uint64_t)e[2 * i] | ((uint64_t)e[2 * i + 1] << 32);
    f_to_words(out, f_pow(f_from_words(a, F), ew, F), F);
}
void orc_fq2_mul(const uint32_t* a, const uint32_t* b, uint32_t* out) {
    Fe2 x = Fq2Ops::load(a), y = Fq2Ops::load(b), o;
    Fq2Ops::mul(o, x, y);
    Fq2Ops::store(out, o);
}
void orc_fq2_inv(const uint32_t* a, uint32_t* out) { Fq2Ops::store(out, Fq2Ops::inv(Fq2Ops::load(a))); }

void orc_root_of_unity(int log_n, uint32_t* out) { f_to_words(out, root_of_unity(log_n), FR()); }

// In-place natural-order NTT of 2^log_n elements (8 words each).
void orc_ntt(uint32_t* data, int log_n, int inverse, int threads) { ntt(data, log_n, inverse != 0, threads); }

// Horner evaluation of sum coeffs[i] x^i over Fr.
void orc_poly_eval(const uint32_t* coeffs, size_t n, const uint32_t* x, uint32_t* out) {
    const Field& F = FR();
    Fe xv = f_from_words(x, F), acc = Fe{{0, 0, 0, 0}};
    for (size_t i = n; i-- > 0;) {
        f_mul(acc, acc, xv, F);
        f_add(acc, acc, f_from_words(coeffs + 8 * i, F), F);
    }
    f_to_words(out, acc, F);
}

// The same evaluation over `threads` threads: P(x) = sum_t x^lo_t P_t(x), each part P_t (the
// coefficients lo_t .. hi_t - 1) by Horner on its own thread.
void orc_poly_eval_threads(const uint32_t* coeffs, size_t n, const uint32_t* x, uint32_t* out, int threads) {
    const Field& F = FR();
    const int t = std::max(1, threads);
    const Fe xv = f_from_words(x, F);
    std::vector<Fe> part(t, Fe{{0, 0, 0, 0}});
    std::vector<std::thread> pool;
    for (int k = 0; k < t; ++k) {
        const size_t lo = n * k / t, hi = n * (k + 1) / t;
        pool.emplace_back([&, k, lo, hi] {
            Fe acc = Fe{{0, 0, 0, 0}};
            for (size_t i = hi; i-- > lo;) {
                f_mul(acc, acc, xv, F);
                f_add(acc, acc, f_from_words(coeffs + 8 * i, F), F);
            }
            const uint64_t e[4] = {(uint64_t)lo, 0, 0, 0};
            Fe shifted;
            f_mul(shifted, acc, f_pow(xv, e, F), F);
            part[k] = shifted;
        });
    }
    for (auto& th : pool) th.join();
    Fe acc = Fe{{0, 0, 0, 0}};
    for (const Fe& v : part) f_add(acc, acc, v, F);
    f_to_words(out, acc, F);
}

// sum a_i b_i mod r (plain words in, plain words out; every a_i, b_i < r). With bases k_i G this is
// the scalar of the closed form sum s_i (k_i G) = (sum k_i s_i) G.
void orc_dot_fr(const uint32_t* a, const uint32_t* b, size_t n, uint32_t* out, int threads) {
    const Field& F = FR();
    const int t = std::max(1, threads);
    std::vector<Fe> part(t, Fe{{0, 0, 0, 0}});
    std::vector<std::thread> pool;
    for (int k = 0; k < t; ++k) {
        const size_t lo = n * k / t, hi = n * (k + 1) / t;
        pool.emplace_back([&, k, lo, hi] {
            Fe acc = Fe{{0, 0, 0, 0}};
            for (size_t i = lo; i < hi; ++i) {
                Fe y, p;
                for (int j = 0; j < 4; ++j) y.l[j] = (uint64_t)b[8 * i + 2 * j] | ((uint64_t)b[8 * i + 2 * j + 1] << 32);
                f_mul(p, f_from_words(a + 8 * i, F), y, F);  // (a R) * b / R = a b, plain
                f_add(acc, acc, p, F);
            }
            part[k] = acc;
        });
    }
    for (auto& th : pool) th.join();
    Fe acc = Fe{{0, 0, 0, 0}};
    for (const Fe& x : part) f_add(acc, acc, x, F);
    for (int j = 0; j < 4; ++j) {
        out[2 * j] = (uint32_t)acc.l[j];
        out[2 * j + 1] = (uint32_t)(acc.l[j] >> 32);
    }
}

// Points: G1 affine 16 words (x, y), G2 affine 32 words (x.a, x.b, y.a, y.b); infinity = zeros.
// group: 1 = G1, 2 = G2. Scalars: 8 words each, any 256-bit value.
void orc_msm(int group, const uint32_t* bases, const uint32_t* scalars, size_t n, uint32_t* out, int threads) {
    if (group == 1) msm<Fq1Ops>(bases, scalars, n, out, threads);
    else msm<Fq2Ops>(bases, scalars, n, out, threads);
}
void orc_msm_naive(int group, const uint32_t* bases, const uint32_t* scalars, size_t n, uint32_t* out) {
    if (group == 1) msm_naive<Fq1Ops>(bases, scalars, n, out);
    else msm_naive<Fq2Ops>(bases, scalars, n, out);
}
// out[i] = ks[i] * point (affine, batch-normalised).
void orc_multiples(int group, const uint32_t* point, const uint32_t* ks, size_t n, uint32_t* out, int threads) {
    if (group == 1) multiples<Fq1Ops>(point, ks, n, out, threads);
    else multiples<Fq2Ops>(point, ks, n, out, threads);
}
// out = a + b (affine, infinity = zeros).
void orc_point_add(int group, const uint32_t* a, const uint32_t* b, uint32_t* out) {
    if (group == 1) {
        Aff<Fq1Ops> p = load_aff<Fq1Ops>(a), q = load_aff<Fq1Ops>(b);
        Jac<Fq1Ops> jp = p.inf ? jac_inf<Fq1Ops>() : Jac<Fq1Ops>{p.x, p.y, Fq1Ops::one()};
        store_aff<Fq1Ops>(out, jac_to_aff(jac_madd(jp, q)));
    } else {
        Aff<Fq2Ops> p = load_aff<Fq2Ops>(a), q = load_aff<Fq2Ops>(b);
        Jac<Fq2Ops> jp = p.inf ? jac_inf<Fq2Ops>() : Jac<Fq2Ops>{p.x, p.y, Fq2Ops::one()};
        store_aff<Fq2Ops>(out, jac_to_aff(jac_madd(jp, q)));
    }
}
// out = s * a
void orc_point_mul(int group, const uint32_t* a, const uint32_t* s, uint32_t* out) {
    if (group == 1) store_aff<Fq1Ops>(out, jac_to_aff(scalar_mul(load_aff<Fq1Ops>(a), s)));
    else store_aff<Fq2Ops>(out, jac_to_aff(scalar_mul(load_aff<Fq2Ops>(a), s)));
}

}  // extern "C"
