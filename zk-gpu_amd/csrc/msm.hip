// Signed-digit Pippenger MSM kernels for gfx950 and the plan that sequences them.
//
// Stages (names follow the reference's profiling contract, src/msm.rs:217-263):
//  count      signed c-bit digits of every scalar (msm_digits.wgsl:19-37 rule: v = bits + carry,
//             v >= 2^(c-1) -> digit v - 2^c with carry 1); per-(window, bucket) point counts with
//             wavefront-aggregated atomics: lanes of a wave that hit the same slot (the common
//             case for repeated scalars such as witness 0/1 values) add with one atomic.
//  scan       exclusive scans of the counts and of the work items per slot, over all windows at
//             once, so the sorted array and the item list are compact.
//  scatter    counting-sort scatter of (point index | sign << 31) into slot order, again with
//             one atomic per group of equal slots in a wave.
//  accumulate one lane per work item (at most CH points of one bucket): a chain of mixed
//             additions of the gathered affine bases into an XYZZ accumulator.
//  merge r    heavy buckets (more than 16 items): items k = 0 mod 4s absorb items k + s, k + 2s,
//             k + 3s (s = 4^r), so a bucket holding most of the points never becomes a serial tail.
//  finalize   each bucket sums its at most 16 remaining items.
//  reduce     S_w = sum_b (b + 1) bucket[w][b]: L buckets per lane with running sums, plus b0 * run.
//  tree sum   per-window sum of the lane results in one workgroup (LDS tree).
//  combine    Horner's rule over the windows from the top, one wave with lane-parallel products.
// With GLV (the default), "count" first splits every scalar into two signed 127-bit halves and
// computes phi(P) = (beta x, y) for every base; the stages then run over 2n virtual points with
// ceil(128 / c) windows instead of ceil(255 / c). A plan over precomputed bases (the bases and
// their 2^64 multiples, msm_precompute_bases) also splits each half at bit 64: 4n virtual points,
// ceil(64 / c) windows.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>

#include "curve.hpp"
#include "field29.hpp"
#include "msm.hpp"
#include "msm_stages.hpp"

namespace zkg {

ZKG_MSM_EXTERN(Fq)
ZKG_MSM_EXTERN(Fq29)
ZKG_MSM_EXTERN(Fq2)
ZKG_MSM_EXTERN(Fq2_29)

uint32_t msm_default_window(size_t n) {
    size_t m = n < 1 ? 1 : n;
    int lg = 63 - __builtin_clzll((unsigned long long)m);
    int c = lg - 5;
    if (c < 11) c = 11;
    if (c > 16) c = 16;
    return (uint32_t)c;
}

uint32_t msm_auto_item_points(size_t n, uint32_t c) {
    // Work items hold about a quarter of the mean bucket (src/msm.rs:99-100).
    uint64_t mean = (uint64_t)(uint32_t)n >> (c - 1);
    if (mean < 1) mean = 1;
    uint64_t q = mean / 4, p = 1;
    while (p < q) p <<= 1;
    if (p < 8) p = 8;
    if (p > 1024) p = 1024;
    return (uint32_t)p;
}

// ---------------------------------------------------------------- small utility kernels

__global__ void fill_u32_kernel(uint32_t* p, uint32_t value, size_t count) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

static void fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, value, count);
}

void fill_words(uint32_t* p, uint32_t value, size_t count, hipStream_t s) {
    fill_u32(p, value, count, s);
    ZKG_HIP(hipGetLastError());
}

__global__ void to_mont_kernel(uint32_t* xy, size_t coords) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= coords) return;
    fp_store(xy + 8 * i, fp_to_mont(fp_load<FqParams>(xy + 8 * i)));
}

void msm_convert_bases(uint32_t* xy, size_t coords, hipStream_t stream) {
    if (!coords) return;
    hipLaunchKernelGGL(to_mont_kernel, dim3(ceil_div(coords, 256)), dim3(256), 0, stream, xy, coords);
    ZKG_HIP(hipGetLastError());
}

// Any 256-bit word -> the same residue below p, in place. The kernels assume canonical
// coordinates; Montgomery-2^256 input (snarkjs / ffjavascript) may hold values up to 2^256 - 1,
// which the reference's upload reduces by its scaling product (src/msm.rs:190-197). 2^256 < 6p,
// so five conditional subtractions reach [0, p).
__global__ void reduce_coords_kernel(uint32_t* xy, size_t coords) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= coords) return;
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = xy[8 * i + k];
#pragma unroll
    for (int rep = 0; rep < 5; ++rep) fp_reduce_once<FqParams>(w);
#pragma unroll
    for (int k = 0; k < 8; ++k) xy[8 * i + k] = w[k];
}

void msm_reduce_coords(uint32_t* xy, size_t coords, hipStream_t stream) {
    if (!coords) return;
    hipLaunchKernelGGL(reduce_coords_kernel, dim3(ceil_div(coords, 256)), dim3(256), 0, stream, xy, coords);
    ZKG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- GLV split
// lambda = 0xb3c4d79d41a917585bfc41088d8daaa78b17ea66b99c90dd is a cube root of unity mod r and
// phi(x, y) = (beta x, y) maps P to lambda P, with beta a cube root of unity mod p:
// beta = 0x59e26bcea0d48bacd4f263f1acdb5c4f5763473177fffffe on G1 and p - 1 minus that on G2.
// The lattice {(a, b) : a + b lambda = 0 mod r} has the reduced basis v1 = (a1, b1),
// v2 = (a2, b2) (extended Euclid on (r, lambda), stopped below sqrt(r); a1 b2 - a2 b1 = r):
//   a1 = b2 = 0x89d3256894d213e3, b1 = -0x6f4d8248eeb859fc8211bbeb7d4f1128,
//   a2 = 0x6f4d8248eeb859fd0be4e1541221250b.
// Babai rounding with c1 = floor(s G1c / 2^256) ~ s b2 / r and c2 = floor(s G2c / 2^256) ~
// -s b1 / r (G1c, G2c = those ratios times 2^256, rounded) gives (k1, k2) = (s, 0) - c1 v1 - c2 v2
// with k1 + k2 lambda = s (mod r); each c_i is within 1.125 of the exact ratio, so
// |k1|, |k2| <= 1.125 (a2 + a1) < 0.98 * 2^127 for every s < r.

struct GlvConst {
    static ZKG_HD constexpr uint32_t g1c(int i) { return i == 0 ? 0xc7e0b3d7u : i == 1 ? 0xd91d232eu : 0x2u; }
    static ZKG_HD constexpr uint32_t g2c(int i) {
        return i == 0 ? 0x391eb18eu : i == 1 ? 0x7a7bd9d4u : i == 2 ? 0xa773d2cfu : i == 3 ? 0x4ccef014u : 0x2u;
    }
    static ZKG_HD constexpr uint32_t a1(int i) { return i == 0 ? 0x94d213e3u : 0x89d32568u; }  // = b2
    static ZKG_HD constexpr uint32_t a2(int i) {
        return i == 0 ? 0x1221250bu : i == 1 ? 0x0be4e154u : i == 2 ? 0xeeb859fdu : 0x6f4d8248u;
    }
    static ZKG_HD constexpr uint32_t nb1(int i) {  // -b1
        return i == 0 ? 0x7d4f1128u : i == 1 ? 0x8211bbebu : i == 2 ? 0xeeb859fcu : 0x6f4d8248u;
    }
    // beta * 2^256 mod p (the bases are in Montgomery form)
    static ZKG_HD constexpr uint32_t beta_g1(int i) {
        return i == 0 ? 0xd782e155u : i == 1 ? 0x71930c11u : i == 2 ? 0xffbe3323u : i == 3 ? 0xa6bb947cu
             : i == 4 ? 0xd4741444u : i == 5 ? 0xaa303344u : i == 6 ? 0x26594943u : 0x2c3b3f0du;
    }
    static ZKG_HD constexpr uint32_t beta_g2(int i) {
        return i == 0 ? 0x13e80b9cu : i == 1 ? 0x3350c88eu : i == 2 ? 0xdb5e56b9u : i == 3 ? 0x7dce557cu
             : i == 4 ? 0xb615564au : i == 5 ? 0x6001b4b8u : i == 6 ? 0x020217e0u : 0x2682e617u;
    }
};

// out[0..NB) = floor(s * g / 2^256), s of 8 words, g of NB words
template <int NB, class G>
ZKG_DEV void glv_mul_high(const uint32_t s[8], G g, uint32_t out[NB]) {
    uint32_t prod[8 + NB];
#pragma unroll
    for (int k = 0; k < 8 + NB; ++k) prod[k] = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const uint64_t t = (uint64_t)s[i] * g(j) + prod[i + j] + carry;
            prod[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + NB] = (uint32_t)carry;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) out[k] = prod[8 + k];
}

// acc -= a * b (mod 2^160); a of NA words, b of NB words (a constant)
template <int NA, int NB, class G>
ZKG_DEV void glv_sub_mul(uint32_t acc[5], const uint32_t* a, G b) {
    uint32_t prod[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (i + j < 5) {
                const uint64_t t = (uint64_t)a[i] * b(j) + prod[i + j] + carry;
                prod[i + j] = (uint32_t)t;
                carry = t >> 32;
            }
        }
        if (i + NB < 5) prod[i + NB] = (uint32_t)carry;
    }
    uint32_t borrow = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) acc[k] = __builtin_subc(acc[k], prod[k], borrow, &borrow);
}

// 160-bit two's complement -> |k| (127 bits) with the sign in bit 127
ZKG_DEV uint4 glv_pack(uint32_t k[5]) {
    const uint32_t neg = k[4] >> 31;
    if (neg) {
        uint32_t carry = 1;
#pragma unroll
        for (int i = 0; i < 5; ++i) k[i] = __builtin_addc(~k[i], 0u, carry, &carry);
    }
    return make_uint4(k[0], k[1], k[2], (k[3] & 0x7fffffffu) | (neg << 31));
}

// Scalar i (any 256-bit value) reduced mod r (2^256 < 6r).
ZKG_DEV void load_scalar_mod_r(const uint32_t* __restrict__ scalars, uint32_t i, uint32_t s[8]) {
    const uint4* p = reinterpret_cast<const uint4*>(scalars + 8 * (size_t)i);
    const uint4 a = p[0], b = p[1];
    s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w; s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
    for (int rep = 0; rep < 5; ++rep) {
        uint32_t d[8], borrow = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) d[k] = __builtin_subc(s[k], FrParams::mod(k), borrow, &borrow);
        if (borrow) break;  // s < r
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = d[k];
    }
}

// The GLV halves (k1, k2) of scalar i, packed as glv_pack does.
ZKG_DEV void glv_split(const uint32_t* __restrict__ scalars, uint32_t i, uint4* h1, uint4* h2) {
    uint32_t s[8];
    load_scalar_mod_r(scalars, i, s);
    uint32_t c1[3], c2[5];
    glv_mul_high<3>(s, GlvConst::g1c, c1);
    glv_mul_high<5>(s, GlvConst::g2c, c2);
    // k1 = s - c1 a1 - c2 a2;  k2 = c1 (-b1) - c2 b2   (mod 2^160; the true values are < 2^127)
    uint32_t k1[5] = {s[0], s[1], s[2], s[3], s[4]};
    glv_sub_mul<3, 2>(k1, c1, GlvConst::a1);
    glv_sub_mul<5, 4>(k1, c2, GlvConst::a2);
    uint32_t k2[5] = {0, 0, 0, 0, 0};
    glv_sub_mul<5, 2>(k2, c2, GlvConst::a1);  // b2 = a1
    uint32_t t[5] = {0, 0, 0, 0, 0};
    glv_sub_mul<3, 4>(t, c1, GlvConst::nb1);  // t = -c1 (-b1)
    uint32_t borrow = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) k2[k] = __builtin_subc(k2[k], t[k], borrow, &borrow);
    *h1 = glv_pack(k1);
    *h2 = glv_pack(k2);
}

__global__ __launch_bounds__(256) void msm_glv_split_kernel(const uint32_t* __restrict__ scalars, uint32_t n,
                                                           uint32_t* __restrict__ split) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 h1, h2;
    glv_split(scalars, i, &h1, &h2);
    *reinterpret_cast<uint4*>(split + 4 * (size_t)i) = h1;
    *reinterpret_cast<uint4*>(split + 4 * ((size_t)n + i)) = h2;
}

void msm_glv_split(const uint32_t* scalars, uint32_t n, uint32_t* split, hipStream_t stream) {
    if (!n) return;
    hipLaunchKernelGGL(msm_glv_split_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, scalars, n, split);
    ZKG_HIP(hipGetLastError());
}

// endo_x[i] = beta * x_i for every base (G2: both Fq components of x times beta)
__global__ __launch_bounds__(256) void msm_endo_kernel(const uint32_t* __restrict__ bases, uint32_t n, int group,
                                                      uint32_t* __restrict__ endo_x) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * group) return;
    const size_t i = t / group, k = t % group;
    Fq beta;
#pragma unroll
    for (int j = 0; j < 8; ++j) beta.w[j] = group == 1 ? GlvConst::beta_g1(j) : GlvConst::beta_g2(j);
    const Fq x = fp_load<FqParams>(bases + 16 * group * i + 8 * k);
    fp_store(endo_x + 8 * t, x * beta);
}

// ---------------------------------------------------------------- GLS split (G2)
// On G2 a second endomorphism is cheap: psi(x, y) = (conj(x) gx, conj(y) gy) with
// gx = xi^((p-1)/3), gy = xi^((p-1)/2), xi = 9 + u (the p-power Frobenius carried to the sextic
// twist), maps Q to mu Q, mu = p mod r, and commutes with phi. So
//   s Q = k0 Q + k1 phi(Q) + k2 psi(Q) + k3 phi(psi(Q))  whenever  k0 + k1 lambda + k2 mu + k3 lambda mu = s (mod r),
// and a 4-dimensional split gives coefficients of 64 bits instead of the GLV split's 127. The
// lattice of (a0..a3) with a0 + a1 lambda + a2 mu + a3 lambda mu = 0 (mod r) has the LLL-reduced
// basis (x = 0x44e992b44a6909f1, the BN parameter of the curve)
//   b0 = (2x, x, x + 1, -x),  b1 = (-x, x, x, 2x + 1),  b2 = (-x - 1, x, -x, -2x),  b3 = (2x + 1, x + 1, -x, x).
// Babai rounding: (s, 0, 0, 0) = sum_j alpha_j b_j with alpha_j = s A_j / r, A = r times the first
// row of the basis' inverse (|A_j| < 2^189; A_1, A_2 < 0). c_j = floor((s G_j + 2^287) / 2^288) with
// G_j = round(|A_j| 2^288 / r) is within 1/2 + 2^-35 of |alpha_j| (s < 2^254), and with d_j =
// sign(A_j) b_j, (k0..k3) = (s, 0, 0, 0) - sum_j c_j d_j:
//   k0 = s - 2x c0 - x c1 - (x + 1) c2 - (2x + 1) c3      k1 = -x c0 + x c1 + x c2 - (x + 1) c3
//   k2 = -(x + 1) c0 + x c1 - x c2 + x c3                 k3 = x c0 + (2x + 1) c1 - 2x c2 - x c3
// |k_i| <= (1/2 + 2^-35) sum_j |b_j,i| <= 3x (1 + 2^-34) < 2^63.7, so the coefficients are exact
// modulo 2^128 and each fits 64 bits.
struct GlsConst {
    static constexpr uint64_t X = 0x44e992b44a6909f1ull;
    // G_j, 7 words each
    static ZKG_HD constexpr uint32_t g(int j, int k) {
        constexpr uint32_t t[4][7] = {
            {0x7d1fff2eu, 0x95d51bb1u, 0x46f4bda9u, 0xfc7184aeu, 0x08e5da66u, 0xb0d92b93u, 0x9e80318au},
            {0x353ccca1u, 0x32e42728u, 0x2dff2915u, 0xa3e5577fu, 0x55b4ca7bu, 0xb0d92b95u, 0x9e80318au},
            {0xa3f693b5u, 0xfac4daffu, 0x071c4c43u, 0xa3e55782u, 0x55b4ca7bu, 0xb0d92b95u, 0x9e80318au},
            {0x23038c2au, 0xcef3cd3fu, 0xc170977du, 0xa3e5577du, 0x55b4ca7bu, 0xb0d92b95u, 0x9e80318au}};
        return t[j][k];
    }
    // psi's coefficients times 2^256 mod p (the bases are in Montgomery form): component c (0 = real,
    // 1 = u) of gx = 0x2fb347984f7911f74c0bec3cf559b143b78cc310c2c3330c99e39557176f553d
    // + 0x16c9e55061ebae204ba4cc8bd75a079432ae2a1d0b7c9dce1665d51c640fcba2 u and of
    // gy = 0x063cf305489af5dcdc5ec698b6e2f9b9dbaae0eda9c95998dc54014671a0135a
    // + 0x07c03cbcac41049a0704b5a7ec796f2b21807dc98fa25bd282d37f632623b0e3 u
    static ZKG_HD constexpr uint32_t gx(int c, int k) {
        constexpr uint32_t t[2][8] = {
            {0x4563ab30u, 0xb5773b10u, 0xa9aa6454u, 0x347f91c8u, 0x242e0991u, 0x7a007127u, 0x118214ecu, 0x1956bcd8u},
            {0xa0aa4757u, 0x6e849f1eu, 0x89f89141u, 0xaa1c7b6du, 0xfae0ca3au, 0xb6e713cdu, 0x4e82ebc3u, 0x26694fbbu}};
        return t[c][k];
    }
    static ZKG_HD constexpr uint32_t gy(int c, int k) {
        constexpr uint32_t t[2][8] = {
            {0x2936b629u, 0xe4bbdd0cu, 0xe133bacbu, 0xbb30f162u, 0xf9645366u, 0x31a9d1b6u, 0xa500f8ddu, 0x253570beu},
            {0x5ffe77c7u, 0xa1d77ce4u, 0x7826d1dbu, 0x07affd11u, 0xbb7edc6bu, 0x6d16bd27u, 0x85defeccu, 0x2c872002u}};
        return t[c][k];
    }
};

typedef unsigned __int128 u128;

// floor((s G_j + 2^287) / 2^288) mod 2^128: words 9..12 of the 15-word sum
ZKG_DEV u128 gls_coeff(const uint32_t s[8], int j) {
    uint32_t prod[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) prod[k] = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const uint64_t t = (uint64_t)s[i] * GlsConst::g(j, k) + prod[i + k] + carry;
            prod[i + k] = (uint32_t)t;
            carry = t >> 32;
        }
        prod[i + 7] = (uint32_t)carry;
    }
    uint32_t c = 0;
    prod[8] = __builtin_addc(prod[8], 0x80000000u, 0u, &c);
#pragma unroll
    for (int k = 9; k < 13; ++k) prod[k] = __builtin_addc(prod[k], 0u, c, &c);
    return ((u128)prod[12] << 96) | ((u128)prod[11] << 64) | ((u128)prod[10] << 32) | prod[9];
}

// The coefficients (k0, k1, k2, k3) of scalar i as magnitudes (< 2^64) and signs.
ZKG_DEV void gls_split(const uint32_t* __restrict__ scalars, uint32_t i, uint64_t mag[4], uint32_t neg[4]) {
    uint32_t s[8];
    load_scalar_mod_r(scalars, i, s);
    const u128 c0 = gls_coeff(s, 0), c1 = gls_coeff(s, 1), c2 = gls_coeff(s, 2), c3 = gls_coeff(s, 3);
    const u128 x = GlsConst::X;
    const u128 p0 = c0 * x, p1 = c1 * x, p2 = c2 * x, p3 = c3 * x;  // mod 2^128
    const u128 sl = ((u128)s[3] << 96) | ((u128)s[2] << 64) | ((u128)s[1] << 32) | s[0];
    u128 k[4];
    k[0] = sl - (p0 + p0 + p1 + p2 + c2 + p3 + p3 + c3);
    k[1] = (p1 + p2) - (p0 + p3 + c3);
    k[2] = (p1 + p3) - (p0 + c0 + p2);
    k[3] = (p0 + p1 + p1 + c1) - (p2 + p2 + p3);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        neg[t] = (uint32_t)(k[t] >> 127);
        mag[t] = (uint64_t)(neg[t] ? (u128)0 - k[t] : k[t]);
    }
}

// GLS plans: table[0..n) = the bases, table[n..2n) = psi of each, endo_x[0..2n) = beta x of every
// table entry (the table and phi images of a plan over 4n virtual points, as a precomputed plan's).
__global__ __launch_bounds__(256) void msm_gls_prepare_kernel(const uint32_t* __restrict__ bases, uint32_t n,
                                                             uint32_t* __restrict__ table,
                                                             uint32_t* __restrict__ endo_x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    typedef FeTraits<Fq2> T2;
    Fq2 gx, gy;
    Fq beta;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        gx.a.w[j] = GlsConst::gx(0, j);
        gx.b.w[j] = GlsConst::gx(1, j);
        gy.a.w[j] = GlsConst::gy(0, j);
        gy.b.w[j] = GlsConst::gy(1, j);
        beta.w[j] = GlvConst::beta_g2(j);
    }
    const uint32_t* src = bases + 32 * (size_t)i;
    const Fq2 x = T2::load(src), y = T2::load(src + 16);
    const Fq zero = fp_zero<FqParams>();
    const Fq2 px = Fq2{x.a, zero - x.b} * gx, py = Fq2{y.a, zero - y.b} * gy;  // infinity (0, 0) stays
    uint32_t* q = table + 32 * (size_t)i;
    uint32_t* pq = table + 32 * ((size_t)n + i);
    T2::store(q, x);
    T2::store(q + 16, y);
    T2::store(pq, px);
    T2::store(pq + 16, py);
    T2::store(endo_x + 16 * (size_t)i, Fq2{x.a * beta, x.b * beta});
    T2::store(endo_x + 16 * ((size_t)n + i), Fq2{px.a * beta, px.b * beta});
}

__global__ __launch_bounds__(256) void msm_gls_split_kernel(const uint32_t* __restrict__ scalars, uint32_t n,
                                                           uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t mag[4];
    uint32_t neg[4];
    gls_split(scalars, i, mag, neg);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        uint32_t* o = out + 12 * (size_t)i + 3 * t;
        o[0] = (uint32_t)mag[t];
        o[1] = (uint32_t)(mag[t] >> 32);
        o[2] = neg[t];
    }
}

void msm_gls_split(const uint32_t* scalars, uint32_t n, uint32_t* out, hipStream_t stream) {
    if (!n) return;
    hipLaunchKernelGGL(msm_gls_split_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, scalars, n, out);
    ZKG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- digits: count and scatter

__device__ __forceinline__ uint32_t window_bits(const uint32_t s[8], uint32_t pos, uint32_t c) {
    const uint32_t word = pos >> 5, off = pos & 31;
    uint32_t v = s[word] >> off;
    if (off + c > 32 && word + 1 < 8) v |= s[word + 1] << (32 - off);
    return v & ((1u << c) - 1);
}

// Scalar i of the digit stages: 8 plain words, or (GLV) a 127-bit magnitude with its sign in
// bit 127; returns the sign.
template <bool GLV>
ZKG_DEV uint32_t load_digit_scalar(const uint32_t* __restrict__ scalars, uint32_t i, uint32_t s[8]) {
    if (GLV) {
        const uint4 a = *reinterpret_cast<const uint4*>(scalars + 4 * (size_t)i);
        s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w & 0x7fffffffu; s[4] = s[5] = s[6] = s[7] = 0;
        return a.w >> 31;
    }
    const uint4* p = reinterpret_cast<const uint4*>(scalars + 8 * (size_t)i);
    const uint4 a = p[0], b = p[1];
    s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w; s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
    return 0;
}

// Signed digit of window value v (bits + carry). Full scalars follow the reference's rule
// (v >= 2^(c-1) -> v - 2^c, carry 1), digits in [-2^(c-1), 2^(c-1)). GLV half-scalars let the
// digit reach +2^(c-1) (v > 2^(c-1) wraps), so ceil(128 / c) windows hold any |k| < 2^127 with no
// carry out of the top window. Either way |digit| <= 2^(c-1): bucket |digit| - 1.
template <bool GLV>
ZKG_DEV void signed_digit(uint32_t v, uint32_t width, uint32_t& mag, uint32_t& neg, uint32_t& carry) {
    const uint32_t half = 1u << (width - 1);
    const bool wrap = GLV ? v > half : v >= half;
    mag = wrap ? (1u << width) - v : v;
    neg = wrap ? 1u : 0u;
    carry = neg;
}

// Bit offset and width of window w (see MsmPlan::wb, nw).
ZKG_DEV uint32_t window_pos(uint32_t w, uint32_t wb, uint32_t nw) { return w * wb + min(w, nw); }
ZKG_DEV uint32_t window_width(uint32_t w, uint32_t wb, uint32_t nw) { return wb + (w < nw ? 1u : 0u); }

template <bool SCATTER, bool GLV>
__global__ __launch_bounds__(256) void msm_digits_kernel(const uint32_t* __restrict__ scalars, uint32_t n, uint32_t c,
                                                         uint32_t W, uint32_t wb, uint32_t nw,
                                                         uint32_t* __restrict__ counter,
                                                         const uint32_t* __restrict__ starts,
                                                         uint32_t* __restrict__ sorted) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const uint32_t B = 1u << (c - 1);
    const uint32_t lane = __lane_id();
    const unsigned long long lt_mask = (1ull << lane) - 1;
    uint32_t s[8], sign = 0;
    if (valid) {
        sign = load_digit_scalar<GLV>(scalars, i, s);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = 0;
    }
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint32_t width = window_width(w, wb, nw);
        const uint32_t v = window_bits(s, window_pos(w, wb, nw), width) + carry;
        uint32_t mag, neg;
        signed_digit<GLV>(v, width, mag, neg, carry);
        neg ^= sign;
        const bool has = valid && mag != 0;
        const uint32_t slot = w * B + mag - 1;
        const unsigned long long active = __ballot(has);
        if (active == 0) continue;
        const int leader = __ffsll((long long)active) - 1;
        const uint32_t lkey = __shfl(slot, leader);
        const bool same = has && slot == lkey;
        const unsigned long long smask = __ballot(same);
        if (SCATTER) {
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&counter[lkey], (uint32_t)__popcll(smask));
            base = __shfl(base, leader);
            if (has) {
                uint32_t pos;
                if (same) pos = base + (uint32_t)__popcll(smask & lt_mask);
                else pos = atomicAdd(&counter[slot], 1u);
                sorted[starts[slot] + pos] = i | (neg << 31);
            }
        } else {
            if ((int)lane == leader) atomicAdd(&counter[lkey], (uint32_t)__popcll(smask));
            else if (has && !same) atomicAdd(&counter[slot], 1u);
        }
    }
}

// ---------------------------------------------------------------- LDS counting sort (c <= 16)
// Device-scope atomics on the W*B global counters cost ~40 ns each at this scale (the counters
// are shared by all XCDs), so for windows up to c = 16 the sort avoids them entirely:
//   digits   one lane per scalar writes its W signed digits, digits[w][i] = bucket | neg << 31
//            (NO_DIGIT for a zero digit), coalesced per window;
//   hist     block (w, blk) builds an LDS histogram of its P points' buckets (LDS atomics) and
//            writes it to table[w][blk][b] (coalesced);
//   colscan  one lane per (w, b) turns its column of the table into exclusive offsets over the
//            blocks and writes the slot total counts[w*B + b];
//   (slot scan of counts -> starts, istarts as in the atomic path)
//   place    block (w, blk) loads cursor[b] = starts[w*B + b] + table[w][blk][b] into LDS and
//            places each point at sorted[atomicAdd_LDS(cursor[b], 1)].

static constexpr uint32_t NO_DIGIT = 0xffffffffu;
static constexpr int SORT_THREADS = 1024;

template <bool GLV>
__global__ __launch_bounds__(256) void msm_digits_store_kernel(const uint32_t* __restrict__ scalars, uint32_t n,
                                                              uint32_t W, uint32_t wb, uint32_t nw,
                                                              uint32_t* __restrict__ digits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s[8];
    const uint32_t sign = load_digit_scalar<GLV>(scalars, i, s);
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint32_t width = window_width(w, wb, nw);
        const uint32_t v = window_bits(s, window_pos(w, wb, nw), width) + carry;
        uint32_t mag, neg;
        signed_digit<GLV>(v, width, mag, neg, carry);
        digits[(size_t)w * n + i] = mag ? ((mag - 1) | ((neg ^ sign) << 31)) : NO_DIGIT;
    }
}

// GLV plans on the LDS sort path, in one pass per point: the split, the digits of both halves
// (virtual points i and n + i) and phi's x-coordinate of base i (what msm_glv_split_kernel,
// msm_endo_kernel and msm_digits_store_kernel<true> do in three launches, without the halves'
// round trip through HBM).
__global__ __launch_bounds__(256) void msm_glv_digits_kernel(const uint32_t* __restrict__ scalars,
                                                            const uint32_t* __restrict__ bases, uint32_t n,
                                                            int group, uint32_t W, uint32_t wb, uint32_t nw,
                                                            uint32_t* __restrict__ endo_x,
                                                            uint32_t* __restrict__ digits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t nv = 2 * (size_t)n;
    uint4 h[2];
    glv_split(scalars, i, &h[0], &h[1]);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        uint32_t s[8] = {h[t].x, h[t].y, h[t].z, h[t].w & 0x7fffffffu, 0, 0, 0, 0};
        const uint32_t sign = h[t].w >> 31, v_idx = i + (t ? n : 0u);
        uint32_t carry = 0;
        for (uint32_t w = 0; w < W; ++w) {
            const uint32_t width = window_width(w, wb, nw);
            const uint32_t v = window_bits(s, window_pos(w, wb, nw), width) + carry;
            uint32_t mag, neg;
            signed_digit<true>(v, width, mag, neg, carry);
            digits[(size_t)w * nv + v_idx] = mag ? ((mag - 1) | ((neg ^ sign) << 31)) : NO_DIGIT;
        }
    }
    Fq beta;
#pragma unroll
    for (int j = 0; j < 8; ++j) beta.w[j] = group == 1 ? GlvConst::beta_g1(j) : GlvConst::beta_g2(j);
    for (int k = 0; k < group; ++k) {
        const Fq x = fp_load<FqParams>(bases + 16 * group * (size_t)i + 8 * k);
        fp_store(endo_x + 8 * ((size_t)group * i + k), x * beta);
    }
}

// Precomputed bases (bases buffer = P_0..P_{n-1}, 2^64 P_0..2^64 P_{n-1}): GLV half-scalar i < 2n
// (k1 of point i for i < n, k2 of point i - n after) has 2W signed digits over its 128 bits, W in
// each 64-bit half with the same widths, the carry running on across bit 64. The low half's
// digits belong to the virtual point of P (phi(P) for k2), the high half's to that of 2^64 P
// (phi(2^64 P)), both in windows 0..W-1. Virtual points: [0, n) P_i, [n, 2n) 2^64 P_i,
// [2n, 3n) phi(P_i), [3n, 4n) phi(2^64 P_i): below 2n an entry of the bases buffer, above it phi of
// entry v - 2n. The top digit takes no carry out: |k| < 2^127 leaves bit 127 clear.
__global__ __launch_bounds__(256) void msm_digits_split_store_kernel(const uint32_t* __restrict__ halves, uint32_t n,
                                                                    uint32_t W, uint32_t wb, uint32_t nw,
                                                                    uint32_t* __restrict__ digits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n) return;
    const size_t nv = 4 * (size_t)n;
    uint32_t s[8];
    const uint32_t sign = load_digit_scalar<true>(halves, i, s);
    const uint32_t vlo = i < n ? i : n + i, vhi = vlo + n;
    uint32_t carry = 0;
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t v_idx = h ? vhi : vlo;
        for (uint32_t w = 0; w < W; ++w) {
            const uint32_t width = window_width(w, wb, nw);
            const uint32_t v = window_bits(s, 64 * h + window_pos(w, wb, nw), width) + carry;
            uint32_t mag, neg;
            signed_digit<true>(v, width, mag, neg, carry);
            digits[(size_t)w * nv + v_idx] = mag ? ((mag - 1) | ((neg ^ sign) << 31)) : NO_DIGIT;
        }
    }
}

// GLS plans (bases table = Q_0..Q_{n-1}, psi(Q_0)..psi(Q_{n-1})): virtual points [0, n) Q_i with
// k0, [n, 2n) psi(Q_i) with k2, [2n, 3n) phi(Q_i) with k1, [3n, 4n) phi(psi(Q_i)) with k3. Each
// coefficient (|k| < 2^64) has W signed digits over 65 bits; the top window (bit 64 clear) takes no
// carry out.
__global__ __launch_bounds__(256) void msm_gls_digits_store_kernel(const uint32_t* __restrict__ scalars, uint32_t n,
                                                                  uint32_t W, uint32_t wb, uint32_t nw,
                                                                  uint32_t* __restrict__ digits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t nv = 4 * (size_t)n;
    uint64_t mag[4];
    uint32_t neg[4];
    gls_split(scalars, i, mag, neg);
    const uint32_t v_idx[4] = {i, 2 * n + i, n + i, 3 * n + i};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        uint32_t carry = 0;
        for (uint32_t w = 0; w < W; ++w) {
            const uint32_t width = window_width(w, wb, nw);
            const uint32_t v = (uint32_t)(mag[t] >> window_pos(w, wb, nw)) & ((1u << width) - 1);
            uint32_t dm, dn;
            signed_digit<true>(v + carry, width, dm, dn, carry);
            digits[(size_t)w * nv + v_idx[t]] = dm ? ((dm - 1) | ((dn ^ neg[t]) << 31)) : NO_DIGIT;
        }
    }
}

__global__ __launch_bounds__(SORT_THREADS) void msm_hist_kernel(const uint32_t* __restrict__ digits, uint32_t n,
                                                               uint32_t B, uint32_t P, uint32_t nblk,
                                                               uint32_t* __restrict__ table) {
    extern __shared__ uint32_t hist[];
    const uint32_t blk = blockIdx.x, w = blockIdx.y;
    for (uint32_t b = threadIdx.x; b < B; b += SORT_THREADS) hist[b] = 0;
    __syncthreads();
    const uint32_t lo = blk * P, hi = min(lo + P, n);
    const uint32_t* d = digits + (size_t)w * n;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) {
        const uint32_t e = d[i];
        if (e != NO_DIGIT) atomicAdd(&hist[e & 0x7fffffffu], 1u);
    }
    __syncthreads();
    uint32_t* row = table + ((size_t)w * nblk + blk) * B;
    for (uint32_t b = threadIdx.x; b < B; b += SORT_THREADS) row[b] = hist[b];
}

// table[w][blk][b] -> exclusive prefix over blk; counts[w*B + b] = column total
__global__ __launch_bounds__(256) void msm_colscan_kernel(uint32_t* __restrict__ table, uint32_t W, uint32_t B,
                                                         uint32_t nblk, uint32_t* __restrict__ counts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= W * B) return;
    const uint32_t w = t / B, b = t % B;
    uint32_t* col = table + (size_t)w * nblk * B + b;
    uint32_t run = 0;
    for (uint32_t k = 0; k < nblk; ++k) {
        const uint32_t v = col[(size_t)k * B];
        col[(size_t)k * B] = run;
        run += v;
    }
    counts[t] = run;
}

__global__ __launch_bounds__(SORT_THREADS) void msm_place_kernel(const uint32_t* __restrict__ digits, uint32_t n,
                                                                uint32_t B, uint32_t P, uint32_t nblk,
                                                                const uint32_t* __restrict__ table,
                                                                const uint32_t* __restrict__ starts,
                                                                uint32_t* __restrict__ sorted) {
    extern __shared__ uint32_t cursor[];
    const uint32_t blk = blockIdx.x, w = blockIdx.y;
    const uint32_t* row = table + ((size_t)w * nblk + blk) * B;
    const uint32_t* st = starts + (size_t)w * B;
    for (uint32_t b = threadIdx.x; b < B; b += SORT_THREADS) cursor[b] = st[b] + row[b];
    __syncthreads();
    const uint32_t lo = blk * P, hi = min(lo + P, n);
    const uint32_t* d = digits + (size_t)w * n;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) {
        const uint32_t e = d[i];
        if (e == NO_DIGIT) continue;
        const uint32_t pos = atomicAdd(&cursor[e & 0x7fffffffu], 1u);
        sorted[pos] = i | (e & 0x80000000u);
    }
}

// ---- two-pass placement (the default for c <= 16 where a packed entry fits 32 bits)
// The one-pass placement above writes each 4-byte entry to its bucket's cursor, and a block holds
// about two points per bucket, so its writes land all over the window's range: rocprof measured
// 716 MB written to HBM for the 75 MB of entries of a G1 2^20 MSM (c = 15), 0.28 ms. Two passes
// keep the writes local: (1) each block writes its points as runs per partition (2^plog
// partitions per window, each a range of 2^sh buckets of the window's own bucket range) of a
// temporary array, one run per (block, partition), reserved with one global atomic; (2) each
// partition is cut into slices of SLICE entries, and one workgroup per slice counts its buckets
// and places its entries, its run of each bucket reserved with one global atomic, so the
// scattered writes of a workgroup stay within a 32 KB range of the final list (at G1 2^20) that
// one CU writes, small enough for the L2 to merge. Slices keep the work of a workgroup bounded
// when a few buckets hold most entries (witness-like scalars: a fifth of a window in one bucket).
// A packed entry holds the point index (31 - kb bits), the bucket within its partition (the kb
// bits below the sign; sh <= kb) and the sign (bit 31). Loops keep four entries per lane in
// flight. The order of entries within a bucket depends on the atomics, as in the one-pass
// placement; sums do not.
// The partition tables (in the plan's `table`): bcount[w][blk][p] (a block's entries per
// partition), pbase[q] (q = w 2^plog + p: exclusive scan of the partition totals, pbase[M] = all
// entries), pcur[q] (the next free run of each partition) and sbase[q] (exclusive scan of the
// slices per partition, sbase[M] = all slices), M = W 2^plog. counts and the plan's `cursors`
// (the next free position of each bucket after its start) are zeroed before.
static constexpr uint32_t PART_MAX = 1024;  // partitions per window, and buckets per partition, at most
static constexpr uint32_t BUCKET_THREADS = 256;
static constexpr uint32_t SLICE = 8192;  // entries per workgroup of the bucket pass
static constexpr int PLACE_PER = 8;       // entries per lane of a staged chunk of the partition pass (8192)
static constexpr int BUCKET_KEYS = 256;   // buckets per partition at most (kb <= 8)

// LDS counters aggregated over the lanes of a wave that share a key: lanes whose key is `hot` (a
// block-uniform guess of the most frequent key) or, else, the key of the wave's first active lane
// use one atomic for all of them (ranks by popcount). Inputs with a hot bucket would otherwise
// serialise one LDS atomic per lane on a single address. Other lanes use their own atomics.
ZKG_DEV uint32_t lds_fetch_inc(uint32_t* cnt, uint32_t key, uint32_t hot) {
    uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
    if (__ballot(key == hot)) k0 = hot;
    const uint64_t same = __ballot(key == k0);
    if (key != k0) return atomicAdd(&cnt[key], 1u);
    const uint32_t lane = __lane_id(), first = (uint32_t)__builtin_ctzll(same);
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(&cnt[k0], (uint32_t)__builtin_popcountll(same));
    base = __shfl(base, (int)first);
    return base + (uint32_t)__builtin_popcountll(same & ((1ull << lane) - 1));
}

ZKG_DEV void lds_inc(uint32_t* cnt, uint32_t key) {
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
    const uint64_t same = __ballot(key == k0);
    if (key != k0) {
        atomicAdd(&cnt[key], 1u);
        return;
    }
    if (__lane_id() == (uint32_t)__builtin_ctzll(same)) atomicAdd(&cnt[k0], (uint32_t)__builtin_popcountll(same));
}

// Bucket bits of partition ranges in window w: the window's bucket range 2^cw (cw = its digit
// width - 1, at most cb) over 2^plog partitions.
ZKG_DEV uint32_t part_shift(uint32_t w, uint32_t wb, uint32_t nw, uint32_t cb, uint32_t plog) {
    const uint32_t cw = min(cb, wb + (w < nw ? 1u : 0u) - 1);
    return cw > plog ? cw - plog : 0u;
}

// LDS of one staged chunk: per-key counts and offsets (NK <= KEYS keys), per-wave totals, and the
// chunk's entries with their keys.
template <int THREADS, int PER, int KEYS>
struct StageLds {
    uint32_t cnt[KEYS], off[KEYS], wsum[THREADS / 64];
    uint32_t val[PER * THREADS];
    uint16_t key[PER * THREADS];
};

// Places one chunk of a workgroup's entries (lane t's entries k: ok[k], key[k], val[k]) at
// dst + cursor[key] onward, cursor[] (LDS, global positions) advancing by the chunk's count of each
// key. The entries are first sorted by key in LDS (ranks from aggregated LDS atomics, offsets from
// a workgroup scan), then written in that order, so consecutive lanes write consecutive positions
// of a key's run: a wave's store covers a few runs instead of up to 64, and the L2 no longer
// writes back partly filled lines (G1 2^20, c = 15: the partition pass wrote 177 MB to HBM for its
// 75 MB of entries, 94 MB staged; the count stage 0.242 -> 0.20 ms). Every thread of the
// workgroup calls it for every chunk (barriers inside).
template <int THREADS, int PER, int KEYS>
ZKG_DEV void stage_chunk(const bool (&ok)[PER], const uint32_t (&key)[PER], const uint32_t (&val)[PER], uint32_t hot,
                         uint32_t nk, StageLds<THREADS, PER, KEYS>& L, uint32_t* cursor, uint32_t* __restrict__ dst) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr uint32_t WAVES = THREADS / 64;
    for (uint32_t k = tid; k < nk; k += THREADS) L.cnt[k] = 0;
    __syncthreads();
    uint32_t rank[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (ok[k]) rank[k] = lds_fetch_inc(L.cnt, key[k], hot);
    __syncthreads();
    // exclusive scan of cnt[0 .. nk) into off (nk <= THREADS)
    const uint32_t c = tid < nk ? L.cnt[tid] : 0u;
    uint32_t x = c;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t a = __shfl_up(x, o);
        if (lane >= o) x += a;
    }
    if (lane == 63) L.wsum[wv] = x;
    __syncthreads();
    if (wv == 0) {
        uint32_t t = lane < WAVES ? L.wsum[lane] : 0u;
#pragma unroll
        for (uint32_t o = 1; o < WAVES; o <<= 1) {
            const uint32_t a = __shfl_up(t, o);
            if (lane >= o) t += a;
        }
        if (lane < WAVES) L.wsum[lane] = t;
    }
    __syncthreads();
    if (tid < nk) L.off[tid] = (wv ? L.wsum[wv - 1] : 0u) + x - c;
    const uint32_t total = L.wsum[WAVES - 1];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (ok[k]) {
            const uint32_t at = L.off[key[k]] + rank[k];
            L.val[at] = val[k];
            L.key[at] = (uint16_t)key[k];
        }
    __syncthreads();
    for (uint32_t at = tid; at < total; at += THREADS) {
        const uint32_t k = L.key[at];
        dst[cursor[k] + (at - L.off[k])] = L.val[at];
    }
    __syncthreads();
    for (uint32_t k = tid; k < nk; k += THREADS) cursor[k] += L.cnt[k];
    __syncthreads();
}

// A block's entries per partition -> bcount, and into the partition totals (zeroed beforehand).
__global__ __launch_bounds__(SORT_THREADS) void msm_part_hist_kernel(const uint32_t* __restrict__ digits, uint32_t n,
                                                                    uint32_t P, uint32_t nblk, uint32_t wb,
                                                                    uint32_t nw, uint32_t cb, uint32_t plog,
                                                                    uint32_t* __restrict__ bcount,
                                                                    uint32_t* __restrict__ ptotal) {
    __shared__ uint32_t hist[PART_MAX];
    const uint32_t blk = blockIdx.x, w = blockIdx.y, NP = 1u << plog, sh = part_shift(w, wb, nw, cb, plog);
    for (uint32_t p = threadIdx.x; p < NP; p += SORT_THREADS) hist[p] = 0;
    __syncthreads();
    const uint32_t lo = blk * P, hi = min(lo + P, n);
    const uint32_t* d = digits + (size_t)w * n;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 4 * SORT_THREADS) {
        uint32_t e[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) e[k] = i + k * SORT_THREADS < hi ? d[i + k * SORT_THREADS] : NO_DIGIT;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (e[k] != NO_DIGIT) lds_inc(hist, (e[k] & 0x7fffffffu) >> sh);
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < NP; p += SORT_THREADS) {
        const uint32_t h = hist[p];
        bcount[((size_t)w * nblk + blk) * NP + p] = h;
        if (h) atomicAdd(&ptotal[w * NP + p], h);
    }
}

// One workgroup: pbase[0 .. M) = exclusive scan of pbase (the partition totals) in place,
// pbase[M] = the total, pcur = pbase; sbase[0 .. M] = exclusive scan of ceil(total / SLICE). Four
// consecutive entries per lane per round of 4096.
__global__ __launch_bounds__(1024) void msm_part_scan_kernel(uint32_t* __restrict__ pbase, uint32_t M,
                                                            uint32_t* __restrict__ pcur,
                                                            uint32_t* __restrict__ sbase) {
    __shared__ uint32_t wsum[2][16];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t carry = 0, scarry = 0;
    for (uint32_t base = 0; base < M; base += 4096) {
        uint32_t v[4], sum = 0, ssum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t idx = base + 4 * threadIdx.x + k;
            v[k] = idx < M ? pbase[idx] : 0u;
            sum += v[k];
            ssum += (v[k] + SLICE - 1) / SLICE;
        }
        uint32_t x = sum, y = ssum;  // inclusive scans over the wave
#pragma unroll
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t a = __shfl_up(x, off), b = __shfl_up(y, off);
            if (lane >= off) {
                x += a;
                y += b;
            }
        }
        if (lane == 63) {
            wsum[0][wv] = x;
            wsum[1][wv] = y;
        }
        __syncthreads();
        if (wv == 0) {  // inclusive scans of the 16 wave totals
            uint32_t t = lane < 16 ? wsum[0][lane] : 0u, u = lane < 16 ? wsum[1][lane] : 0u;
#pragma unroll
            for (uint32_t off = 1; off < 16; off <<= 1) {
                const uint32_t a = __shfl_up(t, off), b = __shfl_up(u, off);
                if (lane >= off) {
                    t += a;
                    u += b;
                }
            }
            if (lane < 16) {
                wsum[0][lane] = t;
                wsum[1][lane] = u;
            }
        }
        __syncthreads();
        uint32_t run = carry + (wv ? wsum[0][wv - 1] : 0u) + x - sum;
        uint32_t srun = scarry + (wv ? wsum[1][wv - 1] : 0u) + y - ssum;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t idx = base + 4 * threadIdx.x + k;
            if (idx < M) {
                pbase[idx] = run;
                pcur[idx] = run;
                sbase[idx] = srun;
            }
            run += v[k];
            srun += (v[k] + SLICE - 1) / SLICE;
        }
        carry += wsum[0][15];
        scarry += wsum[1][15];
        __syncthreads();  // wsum is rewritten next round
    }
    if (threadIdx.x == 0) {
        pbase[M] = carry;
        sbase[M] = scarry;
    }
}

__global__ __launch_bounds__(SORT_THREADS) void msm_part_place_kernel(const uint32_t* __restrict__ digits, uint32_t n,
                                                                     uint32_t P, uint32_t nblk, uint32_t wb,
                                                                     uint32_t nw, uint32_t cb, uint32_t plog,
                                                                     uint32_t kb, const uint32_t* __restrict__ bcount,
                                                                     uint32_t* __restrict__ pcur,
                                                                     uint32_t* __restrict__ temp) {
    __shared__ uint32_t cursor[PART_MAX];
    __shared__ uint32_t hot_s;
    __shared__ StageLds<SORT_THREADS, PLACE_PER, PART_MAX> st;
    const uint32_t blk = blockIdx.x, w = blockIdx.y, NP = 1u << plog, sh = part_shift(w, wb, nw, cb, plog);
    if (threadIdx.x == 0) hot_s = 0;
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < NP; p += SORT_THREADS) {  // reserve this block's run of each partition
        const uint32_t h = bcount[((size_t)w * nblk + blk) * NP + p];
        cursor[p] = h ? atomicAdd(&pcur[w * NP + p], h) : 0u;
        if (h) atomicMax(&hot_s, (h << 10) | p);  // the block's fullest partition (h <= 2^20)
    }
    __syncthreads();
    const uint32_t hot = hot_s & 0x3ffu;
    const uint32_t lo = blk * P, hi = min(lo + P, n), bmask = (1u << sh) - 1;
    const uint32_t* d = digits + (size_t)w * n;
    for (uint32_t base = lo; base < hi; base += PLACE_PER * SORT_THREADS) {  // block-uniform trip count
        bool ok[PLACE_PER];
        uint32_t key[PLACE_PER], val[PLACE_PER];
#pragma unroll
        for (int k = 0; k < PLACE_PER; ++k) {
            const uint32_t i = base + threadIdx.x + k * SORT_THREADS;
            const uint32_t e = i < hi ? d[i] : NO_DIGIT;
            ok[k] = e != NO_DIGIT;
            key[k] = (e & 0x7fffffffu) >> sh;
            val[k] = (e & 0x80000000u) | ((e & bmask) << (31 - kb)) | i;
        }
        stage_chunk(ok, key, val, hot, NP, st, cursor, temp);
    }
}

// The slice of workgroup blockIdx.x: its partition q (binary search of sbase) and entry range.
struct PartSlice {
    uint32_t q, lo, hi;
    bool valid;
};
ZKG_DEV PartSlice part_slice(const uint32_t* __restrict__ pbase, const uint32_t* __restrict__ sbase, uint32_t M) {
    PartSlice r;
    const uint32_t g = blockIdx.x;
    r.valid = g < sbase[M];
    r.q = 0;
    r.lo = r.hi = 0;
    if (!r.valid) return r;
    uint32_t lo = 0, hi = M;  // last q with sbase[q] <= g (partitions with no slice share their successor's sbase)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sbase[mid] <= g) lo = mid;
        else hi = mid;
    }
    r.q = lo;
    r.lo = pbase[lo] + (g - sbase[lo]) * SLICE;
    r.hi = min(r.lo + SLICE, pbase[lo + 1]);
    return r;
}

// Points per bucket of a slice, added into counts (zeroed beforehand).
__global__ __launch_bounds__(BUCKET_THREADS) void msm_bucket_count_kernel(const uint32_t* __restrict__ temp,
                                                                         const uint32_t* __restrict__ pbase,
                                                                         const uint32_t* __restrict__ sbase, uint32_t M,
                                                                         uint32_t wb, uint32_t nw, uint32_t cb,
                                                                         uint32_t plog, uint32_t kb, uint32_t B,
                                                                         uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[PART_MAX];
    const PartSlice sl = part_slice(pbase, sbase, M);
    if (!sl.valid) return;
    const uint32_t w = sl.q >> plog, p = sl.q & ((1u << plog) - 1), sh = part_shift(w, wb, nw, cb, plog);
    const uint32_t nb = 1u << sh, bmask = nb - 1;
    for (uint32_t b = threadIdx.x; b < nb; b += BUCKET_THREADS) hist[b] = 0;
    __syncthreads();
    for (uint32_t i = sl.lo + threadIdx.x; i < sl.hi; i += 4 * BUCKET_THREADS) {
        uint32_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = i + k * BUCKET_THREADS < sl.hi ? temp[i + k * BUCKET_THREADS] : 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * BUCKET_THREADS < sl.hi) lds_inc(hist, (x[k] >> (31 - kb)) & bmask);
    }
    __syncthreads();
    uint32_t* cw = counts + (size_t)w * B + ((size_t)p << sh);
    for (uint32_t b = threadIdx.x; b < nb; b += BUCKET_THREADS)
        if (hist[b]) atomicAdd(&cw[b], hist[b]);
}

// Places a slice: its run of each bucket is reserved with one global atomic on the bucket's
// cursor (positions after the bucket's start), then entries go there as point index | sign << 31.
__global__ __launch_bounds__(BUCKET_THREADS) void msm_bucket_place_kernel(const uint32_t* __restrict__ temp,
                                                                         const uint32_t* __restrict__ pbase,
                                                                         const uint32_t* __restrict__ sbase, uint32_t M,
                                                                         uint32_t wb, uint32_t nw, uint32_t cb,
                                                                         uint32_t plog, uint32_t kb, uint32_t B,
                                                                         const uint32_t* __restrict__ starts,
                                                                         uint32_t* __restrict__ bcur,
                                                                         uint32_t* __restrict__ sorted) {
    __shared__ uint32_t cursor[BUCKET_KEYS];
    __shared__ uint32_t hot_s;
    const PartSlice sl = part_slice(pbase, sbase, M);
    if (!sl.valid) return;
    const uint32_t w = sl.q >> plog, p = sl.q & ((1u << plog) - 1), sh = part_shift(w, wb, nw, cb, plog);
    const uint32_t nb = 1u << sh, bmask = nb - 1, fmask = ((1u << kb) - 1) << (31 - kb);
    for (uint32_t b = threadIdx.x; b < nb; b += BUCKET_THREADS) cursor[b] = 0;
    if (threadIdx.x == 0) hot_s = 0;
    __syncthreads();
    for (uint32_t i = sl.lo + threadIdx.x; i < sl.hi; i += 4 * BUCKET_THREADS) {  // the slice's counts
        uint32_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = i + k * BUCKET_THREADS < sl.hi ? temp[i + k * BUCKET_THREADS] : 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * BUCKET_THREADS < sl.hi) lds_inc(cursor, (x[k] >> (31 - kb)) & bmask);
    }
    __syncthreads();
    const size_t b0 = (size_t)w * B + ((size_t)p << sh);
    for (uint32_t b = threadIdx.x; b < nb; b += BUCKET_THREADS) {
        const uint32_t h = cursor[b];
        if (h) atomicMax(&hot_s, (h << 8) | b);  // the slice's fullest bucket (h <= SLICE)
        cursor[b] = h ? starts[b0 + b] + atomicAdd(&bcur[b0 + b], h) : 0u;
    }
    __syncthreads();
    // (staging this pass through LDS as the partition pass does measured no faster: a slice's
    // writes already stay within one partition's range)
    const uint32_t hot = hot_s & 0xffu;
    for (uint32_t i = sl.lo + threadIdx.x; i < sl.hi; i += 4 * BUCKET_THREADS) {
        uint32_t x[4], pos[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = i + k * BUCKET_THREADS < sl.hi ? temp[i + k * BUCKET_THREADS] : 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * BUCKET_THREADS < sl.hi) pos[k] = lds_fetch_inc(cursor, (x[k] >> (31 - kb)) & bmask, hot);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * BUCKET_THREADS < sl.hi) sorted[pos[k]] = x[k] & ~fmask;
    }
}

// ---------------------------------------------------------------- scan of counts and items
// Three kernels: per-block exclusive scans (4096 slots per block), a single-block scan of the
// block totals, and the add-back of block offsets. Both scans (points and items) run together.

static constexpr int SCAN_THREADS = 1024;
static constexpr int SCAN_PER_THREAD = 4;
static constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_PER_THREAD;

// Work-item slots of a bucket of `count` sorted entries: the most lanes of ch consecutive entries
// of the whole sorted list that a run of `count` entries can meet, ceil((count - 1) / ch) + 1 (the
// G1 accumulation's lanes are cut at multiples of ch, not at bucket starts); kernels that cut
// items at bucket starts use ceil(count / ch) of them and mark the spare slot empty.
__device__ __forceinline__ uint32_t items_of(uint32_t count, uint32_t ch, uint32_t lane_cut) {
    return !count ? 0u : lane_cut ? (count + ch - 2) / ch + 1 : (count + ch - 1) / ch;
}

__global__ __launch_bounds__(SCAN_THREADS) void scan_block_kernel(const uint32_t* __restrict__ counts, uint32_t S,
                                                                 uint32_t ch, uint32_t lane_cut,
                                                                 uint32_t* __restrict__ starts,
                                                                 uint32_t* __restrict__ istarts,
                                                                 uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t sp[SCAN_THREADS], si[SCAN_THREADS];
    __shared__ uint32_t bmax;
    const uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
    uint32_t cp[SCAN_PER_THREAD], tp = 0, ti = 0, mi = 0;
    if (threadIdx.x == 0) bmax = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        cp[k] = base + k < S ? counts[base + k] : 0;
        tp += cp[k];
        ti += items_of(cp[k], ch, lane_cut);
        mi = max(mi, items_of(cp[k], ch, lane_cut));
    }
    sp[threadIdx.x] = tp;
    si[threadIdx.x] = ti;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {  // inclusive Hillis-Steele scan
        uint32_t ap = 0, ai = 0;
        if ((int)threadIdx.x >= off) {
            ap = sp[threadIdx.x - off];
            ai = si[threadIdx.x - off];
        }
        __syncthreads();
        sp[threadIdx.x] += ap;
        si[threadIdx.x] += ai;
        __syncthreads();
    }
    uint32_t rp = sp[threadIdx.x] - tp, ri = si[threadIdx.x] - ti;  // exclusive
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        if (base + k < S) {
            starts[base + k] = rp;
            istarts[base + k] = ri;
        }
        rp += cp[k];
        ri += items_of(cp[k], ch, lane_cut);
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        block_sums[2 * blockIdx.x] = sp[threadIdx.x];
        block_sums[2 * blockIdx.x + 1] = si[threadIdx.x];
    }
    // the most work items of any slot in this block (block_sums[2 nb + block])
    atomicMax(&bmax, mi);
    __syncthreads();
    if (threadIdx.x == 0) block_sums[2 * gridDim.x + blockIdx.x] = bmax;
}

// Exclusive scan of `nb` (points, items) pairs in place; grand totals go to starts[S], istarts[S].
__global__ __launch_bounds__(SCAN_THREADS) void scan_top_kernel(uint32_t* __restrict__ block_sums, uint32_t nb,
                                                               uint32_t S, uint32_t* __restrict__ starts,
                                                               uint32_t* __restrict__ istarts) {
    __shared__ uint32_t sp[SCAN_THREADS], si[SCAN_THREADS];
    __shared__ uint32_t smax;
    const uint32_t per = (nb + SCAN_THREADS - 1) / SCAN_THREADS;
    const uint32_t lo = threadIdx.x * per, hi = min(lo + per, nb);
    uint32_t tp = 0, ti = 0, mi = 0;
    if (threadIdx.x == 0) smax = 0;
    for (uint32_t k = lo; k < hi; ++k) {
        tp += block_sums[2 * k];
        ti += block_sums[2 * k + 1];
        mi = max(mi, block_sums[2 * nb + k]);
    }
    sp[threadIdx.x] = tp;
    si[threadIdx.x] = ti;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        uint32_t ap = 0, ai = 0;
        if ((int)threadIdx.x >= off) {
            ap = sp[threadIdx.x - off];
            ai = si[threadIdx.x - off];
        }
        __syncthreads();
        sp[threadIdx.x] += ap;
        si[threadIdx.x] += ai;
        __syncthreads();
    }
    uint32_t rp = sp[threadIdx.x] - tp, ri = si[threadIdx.x] - ti;
    for (uint32_t k = lo; k < hi; ++k) {
        const uint32_t p = block_sums[2 * k], q = block_sums[2 * k + 1];
        block_sums[2 * k] = rp;
        block_sums[2 * k + 1] = ri;
        rp += p;
        ri += q;
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        starts[S] = sp[threadIdx.x];
        istarts[S] = si[threadIdx.x];
    }
    // block_sums[3 nb] = the most work items of any slot: merge rounds that no bucket needs exit
    atomicMax(&smax, mi);
    __syncthreads();
    if (threadIdx.x == 0) block_sums[3 * nb] = smax;
}

__global__ __launch_bounds__(256) void scan_add_kernel(const uint32_t* __restrict__ block_sums, uint32_t S,
                                                      uint32_t* __restrict__ starts, uint32_t* __restrict__ istarts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const uint32_t b = i / SCAN_BLOCK;
    starts[i] += block_sums[2 * b];
    istarts[i] += block_sums[2 * b + 1];
}

// ---------------------------------------------------------------- plan

uint32_t msm_plan_item_points(int group, size_t n, uint32_t c) { return msm_shape(group, n, c, 0).CH; }

MsmShape msm_shape(int group, size_t npts, uint32_t c, uint32_t item_points, uint32_t sort_points,
                   bool precomputed) {
    if (npts < 1) throw Error(ZKG_ERR_INVALID, "an MSM needs at least one point");
    if (c < 9 || c > 20) throw Error(ZKG_ERR_INVALID, "window size c must be in 9..=20, got " + std::to_string(c));
    if (group != 1 && group != 2) throw Error(ZKG_ERR_INVALID, "group must be G1 (1) or G2 (2)");
    if (npts > 0x7fffffffu) throw Error(ZKG_ERR_INVALID, "an MSM holds at most 2^31 - 1 points");
    MsmShape sh;
    sh.pre = precomputed;
    sh.glv = msm_glv_enabled(npts);
    if (sh.pre) {
        if (c > 16)
            throw Error(ZKG_ERR_INVALID, "a plan over precomputed bases takes c in 9..=16, got " + std::to_string(c));
        // 4n virtual point indices must fit the 31-bit index field of a sorted entry
        if (!sh.glv || npts > 0x1fffffffu)
            throw Error(ZKG_ERR_INVALID, "a plan over precomputed bases needs the GLV split and n < 2^29");
    }
    // GLS for G2 at every c <= 16 (the windows of the LDS-sort path) and n <= 2^18: the accumulation
    // gathers from a table of 2n points and 2n phi x-coordinates (384 bytes per base, twice the GLV
    // plan's), and beyond 2^18 bases those gathers fall out of the caches (2^19 at c = 14: 5.63 ms
    // against 5.02 for the GLV plan; 2^20 accumulate 8.99 ms against 6.49).
    sh.gls = group == 2 && sh.glv && !sh.pre && c <= 16 && npts <= ((size_t)1 << 18) && msm_gls_enabled();
    const uint64_t n = npts;
    const uint64_t nv = sh.pre || sh.gls ? 4 * n : sh.glv ? 2 * n : n;
    // bits per (half-)scalar the windows cover: |k| < 2^63.7 for GLS, so 65 bits of signed digits
    const uint32_t sbits = sh.pre ? 64 : sh.gls ? 65 : 128;
    sh.nv = (uint32_t)nv;
    // GLS windows: the balanced layouts of 65 bits have 9, 10, 11 or 13 digit bits (8, 7, 6 or 5
    // windows); c up to 11 takes its own width, c from 12 on five windows of 13 bits. Measured (G2,
    // uniform scalars, against the GLV plan at the same c): 2^12 c = 10 0.96 vs 1.23 ms, 2^16
    // c = 13 1.44 vs 1.71, 2^17 c = 12 2.03 vs 2.45 (with six windows of 11 bits 2.72: the reason
    // c = 12 takes the 13-bit layout), 2^18 c = 13 2.94 vs 3.19, c = 15 2.91 vs 3.25, c = 16 2.93
    // vs 3.74.
    sh.W = sh.gls ? (c <= 11 ? (sbits + c - 1) / c : 5) : sh.glv ? (sbits + c - 1) / c : (255 + c - 1) / c;
    sh.wb = sh.glv ? sbits / sh.W : c;
    sh.nw = sh.glv ? sbits - sh.wb * sh.W : 0;
    const uint32_t cw = sh.gls ? sh.wb + (sh.nw ? 1 : 0) : c;
    sh.cw = cw;
    sh.B = 1u << (cw - 1);
    // sorted entries are addressed by 32-bit positions (the slot scan's starts) and hold W of
    // them per virtual point
    if ((uint64_t)sh.W * nv > 0xffffffffull)
        throw Error(ZKG_ERR_INVALID, "an MSM of " + std::to_string(n) + " points at c = " + std::to_string(c) +
                                         " needs " + std::to_string((uint64_t)sh.W * nv) +
                                         " sorted entries; at most 2^32 - 1 fit one plan");
    // GLS: the item size of the GLV plan at the same c. Its buckets hold twice the points, but the
    // rule's larger items halve the work items, and the G2 accumulation (two waves per SIMD) then
    // no longer fills the GPU (G2 2^18, c = 13: accumulate 2.50 ms at 64 points per item, 1.91 at 32).
    sh.CH = item_points ? item_points : msm_auto_item_points(sh.gls ? nv / 2 : nv, cw);
    sh.L = msm_reduce_chunk(group, cw, sh.W);
    sh.T = sh.B / sh.L;
    sh.aw = group == 1 ? 16 : 32;
    sh.xw = 2 * sh.aw;
    const uint64_t S = (uint64_t)sh.W * sh.B;
    // slots: sum over buckets of ceil((count - 1) / CH) + 1 <= W nv / CH + 2 W B
    sh.max_items = (size_t)sh.W * ((nv + sh.CH - 1) / sh.CH + 2 * (size_t)sh.B);
    if (sh.max_items >= 0xffffffffull) throw Error(ZKG_ERR_INVALID, "too many work items for one plan");
    // the most slots one bucket can hold (all nv virtual points in one bucket)
    for (uint64_t m = (nv + sh.CH - 2) / sh.CH + 1; m > FINAL_GROUP; m = (m + SUM_GROUP - 1) / SUM_GROUP)
        ++sh.merge_rounds;
    sh.scan_blocks = (S + SCAN_BLOCK - 1) / SCAN_BLOCK;
    sh.lds_sort = cw <= 16;
    // the GLS digits exist only as the stored digits of the LDS-sort path
    if (sh.gls && !sh.lds_sort) throw Error(ZKG_ERR_INTERNAL, "GLS plan without the LDS sort");
    sh.sort_pts = sort_points ? std::max<uint32_t>(sort_points, sh.B) : msm_default_sort_points(nv, cw);
    sh.nblk = (uint32_t)((nv + sh.sort_pts - 1) / sh.sort_pts);
    return sh;
}

// Bytes of the allocations MsmPlan::MsmPlan makes for this shape (n real points).
size_t MsmShape::workspace_bytes(size_t n) const {
    const size_t S = (size_t)W * B;
    const size_t words = (glv && !gls ? 8 * n : 0)                            // split
                         + (glv ? (pre || gls ? 2 : 1) * n * xw / 2 : 0)      // endo_x
                         + (gls ? 2 * n * aw : 0)                             // gls_bases
                         + 4 * S + 2                                      // counts, cursors, starts, istarts
                         + 3 * scan_blocks + 1                            // scan_tmp
                         + (size_t)W * nv * (lds_sort ? 2 : 1)            // sorted (+ digits)
                         + (lds_sort ? (size_t)W * nblk * B : 0)          // table
                         + max_items * (2 + (size_t)xw)                   // item_info, item_pts
                         + S * xw + (size_t)W * T * xw + (size_t)W * xw + xw;  // buckets, partials, wsums, result
    return 4 * words;
}

MsmPlan::MsmPlan(Device* d, int grp, size_t npts, uint32_t win, uint32_t item_points, uint32_t sort_points,
                 bool precomputed)
    : dev(d), group(grp), pre(precomputed) {
    const MsmShape sh = msm_shape(grp, npts, win, item_points, sort_points, precomputed);
    n = (uint32_t)npts;
    c = win;
    glv = sh.glv;
    gls = sh.gls;
    cw = sh.cw;
    f29 = msm_f29_enabled();
    lanes = msm_lanes_enabled((uint64_t)sh.W * sh.nv);
    part_bits = msm_part_bits(sh, &part_log);
    nv = sh.nv;
    W = sh.W;
    wb = sh.wb;
    nw = sh.nw;
    B = sh.B;
    CH = sh.CH;
    L = sh.L;
    T = sh.T;
    aw = sh.aw;
    xw = sh.xw;
    max_items = sh.max_items;
    merge_rounds = sh.merge_rounds;
    lds_sort = sh.lds_sort;
    sort_pts = sh.sort_pts;
    nblk = sh.nblk;
    const size_t S = (size_t)W * B;
    const size_t nb = sh.scan_blocks;
    device_bind(d);
    if (lds_sort) {
        // 4 * B bytes of LDS per sort block, up to 128 KB at c = 16. The limit is a property of
        // the kernel, shared by every plan: always raise it to the largest window's need, so a
        // plan with a small window never lowers it under a plan with a large one.
        const int lds_max = 4 << 15;
        ZKG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(msm_hist_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
        ZKG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(msm_place_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    }
    try {
        ZKG_HIP(hipMalloc(&counts, 4 * S));
        ZKG_HIP(hipMalloc(&cursors, 4 * S));
        ZKG_HIP(hipMalloc(&starts, 4 * (S + 1)));
        ZKG_HIP(hipMalloc(&istarts, 4 * (S + 1)));
        ZKG_HIP(hipMalloc(&scan_tmp, 4 * (3 * (size_t)nb + 1)));
        max_items_dev = scan_tmp + 3 * nb;
        if (glv && !gls) ZKG_HIP(hipMalloc(&split, 32 * (size_t)n));
        if (glv) ZKG_HIP(hipMalloc(&endo_x, 2 * (gls ? 2 * (size_t)n : base_points()) * aw));
        if (gls) ZKG_HIP(hipMalloc(&gls_bases, 4 * 2 * (size_t)n * aw));
        ZKG_HIP(hipMalloc(&sorted, 4 * (size_t)W * nv));
        if (lds_sort) {
            ZKG_HIP(hipMalloc(&digits, 4 * (size_t)W * nv));
            ZKG_HIP(hipMalloc(&table, 4 * (size_t)W * nblk * B));
        }
        ZKG_HIP(hipMalloc(&item_info, 8 * max_items));
        ZKG_HIP(hipMalloc(&item_pts, 4 * max_items * xw));
        ZKG_HIP(hipMalloc(&buckets, 4 * S * xw));
        ZKG_HIP(hipMalloc(&partials, 4 * (size_t)W * T * xw));
        ZKG_HIP(hipMalloc(&wsums, 4 * (size_t)W * xw));
        ZKG_HIP(hipMalloc(&result, 4 * (size_t)xw));
        // Zeroed before the constructor returns: a plan may next run on another stream (a batch's
        // or a server slot's), which nothing would order after a queued clear.
        ZKG_HIP(hipMemsetAsync(result, 0, 4 * (size_t)xw, d->stream));
        ZKG_HIP(hipStreamSynchronize(d->stream));
    } catch (...) {
        this->~MsmPlan();
        throw;
    }
}

// The workspace arrays of a plan in one region: each array's byte size, in a fixed order, each
// starting at a 256-byte boundary. `base` = nullptr only sizes the region.
static size_t carve(const MsmShape& sh, size_t n, uint8_t* base, MsmPlan* p) {
    const size_t S = (size_t)sh.W * sh.B;
    size_t off = 0;
    auto take = [&](size_t bytes) -> uint8_t* {
        uint8_t* at = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    uint32_t* counts = (uint32_t*)take(4 * S);
    uint32_t* cursors = (uint32_t*)take(4 * S);
    uint32_t* starts = (uint32_t*)take(4 * (S + 1));
    uint32_t* istarts = (uint32_t*)take(4 * (S + 1));
    uint32_t* scan_tmp = (uint32_t*)take(4 * (3 * sh.scan_blocks + 1));
    uint32_t* split = sh.glv && !sh.gls ? (uint32_t*)take(32 * n) : nullptr;
    uint32_t* endo_x = sh.glv ? (uint32_t*)take(2 * (sh.pre || sh.gls ? 2 * n : n) * sh.aw) : nullptr;
    uint32_t* gls_bases = sh.gls ? (uint32_t*)take(4 * 2 * n * sh.aw) : nullptr;
    uint32_t* sorted = (uint32_t*)take(4 * (size_t)sh.W * sh.nv);
    uint32_t* digits = sh.lds_sort ? (uint32_t*)take(4 * (size_t)sh.W * sh.nv) : nullptr;
    uint32_t* table = sh.lds_sort ? (uint32_t*)take(4 * (size_t)sh.W * sh.nblk * sh.B) : nullptr;
    uint2* item_info = (uint2*)take(8 * sh.max_items);
    uint32_t* item_pts = (uint32_t*)take(4 * sh.max_items * sh.xw);
    uint32_t* buckets = (uint32_t*)take(4 * S * sh.xw);
    uint32_t* partials = (uint32_t*)take(4 * (size_t)sh.W * sh.T * sh.xw);
    uint32_t* wsums = (uint32_t*)take(4 * (size_t)sh.W * sh.xw);
    if (p) {
        p->counts = counts;
        p->cursors = cursors;
        p->starts = starts;
        p->istarts = istarts;
        p->scan_tmp = scan_tmp;
        p->max_items_dev = scan_tmp + 3 * sh.scan_blocks;
        p->split = split;
        p->endo_x = endo_x;
        p->gls_bases = gls_bases;
        p->sorted = sorted;
        p->digits = digits;
        p->table = table;
        p->item_info = item_info;
        p->item_pts = item_pts;
        p->buckets = buckets;
        p->partials = partials;
        p->wsums = wsums;
    }
    return off;
}

size_t msm_carved_bytes(const MsmShape& sh, size_t n) { return carve(sh, n, nullptr, nullptr); }

// The static scratch of the no-allocation MSM (zkg_msm_run_noalloc): each GPU the library runs
// on holds its own copy until the process ends. Host code never passes `scratch` itself to a
// kernel or a copy (on the host the name is a stand-in, not a device address): it uses the address
// msm_scratch_address returns for the current device, which device_open records per device.
static __device__ __align__(256) uint8_t scratch[MSM_SCRATCH_BYTES];

void* msm_scratch_address() {
    void* ptr = nullptr;
    ZKG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    return ptr;
}

MsmPlan::MsmPlan(Device* d, int grp, size_t npts, uint32_t win, const MsmShape& sh, void* workspace,
                 uint32_t* result_words)
    : dev(d), group(grp), pre(sh.pre), owned(false) {
    n = (uint32_t)npts;
    c = win;
    glv = sh.glv;
    gls = sh.gls;
    cw = sh.cw;
    f29 = msm_f29_enabled();
    lanes = msm_lanes_enabled((uint64_t)sh.W * sh.nv);
    part_bits = msm_part_bits(sh, &part_log);
    nv = sh.nv;
    W = sh.W;
    wb = sh.wb;
    nw = sh.nw;
    B = sh.B;
    CH = sh.CH;
    L = sh.L;
    T = sh.T;
    aw = sh.aw;
    xw = sh.xw;
    max_items = sh.max_items;
    merge_rounds = sh.merge_rounds;
    lds_sort = sh.lds_sort;
    sort_pts = sh.sort_pts;
    nblk = sh.nblk;
    if (lds_sort) {  // the kernels' LDS limit, as the allocating constructor raises it (no allocation)
        const int lds_max = 4 << 15;
        ZKG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(msm_hist_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
        ZKG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(msm_place_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    }
    carve(sh, npts, static_cast<uint8_t*>(workspace), this);
    result = result_words;
}

MsmPlan::~MsmPlan() {
    if (!owned) return;  // carved from a caller's region: nothing to free
    for (uint32_t* p : {split, endo_x, gls_bases, counts, cursors, starts, istarts, scan_tmp, sorted, digits, table,
                        item_pts, buckets, partials, wsums, result})
        if (p) (void)hipFree(p);
    if (item_info) (void)hipFree(item_info);
    split = endo_x = gls_bases = counts = cursors = starts = istarts = scan_tmp = sorted = digits = table = item_pts = buckets =
        partials = wsums = result = nullptr;
    max_items_dev = nullptr;  // points into scan_tmp
    item_info = nullptr;
}

size_t MsmPlan::workspace_bytes() const { return msm_workspace_bytes(group, n, c, CH, sort_pts, pre); }

bool msm_glv_enabled(size_t n) {
    // virtual point indices (2n of them) must fit the 31-bit index field of a sorted entry
    if (n > 0x3fffffffu) return false;
    const char* e = std::getenv("ZKGPU_MSM_GLV");
    return !(e && e[0] == '0');
}

bool msm_gls_enabled() {
    const char* e = std::getenv("ZKGPU_MSM_GLS");
    return !(e && e[0] == '0');
}

// Points per counting-sort block when none is given: at least one bucket range (B) and 8192, and
// nv / 64 (fewer blocks means a smaller per-block cursor table to build, scan and read; measured
// on MI355X at 2^16 .. 2^22 within 2% of the best block size, which was 2B .. 8B). Sizing the
// blocks so that all W windows' blocks fit the GPU in one round of workgroups measured slower
// (G1 2^20, c = 15: 504 blocks of 37888 points 0.326 ms against 576 of 32768 points 0.288 ms):
// the placement's time follows the points per block.
uint32_t msm_default_sort_points(size_t nv, uint32_t c) {
    const size_t B = (size_t)1 << (c - 1);
    return (uint32_t)std::max<size_t>(std::max<size_t>(B, 8192), std::min<size_t>(nv / 64, (size_t)1 << 20));
}

// Buckets per lane of the bucket reduction: its serial chain is 2L additions plus the workgroup
// tree, against W B / L lanes in all. Measured on MI355X: L = 4 for G1 (2^20: 2.90 ms, 3.00 with
// L = 2; over precomputed bases at c = 16, 131072 buckets, L = 2 and 4 tie; on lane pairs at 2^20,
// c = 15, L = 2 doubles the busy waves to 1408 for 1024 SIMDs: reduce 0.19-0.20 -> 0.28-0.31 ms). G2 runs the
// reduction on four lanes per point, and L = 4 there too (2^18, c = 13: reduce 0.46 -> 0.35 ms
// against L = 2, which put a second wave on some SIMDs; precomputed 2^18 at c = 16: 1.05 -> 0.65
// ms).
uint32_t msm_reduce_chunk(int group, uint32_t c, uint32_t windows) {
    // G2 with at most 2^15 buckets (GLS plans at c <= 13): two buckets per lane keep every busy
    // wave on a SIMD of its own and shorten the chain (GLS 2^18, c = 13: reduce 0.32-0.34 -> 0.29 ms,
    // window sums +0.01)
    if (group == 2 && ((uint64_t)windows << (c - 1)) <= 32768) return 2;
    return 4;
}

uint32_t msm_part_bits(const MsmShape& sh, uint32_t* part_log) {
    *part_log = 0;
    if (!sh.lds_sort) return 0;
    const char* e = std::getenv("ZKGPU_MSM_PARTSORT");
    if (e && e[0] == '0') return 0;
    uint32_t cb = 0, lgnv = 0;  // bucket bits (B = 2^cb); point-index bits (nv <= 2^lgnv)
    while ((1u << cb) < sh.B) ++cb;
    while ((1ull << lgnv) < sh.nv) ++lgnv;
    // 2^8 partitions per window (at least four buckets each), more where the point index needs the
    // bits of the bucket-within-partition field (G1 2^24 at c = 16: 2^9)
    uint32_t plog = std::min<uint32_t>(8, cb - 2);
    if (cb - plog + lgnv > 31) plog = cb + lgnv - 31;
    if (plog > cb || (1u << plog) > PART_MAX) return 0;
    const uint32_t kb = cb - plog;  // bucket bits within a partition, at most 8
    if (kb > 8 || (uint64_t)sh.nv > (1ull << (31 - kb))) return 0;
    // bcount, pbase (M + 1), pcur (M) and sbase (M + 1) must fit the plan's table of W nblk B words
    const uint64_t M = (uint64_t)sh.W << plog;
    if (M * sh.nblk + 3 * M + 2 > (uint64_t)sh.W * sh.nblk * sh.B) return 0;
    *part_log = plog;
    return kb;
}

bool msm_lanes_enabled(uint64_t entries) {
    const char* e = std::getenv("ZKGPU_MSM_LANES");
    if (e && e[0] == '1') return true;
    if (e && e[0] == '0') return false;
    return entries >= (1ull << 23);
}

bool msm_f29_enabled() {
    const char* e = std::getenv("ZKGPU_MSM_F29");
    return !(e && e[0] == '0');
}

size_t msm_workspace_bytes(int group, size_t n, uint32_t c, uint32_t item_points, uint32_t sort_points,
                           bool precomputed) {
    return msm_shape(group, n, c, item_points, sort_points, precomputed).workspace_bytes(n);
}

bool poison_enabled() {
    const char* e = std::getenv("ZKGPU_POISON");
    return e && e[0] == '1';
}

// Debug mode (ZKGPU_POISON=1): before a run, every workspace array of the plan and its result are
// filled with 0xA5A5A5A5 by a kernel, so a stage that reads a word no earlier stage of the same
// run wrote gives a wrong result instead of a stale right one.
static void poison_workspace(const MsmPlan& p, hipStream_t s) {
    const size_t S = (size_t)p.W * p.B, nb = (S + SCAN_BLOCK - 1) / SCAN_BLOCK;
    const struct {
        void* ptr;
        size_t bytes;
    } arrays[] = {
        {p.counts, 4 * S}, {p.cursors, 4 * S}, {p.starts, 4 * (S + 1)}, {p.istarts, 4 * (S + 1)},
        {p.scan_tmp, 4 * (3 * nb + 1)}, {p.split, p.glv ? 32 * (size_t)p.n : 0},
        {p.endo_x, p.glv ? 2 * (p.gls ? 2 * (size_t)p.n : p.base_points()) * p.aw : 0},
        {p.gls_bases, p.gls ? 4 * 2 * (size_t)p.n * p.aw : 0}, {p.sorted, 4 * (size_t)p.W * p.nv},
        {p.digits, p.lds_sort ? 4 * (size_t)p.W * p.nv : 0}, {p.table, p.lds_sort ? 4 * (size_t)p.W * p.nblk * p.B : 0},
        {p.item_info, 8 * p.max_items}, {p.item_pts, 4 * p.max_items * p.xw}, {p.buckets, 4 * S * p.xw},
        {p.partials, 4 * (size_t)p.W * p.T * p.xw}, {p.wsums, 4 * (size_t)p.W * p.xw}, {p.result, 4 * (size_t)p.xw}};
    for (const auto& a : arrays)
        if (a.ptr && a.bytes) fill_u32(static_cast<uint32_t*>(a.ptr), 0xA5A5A5A5u, a.bytes / 4, s);
    ZKG_HIP(hipGetLastError());
}

template <class Fe>
// prof: wait around every stage and record its time. yield_heavy: wait after the heavy stages
// (accumulate, finalize, reduce, combine) so that other users of the GPU get it in between
// (the reference's run_yielding, src/msm.rs:300-318).
static void run_stages(const MsmPlan& p, const uint32_t* bases, const uint32_t* scalars, hipStream_t s,
                       std::vector<std::pair<std::string, double>>* prof, bool yield_heavy = false) {
    const uint32_t S = p.W * p.B;
    const uint32_t nb = (S + SCAN_BLOCK - 1) / SCAN_BLOCK;
    auto t0 = std::chrono::steady_clock::now();
    auto stage = [&](const char* name) {
        ZKG_HIP(hipGetLastError());
        if (yield_heavy && (!strcmp(name, "accumulate") || !strcmp(name, "finalize") ||
                            !strcmp(name, "reduce chunks") || !strcmp(name, "combine")))
            ZKG_HIP(hipStreamSynchronize(s));
        if (!prof) return;
        ZKG_HIP(hipStreamSynchronize(s));
        auto t1 = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        bool found = false;
        for (auto& e : *prof)
            if (e.first == name) {
                e.second += ms;
                found = true;
            }
        if (!found) prof->emplace_back(name, ms);
        t0 = t1;
    };
    if (poison_enabled()) poison_workspace(p, s);
    if (prof) {
        ZKG_HIP(hipStreamSynchronize(s));
        t0 = std::chrono::steady_clock::now();
    }
    // GLS: the table of the bases and their psi images, and phi of both; the stages then read the
    // table. GLV: split every scalar into two signed 127-bit halves and compute phi(P_i) = (beta x_i, y_i).
    const uint32_t* dscal = scalars;
    // the plain GLV plan on the LDS sort path splits, writes its digits and phi(x) in one kernel
    const bool glv_fused = p.glv && !p.pre && !p.gls && p.lds_sort;
    if (p.gls) {
        hipLaunchKernelGGL(msm_gls_prepare_kernel, dim3(ceil_div(p.n, 256)), dim3(256), 0, s, bases, p.n, p.gls_bases,
                           p.endo_x);
        bases = p.gls_bases;
    } else if (glv_fused) {
        // below
    } else if (p.glv) {
        hipLaunchKernelGGL(msm_glv_split_kernel, dim3(ceil_div(p.n, 256)), dim3(256), 0, s, scalars, p.n, p.split);
        const size_t nbase = p.base_points();
        hipLaunchKernelGGL(msm_endo_kernel, dim3(ceil_div(nbase * p.group, 256)), dim3(256), 0, s, bases,
                           (uint32_t)nbase, p.group, p.endo_x);
        dscal = p.split;
    }
    const unsigned pg = ceil_div(p.nv, 256);
    // two-pass placement: the digits go to `sorted`'s buffer, the partition runs to `digits`'s,
    // and the final list to `sorted` (both W * nv words)
    const uint32_t kb = p.part_bits, plog = p.part_log, NP = kb ? 1u << plog : 0;
    const uint32_t cb = p.cw - 1;  // bucket bits of a full window
    uint32_t* dig = kb ? p.sorted : p.digits;
    if (p.lds_sort) {
        if (p.gls)
            hipLaunchKernelGGL(msm_gls_digits_store_kernel, dim3(ceil_div(p.n, 256)), dim3(256), 0, s, scalars, p.n,
                               p.W, p.wb, p.nw, dig);
        else if (glv_fused)
            hipLaunchKernelGGL(msm_glv_digits_kernel, dim3(ceil_div(p.n, 256)), dim3(256), 0, s, scalars, bases, p.n,
                               p.group, p.W, p.wb, p.nw, p.endo_x, dig);
        else if (p.pre)
            hipLaunchKernelGGL(msm_digits_split_store_kernel, dim3(ceil_div(2 * (size_t)p.n, 256)), dim3(256), 0, s,
                               dscal, p.n, p.W, p.wb, p.nw, dig);
        else if (p.glv)
            hipLaunchKernelGGL(msm_digits_store_kernel<true>, dim3(pg), dim3(256), 0, s, dscal, p.nv, p.W, p.wb,
                               p.nw, dig);
        else
            hipLaunchKernelGGL(msm_digits_store_kernel<false>, dim3(pg), dim3(256), 0, s, dscal, p.nv, p.W, p.wb,
                               p.nw, dig);
        if (kb) {
            const uint32_t M = p.W * NP;  // partitions in all windows
            uint32_t* bcount = p.table;
            uint32_t* pbase = bcount + (size_t)M * p.nblk;
            uint32_t* pcur = pbase + M + 1;
            uint32_t* sbase = pcur + M;
            const unsigned slices = ceil_div((size_t)p.W * p.nv, SLICE) + M;  // at most
            fill_u32(pbase, 0, M, s);
            fill_u32(p.counts, 0, S, s);
            fill_u32(p.cursors, 0, S, s);
            hipLaunchKernelGGL(msm_part_hist_kernel, dim3(p.nblk, p.W), dim3(SORT_THREADS), 0, s, dig, p.nv,
                               p.sort_pts, p.nblk, p.wb, p.nw, cb, plog, bcount, pbase);
            hipLaunchKernelGGL(msm_part_scan_kernel, dim3(1), dim3(1024), 0, s, pbase, M, pcur, sbase);
            hipLaunchKernelGGL(msm_part_place_kernel, dim3(p.nblk, p.W), dim3(SORT_THREADS), 0, s, dig, p.nv,
                               p.sort_pts, p.nblk, p.wb, p.nw, cb, plog, kb, (const uint32_t*)bcount, pcur, p.digits);
            hipLaunchKernelGGL(msm_bucket_count_kernel, dim3(slices), dim3(BUCKET_THREADS), 0, s,
                               (const uint32_t*)p.digits, (const uint32_t*)pbase, (const uint32_t*)sbase, M, p.wb,
                               p.nw, cb, plog, kb, p.B, p.counts);
        } else {
            hipLaunchKernelGGL(msm_hist_kernel, dim3(p.nblk, p.W), dim3(SORT_THREADS), 4 * p.B, s, dig, p.nv, p.B,
                               p.sort_pts, p.nblk, p.table);
            hipLaunchKernelGGL(msm_colscan_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, s, p.table, p.W, p.B,
                               p.nblk, p.counts);
        }
    } else {
        fill_u32(p.counts, 0, S, s);
        fill_u32(p.cursors, 0, S, s);
        if (p.glv)
            hipLaunchKernelGGL((msm_digits_kernel<false, true>), dim3(pg), dim3(256), 0, s, dscal, p.nv, p.cw, p.W,
                               p.wb, p.nw, p.counts, (const uint32_t*)nullptr, (uint32_t*)nullptr);
        else
            hipLaunchKernelGGL((msm_digits_kernel<false, false>), dim3(pg), dim3(256), 0, s, dscal, p.nv, p.cw, p.W,
                               p.wb, p.nw, p.counts, (const uint32_t*)nullptr, (uint32_t*)nullptr);
    }
    stage("count");
    hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(SCAN_THREADS), 0, s, p.counts, S, p.CH,
                       p.lane_cut() ? 1u : 0u, p.starts, p.istarts, p.scan_tmp);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(SCAN_THREADS), 0, s, p.scan_tmp, nb, S, p.starts, p.istarts);
    hipLaunchKernelGGL(scan_add_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, s, p.scan_tmp, S, p.starts, p.istarts);
    stage("scan");
    if (p.lds_sort && kb) {
        const uint32_t M = p.W * NP;
        const uint32_t* pbase = p.table + (size_t)M * p.nblk;
        const uint32_t* sbase = pbase + 2 * M + 1;
        const unsigned slices = ceil_div((size_t)p.W * p.nv, SLICE) + M;
        hipLaunchKernelGGL(msm_bucket_place_kernel, dim3(slices), dim3(BUCKET_THREADS), 0, s,
                           (const uint32_t*)p.digits, pbase, sbase, M, p.wb, p.nw, cb, plog, kb, p.B,
                           (const uint32_t*)p.starts, p.cursors, p.sorted);
    } else if (p.lds_sort) {
        hipLaunchKernelGGL(msm_place_kernel, dim3(p.nblk, p.W), dim3(SORT_THREADS), 4 * p.B, s, p.digits, p.nv, p.B,
                           p.sort_pts, p.nblk, p.table, p.starts, p.sorted);
    } else if (p.glv) {
        hipLaunchKernelGGL((msm_digits_kernel<true, true>), dim3(pg), dim3(256), 0, s, dscal, p.nv, p.cw, p.W,
                           p.wb, p.nw, p.cursors, p.starts, p.sorted);
    } else {
        hipLaunchKernelGGL((msm_digits_kernel<true, false>), dim3(pg), dim3(256), 0, s, dscal, p.nv, p.cw, p.W,
                           p.wb, p.nw, p.cursors, p.starts, p.sorted);
    }
    stage("scatter");
    launch_accumulate<Fe>(p, bases, s);
    stage("accumulate");
    static const char* merge_names[6] = {"merge 1", "merge 2", "merge 3", "merge 4", "merge 5", "merge 6"};
    uint32_t stride = 1;
    for (uint32_t r = 0; r < p.merge_rounds; ++r, stride *= SUM_GROUP) {
        launch_merge<Fe>(p, stride, s);
        stage(merge_names[std::min<uint32_t>(r, 5)]);
    }
    launch_finalize<Fe>(p, s);
    stage("finalize");
    launch_reduce<Fe>(p, s);
    stage("reduce chunks");
    launch_tree<Fe>(p, s);
    stage("tree sum");
    launch_combine<Fe>(p, s);
    stage("combine");
}

void MsmPlan::run(const uint32_t* bases, const uint32_t* scalars, hipStream_t stream,
                  std::vector<std::pair<std::string, double>>* stage_ms) const {
    if (group == 2 && f29) run_stages<Fq2_29>(*this, bases, scalars, stream, stage_ms);
    else if (group == 2) run_stages<Fq2>(*this, bases, scalars, stream, stage_ms);
    else if (f29) run_stages<Fq29>(*this, bases, scalars, stream, stage_ms);
    else run_stages<Fq>(*this, bases, scalars, stream, stage_ms);
}

void MsmPlan::run_yielding(const uint32_t* bases, const uint32_t* scalars, hipStream_t stream) const {
    if (group == 2 && f29) run_stages<Fq2_29>(*this, bases, scalars, stream, nullptr, true);
    else if (group == 2) run_stages<Fq2>(*this, bases, scalars, stream, nullptr, true);
    else if (f29) run_stages<Fq29>(*this, bases, scalars, stream, nullptr, true);
    else run_stages<Fq>(*this, bases, scalars, stream, nullptr, true);
}

}  // namespace zkg
