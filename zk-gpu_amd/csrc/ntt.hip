// NTT kernels for gfx950 and the plan that sequences them.
//
// Algorithm (same mathematics as the reference's ntt.wgsl:34-74, re-laid out for CDNA4):
// radix-2 decimation-in-frequency, log_n stages split into P passes of at most TILE_LOG stages.
// A pass loads a tile of 2^TILE_LOG elements into LDS (up to 128 KB of the 160 KB per CU), runs
// its stages there two at a time (radix-4 rounds: each lane keeps four elements in registers for
// two stages, so one LDS round trip and one barrier per two stages), and writes the tile back.
// Elements stay in [0, 2r) between the first load and the last store (lazy reduction), which
// removes the final subtraction from every butterfly's Montgomery product. Element (row, column)
// of a tile sits at  ((c >> lm) << (lm + s)) + (c & (2^lm - 1)) + (row << lm)  for global column
// c, so columns with consecutive c are adjacent in memory: loads/stores of a tile are runs of
// whole 32-byte elements, coalesced across the wavefront.
//
// The DIF output is bit-reversed. Instead of a separate permutation, the last pass (lm = 0,
// tiles are contiguous blocks) picks its blocks so that block ids are consecutive after bit
// reversal: block c = bitrev(wg*G + col) lands at  bitrev_s(row) * 2^(log_n - s) + wg*G + col,
// so each output row is again a coalesced run of G elements. That pass is the only out-of-place
// one (its write set differs from its read set); all earlier passes run in place, so the plan
// needs one n-element scratch buffer regardless of the pass count.
//
// Twiddles w^i (i < n/2) are precomputed in Montgomery form while data stays plain, so the
// Montgomery product (a - b) * tw yields the plain (a - b) * w^i. The inverse transform uses
// w^-i twiddles and multiplies by n^-1 in the last pass's store.
#include "field.hpp"
#include "field29.hpp"
#include "ntt.hpp"
#include "tune.hpp"
#include "host_math.hpp"

#include <algorithm>

namespace zkg {

__device__ __forceinline__ uint32_t bitrev_bits(uint32_t x, uint32_t bits) {
    return bits == 0 ? 0u : (__builtin_bitreverse32(x) >> (32u - bits));
}

// Tile slot -> global element index (non-last passes; in place).
__device__ __forceinline__ size_t tile_addr(uint32_t wg, uint32_t row, uint32_t col, uint32_t lm,
                                            uint32_t s, uint32_t lg) {
    const size_t c = ((size_t)wg << lg) + col;
    return ((c >> lm) << (lm + s)) + (c & ((1ull << lm) - 1)) + ((size_t)row << lm);
}

// ---- lazy reduction: between the first load and the last store, elements live in [0, 2r).
// x + y < 4r and x + 2r - y < 4r both fit 256 bits (r < 2^254). A Montgomery product whose
// first factor is < 4r and second (a twiddle or n^-1) < r ends below 2r (mul_lazy states the
// bound), so it needs no final subtraction.

// word i of 2r
ZKG_DEV uint32_t twice_r(int i) { return 2u * FrParams::mod(i) + (i ? FrParams::mod(i - 1) >> 31 : 0u); }

// (x + y) reduced to [0, 2r)
ZKG_DEV Fr add_lazy(const Fr& x, const Fr& y) {
    Fr s, d;
    unsigned c = 0, c2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s.w[i] = __builtin_addc(x.w[i], y.w[i], c, &c2);
        c = c2;
    }
    unsigned b = 0, b2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        d.w[i] = __builtin_subc(s.w[i], twice_r(i), b, &b2);
        b = b2;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) s.w[i] = b ? s.w[i] : d.w[i];
    return s;
}

// x in [0, 4r) -> [0, 2r)
ZKG_DEV Fr reduce_lazy(const Fr& x) {
    Fr d, s = x;
    unsigned b = 0, b2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        d.w[i] = __builtin_subc(s.w[i], twice_r(i), b, &b2);
        b = b2;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) s.w[i] = b ? s.w[i] : d.w[i];
    return s;
}

// x + 2r - y in (0, 4r)
ZKG_DEV Fr sub_lazy(const Fr& x, const Fr& y) {
    Fr s;
    unsigned c = 0, c2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s.w[i] = __builtin_addc(x.w[i], twice_r(i), c, &c2);
        c = c2;
    }
    unsigned b = 0, b2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s.w[i] = __builtin_subc(s.w[i], y.w[i], b, &b2);
        b = b2;
    }
    return s;
}

// Montgomery a * b * 2^-256 for a < 4r, b < r (R = 2^256 form, as the twiddle and scale tables
// hold it), computed in nine 29-bit limbs: a's limbs are a itself, b's words shifted left by five
// bits are b in R = 2^261 form (value 32 b < 32 r), and the 29-bit product
// a (32 b) 2^-261 = a b 2^-256 needs v(a) v(32 b) < 128 r^2 < r 2^261 (2^261 / r is about 169),
// then returns a normalised value below 128 r^2 / 2^261 + r < 2r, packed back into eight words.
// With the conversions about 290 instructions per product against about 390 for a CIOS product in
// 32-bit limbs (which needs an add-with-carry and a move besides each v_mad_u64_u32). A twiddle
// shared by two products is converted once (the compiler merges the identical conversions).
ZKG_DEV Fr mul_lazy(const Fr& a, const Fr& b) {
    const F29 t = f29_mul<Fr29Mod>(f29_unpack(a.w), f29_from_r256_words(b.w));
    Fr r;
    f29_to_words(t, r.w);
    return r;
}

ZKG_DEV Fr tile_get(const uint4* tile, uint32_t slot) {
    const uint4 x0 = tile[2 * slot], x1 = tile[2 * slot + 1];
    Fr x;
    x.w[0] = x0.x; x.w[1] = x0.y; x.w[2] = x0.z; x.w[3] = x0.w;
    x.w[4] = x1.x; x.w[5] = x1.y; x.w[6] = x1.z; x.w[7] = x1.w;
    return x;
}

ZKG_DEV void tile_put(uint4* tile, uint32_t slot, const Fr& x) {
    tile[2 * slot] = make_uint4(x.w[0], x.w[1], x.w[2], x.w[3]);
    tile[2 * slot + 1] = make_uint4(x.w[4], x.w[5], x.w[6], x.w[7]);
}

// One pass on workgroup `wg`'s tile; `tile` is the workgroup's LDS (2 uint4 per element).
template <int TILE_LOG, int THREADS>
__device__ __forceinline__ void ntt_pass_body(const NttPassArgs& a, const uint32_t wg, uint4* __restrict__ tile) {
    const uint32_t s = a.s, lm = a.lm, lg = a.lg, L = a.log_n;
    const uint32_t tlog = s + lg;  // log2 of elements in this tile (<= TILE_LOG)
    const uint32_t count = 1u << tlog;
    const uint32_t smask = (1u << s) - 1, gmask = (1u << lg) - 1;

    // ---- load (coalesced: consecutive e -> consecutive global addresses)
    if (a.last) {
        for (uint32_t e = threadIdx.x; e < count; e += THREADS) {
            const uint32_t row = e & smask, col = e >> s;
            const size_t c = bitrev_bits((wg << lg) + col, L - s);
            const uint4* p = reinterpret_cast<const uint4*>(a.src + 8 * ((c << s) + row));
            const uint32_t slot = (row << lg) + col;
            tile[2 * slot] = p[0];
            tile[2 * slot + 1] = p[1];
        }
    } else {
        const uint32_t gj = min(lg, lm);  // log2 of adjacent columns sharing one row run
        for (uint32_t e = threadIdx.x; e < count; e += THREADS) {
            const uint32_t lo = e & ((1u << gj) - 1), row = (e >> gj) & smask, hi = e >> (gj + s);
            const uint32_t col = (hi << gj) + lo;
            const uint4* p = reinterpret_cast<const uint4*>(a.src + 8 * tile_addr(wg, row, col, lm, s, lg));
            const uint32_t slot = (row << lg) + col;
            tile[2 * slot] = p[0];
            tile[2 * slot + 1] = p[1];
        }
    }
    __syncthreads();

    // ---- s DIF stages in LDS: pairs of stages as radix-4 rounds (four elements per lane, two
    // stages in registers, one barrier), after one radix-2 round when s is odd. Stage st pairs
    // rows that differ in bit lh = s-1-st of the row index; its twiddle exponent is
    // ((row mod 2^lh) * 2^lm + j0) * 2^(L-1-lh-lm), with j0 the tile column's offset inside its
    // group of 2^lm columns. In the last pass lm = j0 = 0, so stage lh = 0 has exponent 0 and
    // stage lh = 1 has exponents 0 and 2^(L-2): the final round multiplies only the rows with
    // exponent 2^(L-2), and every product by w^0 = 1 becomes a reduction to [0, 2r).
    const size_t lmmask = (1ull << lm) - 1;
    // 1024-lane tiles stay within 128 VGPRs with radix-2 rounds only
    constexpr bool RADIX4 = THREADS <= 512;
    const uint32_t r4_from = RADIX4 ? (s & 1) : s;  // stages before it run as radix-2 rounds
    for (uint32_t st = 0; st < r4_from; ++st) {
        const uint32_t lh = s - 1 - st;
        if (a.last && lh == 0) {  // every twiddle is w^0
            for (uint32_t q = threadIdx.x; q < count / 2; q += THREADS) {
                const uint32_t sa = ((q >> lg) << (lg + 1)) + (q & gmask), sb = sa + (1u << lg);
                const Fr x = tile_get(tile, sa), y = tile_get(tile, sb);
                tile_put(tile, sa, add_lazy(x, y));
                tile_put(tile, sb, reduce_lazy(sub_lazy(x, y)));
            }
            __syncthreads();
            continue;
        }
        for (uint32_t q = threadIdx.x; q < count / 2; q += THREADS) {
            const uint32_t col = q & gmask, r = q >> lg;
            const uint32_t r_lo = r & ((1u << lh) - 1);
            const uint32_t sa = ((((r >> lh) << (lh + 1)) + r_lo) << lg) + col, sb = sa + (1u << (lh + lg));
            const size_t j0 = a.last ? 0 : ((((size_t)wg << lg) + col) & lmmask);
            const size_t e = (((size_t)r_lo << lm) + j0) << (L - 1 - lh - lm);
            const Fr t = fp_load<FrParams>(a.tw + 8 * e);
            const Fr x = tile_get(tile, sa), y = tile_get(tile, sb);
            tile_put(tile, sa, add_lazy(x, y));
            tile_put(tile, sb, mul_lazy(sub_lazy(x, y), t));
        }
        __syncthreads();
    }
    for (uint32_t st = r4_from; RADIX4 && st + 1 < s; st += 2) {
        const uint32_t lh1 = s - 1 - st, lh2 = lh1 - 1;
        if (a.last && lh2 == 0) {
            // final round of the last pass: only stage lh = 1's odd rows carry a twiddle
            const Fr w4 = fp_load<FrParams>(a.tw + 8 * ((size_t)1 << (L - 2)));
            for (uint32_t q = threadIdx.x; q < count / 4; q += THREADS) {
                const uint32_t col = q & gmask, r = q >> lg;
                const uint32_t s00 = ((r << 2) << lg) + col, s01 = s00 + (1u << lg);
                const uint32_t s10 = s00 + (2u << lg), s11 = s10 + (1u << lg);
                const Fr x00 = tile_get(tile, s00), x01 = tile_get(tile, s01);
                const Fr x10 = tile_get(tile, s10), x11 = tile_get(tile, s11);
                const Fr a0 = add_lazy(x00, x10), a1 = reduce_lazy(sub_lazy(x00, x10));
                const Fr b0 = add_lazy(x01, x11), b1 = mul_lazy(sub_lazy(x01, x11), w4);
                tile_put(tile, s00, add_lazy(a0, b0));
                tile_put(tile, s01, reduce_lazy(sub_lazy(a0, b0)));
                tile_put(tile, s10, add_lazy(a1, b1));
                tile_put(tile, s11, reduce_lazy(sub_lazy(a1, b1)));
            }
            __syncthreads();
            continue;
        }
        for (uint32_t q = threadIdx.x; q < count / 4; q += THREADS) {
            const uint32_t col = q & gmask, r = q >> lg;
            const uint32_t lo = r & ((1u << lh2) - 1), hi = r >> lh2;
            const uint32_t row00 = (hi << (lh1 + 1)) + lo;
            const uint32_t s00 = (row00 << lg) + col, s01 = s00 + (1u << (lh2 + lg));
            const uint32_t s10 = s00 + (1u << (lh1 + lg)), s11 = s10 + (1u << (lh2 + lg));
            const size_t j0 = a.last ? 0 : ((((size_t)wg << lg) + col) & lmmask);
            const size_t e1a = (((size_t)lo << lm) + j0) << (L - 1 - lh1 - lm);
            const size_t e1b = ((((size_t)lo + (1u << lh2)) << lm) + j0) << (L - 1 - lh1 - lm);
            const size_t e2 = (((size_t)lo << lm) + j0) << (L - 1 - lh2 - lm);
            const Fr t1a = fp_load<FrParams>(a.tw + 8 * e1a), t1b = fp_load<FrParams>(a.tw + 8 * e1b);
            const Fr t2 = fp_load<FrParams>(a.tw + 8 * e2);
            const Fr x00 = tile_get(tile, s00), x01 = tile_get(tile, s01);
            const Fr x10 = tile_get(tile, s10), x11 = tile_get(tile, s11);
            // stage st: (x0b, x1b) for b = 0, 1
            const Fr a0 = add_lazy(x00, x10), a1 = mul_lazy(sub_lazy(x00, x10), t1a);
            const Fr b0 = add_lazy(x01, x11), b1 = mul_lazy(sub_lazy(x01, x11), t1b);
            // stage st + 1: (a0, b0) -> rows 00, 01 and (a1, b1) -> rows 10, 11
            tile_put(tile, s00, add_lazy(a0, b0));
            tile_put(tile, s01, mul_lazy(sub_lazy(a0, b0), t2));
            tile_put(tile, s10, add_lazy(a1, b1));
            tile_put(tile, s11, mul_lazy(sub_lazy(a1, b1), t2));
        }
        __syncthreads();
    }

    // ---- store
    if (a.last) {
        Fr scale;
#pragma unroll
        for (int i = 0; i < 8; ++i) scale.w[i] = a.scale[i];
        for (uint32_t e = threadIdx.x; e < count; e += THREADS) {
            const uint32_t col = e & gmask, orow = e >> lg;
            const uint32_t row = bitrev_bits(orow, s);
            const uint32_t slot = (row << lg) + col;
            Fr v = tile_get(tile, slot);  // < 2r
            const size_t out = ((size_t)orow << (L - s)) + ((size_t)wg << lg) + col;
            if (a.scale_tab) v = mul_lazy(v, fp_load<FrParams>(a.scale_tab + 8 * out));  // per-element scale
            else if (a.scale_flag) v = mul_lazy(v, scale);                                // < 2r
            fp_reduce_once<FrParams>(v.w);                                              // canonical output
            uint4* p = reinterpret_cast<uint4*>(a.dst + 8 * out);
            p[0] = make_uint4(v.w[0], v.w[1], v.w[2], v.w[3]);
            p[1] = make_uint4(v.w[4], v.w[5], v.w[6], v.w[7]);
        }
    } else {
        const uint32_t gj = min(lg, lm);
        for (uint32_t e = threadIdx.x; e < count; e += THREADS) {
            const uint32_t lo = e & ((1u << gj) - 1), row = (e >> gj) & smask, hi = e >> (gj + s);
            const uint32_t col = (hi << gj) + lo;
            const uint32_t slot = (row << lg) + col;
            uint4* p = reinterpret_cast<uint4*>(a.dst + 8 * tile_addr(wg, row, col, lm, s, lg));
            p[0] = tile[2 * slot];
            p[1] = tile[2 * slot + 1];
        }
    }
}

template <int TILE_LOG, int THREADS>
__global__ __launch_bounds__(THREADS) void ntt_pass_kernel(NttPassArgs a) {
    __shared__ uint4 tile[2 << TILE_LOG];  // 32 bytes per element
    ntt_pass_body<TILE_LOG, THREADS>(a, blockIdx.x, tile);
}

// Many transforms in one launch: workgroup g runs tile g - wg_start[j] of job j, where j is the
// last job with wg_start[j] <= g. Jobs may differ in size, direction and pass parameters.
template <int TILE_LOG, int THREADS>
__global__ __launch_bounds__(THREADS) void ntt_pass_batch_kernel(const NttPassArgs* __restrict__ jobs,
                                                                const uint32_t* __restrict__ wg_start,
                                                                uint32_t njobs) {
    __shared__ uint4 tile[2 << TILE_LOG];
    const uint32_t g = blockIdx.x;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (wg_start[mid] <= g) lo = mid;
        else hi = mid;
    }
    const NttPassArgs a = jobs[lo];
    ntt_pass_body<TILE_LOG, THREADS>(a, g - wg_start[lo], tile);
}

// tw[i] = hi[i >> lb] * lo[i & (2^lb - 1)] (Montgomery), i < count
__global__ void ntt_twiddle_kernel(const uint32_t* __restrict__ hi, const uint32_t* __restrict__ lo,
                                   uint32_t* __restrict__ out, uint32_t lb, size_t count) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Fr h = fp_load<FrParams>(hi + 8 * (i >> lb));
    const Fr l = fp_load<FrParams>(lo + 8 * (i & ((1ull << lb) - 1)));
    fp_store(out + 8 * i, h * l);
}

// ---------------------------------------------------------------- plan

// Measured on MI355X (forward and inverse, every log_n in 12..26, tiles 2^10, 2^11, 2^12):
// 2^10-element tiles (256-lane workgroups, 130 VGPRs, three waves per SIMD) are fastest up to
// 2^18 and at 2^22; 2^11 (512 lanes, 127 VGPRs, four waves per SIMD) from 2^19 on, by 4-17%
// (2^24: 1.99 against 2.25 ms). 2^12 tiles run radix-2 rounds only and are never ahead.
int ntt_default_tile_log(uint32_t log_n) { return log_n >= 19 && log_n != 22 ? 11 : 10; }

// The tile a plan uses when none is given: the tuning table's, else the default.
static int ntt_tile_for(const Device* d, uint32_t log_n) {
    NttTune t;
    if (d && tune_lookup_ntt(d->name, log_n, &t)) return t.tile_log;
    return ntt_default_tile_log(log_n);
}

static void launch_pass(int tile_log, const NttPassArgs& a, unsigned wgs, hipStream_t stream) {
    switch (tile_log) {
        case 10: hipLaunchKernelGGL((ntt_pass_kernel<10, 256>), dim3(wgs), dim3(256), 0, stream, a); break;
        case 11: hipLaunchKernelGGL((ntt_pass_kernel<11, 512>), dim3(wgs), dim3(512), 0, stream, a); break;
        case 12: hipLaunchKernelGGL((ntt_pass_kernel<12, 1024>), dim3(wgs), dim3(1024), 0, stream, a); break;
        default: throw Error(ZKG_ERR_INTERNAL, "unsupported NTT tile size");
    }
    ZKG_HIP(hipGetLastError());
}

// out[i] = first * base^i (Montgomery form), i < count: out[i] = hi[i >> lb] * lo[i & (2^lb - 1)]
// from two ~sqrt(count) seed tables built on the host.
static void build_geometric(uint32_t* out, size_t count, const host::U256& base, const host::U256& first,
                            hipStream_t stream) {
    const host::Modulus& F = host::fr();
    uint32_t lb = 0;
    while (((size_t)1 << (2 * lb)) < count) ++lb;  // lo_len^2 >= count
    const size_t lo_len = (size_t)1 << lb;
    const size_t hi_len = (count + lo_len - 1) / lo_len;
    std::vector<uint32_t> lo(8 * lo_len), hi(8 * hi_len);
    const host::U256 bm = F.to_mont(base);
    host::U256 x = F.one;  // Montgomery 1
    for (size_t i = 0; i < lo_len; ++i) {
        host::to_words32(&lo[8 * i], x);
        x = F.mont(x, bm);
    }
    const host::U256 step = x;  // base^lo_len (Montgomery)
    x = F.to_mont(first);
    for (size_t i = 0; i < hi_len; ++i) {
        host::to_words32(&hi[8 * i], x);
        x = F.mont(x, step);
    }
    uint32_t *dlo = nullptr, *dhi = nullptr;
    ZKG_HIP(hipMalloc(&dlo, lo.size() * 4));
    ZKG_HIP(hipMalloc(&dhi, hi.size() * 4));
    ZKG_HIP(hipMemcpyAsync(dlo, lo.data(), lo.size() * 4, hipMemcpyHostToDevice, stream));
    ZKG_HIP(hipMemcpyAsync(dhi, hi.data(), hi.size() * 4, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(ntt_twiddle_kernel, dim3(ceil_div(count, 256)), dim3(256), 0, stream, dhi, dlo, out, lb, count);
    ZKG_HIP(hipGetLastError());
    // The host vectors go out of scope: wait for the copies and the kernel before freeing.
    ZKG_HIP(hipStreamSynchronize(stream));
    ZKG_HIP(hipFree(dlo));
    ZKG_HIP(hipFree(dhi));
}

static void build_twiddles(uint32_t* out, uint32_t log_n, const host::U256& root, hipStream_t stream) {
    host::U256 one = {{1, 0, 0, 0}};
    build_geometric(out, (size_t)1 << (log_n - 1), root, one, stream);
}

void ntt_geometric_table(uint32_t* out, size_t count, const uint32_t base[8], const uint32_t first[8],
                         hipStream_t stream) {
    build_geometric(out, count, host::from_words32(base), host::from_words32(first), stream);
}

NttPlan::NttPlan(Device* d, uint32_t ln, NttShape shape, bool with_scratch) : dev(d), log_n(ln) {
    if (ln < 10 || ln > 27) throw Error(ZKG_ERR_INVALID, "log_n must be in 10..=27");
    tile_log = shape.tile_log ? shape.tile_log : ntt_tile_for(d, ln);
    const size_t n = (size_t)1 << ln;
    device_bind(d);
    try {
        ZKG_HIP(hipMalloc(&tw_fwd, 32 * (n / 2)));
        ZKG_HIP(hipMalloc(&tw_inv, 32 * (n / 2)));
        if (with_scratch) ZKG_HIP(hipMalloc(&scratch, 32 * n));
        const host::Modulus& F = host::fr();
        const host::U256 w = host::root_of_unity(ln);
        build_twiddles(tw_fwd, ln, w, d->stream);
        build_twiddles(tw_inv, ln, F.inverse(w), d->stream);
        host::U256 nn = {{0, 0, 0, 0}};
        nn.l[ln / 64] = 1ull << (ln % 64);
        host::to_words32(scale, F.to_mont(F.inverse(nn)));
    } catch (...) {
        if (tw_fwd) (void)hipFree(tw_fwd);
        if (tw_inv) (void)hipFree(tw_inv);
        if (scratch) (void)hipFree(scratch);
        throw;
    }
    // Split log_n stages over P passes of at most tile_log stages, as evenly as possible.
    const uint32_t P = (ln + tile_log - 1) / tile_log;
    uint32_t done = 0;
    for (uint32_t i = 0; i < P; ++i) {
        NttPass p;
        p.s = ln / P + (i < ln % P ? 1 : 0);
        done += p.s;
        p.lm = ln - done;
        p.last = (i == P - 1);
        p.lg = std::min<uint32_t>(tile_log - p.s, ln - p.s);
        p.wgs = 1u << (ln - p.s - p.lg);
        p.in_place = !p.last || P == 1;
        passes.push_back(p);
    }
}

NttPlan::~NttPlan() {
    (void)hipFree(tw_fwd);
    (void)hipFree(tw_inv);
    (void)hipFree(scratch);
}

size_t NttPlan::workspace_bytes() const { return (size_t)64 << log_n; }

NttPassArgs NttPlan::pass_args(size_t i, uint32_t* data, uint32_t* scr, bool inverse, const uint32_t* scale_tab) const {
    const size_t P = passes.size();
    const NttPass& p = passes[i];
    NttPassArgs a;
    // passes 0..P-3 in place on data; pass P-2 data -> scratch; pass P-1 scratch -> data
    if (P == 1 || i + 2 < P) {
        a.src = data;
        a.dst = data;
    } else if (i + 2 == P) {
        a.src = data;
        a.dst = scr;
    } else {
        a.src = scr;
        a.dst = data;
    }
    a.tw = inverse ? tw_inv : tw_fwd;
    a.lm = p.lm;
    a.s = p.s;
    a.lg = p.lg;
    a.log_n = log_n;
    a.last = p.last;
    a.scale_flag = (inverse && p.last) ? 1 : 0;
    a.scale_tab = p.last ? scale_tab : nullptr;
    for (int k = 0; k < 8; ++k) a.scale[k] = scale[k];
    return a;
}

void NttPlan::run(uint32_t* data, bool inverse, hipStream_t stream, const uint32_t* scale_tab) const {
    run_with(data, inverse, stream, scale_tab, scratch);
}

void NttPlan::run_with(uint32_t* data, bool inverse, hipStream_t stream, const uint32_t* scale_tab, uint32_t* scr) const {
    if (!scr && passes.size() > 1) throw Error(ZKG_ERR_INTERNAL, "NTT plan run without scratch");
    if (scr && poison_enabled()) fill_words(scr, 0xA5A5A5A5u, (size_t)8 << log_n, stream);
    for (size_t i = 0; i < passes.size(); ++i)
        launch_pass(tile_log, pass_args(i, data, scr, inverse, scale_tab), passes[i].wgs, stream);
}

// ---------------------------------------------------------------- batch

NttBatch::NttBatch(Device* d) : dev_(d) {
    device_bind(d);
    ZKG_HIP(hipEventCreateWithFlags(&staged_, hipEventDisableTiming));
}

NttBatch::~NttBatch() {
    (void)hipSetDevice(dev_->ordinal);
    (void)hipEventSynchronize(staged_);
    (void)hipEventDestroy(staged_);
    if (staging_) (void)hipHostFree(staging_);
    scratch_.release();
    table_.release();
}

int NttBatch::tile_for(const Device* d, uint32_t log_n) { return ntt_tile_for(d, log_n); }

const NttPlan& NttBatch::plan(uint32_t log_n, int tile) {
    auto it = plans_.find(log_n);
    if (it == plans_.end() || it->second->tile_log != tile) {
        NttShape shape;
        shape.tile_log = tile;
        auto p = std::make_shared<NttPlan>(dev_, log_n, shape, false);
        if (it == plans_.end()) it = plans_.emplace(log_n, p).first;
        else it->second = p;
    }
    return *it->second;
}

size_t NttBatch::add(uint32_t log_n, uint32_t* data, bool inverse) {
    if (log_n < 10 || log_n > 27) throw Error(ZKG_ERR_INVALID, "log_n must be in 10..=27");
    prepared_ = false;
    jobs_.push_back(Job{log_n, data, inverse});
    return jobs_.size() - 1;
}

void NttBatch::clear() {
    jobs_.clear();
    prepared_ = false;
}

void NttBatch::prepare(hipStream_t stream) {
    prepared_ = false;
    if (jobs_.empty()) return;
    device_bind(dev_);
    // one kernel per pass phase serves every job, so one tile for the batch: the best tile of
    // its largest transform (the tuning table's, else the default rule)
    uint32_t top = 0;
    for (const Job& j : jobs_) top = std::max(top, j.log_n);
    tile_ = tile_for(dev_, top);
    size_t scratch_elems = 0, max_passes = 0;
    std::vector<size_t> offs(jobs_.size());
    for (size_t j = 0; j < jobs_.size(); ++j) {
        offs[j] = scratch_elems;
        scratch_elems += (size_t)1 << jobs_[j].log_n;
        max_passes = std::max(max_passes, plan(jobs_[j].log_n, tile_).passes.size());
    }
    // the previous run's upload must be done before the staging memory is rewritten
    ZKG_HIP(hipEventSynchronize(staged_));
    scratch_.ensure(32 * scratch_elems);
    // table layout per phase: njobs NttPassArgs, then (njobs + 1) u32 workgroup starts (padded)
    const size_t nj = jobs_.size();
    const size_t start_slots = (4 * (nj + 1) + sizeof(NttPassArgs) - 1) / sizeof(NttPassArgs);
    per_phase_ = nj + start_slots;
    const size_t need = per_phase_ * max_passes;
    if (need > staging_cap_) {
        if (staging_) ZKG_HIP(hipHostFree(staging_));
        staging_ = nullptr;
        ZKG_HIP(hipHostMalloc(reinterpret_cast<void**>(&staging_), need * sizeof(NttPassArgs)));
        staging_cap_ = need;
    }
    table_.ensure(need * sizeof(NttPassArgs));
    phase_jobs_.assign(max_passes, 0);
    phase_wgs_.assign(max_passes, 0);
    for (size_t k = 0; k < max_passes; ++k) {
        NttPassArgs* args = staging_ + k * per_phase_;
        uint32_t* starts = reinterpret_cast<uint32_t*>(args + nj);
        uint32_t cnt = 0, wgs = 0;
        for (size_t j = 0; j < nj; ++j) {
            const NttPlan& pl = *plans_.at(jobs_[j].log_n);
            if (pl.passes.size() <= k) continue;
            uint32_t* scr = static_cast<uint32_t*>(scratch_.ptr) + 8 * offs[j];
            args[cnt] = pl.pass_args(k, jobs_[j].data, scr, jobs_[j].inverse);
            starts[cnt] = wgs;
            wgs += pl.passes[k].wgs;
            ++cnt;
        }
        starts[cnt] = wgs;
        phase_jobs_[k] = cnt;
        phase_wgs_[k] = wgs;
    }
    ZKG_HIP(hipMemcpyAsync(table_.ptr, staging_, need * sizeof(NttPassArgs), hipMemcpyHostToDevice, stream));
    ZKG_HIP(hipEventRecord(staged_, stream));
    prepared_ = true;
}

void NttBatch::launch(hipStream_t stream) const {
    if (jobs_.empty()) return;
    if (!prepared_) throw Error(ZKG_ERR_INTERNAL, "NTT batch launched before it was prepared");
    const size_t nj = jobs_.size();
    for (size_t k = 0; k < phase_jobs_.size(); ++k) {
        if (!phase_jobs_[k]) continue;
        const NttPassArgs* args = static_cast<const NttPassArgs*>(table_.ptr) + k * per_phase_;
        const uint32_t* starts = reinterpret_cast<const uint32_t*>(args + nj);
        switch (tile_) {
            case 10:
                hipLaunchKernelGGL((ntt_pass_batch_kernel<10, 256>), dim3(phase_wgs_[k]), dim3(256), 0, stream, args,
                                   starts, phase_jobs_[k]);
                break;
            case 11:
                hipLaunchKernelGGL((ntt_pass_batch_kernel<11, 512>), dim3(phase_wgs_[k]), dim3(512), 0, stream, args,
                                   starts, phase_jobs_[k]);
                break;
            case 12:
                hipLaunchKernelGGL((ntt_pass_batch_kernel<12, 1024>), dim3(phase_wgs_[k]), dim3(1024), 0, stream, args,
                                   starts, phase_jobs_[k]);
                break;
            default: throw Error(ZKG_ERR_INTERNAL, "unsupported NTT tile size");
        }
        ZKG_HIP(hipGetLastError());
    }
}

void NttBatch::run(hipStream_t stream) {
    prepare(stream);
    if (poison_enabled() && scratch_.ptr) fill_words(static_cast<uint32_t*>(scratch_.ptr), 0xA5A5A5A5u, scratch_.bytes / 4, stream);
    launch(stream);
}

}  // namespace zkg
