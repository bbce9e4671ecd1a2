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
//  merge r    heavy buckets: items k = 0 mod s*G absorb items k + s .. k + (G-1)s (s = G^r), so a
//             bucket holding most of the points never becomes a serial tail.
//  finalize   each bucket sums its at most G remaining items.
//  reduce     S_w = sum_b (b + 1) bucket[w][b]: L buckets per lane with running sums, plus b0 * run.
//  tree sum   per-window sum of the lane results in one workgroup (LDS tree).
//  combine    lane w scales S_w by c*w doublings, then a wave-wide LDS tree adds the W results.
#include <algorithm>
#include <chrono>

#include "curve.hpp"
#include "msm.hpp"
#include "msm_stages.hpp"

namespace zkg {

ZKG_MSM_EXTERN(Fq)
ZKG_MSM_EXTERN(Fq2)

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

// ---------------------------------------------------------------- digits: count and scatter

__device__ __forceinline__ uint32_t window_bits(const uint32_t s[8], uint32_t pos, uint32_t c) {
    const uint32_t word = pos >> 5, off = pos & 31;
    uint32_t v = s[word] >> off;
    if (off + c > 32 && word + 1 < 8) v |= s[word + 1] << (32 - off);
    return v & ((1u << c) - 1);
}

template <bool SCATTER>
__global__ __launch_bounds__(256) void msm_digits_kernel(const uint32_t* __restrict__ scalars, uint32_t n, uint32_t c,
                                                         uint32_t W, uint32_t* __restrict__ counter,
                                                         const uint32_t* __restrict__ starts,
                                                         uint32_t* __restrict__ sorted) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const uint32_t B = 1u << (c - 1);
    const uint32_t lane = __lane_id();
    const unsigned long long lt_mask = (1ull << lane) - 1;
    uint32_t s[8];
    if (valid) {
        const uint4* p = reinterpret_cast<const uint4*>(scalars + 8 * (size_t)i);
        const uint4 a = p[0], b = p[1];
        s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w; s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] = 0;
    }
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint32_t v = window_bits(s, w * c, c) + carry;
        uint32_t mag = v, neg = 0;
        carry = 0;
        if (v >= B) {  // digit v - 2^c < 0
            mag = (1u << c) - v;
            neg = 1;
            carry = 1;
        }
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

__global__ __launch_bounds__(256) void msm_digits_store_kernel(const uint32_t* __restrict__ scalars, uint32_t n,
                                                              uint32_t c, uint32_t W, uint32_t* __restrict__ digits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t B = 1u << (c - 1);
    uint32_t s[8];
    const uint4* p = reinterpret_cast<const uint4*>(scalars + 8 * (size_t)i);
    const uint4 a = p[0], b = p[1];
    s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w; s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint32_t v = window_bits(s, w * c, c) + carry;
        uint32_t mag = v, neg = 0;
        carry = 0;
        if (v >= B) {
            mag = (1u << c) - v;
            neg = 1;
            carry = 1;
        }
        digits[(size_t)w * n + i] = mag ? ((mag - 1) | (neg << 31)) : NO_DIGIT;
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

// ---------------------------------------------------------------- scan of counts and items
// Three kernels: per-block exclusive scans (4096 slots per block), a single-block scan of the
// block totals, and the add-back of block offsets. Both scans (points and items) run together.

static constexpr int SCAN_THREADS = 1024;
static constexpr int SCAN_PER_THREAD = 4;
static constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_PER_THREAD;

__device__ __forceinline__ uint32_t items_of(uint32_t count, uint32_t ch) { return (count + ch - 1) / ch; }

__global__ __launch_bounds__(SCAN_THREADS) void scan_block_kernel(const uint32_t* __restrict__ counts, uint32_t S,
                                                                 uint32_t ch, uint32_t* __restrict__ starts,
                                                                 uint32_t* __restrict__ istarts,
                                                                 uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t sp[SCAN_THREADS], si[SCAN_THREADS];
    const uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
    uint32_t cp[SCAN_PER_THREAD], tp = 0, ti = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        cp[k] = base + k < S ? counts[base + k] : 0;
        tp += cp[k];
        ti += items_of(cp[k], ch);
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
        ri += items_of(cp[k], ch);
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        block_sums[2 * blockIdx.x] = sp[threadIdx.x];
        block_sums[2 * blockIdx.x + 1] = si[threadIdx.x];
    }
}

// Exclusive scan of `nb` (points, items) pairs in place; grand totals go to starts[S], istarts[S].
__global__ __launch_bounds__(SCAN_THREADS) void scan_top_kernel(uint32_t* __restrict__ block_sums, uint32_t nb,
                                                               uint32_t S, uint32_t* __restrict__ starts,
                                                               uint32_t* __restrict__ istarts) {
    __shared__ uint32_t sp[SCAN_THREADS], si[SCAN_THREADS];
    const uint32_t per = (nb + SCAN_THREADS - 1) / SCAN_THREADS;
    const uint32_t lo = threadIdx.x * per, hi = min(lo + per, nb);
    uint32_t tp = 0, ti = 0;
    for (uint32_t k = lo; k < hi; ++k) {
        tp += block_sums[2 * k];
        ti += block_sums[2 * k + 1];
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

MsmPlan::MsmPlan(Device* d, int grp, size_t npts, uint32_t win, uint32_t item_points) : dev(d), group(grp) {
    if (npts < 1) throw Error(ZKG_ERR_INVALID, "an MSM needs at least one point");
    if (win < 9 || win > 20) throw Error(ZKG_ERR_INVALID, "window size c must be in 9..=20, got " + std::to_string(win));
    if (grp != 1 && grp != 2) throw Error(ZKG_ERR_INVALID, "group must be G1 (1) or G2 (2)");
    if (npts > 0x7fffffffu) throw Error(ZKG_ERR_INVALID, "an MSM holds at most 2^31 - 1 points");
    n = (uint32_t)npts;
    c = win;
    W = (255 + c - 1) / c;
    B = 1u << (c - 1);
    CH = item_points ? item_points : msm_auto_item_points(n, c);
    L = 4;
    T = B / L;
    aw = group == 1 ? 16 : 32;
    xw = 2 * aw;
    const size_t S = (size_t)W * B;
    max_items = (size_t)W * (((size_t)n + CH - 1) / CH + B);
    if (max_items >= 0xffffffffull) throw Error(ZKG_ERR_INVALID, "too many work items for one plan");
    for (uint64_t s = 1, m = ((uint64_t)n + CH - 1) / CH; s < m; s *= SUM_GROUP) ++merge_rounds;
    const size_t nb = (S + SCAN_BLOCK - 1) / SCAN_BLOCK;
    lds_sort = c <= 16;
    sort_pts = std::max<uint32_t>(B, 8192);
    nblk = (n + sort_pts - 1) / sort_pts;
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
        ZKG_HIP(hipMalloc(&scan_tmp, 8 * nb));
        ZKG_HIP(hipMalloc(&sorted, 4 * (size_t)W * n));
        if (lds_sort) {
            ZKG_HIP(hipMalloc(&digits, 4 * (size_t)W * n));
            ZKG_HIP(hipMalloc(&table, 4 * (size_t)W * nblk * B));
        }
        ZKG_HIP(hipMalloc(&item_slot, 4 * max_items));
        ZKG_HIP(hipMalloc(&item_pts, 4 * max_items * xw));
        ZKG_HIP(hipMalloc(&buckets, 4 * S * xw));
        ZKG_HIP(hipMalloc(&partials, 4 * (size_t)W * T * xw));
        ZKG_HIP(hipMalloc(&wsums, 4 * (size_t)W * xw));
        ZKG_HIP(hipMalloc(&result, 4 * (size_t)xw));
        ZKG_HIP(hipMemsetAsync(result, 0, 4 * (size_t)xw, d->stream));
    } catch (...) {
        this->~MsmPlan();
        throw;
    }
}

MsmPlan::~MsmPlan() {
    for (uint32_t* p : {counts, cursors, starts, istarts, scan_tmp, sorted, digits, table, item_slot, item_pts, buckets,
                        partials, wsums, result})
        if (p) (void)hipFree(p);
    counts = cursors = starts = istarts = scan_tmp = sorted = digits = table = item_slot = item_pts = buckets =
        partials = wsums = result = nullptr;
}

size_t MsmPlan::workspace_bytes() const { return msm_workspace_bytes(group, n, c, CH); }

size_t msm_workspace_bytes(int group, size_t n, uint32_t c, uint32_t item_points) {
    const uint32_t W = (255 + c - 1) / c, B = 1u << (c - 1);
    const uint32_t CH = item_points ? item_points : msm_auto_item_points(n, c);
    const uint32_t xw = group == 1 ? 32 : 64, T = B / 4;
    const size_t S = (size_t)W * B, max_items = (size_t)W * ((n + CH - 1) / CH + B);
    const bool lds = c <= 16;
    const size_t P = std::max<uint32_t>(B, 8192), nblk = (n + P - 1) / P;
    const size_t nb = (S + SCAN_BLOCK - 1) / SCAN_BLOCK;
    return 4 * (4 * S + 2 + 2 * nb + (size_t)W * n * (lds ? 2 : 1) + (lds ? (size_t)W * nblk * B : 0) +
                max_items * (1 + (size_t)xw) + S * xw + (size_t)W * T * xw + (size_t)W * xw + xw);
}

template <class Fe>
static void run_stages(const MsmPlan& p, const uint32_t* bases, const uint32_t* scalars, hipStream_t s,
                       std::vector<std::pair<std::string, double>>* prof) {
    const uint32_t S = p.W * p.B;
    const uint32_t nb = (S + SCAN_BLOCK - 1) / SCAN_BLOCK;
    auto t0 = std::chrono::steady_clock::now();
    auto stage = [&](const char* name) {
        ZKG_HIP(hipGetLastError());
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
    if (prof) {
        ZKG_HIP(hipStreamSynchronize(s));
        t0 = std::chrono::steady_clock::now();
    }
    const unsigned pg = ceil_div(p.n, 256);
    if (p.lds_sort) {
        hipLaunchKernelGGL(msm_digits_store_kernel, dim3(pg), dim3(256), 0, s, scalars, p.n, p.c, p.W, p.digits);
        hipLaunchKernelGGL(msm_hist_kernel, dim3(p.nblk, p.W), dim3(SORT_THREADS), 4 * p.B, s, p.digits, p.n, p.B,
                           p.sort_pts, p.nblk, p.table);
        hipLaunchKernelGGL(msm_colscan_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, s, p.table, p.W, p.B, p.nblk,
                           p.counts);
    } else {
        fill_u32(p.counts, 0, S, s);
        fill_u32(p.cursors, 0, S, s);
        hipLaunchKernelGGL((msm_digits_kernel<false>), dim3(pg), dim3(256), 0, s, scalars, p.n, p.c, p.W, p.counts,
                           (const uint32_t*)nullptr, (uint32_t*)nullptr);
    }
    stage("count");
    hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(SCAN_THREADS), 0, s, p.counts, S, p.CH, p.starts, p.istarts,
                       p.scan_tmp);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(SCAN_THREADS), 0, s, p.scan_tmp, nb, S, p.starts, p.istarts);
    hipLaunchKernelGGL(scan_add_kernel, dim3(ceil_div(S, 256)), dim3(256), 0, s, p.scan_tmp, S, p.starts, p.istarts);
    stage("scan");
    if (p.lds_sort) {
        hipLaunchKernelGGL(msm_place_kernel, dim3(p.nblk, p.W), dim3(SORT_THREADS), 4 * p.B, s, p.digits, p.n, p.B,
                           p.sort_pts, p.nblk, p.table, p.starts, p.sorted);
    } else {
        hipLaunchKernelGGL((msm_digits_kernel<true>), dim3(pg), dim3(256), 0, s, scalars, p.n, p.c, p.W, p.cursors,
                           p.starts, p.sorted);
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
    if (group == 1) run_stages<Fq>(*this, bases, scalars, stream, stage_ms);
    else run_stages<Fq2>(*this, bases, scalars, stream, stage_ms);
}

}  // namespace zkg
