// Curve-arithmetic stages of the MSM (accumulate, merge, finalize, reduce, tree sum,
// combine) as kernel templates over the coordinate field, with their launchers. Included only
// by the translation units that instantiate them (msm_g1.hip, msm_g2_*.hip).
#pragma once
#include <algorithm>
#include <type_traits>

#include "curve.hpp"
#include "field29.hpp"
#include "msm.hpp"
#include "msm_stages.hpp"

namespace zkg {

// ---------------------------------------------------------------- bucket accumulation

// Sorted entries name virtual points v | sign << 31: v < n_real is base v, v >= n_real (GLV) is
// phi(base v - n_real) = (endo_x[v - n_real], y). Without GLV n_real is 2^32 - 1.
// Addresses of the coordinates of sorted entry e: base v (x from the bases, or phi(P)'s x from
// endo_x for v >= n_real) with WORDS words per coordinate.
template <int WORDS>
ZKG_DEV void point_address(uint32_t e, uint32_t n_real, const uint32_t* bases, const uint32_t* endo_x,
                           const uint32_t** xp, const uint32_t** yp) {
    const uint32_t v = e & 0x7fffffffu;
    const bool endo = v >= n_real;
    const uint32_t idx = endo ? v - n_real : v;
    *yp = bases + (size_t)(2 * WORDS) * idx + WORDS;
    *xp = endo ? endo_x + (size_t)WORDS * idx : *yp - WORDS;
}

// Out: the coordinate type of the later stages; item sums are stored in its memory form (the
// 32-bit G2 accumulation feeding the Fq2_29 stages converts each coordinate once per item).
template <class Fe, class Out = Fe>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void msm_accumulate_kernel(
                                                            const uint32_t* __restrict__ bases,
                                                            const uint32_t* __restrict__ endo_x, uint32_t n_real,
                                                            const uint32_t* __restrict__ sorted,
                                                            const uint32_t* __restrict__ starts,
                                                            const uint32_t* __restrict__ istarts,
                                                            const uint32_t* __restrict__ counts, uint32_t S,
                                                            uint32_t ch, uint2* __restrict__ item_info,
                                                            uint32_t* __restrict__ item_pts) {
    typedef FeTraits<Fe> Tr;
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= istarts[S]) return;
    // last slot whose first item is <= item (empty slots before it share its istarts value)
    uint32_t lo = 0, hi = S;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (istarts[mid] <= item) lo = mid;
        else hi = mid;
    }
    const uint32_t slot = lo;
    // items cut at the bucket's start: ceil(count / ch) of the bucket's slots; the spare one (the
    // slot layout leaves room for lanes cut at multiples of ch instead) holds no items
    const uint32_t idx = item - istarts[slot], nitems = (counts[slot] + ch - 1) / ch;
    if (idx >= nitems) {
        item_info[item] = make_uint2(0u, 0u);
        return;
    }
    const uint32_t skip = idx * ch;
    const uint32_t len = min(ch, counts[slot] - skip);
    const uint32_t* entries = sorted + starts[slot] + skip;
    Xyzz<Fe> acc = xyzz_inf<Fe>();
    if constexpr (Tr::WORDS == 8) {
        // G1: point j + 1 is read while point j is added (software pipelining of the gathered
        // reads; G2's registers have no room for the extra point: it spilled and ran 9% slower)
        uint32_t e_next = 0;
        Affine<Fe> next;
        if (len) {
            e_next = entries[0];
            const uint32_t* xp;
            const uint32_t* yp;
            point_address<Tr::WORDS>(e_next, n_real, bases, endo_x, &xp, &yp);
            next = Affine<Fe>{Tr::load(xp), Tr::load(yp)};
        }
        for (uint32_t j = 0; j < len; ++j) {
            const uint32_t e = e_next;
            Affine<Fe> p = next;
            if (j + 1 < len) {
                e_next = entries[j + 1];
                const uint32_t* xp;
                const uint32_t* yp;
                point_address<Tr::WORDS>(e_next, n_real, bases, endo_x, &xp, &yp);
                next = Affine<Fe>{Tr::load(xp), Tr::load(yp)};
            }
            if (e >> 31) p.y = Tr::zero() - p.y;
            acc = xyzz_madd(acc, p);
        }
    } else {
        for (uint32_t j = 0; j < len; ++j) {
            const uint32_t e = entries[j];
            const uint32_t* xp;
            const uint32_t* yp;
            point_address<Tr::WORDS>(e, n_real, bases, endo_x, &xp, &yp);
            Affine<Fe> p{Tr::load(xp), Tr::load(yp)};
            if (e >> 31) p.y = Tr::zero() - p.y;
            acc = xyzz_madd(acc, p);
        }
    }
    if constexpr (std::is_same<Fe, Out>::value) {
        store_xyzz(item_pts + (size_t)(4 * Tr::WORDS) * item, acc);
    } else {
        static_assert(4 * FeTraits<Out>::WORDS == 4 * Tr::WORDS, "same memory layout");
        const Out* tag = nullptr;
        Xyzz<Out> o = xyzz_inf<Out>();
        if (!xyzz_is_inf(acc))
            o = Xyzz<Out>{to_storage_field(acc.x, tag), to_storage_field(acc.y, tag), to_storage_field(acc.zz, tag),
                          to_storage_field(acc.zzz, tag)};
        store_xyzz(item_pts + (size_t)(4 * Tr::WORDS) * item, o);
    }
    // (position in the bucket, items in the bucket) for the merge rounds and the finalize
    item_info[item] = make_uint2(idx, nitems);
}

// The accumulation of msm_accumulate_kernel with lanes cut at multiples of ch of the whole sorted
// list, as msm_accumulate29_kernel does for the G1 29-bit pipeline (there: why, and the layout of
// the partial sums): every lane of a wave runs ch additions; a lane stashes its finished sum in
// LDS where its range enters the next bucket and stores both after the loop.
template <class Fe, class Out = Fe>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void msm_accumulate_lanes_kernel(
                                                            const uint32_t* __restrict__ bases,
                                                            const uint32_t* __restrict__ endo_x, uint32_t n_real,
                                                            const uint32_t* __restrict__ sorted,
                                                            const uint32_t* __restrict__ starts,
                                                            const uint32_t* __restrict__ istarts, uint32_t S,
                                                            uint32_t ch, uint2* __restrict__ item_info,
                                                            uint32_t* __restrict__ item_pts) {
    typedef FeTraits<Fe> Tr;
    constexpr int XW = 4 * Tr::WORDS;  // words of one XYZZ point (the same in Out's memory form)
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = starts[S];
    const uint64_t pos0 = (uint64_t)lane * ch;
    if (pos0 >= total) return;
    const uint32_t p0 = (uint32_t)pos0, end = (uint32_t)min<uint64_t>(pos0 + ch, total);
    auto bucket_of = [&](uint32_t p, uint32_t from) {  // last slot in [from, S) starting at or before p
        uint32_t lo = from, hi = S;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (starts[mid] <= p) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    uint32_t slot = bucket_of(p0, 0), bend = starts[slot + 1];
    uint32_t nslot = slot, nbend = bend;
    if (bend < end) {
        nslot = bucket_of(bend, slot + 1);
        nbend = starts[nslot + 1];
    }
    auto emit = [&](uint32_t b, const Xyzz<Fe>& acc) {
        const uint32_t first = starts[b] / ch, last = (starts[b + 1] - 1) / ch;
        const uint32_t k = lane - first, nitems = last - first + 1, out = istarts[b] + k;
        if constexpr (std::is_same<Fe, Out>::value) {
            store_xyzz(item_pts + (size_t)XW * out, acc);
        } else {
            const Out* tag = nullptr;
            Xyzz<Out> o = xyzz_inf<Out>();
            if (!xyzz_is_inf(acc))
                o = Xyzz<Out>{to_storage_field(acc.x, tag), to_storage_field(acc.y, tag),
                              to_storage_field(acc.zz, tag), to_storage_field(acc.zzz, tag)};
            store_xyzz(item_pts + (size_t)XW * out, o);
        }
        item_info[out] = make_uint2(k, nitems);
        if (lane == last && istarts[b] + nitems < istarts[b + 1]) item_info[istarts[b] + nitems] = make_uint2(0u, 0u);
    };
    // the stashed sum of the previous bucket, word i at stash[256 i] (lanes on distinct banks)
    __shared__ uint32_t stash_lds[XW * 256];
    uint32_t* const stash = stash_lds + threadIdx.x;
    auto put = [&](const Xyzz<Fe>& v) {
        uint32_t w[XW];
        store_xyzz(w, v);
#pragma unroll
        for (int i = 0; i < XW; ++i) stash[256 * i] = w[i];
    };
    auto get = [&]() {
        uint32_t w[XW];
#pragma unroll
        for (int i = 0; i < XW; ++i) w[i] = stash[256 * i];
        return load_xyzz<Fe>(w);
    };
    Xyzz<Fe> acc = xyzz_inf<Fe>();
    bool stashed = false;
    uint32_t s_slot = 0;
    for (uint32_t pos = p0; pos < end; ++pos) {
        if (pos == bend) {  // entry pos starts the next non-empty bucket
            if (stashed) emit(s_slot, get());  // only with buckets shorter than ch
            put(acc);
            s_slot = slot;
            stashed = true;
            acc = xyzz_inf<Fe>();
            slot = nslot;
            bend = nbend;
            if (bend < end) {  // a third bucket in this lane's range
                nslot = bucket_of(bend, slot + 1);
                nbend = starts[nslot + 1];
            }
        }
        const uint32_t e = sorted[pos];
        const uint32_t* xp;
        const uint32_t* yp;
        point_address<Tr::WORDS>(e, n_real, bases, endo_x, &xp, &yp);
        Affine<Fe> p{Tr::load(xp), Tr::load(yp)};
        if (e >> 31) p.y = Tr::zero() - p.y;
        acc = xyzz_madd(acc, p);
    }
    if (stashed) emit(s_slot, get());
    emit(slot, acc);
}

// G1 accumulation with the field in nine 29-bit limbs (field29.hpp): the same work split and the
// same point law (madd-2008-s with every special case) as msm_accumulate_kernel<Fq>; the item sums
// are written in Fq29's memory form for the Fq29 instantiation of the later stages. Bases are read in R = 2^256 form and taken as 32x their
// value, which is the R = 2^261 form unreduced (below 32p). Value bounds in units of p, kept
// between points: X < 9, Y < 5, ZZ < 2, ZZZ < 2; every product's operand bounds multiply to at
// most 144 (the multiplication allows 169). A template so that the header defines it once per
// instantiation (only Fe = Fq29, the G1 pipeline in 29-bit limbs).
template <class Fe>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void msm_accumulate29_kernel(
                                                            const uint32_t* __restrict__ bases,
                                                            const uint32_t* __restrict__ endo_x, uint32_t n_real,
                                                            const uint32_t* __restrict__ sorted,
                                                            const uint32_t* __restrict__ starts,
                                                            const uint32_t* __restrict__ istarts, uint32_t S,
                                                            uint32_t ch, uint2* __restrict__ item_info,
                                                            uint32_t* __restrict__ item_pts) {
    static_assert(std::is_same<Fe, Fq29>::value, "29-bit limbs are for the base field of G1");
    // Lane k adds the sorted entries [k ch, (k + 1) ch) of all windows' buckets together: every lane
    // of a wave runs the same number of additions (cutting lanes at bucket starts instead leaves each
    // bucket's last, shorter item idle for the rest of its wave). Where its range crosses into the
    // next bucket the lane keeps the finished sum in a second accumulator and starts again; both
    // are stored after the loop (a lane that crosses twice, only possible with buckets shorter than
    // ch, stores the older one in the loop). A bucket's partial sums are those of the lanes its
    // entries meet, at slots istarts[b] + (k - the lane of its first entry).
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = starts[S];
    const uint64_t pos0 = (uint64_t)lane * ch;
    if (pos0 >= total) return;
    const uint32_t p0 = (uint32_t)pos0, end = (uint32_t)min<uint64_t>(pos0 + ch, total);
    // the bucket holding entry p: the last slot in [from, S) whose start is <= p (empty slots share
    // their start with the next, so the last one among equals is the non-empty bucket holding p).
    // A binary search: after a window whose digits use only part of its slots, thousands of
    // empty slots may separate two non-empty buckets.
    auto bucket_of = [&](uint32_t p, uint32_t from) {
        uint32_t lo = from, hi = S;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (starts[mid] <= p) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    uint32_t slot = bucket_of(p0, 0), bend = starts[slot + 1];
    // the next non-empty bucket and its end, found before the loop: a load inside the loop would
    // make the wait for it also wait for the point read ahead (lanes meeting three buckets, only
    // possible with buckets shorter than ch, look the third one up in the loop)
    uint32_t nslot = slot, nbend = bend;
    if (bend < end) {
        nslot = bucket_of(bend, slot + 1);
        nbend = starts[nslot + 1];
    }
    // stores the partial sum of bucket b held in limbs (X, Y, ZZ, ZZZ; infinity when inf)
    auto emit = [&](uint32_t b, const F29& X, const F29& Y, const F29& ZZ, const F29& ZZZ, bool inf) {
        const uint32_t s0 = starts[b], first = s0 / ch, last = (starts[b + 1] - 1) / ch;
        const uint32_t k = lane - first, nitems = last - first + 1, out = istarts[b] + k;
        Xyzz<Fq29> o = xyzz_inf<Fq29>();
        if (!inf) {
            const F29 one = f29_const_one();
            const F29 x = f29_mul(X, one), y = f29_mul(Y, one);  // value-preserving, below 2p
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                o.x.l[i] = x.l[i];
                o.y.l[i] = y.l[i];
                o.zz.l[i] = ZZ.l[i];
                o.zzz.l[i] = ZZZ.l[i];
            }
        }
        store_xyzz(item_pts + (size_t)32 * out, o);
        item_info[out] = make_uint2(k, nitems);
        // the lane of the bucket's last entry marks the bucket's spare slot, if any, as empty
        if (lane == last && istarts[b] + nitems < istarts[b + 1]) item_info[istarts[b] + nitems] = make_uint2(0u, 0u);
    };
    const F29 one = f29_const_one();
    F29 X = one, Y = one, ZZ = one, ZZZ = one;
    bool inf = true;
    // the stashed sum of the previous bucket: in LDS (36 words per lane), so that it takes no
    // registers from the additions
    __shared__ uint32_t stash_lds[36 * 256];
    uint32_t* const stash = stash_lds + threadIdx.x;  // word i at stash[256 i]: lanes on distinct banks
    bool s_inf = true, stashed = false;
    uint32_t s_slot = 0;
    auto put = [&](const F29& v, int at) {
#pragma unroll
        for (int i = 0; i < 9; ++i) stash[256 * (at + i)] = v.l[i];
    };
    auto get = [&](int at) {
        F29 v;
#pragma unroll
        for (int i = 0; i < 9; ++i) v.l[i] = stash[256 * (at + i)];
        return v;
    };
    // software pipelining: entry j + 1 is read while entry j is added (the gathered 64-byte reads
    // would otherwise wait behind each addition)
    uint32_t e_next = sorted[p0];
    Fq nx, ny;
    {
        const uint32_t* xp;
        const uint32_t* yp;
        point_address<8>(e_next, n_real, bases, endo_x, &xp, &yp);
        nx = fp_load<FqParams>(xp);
        ny = fp_load<FqParams>(yp);
    }
    for (uint32_t pos = p0; pos < end; ++pos) {
        const uint32_t e = e_next;
        const Fq qx = nx, qy = ny;
        if (pos + 1 < end) {
            e_next = sorted[pos + 1];
            const uint32_t* xp;
            const uint32_t* yp;
            point_address<8>(e_next, n_real, bases, endo_x, &xp, &yp);
            nx = fp_load<FqParams>(xp);
            ny = fp_load<FqParams>(yp);
        }
        if (pos == bend) {  // entry pos starts the next non-empty bucket
            if (stashed) emit(s_slot, get(0), get(9), get(18), get(27), s_inf);  // only with buckets shorter than ch
            put(X, 0);
            put(Y, 9);
            put(ZZ, 18);
            put(ZZZ, 27);
            s_inf = inf;
            s_slot = slot;
            stashed = true;
            inf = true;
            ZZ = one;
            ZZZ = one;
            slot = nslot;
            bend = nbend;
            if (bend < end) {  // a third bucket in this lane's range
                nslot = bucket_of(bend, slot + 1);
                nbend = starts[nslot + 1];
            }
        }
        if (fp_is_zero(qx) && fp_is_zero(qy)) continue;  // the point at infinity
        const F29 xs = f29_from_r256_words(qx.w), ys = f29_from_r256_words(qy.w);  // < 32
        const bool neg = e >> 31;
        // ZZ = ZZZ = 1 while the sum is empty, so these two products are then the point itself,
        // reduced: a lane starting a bucket shares the addition's first products with the others
        // instead of running a path of its own
        const F29 u2 = f29_mul(xs, ZZ);    // 32 * 2 -> < 2
        F29 s2 = f29_mul(ys, ZZZ);         // < 2
        if (neg) s2 = f29_neg<3>(s2);      // < 3
        if (inf) {
            X = u2;  // < 2
            Y = s2;  // < 3
            inf = false;
            continue;
        }
        const F29 pp = f29_sub<10>(u2, X);  // < 12
        const F29 r = f29_sub<6>(s2, Y);    // < 9
        if (f29_is_zero_below12p(pp)) {
            if (f29_is_zero_below12p(r)) {  // the same point: double (rare; in the 32-bit field)
                Xyzz<Fq> a{f29_to_fq(X), f29_to_fq(Y), f29_to_fq(ZZ), f29_to_fq(ZZZ)};
                a = xyzz_dbl(a);
                X = f29_from_fq(a.x);
                Y = f29_from_fq(a.y);
                ZZ = f29_from_fq(a.zz);
                ZZZ = f29_from_fq(a.zzz);
            } else {
                inf = true;  // P + (-P)
                ZZ = one;
                ZZZ = one;
            }
            continue;
        }
        const F29 ppsq = f29_mul(pp, pp);              // 12 * 12 -> < 2
        const F29 ppp = f29_mul(pp, ppsq);             // 12 * 2 -> < 2
        const F29 q = f29_mul(X, ppsq);                // 9 * 2 -> < 2
        const F29 rr = f29_mul(r, r);                  // 9 * 9 -> < 2
        const F29 t = f29_add(ppp, f29_add(q, q));     // < 6, limbs < 3 * 2^29
        const F29 x3 = f29_sub<7>(rr, t);              // < 9
        const F29 d = f29_sub<10>(q, x3);              // < 12
        const F29 y3 = f29_mul2<true>(r, d, Y, ppp, true);  // R D - Y PPP: (9 * 12 + 5 * 2) / 169 + 2 -> < 2.8
        ZZ = f29_mul(ZZ, ppsq);                        // < 2
        ZZZ = f29_mul(ZZZ, ppp);                       // < 2
        X = x3;
        Y = y3;
    }
    // item sums in the Fq29 memory form (R = 2^261 value below 2p), read by the Fq29 stages
    if (stashed) emit(s_slot, get(0), get(9), get(18), get(27), s_inf);
    emit(slot, X, Y, ZZ, ZZZ, inf);
}

// The same accumulation with work items cut at bucket starts (ceil(count / CH) items per bucket,
// one lane each): the default for plans sorting fewer than 2^23 entries, where the lane-cut kernel's
// second stored partial per lane and the finalize's extra partials outweigh the fuller waves.
template <class Fe>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void msm_accumulate29_bucket_kernel(
                                                            const uint32_t* __restrict__ bases,
                                                            const uint32_t* __restrict__ endo_x, uint32_t n_real,
                                                            const uint32_t* __restrict__ sorted,
                                                            const uint32_t* __restrict__ starts,
                                                            const uint32_t* __restrict__ istarts,
                                                            const uint32_t* __restrict__ counts, uint32_t S,
                                                            uint32_t ch, uint2* __restrict__ item_info,
                                                            uint32_t* __restrict__ item_pts) {
    static_assert(std::is_same<Fe, Fq29>::value, "29-bit limbs are for the base field of G1");
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= istarts[S]) return;
    uint32_t lo = 0, hi = S;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (istarts[mid] <= item) lo = mid;
        else hi = mid;
    }
    const uint32_t slot = lo;
    const uint32_t skip = (item - istarts[slot]) * ch;
    const uint32_t len = min(ch, counts[slot] - skip);
    const uint32_t* entries = sorted + starts[slot] + skip;
    const F29 one = f29_const_one();
    F29 X = one, Y = one, ZZ = one, ZZZ = one;
    bool inf = true;
    // software pipelining: point j + 1 is read while point j is added (the gathered 64-byte
    // reads would otherwise wait behind each addition; accumulate 1.63 -> 1.60 ms at c = 15,
    // 1.42 -> 1.34 ms at c = 16 for G1 2^20)
    uint32_t e_next = 0;
    Fq nx, ny;
    if (len) {
        e_next = entries[0];
        const uint32_t* xp;
        const uint32_t* yp;
        point_address<8>(e_next, n_real, bases, endo_x, &xp, &yp);
        nx = fp_load<FqParams>(xp);
        ny = fp_load<FqParams>(yp);
    }
    for (uint32_t j = 0; j < len; ++j) {
        const uint32_t e = e_next;
        const Fq qx = nx, qy = ny;
        if (j + 1 < len) {
            e_next = entries[j + 1];
            const uint32_t* xp;
            const uint32_t* yp;
            point_address<8>(e_next, n_real, bases, endo_x, &xp, &yp);
            nx = fp_load<FqParams>(xp);
            ny = fp_load<FqParams>(yp);
        }
        if (fp_is_zero(qx) && fp_is_zero(qy)) continue;  // the point at infinity
        const F29 xs = f29_from_r256_words(qx.w), ys = f29_from_r256_words(qy.w);  // < 32
        const bool neg = e >> 31;
        if (inf) {
            X = f29_mul(xs, one);  // < 2
            Y = f29_mul(ys, one);  // < 2
            if (neg) Y = f29_neg<3>(Y);  // < 3
            ZZ = one;
            ZZZ = one;
            inf = false;
            continue;
        }
        const F29 u2 = f29_mul(xs, ZZ);    // 32 * 2 -> < 2
        F29 s2 = f29_mul(ys, ZZZ);         // < 2
        if (neg) s2 = f29_neg<3>(s2);      // < 3
        const F29 pp = f29_sub<10>(u2, X);  // < 12
        const F29 r = f29_sub<6>(s2, Y);    // < 9
        if (f29_is_zero_below12p(pp)) {
            if (f29_is_zero_below12p(r)) {  // the same point: double (rare; in the 32-bit field)
                Xyzz<Fq> a{f29_to_fq(X), f29_to_fq(Y), f29_to_fq(ZZ), f29_to_fq(ZZZ)};
                a = xyzz_dbl(a);
                X = f29_from_fq(a.x);
                Y = f29_from_fq(a.y);
                ZZ = f29_from_fq(a.zz);
                ZZZ = f29_from_fq(a.zzz);
            } else {
                inf = true;  // P + (-P)
            }
            continue;
        }
        const F29 ppsq = f29_mul(pp, pp);              // 12 * 12 -> < 2
        const F29 ppp = f29_mul(pp, ppsq);             // 12 * 2 -> < 2
        const F29 q = f29_mul(X, ppsq);                // 9 * 2 -> < 2
        const F29 rr = f29_mul(r, r);                  // 9 * 9 -> < 2
        const F29 t = f29_add(ppp, f29_add(q, q));     // < 6, limbs < 3 * 2^29
        const F29 x3 = f29_sub<7>(rr, t);              // < 9
        const F29 d = f29_sub<10>(q, x3);              // < 12
        const F29 y3 = f29_mul2<true>(r, d, Y, ppp, true);  // R D - Y PPP: (9 * 12 + 5 * 2) / 169 + 2 -> < 2.8
        ZZ = f29_mul(ZZ, ppsq);                        // < 2
        ZZZ = f29_mul(ZZZ, ppp);                       // < 2
        X = x3;
        Y = y3;
    }
    // item sums in the Fq29 memory form (R = 2^261 value below 2p), read by the Fq29 stages
    Xyzz<Fq29> out = xyzz_inf<Fq29>();
    if (!inf) {
        const F29 x = f29_mul(X, one), y = f29_mul(Y, one);  // value-preserving, below 2p
        Fq29 o[4];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            o[0].l[i] = x.l[i];
            o[1].l[i] = y.l[i];
            o[2].l[i] = ZZ.l[i];
            o[3].l[i] = ZZZ.l[i];
        }
        out = Xyzz<Fq29>{o[0], o[1], o[2], o[3]};
    }
    store_xyzz(item_pts + (size_t)32 * item, out);
    item_info[item] = make_uint2(item - istarts[slot], istarts[slot + 1] - istarts[slot]);
}

// Round r (stride s = G^r): in a bucket with more than FINAL_GROUP partial sums left, item
// k = 0 mod s*G absorbs items k + s .. k + (G-1)s. G = 4 keeps each round's chain at three
// additions; typical buckets (up to FINAL_GROUP items) skip the rounds. The grid is bounded
// (MERGE_BLOCKS workgroups striding over the items), so a round that no bucket needs costs one
// small launch: at one lane per item slot, the idle rounds of a G1 2^20 plan took 0.09 ms.
template <class Fe>
__global__ __launch_bounds__(256) void msm_merge_kernel(uint32_t* __restrict__ item_pts,
                                                       const uint2* __restrict__ item_info,
                                                       const uint32_t* __restrict__ istarts, uint32_t S,
                                                       uint32_t stride, const uint32_t* __restrict__ max_items) {
    typedef FeTraits<Fe> Tr;
    // no bucket needs this round (the scan recorded the most items of any bucket): exit before
    // reading any item
    if ((max_items[0] + stride - 1) / stride <= FINAL_GROUP) return;
    constexpr uint32_t LN = FeLanes<Fe>::value;  // the lanes of a pair take the same item
    const uint32_t nitems = istarts[S], step = gridDim.x * blockDim.x / LN;
    const size_t xw = 4 * Tr::WORDS;
    for (uint32_t item = (blockIdx.x * blockDim.x + threadIdx.x) / LN; item < nitems; item += step) {
        const uint2 info = item_info[item];
        const uint32_t k = info.x, total = info.y;
        if ((total + stride - 1) / stride <= FINAL_GROUP) continue;
        if (k % (stride * SUM_GROUP) != 0) continue;
        Xyzz<Fe> acc = load_xyzz<Fe>(item_pts + xw * item);
        for (uint32_t j = 1; j < SUM_GROUP; ++j) {
            if (k + j * stride >= total) break;
            acc = xyzz_add(acc, load_xyzz<Fe>(item_pts + xw * (item + j * stride)));
        }
        store_xyzz(item_pts + xw * item, acc);
    }
}

template <class Fe>
struct LaneOps;

// 2^split lanes per bucket (split 0, 1 or 2): lane 2^split slot + h sums part h of the bucket's
// at most 16 remaining item sums, and the lanes of a bucket add their parts through lane shuffles
// (split rounds), which shortens the serial chain and multiplies the waves.
template <class Fe>
__global__ __launch_bounds__(256) void msm_finalize_kernel(const uint32_t* __restrict__ item_pts,
                                                          const uint32_t* __restrict__ istarts,
                                                          const uint2* __restrict__ item_info, uint32_t S,
                                                          uint32_t split, uint32_t* __restrict__ buckets) {
    typedef FeTraits<Fe> Tr;
    constexpr uint32_t LN = FeLanes<Fe>::value;  // lanes per element (the lanes of a pair stay together)
    const uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) / LN;
    const uint32_t slot = t >> split, part = t & ((1u << split) - 1);
    const bool valid = slot < S;  // every lane reaches the shuffles
    const size_t xw = 4 * Tr::WORDS;
    Xyzz<Fe> acc = xyzz_inf<Fe>();
    if (valid) {
        // the bucket's item count (its slots may include a spare one): its first slot's info
        const uint32_t first = istarts[slot], total = istarts[slot + 1] > first ? item_info[first].y : 0u;
        uint32_t stride = 1;
        while ((total + stride - 1) / stride > FINAL_GROUP) stride *= SUM_GROUP;
        const uint32_t m = (total + stride - 1) / stride;        // terms k * stride, k < m
        const uint32_t share = (m + (1u << split) - 1) >> split;  // terms per part
        const uint32_t lo = min(m, part * share), hi = min(m, lo + share);
        if (lo < hi) {
            // item k + 1 is read while item k is added (half the kernel's wave cycles waited on the
            // item reads with one read per addition)
            acc = load_xyzz<Fe>(item_pts + xw * (first + (size_t)lo * stride));
            Xyzz<Fe> next = acc;
            if (lo + 1 < hi) next = load_xyzz<Fe>(item_pts + xw * (first + (size_t)(lo + 1) * stride));
            for (uint32_t k = lo + 1; k < hi; ++k) {
                const Xyzz<Fe> cur = next;
                if (k + 1 < hi) next = load_xyzz<Fe>(item_pts + xw * (first + (size_t)(k + 1) * stride));
                acc = xyzz_add(acc, cur);
            }
        }
    }
    for (uint32_t r = 0; r < split; ++r) {  // uniform over the grid
        typedef LaneOps<Fe> Lx;
        const int src = (int)(__lane_id() ^ (LN << r));
        const Xyzz<Fe> other{Lx::shfl(acc.x, src), Lx::shfl(acc.y, src), Lx::shfl(acc.zz, src),
                             Lx::shfl(acc.zzz, src)};
        acc = xyzz_add(acc, other);
    }
    if (valid && part == 0) store_xyzz(buckets + xw * slot, acc);
}

// ---------------------------------------------------------------- bucket reduction

// Waves per SIMD the reduction kernel is compiled for (1: no constraint).
template <class Fe>
struct ReduceWaves {
    static constexpr int value = 1;
};
template <>
struct ReduceWaves<Fq2P> {
    static constexpr int value = 2;
};
template <>
struct ReduceWaves<Fq2Q> {
    static constexpr int value = 2;
};

// Lane t of window w takes buckets b0 .. b0+L-1 (b0 = j*L) and computes
// sum_k (k+1) bucket[b0+k] + b0 * sum_k bucket[b0+k]  =  sum_k (b0+k+1) bucket[b0+k];
// the lanes of a workgroup (all in one window) then add their results in an LDS tree, leaving
// one partial sum per workgroup. Each loop has a single point-addition site (operands chosen by
// selects), which keeps the fully inlined kernel small.
// The grid holds only workgroups with buckets: windows w < nw (wb + 1 digit bits, 2^wb buckets)
// take gh workgroups each, the others (2^(wb-1) buckets, half of B with the GLV split's balanced
// widths) gl; window w's partials sit at w * P. Workgroups of bucket ranges beyond a window's
// digits would be idle, and placed round-robin they doubled up busy ones on a CU.
template <class Fe>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ReduceWaves<Fe>::value))) void msm_reduce_kernel(const uint32_t* __restrict__ buckets, uint32_t W,
                                                        uint32_t B, uint32_t L, uint32_t c, uint32_t wb, uint32_t nw,
                                                        uint32_t gh, uint32_t gl, uint32_t P,
                                                        uint32_t* __restrict__ partials) {
    typedef FeTraits<Fe> Tr;
    constexpr int XW = 4 * Tr::WORDS;
    // LDS slots padded to XW + 4 words: at XW words (a multiple of the 64 banks' half) the lanes of
    // a wave storing or loading their points would share banks in pairs of lanes two apart
    constexpr int XP = XW + 4;
    __shared__ uint32_t lds[128 * XP];
    constexpr uint32_t LN = FeLanes<Fe>::value;      // lanes per element
    const uint32_t tid = threadIdx.x / LN, nthr = blockDim.x / LN;  // logical threads
    const uint32_t hi_wgs = nw * gh;
    const uint32_t w = blockIdx.x < hi_wgs ? blockIdx.x / gh : nw + (blockIdx.x - hi_wgs) / gl;
    const uint32_t g = blockIdx.x < hi_wgs ? blockIdx.x % gh : (blockIdx.x - hi_wgs) % gl;
    const uint32_t j = g * nthr + tid, b0 = j * L;
    Xyzz<Fe> r = xyzz_inf<Fe>();
    if (b0 < (1u << (wb + (w < nw ? 1u : 0u) - 1))) {  // else beyond this window's digit range: empty
        const uint32_t* src = buckets + (size_t)XW * ((size_t)w * B + b0);
        Xyzz<Fe> run = xyzz_inf<Fe>(), tot = xyzz_inf<Fe>();
        // steps 2k: run += bucket[L-1-k]; steps 2k+1: tot += run
        for (uint32_t step = 0; step < 2 * L; ++step) {
            const bool odd = step & 1;
            const Xyzz<Fe> x = select_xyzz(odd, tot, run);
            const Xyzz<Fe> y = select_xyzz(odd, run, load_xyzz<Fe>(src + XW * (L - 1 - step / 2)));
            const Xyzz<Fe> z = xyzz_add(x, y);
            run = select_xyzz(odd, run, z);
            tot = select_xyzz(odd, z, tot);
        }
        // r = b0 * run by double-and-add over c bits, then r + tot in one extra iteration
        for (int k = (int)c - 1; k >= -1; --k) {
            if (k >= 0 && !xyzz_is_inf(r)) r = xyzz_dbl(r);
            if (k < 0 || ((b0 >> k) & 1u)) r = xyzz_add(r, select_xyzz(k < 0, tot, run));
        }
    }
    for (uint32_t off = nthr / 2; off > 0; off >>= 1) {
        if (tid >= off && tid < 2 * off) store_xyzz(lds + XP * (tid - off), r);
        __syncthreads();
        if (tid < off) r = xyzz_add(r, load_xyzz<Fe>(lds + XP * tid));
        __syncthreads();
    }
    if (tid == 0) store_xyzz(partials + (size_t)XW * ((size_t)w * P + g), r);
}

// One wave per window adds the window's workgroup partials (gh for windows w < nw, else gl; at
// most 64 at c <= 16 needs no serial part, more add several per lane first), then an LDS tree of
// `top` levels. Window w's partials start at w * P.
template <class Fe>
__global__ __launch_bounds__(64) void msm_window_sum_kernel(const uint32_t* __restrict__ partials, uint32_t P,
                                                           uint32_t nw, uint32_t gh, uint32_t gl, uint32_t top,
                                                           uint32_t* __restrict__ wsums) {
    typedef FeTraits<Fe> Tr;
    constexpr int XW = 4 * Tr::WORDS;
    constexpr int XP = XW + 4;  // padded LDS slots, as in msm_reduce_kernel
    __shared__ uint32_t lds[32 * XP];
    constexpr uint32_t LN = FeLanes<Fe>::value;                       // lanes per element
    const uint32_t tid = threadIdx.x / LN, nthr = blockDim.x / LN;  // logical threads
    const uint32_t w = blockIdx.x, pw = w < nw ? gh : gl;
    const uint32_t* src = partials + (size_t)XW * w * P;
    Xyzz<Fe> acc = tid < pw ? load_xyzz<Fe>(src + XW * tid) : xyzz_inf<Fe>();
    for (uint32_t k = tid + nthr; k < pw; k += nthr) acc = xyzz_add(acc, load_xyzz<Fe>(src + XW * k));
    for (uint32_t off = top; off > 0; off >>= 1) {
        if (tid >= off && tid < 2 * off) store_xyzz(lds + XP * (tid - off), acc);
        __syncthreads();
        if (tid < off) acc = xyzz_add(acc, load_xyzz<Fe>(lds + XP * tid));
        __syncthreads();
    }
    if (tid == 0) store_xyzz(wsums + (size_t)XW * w, acc);
}

// ---------------------------------------------------------------- window combine
// result = sum_w 2^(pos_w) S_w (pos_w = bit offset of window w) by Horner's rule from the top
// window: pos_(W-1) doublings (240 for c = 15 without GLV, 114 with it) in one
// dependency chain, which no amount of parallelism across points can shorten. What can be
// shortened is the latency of each step: the multiplications of one doubling fall into three
// levels of independent products (V = U^2, X^2 | W = UV, S = XV, ZZ3 = V ZZ, M^2 |
// M (S - X3), W Y, W ZZZ), so lanes of the wave each compute one product per level (for G2 three
// lanes per Fq2 product, one Fq product of Karatsuba's formula each) and the results are
// broadcast with lane shuffles. A doubling then costs three multiplication latencies instead of
// nine; the additions of the chain use four levels of up to six products.

template <class Fe>
struct LaneOps;

// LaneOps<Fe>: SPREAD lanes compute one Fe product (1 for Fq, 3 for Fq2: the three Fq products
// of Karatsuba's formula on three lanes); partial(j, a, b) is lane j's share and gather(m, k)
// assembles product k on every lane from the shares of lanes SPREAD k .. SPREAD k + SPREAD - 1.
template <class Fp1>
struct LaneOps1 {
    static constexpr int SPREAD = 1;
    static ZKG_DEV Fp1 partial(uint32_t, const Fp1& a, const Fp1& b) { return a * b; }
    static ZKG_DEV Fp1 gather(const Fp1& m, int k) { return LaneOps<Fp1>::shfl(m, k); }
};

template <class Fp1, class Fp2>
struct LaneOps2 {
    static constexpr int SPREAD = 3;
    // lane share j of (a.a + a.b u)(b.a + b.b u): a.a b.a, a.b b.b or (a.a + a.b)(b.a + b.b), in .a
    static ZKG_DEV Fp2 partial(uint32_t j, const Fp2& a, const Fp2& b) {
        const Fp1 x = fe_select(j == 0, a.a, fe_select(j == 1, a.b, a.a + a.b));
        const Fp1 y = fe_select(j == 0, b.a, fe_select(j == 1, b.b, b.a + b.b));
        return Fp2{x * y, x};
    }
    static ZKG_DEV Fp2 gather(const Fp2& m, int k) {
        const Fp1 t0 = LaneOps<Fp1>::shfl(m.a, 3 * k), t1 = LaneOps<Fp1>::shfl(m.a, 3 * k + 1);
        const Fp1 t2 = LaneOps<Fp1>::shfl(m.a, 3 * k + 2);
        return Fp2{t0 - t1, (t2 - t0) - t1};
    }
};

template <>
struct LaneOps<Fq> : LaneOps1<Fq> {
    static ZKG_DEV Fq shfl(const Fq& x, int src) {
        Fq r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r.w[i] = __shfl(x.w[i], src);
        return r;
    }
};

template <>
struct LaneOps<Fq29> : LaneOps1<Fq29> {
    static ZKG_DEV Fq29 shfl(const Fq29& x, int src) {
        Fq29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.l[i] = __shfl(x.l[i], src);
        return r;
    }
};

template <>
struct LaneOps<Fq2_29> : LaneOps2<Fq29, Fq2_29> {
    static ZKG_DEV Fq2_29 shfl(const Fq2_29& x, int src) {
        return Fq2_29{LaneOps<Fq29>::shfl(x.a, src), LaneOps<Fq29>::shfl(x.b, src)};
    }
};

template <>
struct LaneOps<Fq2P> {
    // lane shuffles of the G2 tail stages: each lane moves its own component
    static ZKG_DEV Fq2P shfl(const Fq2P& x, int src) { return Fq2P{LaneOps<Fq29>::shfl(x.v, src)}; }
    // products spread over lane pairs: pair k (lanes 2k, 2k + 1) computes product k, each lane its
    // own component (two Fq products), and every lane fetches its component of product k from
    // the lane of its parity in pair k with one shuffle
    static constexpr int SPREAD = 2;
    static ZKG_DEV Fq2P partial(uint32_t, const Fq2P& a, const Fq2P& b) { return a * b; }
    static ZKG_DEV Fq2P gather(const Fq2P& m, int k) {  // k wave-uniform: two lane reads per word
        const bool odd = (__lane_id() & 1u) != 0;
        Fq29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)m.v.l[i], 2 * k);
            const uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)m.v.l[i], 2 * k + 1);
            r.l[i] = odd ? o : e;
        }
        return Fq2P{r};
    }
};

template <>
struct LaneOps<Fq29Pair> {
    // lane shuffles between pairs (sources at the same position of their pair)
    static ZKG_DEV Fq29Pair shfl(const Fq29Pair& x, int src) { return Fq29Pair{LaneOps<Fq29>::shfl(x.v, src)}; }
};

template <>
struct LaneOps<Fq2Q> {
    // lane shuffles between groups of four lanes (sources at the same position of their group)
    static ZKG_DEV Fq2Q shfl(const Fq2Q& x, int src) { return Fq2Q{LaneOps<Fq2P>::shfl(x.v, src)}; }
};

template <>
struct LaneOps<Fq2> : LaneOps2<Fq, Fq2> {
    static ZKG_DEV Fq2 shfl(const Fq2& x, int src) { return Fq2{LaneOps<Fq>::shfl(x.a, src), LaneOps<Fq>::shfl(x.b, src)}; }
};

// pick(lane, a0, a1, ...) = a_lane (the last candidate for lanes beyond the list)
template <class Fe>
ZKG_DEV Fe pick(uint32_t lane, const Fe& a0, const Fe& a1) { return fe_select(lane == 0, a0, a1); }
template <class Fe>
ZKG_DEV Fe pick(uint32_t lane, const Fe& a0, const Fe& a1, const Fe& a2) {
    return fe_select(lane == 0, a0, fe_select(lane == 1, a1, a2));
}
template <class Fe>
ZKG_DEV Fe pick(uint32_t lane, const Fe& a0, const Fe& a1, const Fe& a2, const Fe& a3) {
    return fe_select(lane == 0, a0, fe_select(lane == 1, a1, fe_select(lane == 2, a2, a3)));
}
template <class Fe>
ZKG_DEV Fe pick(uint32_t lane, const Fe& a0, const Fe& a1, const Fe& a2, const Fe& a3, const Fe& a4, const Fe& a5) {
    return fe_select(lane == 0, a0,
                     fe_select(lane == 1, a1, fe_select(lane == 2, a2, fe_select(lane == 3, a3, fe_select(lane == 4, a4, a5)))));
}

// 2P with the products of each level spread over lanes 0..3 (all lanes hold P).
template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_dbl_lanes(const Xyzz<Fe>& p, uint32_t lane) {
    typedef FeTraits<Fe> T;
    typedef LaneOps<Fe> Lx;
    const uint32_t k = lane / Lx::SPREAD, j = lane % Lx::SPREAD;  // product k, share j
    const Fe u = p.y + p.y;
    Fe a = pick(k, u, p.x);
    Fe m = Lx::partial(j, a, a);
    const Fe v = Lx::gather(m, 0), x2 = Lx::gather(m, 1);
    const Fe mm = (x2 + x2) + x2;
    a = pick(k, u, p.x, v, mm);
    Fe b = pick(k, v, v, p.zz, mm);
    m = Lx::partial(j, a, b);
    const Fe w = Lx::gather(m, 0), s = Lx::gather(m, 1), zz3 = Lx::gather(m, 2), m2 = Lx::gather(m, 3);
    const Fe x3 = m2 - (s + s);
    a = pick(k, mm, w, w);
    b = pick(k, s - x3, p.y, p.zzz);
    m = Lx::partial(j, a, b);
    const Fe y3 = Lx::gather(m, 0) - Lx::gather(m, 1), zzz3 = Lx::gather(m, 2);
    (void)T::WORDS;
    return Xyzz<Fe>{x3, y3, zz3, zzz3};
}

// P + Q (both XYZZ, all lanes hold both) with every special case, products spread over lanes.
template <class Fe>
ZKG_DEV Xyzz<Fe> xyzz_add_lanes(const Xyzz<Fe>& p, const Xyzz<Fe>& q, uint32_t lane) {
    typedef FeTraits<Fe> T;
    typedef LaneOps<Fe> Lx;
    const uint32_t k = lane / Lx::SPREAD, j = lane % Lx::SPREAD;  // product k, share j
    if (xyzz_is_inf(q)) return p;
    if (xyzz_is_inf(p)) return q;
    Fe a = pick(k, p.x, q.x, p.y, q.y, p.zz, p.zzz);
    Fe b = pick(k, q.zz, p.zz, q.zzz, p.zzz, q.zz, q.zzz);
    Fe m = Lx::partial(j, a, b);
    const Fe u1 = Lx::gather(m, 0), u2 = Lx::gather(m, 1), s1 = Lx::gather(m, 2), s2 = Lx::gather(m, 3);
    const Fe zz12 = Lx::gather(m, 4), zzz12 = Lx::gather(m, 5);
    const Fe pp = u2 - u1, r = s2 - s1;
    if (T::is_zero(pp)) {
        if (T::is_zero(r)) return xyzz_dbl_lanes(p, lane);
        return xyzz_inf<Fe>();
    }
    a = pick(k, pp, r);
    m = Lx::partial(j, a, a);
    const Fe ppsq = Lx::gather(m, 0), r2 = Lx::gather(m, 1);
    a = pick(k, pp, u1, zz12);
    m = Lx::partial(j, a, ppsq);
    const Fe ppp = Lx::gather(m, 0), qq = Lx::gather(m, 1), zz3 = Lx::gather(m, 2);
    const Fe x3 = (r2 - ppp) - (qq + qq);
    a = pick(k, r, s1, zzz12);
    b = pick(k, qq - x3, ppp, ppp);
    m = Lx::partial(j, a, b);
    const Fe y3 = Lx::gather(m, 0) - Lx::gather(m, 1), zzz3 = Lx::gather(m, 2);
    return Xyzz<Fe>{x3, y3, zz3, zzz3};
}

// The field the combine's doubling chain runs in: G2's window sums (Fq2_29) are converted to the
// 32-bit Fq2 on load, whose Karatsuba products spread over three lanes make the chain shorter
// (G2 2^18: 0.94 ms, against 1.05 ms in Fq2_29 and 1.16 ms for Fq2_29 spread over lanes).
template <class Fe>
struct CombineField {
    using type = Fe;
    static ZKG_DEV Fe conv(const Fe& x) { return x; }
};

template <class Fe>
ZKG_DEV Xyzz<typename CombineField<Fe>::type> load_combine(const uint32_t* p) {
    return load_xyzz<typename CombineField<Fe>::type>(p);
}

// The chain of the 32-bit pipelines (ZKGPU_MSM_F29=0: Fq, and Fq2 with three lanes per product);
// the 29-bit pipelines run msm_combine29_kernel (G1) and msm_combine_g2_lanes_kernel (G2).
template <class Fe>
__global__ __launch_bounds__(64) void msm_combine_kernel(const uint32_t* __restrict__ wsums, uint32_t W, uint32_t wb,
                                                        uint32_t nw, uint32_t* __restrict__ result) {
    typedef FeTraits<Fe> Tr;
    typedef typename CombineField<Fe>::type Fc;
    constexpr int XW = 4 * Tr::WORDS;
    const uint32_t lane = threadIdx.x;
    // every lane loads the same sums through a lane-dependent address (vector registers)
    const uint32_t zero = lane >> 6;
    Xyzz<Fc> acc = load_combine<Fe>(wsums + (size_t)XW * (W - 1 + zero));
    for (int w = (int)W - 2; w >= 0; --w) {
        const uint32_t width = wb + ((uint32_t)w < nw ? 1u : 0u);  // bits between window w and w + 1
        for (uint32_t k = 0; k < width; ++k)
            if (!xyzz_is_inf(acc)) acc = xyzz_dbl_lanes(acc, lane);
        acc = xyzz_add_lanes(acc, load_combine<Fe>(wsums + (size_t)XW * (w + zero)), lane);
    }
    if (lane < FeLanes<Fc>::value) store_xyzz_result(result, acc);
}

// ---------------------------------------------------------------- G1 combine on lazy 29-bit limbs
// The same Horner chain for the 29-bit G1 pipeline, with the values kept lazily reduced between
// products as in msm_accumulate29_kernel (no renormalisation after each addition or subtraction,
// which the reduced Fq29 type pays for). Bounds in units of p, the product's limit being 169: the
// running point keeps X < 9, Y < 5, ZZ, ZZZ < 2 through every doubling and addition; window sums
// (Fq29 storage) are below 2.

// limbs 0..7 below 2^29 again (the value is unchanged)
ZKG_DEV F29 f29_carry(F29 a) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a.l[i + 1] += a.l[i] >> 29;
        a.l[i] &= f29::MASK;
    }
    return a;
}

// ---------------------------------------------------------------- limb-per-lane point chains
// The combines' Horner chains (and the window sums' short chains of additions) run one dependent
// chain per wave, so their cost is that chain's instruction count. Here each field element is one
// VGPR: lane i of a 16-lane row holds limb i (lanes 9..15 hold 0), and the four rows of the wave
// carry four independent products. A product then costs about 115 instructions for four (nine
// multiply-adds per column, the operands moved by DPP row shifts and row broadcasts) instead of
// about 230 for each product held on one lane.
//
// Product (Montgomery, R = 2^261, same contract as f29_mul: limbs < 2^30 + 2^8, v(a) v(b) below
// 169 p^2, result below a b / 2^261 + p (1 + 2^-50), limbs 0..7 below 2^29 + 2^7 + 1):
//   T = a b as seventeen 64-bit columns, column j on lane j (j < 16) and column 16 as H, a value
//       every lane of the row holds (each column at most nine products below 2^60.1);
//   q = (T mod 2^261) (-p^-1) mod 2^261, from T's low nine columns after two carry rounds;
//   U = T + q p, whose low 261 bits are 0 mod 2^261; after two carry rounds (lane 15's carries
//       go into H) limbs 0..8 are below 2^30, so their value is 0 or exactly 2^261, and it is
//       2^261 iff one of them is non-zero: that bit is the carry into the result's limb 0;
//   result = (U_9 .. U_15 moved down nine lanes, H's low 29 bits, H's high bits) + carry.
// H stays below 2^53 because H 2^203 is at most the result (below 3p). Every bound was checked on
// a lane-level model of these steps (tests/test_combine_lanes_model.py: random and extreme limbs
// against integer arithmetic, and the point laws below against the textbook ones).
namespace lpl {

template <int CTRL>
ZKG_DEV uint32_t dpp(uint32_t v) {  // lanes with no source lane in the row read 0 (bound_ctrl)
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, true);
}
template <int K>
ZKG_DEV uint32_t shr(uint32_t v) {  // lane i of a row takes lane i - K (row_shr)
    if constexpr (K == 0) return v;
    else return dpp<0x110 + K>(v);
}
template <int K>
ZKG_DEV uint32_t bc(uint32_t v) {  // every lane of a row takes the row's lane K (row_newbcast)
    return dpp<0x150 + K>(v);
}
ZKG_DEV uint64_t shr1_64(uint64_t v) {
    return ((uint64_t)shr<1>((uint32_t)(v >> 32)) << 32) | shr<1>((uint32_t)v);
}
ZKG_DEV uint64_t bc15_64(uint64_t v) {
    return ((uint64_t)bc<15>((uint32_t)(v >> 32)) << 32) | bc<15>((uint32_t)v);
}

// -p^-1 mod 2^261 in 29-bit limbs
ZKG_DEV constexpr uint32_t PINV(int i) {
    return i == 0 ? 0x04866389u : i == 1 ? 0x1e903c17u : i == 2 ? 0x129ab261u : i == 3 ? 0x1cfaca3du
         : i == 4 ? 0x1da809edu : i == 5 ? 0x05e80c19u : i == 6 ? 0x11af62bfu : i == 7 ? 0x16f23111u : 0x0ff57a22u;
}

// acc + sum over k of shr<k>(x) * y_k, with y_k = bc<k>(b) or the constant limbs C(k) (one chain of
// multiply-adds: three independent chains measured no faster)
template <int K = 0, class Y>
ZKG_DEV uint64_t columns(uint64_t acc, uint32_t x, const Y& y) {
    acc += (uint64_t)shr<K>(x) * y.template get<K>();
    if constexpr (K < 8) return columns<K + 1>(acc, x, y);
    else return acc;
}
struct BcastOf {
    uint32_t b;
    template <int K>
    ZKG_DEV uint32_t get() const { return bc<K>(b); }
};
struct PinvLimbs {
    template <int K>
    ZKG_DEV uint32_t get() const { return PINV(K); }
};
struct PLimbs {
    template <int K>
    ZKG_DEV uint32_t get() const { return f29::P(K); }
};

// two carry rounds of 64-bit columns into limbs below 2^29 + 2^7 (carries out of lane 15 dropped)
ZKG_DEV uint32_t carry2(uint64_t c) {
    const uint64_t c1 = (c & f29::MASK) + shr1_64(c >> 29);  // < 2^29 + 2^35
    return ((uint32_t)c1 & f29::MASK) + shr<1>((uint32_t)(c1 >> 29));
}

// Where a lane sits: its row (0..3) and its lane within the row, as the flags the selects use
// (written as two-way selects so that the compiler keeps them branch-free).
struct Where {
    uint32_t lane, rl;
    bool row_b0, row_b1, low8, is8, is7, is0;
    ZKG_DEV explicit Where(uint32_t l)
        : lane(l), rl(l & 15), row_b0((l & 16) != 0), row_b1((l & 32) != 0), low8((l & 15) < 8), is8((l & 15) == 8),
          is7((l & 15) == 7), is0((l & 15) == 0) {}
};

// Montgomery reduction of the 64-bit columns t (lane j: column j) and h (column 16, the same on
// every lane of a row): limbs 0..7 of the result below 2^29 + 2^7 + 1, limb 8 below 2^24
ZKG_DEV uint32_t reduce(const uint64_t t, uint64_t h, const Where& at) {
    const uint32_t tn = carry2(t);                             // = T mod 2^261 on lanes 0..8
    uint32_t q = carry2(columns(0, tn, PinvLimbs{}));
    q = at.low8 ? q : (at.is8 ? q & f29::MASK : 0u);           // q mod 2^261, lanes 9..15 cleared
    const uint64_t u = columns(t, q, PLimbs{});
    h += (uint64_t)bc<8>(q) * f29::P(8);
    const uint64_t hi = u >> 29;
    const uint64_t c1 = (u & f29::MASK) + shr1_64(hi);
    h += bc15_64(hi);
    const uint32_t hi2 = (uint32_t)(c1 >> 29);
    const uint32_t uu = ((uint32_t)c1 & f29::MASK) + shr<1>(hi2);
    h += bc<15>(hi2);
    // per row: is one of lanes 0..8 non-zero? (bit 9 of each row's 16-bit group after adding 0x1ff)
    constexpr uint64_t LOW9 = 0x01ff01ff01ff01ffull;
    const uint64_t nz = (__builtin_amdgcn_ballot_w64(uu != 0) & LOW9) + LOW9;
    const uint32_t carry = at.is0 ? (uint32_t)((nz >> 9) >> at.lane) & 1u : 0u;
    const uint32_t r = dpp<0x100 + 9>(uu) + carry;             // row_shl:9: limbs 9..15 to lanes 0..6
    const uint32_t r78 = at.is7 ? (uint32_t)h & f29::MASK : (uint32_t)(h >> 29);
    return at.is7 || at.is8 ? r78 : r;
}

ZKG_DEV uint32_t mul(uint32_t a, uint32_t b, const Where& at) {
    return reduce(columns(0, a, BcastOf{b}), (uint64_t)bc<8>(a) * bc<8>(b), at);
}

ZKG_DEV uint32_t carry1(uint32_t v) {  // one carry round (no carry leaves lane 8 for values < 2^261)
    return (v & f29::MASK) + shr<1>(v >> 29);
}
template <int C>
ZKG_DEV uint32_t kp(uint32_t rl) {  // lane rl's limb of C p in borrowed limbs (f29::K)
    uint32_t k = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) k = rl == (uint32_t)i ? f29::K<C>(i) : k;
    return k;
}

struct Rows {
    uint32_t r0, r1, r2, r3;
};
// each row's value replicated over all four rows
ZKG_DEV Rows rows(uint32_t v) {
    const auto s = __builtin_amdgcn_permlane16_swap(v, v, false, false);  // rows (0 0 2 2), (1 1 3 3)
    const auto e = __builtin_amdgcn_permlane32_swap(s[0], s[0], false, false);
    const auto o = __builtin_amdgcn_permlane32_swap(s[1], s[1], false, false);
    return Rows{e[0], o[0], e[1], o[1]};
}
ZKG_DEV uint32_t pick(const Where& at, uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3) {  // row r: a_r
    const uint32_t lo = at.row_b0 ? a1 : a0, hi = at.row_b0 ? a3 : a2;
    return at.row_b1 ? hi : lo;
}

struct SubConsts {  // C p in borrowed limbs for the subtractions of the doubling and the addition
    uint32_t k3, k6, k7, k10;
    ZKG_DEV explicit SubConsts(uint32_t rl) : k3(kp<3>(rl)), k6(kp<6>(rl)), k7(kp<7>(rl)), k10(kp<10>(rl)) {}
};

struct Point {  // XYZZ, every coordinate replicated over the four rows
    uint32_t x, y, zz, zzz;
};

// 2P (dbl-2008-s-1) for P finite with X < 9, Y < 5, ZZ, ZZZ < 2 (2P: X < 8, Y < 5, ZZ, ZZZ < 2). W Y
// and ZZZ3 are taken as V (U Y) and V (U ZZZ), so that three rounds of four products suffice:
//   round 1: V = U^2 (U = 2Y), X^2, U Y, U ZZZ        (10 * 10, 8 * 8, 10 * 5, 10 * 2)
//   round 2: S = X V, ZZ3 = V ZZ, M^2 (M = 3X^2), W Y  (8 * 2, 2 * 2, 6 * 6, 2 * 2)
//   round 3: M D (D = S - X3), ZZZ3 = V (U ZZZ)       (6 * 12, 2 * 2)
ZKG_DEV Point dbl(const Point& p, const Where& at, const SubConsts& k) {
    const uint32_t u = p.y + p.y;                                      // < 10, limbs < 2^30 + 2^8
    const Rows m1 = rows(mul(pick(at, u, p.x, u, u), pick(at, u, p.x, p.y, p.zzz), at));
    const uint32_t v = m1.r0, uy = m1.r2, uz = m1.r3;
    const uint32_t m = carry1(m1.r1 + m1.r1 + m1.r1);                  // M = 3 X^2 < 6
    const Rows m2 = rows(mul(pick(at, p.x, v, m, v), pick(at, v, p.zz, m, uy), at));
    const uint32_t s = m2.r0;
    const uint32_t x3 = carry1(m2.r2 + k.k6 - (s + s));                // M^2 + 6p - 2S < 8
    const uint32_t d = carry1(s + k.k10 - x3);                         // < 12
    const Rows m3 = rows(mul(at.row_b0 ? v : m, at.row_b0 ? uz : d, at));
    return Point{x3, carry1(m3.r0 + k.k3 - m2.r3), m2.r1, m3.r1};      // Y3 = M D + 3p - W Y < 5
}

ZKG_DEV uint32_t from_f29(const F29& a, uint32_t rl) {  // a's limbs (the same on every lane)
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) v = rl == (uint32_t)i ? a.l[i] : v;
    return v;
}
ZKG_DEV F29 to_f29(uint32_t v, int row = 0) {  // row's value, limbs 0..7 below 2^29 again, on every lane
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = (uint32_t)__builtin_amdgcn_readlane((int)v, 16 * row + i);
    return f29_carry(r);
}

// P + Q (add-2008-s with its special cases) for P with X < 9, Y < 5, ZZ, ZZZ < 2 and Q with
// coordinates below 2 (the sum: X < 9, Y < 5, ZZ, ZZZ < 2), in four rounds of products:
//   round 1: U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1      (9 * 2, 2 * 2, 5 * 2, 2 * 2)
//   round 2: PP = P^2, R^2, ZZ1 ZZ2, ZZZ1 ZZZ2  (P = U2 - U1, R = S2 - S1)   (5 * 5, 5 * 5, 2 * 2, 2 * 2)
//   round 3: PPP = P PP, Q = U1 PP, ZZ3 = ZZ1 ZZ2 PP                   (5 * 2, 2 * 2, 2 * 2)
//   round 4: R (Q - X3), S1 PPP, ZZZ3 = ZZZ1 ZZZ2 PPP                  (5 * 12, 2 * 2, 2 * 2)
// Returns 0 with the sum in *out, 1 if the sum is infinity (Q = -P), 2 if Q = P (the caller
// doubles P).
ZKG_DEV int add(const Point& p, const Point& q, const Where& at, const SubConsts& k, Point* out) {
    const Rows m1 = rows(mul(pick(at, p.x, q.x, p.y, q.y), pick(at, q.zz, p.zz, q.zzz, p.zzz), at));
    const uint32_t pp = carry1(m1.r1 + k.k3 - m1.r0);                     // P = U2 - U1 < 5
    const uint32_t r = carry1(m1.r3 + k.k3 - m1.r2);                      // R = S2 - S1 < 5
    if (f29_is_zero_below12p(to_f29(pp))) return f29_is_zero_below12p(to_f29(r)) ? 2 : 1;
    const Rows m2 = rows(mul(pick(at, pp, r, p.zz, p.zzz), pick(at, pp, r, q.zz, q.zzz), at));
    const uint32_t pp2 = m2.r0;
    const Rows m3 = rows(mul(pick(at, pp, m1.r0, m2.r2, pp), pp2, at));
    const uint32_t x3 = carry1(m2.r1 + k.k7 - (m3.r0 + m3.r1 + m3.r1));  // R^2 - PPP - 2Q < 9
    const uint32_t d = carry1(m3.r1 + k.k10 - x3);                        // Q - X3 < 12
    const Rows m4 = rows(mul(pick(at, r, m1.r2, m2.r3, r), pick(at, d, m3.r0, m3.r0, d), at));
    *out = Point{x3, carry1(m4.r0 + k.k3 - m4.r1), m3.r2, m4.r2};          // Y3 = R D - S1 PPP < 5
    return 0;
}

// ---- G2: Fq2 = Fq[u]/(u^2 + 1), each component one limb per lane
// An Fq2 product puts its four Fq products on the four rows (a0 b0, a1 b1, a0 b1, a1 b0) and
// combines them on the 64-bit columns before one reduction per component: rows 0, 1 take
// a0 b0 + p 2^261 - a1 b1 and rows 2, 3 a0 b1 + a1 b0 (the rows' columns exchanged by a row
// swap). p 2^261 is added in borrowed columns (every column 0..15 at least 2^62 - 2^33, column 16
// p's bits from 203), so that no column goes negative for a1 b1 < p 2^261 (about 169 p^2) with
// limbs below 2^29 + 2^8: the real part is below a0 b0 / 2^261 + 2p and the imaginary part below
// (a0 b1 + a1 b0) / 2^261 + p, which keeps the chain's bounds closed where separately reduced
// products (each carrying its own + p) did not.
struct Fq2L {
    uint32_t a, b;
};
struct Point2 {  // XYZZ over Fq2, every component replicated over the four rows
    Fq2L x, y, zz, zzz;
};

// lane rl's column of p 2^261 in borrowed 64-bit columns (columns 0..15; column 16 is M16)
ZKG_DEV uint64_t mcol(uint32_t rl) {
    uint32_t pl = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i) pl = rl == (uint32_t)(9 + i) ? f29::P(i) : pl;
    return (1ull << 62) - (rl ? (1ull << 33) : 0ull) + pl;
}
constexpr uint64_t M16 = (uint64_t)f29::P(7) + ((uint64_t)f29::P(8) << 29) - (1ull << 33);

ZKG_DEV uint64_t swap16_64(uint64_t v, bool second) {
    const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(v >> 32), (uint32_t)(v >> 32), false, false);
    return second ? ((uint64_t)hi[1] << 32) | lo[1] : ((uint64_t)hi[0] << 32) | lo[0];
}

// x y for inputs with limbs below 2^29 + 2^8 and x.b y.b < 169 p^2
ZKG_DEV Fq2L mul2(const Fq2L& x, const Fq2L& y, const Where& at, uint64_t mc) {
    const uint32_t a = pick(at, x.a, x.b, x.a, x.b), b = pick(at, y.a, y.b, y.b, y.a);
    const uint64_t t = columns(0, a, BcastOf{b});
    const uint64_t h = (uint64_t)bc<8>(a) * bc<8>(b);
    const uint64_t t0 = swap16_64(t, false), t1 = swap16_64(t, true);  // rows (0 0 2 2), (1 1 3 3)
    const uint64_t h0 = swap16_64(h, false), h1 = swap16_64(h, true);
    const uint64_t u = at.row_b1 ? t0 + t1 : t0 + mc - t1;
    const uint64_t hu = at.row_b1 ? h0 + h1 : h0 + M16 - h1;
    const uint32_t r = reduce(u, hu, at);                              // rows 0, 1: real; 2, 3: imaginary
    const auto e = __builtin_amdgcn_permlane32_swap(r, r, false, false);
    return Fq2L{e[0], e[1]};
}

ZKG_DEV Fq2L add2(const Fq2L& x, const Fq2L& y) { return Fq2L{x.a + y.a, x.b + y.b}; }
ZKG_DEV Fq2L norm2(const Fq2L& x) { return Fq2L{carry1(x.a), carry1(x.b)}; }
ZKG_DEV Fq2L sub2(const Fq2L& x, const Fq2L& y, uint32_t k) {  // x + C p - y, k = C p's limb
    return Fq2L{carry1(x.a + k - y.a), carry1(x.b + k - y.b)};
}
ZKG_DEV bool is_zero2(const Fq2L& x) { return f29_is_zero_below12p(to_f29(x.a)) && f29_is_zero_below12p(to_f29(x.b)); }

struct SubConsts2 {  // C p in borrowed limbs for the G2 chain's subtractions, and p 2^261's column
    uint32_t k4, k6, k8, k10, k12;
    uint64_t mc;
    ZKG_DEV explicit SubConsts2(uint32_t rl)
        : k4(kp<4>(rl)), k6(kp<6>(rl)), k8(kp<8>(rl)), k10(kp<10>(rl)), k12(kp<12>(rl)), mc(mcol(rl)) {}
};

// 2P (dbl-2008-s-1) over Fq2, W Y and ZZZ3 taken as V (U Y) and V (U ZZZ) as in dbl. Bounds in
// units of p per component (real, imaginary), a box closed under this doubling and add_g2: running
// point X < (10.3, 9.4), Y < (6.6, 6), ZZ, ZZZ < (2.1, 2); window sums below 2.
ZKG_DEV Point2 dbl_g2(const Point2& p, const Where& at, const SubConsts2& k) {
    const Fq2L u = norm2(add2(p.y, p.y));                                   // < (13.2, 12)
    const Fq2L v = mul2(u, u, at, k.mc), x2 = mul2(p.x, p.x, at, k.mc);
    const Fq2L uy = mul2(u, p.y, at, k.mc), uz = mul2(u, p.zzz, at, k.mc);
    const Fq2L m = norm2(add2(add2(x2, x2), x2));                           // M = 3 X^2 < (7.9, 6.4)
    const Fq2L s = mul2(p.x, v, at, k.mc), zz3 = mul2(v, p.zz, at, k.mc);
    const Fq2L m2 = mul2(m, m, at, k.mc), wy = mul2(v, uy, at, k.mc);
    const Fq2L x3 = sub2(m2, add2(s, s), k.k6);                            // M^2 + 6p - 2S
    const Fq2L d = sub2(s, x3, k.k10);                                      // S + 10p - X3
    const Fq2L md = mul2(m, d, at, k.mc), zzz3 = mul2(v, uz, at, k.mc);
    return Point2{x3, sub2(md, wy, k.k4), zz3, zzz3};                       // Y3 = M D + 4p - W Y
}

// P + Q (add-2008-s with its special cases) over Fq2, as add; the same return codes.
ZKG_DEV int add_g2(const Point2& p, const Point2& q, const Where& at, const SubConsts2& k, Point2* out) {
    const Fq2L u1 = mul2(p.x, q.zz, at, k.mc), u2 = mul2(q.x, p.zz, at, k.mc);
    const Fq2L s1 = mul2(p.y, q.zzz, at, k.mc), s2 = mul2(q.y, p.zzz, at, k.mc);
    const Fq2L pp = sub2(u2, u1, k.k4), r = sub2(s2, s1, k.k4);             // P, R < (6.1, 5.1)
    if (is_zero2(pp)) return is_zero2(r) ? 2 : 1;
    const Fq2L pp2 = mul2(pp, pp, at, k.mc), r2 = mul2(r, r, at, k.mc);
    const Fq2L zz12 = mul2(p.zz, q.zz, at, k.mc), zzz12 = mul2(p.zzz, q.zzz, at, k.mc);
    const Fq2L ppp = mul2(pp, pp2, at, k.mc), qq = mul2(u1, pp2, at, k.mc);
    const Fq2L zz3 = mul2(zz12, pp2, at, k.mc);
    const Fq2L x3 = sub2(r2, add2(ppp, add2(qq, qq)), k.k8);               // R^2 + 8p - PPP - 2Q
    const Fq2L d = sub2(qq, x3, k.k12);                                     // Q + 12p - X3
    const Fq2L rd = mul2(r, d, at, k.mc), s1ppp = mul2(s1, ppp, at, k.mc);
    const Fq2L zzz3 = mul2(zzz12, ppp, at, k.mc);
    *out = Point2{x3, sub2(rd, s1ppp, k.k4), zz3, zzz3};                    // Y3 = R D + 4p - S1 PPP
    return 0;
}

}  // namespace lpl

template <class Fe>  // Fq29 (a template, so that every translation unit may instantiate it)
__global__ __launch_bounds__(64) void msm_combine29_kernel(const uint32_t* __restrict__ wsums, uint32_t W,
                                                          uint32_t wb, uint32_t nw, uint32_t* __restrict__ result) {
    constexpr int XW = 4 * FeTraits<Fq29>::WORDS;
    const uint32_t lane = threadIdx.x;
    const lpl::Where at(lane);
    const lpl::SubConsts kp(at.rl);
    // window w's sum (Fq29 storage, below 2p) with one limb per lane; every lane reads the whole
    // sum through a lane-dependent address (vector registers) and keeps its limb
    const uint32_t zero = lane >> 6;
    auto limbs = [&](const Fq29& v) {
        F29 f;
#pragma unroll
        for (int i = 0; i < 9; ++i) f.l[i] = v.l[i];
        return lpl::from_f29(f, at.rl);
    };
    auto load = [&](uint32_t w, bool* is_inf) {
        const Xyzz<Fq29> v = load_xyzz<Fq29>(wsums + (size_t)XW * (w + zero));
        *is_inf = fq29_is_zero(v.zz);
        return lpl::Point{limbs(v.x), limbs(v.y), limbs(v.zz), limbs(v.zzz)};
    };
    bool inf;
    lpl::Point acc = load(W - 1, &inf);
    for (int w = (int)W - 2; w >= 0; --w) {
        const uint32_t width = wb + ((uint32_t)w < nw ? 1u : 0u);  // bits between window w and w + 1
        if (!inf)  // no 2-torsion: a finite point stays finite
            for (uint32_t k = 0; k < width; ++k) acc = lpl::dbl(acc, at, kp);
        bool q_inf;
        const lpl::Point q = load(w, &q_inf);
        if (q_inf) continue;
        if (inf) {
            acc = q;
            inf = false;
            continue;
        }
        lpl::Point sum;
        const int how = lpl::add(acc, q, at, kp, &sum);
        if (how == 0) acc = sum;
        else if (how == 1) inf = true;
        else acc = lpl::dbl(acc, at, kp);
    }
    Xyzz<Fq29> out;
    if (!inf) {
        const F29 x = lpl::to_f29(acc.x), y = lpl::to_f29(acc.y), zz = lpl::to_f29(acc.zz), zzz = lpl::to_f29(acc.zzz);
#pragma unroll
        for (int i = 0; i < 9; ++i) {  // values below 12 p: store_result reduces them to canonical words
            out.x.l[i] = x.l[i];
            out.y.l[i] = y.l[i];
            out.zz.l[i] = zz.l[i];
            out.zzz.l[i] = zzz.l[i];
        }
    }
    if (lane != 0) return;
    store_xyzz_result(result, inf ? xyzz_inf<Fq29>() : out);
}

// The G2 (Fq2_29 pipeline) Horner chain on the same limb-per-lane products: an Fq2 product per
// instruction stream, the independent Fq2 products of a doubling or addition round interleaved by
// the compiler. Window sums (Fq2_29 storage) are below 2p per component; the result, below 12p per
// component, is stored canonical.
template <class Fe>  // Fq2_29
__global__ __launch_bounds__(64) void msm_combine_g2_lanes_kernel(const uint32_t* __restrict__ wsums, uint32_t W,
                                                                 uint32_t wb, uint32_t nw,
                                                                 uint32_t* __restrict__ result) {
    constexpr int XW = 4 * FeTraits<Fq2_29>::WORDS;
    const uint32_t lane = threadIdx.x;
    const lpl::Where at(lane);
    const lpl::SubConsts2 kp(at.rl);
    const uint32_t zero = lane >> 6;  // every lane reads the whole sum (vector registers) and keeps its limb
    auto limbs = [&](const Fq29& v) {
        F29 f;
#pragma unroll
        for (int i = 0; i < 9; ++i) f.l[i] = v.l[i];
        return lpl::from_f29(f, at.rl);
    };
    auto load = [&](uint32_t w, bool* is_inf) {
        const Xyzz<Fq2_29> v = load_xyzz<Fq2_29>(wsums + (size_t)XW * (w + zero));
        *is_inf = FeTraits<Fq2_29>::is_zero(v.zz);
        return lpl::Point2{{limbs(v.x.a), limbs(v.x.b)}, {limbs(v.y.a), limbs(v.y.b)},
                           {limbs(v.zz.a), limbs(v.zz.b)}, {limbs(v.zzz.a), limbs(v.zzz.b)}};
    };
    bool inf;
    lpl::Point2 acc = load(W - 1, &inf);
    for (int w = (int)W - 2; w >= 0; --w) {
        const uint32_t width = wb + ((uint32_t)w < nw ? 1u : 0u);  // bits between window w and w + 1
        if (!inf)
            for (uint32_t k = 0; k < width; ++k) acc = lpl::dbl_g2(acc, at, kp);
        bool q_inf;
        const lpl::Point2 q = load(w, &q_inf);
        if (q_inf) continue;
        if (inf) {
            acc = q;
            inf = false;
            continue;
        }
        lpl::Point2 sum;
        const int how = lpl::add_g2(acc, q, at, kp, &sum);
        if (how == 0) acc = sum;
        else if (how == 1) inf = true;
        else acc = lpl::dbl_g2(acc, at, kp);
    }
    auto fq29 = [](uint32_t v) {
        const F29 f = lpl::to_f29(v);
        Fq29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.l[i] = f.l[i];
        return r;
    };
    Xyzz<Fq2_29> out;
    if (!inf)
        out = Xyzz<Fq2_29>{{fq29(acc.x.a), fq29(acc.x.b)}, {fq29(acc.y.a), fq29(acc.y.b)},
                           {fq29(acc.zz.a), fq29(acc.zz.b)}, {fq29(acc.zzz.a), fq29(acc.zzz.b)}};
    if (lane != 0) return;
    store_xyzz_result(result, inf ? xyzz_inf<Fq2_29>() : out);  // below 12p: reduced to canonical words
}

// G1 window sums with one limb per lane: a workgroup of NW waves per window; wave i adds its
// window's partials i, i + NW, ... (each an lpl::add, both operands in the running bounds X < 9p,
// Y < 5p, ZZ, ZZZ < 2p), then the waves fold their sums through LDS in log2 NW rounds. The sum is
// multiplied by Montgomery's one (the four coordinates in one round of products) to bring it below
// 2p, the form the window sums are stored in.
template <int NW>
__global__ __launch_bounds__(64 * NW) void msm_window_sum29_lanes_kernel(const uint32_t* __restrict__ partials,
                                                                        uint32_t P, uint32_t nw, uint32_t gh,
                                                                        uint32_t gl, uint32_t* __restrict__ wsums) {
    constexpr int XW = 4 * FeTraits<Fq29>::WORDS;
    __shared__ uint32_t pts[NW / 2][4][64];  // a wave's point, one limb per lane
    __shared__ uint32_t pinf[NW / 2];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const lpl::Where at(lane);
    const lpl::SubConsts kp(at.rl);
    const uint32_t w = blockIdx.x, pw = w < nw ? gh : gl;
    const uint32_t* src = partials + (size_t)XW * w * P;
    const uint32_t zero = threadIdx.x >> 16;  // a lane-dependent address: vector loads of the whole point
    auto limbs = [&](const Fq29& v) {
        F29 f;
#pragma unroll
        for (int i = 0; i < 9; ++i) f.l[i] = v.l[i];
        return lpl::from_f29(f, at.rl);
    };
    bool inf = true;
    lpl::Point acc{0, 0, 0, 0};
    auto fold = [&](const lpl::Point& q, bool q_inf) {
        if (q_inf) return;
        if (inf) {
            acc = q;
            inf = false;
            return;
        }
        lpl::Point sum;
        const int how = lpl::add(acc, q, at, kp, &sum);
        if (how == 0) acc = sum;
        else if (how == 1) inf = true;
        else acc = lpl::dbl(acc, at, kp);
    };
    for (uint32_t g = wave; g < pw; g += NW) {
        const Xyzz<Fq29> v = load_xyzz<Fq29>(src + (size_t)XW * (g + zero));
        fold(lpl::Point{limbs(v.x), limbs(v.y), limbs(v.zz), limbs(v.zzz)}, fq29_is_zero(v.zz));
    }
    for (uint32_t off = NW / 2; off > 0; off >>= 1) {
        if (wave >= off && wave < 2 * off) {
            uint32_t(*d)[64] = pts[wave - off];
            d[0][lane] = acc.x;
            d[1][lane] = acc.y;
            d[2][lane] = acc.zz;
            d[3][lane] = acc.zzz;
            if (lane == 0) pinf[wave - off] = inf ? 1u : 0u;
        }
        __syncthreads();
        if (wave < off) {
            const uint32_t(*d)[64] = pts[wave];
            fold(lpl::Point{d[0][lane], d[1][lane], d[2][lane], d[3][lane]}, pinf[wave] != 0);
        }
        __syncthreads();
    }
    if (wave != 0) return;
    Xyzz<Fq29> out = xyzz_inf<Fq29>();
    if (!inf) {
        // rows x, y, zz, zzz times Montgomery's one: each coordinate below (9 / 169 + 1) p < 2p
        const uint32_t r = lpl::mul(lpl::pick(at, acc.x, acc.y, acc.zz, acc.zzz), lpl::from_f29(f29_const_one(), at.rl), at);
        auto fq29 = [&](int row) {
            const F29 f = lpl::to_f29(r, row);
            Fq29 o;
#pragma unroll
            for (int i = 0; i < 9; ++i) o.l[i] = f.l[i];
            return o;
        };
        out = Xyzz<Fq29>{fq29(0), fq29(1), fq29(2), fq29(3)};
    }
    if (lane == 0) store_xyzz(wsums + (size_t)XW * w, out);
}

// G2 window sums (Fq2_29 pipeline) the same way, with lpl::add_g2 (sums added to sums stay in the
// G2 chain's bounds); the sum's eight components are multiplied by Montgomery's one in two rounds
// of products, which brings each below 2p for storage.
template <int NW>
__global__ __launch_bounds__(64 * NW) void msm_window_sum_g2_lanes_kernel(const uint32_t* __restrict__ partials,
                                                                         uint32_t P, uint32_t nw, uint32_t gh,
                                                                         uint32_t gl, uint32_t* __restrict__ wsums) {
    constexpr int XW = 4 * FeTraits<Fq2_29>::WORDS;
    __shared__ uint32_t pts[NW / 2][8][64];  // a wave's point, one limb per lane
    __shared__ uint32_t pinf[NW / 2];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const lpl::Where at(lane);
    const lpl::SubConsts2 kp(at.rl);
    const uint32_t w = blockIdx.x, pw = w < nw ? gh : gl;
    const uint32_t* src = partials + (size_t)XW * w * P;
    const uint32_t zero = threadIdx.x >> 16;  // a lane-dependent address: vector loads of the whole point
    auto limbs = [&](const Fq29& v) {
        F29 f;
#pragma unroll
        for (int i = 0; i < 9; ++i) f.l[i] = v.l[i];
        return lpl::from_f29(f, at.rl);
    };
    bool inf = true;
    lpl::Point2 acc{};
    auto fold = [&](const lpl::Point2& q, bool q_inf) {
        if (q_inf) return;
        if (inf) {
            acc = q;
            inf = false;
            return;
        }
        lpl::Point2 sum;
        const int how = lpl::add_g2(acc, q, at, kp, &sum);
        if (how == 0) acc = sum;
        else if (how == 1) inf = true;
        else acc = lpl::dbl_g2(acc, at, kp);
    };
    for (uint32_t g = wave; g < pw; g += NW) {
        const Xyzz<Fq2_29> v = load_xyzz<Fq2_29>(src + (size_t)XW * (g + zero));
        fold(lpl::Point2{{limbs(v.x.a), limbs(v.x.b)}, {limbs(v.y.a), limbs(v.y.b)},
                         {limbs(v.zz.a), limbs(v.zz.b)}, {limbs(v.zzz.a), limbs(v.zzz.b)}},
             FeTraits<Fq2_29>::is_zero(v.zz));
    }
    for (uint32_t off = NW / 2; off > 0; off >>= 1) {
        if (wave >= off && wave < 2 * off) {
            uint32_t(*d)[64] = pts[wave - off];
            d[0][lane] = acc.x.a;
            d[1][lane] = acc.x.b;
            d[2][lane] = acc.y.a;
            d[3][lane] = acc.y.b;
            d[4][lane] = acc.zz.a;
            d[5][lane] = acc.zz.b;
            d[6][lane] = acc.zzz.a;
            d[7][lane] = acc.zzz.b;
            if (lane == 0) pinf[wave - off] = inf ? 1u : 0u;
        }
        __syncthreads();
        if (wave < off) {
            const uint32_t(*d)[64] = pts[wave];
            fold(lpl::Point2{{d[0][lane], d[1][lane]}, {d[2][lane], d[3][lane]}, {d[4][lane], d[5][lane]},
                             {d[6][lane], d[7][lane]}},
                 pinf[wave] != 0);
        }
        __syncthreads();
    }
    if (wave != 0) return;
    Xyzz<Fq2_29> out = xyzz_inf<Fq2_29>();
    if (!inf) {
        // components times Montgomery's one, four per round: each below (10.3 / 169 + 1) p < 2p
        const uint32_t one = lpl::from_f29(f29_const_one(), at.rl);
        const uint32_t r0 = lpl::mul(lpl::pick(at, acc.x.a, acc.x.b, acc.y.a, acc.y.b), one, at);
        const uint32_t r1 = lpl::mul(lpl::pick(at, acc.zz.a, acc.zz.b, acc.zzz.a, acc.zzz.b), one, at);
        auto fq29 = [&](uint32_t r, int row) {
            const F29 f = lpl::to_f29(r, row);
            Fq29 o;
#pragma unroll
            for (int i = 0; i < 9; ++i) o.l[i] = f.l[i];
            return o;
        };
        out = Xyzz<Fq2_29>{{fq29(r0, 0), fq29(r0, 1)}, {fq29(r0, 2), fq29(r0, 3)},
                           {fq29(r1, 0), fq29(r1, 1)}, {fq29(r1, 2), fq29(r1, 3)}};
    }
    if (lane == 0) store_xyzz(wsums + (size_t)XW * w, out);
}

// ---------------------------------------------------------------- stage launchers
// Each group's stages are instantiated in their own translation units (msm_g1.hip,
// msm_g2_*.hip) so that the large, fully inlined G2 kernels compile in parallel.

template <class Fe>
void launch_accumulate(const MsmPlan& p, const uint32_t* bases, hipStream_t s) {
    const size_t lanes = ((size_t)p.W * p.nv + p.CH - 1) / p.CH;  // lane-cut: one lane per CH sorted entries
    if (p.lane_cut() && !std::is_same<Fe, Fq29>::value) {
        if constexpr (std::is_same<Fe, Fq2_29>::value)
            hipLaunchKernelGGL((msm_accumulate_lanes_kernel<Fq2, Fq2_29>), dim3(ceil_div(lanes, 256)), dim3(256), 0, s,
                               bases, (const uint32_t*)p.endo_x, p.n_real(), p.sorted, p.starts, p.istarts,
                               p.W * p.B, p.CH, p.item_info, p.item_pts);
        else if constexpr (!std::is_same<Fe, Fq29>::value)
            hipLaunchKernelGGL(msm_accumulate_lanes_kernel<Fe>, dim3(ceil_div(lanes, 256)), dim3(256), 0, s, bases,
                               (const uint32_t*)p.endo_x, p.n_real(), p.sorted, p.starts, p.istarts, p.W * p.B,
                               p.CH, p.item_info, p.item_pts);
        return;
    }
    if constexpr (std::is_same<Fe, Fq2_29>::value) {
        hipLaunchKernelGGL((msm_accumulate_kernel<Fq2, Fq2_29>), dim3(ceil_div(p.max_items, 256)), dim3(256), 0, s,
                           bases, (const uint32_t*)p.endo_x, p.n_real(), p.sorted, p.starts, p.istarts,
                           p.counts, p.W * p.B, p.CH, p.item_info, p.item_pts);
    } else if constexpr (std::is_same<Fe, Fq29>::value) {
        if (!p.lane_cut()) {
            hipLaunchKernelGGL(msm_accumulate29_bucket_kernel<Fq29>, dim3(ceil_div(p.max_items, 256)), dim3(256), 0, s,
                               bases, (const uint32_t*)p.endo_x, p.n_real(), p.sorted, p.starts, p.istarts, p.counts,
                               p.W * p.B, p.CH, p.item_info, p.item_pts);
            return;
        }
        hipLaunchKernelGGL(msm_accumulate29_kernel<Fq29>, dim3(ceil_div(lanes, 256)), dim3(256), 0, s, bases,
                           (const uint32_t*)p.endo_x, p.n_real(), p.sorted, p.starts, p.istarts, p.W * p.B, p.CH,
                           p.item_info, p.item_pts);
    } else {
        hipLaunchKernelGGL(msm_accumulate_kernel<Fe>, dim3(ceil_div(p.max_items, 256)), dim3(256), 0, s, bases,
                           (const uint32_t*)p.endo_x, p.n_real(), p.sorted, p.starts, p.istarts,
                           p.counts, p.W * p.B, p.CH, p.item_info, p.item_pts);
    }
}

// The field of the merge, finalize, reduce and tree-sum kernels: G2's Fq2_29 stages run on lane
// pairs (Fq2P, the same memory form), which halves their chains' latency per product and their
// registers per lane; the accumulation and the combine keep their own fields.
template <class Fe>
struct TailField {
    using type = Fe;
};
template <>
struct TailField<Fq2_29> {
    using type = Fq2P;
};

// The field of the bucket reduction and window sums: G1 (Fq29) runs them on lane pairs (Fq29Pair:
// each point operation's products split between the two lanes), G2 on groups of four lanes (Fq2Q:
// split between two lane pairs that each hold one Fq2 element).
// Their serial chains of additions otherwise leave most SIMDs idle: the G1 2^20 reduction had one
// wave per two SIMDs. Measured on MI355X at G1 2^20, c = 15 (kernel trace): 0.296 ms on one lane
// per chain, 0.279 ms on pairs over the full bucket range, 0.20 ms with the grid below.
template <class Fe>
struct ReduceField {
    using type = typename TailField<Fe>::type;
};
template <>
struct ReduceField<Fq29> {
    using type = Fq29Pair;
};
template <>
struct ReduceField<Fq2_29> {
    using type = Fq2Q;
};

template <class Fe>
void launch_merge(const MsmPlan& p, uint32_t stride, hipStream_t s) {
    // (the reduction's lane pairs and quads measured no faster here: G2 2^18 merge 1 0.070 ms on
    // pairs, 0.072 on quads)
    using Ft = typename TailField<Fe>::type;
    constexpr uint32_t LN = FeLanes<Ft>::value;
    const size_t blocks = std::min<size_t>(ceil_div(p.max_items * LN, 256), MERGE_BLOCKS);
    hipLaunchKernelGGL(msm_merge_kernel<Ft>, dim3(blocks), dim3(256), 0, s, p.item_pts,
                       p.item_info, p.istarts, p.W * p.B, stride, (const uint32_t*)p.max_items_dev);
}

// The field of the finalize: the tail field, except G2 on four lanes per point (Fq2Q; plain G2
// 2^18: 0.215 -> 0.18 ms). G1 on lane pairs measured slower (2^20: 0.12 -> 0.15 ms).
template <class Fe>
struct FinalizeField {
    using type = typename TailField<Fe>::type;
};
template <>
struct FinalizeField<Fq2_29> {
    using type = Fq2Q;
};

template <class Fe>
void launch_finalize(const MsmPlan& p, hipStream_t s) {
    const uint32_t S = p.W * p.B;
    // Several lanes per bucket while the buckets alone leave the GPU short of waves. Measured on
    // MI355X: G1 2^20 (147456 buckets, two lanes) 0.20 -> 0.16 ms, G2 2^20 0.55 -> 0.47 ms; G2 2^16
    // (12288 buckets) 0.39 -> 0.27 ms with two lanes, 0.14 ms with four; at 2^18 buckets and more
    // one lane each is faster (G1 2^22 0.18 against 0.20, G2 0.62 against 0.80 with two).
    // G2 runs on four lanes per point (Fq2Q): its thresholds are halved (precomputed 2^18 at c = 16,
    // 131072 buckets: 0.44 ms on pairs, 0.47 with two parts of four lanes per bucket)
    using Ft = typename FinalizeField<Fe>::type;
    constexpr uint32_t LN = FeLanes<Ft>::value;
    constexpr uint32_t hs = LN >= 4 ? 1u : 0u;
    const uint32_t split = S < (1u << (15 - hs)) ? 2u : S < (1u << (18 - hs)) ? 1u : 0u;
    hipLaunchKernelGGL(msm_finalize_kernel<Ft>, dim3(ceil_div(((size_t)S << split) * LN, 256)), dim3(256), 0, s,
                       p.item_pts, p.istarts, (const uint2*)p.item_info, S, split, p.buckets);
}

// Logical threads (one per element) per workgroup of the bucket reduction: 256 lanes.
template <class Ft>
constexpr uint32_t reduce_block() {
    return 256 / FeLanes<Ft>::value;
}

// Workgroups of the bucket reduction: rb logical threads each (a workgroup never spans two
// windows); windows below nw (2^wb buckets) take gh of them, the rest (2^(wb-1) buckets) gl, and
// each window's partials are P = T / rb apart.
template <class Ft>
struct ReduceGrid {
    uint32_t rb, gh, gl, P;
    explicit ReduceGrid(const MsmPlan& p) {
        rb = std::min<uint32_t>(reduce_block<Ft>(), p.T);
        P = p.T / rb;
        const uint32_t bh = std::min<uint32_t>(p.B, 1u << p.wb), bl = std::min<uint32_t>(p.B, 1u << (p.wb - 1));
        gh = ceil_div(ceil_div(bh, p.L), rb);
        gl = ceil_div(ceil_div(bl, p.L), rb);
    }
};

template <class Fe>
void launch_reduce(const MsmPlan& p, hipStream_t s) {
    using Ft = typename ReduceField<Fe>::type;
    constexpr uint32_t LN = FeLanes<Ft>::value;
    const ReduceGrid<Ft> g(p);
    const size_t wgs = (size_t)p.nw * g.gh + (size_t)(p.W - p.nw) * g.gl;
    hipLaunchKernelGGL(msm_reduce_kernel<Ft>, dim3(wgs), dim3(g.rb * LN), 0, s, p.buckets, p.W, p.B, p.L, p.cw,
                       p.wb, p.nw, g.gh, g.gl, g.P, p.partials);
}

template <class Fe>
void launch_tree(const MsmPlan& p, hipStream_t s) {
    using Ft = typename ReduceField<Fe>::type;
    constexpr uint32_t LN = FeLanes<Ft>::value;
    const ReduceGrid<Ft> g(p);
    if constexpr (std::is_same<Fe, Fq29>::value) {
        hipLaunchKernelGGL(msm_window_sum29_lanes_kernel<16>, dim3(p.W), dim3(64 * 16), 0, s, p.partials, g.P, p.nw,
                           g.gh, g.gl, p.wsums);
    } else if constexpr (std::is_same<Fe, Fq2_29>::value) {
        hipLaunchKernelGGL(msm_window_sum_g2_lanes_kernel<16>, dim3(p.W), dim3(64 * 16), 0, s, p.partials, g.P,
                           p.nw, g.gh, g.gl, p.wsums);
    } else {
        uint32_t top = 1;
        while (top < std::min<uint32_t>(std::max(g.gh, g.gl), 64 / LN)) top <<= 1;
        hipLaunchKernelGGL(msm_window_sum_kernel<Ft>, dim3(p.W), dim3(64), 0, s, p.partials, g.P, p.nw, g.gh, g.gl,
                           top / 2, p.wsums);
    }
}

template <class Fe>
void launch_combine(const MsmPlan& p, hipStream_t s) {
    if constexpr (std::is_same<Fe, Fq29>::value)
        hipLaunchKernelGGL(msm_combine29_kernel<Fe>, dim3(1), dim3(64), 0, s, p.wsums, p.W, p.wb, p.nw, p.result);
    else if constexpr (std::is_same<Fe, Fq2_29>::value)
        hipLaunchKernelGGL(msm_combine_g2_lanes_kernel<Fe>, dim3(1), dim3(64), 0, s, p.wsums, p.W, p.wb, p.nw,
                           p.result);
    else
        hipLaunchKernelGGL(msm_combine_kernel<Fe>, dim3(1), dim3(64), 0, s, p.wsums, p.W, p.wb, p.nw, p.result);
}

}  // namespace zkg
