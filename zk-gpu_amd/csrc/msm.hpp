// BN254 multi-scalar multiplication (signed-digit Pippenger) for G1 and G2 on gfx950.
//
// Drop-in for the reference's Msm (src/msm.rs:25-342): same validation (n >= 1, 9 <= c <= 20),
// same default window and work-item sizing, same result contract (one XYZZ point in the
// library's Montgomery form, plan-owned, read through xyzz_to_affine).
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "runtime.hpp"

namespace zkg {

uint32_t msm_default_window(size_t n);          // clamp(ilog2(max(n,1)) - 5, 11, 16)
uint32_t msm_auto_item_points(size_t n, uint32_t c);
// The work-item size a plan for (group, n, c) picks when none is given.
uint32_t msm_plan_item_points(int group, size_t n, uint32_t c);
// Everything a plan for (group, n, c, item_points, sort_points, precomputed) derives from its
// arguments (0 = automatic), with the argument checks of Msm::for_group (src/msm.rs:96-98) and the
// index-range checks of this library's kernels: positions into the sorted point list, work-item
// numbers and virtual point indices are 32-bit, so a plan whose W * nv entries or work items
// reach 2^32 is rejected here (an Error with a message), before anything is allocated.
struct MsmShape {
    bool glv = false, pre = false, gls = false, lds_sort = false;
    // cw: digit bits of the widest window (B = 2^(cw - 1)); c, except for GLS plans (65 bits in
    // windows of 9, 10, 11 or 13 bits)
    uint32_t nv = 0, W = 0, wb = 0, nw = 0, cw = 0, B = 0, CH = 0, L = 0, T = 0, aw = 0, xw = 0;
    size_t max_items = 0;
    uint32_t merge_rounds = 0, sort_pts = 0, nblk = 0;
    size_t scan_blocks = 0;  // blocks of the slot scan
    size_t workspace_bytes(size_t n) const;
};
MsmShape msm_shape(int group, size_t n, uint32_t c, uint32_t item_points, uint32_t sort_points = 0,
                   bool precomputed = false);
// Bytes of one region holding every workspace array of a plan of shape `sh` over n points, each
// array starting at a 256-byte boundary (the no-allocation plans carve their arrays from it).
size_t msm_carved_bytes(const MsmShape& sh, size_t n);
// Size of the static device array that zkg_msm_run_noalloc carves its workspace from (one per GPU,
// held until the process ends), and its device address on the calling thread's current device.
static constexpr size_t MSM_SCRATCH_BYTES = (size_t)64 << 20;
void* msm_scratch_address();
// Device bytes a plan for (group, n, c, item_points, sort_points, precomputed) allocates
// (0 = automatic); throws as msm_shape for arguments no plan accepts.
size_t msm_workspace_bytes(int group, size_t n, uint32_t c, uint32_t item_points, uint32_t sort_points = 0,
                           bool precomputed = false);
// Whether plans use the GLV split (default on; ZKGPU_MSM_GLV=0 turns it off for comparisons).
bool msm_glv_enabled(size_t n);
// Whether G2 plans may use the GLS split (MsmPlan::gls; default on, ZKGPU_MSM_GLS=0 turns it off
// for comparisons).
bool msm_gls_enabled();
// Whether plans use the 29-bit-limb field (field29.hpp): the whole G1 pipeline, and the G2 stages
// after the bucket accumulation (default on; ZKGPU_MSM_F29=0 selects the 32-bit-limb kernels,
// for comparisons). Same results either way.
bool msm_f29_enabled();
// Whether a plan sorting `entries` (W nv) entries cuts the G1 accumulation's lanes at multiples of
// its item size (see MsmPlan::lanes).
bool msm_lanes_enabled(uint64_t entries);
// Bucket bits within a partition of the two-pass placement for a plan of this shape, or 0 where it
// does not apply (c > 16, or a packed entry would not hold the point index; ZKGPU_MSM_PARTSORT=0
// turns it off for comparisons).
uint32_t msm_part_bits(const MsmShape& sh, uint32_t* part_log);
// Counting-sort block size (points) a plan over nv virtual points with window c uses by default.
uint32_t msm_default_sort_points(size_t nv, uint32_t c);
// Buckets per lane of the bucket-reduction stage (windows of 2^(c-1) buckets each).
uint32_t msm_reduce_chunk(int group, uint32_t c, uint32_t windows);

struct MsmPlan {
    Device* dev = nullptr;
    int group = 1;           // 1 = G1 (Fq), 2 = G2 (Fq2)
    uint32_t n = 0, c = 0;
    // GLV: s = k1 + k2*lambda (mod r) with |k1|, |k2| < 2^127, and lambda*P = phi(P) = (beta*x, y),
    // so the MSM runs over nv = 2n virtual points (P_i with k1_i, phi(P_i) with k2_i) with
    // 128-bit scalars: half the windows, half the bucket reduction and half the doubling chain.
    bool glv = false;
    // Precomputed bases (`pre`): the bases buffer holds P_0..P_{n-1} and then 2^64 P_0 ..
    // 2^64 P_{n-1} (msm_precompute_bases), so each 128-bit half-scalar splits at bit 64 into two
    // 64-bit ones: nv = 4n virtual points (P, 2^64 P, phi(P), phi(2^64 P)) over ceil(64 / c)
    // windows. The same bucket additions as GLV alone, but half the windows: half the buckets to
    // finalize and reduce and a doubling chain of about 64 - c instead of 128 - c.
    bool pre = false;
    // GLS (G2): psi(x, y) = (conj(x) gx, conj(y) gy) acts on G2 as mu = p mod r, so each scalar splits
    // into four signed coefficients below 2^64 (msm.hip, GLS split) of Q, phi(Q), psi(Q) and
    // phi(psi(Q)): nv = 4n virtual points over 8, 7, 6 or 5 windows of 65 bits (9, 10, 11 or 13
    // digit bits: c up to 11 its own, c from 12 on 13). The run first writes a table of the bases
    // and their psi images (gls_bases) and the phi images of both (endo_x), and the stages then
    // run as over a precomputed table: about as many bucket additions as the GLV split, half the
    // buckets to finalize and reduce, and a doubling chain of about 52 bits instead of 115. Used
    // for G2 at c <= 16 and n <= 2^18, where the gathers from its larger table stay in the caches
    // (msm_shape).
    bool gls = false;
    bool f29 = false;        // 29-bit-limb field (field29.hpp)
    uint32_t nv = 0;         // virtual points: 4n precomputed or GLS, 2n with GLV, else n
    uint32_t W = 0;          // windows: ceil(64 / c) precomputed, 8..5 GLS, ceil(128 / c) GLV, else ceil(255 / c)
    uint32_t cw = 0;         // digit bits of the widest window (c, or less for GLS: MsmShape::cw)
    // Window w covers bits [w*wb + min(w, nw), +wb + (w < nw)): without GLV wb = c, nw = 0; with
    // GLV the 128 bits are split evenly (c = 15: two windows of 15 bits and seven of 14), so no
    // window is left with a few bits and a few heavy buckets.
    uint32_t wb = 0, nw = 0;
    uint32_t B = 0;          // buckets per window, 2^(cw-1)
    uint32_t CH = 0;         // points per work item
    uint32_t L = 0;          // buckets per running-sum thread
    uint32_t T = 0;          // running-sum threads per window, B / L
    uint32_t xw = 0;         // u32 words per XYZZ point (32 or 64)
    uint32_t aw = 0;         // u32 words per affine point (16 or 32)
    size_t max_items = 0;    // upper bound on work items over all windows
    uint32_t merge_rounds = 0;
    bool lds_sort = false;   // c <= 16: LDS counting sort (no global atomics); else atomic counting sort
    uint32_t sort_pts = 0;   // points per sort block
    uint32_t nblk = 0;       // sort blocks per window

    // workspace (device)
    uint32_t* split = nullptr;       // 2n * 4  GLV half-scalars: |k| in 127 bits, sign in bit 127 (not GLS)
    uint32_t* endo_x = nullptr;      // n * aw / 2 (2n * aw / 2 precomputed or GLS)  beta * x of every base (Montgomery form)
    uint32_t* gls_bases = nullptr;   // 2n * aw  GLS: the bases, then psi of each
    uint32_t* counts = nullptr;      // W*B  points per (window, bucket)
    uint32_t* cursors = nullptr;     // W*B  scatter cursors
    uint32_t* starts = nullptr;      // W*B + 1  exclusive scan of counts (global over windows)
    uint32_t* istarts = nullptr;     // W*B + 1  exclusive scan of items per slot
    uint32_t* scan_tmp = nullptr;    // per-block partial sums of the scan, per-block item maxima, the maximum
    uint32_t* max_items_dev = nullptr;  // scan_tmp's last word: the most work items of any bucket
    uint32_t* sorted = nullptr;      // W*nv  point index | sign << 31, grouped by slot
    uint32_t* digits = nullptr;      // W*nv  bucket | sign << 31 per (window, point)  (LDS sort only)
    uint32_t* table = nullptr;       // W*nblk*B  per-block bucket counts -> offsets    (LDS sort only)
    uint2* item_info = nullptr;      // max_items  (position in its bucket, items in its bucket)
    uint32_t* item_pts = nullptr;    // max_items * xw  per-item partial sums
    uint32_t* buckets = nullptr;     // W*B * xw
    uint32_t* partials = nullptr;    // W*T * xw  running-sum results
    uint32_t* wsums = nullptr;       // W * xw    window sums
    uint32_t* result = nullptr;      // xw        final XYZZ point
    void* result_handle = nullptr;   // zkg_buffer* wrapping `result` (owned by the C ABI layer)

    bool owned = true;       // workspace from hipMalloc (freed with the plan) or carved from a caller's region

    // item_points / sort_points 0 = automatic (a quarter of the mean bucket / max(2^(c-1), 8192));
    // precomputed: a plan over msm_precompute_bases tables (c <= 16, n < 2^29)
    MsmPlan(Device* d, int group, size_t n, uint32_t c, uint32_t item_points, uint32_t sort_points = 0,
            bool precomputed = false);
    // A plan that allocates nothing: its workspace is carved from `workspace` (256-byte aligned,
    // at least msm_carved_bytes(sh) bytes, where sh = msm_shape of the same arguments) and its
    // result goes to `result_words` (xw words). Nothing is freed when it is destroyed.
    MsmPlan(Device* d, int group, size_t n, uint32_t c, const MsmShape& sh, void* workspace, uint32_t* result_words);
    // Virtual points below this index are entries of the bases buffer, the others phi of one
    // (2^32 - 1 without GLV).
    uint32_t n_real() const { return pre || gls ? 2 * n : glv ? n : 0xffffffffu; }
    // Points the bases buffer must hold: 2n for a precomputed plan, else n.
    size_t base_points() const { return pre ? 2 * (size_t)n : n; }
    // Whether the bucket accumulation cuts its lanes at multiples of CH of the sorted list (a
    // bucket then takes up to ceil((count - 1) / CH) + 1 slots) rather than at bucket starts
    // (ceil(count / CH) slots).
    bool lane_cut() const { return lanes; }
    // Lanes cut at multiples of CH for the bucket accumulation: by default when the plan sorts at
    // least 2^23 entries (W nv), where it measured faster (G1 2^20 c = 15 accumulate 1.52 -> 1.46
    // ms, 2^22 c = 16 5.98 -> 5.81 ms; G1 2^16 0.21 -> 0.22 ms); ZKGPU_MSM_LANES=1 or 0 forces it
    // on or off.
    bool lanes = false;
    // Two-pass placement (msm.hip): bits of the bucket-within-partition field of a packed entry, or
    // 0 for the one-pass placement; 2^part_log partitions per window.
    uint32_t part_bits = 0, part_log = 0;
    ~MsmPlan();
    // Queue the MSM on `stream`. With `stage_ms` set, wait around every stage and record
    // (stage name, milliseconds) as the reference's Msm::profile does.
    void run(const uint32_t* bases, const uint32_t* scalars, hipStream_t stream,
             std::vector<std::pair<std::string, double>>* stage_ms = nullptr) const;
    // As run, but waits for the GPU after the heavy stages (accumulate, finalize, reduce,
    // combine) so that a display sharing the GPU gets it in between.
    void run_yielding(const uint32_t* bases, const uint32_t* scalars, hipStream_t stream) const;
    size_t workspace_bytes() const;
};

// GLV split of n scalars (any 256-bit values) into split[0..4n) = k1 of each, split[4n..8n) = k2
// of each: |k| in the low 127 bits, its sign in bit 127, s = k1 + k2 lambda (mod r).
void msm_glv_split(const uint32_t* scalars, uint32_t n, uint32_t* split, hipStream_t stream);
// GLS split of n scalars (any 256-bit values): out[12 i + 3 t ..] = |k_t| (2 words), sign of k_t,
// t = 0..3, with s = k0 + k1 lambda + k2 mu + k3 lambda mu (mod r), mu = p mod r, |k_t| < 2^64.
void msm_gls_split(const uint32_t* scalars, uint32_t n, uint32_t* out, hipStream_t stream);

// Plain affine coordinates -> Montgomery form in place (`coords` elements of 8 words).
void msm_convert_bases(uint32_t* xy, size_t coords, hipStream_t stream);
// Checks n bases (library form) of `group`: bad = how many are neither infinity nor on the curve
// and (G2) in the order-r subgroup, first_bad = the smallest such index (0 when none). Waits for
// the device.
void msm_check_bases(Device* d, int group, const uint32_t* bases, size_t n, uint64_t* bad, uint64_t* first_bad);
// Montgomery-2^256 coordinates of any 256-bit value -> canonical (below p) in place.
void msm_reduce_coords(uint32_t* xy, size_t coords, hipStream_t stream);

// The bases table of a precomputed plan: out[0..n) = bases[0..n), out[n..2n) = 2^64 bases[i]
// (library form; `out` holds 2n points of `group`).
void msm_precompute_bases(int group, const uint32_t* bases, size_t n, uint32_t* out, hipStream_t stream);

}  // namespace zkg
