// In-place, natural-order NTT over the BN254 scalar field Fr.
//
// Drop-in for the reference's Ntt (src/ntt.rs:16-68): Ntt::new(log_n) with 10 <= log_n <= 27,
// forward = X_j = sum_i x_i w^(ij) with w = 5^((r-1) >> log_n), inverse uses w^-1 and folds in
// the 1/n scaling; data is 8 little-endian u32 words per element, plain (not Montgomery) form.
#pragma once
#include <map>
#include <memory>
#include <vector>

#include "runtime.hpp"

namespace zkg {

// Parameters of one NTT pass over one transform (kernel argument, or an entry of a batch table).
struct NttPassArgs {
    const uint32_t* src;
    uint32_t* dst;
    const uint32_t* tw;
    uint32_t lm, s, lg, log_n;
    uint32_t last, scale_flag;
    uint32_t scale[8];
    const uint32_t* scale_tab;  // last pass: multiply output element i by scale_tab[i] (Montgomery), else scale
};

// One pass = `s` radix-2 decimation-in-frequency stages on LDS-resident tiles.
struct NttPass {
    uint32_t lm;     // log2 of the element stride inside a group
    uint32_t s;      // stages in this pass
    uint32_t lg;     // log2 of groups (columns) per workgroup
    uint32_t last;   // last pass: bit-reversed, coalesced scatter (+ optional 1/n scale)
    uint32_t wgs;    // workgroups launched
    int in_place;    // reads and writes the same buffer
};

// Launch-shape choice for one transform size; the autotuner may override the default.
struct NttShape {
    int tile_log = 0;  // log2 of elements per workgroup tile (0 = library default)
};

struct NttPlan {
    Device* dev = nullptr;
    uint32_t log_n = 0;
    int tile_log = 11;
    uint32_t* tw_fwd = nullptr;   // w^i * 2^256 mod r, i < n/2
    uint32_t* tw_inv = nullptr;   // w^-i * 2^256 mod r
    uint32_t* scratch = nullptr;  // n elements, for the out-of-place last pass
    uint32_t scale[8];            // n^-1 * 2^256 mod r (so a Montgomery product multiplies by n^-1)
    std::vector<NttPass> passes;

    // with_scratch = false: twiddles only (a batch supplies per-job scratch)
    NttPlan(Device* d, uint32_t log_n, NttShape shape = NttShape(), bool with_scratch = true);
    ~NttPlan();
    // Queue a transform of `data` (device pointer holding at least 2^log_n elements) on `stream`.
    // scale_tab (optional, 2^log_n Montgomery elements) replaces the inverse's n^-1 by a
    // per-element factor applied in the last pass's store (either direction).
    void run(uint32_t* data, bool inverse, hipStream_t stream, const uint32_t* scale_tab = nullptr) const;
    // The same with the caller's n-element scratch (plans built without one share their twiddles).
    void run_with(uint32_t* data, bool inverse, hipStream_t stream, const uint32_t* scale_tab, uint32_t* scr) const;
    // Pass i of a transform of `data` using `scr` as the out-of-place scratch.
    NttPassArgs pass_args(size_t i, uint32_t* data, uint32_t* scr, bool inverse,
                          const uint32_t* scale_tab = nullptr) const;
    size_t workspace_bytes() const;
};

// Many NTTs of mixed sizes and directions in one grouped launch per pass phase: phase k runs
// pass k of every job that has more than k passes, each workgroup finding its job in a prefix
// table. Twiddle tables are shared per size; every job gets its own scratch region.
class NttBatch {
  public:
    explicit NttBatch(Device* d);
    ~NttBatch();
    size_t add(uint32_t log_n, uint32_t* data, bool inverse);  // returns the job index
    void clear();
    size_t size() const { return jobs_.size(); }
    void run(hipStream_t stream);  // prepare + launch
    // Build the pass tables, size the scratch and queue the table upload on `stream`
    // (synchronises with the previous upload; allocates when the batch grows).
    void prepare(hipStream_t stream);
    // Only the kernel launches of a prepared batch (no allocation, copy or host wait: capturable).
    void launch(hipStream_t stream) const;
    // The plan (twiddles and pass structure) of one size at one tile.
    const NttPlan& plan(uint32_t log_n, int tile);
    // Use an existing twiddle plan for its size (rebuilt if a run needs another tile).
    void share_plan(const std::shared_ptr<NttPlan>& p) { plans_[p->log_n] = p; }
    // The tile a batch whose largest transform is 2^log_n runs with.
    static int tile_for(const Device* d, uint32_t log_n);

  private:
    struct Job {
        uint32_t log_n;
        uint32_t* data;
        bool inverse;
    };
    Device* dev_;
    std::vector<Job> jobs_;
    std::map<uint32_t, std::shared_ptr<NttPlan>> plans_;
    Scratch scratch_;                 // sum of the jobs' sizes, 32 bytes per element
    Scratch table_;                   // device copy of the pass tables of all phases
    NttPassArgs* staging_ = nullptr;  // pinned host staging of the tables
    size_t staging_cap_ = 0;
    hipEvent_t staged_ = nullptr;     // the last upload from staging has completed
    bool prepared_ = false;
    int tile_ = 0;
    size_t per_phase_ = 0;
    std::vector<uint32_t> phase_jobs_, phase_wgs_;
};

int ntt_default_tile_log(uint32_t log_n);
// out[i] = first * base^i in Montgomery form (base, first: plain Fr words), i < count.
void ntt_geometric_table(uint32_t* out, size_t count, const uint32_t base[8], const uint32_t first[8],
                         hipStream_t stream);

}  // namespace zkg
