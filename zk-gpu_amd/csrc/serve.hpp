// Device-resident proof-job server with continuous batching.
//
// No counterpart in the reference (one call at a time). Proving keys (four G1 base sets A, B1,
// C, H and one G2 set B2) stay resident in HBM. A job names a key and brings seven polynomials
// of 2^log_n elements (transformed in place, each in its own direction) and a witness (scalars of
// the A, B1 and B2 MSMs; C's start at an offset into it); the H MSM takes the first outputs of the
// seventh transform as its scalars, as in a Groth16 prover. A key's shape gives each MSM's point
// count: n, n, n, n - 1, n by default, or a snarkjs key's nVars, nVars, nVars - nPublic - 1,
// domainSize, nVars with C's scalars from witness element nPublic + 1. A scheduler thread admits jobs in arrival order as soon as a workspace slot and
// the HBM budget allow, so jobs that arrive while others run overlap on the device: every job
// runs on its own streams (NTTs grouped into one launch per pass phase, then H; A, B1, C and B2
// concurrently). Results (five XYZZ points) land in pinned host memory.
#pragma once
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "h_poly.hpp"
#include "msm.hpp"
#include "ntt.hpp"

namespace zkg {

static constexpr int SERVE_POLYS = 7;
static constexpr int SERVE_MSMS = 5;                  // A, B1, C, H (G1) and B2 (G2)
static constexpr int SERVE_RESULT_WORDS = 4 * 32 + 64;  // XYZZ words: four G1 points, then the G2 point

struct Buffer;

struct KeyShape {
    uint32_t log_n = 0;               // polynomial size 2^log_n
    size_t points[SERVE_MSMS] = {};   // A, B1, C, H (G1) and B2 (G2)
    size_t c_offset = 0;              // C's scalars start at this witness element
    HForm h_form = HForm::coefficients;  // what quotient jobs (submit_groth16) feed the H MSM
    // slots (workspaces) depend on the sizes only
    bool operator==(const KeyShape& o) const {
        for (int k = 0; k < SERVE_MSMS; ++k)
            if (points[k] != o.points[k]) return false;
        return log_n == o.log_n && c_offset == o.c_offset;
    }
    // scalars a job's witness must hold
    size_t witness_len() const;
    // n, n, n, n - 1, n points with n = 2^log_n
    static KeyShape groth16(uint32_t log_n);
    // checks the ranges; throws on violation
    void validate() const;
};

struct ServerStats {
    uint64_t submitted = 0, completed = 0;
    double p50_ms = 0, p99_ms = 0, mean_ms = 0, jobs_per_s = 0;
    size_t bytes_in_use = 0, peak_bytes = 0;
    uint32_t inflight = 0, peak_inflight = 0;
};

class Server {
  public:
    // mem_budget: HBM bytes the server may hold (keys + workspaces); max_inflight: jobs on the
    // device at once. MSMs use the tuned window for their size (default_window(n) untuned).
    Server(Device* d, size_t mem_budget, uint32_t max_inflight);
    ~Server();

    // bases: device buffers in the library's Montgomery form holding shape.points[] points;
    // g1[0..3] = A, B1, C, H; g2 = B2. `owned` buffers are freed with the server. Returns the key id.
    int add_key(const KeyShape& shape, const uint32_t* const g1[4], const uint32_t* g2, size_t key_bytes,
                std::vector<Buffer*> owned = {});
    KeyShape key_shape(int key);
    // Queue a job (returns its id); the buffers must stay valid until the job completes.
    uint64_t submit(int key, uint32_t* const polys[SERVE_POLYS], uint32_t inverse_mask, const uint32_t* witness);
    // A quotient job: a, b, c are A, B, C's evaluations on the domain (2^log_n each, clobbered);
    // the job computes the key's H form into a (h_poly.hpp), runs the H MSM on it and A, B1, C, B2
    // on the witness.
    uint64_t submit_groth16(int key, uint32_t* a, uint32_t* b, uint32_t* c, const uint32_t* witness);
    // Block until the job is done; copies its SERVE_RESULT_WORDS result words and latency.
    void wait(uint64_t job, uint32_t* results, double* latency_ms);
    bool done(uint64_t job);
    ServerStats stats();
    void reset_stats();
    // Per-job device footprint (workspace slot) for a key of this shape.
    size_t slot_bytes(const KeyShape& shape);

  private:
    struct Key {
        KeyShape shape;
        const uint32_t* g1[4];
        const uint32_t* g2;
        std::vector<Buffer*> owned;
    };
    struct Slot {
        KeyShape shape;
        std::unique_ptr<MsmPlan> plans[SERVE_MSMS];
        std::unique_ptr<NttBatch> ntt;
        Scratch h_scratch;  // n elements for the quotient's transforms
        hipStream_t streams[SERVE_MSMS] = {};
        hipEvent_t ev_ntt = nullptr, ev_done[SERVE_MSMS] = {};
        uint32_t* host = nullptr;  // pinned results
        size_t bytes = 0;
        bool busy = false;
        uint64_t last_use = 0;
    };
    struct Job {
        uint64_t id;
        int key;
        uint32_t* polys[SERVE_POLYS];
        uint32_t inverse_mask;
        bool groth16 = false;  // polys[0..2] = a, b, c of a quotient job
        const uint32_t* witness;
        hipEvent_t ready = nullptr;  // recorded on the device stream at submit
        double t_submit = 0, t_done = 0;
        Slot* slot = nullptr;
        bool finished = false;
        uint32_t results[SERVE_RESULT_WORDS];
    };

    void loop();
    bool try_admit(Job& j);           // called with mu_ held
    void launch(Job& j);              // called with mu_ held
    Slot* acquire_slot(const KeyShape& shape);
    std::unique_ptr<Slot> make_slot(const KeyShape& shape);
    void destroy_slot(Slot* s);

    Device* dev_;
    size_t budget_, key_bytes_ = 0, slot_bytes_total_ = 0, peak_bytes_ = 0;
    uint32_t max_inflight_, inflight_ = 0, peak_inflight_ = 0;
    std::vector<Key> keys_;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::map<uint32_t, std::shared_ptr<NttPlan>> twiddles_;
    std::map<std::pair<uint32_t, int>, std::unique_ptr<HTables>> h_tables_;  // per (log_n, form)
    uint64_t enqueue(std::unique_ptr<Job> j);  // called with mu_ held
    std::map<uint64_t, std::unique_ptr<Job>> jobs_;
    std::deque<uint64_t> pending_;
    std::vector<uint64_t> running_;
    std::vector<double> latencies_;
    uint64_t next_id_ = 1, use_clock_ = 0, submitted_ = 0;
    double first_submit_ = -1, last_done_ = 0;
    std::string error_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    bool stop_ = false;
    std::thread worker_;
};

}  // namespace zkg
