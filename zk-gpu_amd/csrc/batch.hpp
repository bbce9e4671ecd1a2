// Batched execution of many MSMs and NTTs of uneven size.
//
// No counterpart in the reference (it runs one call at a time). A batch forks from the device's
// stream: all NTT jobs run as grouped launches (one per pass phase) on one stream, every MSM job
// runs on its own stream, and the device stream waits for all of them, so a batch is ordered
// like a single call. MSM plans are not re-entrant, so every MSM job needs its own plan.
#pragma once
#include <vector>

#include "msm.hpp"
#include "ntt.hpp"

namespace zkg {

class Batch {
  public:
    explicit Batch(Device* d);
    ~Batch();
    size_t add_ntt(uint32_t log_n, uint32_t* data, bool inverse);
    size_t add_msm(const MsmPlan* plan, const uint32_t* bases, const uint32_t* scalars);
    void clear();
    // Queue everything; work queued on the device stream afterwards runs after the whole batch.
    // With a current captured graph, run() replays it with one launch.
    void run();
    // Capture the whole batch (every kernel on every fork stream) into one hipGraph; later runs
    // replay it until the batch changes. The NTT tables are prepared and uploaded first.
    void capture();
    bool captured() const { return exec_ != nullptr; }
    size_t ntt_jobs() const { return ntt_.size(); }
    size_t msm_jobs() const { return msm_.size(); }

  private:
    struct MsmJob {
        const MsmPlan* plan;
        const uint32_t* bases;
        const uint32_t* scalars;
    };
    Device* dev_;
    NttBatch ntt_;
    std::vector<MsmJob> msm_;
    std::vector<hipStream_t> streams_;  // [0] NTTs, [1..] MSMs
    std::vector<hipEvent_t> done_;
    hipEvent_t fork_ = nullptr;
    hipStream_t origin_ = nullptr;  // capture origin
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    void ensure_streams(size_t k);
    void enqueue(hipStream_t from);  // fork from `from`, queue every job, join back into `from`
    void drop_graph();
};

}  // namespace zkg
