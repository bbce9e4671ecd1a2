// Batch executor (see batch.hpp).
#include "batch.hpp"

#include <algorithm>

namespace zkg {

static constexpr size_t MAX_STREAMS = 8;

Batch::Batch(Device* d) : dev_(d), ntt_(d) {
    device_bind(d);
    ZKG_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
}

Batch::~Batch() {
    (void)hipSetDevice(dev_->ordinal);
    drop_graph();
    if (origin_) {
        (void)hipStreamSynchronize(origin_);
        (void)hipStreamDestroy(origin_);
    }
    for (hipStream_t s : streams_) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (hipEvent_t e : done_) (void)hipEventDestroy(e);
    (void)hipEventDestroy(fork_);
}

void Batch::ensure_streams(size_t k) {
    while (streams_.size() < k) {
        hipStream_t s;
        ZKG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
        streams_.push_back(s);
        hipEvent_t e;
        ZKG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        done_.push_back(e);
    }
}

size_t Batch::add_ntt(uint32_t log_n, uint32_t* data, bool inverse) {
    drop_graph();
    return ntt_.add(log_n, data, inverse);
}

size_t Batch::add_msm(const MsmPlan* plan, const uint32_t* bases, const uint32_t* scalars) {
    for (const MsmJob& j : msm_)
        if (j.plan == plan) throw Error(ZKG_ERR_INVALID, "an MSM plan can appear only once in a batch (plans are not re-entrant)");
    drop_graph();
    msm_.push_back(MsmJob{plan, bases, scalars});
    return msm_.size() - 1;
}

void Batch::clear() {
    drop_graph();
    ntt_.clear();
    msm_.clear();
}

void Batch::drop_graph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
}

void Batch::enqueue(hipStream_t from) {
    const size_t lanes = std::min(MAX_STREAMS, 1 + msm_.size());
    ensure_streams(lanes);
    ZKG_HIP(hipEventRecord(fork_, from));
    std::vector<bool> used(lanes, false);
    if (ntt_.size()) {
        ZKG_HIP(hipStreamWaitEvent(streams_[0], fork_, 0));
        ntt_.launch(streams_[0]);
        used[0] = true;
    }
    for (size_t j = 0; j < msm_.size(); ++j) {
        const size_t k = lanes == 1 ? 0 : 1 + j % (lanes - 1);
        if (!used[k]) ZKG_HIP(hipStreamWaitEvent(streams_[k], fork_, 0));
        msm_[j].plan->run(msm_[j].bases, msm_[j].scalars, streams_[k]);
        used[k] = true;
    }
    for (size_t k = 0; k < lanes; ++k) {
        if (!used[k]) continue;
        ZKG_HIP(hipEventRecord(done_[k], streams_[k]));
        ZKG_HIP(hipStreamWaitEvent(from, done_[k], 0));
    }
}

void Batch::run() {
    device_bind(dev_);
    if (exec_) {
        ZKG_HIP(hipGraphLaunch(exec_, dev_->stream));
        return;
    }
    if (ntt_.size()) ntt_.prepare(dev_->stream);  // the table upload precedes the fork
    enqueue(dev_->stream);
}

void Batch::capture() {
    device_bind(dev_);
    drop_graph();
    if (!origin_) ZKG_HIP(hipStreamCreateWithFlags(&origin_, hipStreamNonBlocking));  // default priority
    if (ntt_.size()) {
        ntt_.prepare(dev_->stream);
        ZKG_HIP(hipStreamSynchronize(dev_->stream));  // tables resident before the graph refers to them
    }
    ensure_streams(std::min(MAX_STREAMS, 1 + msm_.size()));
    ZKG_HIP(hipStreamBeginCapture(origin_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue(origin_);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(origin_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    ZKG_HIP(hipStreamEndCapture(origin_, &graph_));
    ZKG_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

}  // namespace zkg
