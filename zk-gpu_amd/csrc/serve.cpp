// Proof-job server (see serve.hpp).
#include "serve.hpp"
#include "tune.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace zkg {

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

size_t KeyShape::witness_len() const {
    return std::max(std::max(points[0], points[1]), std::max(points[4], c_offset + points[2]));
}

KeyShape KeyShape::groth16(uint32_t log_n) {
    KeyShape k;
    k.log_n = log_n;
    const size_t n = (size_t)1 << log_n;
    for (int m = 0; m < SERVE_MSMS; ++m) k.points[m] = m == 3 ? n - 1 : n;
    return k;
}

void KeyShape::validate() const {
    if (log_n < 10 || log_n > 27) throw Error(ZKG_ERR_INVALID, "log_n must be in 10..=27");
    for (int m = 0; m < SERVE_MSMS; ++m) {
        if (points[m] < 1) throw Error(ZKG_ERR_INVALID, "an MSM needs at least one point");
        if (points[m] > 0x7fffffffu) throw Error(ZKG_ERR_INVALID, "an MSM holds at most 2^31 - 1 points");
    }
    if (points[3] > ((size_t)1 << log_n))
        throw Error(ZKG_ERR_INVALID, "the H MSM takes its scalars from one transform: at most 2^log_n points");
}

Server::Server(Device* d, size_t mem_budget, uint32_t max_inflight)
    : dev_(d), budget_(mem_budget), max_inflight_(max_inflight ? max_inflight : 1) {
    worker_ = std::thread([this] { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
        pending_.clear();
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    worker_.join();
    (void)hipSetDevice(dev_->ordinal);
    for (auto& s : slots_) destroy_slot(s.get());
    slots_.clear();
    for (auto& j : jobs_)
        if (j.second->ready) (void)hipEventDestroy(j.second->ready);
    for (auto& k : keys_)
        for (Buffer* b : k.owned) buffer_free(b);
}

size_t Server::slot_bytes(const KeyShape& shape) {
    // NTT batch scratch (7 n elements) and the quotient's scratch (n)
    size_t bytes = (size_t)(SERVE_POLYS + 1) * 32 * ((size_t)1 << shape.log_n);
    for (int m = 0; m < SERVE_MSMS; ++m) {
        const int group = m == 4 ? 2 : 1;
        const MsmTune t = msm_config(dev_, group, shape.points[m]);
        bytes += msm_workspace_bytes(group, shape.points[m], t.c, t.item_points, t.sort_points);
    }
    return bytes;
}

int Server::add_key(const KeyShape& shape, const uint32_t* const g1[4], const uint32_t* g2, size_t key_bytes,
                    std::vector<Buffer*> owned) {
    shape.validate();
    std::lock_guard<std::mutex> g(mu_);
    if (key_bytes_ + key_bytes > budget_) throw Error(ZKG_ERR_NOMEM, "proving key does not fit the server's HBM budget");
    Key k;
    k.shape = shape;
    for (int i = 0; i < 4; ++i) k.g1[i] = g1[i];
    k.g2 = g2;
    k.owned = std::move(owned);
    keys_.push_back(std::move(k));
    key_bytes_ += key_bytes;
    peak_bytes_ = std::max(peak_bytes_, key_bytes_ + slot_bytes_total_);
    return (int)keys_.size() - 1;
}

uint64_t Server::submit(int key, uint32_t* const polys[SERVE_POLYS], uint32_t inverse_mask, const uint32_t* witness) {
    std::unique_ptr<Job> j(new Job());
    for (int p = 0; p < SERVE_POLYS; ++p) j->polys[p] = polys[p];
    j->key = key;
    j->inverse_mask = inverse_mask;
    j->witness = witness;
    std::lock_guard<std::mutex> g(mu_);
    return enqueue(std::move(j));
}

uint64_t Server::submit_groth16(int key, uint32_t* a, uint32_t* b, uint32_t* c, const uint32_t* witness) {
    std::unique_ptr<Job> j(new Job());
    for (int p = 0; p < SERVE_POLYS; ++p) j->polys[p] = nullptr;
    j->polys[0] = a;
    j->polys[1] = b;
    j->polys[2] = c;
    j->groth16 = true;
    j->key = key;
    j->inverse_mask = 0;
    j->witness = witness;
    std::lock_guard<std::mutex> g(mu_);
    return enqueue(std::move(j));
}

uint64_t Server::enqueue(std::unique_ptr<Job> j) {
    {
        const int key = j->key;
        if (key < 0 || (size_t)key >= keys_.size()) throw Error(ZKG_ERR_INVALID, "unknown proving key");
        if (!error_.empty()) throw Error(ZKG_ERR_DEVICE, "server stopped after an error: " + error_);
        const KeyShape& shape = keys_[key].shape;
        const size_t need = slot_bytes(shape);
        // a quotient job needs the scale tables of its size and form, made at the first such job
        // and shared by every key of that size and form (counted with the keys)
        const auto tkey = std::make_pair(shape.log_n, (int)shape.h_form);
        const size_t tbytes = j->groth16 && !h_tables_.count(tkey)
                                  ? (size_t)(shape.h_form == HForm::coefficients ? 64 : 32) << shape.log_n
                                  : 0;
        if (key_bytes_ + tbytes + need > budget_)
            throw Error(ZKG_ERR_NOMEM, "a job for this key needs " + std::to_string(need + tbytes) +
                                           " bytes of workspace, more than the server's HBM budget leaves");
        if (tbytes) {
            h_tables_[tkey].reset(new HTables(dev_, shape.log_n, shape.h_form));
            key_bytes_ += tbytes;
            peak_bytes_ = std::max(peak_bytes_, key_bytes_ + slot_bytes_total_);
        }
        j->id = next_id_++;
        j->t_submit = now_ms();
        // the job's inputs may still be written by work queued on the device stream
        ZKG_HIP(hipEventCreateWithFlags(&j->ready, hipEventDisableTiming));
        ZKG_HIP(hipEventRecord(j->ready, dev_->stream));
        if (first_submit_ < 0) first_submit_ = j->t_submit;
        ++submitted_;
        const uint64_t id = j->id;
        jobs_[id] = std::move(j);
        pending_.push_back(id);
        cv_work_.notify_all();
        return id;
    }
}

bool Server::done(uint64_t job) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) throw Error(ZKG_ERR_INVALID, "unknown job");
    return it->second->finished;
}

void Server::wait(uint64_t job, uint32_t* results, double* latency_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(job);
    if (it == jobs_.end()) throw Error(ZKG_ERR_INVALID, "unknown job");
    cv_done_.wait(lk, [&] { return it->second->finished || !error_.empty() || stop_; });
    if (!it->second->finished)
        throw Error(ZKG_ERR_DEVICE, stop_ ? std::string("server is shutting down") : "server stopped after an error: " + error_);
    if (results) std::memcpy(results, it->second->results, sizeof(it->second->results));
    if (latency_ms) *latency_ms = it->second->t_done - it->second->t_submit;
    if (it->second->ready) (void)hipEventDestroy(it->second->ready);
    jobs_.erase(it);
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> g(mu_);
    ServerStats s;
    s.submitted = submitted_;
    s.completed = latencies_.size();
    if (!latencies_.empty()) {
        std::vector<double> v = latencies_;
        std::sort(v.begin(), v.end());
        auto q = [&](double f) { return v[std::min(v.size() - 1, (size_t)(f * (v.size() - 1) + 0.5))]; };
        s.p50_ms = q(0.50);
        s.p99_ms = q(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s.mean_ms = sum / v.size();
        if (last_done_ > first_submit_) s.jobs_per_s = 1e3 * v.size() / (last_done_ - first_submit_);
    }
    s.bytes_in_use = key_bytes_ + slot_bytes_total_;
    s.peak_bytes = peak_bytes_;
    s.inflight = inflight_;
    s.peak_inflight = peak_inflight_;
    return s;
}

void Server::reset_stats() {
    std::lock_guard<std::mutex> g(mu_);
    latencies_.clear();
    submitted_ = 0;
    first_submit_ = -1;
    last_done_ = 0;
    peak_inflight_ = inflight_;
}

KeyShape Server::key_shape(int key) {
    std::lock_guard<std::mutex> g(mu_);
    if (key < 0 || (size_t)key >= keys_.size()) throw Error(ZKG_ERR_INVALID, "unknown proving key");
    return keys_[key].shape;
}

std::unique_ptr<Server::Slot> Server::make_slot(const KeyShape& shape) {
    std::unique_ptr<Slot> s(new Slot());
    s->shape = shape;
    const uint32_t log_n = shape.log_n;
    // window, item and sort sizes from the tuning table (default_window when untuned)
    for (int m = 0; m < SERVE_MSMS; ++m) {
        const int group = m == 4 ? 2 : 1;
        const MsmTune t = msm_config(dev_, group, shape.points[m]);
        s->plans[m].reset(new MsmPlan(dev_, group, shape.points[m], t.c, t.item_points, t.sort_points));
    }
    s->ntt.reset(new NttBatch(dev_));
    s->h_scratch.ensure((size_t)32 << log_n);
    auto tw = twiddles_.find(log_n);
    if (tw == twiddles_.end()) {
        NttShape ns;
        ns.tile_log = NttBatch::tile_for(dev_, log_n);
        tw = twiddles_.emplace(log_n, std::make_shared<NttPlan>(dev_, log_n, ns, false)).first;
    }
    s->ntt->share_plan(tw->second);
    for (int k = 0; k < SERVE_MSMS; ++k) {
        ZKG_HIP(hipStreamCreateWithFlags(&s->streams[k], hipStreamNonBlocking));  // default priority
        ZKG_HIP(hipEventCreateWithFlags(&s->ev_done[k], hipEventDisableTiming));
    }
    ZKG_HIP(hipEventCreateWithFlags(&s->ev_ntt, hipEventDisableTiming));
    ZKG_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->host), 4 * SERVE_RESULT_WORDS));
    s->bytes = slot_bytes(shape);
    return s;
}

void Server::destroy_slot(Slot* s) {
    for (int k = 0; k < SERVE_MSMS; ++k) {
        if (s->streams[k]) {
            (void)hipStreamSynchronize(s->streams[k]);
            (void)hipStreamDestroy(s->streams[k]);
        }
        if (s->ev_done[k]) (void)hipEventDestroy(s->ev_done[k]);
    }
    if (s->ev_ntt) (void)hipEventDestroy(s->ev_ntt);
    s->h_scratch.release();
    if (s->host) (void)hipHostFree(s->host);
    s->host = nullptr;
}

Server::Slot* Server::acquire_slot(const KeyShape& shape) {
    for (auto& s : slots_)
        if (!s->busy && s->shape == shape) return s.get();
    const size_t need = slot_bytes(shape);
    // evict idle slots of other sizes, least recently used first, until the new one fits
    while (key_bytes_ + slot_bytes_total_ + need > budget_) {
        Slot* victim = nullptr;
        for (auto& s : slots_)
            if (!s->busy && (!victim || s->last_use < victim->last_use)) victim = s.get();
        if (!victim) return nullptr;
        destroy_slot(victim);
        slot_bytes_total_ -= victim->bytes;
        slots_.erase(std::find_if(slots_.begin(), slots_.end(), [&](const std::unique_ptr<Slot>& p) { return p.get() == victim; }));
    }
    slots_.push_back(make_slot(shape));
    slot_bytes_total_ += slots_.back()->bytes;
    peak_bytes_ = std::max(peak_bytes_, key_bytes_ + slot_bytes_total_);
    return slots_.back().get();
}

bool Server::try_admit(Job& j) {
    if (inflight_ >= max_inflight_) return false;
    Slot* s = acquire_slot(keys_[j.key].shape);
    if (!s) return false;
    s->busy = true;
    s->last_use = ++use_clock_;
    j.slot = s;
    return true;
}

void Server::launch(Job& j) {
    Slot* s = j.slot;
    const Key& k = keys_[j.key];
    hipStream_t s0 = s->streams[0];
    for (int t = 0; t < SERVE_MSMS; ++t) ZKG_HIP(hipStreamWaitEvent(s->streams[t], j.ready, 0));
    const uint32_t* h_scalars;
    if (j.groth16) {
        // the quotient (six or seven transforms and the pointwise step), then H over it
        const HTables& t = *h_tables_.at(std::make_pair(k.shape.log_n, (int)k.shape.h_form));
        h_compute(*twiddles_.at(k.shape.log_n), static_cast<uint32_t*>(s->h_scratch.ptr), t, j.polys[0], j.polys[1],
                  j.polys[2], s0);
        h_scalars = j.polys[0];
    } else {
        // the NTT bundle, then H (its scalars are the first outputs of the seventh transform)
        s->ntt->clear();
        for (int p = 0; p < SERVE_POLYS; ++p) s->ntt->add(k.shape.log_n, j.polys[p], (j.inverse_mask >> p) & 1);
        s->ntt->run(s0);
        h_scalars = j.polys[SERVE_POLYS - 1];
    }
    s->plans[3]->run(k.g1[3], h_scalars, s0);
    ZKG_HIP(hipMemcpyAsync(s->host + 3 * 32, s->plans[3]->result, 4 * 32, hipMemcpyDeviceToHost, s0));
    // A, B1, C and B2 on their own streams
    const int order[4] = {0, 1, 2, 4};
    const size_t offset[4] = {0, 32, 64, 128};
    for (int t = 0; t < 4; ++t) {
        const int m = order[t];
        hipStream_t st = s->streams[t + 1];
        const uint32_t* bases = m == 4 ? k.g2 : k.g1[m];
        s->plans[m]->run(bases, j.witness + (m == 2 ? 8 * k.shape.c_offset : 0), st);
        ZKG_HIP(hipMemcpyAsync(s->host + offset[t], s->plans[m]->result, 4 * (size_t)s->plans[m]->xw,
                               hipMemcpyDeviceToHost, st));
        ZKG_HIP(hipEventRecord(s->ev_done[t + 1], st));
        ZKG_HIP(hipStreamWaitEvent(s0, s->ev_done[t + 1], 0));
    }
    ZKG_HIP(hipEventRecord(s->ev_done[0], s0));
}

void Server::loop() {
    (void)hipSetDevice(dev_->ordinal);
    std::unique_lock<std::mutex> lk(mu_);
    while (true) {
        try {
            // completions
            for (size_t i = 0; i < running_.size();) {
                Job& j = *jobs_.at(running_[i]);
                hipError_t q = hipEventQuery(j.slot->ev_done[0]);
                if (q == hipErrorNotReady) {
                    ++i;
                    continue;
                }
                if (q != hipSuccess) throw Error(ZKG_ERR_DEVICE, std::string("job failed: ") + hipGetErrorString(q), true);
                std::memcpy(j.results, j.slot->host, sizeof(j.results));
                j.t_done = now_ms();
                j.finished = true;
                j.slot->busy = false;
                j.slot = nullptr;
                latencies_.push_back(j.t_done - j.t_submit);
                last_done_ = j.t_done;
                --inflight_;
                running_.erase(running_.begin() + i);
                cv_done_.notify_all();
            }
            // admissions, in arrival order
            while (!stop_ && !pending_.empty()) {
                Job& j = *jobs_.at(pending_.front());
                if (!try_admit(j)) break;
                launch(j);
                running_.push_back(j.id);
                pending_.pop_front();
                ++inflight_;
                peak_inflight_ = std::max(peak_inflight_, inflight_);
            }
        } catch (const std::exception& e) {
            // a device error ends the server: no job is reported as done after it
            error_ = e.what();
            for (uint64_t id : running_) {
                Job& j = *jobs_.at(id);
                if (j.slot) j.slot->busy = false;
                j.slot = nullptr;
            }
            running_.clear();
            inflight_ = 0;
            cv_done_.notify_all();
            // no call waits for this error: hand it to the device's error callback, unlocked
            const Error* ze = dynamic_cast<const Error*>(&e);
            if (ze && ze->gpu) {
                const std::string msg = error_;
                lk.unlock();
                report_gpu_error(dev_, msg);
                lk.lock();
            }
        }
        if (stop_ && running_.empty()) break;
        if (running_.empty() && (pending_.empty() || !error_.empty())) {
            cv_work_.wait(lk);
        } else {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        }
    }
}

}  // namespace zkg
