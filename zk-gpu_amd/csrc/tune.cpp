// Autotuner (see tune.hpp).
#include "tune.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <tuple>
#include <vector>

#include "msm.hpp"
#include "ntt.hpp"

namespace zkg {

// Synthetic MSM inputs (util_kernels.hip): bases k_i G in the library's form and scalars.
void synthetic_msm_inputs(Device* d, int group, size_t n, uint64_t seed, uint32_t* bases, uint32_t* scalars);

namespace {

struct Table {
    std::mutex mu;
    std::map<std::tuple<std::string, int, uint32_t>, MsmTune> msm;
    std::map<std::pair<std::string, uint32_t>, NttTune> ntt;
    bool env_loaded = false;
};

Table& table() {
    static Table t;
    return t;
}

void load_locked(Table& t, const std::string& path, bool must_exist) {
    std::ifstream f(path);
    if (!f) {
        if (must_exist) throw Error(ZKG_ERR_INVALID, "cannot read tuning file " + path);
        return;
    }
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        if (line.empty() || line[0] == '#') continue;
        std::istringstream in(line);
        std::string kind, arch;
        in >> kind >> arch;
        if (kind == "msm") {
            int group = 0;
            uint32_t key = 0;
            MsmTune m;
            in >> group >> key >> m.c >> m.item_points >> m.sort_points >> m.ms;
            if (!in || (group != 1 && group != 2) || m.c < 9 || m.c > 20)
                throw Error(ZKG_ERR_INVALID, path + ":" + std::to_string(lineno) + ": bad msm entry");
            t.msm[std::make_tuple(arch, group, key)] = m;
        } else if (kind == "ntt") {
            uint32_t log_n = 0;
            NttTune k;
            in >> log_n >> k.tile_log >> k.ms;
            if (!in || k.tile_log < 10 || k.tile_log > 12)
                throw Error(ZKG_ERR_INVALID, path + ":" + std::to_string(lineno) + ": bad ntt entry");
            t.ntt[std::make_pair(arch, log_n)] = k;
        } else {
            throw Error(ZKG_ERR_INVALID, path + ":" + std::to_string(lineno) + ": unknown entry kind '" + kind + "'");
        }
    }
}

Table& ready_table() {
    Table& t = table();
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.env_loaded) {
        t.env_loaded = true;
        const char* e = std::getenv("ZKGPU_TUNE_FILE");
        if (e && *e) load_locked(t, e, false);
    }
    return t;
}

struct DevWords {
    uint32_t* ptr = nullptr;
    explicit DevWords(size_t words) { ZKG_HIP(hipMalloc(&ptr, 4 * words)); }
    ~DevWords() {
        if (ptr) (void)hipFree(ptr);
    }
    DevWords(const DevWords&) = delete;
    DevWords& operator=(const DevWords&) = delete;
};

struct HipTimer {
    hipEvent_t a = nullptr, b = nullptr;
    HipTimer() {
        ZKG_HIP(hipEventCreate(&a));
        ZKG_HIP(hipEventCreate(&b));
    }
    ~HipTimer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

// min over `reps` timed runs after one warm-up, in milliseconds
template <class F>
double time_min(Device* d, F run, int reps) {
    HipTimer t;
    run();
    double best = 1e30;
    for (int r = 0; r < reps; ++r) {
        ZKG_HIP(hipEventRecord(t.a, d->stream));
        run();
        ZKG_HIP(hipEventRecord(t.b, d->stream));
        ZKG_HIP(hipEventSynchronize(t.b));
        float ms = 0;
        ZKG_HIP(hipEventElapsedTime(&ms, t.a, t.b));
        best = std::min(best, (double)ms);
    }
    return best;
}

}  // namespace

uint32_t tune_size_key(size_t n) {
    uint32_t k = 0;
    while (k < 63 && ((size_t)1 << k) < n) ++k;
    return k;
}

bool tune_lookup_msm(const std::string& arch, int group, uint32_t key, MsmTune* out) {
    Table& t = ready_table();
    std::lock_guard<std::mutex> g(t.mu);
    auto it = t.msm.find(std::make_tuple(arch, group, key));
    if (it == t.msm.end()) return false;
    if (out) *out = it->second;
    return true;
}

bool tune_lookup_ntt(const std::string& arch, uint32_t log_n, NttTune* out) {
    Table& t = ready_table();
    std::lock_guard<std::mutex> g(t.mu);
    auto it = t.ntt.find(std::make_pair(arch, log_n));
    if (it == t.ntt.end()) return false;
    if (out) *out = it->second;
    return true;
}

void tune_store_msm(const std::string& arch, int group, uint32_t key, const MsmTune& m) {
    Table& t = ready_table();
    std::lock_guard<std::mutex> g(t.mu);
    t.msm[std::make_tuple(arch, group, key)] = m;
}

void tune_store_ntt(const std::string& arch, uint32_t log_n, const NttTune& k) {
    Table& t = ready_table();
    std::lock_guard<std::mutex> g(t.mu);
    t.ntt[std::make_pair(arch, log_n)] = k;
}

void tune_load(const std::string& path) {
    Table& t = ready_table();
    std::lock_guard<std::mutex> g(t.mu);
    load_locked(t, path, true);
}

void tune_save(const std::string& path) {
    Table& t = ready_table();
    std::lock_guard<std::mutex> g(t.mu);
    std::ofstream f(path);
    if (!f) throw Error(ZKG_ERR_INVALID, "cannot write tuning file " + path);
    f << "# zkgpu tuning table v1\n"
      << "# msm <arch> <group> <ceil log2 n> <c> <item points> <sort points> <ms>\n"
      << "# ntt <arch> <log_n> <tile log2> <ms>\n";
    for (const auto& e : t.msm) {
        const MsmTune& m = e.second;
        f << "msm " << std::get<0>(e.first) << ' ' << std::get<1>(e.first) << ' ' << std::get<2>(e.first) << ' ' << m.c
          << ' ' << m.item_points << ' ' << m.sort_points << ' ' << m.ms << '\n';
    }
    for (const auto& e : t.ntt)
        f << "ntt " << e.first.first << ' ' << e.first.second << ' ' << e.second.tile_log << ' ' << e.second.ms << '\n';
    if (!f) throw Error(ZKG_ERR_INVALID, "cannot write tuning file " + path);
}

void tune_clear() {
    Table& t = ready_table();
    std::lock_guard<std::mutex> g(t.mu);
    t.msm.clear();
    t.ntt.clear();
}

MsmTune msm_config(const Device* d, int group, size_t n) {
    MsmTune m;
    if (d && tune_lookup_msm(d->name, group, tune_size_key(n), &m)) return m;
    m = MsmTune();
    m.c = msm_default_window(n);
    return m;
}

MsmTune tune_msm(Device* d, int group, size_t n) {
    if (n < 1 || n > 0x7fffffffu) throw Error(ZKG_ERR_INVALID, "an MSM holds 1 .. 2^31 - 1 points");
    if (group != 1 && group != 2) throw Error(ZKG_ERR_INVALID, "group must be G1 (1) or G2 (2)");
    device_bind(d);
    const size_t aw = group == 1 ? 16 : 32;
    DevWords inputs(aw * n + 8 * n);
    uint32_t* bases = inputs.ptr;
    uint32_t* scalars = inputs.ptr + aw * n;
    synthetic_msm_inputs(d, group, n, 0x5eed0000u + (uint64_t)group, bases, scalars);
    const int reps = n >= ((size_t)1 << 22) ? 2 : 3;
    auto measure = [&](const MsmTune& cand) {
        MsmPlan plan(d, group, n, cand.c, cand.item_points, cand.sort_points);
        return time_min(d, [&] { plan.run(bases, scalars, d->stream); }, reps);
    };
    const uint32_t c0 = msm_default_window(n);
    MsmTune best;
    best.c = c0;
    best.ms = measure(best);
    // 1. window (the LDS counting sort covers c <= 16)
    for (int dc : {-2, -1, 1, 2}) {
        const int c = (int)c0 + dc;
        if (c < 9 || c > 16) continue;
        MsmTune cand;
        cand.c = (uint32_t)c;
        cand.ms = measure(cand);
        if (cand.ms < best.ms) best = cand;
    }
    // 2. work-item size around the automatic one
    const uint32_t auto_ch = msm_plan_item_points(group, n, best.c);
    for (uint32_t ch : {auto_ch / 2, auto_ch * 2}) {
        if (ch < 4 || ch > 1024) continue;
        MsmTune cand = best;
        cand.item_points = ch;
        cand.ms = measure(cand);
        if (cand.ms < best.ms) best = cand;
    }
    // 3. points per counting-sort block around the default one
    const MsmShape best_sh = msm_shape(group, n, best.c, 0);
    const uint32_t B = best_sh.B, auto_sp = best_sh.sort_pts;
    for (uint32_t sp : {auto_sp / 2, auto_sp * 2, auto_sp * 4}) {
        if (sp < B) continue;
        MsmTune cand = best;
        cand.sort_points = sp;
        cand.ms = measure(cand);
        if (cand.ms < best.ms) best = cand;
    }
    tune_store_msm(d->name, group, tune_size_key(n), best);
    return best;
}

NttTune tune_ntt(Device* d, uint32_t log_n) {
    if (log_n < 10 || log_n > 27) throw Error(ZKG_ERR_INVALID, "log_n must be in 10..=27");
    device_bind(d);
    const size_t n = (size_t)1 << log_n;
    DevWords data(8 * n);
    ZKG_HIP(hipMemsetAsync(data.ptr, 0, 32 * n, d->stream));
    NttTune best;
    best.ms = 1e30;
    for (int tile : {10, 11, 12}) {
        if ((uint32_t)tile > log_n) continue;
        NttShape shape;
        shape.tile_log = tile;
        NttPlan plan(d, log_n, shape);
        const double ms = time_min(d, [&] { plan.run(data.ptr, false, d->stream); }, 3);
        if (ms < best.ms) {
            best.ms = ms;
            best.tile_log = tile;
        }
    }
    tune_store_ntt(d->name, log_n, best);
    return best;
}

}  // namespace zkg
