// Autotuner: picks, per GPU architecture and problem size, the MSM window c, work-item size and
// counting-sort block size, and the NTT's LDS tile, by timing candidates on the device with HIP
// events. Results live in a process-wide table that plans consult, and can be saved to and
// loaded from a text file (ZKGPU_TUNE_FILE names one to load at first use).
//
// No counterpart in the reference (its window comes from default_window and its tile sizes are
// fixed). Every tuned choice changes only how the work is split: results are bit-identical.
#pragma once
#include <string>

#include "runtime.hpp"

namespace zkg {

struct MsmTune {
    uint32_t c = 0, item_points = 0, sort_points = 0;
    double ms = 0;  // measured time of the chosen configuration
};

struct NttTune {
    int tile_log = 0;
    double ms = 0;
};

// ceil(log2(n)): the table key for an MSM of n points (n and n - 1 share a key)
uint32_t tune_size_key(size_t n);

bool tune_lookup_msm(const std::string& arch, int group, uint32_t key, MsmTune* out);
bool tune_lookup_ntt(const std::string& arch, uint32_t log_n, NttTune* out);
void tune_store_msm(const std::string& arch, int group, uint32_t key, const MsmTune& t);
void tune_store_ntt(const std::string& arch, uint32_t log_n, const NttTune& t);
void tune_load(const std::string& path);  // merges the file's entries into the table
void tune_save(const std::string& path);
void tune_clear();

// Searches (coordinate descent: c, then item size, then sort block) and records the winner.
MsmTune tune_msm(Device* d, int group, size_t n);
NttTune tune_ntt(Device* d, uint32_t log_n);

// The configuration a plan for (group, n) uses: the table's entry, else default_window(n) with
// automatic item and sort sizes.
MsmTune msm_config(const Device* d, int group, size_t n);

}  // namespace zkg
