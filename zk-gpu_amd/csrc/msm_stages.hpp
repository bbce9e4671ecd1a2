// Declarations of the per-group MSM stage launchers (defined in msm_kernels.hpp and
// instantiated per group in msm_g1.hip and msm_g2_*.hip).
#pragma once
#include "curve.hpp"
#include "msm.hpp"

namespace zkg {

static constexpr uint32_t SUM_GROUP = 4;    // fan-in of the heavy-bucket merge rounds
static constexpr uint32_t FINAL_GROUP = 16;  // partial sums per bucket that finalize adds directly
static constexpr uint32_t MERGE_BLOCKS = 256;  // workgroups of a merge round (one per CU; they stride)

template <class Fe> void launch_accumulate(const MsmPlan& p, const uint32_t* bases, hipStream_t s);
template <class Fe> void launch_merge(const MsmPlan& p, uint32_t stride, hipStream_t s);
template <class Fe> void launch_finalize(const MsmPlan& p, hipStream_t s);
template <class Fe> void launch_reduce(const MsmPlan& p, hipStream_t s);
template <class Fe> void launch_tree(const MsmPlan& p, hipStream_t s);
template <class Fe> void launch_combine(const MsmPlan& p, hipStream_t s);

#define ZKG_MSM_EXTERN(Fe)                                                                  \
    extern template void launch_accumulate<Fe>(const MsmPlan&, const uint32_t*, hipStream_t); \
    extern template void launch_merge<Fe>(const MsmPlan&, uint32_t, hipStream_t);             \
    extern template void launch_finalize<Fe>(const MsmPlan&, hipStream_t);                    \
    extern template void launch_reduce<Fe>(const MsmPlan&, hipStream_t);                      \
    extern template void launch_tree<Fe>(const MsmPlan&, hipStream_t);                        \
    extern template void launch_combine<Fe>(const MsmPlan&, hipStream_t);

}  // namespace zkg
