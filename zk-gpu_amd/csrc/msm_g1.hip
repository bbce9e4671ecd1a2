// G1 (coordinates in Fq) instantiation of the MSM curve stages.
#include "msm_kernels.hpp"

namespace zkg {
template void launch_accumulate<Fq>(const MsmPlan&, const uint32_t*, hipStream_t);
template void launch_merge<Fq>(const MsmPlan&, uint32_t, hipStream_t);
template void launch_finalize<Fq>(const MsmPlan&, hipStream_t);
template void launch_reduce<Fq>(const MsmPlan&, hipStream_t);
template void launch_tree<Fq>(const MsmPlan&, hipStream_t);
template void launch_combine<Fq>(const MsmPlan&, hipStream_t);
}  // namespace zkg
