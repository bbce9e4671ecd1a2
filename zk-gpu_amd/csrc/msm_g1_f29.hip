// G1 instantiation of the MSM curve stages with the field in nine 29-bit limbs (Fq29, field29.hpp):
// the default G1 pipeline (ZKGPU_MSM_F29=0 selects the Fq instantiation in msm_g1.hip).
#include "msm_kernels.hpp"

namespace zkg {
template void launch_accumulate<Fq29>(const MsmPlan&, const uint32_t*, hipStream_t);
template void launch_merge<Fq29>(const MsmPlan&, uint32_t, hipStream_t);
template void launch_finalize<Fq29>(const MsmPlan&, hipStream_t);
template void launch_reduce<Fq29>(const MsmPlan&, hipStream_t);
template void launch_tree<Fq29>(const MsmPlan&, hipStream_t);
template void launch_combine<Fq29>(const MsmPlan&, hipStream_t);
}  // namespace zkg
