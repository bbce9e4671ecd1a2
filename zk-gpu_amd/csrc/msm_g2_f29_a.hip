// G2 instantiation of the MSM curve stages with Fq2 over 29-bit limbs (Fq2_29, field29.hpp): the
// default G2 pipeline after the 32-bit bucket accumulation (ZKGPU_MSM_F29=0 selects msm_g2_*.hip).
#include "msm_kernels.hpp"

namespace zkg {
template void launch_accumulate<Fq2_29>(const MsmPlan&, const uint32_t*, hipStream_t);
}  // namespace zkg
