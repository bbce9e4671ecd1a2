// G2 (coordinates in Fq2) instantiation of the MSM curve stages, part 1: bucket accumulation.
#include "msm_kernels.hpp"

namespace zkg {
template void launch_accumulate<Fq2>(const MsmPlan&, const uint32_t*, hipStream_t);
}  // namespace zkg
