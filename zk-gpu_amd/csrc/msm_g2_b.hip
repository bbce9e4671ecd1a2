// G2 (coordinates in Fq2) instantiation of the MSM curve stages, part 2: heavy-bucket merge
// rounds and bucket finalisation.
#include "msm_kernels.hpp"

namespace zkg {
template void launch_merge<Fq2>(const MsmPlan&, uint32_t, hipStream_t);
template void launch_finalize<Fq2>(const MsmPlan&, hipStream_t);
}  // namespace zkg
