// G2 (coordinates in Fq2) instantiation of the MSM curve stages, part 3: running-sum bucket
// reduction, per-window tree sum and the window combine.
#include "msm_kernels.hpp"

namespace zkg {
template void launch_reduce<Fq2>(const MsmPlan&, hipStream_t);
template void launch_tree<Fq2>(const MsmPlan&, hipStream_t);
template void launch_combine<Fq2>(const MsmPlan&, hipStream_t);
}  // namespace zkg
