// G2 instances of the point-law test kernels, the generator-multiples kernel and the 2^64-multiple
// kernel of precomputed MSM bases (kept in their
// own translation unit so the large Fq2 kernels compile in parallel with the rest).
#include "util_point.hpp"

namespace zkg {
template void launch_point_op<Fq2>(int, const uint32_t*, const uint32_t*, uint32_t*, size_t, hipStream_t);
template void launch_generator_multiples<Fq2>(const uint32_t*, uint32_t*, size_t, hipStream_t);
template void launch_shift64<Fq2>(const uint32_t*, uint32_t*, size_t, hipStream_t);
template void launch_check_bases<Fq2>(const uint32_t*, size_t, unsigned long long*, hipStream_t);
}  // namespace zkg
