// ssg_beam_v1_4.hip -- beam-search kernel instantiations for dp = 16, dp = 32, dp = 48, dp = 64.
#include "ssg_beam.h"

namespace ssg {
template hipError_t launch_v<1>(const BeamArgs&, int, hipStream_t);
template hipError_t launch_v<2>(const BeamArgs&, int, hipStream_t);
template hipError_t launch_v<3>(const BeamArgs&, int, hipStream_t);
template hipError_t launch_v<4>(const BeamArgs&, int, hipStream_t);
}  // namespace ssg
