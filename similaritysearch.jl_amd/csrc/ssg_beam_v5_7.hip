// ssg_beam_v5_7.hip -- beam-search kernel instantiations for dp = 80, dp = 96, dp = 112.
#include "ssg_beam.h"

namespace ssg {
template hipError_t launch_v<5>(const BeamArgs&, int, hipStream_t);
template hipError_t launch_v<6>(const BeamArgs&, int, hipStream_t);
template hipError_t launch_v<7>(const BeamArgs&, int, hipStream_t);
}  // namespace ssg
