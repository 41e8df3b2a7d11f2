// ssg_beam_v12_16.hip -- beam-search kernel instantiations for dp = 192, dp = 256.
#include "ssg_beam.h"

namespace ssg {
template hipError_t launch_v<12>(const BeamArgs&, int, hipStream_t);
template hipError_t launch_v<16>(const BeamArgs&, int, hipStream_t);
}  // namespace ssg
