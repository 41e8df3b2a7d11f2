// ssg_beam_v24_32.hip -- beam-search kernel instantiations for dp = 384, dp = 512.
#include "ssg_beam.h"

namespace ssg {
template hipError_t launch_v<24>(const BeamArgs&, int, hipStream_t);
template hipError_t launch_v<32>(const BeamArgs&, int, hipStream_t);
}  // namespace ssg
