// ssg_beam_v8.hip -- beam-search kernel instantiations for dp = 128.
#include "ssg_beam.h"

namespace ssg {
template hipError_t launch_v<8>(const BeamArgs&, int, hipStream_t);
}  // namespace ssg
