// ssg_beam_v48_64.hip -- beam-search kernel instantiations for dp = 768, dp = 1024.
#include "ssg_beam.h"

namespace ssg {
template hipError_t launch_v<48>(const BeamArgs&, int, hipStream_t);
template hipError_t launch_v<64>(const BeamArgs&, int, hipStream_t);
}  // namespace ssg
