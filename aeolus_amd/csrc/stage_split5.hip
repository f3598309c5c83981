// Legendre stage kernel instances at precision NS = 5 (0 = FP64 contractions,
// k > 0 = k INT8 digit slices); see stage_launch.hpp.
#include "stage_launch_impl.hpp"

namespace aeo {
AEO_STAGE_INSTANCES(5)
}  // namespace aeo
