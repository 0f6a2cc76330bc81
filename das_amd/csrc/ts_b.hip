// Instantiations of the z-streaming blocked kernel (see ts.h), split over
// two files so that they compile in parallel.
#include "ts.h"

namespace tsw {
TSW_TS_CONFIGS_B(TSW_TS_INSTANTIATE)
}  // namespace tsw
