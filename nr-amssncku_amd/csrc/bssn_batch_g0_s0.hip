// bssn_batch_g0_s0.hip -- tile-list kernels for GAUGE 0, tile shape 0.
#include "bssn_batch_impl.h"

namespace amss {
AMSS_BATCH_TU(0, 0)
}  // namespace amss
