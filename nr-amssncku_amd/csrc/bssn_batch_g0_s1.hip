// bssn_batch_g0_s1.hip -- tile-list kernels for GAUGE 0, tile shape 1.
#include "bssn_batch_impl.h"

namespace amss {
AMSS_BATCH_TU(0, 1)
}  // namespace amss
