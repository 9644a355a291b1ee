// bssn_batch_g0_s2.hip -- tile-list kernels for GAUGE 0, tile shape 2.
#include "bssn_batch_impl.h"

namespace amss {
AMSS_BATCH_TU(0, 2)
}  // namespace amss
