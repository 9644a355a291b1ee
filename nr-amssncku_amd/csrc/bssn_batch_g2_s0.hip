// bssn_batch_g2_s0.hip -- tile-list kernels for GAUGE 2, tile shape 0.
#include "bssn_batch_impl.h"

namespace amss {
AMSS_BATCH_TU(2, 0)
}  // namespace amss
