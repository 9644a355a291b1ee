// bssn_batch_g2_s2.hip -- tile-list kernels for GAUGE 2, tile shape 2.
#include "bssn_batch_impl.h"

namespace amss {
AMSS_BATCH_TU(2, 2)
}  // namespace amss
