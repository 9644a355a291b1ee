// bssn_batch_g2_s1.hip -- tile-list kernels for GAUGE 2, tile shape 1.
#include "bssn_batch_impl.h"

namespace amss {
AMSS_BATCH_TU(2, 1)
}  // namespace amss
