// bssn_batch_g0_s3.hip -- tile-list kernels for GAUGE 0, tile shape 3.
#include "bssn_batch_impl.h"

namespace amss {
AMSS_BATCH_TU(0, 3)
}  // namespace amss
