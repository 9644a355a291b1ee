// bssn_batch_g2_s3.hip -- tile-list kernels for GAUGE 2, tile shape 3.
#include "bssn_batch_impl.h"

namespace amss {
AMSS_BATCH_TU(2, 3)
}  // namespace amss
