// bssn_edge_g0_res.hip -- the LDS boundary stencil kernel for GAUGE 0 with the constraint residuals.
#include "bssn_edge_lds.h"

namespace amss {

template void launch_edge_lds_rhs<0, true>(const FieldsIn&, const AuxOut&, const Box&, const EdgeTileSet&,
                                              const Coef&, double*, int*, hipStream_t);

}  // namespace amss
