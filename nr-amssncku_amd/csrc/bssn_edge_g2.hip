// bssn_edge_g2.hip -- the LDS boundary stencil kernel for GAUGE 2.
#include "bssn_edge_lds.h"

namespace amss {

template void launch_edge_lds_rhs<2, false>(const FieldsIn&, const AuxOut&, const Box&, const EdgeTileSet&,
                                              const Coef&, double*, int*, hipStream_t);

}  // namespace amss
