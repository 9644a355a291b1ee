// bssn_edge_g0.hip -- the LDS boundary stencil kernel for GAUGE 0.
#include "bssn_edge_lds.h"

namespace amss {

template void launch_edge_lds_rhs<0, false>(const FieldsIn&, const AuxOut&, const Box&, const EdgeTileSet&,
                                              const Coef&, double*, int*, hipStream_t);

}  // namespace amss
