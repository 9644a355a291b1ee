// bssn_edge_update.hip -- the RK4 update of the LDS boundary tiles' points.
#include "bssn_edge_lds.h"

namespace amss {

void launch_edge_lds_update(const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out,
                                   const AuxOut& aux, const Box& box, const EdgeTileSet& T, const double* K,
                                   const Rk4Param& R, hipStream_t s) {
  const unsigned gr = (unsigned)T.tiles();
  switch (R.stage) {
    case 0: edge_update_kernel<0><<<gr, kEdgeThreads, 0, s>>>(u0, acc, out, aux, box, T, K, R); break;
    case 1: edge_update_kernel<1><<<gr, kEdgeThreads, 0, s>>>(u0, acc, out, aux, box, T, K, R); break;
    case 2: edge_update_kernel<2><<<gr, kEdgeThreads, 0, s>>>(u0, acc, out, aux, box, T, K, R); break;
    default: edge_update_kernel<3><<<gr, kEdgeThreads, 0, s>>>(u0, acc, out, aux, box, T, K, R); break;
  }
}

}  // namespace amss
