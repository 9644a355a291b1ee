// bssn_edge_update.hip -- the RK4 update of the LDS boundary tiles' points.
#include "bssn_edge_lds.h"

namespace amss {

void launch_edge_lds_update(const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out,
                                   const AuxOut& aux, const Box& box, const EdgeTileSet& T, const double* K,
                                   const Rk4Param& R, hipStream_t s) {
  const unsigned gr = (unsigned)T.tiles();
#define AMSS_UPD(ST) \
  edge_update_kernel<ST><<<dim3(gr, UpdateItems<ST>::kCount), kEdgeThreads, 0, s>>>(u0, acc, out, aux, box, T, K, R)
  switch (R.stage) {
    case 0: AMSS_UPD(0); break;
    case 1: AMSS_UPD(1); break;
    case 2: AMSS_UPD(2); break;
    default: AMSS_UPD(3); break;
  }
#undef AMSS_UPD
}

}  // namespace amss
