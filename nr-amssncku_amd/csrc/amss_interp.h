// amss_interp.h -- the 6-point polynomial interpolation the reference's
// interpolating routines share (polint / polin3, fmisc.f90:1121-1260).
#pragma once
#include <hip/hip_runtime.h>

namespace amss {

// Neville's tableau over the samples y[0..5] at nodes 0..5, evaluated at x
// (polint, fmisc.f90:1121-1179): start from the sample nearest x, then add
// the corrections c or d that keep the path centred on it.
__device__ __forceinline__ double neville6(const double y[6], double x) {
  double c[6], d[6], ho[6];
  int ns = 0;
  double dif = fabs(x);
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    c[m] = d[m] = y[m];
    ho[m] = (double)m - x;
    const double t = fabs(x - (double)m);
    if (t < dif) {
      ns = m;
      dif = t;
    }
  }
  double v = y[ns];
  // ns counts the tableau entries left of the path (1-based index minus one)
#pragma unroll
  for (int m = 1; m < 6; ++m) {
#pragma unroll
    for (int i = 0; i < 6 - m; ++i) {
      const double w = (c[i + 1] - d[i]) / (ho[i] - ho[i + m]);
      d[i] = ho[i + m] * w;
      c[i] = ho[i] * w;
    }
    double dy;
    if (2 * ns < 6 - m) {
      dy = c[ns];
    } else {
      dy = d[ns - 1];
      ns -= 1;
    }
    v = v + dy;
  }
  return v;
}

}  // namespace amss
