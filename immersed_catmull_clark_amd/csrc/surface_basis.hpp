// Quadratic B-spline basis at a point, shared by the host (scalar_basis_at)
// and the surface kernels: the span search of rust/src/iga/basis.rs:11-15 and
// the Cox-de Boor values of basis.rs:20-50 for degree 2, in the same
// operation order, so host and device give the same bits.
#pragma once
#include "icc_internal.hpp"

namespace icc {

// basis.rs:11-15 find_span: number of knots <= x, minus one, clamped to [p, ndof - 1]
ICC_HD inline int surf_span(const double* knots, int nk, double x) {
  int lo = 0, hi = nk;  // first index with knots[i] > x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (knots[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  int s = lo - 1;
  const int ndof = nk - 3;
  s = s < 2 ? 2 : s;
  s = s > ndof - 1 ? ndof - 1 : s;
  return s;
}

// basis.rs:20-50 for degree 2: values of functions s - 2 .. s at x
ICC_HD inline void surf_basis_p2(const double* knots, int s, double x, double v[3]) {
  double b1[2];
  {
    // k = 1: functions s - 1 (r = 0) and s (r = 1)
    double val0 = 0.0, val1 = 0.0;
    const double d = knots[s + 1] - knots[s];
    if (d > 0.0) {
      val0 += (knots[s + 1] - x) / d * 1.0;
      val1 += (x - knots[s]) / d * 1.0;
    }
    b1[0] = val0;
    b1[1] = val1;
  }
  for (int r = 0; r <= 2; ++r) {
    const int i = s + r - 2;
    double val = 0.0;
    if (r > 0) {
      const double d = knots[i + 2] - knots[i];
      if (d > 0.0) val += (x - knots[i]) / d * b1[r - 1];
    }
    if (r < 2) {
      const double d = knots[i + 3] - knots[i + 1];
      if (d > 0.0) val += (knots[i + 3] - x) / d * b1[r];
    }
    v[r] = val;
  }
}

}  // namespace icc
