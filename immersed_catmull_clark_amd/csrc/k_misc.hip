// Filters and small conversion kernels.
//
// Cartesian filter on element centres (topopt_iga.rs:100-126 semantics):
// r = rmin * mean(h); neighbour f of e iff d2 <= r^2 with d2 computed from the
// element centres (s + 0.5) h exactly as the reference does, weight
// max(r - sqrt(d2), 0); neighbours visited in increasing linear index, so the
// sums match the reference's row order.  This file is built with
// -ffp-contract=off so no multiply-add is fused.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

// mode 0: sensitivity filter  out_e = sum_f w x_f v_f / (Hs_e max(x_e, 1e-3))
// mode 1: density forward     out_e = sum_f w v_f / Hs_e
// mode 2: density backward    out_e = sum_f w (v_f / Hs_f)     (H symmetric)
// mode 3: row sums            out_e = Hs_e
__global__ __launch_bounds__(256) void filter_kernel(int n0, int n1, int n2, double h0, double h1, double h2, double r,
                                                     int R0, int R1, int R2, int mode, const double* __restrict__ x,
                                                     const double* __restrict__ v, const double* __restrict__ hs,
                                                     double* __restrict__ out) {
  const int nel = n0 * n1 * n2;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nel; e += gridDim.x * blockDim.x) {
    const int s0 = e % n0, s1 = (e / n0) % n1, s2 = e / (n0 * n1);
    const double c0 = ((double)s0 + 0.5) * h0, c1 = ((double)s1 + 0.5) * h1, c2 = ((double)s2 + 0.5) * h2;
    const double rr = r * r;
    double acc = 0.0, wsum = 0.0;
    const int k_lo = s2 - R2 < 0 ? 0 : s2 - R2, k_hi = s2 + R2 > n2 - 1 ? n2 - 1 : s2 + R2;
    const int j_lo = s1 - R1 < 0 ? 0 : s1 - R1, j_hi = s1 + R1 > n1 - 1 ? n1 - 1 : s1 + R1;
    const int i_lo = s0 - R0 < 0 ? 0 : s0 - R0, i_hi = s0 + R0 > n0 - 1 ? n0 - 1 : s0 + R0;
    for (int k = k_lo; k <= k_hi; ++k) {
      const double f2 = ((double)k + 0.5) * h2;
      const double dz = c2 - f2;
      for (int j = j_lo; j <= j_hi; ++j) {
        const double f1 = ((double)j + 0.5) * h1;
        const double dy = c1 - f1;
        for (int i = i_lo; i <= i_hi; ++i) {
          const double f0 = ((double)i + 0.5) * h0;
          const double dx = c0 - f0;
          const double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 <= rr) {
            const double w = fmax(r - sqrt(d2), 0.0);
            const int f = i + n0 * (j + n1 * k);
            if (mode == 0)
              acc += w * x[f] * v[f];
            else if (mode == 1)
              acc += w * v[f];
            else if (mode == 2)
              acc += w * (v[f] / hs[f]);
            wsum += w;
          }
        }
      }
    }
    if (mode == 0)
      out[e] = acc / (wsum * fmax(x[e], 1e-3));
    else if (mode == 1)
      out[e] = acc / wsum;
    else if (mode == 2)
      out[e] = acc;
    else
      out[e] = wsum;
  }
}

__global__ __launch_bounds__(256) void to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

hipError_t launch_filter(const icc_filter* f, int mode, const double* x, const double* v, double* out,
                         hipStream_t st) {
  const auto& s = f->g->s;
  const int nel = (int)s.nel_tot;
  int nb = (nel + 255) / 256;
  if (nb > 16384) nb = 16384;
  const double h0 = s.len[0] / (double)s.nel[0], h1 = s.len[1] / (double)s.nel[1], h2 = s.len[2] / (double)s.nel[2];
  hipLaunchKernelGGL(filter_kernel, dim3(nb), dim3(256), 0, st, (int)s.nel[0], (int)s.nel[1], (int)s.nel[2], h0, h1,
                     h2, f->r, f->rad[0], f->rad[1], f->rad[2], mode, x, v, f->d_hs, out);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double* a, float* b, int64_t n, hipStream_t st) {
  int nb = vec_blocks(n);
  hipLaunchKernelGGL(to_f32_kernel, dim3(nb), dim3(256), 0, st, a, b, n);
  return hipGetLastError();
}

}  // namespace icc
