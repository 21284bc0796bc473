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

// WQ diagonal (see wq.cpp): gather over the <= 27 elements of each control
// point with the per-direction tables A0 = sum W00 B, A1 = sum W11 B'.
// tab layout: [A0x, A1x, A0y, A1y, A0z, A1z], each [ncp_d * 3].
__global__ __launch_bounds__(256) void wq_diag_kernel(const double* __restrict__ tab, int ncp0, int ncp1, int ncp2,
                                                      int nel0, int nel1, int nel2, const double* __restrict__ xdens,
                                                      double penal, double emin, double cf0, double cf1, double cf2,
                                                      double lam, double mu, double* __restrict__ d) {
  const int nsc = ncp0 * ncp1 * ncp2;
  const double* A0x = tab;
  const double* A1x = A0x + 3 * ncp0;
  const double* A0y = A1x + 3 * ncp0;
  const double* A1y = A0y + 3 * ncp1;
  const double* A0z = A1y + 3 * ncp1;
  const double* A1z = A0z + 3 * ncp2;
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < nsc; P += gridDim.x * blockDim.x) {
    const int pi = P % ncp0, pj = (P / ncp0) % ncp1, pk = P / (ncp0 * ncp1);
    const int fi = pi - 2 < 0 ? 0 : pi - 2, fj = pj - 2 < 0 ? 0 : pj - 2, fk = pk - 2 < 0 ? 0 : pk - 2;
    const int li = pi < nel0 - 1 ? pi : nel0 - 1, lj = pj < nel1 - 1 ? pj : nel1 - 1, lk = pk < nel2 - 1 ? pk : nel2 - 1;
    double T0 = 0.0, T1 = 0.0, T2 = 0.0;
    for (int ek = fk; ek <= lk; ++ek) {
      const double a0z = A0z[pk * 3 + ek - fk], a1z = A1z[pk * 3 + ek - fk];
      for (int ej = fj; ej <= lj; ++ej) {
        const double a0y = A0y[pj * 3 + ej - fj], a1y = A1y[pj * 3 + ej - fj];
        for (int ei = fi; ei <= li; ++ei) {
          const double a0x = A0x[pi * 3 + ei - fi], a1x = A1x[pi * 3 + ei - fi];
          double s = 1.0;
          if (xdens != nullptr) s = emin + pow(xdens[ei + nel0 * (ej + nel1 * ek)], penal) * (1.0 - emin);
          T0 += s * (a1x * a0y * a0z);
          T1 += s * (a0x * a1y * a0z);
          T2 += s * (a0x * a0y * a1z);
        }
      }
    }
    const double T[3] = {T0, T1, T2}, cf[3] = {cf0, cf1, cf2};
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int k = 0; k < 3; ++k) v += cf[k] * (mu + (c == k ? lam + mu : 0.0)) * T[k];
      d[c * nsc + P] = v;
    }
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

hipError_t launch_wq_diag(const icc_grid* g, const double* tab, const double* x, double penal, double emin,
                          double* d, hipStream_t st) {
  const auto& s = g->s;
  const double det = s.len[0] * s.len[1] * s.len[2];
  int nb = (int)((s.ndof_sc + 255) / 256);
  if (nb > 16384) nb = 16384;
  hipLaunchKernelGGL(wq_diag_kernel, dim3(nb), dim3(256), 0, st, tab, (int)s.ncp[0], (int)s.ncp[1], (int)s.ncp[2],
                     (int)s.nel[0], (int)s.nel[1], (int)s.nel[2], x, penal, emin, det / (s.len[0] * s.len[0]),
                     det / (s.len[1] * s.len[1]), det / (s.len[2] * s.len[2]), s.lambda, s.mu, d);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double* a, float* b, int64_t n, hipStream_t st) {
  int nb = vec_blocks(n);
  hipLaunchKernelGGL(to_f32_kernel, dim3(nb), dim3(256), 0, st, a, b, n);
  return hipGetLastError();
}

}  // namespace icc
