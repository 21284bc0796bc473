// Weighted-quadrature (WQ) formation of the elasticity stiffness on the GPU:
// the sum-factorised row assembly of wq_stiffness_layout
// (rust/src/wq.rs:329-514, interior or Calabro point layout, p = 2), written into a stencil
// matrix (every scalar control point couples to the 5 x 5 x 5 box of control
// points around it), then symmetrised in place, 0.5 (K + K^T), as the
// reference does with its mirrored triplets.
//
// Row of test function P = (i0, i1, i2), components (i, j):
//   K[(i,P),(j,Q)] = sum_{k1,k2} coef(i,j,k1,k2) R_{k1k2}[Q - P],
//   R_{k1k2} = (B_0^{t0} x B_1^{t1} x B_2^{t2}) s_loc,
// with s_loc the SIMP factor on the row's 9 x 9 x 9 WQ points (the row's
// points in each direction are contiguous from pstart, at most 9) and B_d^t the
// 5 x 9 per-direction tables of wq.cpp (t = 0: W00 B, 1: W10 B, 2: W01 B',
// 3: W11 B'), t_d = 3 on k1 = k2 = d, 1 on d = k1 != k2, 2 on d = k2 != k1,
// else 0; coef(i,j,k1,k2) = det / (L_k1 L_k2) (lambda d_{i k1} d_{j k2} +
// mu (d_ij d_{k1 k2} + d_{i k2} d_{j k1})).
//
// One wave per row: s_loc and the three 4 x 5 x 9 tables are staged in
// LDS; the first mode product is shared by the (k1, k2) pairs that use the
// same direction-0 table (four groups, one 405-value product held at a
// time), the second and third are done per (k1, k2), and each lane keeps the
// 9 component blocks of its 2 stencil entries in registers.
//
// Storage: S[(3 i + j) * 125 + o][P] with o = (dx+2) + 5 (dy+2) + 25 (dz+2),
// coalesced over P; entries whose column falls outside the grid are zero.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

constexpr int WQ_WAVES = 2;  // rows (waves) per workgroup
constexpr int WQ_TAB = 4 * 5 * 9;

struct WqCoef {
  double c[81];  // coef(i, j, k1, k2) at ((i * 3 + j) * 3 + k1) * 3 + k2
};

// simp factor at every WQ point (points x fastest; np_d points in direction d)
// kind 0: 1; kind 1: element densities x[nel], the element of a point from
// pbin (per direction, concatenated); kind 2: spline densities x[ndof_sc]
// evaluated with the per-direction point bases (first index pfirst, 3 values pval).
__global__ __launch_bounds__(256) void wq_simp_kernel(int kind, int np0, int np1, int np2, int nel0, int nel1,
                                                      int ncp0, int ncp1, const double* __restrict__ x, double penal,
                                                      double emin, const int* __restrict__ pbin,
                                                      const int* __restrict__ pfirst, const double* __restrict__ pval,
                                                      double* __restrict__ simp) {
  const int64_t npt = (int64_t)np0 * np1 * np2;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < npt; g += (int64_t)gridDim.x * blockDim.x) {
    const int p0 = (int)(g % np0), p1 = (int)((g / np0) % np1), p2 = (int)(g / ((int64_t)np0 * np1));
    double v = 1.0;
    if (kind == 1) {
      const double xe = x[pbin[p0] + nel0 * (pbin[np0 + p1] + nel1 * pbin[np0 + np1 + p2])];
      v = emin + pow(xe, penal) * (1.0 - emin);
    } else if (kind == 2) {
      const int q[3] = {p0, p1 + np0, p2 + np0 + np1};
      int f[3];
      double b[3][3];
      for (int d = 0; d < 3; ++d) {
        f[d] = pfirst[q[d]];
        for (int r = 0; r < 3; ++r) b[d][r] = pval[3 * q[d] + r];
      }
      double rho = 0.0;
      for (int c = 0; c < 3; ++c)
        for (int bb = 0; bb < 3; ++bb)
          for (int a = 0; a < 3; ++a)
            rho += b[0][a] * b[1][bb] * b[2][c] * x[(f[0] + a) + ncp0 * ((f[1] + bb) + ncp1 * (f[2] + c))];
      v = emin + pow(rho, penal) * (1.0 - emin);
    }
    simp[g] = v;
  }
}

// The (k1, k2) pairs grouped by the table of direction 0 they use (t0), so
// that only one t0's mode-0 product is held at a time: t0 = 3 for (0, 0);
// 0 for (1, 1), (2, 2), (1, 2), (2, 1); 1 for (0, 1), (0, 2); 2 for (1, 0),
// (2, 0).  Entries: k1 * 3 + k2, groups at WQ_GRP.
__constant__ int WQ_PAIRS[9] = {0, 4, 8, 5, 7, 1, 2, 3, 6};
__constant__ int WQ_GRP[5] = {0, 1, 5, 7, 9};
__constant__ int WQ_GRP_T0[4] = {3, 0, 1, 2};

__global__ __launch_bounds__(64 * WQ_WAVES) void wq_form_kernel(int ncp0, int ncp1, int ncp2, int np0, int np1,
                                                                int np2, const int* __restrict__ pstart,
                                                                const double* __restrict__ tab0,
                                                                const double* __restrict__ tab1,
                                                                const double* __restrict__ tab2,
                                                                const double* __restrict__ simp, WqCoef cf,
                                                                double* __restrict__ S) {
  // 15 KB of LDS per wave (one t0's mode-0 product at a time), so ten
  // waves fit on a CU
  __shared__ double sm[WQ_WAVES][729];
  __shared__ double m0[WQ_WAVES][405];
  __shared__ double tm[WQ_WAVES][225];
  __shared__ double tb[WQ_WAVES][3][WQ_TAB];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t nsc = (int64_t)ncp0 * ncp1 * ncp2;
  const int64_t P = (int64_t)blockIdx.x * WQ_WAVES + w;
  const bool live = P < nsc;
  const int i0 = live ? (int)(P % ncp0) : 0, i1 = live ? (int)((P / ncp0) % ncp1) : 0,
            i2 = live ? (int)(P / ((int64_t)ncp0 * ncp1)) : 0;
  // first WQ point of the row in each direction
  const int f0 = pstart[i0], f1 = pstart[ncp0 + i1], f2 = pstart[ncp0 + ncp1 + i2];
  for (int k = l; k < 729; k += 64) {
    const int q0 = k % 9, q1 = (k / 9) % 9, q2 = k / 81;
    const int p0 = f0 + q0, p1 = f1 + q1, p2 = f2 + q2;
    double v = 0.0;
    if (live && p0 < np0 && p1 < np1 && p2 < np2) v = simp[p0 + (int64_t)np0 * (p1 + (int64_t)np1 * p2)];
    sm[w][k] = v;
  }
  for (int k = l; k < WQ_TAB; k += 64) {
    tb[w][0][k] = tab0[(int64_t)i0 * WQ_TAB + k];
    tb[w][1][k] = tab1[(int64_t)i1 * WQ_TAB + k];
    tb[w][2][k] = tab2[(int64_t)i2 * WQ_TAB + k];
  }
  __syncthreads();
  double acc[9][2];
#pragma unroll
  for (int a = 0; a < 9; ++a) acc[a][0] = acc[a][1] = 0.0;
  for (int gi = 0; gi < 4; ++gi) {
    const int t0 = WQ_GRP_T0[gi];
    // mode 0: m0[jr0][q1][q2] = sum_q0 B0[t0][jr0][q0] s[q0][q1][q2]
    for (int k = l; k < 405; k += 64) {
      const int jr0 = k / 81, q12 = k - 81 * jr0;
      const double* B = &tb[w][0][(t0 * 5 + jr0) * 9];
      const double* sv = &sm[w][q12 * 9];
      double a = 0.0;
#pragma unroll
      for (int q0 = 0; q0 < 9; ++q0) a += B[q0] * sv[q0];
      m0[w][k] = a;
    }
    __syncthreads();
    for (int pi = WQ_GRP[gi]; pi < WQ_GRP[gi + 1]; ++pi) {
      const int k1 = WQ_PAIRS[pi] / 3, k2 = WQ_PAIRS[pi] % 3;
      int t[3];
      for (int d = 0; d < 3; ++d)
        t[d] = (k1 == k2) ? (d == k1 ? 3 : 0) : (d == k1 ? 1 : (d == k2 ? 2 : 0));
      // mode 1: tm[jr0][jr1][q2] = sum_q1 B1[t1][jr1][q1] m0[jr0][q1][q2]
      for (int k = l; k < 225; k += 64) {
        const int jr0 = k % 5, jr1 = (k / 5) % 5, q2 = k / 25;
        const double* B = &tb[w][1][(t[1] * 5 + jr1) * 9];
        const double* m = &m0[w][jr0 * 81 + q2 * 9];
        double a = 0.0;
#pragma unroll
        for (int q1 = 0; q1 < 9; ++q1) a += B[q1] * m[q1];
        tm[w][k] = a;
      }
      __syncthreads();
      // mode 2 for this lane's stencil entries o = l, l + 64
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int o = l + 64 * h;
        if (o < 125) {
          const int jr01 = o % 25, jr2 = o / 25;
          const double* B = &tb[w][2][(t[2] * 5 + jr2) * 9];
          double r = 0.0;
#pragma unroll
          for (int q2 = 0; q2 < 9; ++q2) r += B[q2] * tm[w][jr01 + 25 * q2];
#pragma unroll
          for (int ij = 0; ij < 9; ++ij) acc[ij][h] += cf.c[(ij * 3 + k1) * 3 + k2] * r;
        }
      }
      __syncthreads();
    }
  }
  if (!live) return;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int o = l + 64 * h;
    if (o < 125)
#pragma unroll
      for (int ij = 0; ij < 9; ++ij) S[((int64_t)ij * 125 + o) * nsc + P] = acc[ij][h];
  }
}

// In-place 0.5 (K + K^T): entry (ij, o, P) pairs with (ji, 124 - o, P + D(o));
// the lexicographically smaller of the two updates both.
__global__ __launch_bounds__(256) void wq_sym_kernel(int ncp0, int ncp1, int ncp2, double* __restrict__ S) {
  const int64_t nsc = (int64_t)ncp0 * ncp1 * ncp2;
  const int64_t total = 9 * 125 * nsc;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t P = g % nsc;
    const int so = (int)(g / nsc);
    const int ij = so / 125, o = so - 125 * ij;
    const int i = ij / 3, j = ij - 3 * i;
    const int ji = 3 * j + i, o2 = 124 - o;
    if (!(ij < ji || (ij == ji && o < o2))) continue;
    const int dx = o % 5 - 2, dy = (o / 5) % 5 - 2, dz = o / 25 - 2;
    const int x0 = (int)(P % ncp0), y0 = (int)((P / ncp0) % ncp1), z0 = (int)(P / ((int64_t)ncp0 * ncp1));
    const int x1 = x0 + dx, y1 = y0 + dy, z1 = z0 + dz;
    const double a = S[g];
    if (x1 < 0 || y1 < 0 || z1 < 0 || x1 >= ncp0 || y1 >= ncp1 || z1 >= ncp2) {
      S[g] = 0.5 * a;
      continue;
    }
    const int64_t Q = x1 + (int64_t)ncp0 * (y1 + (int64_t)ncp1 * z1);
    const int64_t g2 = ((int64_t)ji * 125 + o2) * nsc + Q;
    const double v = 0.5 * (a + S[g2]);
    S[g] = v;
    S[g2] = v;
  }
}

// y[i][P] = sum_j sum_o S[(3i+j)*125+o][P] x[j][P + D(o)]
__global__ __launch_bounds__(256) void wq_spmv_kernel(int ncp0, int ncp1, int ncp2, const double* __restrict__ S,
                                                      const double* __restrict__ x, double* __restrict__ y) {
  const int64_t nsc = (int64_t)ncp0 * ncp1 * ncp2;
  for (int64_t P = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; P < nsc; P += (int64_t)gridDim.x * blockDim.x) {
    const int x0 = (int)(P % ncp0), y0 = (int)((P / ncp0) % ncp1), z0 = (int)(P / ((int64_t)ncp0 * ncp1));
    double acc[3] = {0.0, 0.0, 0.0};
    for (int o = 0; o < 125; ++o) {
      const int x1 = x0 + o % 5 - 2, y1 = y0 + (o / 5) % 5 - 2, z1 = z0 + o / 25 - 2;
      if (x1 < 0 || y1 < 0 || z1 < 0 || x1 >= ncp0 || y1 >= ncp1 || z1 >= ncp2) continue;
      const int64_t Q = x1 + (int64_t)ncp0 * (y1 + (int64_t)ncp1 * z1);
      const double xv[3] = {x[Q], x[nsc + Q], x[2 * nsc + Q]};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[i] += S[((int64_t)(3 * i + j) * 125 + o) * nsc + P] * xv[j];
    }
    for (int i = 0; i < 3; ++i) y[i * nsc + P] = acc[i];
  }
}

static int grid_blocks(int64_t n) {
  int64_t nb = (n + 255) / 256;
  return (int)(nb > 65536 ? 65536 : (nb < 1 ? 1 : nb));
}

hipError_t launch_wq_simp(const icc_grid* g, int kind, const int np[3], const double* x, double penal, double emin,
                          const int* pbin, const int* pfirst, const double* pval, double* simp, hipStream_t st) {
  const auto& s = g->s;
  const int64_t npt = (int64_t)np[0] * np[1] * np[2];
  hipLaunchKernelGGL(wq_simp_kernel, dim3(grid_blocks(npt)), dim3(256), 0, st, kind, np[0], np[1], np[2],
                     (int)s.nel[0], (int)s.nel[1], (int)s.ncp[0], (int)s.ncp[1], x, penal, emin, pbin, pfirst, pval,
                     simp);
  return hipGetLastError();
}

hipError_t launch_wq_form(const icc_grid* g, const int np[3], const int* pstart, const double* tabs,
                          const double* simp, double* S, hipStream_t st) {
  const auto& s = g->s;
  WqCoef cf;
  const double det = s.len[0] * s.len[1] * s.len[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k1 = 0; k1 < 3; ++k1)
        for (int k2 = 0; k2 < 3; ++k2) {
          const double dl = (i == k1 && j == k2) ? 1.0 : 0.0;
          const double dm = ((i == j && k1 == k2) ? 1.0 : 0.0) + ((i == k2 && j == k1) ? 1.0 : 0.0);
          cf.c[((i * 3 + j) * 3 + k1) * 3 + k2] = det / (s.len[k1] * s.len[k2]) * (s.lambda * dl + s.mu * dm);
        }
  const double* t0 = tabs;
  const double* t1 = t0 + s.ncp[0] * WQ_TAB;
  const double* t2 = t1 + s.ncp[1] * WQ_TAB;
  const int64_t nsc = s.ndof_sc;
  const int nb = (int)((nsc + WQ_WAVES - 1) / WQ_WAVES);
  hipLaunchKernelGGL(wq_form_kernel, dim3(nb), dim3(64 * WQ_WAVES), 0, st, (int)s.ncp[0], (int)s.ncp[1],
                     (int)s.ncp[2], np[0], np[1], np[2], pstart, t0, t1, t2, simp, cf, S);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(wq_sym_kernel, dim3(grid_blocks(9 * 125 * nsc)), dim3(256), 0, st, (int)s.ncp[0],
                     (int)s.ncp[1], (int)s.ncp[2], S);
  return hipGetLastError();
}

hipError_t launch_wq_spmv(const icc_grid* g, const double* S, const double* x, double* y, hipStream_t st) {
  const auto& s = g->s;
  hipLaunchKernelGGL(wq_spmv_kernel, dim3(grid_blocks(s.ndof_sc)), dim3(256), 0, st, (int)s.ncp[0], (int)s.ncp[1],
                     (int)s.ncp[2], S, x, y);
  return hipGetLastError();
}

}  // namespace icc
