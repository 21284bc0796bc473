// Design-update kernels of the immersed SIMP loop (rust/src/topopt_immersed.rs):
//
// * imm_filter_kernel: the sensitivity filter on the design domain
//   (topopt_immersed.rs:117-151, 213-215).  Offsets (di, dj, dk) with
//   dist = |(di h0, dj h1, dk h2)| <= r, weight r - dist, are tabulated on the
//   host in the reference's (dk, dj, di) order; only active in-range
//   neighbours of an active cell count, an inactive cell has an empty row
//   (filtered value 0, row sum 1), and a row sum below 1e-6 is replaced by 1.
//   Sums run in the table order with separate multiplies and adds (this file
//   is built with -ffp-contract=off), so for equal inputs the result is the
//   reference's bit for bit.
// * OC bisection on lambda with the CAD-weighted volume constraint
//   (topopt_immersed.rs:216-238), all on the device: one launch per
//   bisection step.  Step k reads the bisection state and the block partials
//   of sum_e w_e x_e(lambda) written by step k - 1 (double-buffered by the
//   parity of k), so every block takes the same decision from the same
//   ordered sum; block 0 publishes the new state.  Once the interval test
//   fails the state is marked done and later steps return at once; x_new then
//   holds the candidate of the last evaluated lambda, as in the reference.
// * stats_kernel: per-block partials of F.u, max |x - x_old| and sum w x
//   over the active cells.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

constexpr int TB = 256;

__device__ __forceinline__ double tb_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < TB / 64; ++i) s += sh[i];
  __syncthreads();
  return s;  // valid on thread 0
}

__device__ __forceinline__ double tb_max(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < TB / 64; ++i) s = fmax(s, sh[i]);
  __syncthreads();
  return s;  // valid on thread 0
}

// mode 0: out_e = sum_f w x_f v_f / (Hs_e max(x_e, 1e-3));  mode 1: out_e = Hs_e;
// mode 2: out_e = sum_f w v_f / Hs_e (density forward);
// mode 3: out_e = sum_f w v_f / Hs_f (its adjoint: the weights are symmetric)
__global__ __launch_bounds__(TB) void imm_filter_kernel(int mode, int n0, int n1, int n2,
                                                        const int32_t* __restrict__ off,
                                                        const double* __restrict__ ow, int noff,
                                                        const uint8_t* __restrict__ act,
                                                        const double* __restrict__ x, const double* __restrict__ v,
                                                        const double* __restrict__ hs, double* __restrict__ out) {
  const int nel = n0 * n1 * n2;
  for (int e = blockIdx.x * TB + threadIdx.x; e < nel; e += gridDim.x * TB) {
    double s = 0.0;
    if (act[e]) {
      const int i = e % n0, j = (e / n0) % n1, k = e / (n0 * n1);
      for (int o = 0; o < noff; ++o) {
        const int ni = i + off[3 * o], nj = j + off[3 * o + 1], nk = k + off[3 * o + 2];
        if (ni < 0 || nj < 0 || nk < 0 || ni >= n0 || nj >= n1 || nk >= n2) continue;
        const int f = ni + n0 * (nj + n1 * nk);
        if (!act[f]) continue;
        if (mode == 0)
          s += ow[o] * x[f] * v[f];
        else if (mode == 1)
          s += ow[o];
        else if (mode == 2)
          s += ow[o] * v[f];
        else
          s += ow[o] * (v[f] / hs[f]);
      }
    }
    if (mode == 0)
      out[e] = s / (hs[e] * fmax(x[e], 1e-3));
    else if (mode == 1)
      out[e] = s < 1e-6 ? 1.0 : s;
    else if (mode == 2)
      out[e] = s / hs[e];
    else
      out[e] = s;
  }
}

// x_new = active ? x_old : 0 (the candidate before the first bisection step)
// and per-block max |dc| over the active cells. act == nullptr: every cell is active.
__global__ __launch_bounds__(TB) void oc_init_kernel(int nel, const uint8_t* __restrict__ act,
                                                     const double* __restrict__ dc, const double* __restrict__ xold,
                                                     double* __restrict__ xnew, double* __restrict__ parts_max) {
  __shared__ double sh[TB / 64];
  double m = 0.0;
  for (int e = blockIdx.x * TB + threadIdx.x; e < nel; e += gridDim.x * TB) {
    const bool a = !act || act[e] != 0;
    xnew[e] = a ? xold[e] : 0.0;
    if (a) m = fmax(m, fabs(dc[e]));
  }
  m = tb_max(m, sh);
  if (threadIdx.x == 0) parts_max[blockIdx.x] = m;
}

// One bisection step k: update [l1, l2] from step k-1's volume, stop on the
// interval test, else x_new = clamp(x_old sqrt(-dc / (dvs lmid))) and the
// per-block volume sum of w x_new. act, w and dvs may be null (all active,
// unit weights, no divisor).
__global__ __launch_bounds__(TB) void oc_step_kernel(int k, int nel, const uint8_t* __restrict__ act,
                                                     const double* __restrict__ w, const double* __restrict__ dvs,
                                                     const double* __restrict__ dc,
                                                     const double* __restrict__ xold, double* __restrict__ xnew,
                                                     double target, double move,
                                                     const double* __restrict__ parts_max,
                                                     double* __restrict__ parts, OcState* __restrict__ st) {
  __shared__ double sh[TB / 64];
  __shared__ double s_lmid;
  __shared__ int s_go;
  const int G = gridDim.x;
  const double* pin = parts + (k & 1) * G;
  double* pout = parts + ((k + 1) & 1) * G;
  // every block derives the same reduction of the previous step's partials
  double red;
  if (k == 0) {
    double m = 0.0;
    for (int i = threadIdx.x; i < G; i += TB) m = fmax(m, parts_max[i]);
    red = tb_max(m, sh);
  } else {
    double s = 0.0;
    for (int i = threadIdx.x; i < G; i += TB) s += pin[i];
    red = tb_sum(s, sh);
  }
  if (threadIdx.x == 0) {
    OcState cur;
    int go = 1;
    if (k == 0) {
      cur.l1 = 0.0;
      cur.l2 = 2.0 * red;
      cur.done = 0;
      cur.steps = 0;
    } else {
      cur = st[k & 1];
      if (cur.done) {
        go = 0;
      } else {
        const double lmid = 0.5 * (cur.l1 + cur.l2);
        if (red > target)
          cur.l1 = lmid;
        else
          cur.l2 = lmid;
      }
    }
    if (go && !((cur.l2 - cur.l1) / (cur.l1 + cur.l2 + 1e-10) > 1e-4)) {
      cur.done = 1;
      go = 0;
    }
    if (go) cur.steps = k + 1;
    if (blockIdx.x == 0) st[(k + 1) & 1] = cur;
    s_go = go;
    s_lmid = 0.5 * (cur.l1 + cur.l2);
  }
  __syncthreads();
  if (!s_go) return;
  const double lmid = s_lmid;
  double acc = 0.0;
  for (int e = blockIdx.x * TB + threadIdx.x; e < nel; e += G * TB) {
    if (act && !act[e]) continue;
    const double xo = xold[e];
    const double be = sqrt(-dc[e] / (dvs ? dvs[e] * lmid : lmid));
    const double c = fmax(fmax(fmin(fmin(xo * be, xo + move), 1.0), xo - move), 0.001);
    xnew[e] = c;
    acc += w ? w[e] * c : c;
  }
  acc = tb_sum(acc, sh);
  if (threadIdx.x == 0) pout[blockIdx.x] = acc;
}

// parts[0][b] = sum f.u, parts[1][b] = max |x - x_old| (active), parts[2][b] = sum w x (active);
// act and w may be null (all active, unit weights)
__global__ __launch_bounds__(TB) void stats_kernel(int64_t n, const double* __restrict__ f,
                                                   const double* __restrict__ u, int nel,
                                                   const uint8_t* __restrict__ act, const double* __restrict__ w,
                                                   const double* __restrict__ x, const double* __restrict__ xold,
                                                   double* __restrict__ parts) {
  __shared__ double sh[TB / 64];
  const int G = gridDim.x;
  double c = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)TB + threadIdx.x; i < n; i += (int64_t)G * TB) c += f[i] * u[i];
  double ch = 0.0, vol = 0.0;
  for (int e = blockIdx.x * TB + threadIdx.x; e < nel; e += G * TB)
    if (!act || act[e]) {
      ch = fmax(ch, fabs(x[e] - xold[e]));
      vol += w ? w[e] * x[e] : x[e];
    }
  c = tb_sum(c, sh);
  ch = tb_max(ch, sh);
  vol = tb_sum(vol, sh);
  if (threadIdx.x == 0) {
    parts[blockIdx.x] = c;
    parts[G + blockIdx.x] = ch;
    parts[2 * G + blockIdx.x] = vol;
  }
}

__global__ __launch_bounds__(TB) void fill_active_kernel(int nel, const uint8_t* __restrict__ act, double v,
                                                         double* __restrict__ x) {
  for (int e = blockIdx.x * TB + threadIdx.x; e < nel; e += gridDim.x * TB) x[e] = act[e] ? v : 0.0;
}

int topopt_blocks(int64_t nel) {
  int64_t nb = (nel + TB * 4 - 1) / (TB * 4);
  if (nb < 1) nb = 1;
  return (int)(nb > VEC_MAX_BLOCKS ? VEC_MAX_BLOCKS : nb);
}

hipError_t launch_imm_filter(int mode, const int64_t nel[3], const int32_t* off, const double* ow, int noff,
                             const uint8_t* act, const double* x, const double* v, const double* hs, double* out,
                             hipStream_t st) {
  const int64_t ne = nel[0] * nel[1] * nel[2];
  int nb = (int)((ne + TB - 1) / TB);
  if (nb > 16384) nb = 16384;
  hipLaunchKernelGGL(imm_filter_kernel, dim3(nb), dim3(TB), 0, st, mode, (int)nel[0], (int)nel[1], (int)nel[2], off,
                     ow, noff, act, x, v, hs, out);
  return hipGetLastError();
}

hipError_t launch_oc_init(int64_t nel, const uint8_t* act, const double* dc, const double* xold, double* xnew,
                          double* parts_max, hipStream_t st) {
  hipLaunchKernelGGL(oc_init_kernel, dim3(topopt_blocks(nel)), dim3(TB), 0, st, (int)nel, act, dc, xold, xnew,
                     parts_max);
  return hipGetLastError();
}

hipError_t launch_oc_step(int k, int64_t nel, const uint8_t* act, const double* w, const double* dvs,
                          const double* dc, const double* xold, double* xnew, double target, double move,
                          const double* parts_max, double* parts, OcState* st, hipStream_t s) {
  hipLaunchKernelGGL(oc_step_kernel, dim3(topopt_blocks(nel)), dim3(TB), 0, s, k, (int)nel, act, w, dvs, dc, xold,
                     xnew, target, move, parts_max, parts, st);
  return hipGetLastError();
}

hipError_t launch_topopt_stats(int64_t n, const double* f, const double* u, int64_t nel, const uint8_t* act,
                               const double* w, const double* x, const double* xold, double* parts, hipStream_t st) {
  hipLaunchKernelGGL(stats_kernel, dim3(topopt_blocks(nel)), dim3(TB), 0, st, n, f, u, (int)nel, act, w, x, xold,
                     parts);
  return hipGetLastError();
}

hipError_t launch_fill_active(int64_t nel, const uint8_t* act, double v, double* x, hipStream_t st) {
  hipLaunchKernelGGL(fill_active_kernel, dim3(topopt_blocks(nel)), dim3(TB), 0, st, (int)nel, act, v, x);
  return hipGetLastError();
}

}  // namespace icc
