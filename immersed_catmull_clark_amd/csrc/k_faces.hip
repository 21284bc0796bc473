// Matrix-free cut-cell face stabilisation: the legacy "ghost penalty" term
// of rust/src/immersed.rs:168-222 and the consistent ghost penalty of
// immersed.rs:234-370, applied without assembling either.
//
// For each face (e1, e2 = e1 + unit step along `dir`) between two active
// cells of which at least one is cut, and each displacement component, the
// reference adds
//   * a block on the 18 control points the two cells share whose triplets
//     all land on diagonal entries and sum to zero per entry (it contributes
//     nothing beyond round-off, as the reference itself documents), and
//   * springs of stiffness alpha = gamma * mean(h) between the 9 control
//     points owned by e1 only (layer 0 of the 4 x 3 x 3 union along dir) and
//     the 9 owned by e2 only (layer 3), paired in ascending global index:
//     both lists are then ordered by the same tangential index, so each
//     spring joins the two points of one line along dir.
// The operator therefore applies the springs: y_0 += alpha (x_0 - x_3),
// y_3 += alpha (x_3 - x_0) on masked values, scattered to free DOFs only.
// One thread per (face, component, tangential point): two gathers, two
// hardware FP atomics, and with pap != null the quadratic form
// alpha (x_0 - x_3)^2 reduced to one partial per block, so the PCG's p.Ap
// needs no extra pass.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

constexpr int FB = 256;

template <typename T>
__global__ __launch_bounds__(FB) void legacy_faces_kernel(const int32_t* __restrict__ faces, int nf, int nel0,
                                                          int nel1, int ncp0, int ncp1, int nsc, double alpha,
                                                          const uint8_t* __restrict__ freem,
                                                          const T* __restrict__ x, T* __restrict__ y,
                                                          double* __restrict__ pap, const PcgState* pst) {
  if (pst != nullptr && pst->done) return;
  double q = 0.0;
  const int items = 27 * nf;  // (component, tangential point) major, face minor
  for (int it = blockIdx.x * FB + threadIdx.x; it < items; it += gridDim.x * FB) {
    const int cm = it / nf, f = it - cm * nf;
    const int c = cm / 9, m = cm - 9 * c;
    const int a = m % 3, b = m / 3;
    const int code = faces[f];
    const int e1 = code >> 2, dir = code & 3;
    const int i = e1 % nel0, j = (e1 / nel0) % nel1, k = e1 / (nel0 * nel1);
    const int sy = ncp0, sz = ncp0 * ncp1;
    const int st = dir == 0 ? 1 : (dir == 1 ? sy : sz);
    const int sa = dir == 0 ? sy : 1;
    const int sb = dir == 2 ? sy : sz;
    const int64_t g0 = (int64_t)c * nsc + i + (int64_t)sy * j + (int64_t)sz * k + a * sa + b * sb;
    const int64_t g3 = g0 + 3 * st;
    const bool f0 = freem == nullptr || freem[g0] != 0;
    const bool f3 = freem == nullptr || freem[g3] != 0;
    const double x0 = f0 ? (double)x[g0] : 0.0;
    const double x3 = f3 ? (double)x[g3] : 0.0;
    const double d = alpha * (x0 - x3);
    if (f0) atomicAdd(y + g0, (T)d);
    if (f3) atomicAdd(y + g3, (T)(-d));
    q += d * (x0 - x3);
  }
  if (pap == nullptr) return;
  __shared__ double sh[FB / 64];
  for (int off = 32; off > 0; off >>= 1) q += __shfl_down(q, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = q;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < FB / 64; ++w) s += sh[w];
    pap[blockIdx.x] = s;
  }
}

// Consistent ghost penalty (rust/src/immersed.rs:234-370) on the same faces:
//   j(u, v) = gamma E h^3 sum_F int_F [d_n^2 u] . [d_n^2 v]
// whose face block is coef (j j^T) (x) M_a (x) M_b per component: j the jump
// of the second normal derivative of the 4 functions across the face's break
// (normal direction), M_a, M_b the 1D element mass matrices of the face's
// tangential element indices.  One thread per (face, component): gather the
// 4 x 3 x 3 masked slab X, contract T = j . X (the jump at each tangential
// point), W = M_a T M_b^T, scatter y += coef j (x) W to the free DOFs, and
// with pap != null add coef T : W to the block's p.Ap partial.
struct GhostTabs {
  const double* jump;  // per direction d: [nel_d][4], row k = break between elements k-1 and k
  const double* mass;  // per direction d: [nel_d][9]
  int joff[3], moff[3];
  double coef;
};

template <typename T>
__global__ __launch_bounds__(FB) void ghost_faces_kernel(const int32_t* __restrict__ faces, int nf, int nel0,
                                                         int nel1, int ncp0, int ncp1, int nsc, GhostTabs tb,
                                                         const uint8_t* __restrict__ freem,
                                                         const T* __restrict__ x, T* __restrict__ y,
                                                         double* __restrict__ pap, const PcgState* pst) {
  if (pst != nullptr && pst->done) return;
  double q = 0.0;
  const int items = 3 * nf;  // component major, face minor
  for (int it = blockIdx.x * FB + threadIdx.x; it < items; it += gridDim.x * FB) {
    const int c = it / nf, f = it - c * nf;
    const int code = faces[f];
    const int e1 = code >> 2, dir = code & 3;
    const int e[3] = {e1 % nel0, (e1 / nel0) % nel1, e1 / (nel0 * nel1)};
    const int sy = ncp0, sz = ncp0 * ncp1;
    const int ta = dir == 0 ? 1 : 0, tbd = dir == 2 ? 1 : 2;
    const int st = dir == 0 ? 1 : (dir == 1 ? sy : sz);
    const int sa = dir == 0 ? sy : 1;
    const int sb = dir == 2 ? sy : sz;
    const double* jv = tb.jump + tb.joff[dir] + 4 * (e[dir] + 1);
    const double* Ma = tb.mass + tb.moff[ta] + 9 * e[ta];
    const double* Mb = tb.mass + tb.moff[tbd] + 9 * e[tbd];
    const int64_t base = (int64_t)c * nsc + e[0] + (int64_t)sy * e[1] + (int64_t)sz * e[2];
    int off[9];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) off[3 * b + a] = a * sa + b * sb;
    double Tm[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) Tm[m] = 0.0;
    // layers are walked one at a time (not unrolled) to keep 9, not 36,
    // addresses live
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
      const double jt = jv[t];
      const int64_t lb = base + t * st;
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        const int64_t g = lb + off[m];
        const double xv = (freem == nullptr || freem[g]) ? (double)x[g] : 0.0;
        Tm[m] += jt * xv;
      }
    }
    // U = T M_a^T (rows b), then W = M_b U
    double U[9], W[9];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a)
        U[3 * b + a] = Ma[3 * a] * Tm[3 * b] + Ma[3 * a + 1] * Tm[3 * b + 1] + Ma[3 * a + 2] * Tm[3 * b + 2];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a)
        W[3 * b + a] = Mb[3 * b] * U[a] + Mb[3 * b + 1] * U[3 + a] + Mb[3 * b + 2] * U[6 + a];
    double tw = 0.0;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      W[m] *= tb.coef;
      tw += Tm[m] * W[m];
    }
    q += tw;
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
      const double jt = jv[t];
      const int64_t lb = base + t * st;
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        const int64_t g = lb + off[m];
        if (freem == nullptr || freem[g]) atomicAdd(y + g, (T)(jt * W[m]));
      }
    }
  }
  if (pap == nullptr) return;
  __shared__ double sh[FB / 64];
  for (int off = 32; off > 0; off >>= 1) q += __shfl_down(q, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = q;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < FB / 64; ++w) s += sh[w];
    pap[blockIdx.x] = s;
  }
}

int face_blocks(const icc_op* op) {
  if (op->f_n <= 0) return 0;
  const int64_t items = (op->f_kind == FACE_GHOST ? 3 : 27) * op->f_n;
  const int64_t nb = (items + FB - 1) / FB;
  return (int)(nb < FACE_MAX_BLOCKS ? nb : FACE_MAX_BLOCKS);
}

template <typename T>
hipError_t launch_faces(const icc_op* op, const T* x, T* y, double* pap, const PcgState* pst, hipStream_t st) {
  if (op->f_n <= 0) return hipSuccess;
  const auto& s = op->g->s;
  if (op->f_kind == FACE_GHOST) {
    GhostTabs tb;
    tb.jump = op->d_ftab;
    tb.mass = op->d_ftab + op->ftab_mass;
    for (int d = 0; d < 3; ++d) {
      tb.joff[d] = op->ftab_joff[d];
      tb.moff[d] = op->ftab_moff[d];
    }
    tb.coef = op->f_alpha;
    hipLaunchKernelGGL(ghost_faces_kernel<T>, dim3(face_blocks(op)), dim3(FB), 0, st, op->d_faces, (int)op->f_n,
                       (int)s.nel[0], (int)s.nel[1], (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc, tb, op->d_free,
                       x, y, pap, pst);
  } else {
    hipLaunchKernelGGL(legacy_faces_kernel<T>, dim3(face_blocks(op)), dim3(FB), 0, st, op->d_faces, (int)op->f_n,
                       (int)s.nel[0], (int)s.nel[1], (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc, op->f_alpha,
                       op->d_free, x, y, pap, pst);
  }
  return hipGetLastError();
}

template hipError_t launch_faces<double>(const icc_op*, const double*, double*, double*, const PcgState*,
                                         hipStream_t);
template hipError_t launch_faces<float>(const icc_op*, const float*, float*, double*, const PcgState*, hipStream_t);

}  // namespace icc
