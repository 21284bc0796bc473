// Matrix-free cut-cell face stabilisation: the legacy "ghost penalty" term
// of rust/src/immersed.rs:168-222, applied without assembling it.
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

int face_blocks(const icc_op* op) {
  if (op->f_n <= 0) return 0;
  int64_t nb = (27 * op->f_n + FB - 1) / FB;
  return (int)(nb < FACE_MAX_BLOCKS ? nb : FACE_MAX_BLOCKS);
}

template <typename T>
hipError_t launch_faces(const icc_op* op, const T* x, T* y, double* pap, const PcgState* pst, hipStream_t st) {
  if (op->f_n <= 0) return hipSuccess;
  const auto& s = op->g->s;
  hipLaunchKernelGGL(legacy_faces_kernel<T>, dim3(face_blocks(op)), dim3(FB), 0, st, op->d_faces, (int)op->f_n,
                     (int)s.nel[0], (int)s.nel[1], (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc, op->f_alpha,
                     op->d_free, x, y, pap, pst);
  return hipGetLastError();
}

template hipError_t launch_faces<double>(const icc_op*, const double*, double*, double*, const PcgState*,
                                         hipStream_t);
template hipError_t launch_faces<float>(const icc_op*, const float*, float*, double*, const PcgState*, hipStream_t);

}  // namespace icc
