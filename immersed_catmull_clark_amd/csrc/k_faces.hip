// Matrix-free cut-cell face stabilisation: the legacy "ghost penalty" term
// of rust/src/immersed.rs:168-222 and the consistent ghost penalty of
// immersed.rs:234-370, applied without assembling either.
//
// Faces (e1, e2 = e1 + unit step along `dir`) between two active cells of
// which at least one is cut; each face couples the 4 x 3 x 3 slab of control
// points of its two cells (4 layers along dir), per displacement component:
//   * legacy: springs alpha between layer 0 and layer 3 at each of the 9
//     tangential points (the reference's block on the 18 shared points lands
//     on diagonal entries only and sums to zero in each, so it is left out);
//   * ghost: coef (j j^T) (x) M_a (x) M_b, j the jump of the second normal
//     derivative of the 4 layer functions across the face's break, M_a, M_b
//     the 1D element mass matrices of the two tangential directions:
//     T = j . X (the jump at each tangential point), W = M_a T M_b^T,
//     y += coef j (x) W, quadratic form coef T : W.
//
// Faces are grouped by 4 x 4 x 4 element blocks.  One workgroup takes one
// block: it stages the masked x of the block's (4+3)^3 control-point box in
// LDS, applies every face of the block with LDS double atomics into a box of
// y (one thread per face and component), and flushes the nonzero box
// entries to the free DOFs with one global atomic each (coalesced rows).  Global atomics drop from 18 (legacy) or 36
// (ghost) per face and component to about one per touched DOF and block,
// the cost that dominated the per-face formulation.  With pap != null each
// workgroup also writes its share of x_m^T S x_m, so the PCG's p.Ap needs no
// extra pass.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

constexpr int FB = 256;
constexpr int FBOX = FACE_BLK + 3;
constexpr int FBOX3 = FBOX * FBOX * FBOX;

struct FaceParams {
  const double* jump;  // ghost, per direction d: [nel_d][4], row k = break between elements k-1 and k
  const double* mass;  // ghost, per direction d: [nel_d][9]
  int joff[3], moff[3];
  double coef;  // legacy: alpha; ghost: gamma E mean(h)^3
};

template <typename T, int KIND>
__global__ __launch_bounds__(FB) void faces_blocked_kernel(const int32_t* __restrict__ faces,
                                                           const int32_t* __restrict__ fblk,
                                                           const int32_t* __restrict__ fptr, int nblk, int nbx,
                                                           int nby, int nel0, int nel1, int ncp0, int ncp1, int ncp2,
                                                           int nsc, FaceParams fp, const uint8_t* __restrict__ freem,
                                                           const T* __restrict__ x, T* __restrict__ y,
                                                           double* __restrict__ pap, const PcgState* pst) {
  if (pst != nullptr && pst->done) return;
  __shared__ double xs[3 * FBOX3];
  __shared__ double ys[3 * FBOX3];
  double q = 0.0;
  for (int bi = blockIdx.x; bi < nblk; bi += gridDim.x) {
    const uint32_t bc = (uint32_t)fblk[bi];  // packed block coordinates
    const int b0 = (int)(bc & 1023u) * FACE_BLK, b1 = (int)((bc >> 10) & 1023u) * FACE_BLK,
              b2 = (int)(bc >> 20) * FACE_BLK;
    // staging: all loads of a thread are issued before the LDS writes (the
    // x value and its mask bit are independent loads)
    constexpr int NST = (3 * FBOX3 + FB - 1) / FB;
    double sv[NST];
    bool sf[NST];
#pragma unroll
    for (int n = 0; n < NST; ++n) {
      const int l = threadIdx.x + n * FB;
      const int c = l / FBOX3, r = l - c * FBOX3;
      const int gx = b0 + r % FBOX, gy = b1 + (r / FBOX) % FBOX, gz = b2 + r / (FBOX * FBOX);
      double v = 0.0;
      bool f = false;
      if (l < 3 * FBOX3 && gx < ncp0 && gy < ncp1 && gz < ncp2) {
        const int64_t g = (int64_t)c * nsc + gx + (int64_t)ncp0 * (gy + (int64_t)ncp1 * gz);
        v = (double)x[g];
        f = freem == nullptr || freem[g];
      }
      sv[n] = v;
      sf[n] = f;
    }
#pragma unroll
    for (int n = 0; n < NST; ++n) {
      const int l = threadIdx.x + n * FB;
      if (l < 3 * FBOX3) {
        xs[l] = sf[n] ? sv[n] : 0.0;
        ys[l] = 0.0;
      }
    }
    __syncthreads();
    const int fb = fptr[bi], nfb = fptr[bi + 1] - fb;
    const int items = 3 * nfb;
    for (int it = threadIdx.x; it < items; it += FB) {
      const int c = (it >= nfb) + (it >= 2 * nfb), f = fb + (it - c * nfb);  // component major, face minor
      const uint32_t code = (uint32_t)faces[f];  // ex | ey << 10 | ez << 20 | dir << 30
      const int dir = (int)(code >> 30);
      const int e[3] = {(int)(code & 1023u), (int)((code >> 10) & 1023u), (int)((code >> 20) & 1023u)};
      const int l0 = e[0] - b0, l1 = e[1] - b1, l2 = e[2] - b2;
      const int sa = dir == 0 ? FBOX : 1;
      const int sb = dir == 2 ? FBOX : FBOX * FBOX;
      const int st = dir == 0 ? 1 : (dir == 1 ? FBOX : FBOX * FBOX);
      const int base = c * FBOX3 + l0 + FBOX * (l1 + FBOX * l2);
      if (KIND == FACE_LEGACY) {
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const int p0 = base + a * sa + b * sb, p3 = p0 + 3 * st;
            const double dx = xs[p0] - xs[p3];
            const double d = fp.coef * dx;
            atomicAdd(&ys[p0], d);
            atomicAdd(&ys[p3], -d);
            q += d * dx;
          }
      } else {
        const int ta = dir == 0 ? 1 : 0, tb = dir == 2 ? 1 : 2;
        const double* jv = fp.jump + fp.joff[dir] + 4 * (e[dir] + 1);
        const double* Ma = fp.mass + fp.moff[ta] + 9 * e[ta];
        const double* Mb = fp.mass + fp.moff[tb] + 9 * e[tb];
        double Tm[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) Tm[m] = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const double jt = jv[t];
#pragma unroll
          for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) Tm[3 * b + a] += jt * xs[base + t * st + a * sa + b * sb];
        }
        double U[9], W[9];
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int a = 0; a < 3; ++a)
            U[3 * b + a] = Ma[3 * a] * Tm[3 * b] + Ma[3 * a + 1] * Tm[3 * b + 1] + Ma[3 * a + 2] * Tm[3 * b + 2];
        double tw = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const double v = fp.coef * (Mb[3 * b] * U[a] + Mb[3 * b + 1] * U[3 + a] + Mb[3 * b + 2] * U[6 + a]);
            W[3 * b + a] = v;
            tw += Tm[3 * b + a] * v;
          }
        q += tw;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const double jt = jv[t];
#pragma unroll
          for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) atomicAdd(&ys[base + t * st + a * sa + b * sb], jt * W[3 * b + a]);
        }
      }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < 3 * FBOX3; l += FB) {
      const double v = ys[l];
      if (v == 0.0) continue;
      const int c = l / FBOX3, r = l - c * FBOX3;
      const int gx = b0 + r % FBOX, gy = b1 + (r / FBOX) % FBOX, gz = b2 + r / (FBOX * FBOX);
      if (gx >= ncp0 || gy >= ncp1 || gz >= ncp2) continue;
      const int64_t g = (int64_t)c * nsc + gx + (int64_t)ncp0 * (gy + (int64_t)ncp1 * gz);
      if (freem == nullptr || freem[g]) atomicAdd(y + g, (T)v);
    }
    __syncthreads();
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
  return op->f_nblk < FACE_MAX_BLOCKS ? (int)op->f_nblk : FACE_MAX_BLOCKS;
}

template <typename T>
hipError_t launch_faces(const icc_op* op, const T* x, T* y, double* pap, const PcgState* pst, hipStream_t st) {
  if (op->f_n <= 0) return hipSuccess;
  const auto& s = op->g->s;
  FaceParams fp;
  fp.jump = op->d_ftab;
  fp.mass = op->d_ftab ? op->d_ftab + op->ftab_mass : nullptr;
  for (int d = 0; d < 3; ++d) {
    fp.joff[d] = op->ftab_joff[d];
    fp.moff[d] = op->ftab_moff[d];
  }
  fp.coef = op->f_alpha;
  const int nbx = (int)((s.nel[0] + FACE_BLK - 1) / FACE_BLK), nby = (int)((s.nel[1] + FACE_BLK - 1) / FACE_BLK);
  if (op->f_kind == FACE_GHOST)
    hipLaunchKernelGGL((faces_blocked_kernel<T, FACE_GHOST>), dim3(face_blocks(op)), dim3(FB), 0, st, op->d_faces,
                       op->d_fblk, op->d_fptr, (int)op->f_nblk, nbx, nby, (int)s.nel[0], (int)s.nel[1],
                       (int)s.ncp[0], (int)s.ncp[1], (int)s.ncp[2], (int)s.ndof_sc, fp, op->d_free, x, y, pap, pst);
  else
    hipLaunchKernelGGL((faces_blocked_kernel<T, FACE_LEGACY>), dim3(face_blocks(op)), dim3(FB), 0, st, op->d_faces,
                       op->d_fblk, op->d_fptr, (int)op->f_nblk, nbx, nby, (int)s.nel[0], (int)s.nel[1],
                       (int)s.ncp[0], (int)s.ncp[1], (int)s.ncp[2], (int)s.ndof_sc, fp, op->d_free, x, y, pap, pst);
  return hipGetLastError();
}

template hipError_t launch_faces<double>(const icc_op*, const double*, double*, double*, const PcgState*,
                                         hipStream_t);
template hipError_t launch_faces<float>(const icc_op*, const float*, float*, double*, const PcgState*, hipStream_t);

}  // namespace icc
