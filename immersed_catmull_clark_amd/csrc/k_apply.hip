// K.u apply, element energies and Jacobi diagonal for p = 2 trivariate
// B-spline grids on gfx950.
//
// Work decomposition.  The element grid is cut into tiles of <= 128 elements
// that never straddle an element-type boundary (host: setup.cpp), so every
// element of a tile shares one 81x81 matrix K_t.  One workgroup = one tile,
// six waves, wave w owning the 16-row M tile w of K_t (rows 16w..16w+15; the
// sixth tile holds row 80 only).  A wave keeps its 21 A fragments (K rows x
// 4-deep K steps, pre-permuted on the host into lane order) in VGPRs and
// streams the tile's elements 16 at a time as the MFMA N dimension:
//     D[16 x 16] += A[16 x 4] * B[4 x 16],  B[k][j] = x_local(element j, trial k)
// with the B operand gathered from the LDS copy of the tile's control points
// (x masked by the free mask).  The 4 accumulator rows per lane are scaled by
// s_e and scatter-added into an LDS copy of the tile's output points
// (ds_add), and the tile is flushed to HBM with global atomic adds (y must be
// zero on entry).  The element energy p_e^T K p_e, needed for p.Ap, is
// accumulated on the fly from the same accumulators.
//
// MFMA lane layouts (v_mfma_{f64,f32}_16x16x4): lane l holds A[l%16][l/16] and
// B[l/16][l%16]; accumulator register r of lane l holds D[i][l%16] with
// i = 4 r + l/16 for f64 and i = 4 (l/16) + r for f32 (Mfma<T>::row).
// icc_debug_mfma_layout pins both on the device.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row (within the 16-row tile) held by accumulator register r of row group kq = lane / 16
  static __device__ __forceinline__ int row(int kq, int r) { return 4 * r + kq; }
};
template <>
struct Mfma<float> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kq, int r) { return 4 * kq + r; }
};

// Local offset of trial / test index b (0..80) inside a tile staged as
// [c][pz][py][px]: component block c, then control point (ax, ay, az).
__device__ __forceinline__ int local_offset(int b, int px, int py, int npts) {
  int c = b / NSC;
  int a = b - NSC * c;
  int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
  return c * npts + ax + px * (ay + py * az);
}

// Stages the tile's control points (all 3 components, masked) into xs and
// zeroes ys.  Returns the point count.
template <typename T>
__device__ __forceinline__ void stage_tile(const Tile& t, const T* __restrict__ x, const uint8_t* __restrict__ freem,
                                           int ncp0, int ncp1, int ndof_sc, T* xs, T* ys, int px, int py, int npts) {
  for (int idx = threadIdx.x; idx < 3 * npts; idx += blockDim.x) {
    int c = idx / npts;
    int p = idx - c * npts;
    int lx = p % px;
    int r = p / px;
    int ly = r % py;
    int lz = r / py;
    int g = (t.i0 + lx) + ncp0 * ((t.j0 + ly) + ncp1 * (t.k0 + lz)) + c * ndof_sc;
    T v = x[g];
    if (freem != nullptr && freem[g] == 0) v = T(0);
    xs[idx] = v;
    if (ys) ys[idx] = T(0);
  }
}

// Shift-by-k within 16-lane rows (lane c gets lane c - k, 0 for c < k).
template <int K>
__device__ __forceinline__ float row_shr(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x110 + K, 0xf, 0xf, true));
}
template <int K>
__device__ __forceinline__ double row_shr(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x110 + K, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + K, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// y += M * sum_e s_e K_t x_e (M x) over one tile per workgroup; optional
// per-tile partial of sum_e s_e x_e^T K x_e into pap[blockIdx.x].
//
// Each wave runs two element groups at a time as independent MFMA chains.
// In line mode a group is one x-line of the tile (lane column = x offset);
// accumulator registers 0..2 of a lane hold the ax = 0, 1, 2 rows of one
// (component, ay, az) triplet (perm_row), so the three contributions each
// control point of the line receives are summed in registers with two DPP
// row shifts before a single ds_add per point.
template <typename T>
__global__ __launch_bounds__(384) void apply_tiles(const Tile* __restrict__ tiles, const T* __restrict__ kfrag,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   const T* __restrict__ scale, const uint8_t* __restrict__ freem,
                                                   int ncp0, int ncp1, int ndof_sc, int nel0, int nel1,
                                                   double* __restrict__ pap, const PcgState* st) {
  __shared__ T xs[3 * TILE_MAX_PTS];
  __shared__ T ys[3 * TILE_MAX_PTS];
  __shared__ T ss[TILE_MAX_EL];
  __shared__ double red[6];
  if (st != nullptr && st->done != 0) return;  // solver already finished
  const Tile t = tiles[blockIdx.x];
  const int px = t.nx + 2, py = t.ny + 2, pz = t.nz + 2;
  const int npts = px * py * pz;
  const int nel = t.nx * t.ny * t.nz;
  stage_tile<T>(t, x, freem, ncp0, ncp1, ndof_sc, xs, ys, px, py, npts);
  for (int el = threadIdx.x; el < nel; el += blockDim.x) {
    const int ex = el % t.nx, rr = el / t.nx, ey = rr % t.ny, ez = rr / t.ny;
    ss[el] = scale[(t.i0 + ex) + nel0 * ((t.j0 + ey) + nel1 * (t.k0 + ez))];
  }

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int col = lane & 15;
  const int kq = lane >> 4;

  // A fragments of this wave's M tile, and the trial offsets of this lane
  T a[KS];
  int boff[KS];
  const T* kf = kfrag + ((size_t)(t.type * MT + wave) * KS) * 64 + lane;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    a[ks] = kf[ks * 64];
    int b = 4 * ks + kq;
    boff[ks] = b < NSH ? local_offset(b, px, py, npts) : 0;  // A is zero there
  }
  int roff[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = perm_row(wave, kq, r);
    roff[r] = row >= 0 ? local_offset(row, px, py, npts) : 0;
  }
  const bool r3 = perm_row(wave, kq, 3) >= 0;
  __syncthreads();

  double pap_acc = 0.0;
  const bool lines = t.lines != 0;
  const int ngroups = lines ? t.ny * t.nz : (nel + 15) / 16;
  for (int g0 = 0; g0 < ngroups; g0 += 2) {
    int eb[2], el[2];
    bool valid[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int g = g0 + q;
      if (lines) {
        const int ey = g % t.ny, ez = g / t.ny;
        valid[q] = g < ngroups && col < t.nx;
        el[q] = col + t.nx * g;
        eb[q] = valid[q] ? col + px * (ey + py * ez) : 0;
      } else {
        const int e = 16 * g + col;
        valid[q] = g < ngroups && e < nel;
        int ex = 0, ey = 0, ez = 0;
        if (valid[q]) {
          ex = e % t.nx;
          const int rr = e / t.nx;
          ey = rr % t.ny;
          ez = rr / t.ny;
        }
        el[q] = e;
        eb[q] = ex + px * (ey + py * ez);
      }
    }
    typename Mfma<T>::acc_t acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      acc0 = Mfma<T>::run(a[ks], xs[eb[0] + boff[ks]], acc0);
      acc1 = Mfma<T>::run(a[ks], xs[eb[1] + boff[ks]], acc1);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const typename Mfma<T>::acc_t& acc = q == 0 ? acc0 : acc1;
      const T s = valid[q] ? ss[el[q]] : T(0);
      const T v0 = acc[0] * s, v1 = acc[1] * s, v2 = acc[2] * s, v3 = acc[3] * s;
      if (pap != nullptr && valid[q]) {
        pap_acc += (double)xs[eb[q] + roff[0]] * (double)v0 + (double)xs[eb[q] + roff[1]] * (double)v1 +
                   (double)xs[eb[q] + roff[2]] * (double)v2;
        if (r3) pap_acc += (double)xs[eb[q] + roff[3]] * (double)v3;
      }
      if (lines) {
        // point x = col of the line receives v0[col] + v1[col-1] + v2[col-2];
        // the DPP shifts run on all lanes, before any lane-dependent branch
        const T comb = v0 + row_shr<1>(v1) + row_shr<2>(v2);
        const T e16 = v1 + row_shr<1>(v2);  // at lane 15: point 16 = v1[15] + v2[14]
        const int g = g0 + q;
        if (g < ngroups) {
          const int ey = g % t.ny, ez = g / t.ny;
          const int lb = px * (ey + py * ez) + roff[0];  // point x = 0 of this line, ax = 0 row
          if (col <= t.nx + 1) unsafeAtomicAdd(&ys[lb + col], comb);
          if (col == 15) {  // points 16 and 17 exist only for lines of 15 and 16 elements
            if (t.nx >= 15) unsafeAtomicAdd(&ys[lb + 16], e16);
            if (t.nx >= 16) unsafeAtomicAdd(&ys[lb + 17], v2);
          }
          if (r3 && valid[q]) unsafeAtomicAdd(&ys[eb[q] + roff[3]], v3);
        }
      } else if (valid[q]) {
        unsafeAtomicAdd(&ys[eb[q] + roff[0]], v0);
        unsafeAtomicAdd(&ys[eb[q] + roff[1]], v1);
        unsafeAtomicAdd(&ys[eb[q] + roff[2]], v2);
        if (r3) unsafeAtomicAdd(&ys[eb[q] + roff[3]], v3);
      }
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 3 * npts; idx += blockDim.x) {
    int c = idx / npts;
    int p = idx - c * npts;
    int lx = p % px;
    int r = p / px;
    int ly = r % py;
    int lz = r / py;
    int g = (t.i0 + lx) + ncp0 * ((t.j0 + ly) + ncp1 * (t.k0 + lz)) + c * ndof_sc;
    if (freem == nullptr || freem[g] != 0) unsafeAtomicAdd(&y[g], ys[idx]);
  }
  if (pap != nullptr) {
    // wave reduction, then across the six waves in fixed order
    for (int off = 32; off > 0; off >>= 1) pap_acc += __shfl_down(pap_acc, off, 64);
    if (lane == 0) red[wave] = pap_acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w = 0; w < 6; ++w) s += red[w];
      pap[blockIdx.x] = s;
    }
  }
}

// ee_e = u_e^T K_t u_e (unscaled) for every element of a tile; optionally the
// SIMP sensitivity dc_e = -penal (1 - emin) x_e^(penal-1) ee_e [* w_e].
__global__ __launch_bounds__(384) void energies_tiles(const Tile* __restrict__ tiles, const double* __restrict__ kfrag,
                                                      const double* __restrict__ u, int ncp0, int ncp1, int ndof_sc,
                                                      int nel0, int nel1, double* __restrict__ ee,
                                                      const double* __restrict__ xdens, double penal, double emin,
                                                      const double* __restrict__ wgt, double* __restrict__ dc) {
  __shared__ double xs[3 * TILE_MAX_PTS];
  __shared__ double eacc[TILE_MAX_EL];
  const Tile t = tiles[blockIdx.x];
  const int px = t.nx + 2, py = t.ny + 2, pz = t.nz + 2;
  const int npts = px * py * pz;
  const int nel = t.nx * t.ny * t.nz;
  stage_tile<double>(t, u, nullptr, ncp0, ncp1, ndof_sc, xs, nullptr, px, py, npts);
  for (int i = threadIdx.x; i < TILE_MAX_EL; i += blockDim.x) eacc[i] = 0.0;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int col = lane & 15;
  const int kq = lane >> 4;
  double a[KS];
  int boff[KS];
  const double* kf = kfrag + ((size_t)(t.type * MT + wave) * KS) * 64 + lane;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    a[ks] = kf[ks * 64];
    int b = 4 * ks + kq;
    boff[ks] = b < NSH ? local_offset(b, px, py, npts) : 0;
  }
  int roff[4];
  bool rvalid[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    int row = perm_row(wave, kq, r);
    rvalid[r] = row >= 0;
    roff[r] = rvalid[r] ? local_offset(row, px, py, npts) : 0;
  }
  __syncthreads();
  for (int g0 = 0; g0 < nel; g0 += 16) {
    const int el = g0 + col;
    const bool valid = el < nel;
    int ex = 0, ey = 0, ez = 0;
    if (valid) {
      ex = el % t.nx;
      int rr = el / t.nx;
      ey = rr % t.ny;
      ez = rr / t.ny;
    }
    const int ebase = ex + px * (ey + py * ez);
    d4_t acc = {0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks], xs[ebase + boff[ks]], acc, 0, 0, 0);
    double part = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (rvalid[r]) part += xs[ebase + roff[r]] * acc[r];
    // combine the four row groups of this column (lanes col, col+16, col+32, col+48)
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    if (valid && kq == 0) unsafeAtomicAdd(&eacc[el], part);
  }
  __syncthreads();
  for (int el = threadIdx.x; el < nel; el += blockDim.x) {
    int ex = el % t.nx;
    int rr = el / t.nx;
    int ey = rr % t.ny;
    int ez = rr / t.ny;
    const int ge = (t.i0 + ex) + nel0 * ((t.j0 + ey) + nel1 * (t.k0 + ez));
    const double e = eacc[el];
    if (ee) ee[ge] = e;
    if (dc) {
      double v = -penal * (1.0 - emin) * pow(xdens[ge], penal - 1.0) * e;
      if (wgt) v = -wgt[ge] * (penal * (1.0 - emin) * pow(xdens[ge], penal - 1.0) * e);
      dc[ge] = v;
    }
  }
}

// Jacobi diagonal by gathering over the <= 27 elements around each control
// point: d(c, P) = sum_{e ni P} s_e K_t(e)[(c,a), (c,a)], a = P - corner(e).
// Deterministic (fixed element order), no atomics.
template <typename T>
__global__ void diag_gather(const double* __restrict__ kdiag, const int32_t* __restrict__ tid1,
                            const int32_t* __restrict__ type_of, int maxnel, int T0, int T1,
                            const T* __restrict__ scale, int ncp0, int ncp1, int ncp2, int nel0, int nel1, int nel2,
                            T* __restrict__ d) {
  const int ndof_sc = ncp0 * ncp1 * ncp2;
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < ndof_sc; P += gridDim.x * blockDim.x) {
    const int pi = P % ncp0, pj = (P / ncp0) % ncp1, pk = P / (ncp0 * ncp1);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int az = 2; az >= 0; --az) {
      const int ek = pk - az;
      if (ek < 0 || ek >= nel2) continue;
      for (int ay = 2; ay >= 0; --ay) {
        const int ej = pj - ay;
        if (ej < 0 || ej >= nel1) continue;
        for (int ax = 2; ax >= 0; --ax) {
          const int ei = pi - ax;
          if (ei < 0 || ei >= nel0) continue;
          const int e = ei + nel0 * (ej + nel1 * ek);
          const int type = type_of[tid1[ei] + T0 * (tid1[maxnel + ej] + T1 * tid1[2 * maxnel + ek])];
          const double s = (double)scale[e];
          const int a = ax + 3 * ay + 9 * az;
          const double* kd = kdiag + type * NSH;
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[c] += s * kd[c * NSC + a];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c * ndof_sc + P] = (T)acc[c];
  }
}

// A[i][k] = i + 16 k, B[k][j] = (k == j % 4), so D[i][j] = i + 16 (j % 4):
// out[l * 4 + r] (f64) and out[256 + l * 4 + r] (f32) identify the row and
// column each accumulator register holds.
__global__ void mfma_layout_probe(double* out) {
  const int l = threadIdx.x;
  const double a = (double)((l % 16) + 16 * (l / 16));
  const double b = ((l / 16) == ((l % 16) % 4)) ? 1.0 : 0.0;
  d4_t acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  f4_t accf = {0, 0, 0, 0};
  accf = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, accf, 0, 0, 0);
  for (int r = 0; r < 4; ++r) {
    out[l * 4 + r] = acc[r];
    out[256 + l * 4 + r] = (double)accf[r];
  }
}

// ---- launchers ---------------------------------------------------------------
hipError_t launch_apply_f64(const icc_grid* g, const icc_op* op, const double* x, double* y, double* pap,
                            const PcgState* pst, hipStream_t st) {
  const auto& s = g->s;
  hipLaunchKernelGGL((apply_tiles<double>), dim3(g->ntiles), dim3(384), 0, st, g->d_tiles, g->d_kfrag64, x, y,
                     op->d_scale, op->d_free, (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc, (int)s.nel[0],
                     (int)s.nel[1], pap, pst);
  return hipGetLastError();
}

hipError_t launch_apply_f32(const icc_grid* g, const icc_op* op, const float* x, float* y, double* pap,
                            const PcgState* pst, hipStream_t st) {
  const auto& s = g->s;
  hipLaunchKernelGGL((apply_tiles<float>), dim3(g->ntiles), dim3(384), 0, st, g->d_tiles, g->d_kfrag32, x, y,
                     op->d_scale32, op->d_free, (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc, (int)s.nel[0],
                     (int)s.nel[1], pap, pst);
  return hipGetLastError();
}

hipError_t launch_energies(const icc_grid* g, const double* u, double* ee, const double* x, double penal,
                           double emin, const double* w, double* dc, hipStream_t st) {
  const auto& s = g->s;
  hipLaunchKernelGGL(energies_tiles, dim3(g->ntiles), dim3(384), 0, st, g->d_tiles, g->d_kfrag64, u, (int)s.ncp[0],
                     (int)s.ncp[1], (int)s.ndof_sc, (int)s.nel[0], (int)s.nel[1], ee, x, penal, emin, w, dc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_diag_t(const icc_grid* g, const T* scale, T* d, hipStream_t st) {
  const auto& s = g->s;
  int nb = (int)((s.ndof_sc + 255) / 256);
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL((diag_gather<T>), dim3(nb), dim3(256), 0, st, g->d_kdiag, g->d_tid1, g->d_typeof, g->maxnel,
                     g->T[0], g->T[1], scale, (int)s.ncp[0], (int)s.ncp[1], (int)s.ncp[2], (int)s.nel[0],
                     (int)s.nel[1], (int)s.nel[2], d);
  return hipGetLastError();
}
hipError_t launch_diag_f64(const icc_grid* g, const double* scale, double* d, hipStream_t st) {
  return launch_diag_t<double>(g, scale, d, st);
}
hipError_t launch_diag_f32(const icc_grid* g, const float* scale, float* d, hipStream_t st) {
  return launch_diag_t<float>(g, scale, d, st);
}

hipError_t launch_layout_probe(double* out, hipStream_t st) {
  hipLaunchKernelGGL(mfma_layout_probe, dim3(1), dim3(64), 0, st, out);
  return hipGetLastError();
}

}  // namespace icc
