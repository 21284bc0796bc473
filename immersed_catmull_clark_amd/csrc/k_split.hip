// Split-FP16 K.u apply for p = 2 trivariate B-spline grids on gfx950: the
// element contraction of the mixed-precision inner solve.
//
// Arithmetic.  K_t (scaled by a power of two per type so that its largest
// entry lies in [2^14, 2^15)) and the staged x (scaled the same way per tile) are
// each split into FP16 parts, v = hi + lo with hi = fp16(v) and
// lo = fp16(v - hi) (round to nearest even), and every contraction runs as
// three MFMA passes with FP32 accumulation:
//     acc += K_hi x_hi + K_hi x_lo + K_lo x_hi
// (v_mfma_f32_16x16x32_f16, and v_mfma_f32_16x16x16_f16 for the last 16
// trial columns).  Each part carries 11 significant bits, so the product has
// a relative error of about 2^-21 (the dropped K_lo x_lo term is 2^-22), near
// an FP32 contraction, at five times the FP32-MFMA rate.  The scale factors
// are folded into the element scale s_e when the accumulators are scaled.
//
// B operand without a gather.  The trial columns of an element are ordered
// by triplet (component c, ay, az) with four slots ax = 0..3 per triplet (slot
// 3 has zero K entries).  The staging writes, for every element x-position of
// every staged (c, y, z) row, the window of the four control points
// x .. x + 3 as 16 bytes (hi parts, then lo parts).  A lane's eight K values
// of one 32-deep step are two triplets of its element, so the B operand of a
// step is two 8-byte LDS reads per part, shared by three M tiles and three
// passes.  24 triplets fill three 32-deep steps; the last three take one
// 16-deep step (four slots per lane group).  Executed MFMA work per element:
// 96 rows x 112 columns x 3 passes.
//
// Work decomposition.  One workgroup of twelve waves per CU walks a
// contiguous run of tiles (sub-boxes of constant element type, up to 16
// x-lines of <= 16 elements, or 16-element groups of an x-boundary box):
//   * waves 0..7 (MFMA waves): wave w owns M tiles 3h .. 3h + 2 (h = w / 4)
//     of every element group g = w % 4 (mod 4) of the tile, with its 84 VGPRs
//     of A fragments (both parts) resident; scale, DPP combine along the
//     x-line (rows are permuted so that accumulator registers 0..2 hold the
//     ax = 0, 1, 2 rows of one triplet, as in the FP64 apply) and ds_add_f64
//     into the tile's LDS output box;
//   * waves 8..11 (memory waves): one tile ahead they load the points of
//     tile i + 2 and reduce their maximum, convert and write the windows of
//     tile i + 1, and flush tile i - 1's output box to y with global atomic
//     adds (masked), accumulating the p.Ap partial.
// One barrier per tile.  Windows and output boxes are double-buffered, the
// FP32 copy of the staged points (the p.Ap partial of the flush) is
// triple-buffered.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

typedef _Float16 sp_h8 __attribute__((ext_vector_type(8)));
typedef _Float16 sp_h4 __attribute__((ext_vector_type(4)));
typedef _Float16 sp_h2 __attribute__((ext_vector_type(2)));
typedef float sp_f4 __attribute__((ext_vector_type(4)));

// Exact quotient n / d for 0 <= n < 2^12 and 2 <= d < SP_MAGIC with
// m = ceil(2^32 / d): floor(n m / 2^32).
constexpr int SP_MAGIC = 1024;
struct SpMagic {
  uint32_t v[SP_MAGIC];
  constexpr SpMagic() : v() {
    for (uint32_t d = 2; d < (uint32_t)SP_MAGIC; ++d) v[d] = 0xffffffffu / d + 1u;
  }
};
static __constant__ SpMagic kSpMagic = SpMagic();
static_assert(3 * SP_MAX_PTS < 4096 && SP_MAX_PTS < SP_MAGIC, "tile indices fit the division table");
__device__ __forceinline__ int sp_div(int n, int d) {
  return d == 1 ? n : (int)__umulhi((uint32_t)n, kSpMagic.v[d]);
}

template <int K>
__device__ __forceinline__ float sp_shr(float v) {  // lane c of each 16-lane row gets lane c - K (0 for c < K)
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x110 + K, 0xf, 0xf, true));
}

__device__ __forceinline__ uint32_t sp_pack(_Float16 a, _Float16 b) {
  sp_h2 p = {a, b};
  return __builtin_bit_cast(uint32_t, p);
}

// Triplet t (0..26) of an element's 81 DOFs: component c = t / 9, then
// ay = (t % 9) % 3, az = (t % 9) / 3 (the triplet order of perm_row).
__device__ __forceinline__ void sp_triplet(int t, int& c, int& ay, int& az) {
  c = t / 9;
  const int rem = t - 9 * c;
  ay = rem % 3;
  az = rem / 3;
}

// Staging of one tile by the 256 memory threads: task = (row, quad), row =
// (c, lz, ly) of the staged box, quad q = windows 4q .. 4q + 3 of the row,
// which need points 4q .. 4q + 6.
struct SpLoad {
  float v[SP_TASK_ROUNDS][7];
  float sc;
};

__device__ __forceinline__ void sp_geom(const Tile& t, int& px, int& py, int& pz, int& nq, int& ntask) {
  px = t.nx + 2;
  py = t.ny + 2;
  pz = t.nz + 2;
  nq = (t.nx + 3) >> 2;
  ntask = nq * 3 * py * pz;
}

// Loads tile t's points (masked when `masked`) and element scales into L;
// returns the largest |value| this thread loaded.
__device__ __forceinline__ float sp_load(const Tile& t, bool masked, int tm, const float* __restrict__ x,
                                         const float* __restrict__ scale, const uint8_t* __restrict__ freem,
                                         int ncp0, int ncp1, int ndof_sc, int nel0, int nel1, SpLoad& L) {
  int px, py, pz, nq, ntask;
  sp_geom(t, px, py, pz, nq, ntask);
  float mx = 0.0f;
#pragma unroll
  for (int k = 0; k < SP_TASK_ROUNDS; ++k) {
    const int task = tm + 256 * k;
    if (task < ntask) {
      const int row = sp_div(task, nq), q = task - row * nq;
      const int r2 = sp_div(row, py), ly = row - r2 * py;
      const int c = sp_div(r2, pz), lz = r2 - c * pz;
      const int gb = (t.i0 + 4 * q) + ncp0 * ((t.j0 + ly) + ncp1 * (t.k0 + lz)) + c * ndof_sc;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        float v = 0.0f;
        if (4 * q + j < px) {
          v = x[gb + j];
          if (masked && freem[gb + j] == 0) v = 0.0f;
        }
        L.v[k][j] = v;
        mx = fmaxf(mx, fabsf(v));
      }
    }
  }
  const int nel = t.nx * t.ny * t.nz;
  L.sc = 0.0f;
  if (tm < nel) {
    const int r = sp_div(tm, t.nx), ex = tm - r * t.nx;
    const int ez = sp_div(r, t.ny), ey = r - ez * t.ny;
    L.sc = scale[(t.i0 + ex) + nel0 * ((t.j0 + ey) + nel1 * (t.k0 + ez))];
  }
  return mx;
}

// Power-of-two exponent e with max * 2^-e in [2^(SP_SCALE_BITS - 1),
// 2^SP_SCALE_BITS) (0 for max = 0): the scaled values sit at the top of the
// FP16 range, so the lo part of every value down to 2^-17 of the maximum is
// a normal FP16 number (near 1 the lo parts would be subnormal below 2^-3 of
// the maximum and carry an absolute error of 2^-25).
__device__ __forceinline__ int sp_exponent(uint32_t maxbits) {
  const int E = (int)((maxbits >> 23) & 0xffu);
  return maxbits == 0u ? 0 : E - 126 - SP_SCALE_BITS;
}

// Writes tile t's windows and FP32 point copy from L, scaled by 2^-ex.
__device__ __forceinline__ void sp_stage(const Tile& t, int tm, int ex, const SpLoad& L, uint4* __restrict__ win,
                                         float* __restrict__ xs, float* __restrict__ ss) {
  int px, py, pz, nq, ntask;
  sp_geom(t, px, py, pz, nq, ntask);
#pragma unroll
  for (int k = 0; k < SP_TASK_ROUNDS; ++k) {
    const int task = tm + 256 * k;
    if (task < ntask) {
      const int row = sp_div(task, nq), q = task - row * nq;
      _Float16 hi[7], lo[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const float v = __builtin_amdgcn_ldexpf(L.v[k][j], -ex);
        hi[j] = (_Float16)v;
        lo[j] = (_Float16)(v - (float)hi[j]);
      }
      uint32_t ph[6], pl[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        ph[j] = sp_pack(hi[j], hi[j + 1]);
        pl[j] = sp_pack(lo[j], lo[j + 1]);
      }
      uint4* wrow = win + row * t.nx + 4 * q;
#pragma unroll
      for (int w = 0; w < 4; ++w)
        if (4 * q + w < t.nx) wrow[w] = make_uint4(ph[w], ph[w + 2], pl[w], pl[w + 2]);
      float* xrow = xs + row * px + 4 * q;
      const int jmax = (q == nq - 1) ? 7 : 4;
#pragma unroll
      for (int j = 0; j < 7; ++j)
        if (j < jmax && 4 * q + j < px) xrow[j] = L.v[k][j];
    }
  }
  if (tm < t.nx * t.ny * t.nz) ss[tm] = L.sc;
}

// y += M sum_e s_e K_t(e) x_e (M x) in split FP16; optional per-workgroup
// partials of x^T M K M x into pap[blockIdx.x].  afrag: per type, half h,
// local M tile mm and K step ks (0..3), lane l: uint4 hi parts then uint4 lo
// parts (the 16-deep step uses the first 8 bytes of each).  kexp[type]: the
// power of two K_t was scaled by.
__global__ __launch_bounds__(SP_THREADS, 1) void apply_split_kernel(
    const Tile* __restrict__ tiles, const int* __restrict__ wg_first, const uint4* __restrict__ afrag,
    const int* __restrict__ kexp, const float* __restrict__ x, float* __restrict__ y,
    const float* __restrict__ scale, const uint8_t* __restrict__ freem, const int32_t* __restrict__ tmask,
    int ncp0, int ncp1, int ndof_sc, int nel0, int nel1, double* __restrict__ pap, const PcgState* st) {
  __shared__ uint4 win[2][SP_MAX_WIN];
  __shared__ float xs[3][3 * SP_MAX_PTS];
  __shared__ double ys[2][3 * SP_MAX_PTS];
  __shared__ float ss[2][SP_MAX_EL];
  __shared__ uint32_t pmax[2][4];
  __shared__ int sexp[2];
  __shared__ double red[4];

  if (st != nullptr && st->done != 0) return;  // solver already finished
  int first, last;
  {
    const int G = gridDim.x, b = blockIdx.x;
    const int c = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;  // runs of one XCD adjacent
    first = wg_first[c];
    last = wg_first[c + 1];
  }
  if (first >= last) {
    if (pap != nullptr && threadIdx.x == 0) pap[blockIdx.x] = 0.0;
    return;
  }
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const bool memw = wave >= 8;
  const int tm = threadIdx.x - 512;  // memory thread index (memory waves)
  const auto masked_at = [&](int i) { return tmask != nullptr && tmask[i] != 0; };

  // The two roles run separate loops with the same barrier sequence: two in
  // the prologue, one per tile, and one before the p.Ap reduction.
  if (memw) {
    SpLoad L;
    double pap_acc = 0.0;
    auto publish_max = [&](float mx, int slot) {
      for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
      if (lane == 0) pmax[slot][wave - 8] = __builtin_bit_cast(uint32_t, mx);
    };
    auto stage = [&](int i) {  // tile i from L into its buffers
      const Tile t = tiles[i];
      uint32_t m = pmax[i & 1][0];
      m = max(m, pmax[i & 1][1]);
      m = max(m, pmax[i & 1][2]);
      m = max(m, pmax[i & 1][3]);
      const int ex = sp_exponent(m);
      sp_stage(t, tm, ex, L, win[i & 1], xs[i % 3], ss[i & 1]);
      if (tm == 0) sexp[i & 1] = ex;
    };
    auto flush = [&](int i) {  // tile i's output box to y, cleared
      const Tile t = tiles[i];
      const bool masked = masked_at(i);
      const int px = t.nx + 2, py = t.ny + 2, pz = t.nz + 2;
      const int n3 = 3 * px * py * pz;
      double* yb = ys[i & 1];
      const float* xb = xs[i % 3];
      for (int idx = tm; idx < n3; idx += 256) {
        const double v = yb[idx];
        yb[idx] = 0.0;
        if (pap != nullptr) pap_acc += (double)xb[idx] * v;
        const int r = sp_div(idx, px), lx = idx - r * px;
        const int r2 = sp_div(r, py), ly = r - r2 * py;
        const int c = sp_div(r2, pz), lz = r2 - c * pz;
        const int g = (t.i0 + lx) + ncp0 * ((t.j0 + ly) + ncp1 * (t.k0 + lz)) + c * ndof_sc;
        if (!masked || freem[g] != 0) unsafeAtomicAdd(&y[g], (float)v);
      }
    };
    for (int idx = tm; idx < 3 * SP_MAX_PTS; idx += 256) {
      ys[0][idx] = 0.0;
      ys[1][idx] = 0.0;
    }
    publish_max(sp_load(tiles[first], masked_at(first), tm, x, scale, freem, ncp0, ncp1, ndof_sc, nel0, nel1, L),
                first & 1);
    __syncthreads();
    stage(first);
    if (first + 1 < last)
      publish_max(sp_load(tiles[first + 1], masked_at(first + 1), tm, x, scale, freem, ncp0, ncp1, ndof_sc, nel0,
                          nel1, L),
                  (first + 1) & 1);
    __syncthreads();
    for (int ti = first; ti < last; ++ti) {
      // flush tile ti - 1, stage ti + 1, load and reduce ti + 2
      if (ti > first) flush(ti - 1);
      if (ti + 1 < last) stage(ti + 1);
      if (ti + 2 < last)
        publish_max(sp_load(tiles[ti + 2], masked_at(ti + 2), tm, x, scale, freem, ncp0, ncp1, ndof_sc, nel0, nel1,
                            L),
                    (ti + 2) & 1);
      __syncthreads();
    }
    flush(last - 1);
    if (pap != nullptr) {
      for (int off = 32; off > 0; off >>= 1) pap_acc += __shfl_down(pap_acc, off, 64);
      if (lane == 0) red[wave - 8] = pap_acc;
    }
  } else {
    const int h = wave >> 2;  // M tiles 3h .. 3h + 2
    const int gs = wave & 3;  // element groups gs, gs + 4, ...
    const int col = lane & 15;
    const int kq = lane >> 4;
    sp_h8 ahi[3][3], alo[3][3];
    sp_h4 ahi3[3], alo3[3];
    int cur_type = -1;
    __syncthreads();
    __syncthreads();
    for (int ti = first; ti < last; ++ti) {
      const int b = ti & 1;
      const Tile t = tiles[ti];
      if (t.type != cur_type) {  // wave-uniform: (re)load this wave's A fragments
        const uint4* af = afrag + ((size_t)(t.type * 2 + h) * 3 * 4) * 128 + lane * 2;
#pragma unroll
        for (int mm = 0; mm < 3; ++mm) {
#pragma unroll
          for (int ks = 0; ks < 3; ++ks) {
            const uint4* p = af + (mm * 4 + ks) * 128;
            ahi[mm][ks] = __builtin_bit_cast(sp_h8, p[0]);
            alo[mm][ks] = __builtin_bit_cast(sp_h8, p[1]);
          }
          const uint4* p = af + (mm * 4 + 3) * 128;
          const uint4 a0 = p[0], a1 = p[1];
          ahi3[mm] = __builtin_bit_cast(sp_h4, make_uint2(a0.x, a0.y));
          alo3[mm] = __builtin_bit_cast(sp_h4, make_uint2(a1.x, a1.y));
        }
        cur_type = t.type;
      }
      const int nx = t.nx, px = nx + 2, py = t.ny + 2, pz = t.nz + 2;
      const int nel = nx * t.ny * t.nz;
      const bool lines = t.lines != 0;
      const int ngroups = lines ? t.ny * t.nz : (nel + 15) >> 4;
      const float mult = __builtin_amdgcn_ldexpf(1.0f, sexp[b] + kexp[t.type]);
      const int wplane = nx * py * pz;  // windows per component
      const int pplane = px * py * pz;  // points per component
      // window offsets of this lane's B triplets (K steps 0..2: 8 ks + 2 kq
      // and + 1; step 3: 24 + kq, kq = 3 reads triplet 26 against zero A)
      int woff[7];
#pragma unroll
      for (int ks = 0; ks < 3; ++ks)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          int c, ay, az;
          sp_triplet(8 * ks + 2 * kq + u, c, ay, az);
          woff[2 * ks + u] = c * wplane + nx * (ay + py * az);
        }
      {
        int c, ay, az;
        sp_triplet(kq < 3 ? 24 + kq : 26, c, ay, az);
        woff[6] = c * wplane + nx * (ay + py * az);
      }
      // output offsets of this lane's rows: triplet 4 (3h + mm) + kq, and
      // for h = 0 the fourth register's triplet 24 + kq (kq < 3)
      int roff[3], roff3 = 0;
#pragma unroll
      for (int mm = 0; mm < 3; ++mm) {
        int c, ay, az;
        sp_triplet(4 * (3 * h + mm) + kq, c, ay, az);
        roff[mm] = c * pplane + px * (ay + py * az);
      }
      if (h == 0 && kq < 3) {
        int c, ay, az;
        sp_triplet(24 + kq, c, ay, az);
        roff3 = c * pplane + px * (ay + py * az);
      }
      const uint4* wb = win[b];
      double* yb = ys[b];
      const float* sb = ss[b];
      for (int g = gs; g < ngroups; g += 4) {
        int wbase, pbase, eidx;
        bool valid;
        if (lines) {
          const int lz = sp_div(g, t.ny), ly = g - lz * t.ny;
          valid = col < nx;
          eidx = col + nx * g;
          wbase = valid ? col + nx * (ly + py * lz) : 0;
          pbase = px * (ly + py * lz);
        } else {
          const int e = 16 * g + col;
          valid = e < nel;
          int ex = 0, ey = 0, ez = 0;
          if (valid) {
            const int r = sp_div(e, nx);
            ex = e - r * nx;
            ez = sp_div(r, t.ny);
            ey = r - ez * t.ny;
          }
          eidx = valid ? e : 0;
          wbase = ex + nx * (ey + py * ez);
          pbase = ex + px * (ey + py * ez);
        }
        const float s = valid ? sb[eidx] * mult : 0.0f;
        sp_f4 acc[3];
#pragma unroll
        for (int mm = 0; mm < 3; ++mm) acc[mm] = sp_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) {
          const uint4 w1 = wb[wbase + woff[2 * ks]];
          const uint4 w2 = wb[wbase + woff[2 * ks + 1]];
          const sp_h8 bh = __builtin_bit_cast(sp_h8, make_uint4(w1.x, w1.y, w2.x, w2.y));
          const sp_h8 bl = __builtin_bit_cast(sp_h8, make_uint4(w1.z, w1.w, w2.z, w2.w));
#pragma unroll
          for (int mm = 0; mm < 3; ++mm)
            acc[mm] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[mm][ks], bh, acc[mm], 0, 0, 0);
#pragma unroll
          for (int mm = 0; mm < 3; ++mm)
            acc[mm] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[mm][ks], bl, acc[mm], 0, 0, 0);
#pragma unroll
          for (int mm = 0; mm < 3; ++mm)
            acc[mm] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo[mm][ks], bh, acc[mm], 0, 0, 0);
        }
        {
          const uint4 w3 = wb[wbase + woff[6]];
          const sp_h4 bh = __builtin_bit_cast(sp_h4, make_uint2(w3.x, w3.y));
          const sp_h4 bl = __builtin_bit_cast(sp_h4, make_uint2(w3.z, w3.w));
#pragma unroll
          for (int mm = 0; mm < 3; ++mm) acc[mm] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi3[mm], bh, acc[mm], 0, 0, 0);
#pragma unroll
          for (int mm = 0; mm < 3; ++mm) acc[mm] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi3[mm], bl, acc[mm], 0, 0, 0);
#pragma unroll
          for (int mm = 0; mm < 3; ++mm) acc[mm] = __builtin_amdgcn_mfma_f32_16x16x16f16(alo3[mm], bh, acc[mm], 0, 0, 0);
        }
        if (lines) {
          // point x = col of the line receives v0[col] + v1[col - 1] + v2[col - 2];
          // points 16 and 17 (lines of 15 and 16 elements) come from lane 15
#pragma unroll
          for (int mm = 0; mm < 3; ++mm) {
            const float v0 = acc[mm][0] * s, v1 = acc[mm][1] * s, v2 = acc[mm][2] * s;
            const float comb = v0 + sp_shr<1>(v1) + sp_shr<2>(v2);
            const float e16 = v1 + sp_shr<1>(v2);
            double* yl = yb + pbase + roff[mm];
            if (col <= nx + 1) unsafeAtomicAdd(yl + col, (double)comb);
            if (col == 15) {
              if (nx >= 15) unsafeAtomicAdd(yl + 16, (double)e16);
              if (nx >= 16) unsafeAtomicAdd(yl + 17, (double)v2);
            }
          }
          if (h == 0) {
            const float r0 = acc[0][3] * s, r1 = acc[1][3] * s, r2 = acc[2][3] * s;
            const float comb = r0 + sp_shr<1>(r1) + sp_shr<2>(r2);
            const float e16 = r1 + sp_shr<1>(r2);
            if (kq < 3) {
              double* yl = yb + pbase + roff3;
              if (col <= nx + 1) unsafeAtomicAdd(yl + col, (double)comb);
              if (col == 15) {
                if (nx >= 15) unsafeAtomicAdd(yl + 16, (double)e16);
                if (nx >= 16) unsafeAtomicAdd(yl + 17, (double)r2);
              }
            }
          }
        } else if (valid) {
#pragma unroll
          for (int mm = 0; mm < 3; ++mm) {
            double* ye = yb + pbase + roff[mm];
            unsafeAtomicAdd(ye + 0, (double)(acc[mm][0] * s));
            unsafeAtomicAdd(ye + 1, (double)(acc[mm][1] * s));
            unsafeAtomicAdd(ye + 2, (double)(acc[mm][2] * s));
          }
          if (h == 0 && kq < 3) {
            double* ye = yb + pbase + roff3;
#pragma unroll
            for (int mm = 0; mm < 3; ++mm) unsafeAtomicAdd(ye + mm, (double)(acc[mm][3] * s));
          }
        }
      }
      __syncthreads();
    }
  }
  if (pap != nullptr) {
    __syncthreads();
    if (threadIdx.x == 0) pap[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
  }
}

// Lane-layout check of the two FP16 MFMAs the split apply uses: A[i][k] =
// (7 i + 3 k) % 17 - 8, B[k][n] = (5 k + 11 n) % 13 - 6 (exact in FP16)
// loaded under the assumed layouts (16x16x32: lane l holds A[l%16][8 (l/16)
// + j] and B[8 (l/16) + j][l%16]; 16x16x16: A[l%16][4 (l/16) + j] and
// B[4 (l/16) + j][l%16]); out[4 l + r] and out[256 + 4 l + r] = accumulator
// register r, which should hold D[4 (l/16) + r][l%16].
__global__ void mfma_layout_probe_f16(double* out) {
  const int l = threadIdx.x;
  sp_h8 a, b;
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * (l / 16) + j, i = l % 16, n = l % 16;
    a[j] = (_Float16)(float)((7 * i + 3 * k) % 17 - 8);
    b[j] = (_Float16)(float)((5 * k + 11 * n) % 13 - 6);
  }
  sp_f4 acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
  sp_h4 a4, b4;
  for (int j = 0; j < 4; ++j) {
    const int k = 4 * (l / 16) + j, i = l % 16, n = l % 16;
    a4[j] = (_Float16)(float)((7 * i + 3 * k) % 17 - 8);
    b4[j] = (_Float16)(float)((5 * k + 11 * n) % 13 - 6);
  }
  sp_f4 acc4 = {0, 0, 0, 0};
  acc4 = __builtin_amdgcn_mfma_f32_16x16x16f16(a4, b4, acc4, 0, 0, 0);
  for (int r = 0; r < 4; ++r) {
    out[l * 4 + r] = (double)acc[r];
    out[256 + l * 4 + r] = (double)acc4[r];
  }
}

hipError_t launch_apply_split(const icc_grid* g, const icc_op* op, const float* x, float* y, double* pap,
                              const PcgState* pst, hipStream_t st) {
  const auto& s = g->s;
  hipLaunchKernelGGL(apply_split_kernel, dim3(g->split_G), dim3(SP_THREADS), 0, st, g->d_stiles, g->d_swgfirst,
                     g->d_afrag16, g->d_kexp16, x, y, op->d_scale32, op->d_free, op->d_stmask, (int)s.ncp[0],
                     (int)s.ncp[1], (int)s.ndof_sc, (int)s.nel[0], (int)s.nel[1], pap, pst);
  return hipGetLastError();
}

hipError_t launch_layout_probe_f16(double* out, hipStream_t st) {
  hipLaunchKernelGGL(mfma_layout_probe_f16, dim3(1), dim3(64), 0, st, out);
  return hipGetLastError();
}

}  // namespace icc
