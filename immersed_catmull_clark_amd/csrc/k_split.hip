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
// 3 has zero K entries).  The staged points are kept as FP16 hi and lo rows,
// one row of x-points per (c, y, z), padded to sp_pxp points (at least one
// zero past the row, for the last element's run), each row twice: as is and
// shifted by one point.  A lane's eight K values of one 32-deep step are two
// triplets of its element, i.e. two runs of 4 consecutive x-points; a run
// starting at an even point is a dword-aligned 8-byte read from the first
// copy, one starting at an odd point from the shifted copy.  So the B operand
// of a step is two aligned LDS reads per part, shared by three M tiles and
// three passes.  24 triplets fill three 32-deep steps; the last three take
// one 16-deep step (four slots per lane group).  Executed MFMA work per
// element: 96 rows x 112 columns x 3 passes.
//
// Work decomposition.  One workgroup of eight waves per CU walks a
// contiguous run of tiles (sub-boxes of constant element type, up to 16
// x-lines of <= 16 elements, or 16-element groups of an x-boundary box):
//   * waves 0..3 (MFMA waves, one per SIMD): all six M tiles of element
//     groups w, w + 4, ... of each tile, with the A fragments of both parts
//     resident; scale, DPP combine along the x-line (rows are permuted so
//     that accumulator registers 0..2 hold the ax = 0, 1, 2 rows of one
//     triplet, as in the FP64 apply) and ds_add_f64 into the tile's LDS
//     output box;
//   * waves 4..7 (memory waves): per tile i they issue the loads of tile
//     i + 3, convert and write the windows of tile i + 1, publish tile
//     i + 2's maximum, then flush tile i - 1's output box to y with global
//     atomic adds (masked), accumulating the p.Ap partial.  The atomics are
//     issued after every load the memory waves wait for in that tile, so
//     those waits never wait for the atomics.
// The MFMA waves load the first tile's A fragments while the memory waves
// stage it, and all eight waves flush the last tile.  One barrier per tile.  Windows and output boxes are double-buffered, the
// FP32 copy of the staged points (the p.Ap partial of the flush) and the
// row bases are triple-buffered.
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

// Staging of one tile by the 256 memory threads: task = (row, chunk), row =
// (c, lz, ly) of the staged box, chunk = 8 consecutive x-points of the row
// (rows padded to sp_pxp points, a multiple of 8 with at least one zero past
// the row) plus the next point, which the shifted copy needs.
struct SpLoad {
  float v[SP_TASK_ROUNDS][9];
  float sc;
};

__device__ __forceinline__ void sp_geom(const Tile& t, int& px, int& py, int& pz, int& pxp, int& ntask) {
  px = t.nx + 2;
  py = t.ny + 2;
  pz = t.nz + 2;
  pxp = sp_pxp(t.nx);
  ntask = (pxp >> 3) * 3 * py * pz;
}

// Issues the loads of tile t's points (masked when `masked`) and element
// scales into L; the values are consumed later (sp_absmax, sp_stage), so
// the loads stay in flight meanwhile.  Full chunks load as two 16-byte
// vectors (4-byte aligned: gfx950 supports unaligned global access).
__device__ __forceinline__ void sp_load(const Tile& t, bool masked, int tm, const float* __restrict__ x,
                                        const float* __restrict__ scale, const uint8_t* __restrict__ freem,
                                        int ncp0, int ncp1, int ndof_sc, int nel0, int nel1, SpLoad& L) {
  int px, py, pz, pxp, ntask;
  sp_geom(t, px, py, pz, pxp, ntask);
  const int nch = pxp >> 3;
#pragma unroll
  for (int k = 0; k < SP_TASK_ROUNDS; ++k) {
    const int task = tm + 256 * k;
    if (task < ntask) {
      const int row = sp_div(task, nch), ch = task - row * nch;
      const int r2 = sp_div(row, py), ly = row - r2 * py;
      const int c = sp_div(r2, pz), lz = r2 - c * pz;
      const int gb = (t.i0 + 8 * ch) + ncp0 * ((t.j0 + ly) + ncp1 * (t.k0 + lz)) + c * ndof_sc;
      if (8 * ch + 8 <= px) {
        float4 a, b;
        __builtin_memcpy(&a, x + gb, 16);
        __builtin_memcpy(&b, x + gb + 4, 16);
        L.v[k][0] = a.x;
        L.v[k][1] = a.y;
        L.v[k][2] = a.z;
        L.v[k][3] = a.w;
        L.v[k][4] = b.x;
        L.v[k][5] = b.y;
        L.v[k][6] = b.z;
        L.v[k][7] = b.w;
        L.v[k][8] = (8 * ch + 8 < px) ? x[gb + 8] : 0.0f;
      } else {
#pragma unroll
        for (int j = 0; j < 9; ++j) L.v[k][j] = (8 * ch + j < px) ? x[gb + j] : 0.0f;
      }
      if (masked)
#pragma unroll
        for (int j = 0; j < 9; ++j)
          if (8 * ch + j < px && freem[gb + j] == 0) L.v[k][j] = 0.0f;
    }
  }
  const int nel = t.nx * t.ny * t.nz;
  L.sc = 0.0f;
  if (tm < nel) {
    const int r = sp_div(tm, t.nx), ex = tm - r * t.nx;
    const int ez = sp_div(r, t.ny), ey = r - ez * t.ny;
    L.sc = scale[(t.i0 + ex) + nel0 * ((t.j0 + ey) + nel1 * (t.k0 + ez))];
  }
}

// Largest |value| of the tasks this thread loaded for tile t.
__device__ __forceinline__ float sp_absmax(const Tile& t, int tm, const SpLoad& L) {
  int px, py, pz, pxp, ntask;
  sp_geom(t, px, py, pz, pxp, ntask);
  float mx = 0.0f;
#pragma unroll
  for (int k = 0; k < SP_TASK_ROUNDS; ++k)
    if (tm + 256 * k < ntask)
#pragma unroll
      for (int j = 0; j < 8; ++j) mx = fmaxf(mx, fabsf(L.v[k][j]));  // point 8 belongs to the next chunk
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

// Writes tile t's FP16 hi and lo point rows (both copies, padded layout),
// FP32 point copy (unpadded) and global row bases from L, scaled by 2^-ex.
// Row r's points p, p + 1 (p even) are dword p / 2 of copy 0; points p + 1,
// p + 2 are dword p / 2 of copy 1.
__device__ __forceinline__ void sp_stage(const Tile& t, int tm, int ex, const SpLoad& L, uint4* __restrict__ h0,
                                         uint4* __restrict__ h1, uint4* __restrict__ l0, uint4* __restrict__ l1,
                                         float* __restrict__ xs, float* __restrict__ ss, int* __restrict__ rowg,
                                         int ncp0, int ncp1, int ndof_sc) {
  int px, py, pz, pxp, ntask;
  sp_geom(t, px, py, pz, pxp, ntask);
  const int nch = pxp >> 3;
#pragma unroll
  for (int k = 0; k < SP_TASK_ROUNDS; ++k) {
    const int task = tm + 256 * k;
    if (task < ntask) {
      const int row = sp_div(task, nch), ch = task - row * nch;
      _Float16 hi[9], lo[9];
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const float v = __builtin_amdgcn_ldexpf(L.v[k][j], -ex);
        hi[j] = (_Float16)v;
        lo[j] = (_Float16)(v - (float)hi[j]);
      }
      const int w = row * nch + ch;
      h0[w] = make_uint4(sp_pack(hi[0], hi[1]), sp_pack(hi[2], hi[3]), sp_pack(hi[4], hi[5]), sp_pack(hi[6], hi[7]));
      h1[w] = make_uint4(sp_pack(hi[1], hi[2]), sp_pack(hi[3], hi[4]), sp_pack(hi[5], hi[6]), sp_pack(hi[7], hi[8]));
      l0[w] = make_uint4(sp_pack(lo[0], lo[1]), sp_pack(lo[2], lo[3]), sp_pack(lo[4], lo[5]), sp_pack(lo[6], lo[7]));
      l1[w] = make_uint4(sp_pack(lo[1], lo[2]), sp_pack(lo[3], lo[4]), sp_pack(lo[5], lo[6]), sp_pack(lo[7], lo[8]));
      float* xrow = xs + row * px + 8 * ch;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (8 * ch + j < px) xrow[j] = L.v[k][j];
    }
  }
  const int nrows = 3 * py * pz;
  for (int row = tm; row < nrows; row += 256) {
    const int r2 = sp_div(row, py), ly = row - r2 * py;
    const int c = sp_div(r2, pz), lz = r2 - c * pz;
    rowg[row] = t.i0 + ncp0 * ((t.j0 + ly) + ncp1 * (t.k0 + lz)) + c * ndof_sc;
  }
  if (tm < t.nx * t.ny * t.nz) ss[tm] = L.sc;
}

// Row rotation within 16-lane rows: lane c gets lane (c - K) mod 16.
template <int K>
__device__ __forceinline__ float sp_ror(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x120 + K, 0xf, 0xf, true));
}

// Scatter of one accumulator set (registers 0..2 = ax 0, 1, 2 of one triplet
// per row group) of an x-line into the LDS output box: point x = col
// receives s[col] v0[col] + s[col-1] v1[col-1] + s[col-2] v2[col-2] (s1, s2:
// the element scales shifted like v1, v2); points 16 and 17 (lines of 15 and
// 16 elements) are s[15] v1[15] + s[14] v2[14] and s[15] v2[15], moved to
// lanes 0 and 1 by row rotations (r1 = rotated scale on lane 0 only, r2).
// yl: this lane's point; cmask: col <= nx + 1; emask: lanes 0, 1 of lines
// that reach points 16, 17.
__device__ __forceinline__ void sp_line_scatter(double* yl, bool cmask, bool emask, float s, float s1, float s2,
                                                float r1, float r2, float v0, float v1, float v2) {
  float comb = v0 * s;
  comb = fmaf(sp_shr<1>(v1), s1, comb);
  comb = fmaf(sp_shr<2>(v2), s2, comb);
  float ext = sp_ror<2>(v2) * r2;
  ext = fmaf(sp_ror<1>(v1), r1, ext);
  // the DPP moves read other lanes: they must run with every lane active,
  // so their results are pinned here, before the lane-dependent branches
  asm("" : "+v"(comb), "+v"(ext));
  if (cmask) unsafeAtomicAdd(yl, (double)comb);
  if (emask) unsafeAtomicAdd(yl + 16, (double)ext);
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
  // FP16 hi and lo parts of the staged points: [buffer][copy] (copy 1
  // shifted by one point)
  __shared__ uint4 hh[2][2][SP_MAX_HPTS / 8];
  __shared__ uint4 hl[2][2][SP_MAX_HPTS / 8];
  __shared__ float xs[3][3 * SP_MAX_PTS];
  __shared__ int rowg[3][SP_MAX_ROWS];
  // output boxes accumulate in FP64: ds_add_f64 runs at full rate, while
  // FP32 LDS atomics back up the LDS data FIFO on gfx950 and stall every
  // other LDS access of the workgroup
  __shared__ double ys[2][3 * SP_MAX_PTS];
  __shared__ float ss[2][SP_MAX_EL];
  __shared__ uint32_t pmax[2][4];
  __shared__ int sexp[2];
  __shared__ double red[SP_THREADS / 64];
  __shared__ double sdummy[4][64];

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
  const bool memw = wave >= 4;
  const int tm = threadIdx.x - 256;  // memory thread index (memory waves)
  const auto masked_at = [&](int i) { return tmask != nullptr && tmask[i] != 0; };

  // Flush of tile i's output box to y (masked) with global atomic adds,
  // clearing it, by the memory waves (thread tf = tm); the p.Ap partial is
  // the sum over the staged points of (M x)_i times the tile's contribution
  // to (A x)_i, which summed over tiles is x^T M K M x.
  double pap_acc = 0.0;
  const auto flush = [&](int i, int t0, int nt) {  // points t0, t0 + nt, ... of tile i
    const Tile t = tiles[i];
    const bool masked = masked_at(i);
    const int px = t.nx + 2;
    const int n3 = 3 * px * (t.ny + 2) * (t.nz + 2);
    double* yb = ys[i & 1];
    const float* xb = xs[i % 3];
    const int* rb = rowg[i % 3];
    for (int idx = t0; idx < n3; idx += nt) {
      const double v = yb[idx];
      const float xv = xb[idx];
      const int r = sp_div(idx, px);
      const int g = rb[r] + (idx - r * px);
      yb[idx] = 0.0;
      if (pap != nullptr) pap_acc += (double)xv * v;
      if (!masked || freem[g] != 0) unsafeAtomicAdd(&y[g], (float)v);
    }
  };
  // The two roles run separate loops with the same barrier sequence: two in
  // the prologue, one per tile, and one before the p.Ap reduction.
  if (memw) {
    // Three register sets rotate: at the top of tile i one holds tile i + 1
    // (its maximum published), one tile i + 2 (loaded during tile i - 1) and
    // one is free; the loads of tile i + 3 go into the free set first.  So
    // every load has a whole tile's work before its first use (the maximum,
    // one tile later).
    SpLoad La, Lb, Lc;
    auto publish_max = [&](float mx, int slot) {
      for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
      if (lane == 0) pmax[slot][wave - 4] = __builtin_bit_cast(uint32_t, mx);
    };
    auto load = [&](int i, SpLoad& L) {
      sp_load(tiles[i], masked_at(i), tm, x, scale, freem, ncp0, ncp1, ndof_sc, nel0, nel1, L);
    };
    auto stage = [&](int i, const SpLoad& L) {  // tile i from L into its buffers
      const Tile t = tiles[i];
      uint32_t m = pmax[i & 1][0];
      m = max(m, pmax[i & 1][1]);
      m = max(m, pmax[i & 1][2]);
      m = max(m, pmax[i & 1][3]);
      const int ex = sp_exponent(m);
      sp_stage(t, tm, ex, L, hh[i & 1][0], hh[i & 1][1], hl[i & 1][0], hl[i & 1][1], xs[i % 3], ss[i & 1],
               rowg[i % 3], ncp0, ncp1, ndof_sc);
      if (tm == 0) sexp[i & 1] = ex + kexp[t.type];
    };
    // tile i: load i + 3 into Ll, stage i + 1 from Ls, publish i + 2's
    // maximum from Lp, then flush tile i - 1 (its output box is complete
    // since the barrier that ended tile i - 1, and the MFMA waves write it
    // again only after the barrier that ends tile i).  The flush's atomics
    // are issued after every load this body waits for; the next waits (a
    // barrier later) find them long retired.
    auto body = [&](int i, const SpLoad& Ls, SpLoad& Lp, SpLoad& Ll) {
      if (i + 3 < last) load(i + 3, Ll);
      if (i + 1 < last) stage(i + 1, Ls);
      if (i + 2 < last) publish_max(sp_absmax(tiles[i + 2], tm, Lp), (i + 2) & 1);
      if (i > first) flush(i - 1, tm, 256);
      __syncthreads();
    };
    for (int idx = tm; idx < 3 * SP_MAX_PTS; idx += 256) {
      ys[0][idx] = 0.0;
      ys[1][idx] = 0.0;
    }
    load(first, La);
    if (first + 1 < last) load(first + 1, Lb);
    if (first + 2 < last) load(first + 2, Lc);
    publish_max(sp_absmax(tiles[first], tm, La), first & 1);
    __syncthreads();
    stage(first, La);
    if (first + 1 < last) publish_max(sp_absmax(tiles[first + 1], tm, Lb), (first + 1) & 1);
    __syncthreads();
    // set of tile k: La, Lb, Lc for (k - first) mod 3 = 0, 1, 2
    for (int ti = first; ti < last; ti += 3) {
      body(ti, Lb, Lc, La);
      if (ti + 1 < last) body(ti + 1, Lc, La, Lb);
      if (ti + 2 < last) body(ti + 2, La, Lb, Lc);
    }
  } else {
    // MFMA waves 0..3, one per SIMD: all six M tiles of element groups
    // w, w + 4, ... of each tile.
    const int gs = wave;
    const int col = lane & 15;
    const int kq0 = lane >> 4;
    double* dummy = &sdummy[wave][lane];  // target of the lanes whose output point does not exist
    sp_h8 ahi[6][3], alo[6][3];
    sp_h4 ahi3[6], alo3[6];
    // (re)loads the A fragments of all six M tiles of type `type`
    // (wave-uniform); they are consumed right here, so the loads are waited
    // for on this (rare) path and the group loop carries no vector-memory
    // waits
    const auto load_a = [&](int type) {
#pragma unroll
      for (int mt = 0; mt < 6; ++mt) {
        const uint4* af = afrag + ((size_t)(type * 2 + mt / 3) * 12 + (mt % 3) * 4) * 128 + lane * 2;
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) {
          ahi[mt][ks] = __builtin_bit_cast(sp_h8, af[ks * 128]);
          alo[mt][ks] = __builtin_bit_cast(sp_h8, af[ks * 128 + 1]);
        }
        const uint4 a0 = af[3 * 128], a1 = af[3 * 128 + 1];
        ahi3[mt] = __builtin_bit_cast(sp_h4, make_uint2(a0.x, a0.y));
        alo3[mt] = __builtin_bit_cast(sp_h4, make_uint2(a1.x, a1.y));
      }
#pragma unroll
      for (int mt = 0; mt < 6; ++mt) {
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) asm volatile("" ::"v"(ahi[mt][ks]), "v"(alo[mt][ks]));
        asm volatile("" ::"v"(ahi3[mt]), "v"(alo3[mt]));
      }
    };
    // the first tile's fragments load while the memory waves stage it
    int cur_type = tiles[first].type;
    load_a(cur_type);
    __syncthreads();
    __syncthreads();
    for (int ti = first; ti < last; ++ti) {
      const int b = ti & 1;
      const Tile t = tiles[ti];
      if (t.type != cur_type) {
        load_a(t.type);
        cur_type = t.type;
      }
      const int nx = t.nx, px = nx + 2, py = t.ny + 2, pz = t.nz + 2;
      const int nel = nx * t.ny * t.nz;
      const bool lines = t.lines != 0;
      const int ngroups = lines ? t.ny * t.nz : (nel + 15) >> 4;
      const float mult = __builtin_amdgcn_ldexpf(1.0f, sexp[b]);
      // the row-group index, laundered so that the triplet decomposition
      // below is recomputed per tile instead of being held (and spilled)
      // across the tile loop
      int kq = kq0;
      asm volatile("" : "+v"(kq));
      const int pxp = sp_pxp(nx);        // padded row length of the FP16 parts
      const int wplane = pxp * py * pz;  // padded points per component
      const int pplane = px * py * pz;   // points per component
      // point offsets (halved: dwords of a copy) of this lane's B triplets
      // (K steps 0..2: 8 ks + 2 kq and + 1; step 3: 24 + kq, kq = 3 reads
      // triplet 26 against zero A)
      int woff[7];
#pragma unroll
      for (int ks = 0; ks < 3; ++ks)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          int c, ay, az;
          sp_triplet(8 * ks + 2 * kq + u, c, ay, az);
          woff[2 * ks + u] = (c * wplane + pxp * (ay + py * az)) >> 1;
        }
      {
        int c, ay, az;
        sp_triplet(kq < 3 ? 24 + kq : 26, c, ay, az);
        woff[6] = (c * wplane + pxp * (ay + py * az)) >> 1;
      }
      // output offsets of this lane's rows: triplet 4 mt + kq of M tile mt,
      // and the fourth registers' triplet 24 + kq (kq < 3)
      int roff[6], roff3 = -1;
#pragma unroll
      for (int mt = 0; mt < 6; ++mt) {
        int c, ay, az;
        sp_triplet(4 * mt + kq, c, ay, az);
        roff[mt] = c * pplane + px * (ay + py * az);
      }
      if (kq < 3) {
        int c, ay, az;
        sp_triplet(24 + kq, c, ay, az);
        roff3 = c * pplane + px * (ay + py * az);
      }
      const uint32_t* whb0 = reinterpret_cast<const uint32_t*>(hh[b][0]);
      const uint32_t* whb1 = reinterpret_cast<const uint32_t*>(hh[b][1]);
      const uint32_t* wlb0 = reinterpret_cast<const uint32_t*>(hl[b][0]);
      const uint32_t* wlb1 = reinterpret_cast<const uint32_t*>(hl[b][1]);
      double* yb = ys[b];
      const float* sb = ss[b];
      const bool cmask = col <= nx + 1, emask = col < 2 && nx >= 15 + col;
      // the triplet offsets are even (rows are padded to a multiple of 8), so
      // a group's runs all start with the parity of its element base: the
      // row copy is picked once per group, and a run at point wbase + 2 o is
      // dwords (wbase >> 1) + o .. + 1 of that copy
      auto rd4 = [&](const uint32_t* w, int o) { return make_uint2(w[o], w[o + 1]); };
      auto rd8 = [&](const uint32_t* w, int o1, int o2) {  // two runs as one 8-half operand
        const uint2 a = rd4(w, o1), c = rd4(w, o2);
        return __builtin_bit_cast(sp_h8, make_uint4(a.x, a.y, c.x, c.y));
      };
      for (int g = gs; g < ngroups; g += 4) {
        // element of group g seen by this lane: point-row base, output base,
        // scale index
        int wbase, pb, eidx;
        bool valid;
        if (lines) {
          const int lz = sp_div(g, t.ny), ly = g - lz * t.ny;
          valid = col < nx;
          eidx = col + nx * g;
          wbase = valid ? col + pxp * (ly + py * lz) : 0;
          pb = px * (ly + py * lz);
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
          wbase = ex + pxp * (ey + py * ez);
          pb = ex + px * (ey + py * ez);
        }
        const float s = valid ? sb[eidx] * mult : 0.0f;
        sp_f4 acc[6];
#pragma unroll
        for (int mt = 0; mt < 6; ++mt) acc[mt] = sp_f4{0.0f, 0.0f, 0.0f, 0.0f};
        // M tiles [m0, m0 + 3): three passes per K step, the three tiles'
        // chains interleaved; B is streamed from LDS one K step ahead (and
        // read again for the second half), which keeps 16 of its registers
        // live instead of 28, and work(ks) is issued behind step ks's MFMAs
        auto mfma_half = [&](int m0, auto&& work) {
          int o = wbase >> 1;
          asm volatile("" : "+v"(o));  // no reuse of the first half's B reads
          const uint32_t* wh = ((wbase & 1) ? whb1 : whb0) + o;
          const uint32_t* wl = ((wbase & 1) ? wlb1 : wlb0) + o;
          sp_h8 bh[2], bl[2];
          sp_h4 bh3, bl3;
          bh[0] = rd8(wh, woff[0], woff[1]);
          bl[0] = rd8(wl, woff[0], woff[1]);
#pragma unroll
          for (int ks = 0; ks < 3; ++ks) {
            const int c = ks & 1;
            if (ks < 2) {
              bh[c ^ 1] = rd8(wh, woff[2 * ks + 2], woff[2 * ks + 3]);
              bl[c ^ 1] = rd8(wl, woff[2 * ks + 2], woff[2 * ks + 3]);
            } else {
              bh3 = __builtin_bit_cast(sp_h4, rd4(wh, woff[6]));
              bl3 = __builtin_bit_cast(sp_h4, rd4(wl, woff[6]));
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int m = m0; m < m0 + 3; ++m)
              acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[m][ks], bh[c], acc[m], 0, 0, 0);
#pragma unroll
            for (int m = m0; m < m0 + 3; ++m)
              acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo[m][ks], bh[c], acc[m], 0, 0, 0);
#pragma unroll
            for (int m = m0; m < m0 + 3; ++m)
              acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi[m][ks], bl[c], acc[m], 0, 0, 0);
            work(ks);
            __builtin_amdgcn_sched_barrier(0);
          }
#pragma unroll
          for (int m = m0; m < m0 + 3; ++m) acc[m] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi3[m], bh3, acc[m], 0, 0, 0);
#pragma unroll
          for (int m = m0; m < m0 + 3; ++m) acc[m] = __builtin_amdgcn_mfma_f32_16x16x16f16(alo3[m], bh3, acc[m], 0, 0, 0);
#pragma unroll
          for (int m = m0; m < m0 + 3; ++m) acc[m] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi3[m], bl3, acc[m], 0, 0, 0);
          work(3);
        };
        auto none = [](int) {};
        mfma_half(0, none);
        if (lines) {
          // element scales shifted like the DPP-combined rows (full wave)
          float s1 = sp_shr<1>(s), s2 = sp_shr<2>(s), r2 = sp_ror<2>(s), q1 = sp_ror<1>(s);
          asm("" : "+v"(s1), "+v"(s2), "+v"(r2), "+v"(q1));
          const float r1 = col == 0 ? q1 : 0.0f;
          double* yl = yb + pb + col;
          // point x = col of the line receives s v0[col] + s1 v1[col - 1] +
          // s2 v2[col - 2]: branch-free (lanes past the line add to a dummy
          // slot), so it interleaves with the MFMAs of M tiles 3..5
          auto comb_add = [&](int ro, bool on, float v0, float v1, float v2) {
            float cb = v0 * s;
            cb = fmaf(sp_shr<1>(v1), s1, cb);
            cb = fmaf(sp_shr<2>(v2), s2, cb);
            unsafeAtomicAdd((on && cmask) ? yl + ro : dummy, (double)cb);
          };
          // points 16 and 17 (lines of 15 and 16 elements) from lanes 15 and
          // 14, rotated to lanes 0 and 1
          auto ext_val = [&](float v1, float v2) {
            float ex = sp_ror<2>(v2) * r2;
            return fmaf(sp_ror<1>(v1), r1, ex);
          };
          // M tiles 0..2 (and their fourth registers) scattered behind the
          // K steps of M tiles 3..5
          mfma_half(3, [&](int ks) {
            if (ks < 3)
              comb_add(roff[ks], true, acc[ks][0], acc[ks][1], acc[ks][2]);
            else
              comb_add(roff3, kq < 3, acc[0][3], acc[1][3], acc[2][3]);
          });
#pragma unroll
          for (int m = 3; m < 6; ++m) comb_add(roff[m], true, acc[m][0], acc[m][1], acc[m][2]);
          if (nx >= 15) {  // uniform
            float e[7];
#pragma unroll
            for (int m = 0; m < 6; ++m) e[m] = ext_val(acc[m][1], acc[m][2]);
            e[6] = ext_val(acc[1][3], acc[2][3]);
            asm("" : "+v"(e[0]), "+v"(e[1]), "+v"(e[2]), "+v"(e[3]), "+v"(e[4]), "+v"(e[5]), "+v"(e[6]));
            if (emask) {
#pragma unroll
              for (int m = 0; m < 6; ++m) unsafeAtomicAdd(yl + roff[m] + 16, (double)e[m]);
              if (kq < 3) unsafeAtomicAdd(yl + roff3 + 16, (double)e[6]);
            }
          }
        } else {
          mfma_half(3, none);
          if (valid) {
#pragma unroll
            for (int m = 0; m < 6; ++m) {
              double* ye = yb + pb + roff[m];
              unsafeAtomicAdd(ye + 0, (double)(acc[m][0] * s));
              unsafeAtomicAdd(ye + 1, (double)(acc[m][1] * s));
              unsafeAtomicAdd(ye + 2, (double)(acc[m][2] * s));
            }
            if (kq < 3) {
              double* ye = yb + pb + roff3;
#pragma unroll
              for (int m = 0; m < 3; ++m) unsafeAtomicAdd(ye + m, (double)(acc[m][3] * s));
            }
          }
        }
      }
      __syncthreads();
    }
  }
  // The last tile's output box (complete since the barrier that ended the
  // tile) is flushed by all eight waves: nothing is left to overlap with.
  flush(last - 1, threadIdx.x, SP_THREADS);
  if (pap != nullptr) {
    for (int off = 32; off > 0; off >>= 1) pap_acc += __shfl_down(pap_acc, off, 64);
    if (lane == 0) red[wave] = pap_acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int w = 0; w < SP_THREADS / 64; ++w) sum += red[w];
      pap[blockIdx.x] = sum;
    }
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
