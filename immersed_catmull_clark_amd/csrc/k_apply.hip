// K.u apply, element energies and Jacobi diagonal for p = 2 trivariate
// B-spline grids on gfx950.
//
// Work decomposition.  The element grid is cut into tiles of <= 128 elements
// that never straddle an element-type boundary (host: setup.cpp), so every
// element of a tile shares one 81x81 matrix K_t.  K_t's rows form six 16-row
// M tiles (the sixth holds row 80 only).  A workgroup walks a run of tiles
// (host: balance_tiles) with W waves, each owning one M tile (apply_mtile;
// W = 12 in FP64, two waves per M tile claiming its element groups
// dynamically; W = 8 in FP32 / BF16, apply_pairs).  A wave keeps its 21 A fragments (K rows x 4-deep K
// steps, pre-permuted on the host into lane order) in VGPRs and streams the
// tile's elements 16 at a time as the MFMA N dimension:
//     D[16 x 16] += A[16 x 4] * B[4 x 16],  B[k][j] = x_local(element j, trial k)
// with the B operand gathered from the LDS copy of the tile's control points
// (x masked by the free mask).  The 4 accumulator rows per lane are scaled by
// s_e and scatter-added into an LDS copy of the tile's output points
// (ds_add_f64), and the tile is flushed to HBM with global atomic adds (y
// holds the sum to accumulate into).  The p.Ap partial of PCG is accumulated
// in that flush from the staged points and the tile's contributions.
//
// MFMA lane layouts (v_mfma_{f64,f32}_16x16x4): lane l holds A[l%16][l/16] and
// B[l/16][l%16]; accumulator register r of lane l holds D[i][l%16] with
// i = 4 r + l/16 for f64 and i = 4 (l/16) + r for f32 (Mfma<T>::row).
// icc_debug_mfma_layout pins both on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

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

// Exact quotient n / d for n < 2^12 and 2 <= d < 2^10 with m = ceil(2^32 / d):
// floor(n * m / 2^32) (the error n (m - 2^32/d) / 2^32 stays below 1/d).  The
// multipliers sit in a constant-memory table, so a tile's geometry costs a few
// scalar loads instead of integer divisions.
constexpr int MAGIC_MAX = 1024;
struct MagicTab {
  uint32_t v[MAGIC_MAX];
  constexpr MagicTab() : v() {
    for (uint32_t d = 2; d < (uint32_t)MAGIC_MAX; ++d) v[d] = 0xffffffffu / d + 1u;
  }
};
__constant__ MagicTab kMagic = MagicTab();
static_assert(TILE_MAX_PTS < MAGIC_MAX, "tile point counts index the multiplier table");
__device__ __forceinline__ int qdiv(int n, uint32_t m) { return (int)__umulhi((uint32_t)n, m); }
// n / d for 0 <= n < 2^12, 1 <= d < MAGIC_MAX
__device__ __forceinline__ int sdiv(int n, int d) { return d == 1 ? n : qdiv(n, kMagic.v[d]); }

// Geometry of a staged tile: points [c][lz][ly][lx] of extents px, py, pz.
struct TileGeom {
  int px, py, npts, n3, nel;
  uint32_t m_npts, m_px, m_py;
  __device__ __forceinline__ void init(const Tile& t) {
    px = t.nx + 2;
    py = t.ny + 2;
    npts = px * py * (t.nz + 2);
    n3 = 3 * npts;
    nel = t.nx * t.ny * t.nz;
    m_npts = kMagic.v[npts];
    m_px = kMagic.v[px];
    m_py = kMagic.v[py];
  }
  // global DOF of staged slot idx < n3
  __device__ __forceinline__ int dof(const Tile& t, int idx, int ncp0, int ncp1, int ndof_sc) const {
    const int c = qdiv(idx, m_npts);
    const int p = idx - c * npts;
    const int r = qdiv(p, m_px);
    const int lx = p - r * px;
    const int lz = qdiv(r, m_py);
    const int ly = r - lz * py;
    return (t.i0 + lx) + ncp0 * ((t.j0 + ly) + ncp1 * (t.k0 + lz)) + c * ndof_sc;
  }
  // global element of tile element el < nel
  __device__ __forceinline__ int elem(const Tile& t, int el, int nel0, int nel1) const {
    const int r = sdiv(el, t.nx);
    const int ex = el - r * t.nx;
    const int ez = sdiv(r, t.ny);
    const int ey = r - ez * t.ny;
    return (t.i0 + ex) + nel0 * ((t.j0 + ey) + nel1 * (t.k0 + ez));
  }
};

// Waves per workgroup W, with wave w on SIMD w mod 4 (the dispatch order):
//   W = 12 (FP64): two waves per M tile, wave w owns M tile w mod 6 (the
//     static split below is the fallback; the kernel claims pairs from an
//     LDS counter); every SIMD carries three waves;
//   W = 8 (FP32, BF16): waves 0..3 own M tiles 0..3 over all pairs, waves
//     4..7 own M tiles 4 and 5 (4, 6: tile 4; 5, 7: tile 5) over half of them;
//     every SIMD carries 1.5 M tiles of work, and at <= 128 VGPRs two
//     workgroups share a CU.
// Either way no SIMD holds up the tile's barrier.
template <int W>
__device__ __forceinline__ int apply_mtile(int wave) {
  return W == 12 ? wave % 6 : (wave < 4 ? wave : 4 + (wave & 1));
}
// element-group pairs [lo, hi) of a wave among npairs
template <int W>
__device__ __forceinline__ void apply_pairs(int wave, int npairs, int& lo, int& hi) {
  const int half = (npairs + 1) / 2;
  if (W == 12) {
    lo = wave < 6 ? 0 : half;
    hi = wave < 6 ? half : npairs;
  } else {
    lo = wave < 6 ? 0 : half;
    hi = (wave < 4 || wave >= 6) ? npairs : half;
  }
}
constexpr int BCH = 4;                        // K steps per B-operand chunk
constexpr int NBCH = (KS + BCH - 1) / BCH;    // chunks per element group
// BF16 mode (v_mfma_f32_16x16x32_bf16): lane l holds A[l%16][8 (l/16) + j]
// and B[8 (l/16) + j][l%16], j = 0..7; accumulators as v_mfma_f32_16x16x4.
constexpr int KSB = 3;                        // 32-deep K steps covering 81 trial columns (96 padded)
// trial offsets of a tile: [kq][ks] (4 x KS) in the FP64 / FP32 modes,
// [kq][ks][j] (4 x KSB x 8) in the BF16 mode; then [wave][kq][r] output-row
// offsets (6 x 4 x 4)
template <bool BF>
struct Otab {
  static constexpr int ntrial = BF ? 4 * KSB * 8 : 4 * KS;
  static constexpr int size = ntrial + MT * 16;
};
constexpr int OTAB_MAX = Otab<true>::size;

// Fills the offset table of a tile (threads 0 .. size-1).  Entries for the
// padding trial columns and rows point at offset 0 (their A entries are zero
// and padding rows are never stored).
template <bool BF>
__device__ __forceinline__ void offset_table(const TileGeom& G, int* tab, int i) {
  constexpr int NT = Otab<BF>::ntrial;
  if (i < NT) {
    int b;
    if (BF) {
      const int kq = i / (KSB * 8), rem = i - KSB * 8 * kq, ks = rem / 8, j = rem - 8 * ks;
      b = 32 * ks + 8 * kq + j;
    } else {
      const int kq = i / KS, ks = i - KS * kq;
      b = 4 * ks + kq;
    }
    tab[i] = b < NSH ? local_offset(b, G.px, G.py, G.npts) : 0;
  } else if (i < Otab<BF>::size) {
    const int j = i - NT, w = j >> 4, kq = (j >> 2) & 3, r = j & 3;
    const int row = perm_row(w, kq, r);
    tab[i] = row >= 0 ? local_offset(row, G.px, G.py, G.npts) : 0;
  }
}
// staged slots per thread of a W-wave workgroup
template <int W>
struct Npf {
  static constexpr int value = (3 * TILE_MAX_PTS + 64 * W - 1) / (64 * W);
};

// Tiles [first, last) of workgroup b among G: chunk c of the host's balanced
// partition (wg_first, G + 1 entries), the chunks of the workgroups that share
// an XCD (b mod 8) adjacent, so that neighbouring tiles and their shared halo
// points meet in one L2.
__device__ __forceinline__ void tile_range(const int* __restrict__ wg_first, int& first, int& last) {
  const int G = gridDim.x, b = blockIdx.x;
  const int c = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  first = wg_first[c];
  last = wg_first[c + 1];
}

// y += M * sum_e s_e K_t x_e (M x); optional per-workgroup partials of
// x^T M K M x into pap[blockIdx.x].
//
// Each workgroup walks a contiguous run of tiles through a three-deep
// pipeline with double-buffered LDS (points, scales, offset tables and
// output points): at the top of tile i a wave flushes tile i - 1's output
// points to y with global atomic adds (y must hold the sum to accumulate
// into), stages tile i + 1 from registers and issues the loads of tile
// i + 2, then runs its MFMA loop on tile i; one barrier per tile.  The A
// fragments are reloaded only when the element type changes.
//
// Each wave runs two element groups at a time as independent MFMA chains.
// In line mode a group is one x-line of the tile (lane column = x offset);
// accumulator registers 0..2 of a lane hold the ax = 0, 1, 2 rows of one
// (component, ay, az) triplet (perm_row), so the three contributions each
// control point of the line receives are summed in registers with two DPP
// row shifts before a single ds_add per point.
//
// BF = true (T = float): K and the staged x are rounded to BF16 for
// v_mfma_f32_16x16x32_bf16 (3 K steps of 32), FP32 accumulation; the rest as
// the FP32 mode.
//
// MI = true (a batch of instances): the tiles of all instances form one list;
// tinst[i] gives tile i's instance geometry and vector offsets (dof(), elem()
// add dbase / ebase), and p.Ap goes to one slot per (workgroup run, instance):
// each flush sums its tile's share across the flushing waves, and thread 0
// adds the tile sums in tile order, writing an instance's slot when the run
// moves on.  The tile types index the concatenated A fragments of all
// instances.
template <typename T, bool BF, int W, bool MI = false>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(W == 12 ? 3 : 4))) void apply_tiles(
    const Tile* __restrict__ tiles, const int* __restrict__ wg_first, const void* __restrict__ kfrag_v,
    const T* __restrict__ x,
    T* __restrict__ y, const T* __restrict__ scale, const uint8_t* __restrict__ freem,
    const int32_t* __restrict__ tmask, int ncp0, int ncp1, int ndof_sc, int nel0, int nel1,
    double* __restrict__ pap, const PcgState* st, const TileInst* __restrict__ tinst = nullptr) {
  __shared__ T xs[2][3 * TILE_MAX_PTS];
  // output points accumulate in FP64 for both precisions: ds_add_f64 runs at
  // full rate, while FP32 LDS atomics back up the LDS data FIFO on gfx950
  __shared__ double ys[2][3 * TILE_MAX_PTS];
  __shared__ T ss[2][TILE_MAX_EL];
  __shared__ int otab[2][OTAB_MAX];  // per tile: LDS offsets of the B operands and of the output rows
  constexpr int NTR = Otab<BF>::ntrial;
  __shared__ double red[W];
  __shared__ int claim[2][8];  // FP64: next element-group pair of each M tile, per buffer
  constexpr int NT = 64 * W;
  // FP64 (twelve waves): waves 6..11 do the per-tile flush, staging and
  // prefetch while waves 0..5 start their MFMA loops at once, and the two
  // waves of an M tile claim its element-group pairs from an LDS counter, so
  // that whichever is free takes the next one.  (With eight waves the split
  // gained nothing.)
  constexpr bool SPEC = W == 12;
  constexpr int NBG = SPEC ? 6 * 64 : NT;
  constexpr int NPF = (3 * TILE_MAX_PTS + NBG - 1) / NBG;
  if (st != nullptr && st->done != 0) return;  // solver already finished
  int first, last;
  tile_range(wg_first, first, last);
  if (first >= last) {
    if (!MI && pap != nullptr && threadIdx.x == 0) pap[blockIdx.x] = 0.0;
    return;
  }
  // geometry and vector offsets of tile i
  const auto ctx_of = [&](int i) {
    TileInst c;
    if constexpr (MI) {
      c = tinst[i];
    } else {
      c.dbase = 0;
      c.ebase = 0;
      c.ncp0 = ncp0;
      c.ncp1 = ncp1;
      c.ndof_sc = ndof_sc;
      c.nel0 = nel0;
      c.nel1 = nel1;
      c.slot = 0;
    }
    return c;
  };
  __shared__ double red2[2][W];  // MI: per-wave p.Ap shares of the last two flushed tiles
  double mi_acc = 0.0;           // MI, thread 0: the current instance's sum in this run
  int mi_slot = -1;

  const int wave = threadIdx.x >> 6;
  const int mt = apply_mtile<W>(wave);
  const int lane = threadIdx.x & 63;
  const int col = lane & 15;
  const int kq = lane >> 4;
  const int tb = SPEC ? (int)threadIdx.x - 6 * 64 : (int)threadIdx.x;  // background thread index
  const bool bgw = !SPEC || wave >= 6;

  // Flush of a finished tile tf from output buffer b (masked): every staged
  // point's sum goes to y with a global atomic add and the buffer is cleared.
  // The p.Ap partial is the sum over the staged points of (M x)_i times the
  // tile's contribution to (A x)_i, which summed over tiles is x^T M K M x.
  double pap_acc = 0.0;  // this thread's share, tiles in order
  auto flush = [&](const Tile& tf, int tfi, bool maskf, int b) {
    if (!bgw) return;
    TileGeom Gf;
    Gf.init(tf);
    const TileInst cf = ctx_of(tfi);
    double pt = 0.0;
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int idx = tb + j * NBG;
      if (idx < Gf.n3) {
        const double v = ys[b][idx];
        ys[b][idx] = 0.0;
        if (pap != nullptr) {
          if constexpr (MI)
            pt += (double)xs[b][idx] * v;
          else
            pap_acc += (double)xs[b][idx] * v;
        }
        const int g = Gf.dof(tf, idx, cf.ncp0, cf.ncp1, cf.ndof_sc) + cf.dbase;
        if (!maskf || freem[g] != 0) unsafeAtomicAdd(&y[g], (T)v);
      }
    }
    if constexpr (MI) {
      for (int off = 32; off > 0; off >>= 1) pt += __shfl_down(pt, off, 64);
      if (lane == 0) red2[b][wave] = pt;
    }
  };
  // MI, thread 0: adds flushed tile j's share (written by waves w0..W-1
  // before a barrier the caller has passed) to its instance's slot sum
  auto collect = [&](int j, int w0) {
    if constexpr (MI) {
      const int b = (j - first) & 1;
      double sj = 0.0;
      for (int w = w0; w < W; ++w) sj += red2[b][w];
      const int slot = tinst[j].slot;
      if (slot != mi_slot) {
        if (mi_slot >= 0) pap[mi_slot] = mi_acc;
        mi_acc = 0.0;
        mi_slot = slot;
      }
      mi_acc += sj;
    }
  };
  // loads of tile tl's points and scales into registers (unused slots are not
  // loaded)
  T pv[NPF];
  T psc = T(0);
  auto load_regs = [&](int tli) {
    if (!bgw) return;
    const Tile tl = tiles[tli];
    const TileInst cl = ctx_of(tli);
    TileGeom Gl;
    Gl.init(tl);
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int idx = tb + j * NBG;
      if (idx < Gl.n3) pv[j] = x[Gl.dof(tl, idx, cl.ncp0, cl.ncp1, cl.ndof_sc) + cl.dbase];
    }
    if (tb < Gl.nel) psc = scale[Gl.elem(tl, tb, cl.nel0, cl.nel1) + cl.ebase];
  };

  // A fragments of this wave's M tile mt for element type `type`; consumed
  // by an empty asm each at the point given by the caller, so that the loads
  // are waited for there (never in the MFMA loop, whose waits would also
  // wait for the flush's atomics and the prefetch of tile i + 2)
  T a[BF ? 1 : KS];
  bf8_t ab[BF ? KSB : 1];
  const auto load_a = [&](int type) {
    if constexpr (BF) {
      const bf8_t* kf = (const bf8_t*)kfrag_v + ((size_t)(type * MT + mt) * KSB) * 64 + lane;
#pragma unroll
      for (int ks = 0; ks < KSB; ++ks) ab[ks] = kf[ks * 64];
    } else {
      const T* kf = (const T*)kfrag_v + ((size_t)(type * MT + mt) * KS) * 64 + lane;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) a[ks] = kf[ks * 64];
    }
  };
  const auto use_a = [&]() {
    if constexpr (BF) {
#pragma unroll
      for (int ks = 0; ks < KSB; ++ks) asm volatile("" ::"v"(ab[ks]));
    } else {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) asm volatile("" ::"v"(a[ks]));
    }
  };

  // prologue: stage the first tile into buffer 0, clear both output buffers,
  // load the second tile into registers.  The first tile's A fragments are
  // loaded first: their latency overlaps the staging's.
  Tile t = tiles[first];
  TileGeom G;
  G.init(t);
  int cur_type = t.type;
  load_a(cur_type);
  {
    const bool masked = tmask != nullptr && tmask[first] != 0;
    const TileInst c0 = ctx_of(first);
    for (int idx = threadIdx.x; idx < 3 * TILE_MAX_PTS; idx += NT) {
      ys[0][idx] = 0.0;
      ys[1][idx] = 0.0;
      if (idx < 16) claim[idx >> 3][idx & 7] = 0;
      if (idx < G.n3) {
        const int g = G.dof(t, idx, c0.ncp0, c0.ncp1, c0.ndof_sc) + c0.dbase;
        xs[0][idx] = (!masked || freem[g] != 0) ? x[g] : T(0);
      }
    }
    if (threadIdx.x < G.nel) ss[0][threadIdx.x] = scale[G.elem(t, threadIdx.x, c0.nel0, c0.nel1) + c0.ebase];
  }
  offset_table<BF>(G, otab[0], threadIdx.x);

  // The tile pipeline, three deep.  At the top of tile i the background
  // waves flush tile i - 1 (its output buffer is complete once every wave has
  // passed the barrier that ends tile i - 1), stage tile i + 1 from the
  // registers their loads went to at the top of tile i - 1, and issue the
  // loads of tile i + 2; then they join the MFMA loop on tile i, which the
  // other waves started right after the barrier.  One barrier per tile
  // suffices.  Descriptors and mask flags (tiles whose staged points include
  // fixed DOFs, few, read the free mask) are read where they are used: after
  // their first use they sit in the scalar cache, and no descriptor has to
  // be carried (and spilled) across tiles.
  const auto mask_at = [&](int i) { return tmask != nullptr && tmask[i] != 0; };
  use_a();  // arrived with the staged points
  if (first + 1 < last) load_regs(first + 1);
  __syncthreads();  // buffer 0 staged, output buffers clear

  int buf = 0;
  for (int ti = first; ti < last; ++ti) {
    if (ti > first) {
      t = tiles[ti];
      G.init(t);
    }
    // the A fragments are (re)loaded before any other memory operation of
    // the tile and waited for on this (rare, uniform) path, so that waiting
    // for them never waits for the flush's atomics or the prefetch below (the
    // vector memory counter retires in order)
    if (t.type != cur_type) {
      load_a(t.type);
      use_a();
      cur_type = t.type;
    }
    if (MI && threadIdx.x == 0 && ti - 2 >= first) collect(ti - 2, SPEC ? 6 : 0);
    if (ti > first) flush(tiles[ti - 1], ti - 1, mask_at(ti - 1), buf ^ 1);
    if (bgw && tb < 8) claim[buf ^ 1][tb] = 0;
    if (bgw && ti + 1 < last) {  // stage tile i + 1 into buffer buf ^ 1
      const Tile tn = tiles[ti + 1];
      const bool mask_n = mask_at(ti + 1);
      const TileInst cn = ctx_of(ti + 1);
      TileGeom Gn;
      Gn.init(tn);
#pragma unroll
      for (int j = 0; j < NPF; ++j) {
        const int idx = tb + j * NBG;
        if (idx < Gn.n3)
          xs[buf ^ 1][idx] =
              (!mask_n || freem[Gn.dof(tn, idx, cn.ncp0, cn.ncp1, cn.ndof_sc) + cn.dbase] != 0) ? pv[j] : T(0);
      }
      if (tb < Gn.nel) ss[buf ^ 1][tb] = psc;
      offset_table<BF>(Gn, otab[buf ^ 1], tb);
    }
    if (ti + 2 < last) load_regs(ti + 2);

    const int px = G.px, py = G.py, nel = G.nel;
    const bool r3 = perm_row(mt, kq, 3) >= 0;
    const T* xb = xs[buf];
    const T* sb = ss[buf];
    double* yb = ys[buf];
    constexpr int NB = NTR / 4;  // trial offsets per row group
    int boff[NB];
#pragma unroll
    for (int m = 0; m < NB; ++m) boff[m] = otab[buf][kq * NB + m];
    int roff[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) roff[r] = otab[buf][NTR + (mt * 4 + kq) * 4 + r];

    const bool lines = t.lines != 0;
    const int ngroups = lines ? t.ny * t.nz : (nel + 15) / 16;
    int p_lo, p_hi;
    apply_pairs<W>(wave, (ngroups + 1) / 2, p_lo, p_hi);
    const int npairs = (ngroups + 1) / 2;
    auto next_pair = [&]() {
      int p = 0;
      if (lane == 0) p = atomicAdd(&claim[buf][mt], 1);
      p = __shfl(p, 0, 64);
      return __builtin_amdgcn_readfirstlane(p);
    };
    for (int g0 = SPEC ? 2 * next_pair() : 2 * p_lo; g0 < (SPEC ? 2 * npairs : 2 * p_hi);
         g0 = SPEC ? 2 * next_pair() : g0 + 2) {
      if constexpr (W == 8) {
        // waves with more pairs left issue first, so that the full-M-tile
        // waves (twice the work of the half-tile ones on their SIMD) are not
        // left to finish alone
        const int rem = p_hi - g0 / 2;
        if (rem >= 3)
          __builtin_amdgcn_s_setprio(3);
        else if (rem == 2)
          __builtin_amdgcn_s_setprio(2);
        else
          __builtin_amdgcn_s_setprio(1);
      }
      int eb[2];
      bool valid[2];
      T sc[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int g = g0 + q;
        int e;
        if (lines) {
          const int ez = sdiv(g, t.ny), ey = g - ez * t.ny;
          valid[q] = g < ngroups && col < t.nx;
          e = col + t.nx * g;
          eb[q] = valid[q] ? col + px * (ey + py * ez) : 0;
        } else {
          e = 16 * g + col;
          valid[q] = g < ngroups && e < nel;
          int ex = 0, ey = 0, ez = 0;
          if (valid[q]) {
            const int rr = sdiv(e, t.nx);
            ex = e - rr * t.nx;
            ez = sdiv(rr, t.ny);
            ey = rr - ez * t.ny;
          }
          eb[q] = ex + px * (ey + py * ez);
        }
        // the scale is read before the MFMA chain so that no LDS read sits
        // between the chain and the scatter below
        sc[q] = sb[valid[q] ? e : 0];
        if (!valid[q]) sc[q] = T(0);
      }
      // B operands software-pipelined in register chunks of BCH K steps:
      // the reads of chunk c + 1 are in flight while chunk c runs on MFMA
      // (left to itself the scheduler issues each read right before its MFMA
      // and exposes the full LDS latency)
      typename Mfma<T>::acc_t acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
      if constexpr (BF) {
#pragma unroll
        for (int ks = 0; ks < KSB; ++ks) {
          bf8_t b0, b1;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            b0[j] = (__bf16)xb[eb[0] + boff[ks * 8 + j]];
            b1[j] = (__bf16)xb[eb[1] + boff[ks * 8 + j]];
          }
          acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab[ks], b0, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab[ks], b1, acc1, 0, 0, 0);
        }
      } else {
      T bq0[2][BCH], bq1[2][BCH];
#pragma unroll
      for (int i = 0; i < BCH; ++i) {
        bq0[0][i] = xb[eb[0] + boff[i]];
        bq1[0][i] = xb[eb[1] + boff[i]];
      }
#pragma unroll
      for (int c = 0; c < NBCH; ++c) {
        const int cur = c & 1;
        if (c + 1 < NBCH) {
#pragma unroll
          for (int i = 0; i < BCH; ++i) {
            const int ks = (c + 1) * BCH + i;
            if (ks < KS) {
              bq0[cur ^ 1][i] = xb[eb[0] + boff[ks]];
              bq1[cur ^ 1][i] = xb[eb[1] + boff[ks]];
            }
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < BCH; ++i) {
          const int ks = c * BCH + i;
          if (ks < KS) {
            acc0 = Mfma<T>::run(a[ks], bq0[cur][i], acc0);
            acc1 = Mfma<T>::run(a[ks], bq1[cur][i], acc1);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      }
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const typename Mfma<T>::acc_t& acc = q == 0 ? acc0 : acc1;
        const T s = sc[q];
        const T v0 = acc[0] * s, v1 = acc[1] * s, v2 = acc[2] * s, v3 = acc[3] * s;
        if (lines) {
          // point x = col of the line receives v0[col] + v1[col-1] + v2[col-2];
          // the DPP shifts run on all lanes, before any lane-dependent branch
          const T comb = v0 + row_shr<1>(v1) + row_shr<2>(v2);
          const T e16 = v1 + row_shr<1>(v2);  // at lane 15: point 16 = v1[15] + v2[14]
          const int g = g0 + q;
          if (g < ngroups) {
            const int ez = sdiv(g, t.ny), ey = g - ez * t.ny;
            const int lb = px * (ey + py * ez) + roff[0];  // point x = 0 of this line, ax = 0 row
            if (col <= t.nx + 1) unsafeAtomicAdd(&yb[lb + col], (double)comb);
            if (col == 15) {  // points 16 and 17 exist only for lines of 15 and 16 elements
              if (t.nx >= 15) unsafeAtomicAdd(&yb[lb + 16], (double)e16);
              if (t.nx >= 16) unsafeAtomicAdd(&yb[lb + 17], (double)v2);
            }
            if (r3 && valid[q]) unsafeAtomicAdd(&yb[eb[q] + roff[3]], (double)v3);
          }
        } else if (valid[q]) {
          unsafeAtomicAdd(&yb[eb[q] + roff[0]], (double)v0);
          unsafeAtomicAdd(&yb[eb[q] + roff[1]], (double)v1);
          unsafeAtomicAdd(&yb[eb[q] + roff[2]], (double)v2);
          if (r3) unsafeAtomicAdd(&yb[eb[q] + roff[3]], (double)v3);
        }
      }
    }
    if constexpr (W == 8) __builtin_amdgcn_s_setprio(0);
    __syncthreads();  // tile i's output complete; buffer buf ^ 1 staged
    buf ^= 1;
  }
  {
    // the last tile: nothing is left to overlap with, so every wave flushes
    // (the same point order as flush() when all waves are background waves)
    const int b = buf ^ 1;
    const Tile tf = tiles[last - 1];
    TileGeom Gf;
    Gf.init(tf);
    const TileInst cf = ctx_of(last - 1);
    const bool maskf = mask_at(last - 1);
    double pt = 0.0;
    for (int idx = threadIdx.x; idx < Gf.n3; idx += NT) {
      const double v = ys[b][idx];
      if (pap != nullptr) {
        if constexpr (MI)
          pt += (double)xs[b][idx] * v;
        else
          pap_acc += (double)xs[b][idx] * v;
      }
      const int g = Gf.dof(tf, idx, cf.ncp0, cf.ncp1, cf.ndof_sc) + cf.dbase;
      if (!maskf || freem[g] != 0) unsafeAtomicAdd(&y[g], (T)v);
    }
    if constexpr (MI) {
      for (int off = 32; off > 0; off >>= 1) pt += __shfl_down(pt, off, 64);
      if (lane == 0) red2[b][wave] = pt;
    }
  }
  if constexpr (MI) {
    __syncthreads();
    if (threadIdx.x == 0) {
      if (last - 2 >= first) collect(last - 2, SPEC ? 6 : 0);
      collect(last - 1, 0);
      pap[mi_slot] = mi_acc;
    }
    return;
  }
  if (pap != nullptr) {
    // wave sums, then the waves in fixed order
    for (int off = 32; off > 0; off >>= 1) pap_acc += __shfl_down(pap_acc, off, 64);
    if (lane == 0) red[wave] = pap_acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int w = 0; w < W; ++w) sum += red[w];
      pap[blockIdx.x] = sum;
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

// Deterministic apply, pass 1: every element's 81 outputs s_e K_t x_e
// (x masked) into E[row][e] (row-major by output slot, so that the gather's
// neighbouring control points read neighbouring elements), one workgroup of
// six waves (one per M tile) per tile.  Each entry is written once, so the
// result does not depend on the order in which the workgroups run.
template <typename T>
__global__ __launch_bounds__(384) void det_elem_kernel(const Tile* __restrict__ tiles, const void* __restrict__ kfrag_v,
                                                       const T* __restrict__ x, const T* __restrict__ scale,
                                                       const uint8_t* __restrict__ freem,
                                                       const int32_t* __restrict__ tmask, int ncp0, int ncp1,
                                                       int ndof_sc, int nel0, int nel1, int64_t nel_tot,
                                                       T* __restrict__ E, const PcgState* st) {
  if (st != nullptr && st->done != 0) return;
  __shared__ T xs[3 * TILE_MAX_PTS];
  const Tile t = tiles[blockIdx.x];
  const int px = t.nx + 2, py = t.ny + 2, pz = t.nz + 2;
  const int npts = px * py * pz;
  const int nel = t.nx * t.ny * t.nz;
  const bool masked = tmask != nullptr && tmask[blockIdx.x] != 0;
  stage_tile<T>(t, x, masked ? freem : nullptr, ncp0, ncp1, ndof_sc, xs, nullptr, px, py, npts);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int col = lane & 15;
  const int kq = lane >> 4;
  T a[KS];
  int boff[KS];
  const T* kf = (const T*)kfrag_v + ((size_t)(t.type * MT + wave) * KS) * 64 + lane;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    a[ks] = kf[ks * 64];
    const int b = 4 * ks + kq;
    boff[ks] = b < NSH ? local_offset(b, px, py, npts) : 0;
  }
  int rowi[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) rowi[r] = perm_row(wave, kq, r);
  __syncthreads();
  for (int g0 = 0; g0 < nel; g0 += 16) {
    const int el = g0 + col;
    const bool valid = el < nel;
    int ex = 0, ey = 0, ez = 0;
    if (valid) {
      ex = el % t.nx;
      const int rr = el / t.nx;
      ey = rr % t.ny;
      ez = rr / t.ny;
    }
    const int eb = ex + px * (ey + py * ez);
    typename Mfma<T>::acc_t acc = {0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) acc = Mfma<T>::run(a[ks], xs[eb + boff[ks]], acc);
    if (valid) {
      const int64_t ge = (t.i0 + ex) + (int64_t)nel0 * ((t.j0 + ey) + (int64_t)nel1 * (t.k0 + ez));
      const T sc = scale[ge];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (rowi[r] >= 0) E[rowi[r] * nel_tot + ge] = acc[r] * sc;
    }
  }
}

// Deterministic apply, pass 2: y(c, P) += sum over the <= 27 elements around
// control point P, in a fixed order, of E[(c, a)][e] (masked), accumulated in
// FP64 and rounded once; pap[block] = the block's share of (M x).y in a fixed
// tree.
template <typename T>
__global__ __launch_bounds__(256) void det_gather_kernel(const T* __restrict__ E, const T* __restrict__ x,
                                                         T* __restrict__ y, const uint8_t* __restrict__ freem,
                                                         int ncp0, int ncp1, int ncp2, int nel0, int nel1, int nel2,
                                                         double* __restrict__ pap, const PcgState* st) {
  if (st != nullptr && st->done != 0) return;
  __shared__ double sh[4];
  const int64_t ndof_sc = (int64_t)ncp0 * ncp1 * ncp2;
  const int64_t nel_tot = (int64_t)nel0 * nel1 * nel2;
  double part = 0.0;
  for (int64_t P = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; P < ndof_sc; P += (int64_t)gridDim.x * blockDim.x) {
    const int pi = (int)(P % ncp0), pj = (int)((P / ncp0) % ncp1), pk = (int)(P / ((int64_t)ncp0 * ncp1));
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
          const T* Ee = E + (int64_t)(ax + 3 * ay + 9 * az) * nel_tot + (ei + (int64_t)nel0 * (ej + (int64_t)nel1 * ek));
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[c] += (double)Ee[c * NSC * nel_tot];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t gi = c * ndof_sc + P;
      const bool fr = freem == nullptr || freem[gi] != 0;
      const T v = fr ? (T)acc[c] : T(0);
      y[gi] += v;
      if (fr) part += (double)x[gi] * (double)v;
    }
  }
  if (pap != nullptr) {
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = part;
    __syncthreads();
    if (threadIdx.x == 0) pap[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
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

// Lane-layout check of v_mfma_f32_16x16x32_bf16: the lanes are filled under
// the layout the BF16 apply assumes (lane l holds A[l%16][8 (l/16) + j] and
// B[8 (l/16) + j][l%16]) with small integers A[i][k] = (7 i + 3 k) % 17 - 8,
// B[k][n] = (5 k + 11 n) % 13 - 6 (exact in BF16 and FP32); out[4 l + r] =
// accumulator r, which under the assumed output layout is D[4 (l/16) + r][l%16].
__global__ void mfma_layout_probe_bf16(double* out) {
  const int l = threadIdx.x;
  bf8_t a, b;
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * (l / 16) + j, i = l % 16, n = l % 16;
    a[j] = (__bf16)(float)((7 * i + 3 * k) % 17 - 8);
    b[j] = (__bf16)(float)((5 * k + 11 * n) % 13 - 6);
  }
  f4_t acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[l * 4 + r] = (double)acc[r];
}

// ---- launchers ---------------------------------------------------------------
// Apply grid: wg_per_cu workgroups per CU (1, 2 or 4; the size the tile list
// was partitioned for).  Resident at once: one FP64 workgroup (twelve waves at
// the MFMA loop's VGPR budget) or two FP32 / BF16 ones per CU; the others
// start as those finish, which the hardware dispatcher balances.
int apply_grid(const icc_grid* g) { return g->apply_G; }

hipError_t launch_apply_f64(const icc_grid* g, const icc_op* op, const double* x, double* y, double* pap,
                            const PcgState* pst, hipStream_t st) {
  const auto& s = g->s;
  hipLaunchKernelGGL((apply_tiles<double, false, 12>), dim3(apply_grid(g)), dim3(64 * 12), 0, st, g->d_tiles, g->d_wgfirst,
                     (const void*)g->d_kfrag64, x, y,
                     op->d_scale, op->d_free, op->d_tmask, (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc,
                     (int)s.nel[0], (int)s.nel[1], pap, pst);
  return hipGetLastError();
}

hipError_t launch_apply_f32(const icc_grid* g, const icc_op* op, const float* x, float* y, double* pap,
                            const PcgState* pst, hipStream_t st) {
  const auto& s = g->s;
  hipLaunchKernelGGL((apply_tiles<float, false, 8>), dim3(apply_grid(g)), dim3(64 * 8), 0, st, g->d_tiles, g->d_wgfirst,
                     (const void*)g->d_kfrag32, x, y,
                     op->d_scale32, op->d_free, op->d_tmask, (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc,
                     (int)s.nel[0], (int)s.nel[1], pap, pst);
  return hipGetLastError();
}

hipError_t launch_apply_bf16(const icc_grid* g, const icc_op* op, const float* x, float* y, double* pap,
                             const PcgState* pst, hipStream_t st) {
  const auto& s = g->s;
  hipLaunchKernelGGL((apply_tiles<float, true, 8>), dim3(apply_grid(g)), dim3(64 * 8), 0, st, g->d_tiles,
                     g->d_wgfirst, (const void*)g->d_kfragb, x, y, op->d_scale32, op->d_free, op->d_tmask,
                     (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc, (int)s.nel[0], (int)s.nel[1], pap, pst);
  return hipGetLastError();
}

hipError_t launch_apply_batch_f64(const BatchApply& b, const double* x, double* y, const double* scale,
                                  const uint8_t* freem, double* pap, hipStream_t st) {
  hipLaunchKernelGGL((apply_tiles<double, false, 12, true>), dim3(b.G), dim3(64 * 12), 0, st, b.tiles, b.wg_first,
                     (const void*)b.kfrag64, x, y, scale, freem, b.tmask, 0, 0, 0, 0, 0, pap, nullptr, b.tinst);
  return hipGetLastError();
}

hipError_t launch_apply_batch_f32(const BatchApply& b, const float* x, float* y, const float* scale,
                                  const uint8_t* freem, double* pap, hipStream_t st) {
  hipLaunchKernelGGL((apply_tiles<float, false, 8, true>), dim3(b.G), dim3(64 * 8), 0, st, b.tiles, b.wg_first,
                     (const void*)b.kfrag32, x, y, scale, freem, b.tmask, 0, 0, 0, 0, 0, pap, nullptr, b.tinst);
  return hipGetLastError();
}

int det_grid(const icc_grid* g) {
  const int64_t nb = (g->s.ndof_sc + 255) / 256;
  const int cap = APPLY_MAX_WG_PER_CU * (g->ctx->num_cu > 0 ? g->ctx->num_cu : 256);
  return (int)std::max<int64_t>(1, std::min<int64_t>(nb, cap));
}

template <typename T>
hipError_t launch_apply_det(const icc_grid* g, const icc_op* op, const T* x, T* y, double* pap, const PcgState* pst,
                            T* E, hipStream_t st) {
  const auto& s = g->s;
  const void* kf;
  const T* sc;
  if constexpr (std::is_same<T, double>::value) {
    kf = g->d_kfrag64;
    sc = op->d_scale;
  } else {
    kf = g->d_kfrag32;
    sc = op->d_scale32;
  }
  hipLaunchKernelGGL((det_elem_kernel<T>), dim3(g->ntiles), dim3(384), 0, st, g->d_tiles, kf, x, sc, op->d_free,
                     op->d_tmask, (int)s.ncp[0], (int)s.ncp[1], (int)s.ndof_sc, (int)s.nel[0], (int)s.nel[1],
                     (int64_t)s.nel_tot, E, pst);
  hipLaunchKernelGGL((det_gather_kernel<T>), dim3(det_grid(g)), dim3(256), 0, st, E, x, y, op->d_free, (int)s.ncp[0],
                     (int)s.ncp[1], (int)s.ncp[2], (int)s.nel[0], (int)s.nel[1], (int)s.nel[2], pap, pst);
  return hipGetLastError();
}
template hipError_t launch_apply_det<double>(const icc_grid*, const icc_op*, const double*, double*, double*,
                                             const PcgState*, double*, hipStream_t);
template hipError_t launch_apply_det<float>(const icc_grid*, const icc_op*, const float*, float*, double*,
                                            const PcgState*, float*, hipStream_t);

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

hipError_t launch_layout_probe_bf16(double* out, hipStream_t st) {
  hipLaunchKernelGGL(mfma_layout_probe_bf16, dim3(1), dim3(64), 0, st, out);
  return hipGetLastError();
}

}  // namespace icc
