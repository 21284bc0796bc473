// C ABI of libicc (host side): contexts, grids, operators, solvers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <functional>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <unordered_map>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

#define ICC_API extern "C" __attribute__((visibility("default")))

namespace icc {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return ICC_EHIP;
}

template <typename T>
static int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    g_err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return ICC_ENOMEM;
  }
  return ICC_OK;
}

// BF16 bits of a float, round to nearest even (the conversion the device's
// v_cvt_pk_bf16_f32 applies to the staged x)
static uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

#define ICC_TRY(call)            \
  do {                           \
    int rc__ = (call);           \
    if (rc__ != ICC_OK) return rc__; \
  } while (0)

// Workspace of one operator: solver vectors and reduction scratch.
struct Work {
  int64_t n = 0;
  double *x = nullptr, *r = nullptr, *p = nullptr, *ap = nullptr, *dinv = nullptr, *tmp = nullptr, *f = nullptr;
  double* ubest = nullptr;  // best refinement iterate (icc_solve_ir)
  float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *ap32 = nullptr, *dinv32 = nullptr;
  double* parts = nullptr;  // 2 * VEC_MAX_BLOCKS vector partials
  double* pap = nullptr;    // apply workgroups + csr blocks
  PcgState* st = nullptr;
  PcgState* st_host = nullptr;  // pinned, 2 entries (batch lookahead)
};

static Work* work_of(icc_op* op);

}  // namespace icc

using namespace icc;

struct icc_op_ext {
  Work w;
};

// ---------------------------------------------------------------------------
ICC_API const char* icc_last_error(void) { return g_err.c_str(); }
ICC_API const char* icc_version(void) { return "libicc 0.1 (gfx950)"; }

ICC_API int icc_ctx_create(int device, icc_ctx** out) {
  if (!out) {
    set_error("icc_ctx_create: null out");
    return ICC_EINVAL;
  }
  int ndev = 0;
  ICC_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_error("icc_ctx_create: no such device");
    return ICC_EINVAL;
  }
  ICC_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  ICC_HIP(hipGetDeviceProperties(&prop, device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) {
    set_error(std::string("icc_ctx_create: built for gfx950, device is ") + prop.gcnArchName);
    return ICC_ENOTSUP;
  }
  // the no-allocation PCG's fixed workspace on this device (its failure is
  // reported by that entry point, not here)
  (void)noalloc_init(device);
  auto* c = new icc_ctx();
  c->device = device;
  c->num_cu = prop.multiProcessorCount;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = c;
  return ICC_OK;
}

ICC_API int icc_ctx_destroy(icc_ctx* ctx) {
  if (!ctx) return ICC_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return ICC_OK;
}

ICC_API int icc_ctx_sync(icc_ctx* ctx) {
  ICC_HIP(hipStreamSynchronize(ctx->stream));
  return ICC_OK;
}

ICC_API void* icc_ctx_stream(icc_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ---------------------------------------------------------------------------
namespace icc {
// Partitions the tile list for the apply grid (wg_per_cu workgroups per CU,
// a multiple of 8 so the XCD grouping applies, at most one per tile):
// balanced work (balance_tiles) or equal whole-tile counts; then uploads the
// tiles and the per-workgroup ranges.
static int upload_tiles(icc_grid* g, bool balance) {
  int G = g->wg_per_cu * (g->ctx->num_cu > 0 ? g->ctx->num_cu : 256);
  G -= G % 8;
  if (G > (int)g->s.tiles.size()) G = (int)g->s.tiles.size();
  if (G < 1) G = 1;
  std::vector<int32_t> first;
  if (balance) {
    balance_tiles(g->s, G, first);
  } else {
    first.resize(G + 1);
    for (int b = 0; b <= G; ++b) first[b] = (int32_t)(((int64_t)b * (int64_t)g->s.tiles.size()) / G);
  }
  Tile* dt = nullptr;
  int32_t* df = nullptr;
  int rc;
  if ((rc = dalloc(&dt, g->s.tiles.size())) || (rc = dalloc(&df, first.size()))) {
    hipFree(dt);
    return rc;
  }
  if (hipMemcpy(dt, g->s.tiles.data(), g->s.tiles.size() * sizeof(Tile), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(df, first.data(), first.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(dt);
    hipFree(df);
    set_error("tile upload failed");
    return ICC_EHIP;
  }
  hipFree(g->d_tiles);
  hipFree(g->d_wgfirst);
  g->d_tiles = dt;
  g->d_wgfirst = df;
  g->ntiles = (int)g->s.tiles.size();
  g->apply_G = G;
  return ICC_OK;
}
// Partitions the split-apply tiles over one workgroup per CU (balanced
// work, tiles split where needed) and uploads them.
static int upload_split_tiles(icc_grid* g) {
  int G = g->ctx->num_cu > 0 ? g->ctx->num_cu : 256;
  G -= G % 8;
  if (G > (int)g->s.stiles.size()) G = (int)g->s.stiles.size();
  if (G < 1) G = 1;
  std::vector<int32_t> first;
  partition_split_tiles(g->s.stiles, G, first);
  Tile* dt = nullptr;
  int32_t* df = nullptr;
  int rc;
  if ((rc = dalloc(&dt, g->s.stiles.size())) || (rc = dalloc(&df, first.size()))) {
    hipFree(dt);
    return rc;
  }
  if (hipMemcpy(dt, g->s.stiles.data(), g->s.stiles.size() * sizeof(Tile), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(df, first.data(), first.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(dt);
    hipFree(df);
    set_error("split tile upload failed");
    return ICC_EHIP;
  }
  hipFree(g->d_stiles);
  hipFree(g->d_swgfirst);
  g->d_stiles = dt;
  g->d_swgfirst = df;
  g->nstiles = (int)g->s.stiles.size();
  g->split_G = G;
  return ICC_OK;
}

// A fragments of the split-FP16 apply.  K_t is scaled by 2^-kexp (so that
// max |K_t| 2^-kexp lies in [2^14, 2^15), keeping the lo parts normal) and split
// into hi = fp16(v), lo = fp16(v - hi).  Lane l of (type, half h, local M
// tile mm, K step ks) holds row i = l % 16 of M tile m = 3 h + mm (the K
// row perm_row assigns to FP32 accumulator slot i = 4 kq + r) against
// trial columns k = 8 (l / 16) + j (ks < 3) or 4 (l / 16) + j (ks = 3,
// j < 4): triplet T = 8 ks + 2 (l / 16) + j / 4, slot j % 4 (ks < 3), or
// triplet 24 + l / 16, slot j (ks = 3); slot 3 and triplets past 26 are
// zero columns.
static void split_fragments(const GridSetup& s, std::vector<uint16_t>& frag, std::vector<int32_t>& kexp) {
  const int nt = (int)s.kmat.size();
  frag.assign((size_t)nt * 2 * 3 * 4 * 64 * 2 * 8, 0);
  kexp.assign(nt, 0);
  auto half_bits = [](double v) {
    const _Float16 hv = (_Float16)v;
    uint16_t b;
    std::memcpy(&b, &hv, 2);
    return b;
  };
  for (int t = 0; t < nt; ++t) {
    const double* K = s.kmat[t].data();
    double mx = 0.0;
    for (int q = 0; q < NSH * NSH; ++q) mx = std::max(mx, std::fabs(K[q]));
    int e = 0;
    if (mx > 0.0) std::frexp(mx, &e);
    e -= SP_SCALE_BITS;
    kexp[t] = e;
    for (int h = 0; h < 2; ++h)
      for (int mm = 0; mm < 3; ++mm)
        for (int ks = 0; ks < 4; ++ks)
          for (int l = 0; l < 64; ++l)
            for (int j = 0; j < (ks < 3 ? 8 : 4); ++j) {
              const int i = l % 16;
              const int row = perm_row(3 * h + mm, i / 4, i % 4);
              const int T = ks < 3 ? 8 * ks + 2 * (l / 16) + j / 4 : 24 + l / 16;
              const int slot = ks < 3 ? j % 4 : j;
              int col = -1;
              if (T < 27 && slot < 3) {
                const int c = T / 9, rem = T % 9, ay = rem % 3, az = rem / 3;
                col = c * NSC + slot + 3 * ay + 9 * az;
              }
              const double v = (row >= 0 && col >= 0) ? std::ldexp(K[row * NSH + col], -e) : 0.0;
              const _Float16 hi = (_Float16)v;
              const double lo = v - (double)hi;
              const size_t base = ((((((size_t)t * 2 + h) * 3 + mm) * 4 + ks) * 64 + l) * 2) * 8;
              frag[base + j] = half_bits(v);
              frag[base + 8 + j] = half_bits(lo);
            }
  }
}
}  // namespace icc

ICC_API int icc_element_types(const int64_t nel[3], const double bounds[6], int64_t* ntypes, int64_t* type_id) {
  if (!nel || !bounds || !ntypes) return ICC_EINVAL;
  try {
    GridSetup s;
    const int rc = build_setup(nel, bounds, 1.0, 0.3, s);
    if (rc != ICC_OK) return rc;
    *ntypes = (int64_t)s.rtype_internal.size();
    if (type_id) {
      int64_t e = 0;
      for (int64_t k = 0; k < s.nel[2]; ++k)
        for (int64_t j = 0; j < s.nel[1]; ++j)
          for (int64_t i = 0; i < s.nel[0]; ++i, ++e) type_id[e] = reference_type(s, i, j, k);
    }
  } catch (const std::exception& ex) {
    set_error(std::string("icc_element_types: ") + ex.what());
    return ICC_EINVAL;
  }
  return ICC_OK;
}

ICC_API int icc_grid_create(icc_ctx* ctx, const int64_t nel[3], const double bounds[6], int64_t degree, double young,
                            double poisson, icc_grid** out) {
  if (!ctx || !nel || !bounds || !out) {
    set_error("icc_grid_create: null argument");
    return ICC_EINVAL;
  }
  if (degree != 2) {
    set_error("icc_grid_create: only degree 2 is implemented");
    return ICC_ENOTSUP;
  }
  auto* g = new icc_grid();
  g->ctx = ctx;
  int rc = build_setup(nel, bounds, young, poisson, g->s);
  if (rc != ICC_OK) {
    delete g;
    return rc;
  }
  const GridSetup& s = g->s;
  g->ntypes = (int)s.kmat.size();
  g->ntiles = (int)s.tiles.size();
  for (int d = 0; d < 3; ++d) g->T[d] = (int)s.mats1[d].size();
  g->maxnel = (int)std::max(s.nel[0], std::max(s.nel[1], s.nel[2]));
  // MFMA A fragments: lane l of (type, mt, ks) holds A[i = l%16][k = l/16] =
  // K[row][4 ks + l/16], where row is the K row the accumulator slot i of M
  // tile mt is assigned (perm_row); slot i belongs to register r and row group
  // kq with i = 4 r + kq for v_mfma_f64_16x16x4 and i = 4 kq + r for
  // v_mfma_f32_16x16x4.
  const size_t nfrag = (size_t)g->ntypes * MT * KS * 64;
  std::vector<double> f64(nfrag, 0.0);
  std::vector<float> f32(nfrag, 0.0f);
  // BF16 fragments (v_mfma_f32_16x16x32_bf16): lane l of (type, mt, ks) holds
  // A[i = l%16][k = 32 ks + 8 (l/16) + j], j = 0..7, with the FP32 row slots
  const size_t nfragb = (size_t)g->ntypes * MT * 3 * 64 * 8;
  std::vector<uint16_t> fb(nfragb, 0);
  std::vector<double> kd((size_t)g->ntypes * NSH);
  for (int t = 0; t < g->ntypes; ++t) {
    const double* K = s.kmat[t].data();
    const double* Kq = s.kmat_q32[t].data();  // FP32: zero block row sums, exact in FP32
    for (int mt = 0; mt < MT; ++mt)
      for (int ks = 0; ks < KS; ++ks)
        for (int l = 0; l < 64; ++l) {
          const int i = l % 16, colk = 4 * ks + (l / 16);
          const int row64 = perm_row(mt, i % 4, i / 4), row32 = perm_row(mt, i / 4, i % 4);
          size_t idx = (((size_t)t * MT + mt) * KS + ks) * 64 + l;
          f64[idx] = (row64 >= 0 && colk < NSH) ? K[row64 * NSH + colk] : 0.0;
          f32[idx] = (float)((row32 >= 0 && colk < NSH) ? Kq[row32 * NSH + colk] : 0.0);
        }
    for (int mt = 0; mt < MT; ++mt)
      for (int ks = 0; ks < 3; ++ks)
        for (int l = 0; l < 64; ++l)
          for (int j = 0; j < 8; ++j) {
            const int i = l % 16, colk = 32 * ks + 8 * (l / 16) + j;
            const int row32 = perm_row(mt, i / 4, i % 4);
            const float v = (float)((row32 >= 0 && colk < NSH) ? K[row32 * NSH + colk] : 0.0);
            fb[((((size_t)t * MT + mt) * 3 + ks) * 64 + l) * 8 + j] = bf16_bits(v);
          }
    for (int a = 0; a < NSH; ++a) kd[(size_t)t * NSH + a] = K[a * NSH + a];
  }
  std::vector<int32_t> tid((size_t)3 * g->maxnel, 0);
  for (int d = 0; d < 3; ++d)
    for (int64_t e = 0; e < s.nel[d]; ++e) tid[(size_t)d * g->maxnel + e] = s.tid1[d][e];
  hipStream_t st = ctx->stream;
  if ((rc = dalloc(&g->d_kfrag64, nfrag)) || (rc = dalloc(&g->d_kfrag32, nfrag)) || (rc = dalloc(&g->d_kfragb, nfragb)) ||
      (rc = dalloc(&g->d_kdiag, kd.size())) || (rc = dalloc(&g->d_tid1, tid.size())) ||
      (rc = dalloc(&g->d_typeof, s.type_of.size()))) {
    icc_grid_destroy(g);
    return rc;
  }
  hipError_t e;
  if ((e = hipMemcpyAsync(g->d_kfrag64, f64.data(), nfrag * 8, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(g->d_kfrag32, f32.data(), nfrag * 4, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(g->d_kfragb, fb.data(), nfragb * 2, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(g->d_kdiag, kd.data(), kd.size() * 8, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(g->d_tid1, tid.data(), tid.size() * 4, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(g->d_typeof, s.type_of.data(), s.type_of.size() * 4, hipMemcpyHostToDevice, st)) ||
      (e = hipStreamSynchronize(st))) {
    icc_grid_destroy(g);
    return hip_fail(e, "icc_grid_create upload");
  }
  if ((rc = upload_tiles(g, true)) != ICC_OK) {
    icc_grid_destroy(g);
    return rc;
  }
  {
    std::vector<uint16_t> f16;
    std::vector<int32_t> kx;
    split_fragments(s, f16, kx);
    if ((rc = dalloc(&g->d_afrag16, f16.size() / 8)) || (rc = dalloc(&g->d_kexp16, kx.size()))) {
      icc_grid_destroy(g);
      return rc;
    }
    if ((e = hipMemcpy(g->d_afrag16, f16.data(), f16.size() * 2, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(g->d_kexp16, kx.data(), kx.size() * 4, hipMemcpyHostToDevice))) {
      icc_grid_destroy(g);
      return hip_fail(e, "icc_grid_create upload");
    }
    if ((rc = upload_split_tiles(g)) != ICC_OK) {
      icc_grid_destroy(g);
      return rc;
    }
  }
  *out = g;
  return ICC_OK;
}

ICC_API int icc_grid_destroy(icc_grid* g) {
  if (!g) return ICC_OK;
  hipFree(g->d_kfrag64);
  hipFree(g->d_kfrag32);
  hipFree(g->d_kfragb);
  hipFree(g->d_kdiag);
  hipFree(g->d_tid1);
  hipFree(g->d_typeof);
  hipFree(g->d_tiles);
  hipFree(g->d_wgfirst);
  hipFree(g->d_mats1);
  hipFree(g->d_wq);
  hipFree(g->d_afrag16);
  hipFree(g->d_kexp16);
  hipFree(g->d_stiles);
  hipFree(g->d_swgfirst);
  delete g;
  return ICC_OK;
}

ICC_API int icc_grid_sizes(const icc_grid* g, int64_t out[11]) {
  if (!g || !out) return ICC_EINVAL;
  const auto& s = g->s;
  out[0] = s.nel_tot;
  out[1] = s.ndof_sc;
  out[2] = s.ndof;
  out[3] = (int64_t)g->s.rtype_internal.size();  // the reference's type count
  out[4] = NSH;
  for (int d = 0; d < 3; ++d) {
    out[5 + d] = s.ncp[d];
    out[8 + d] = s.nel[d];
  }
  return ICC_OK;
}

ICC_API int icc_grid_connectivity(const icc_grid* g, int64_t* out) {
  if (!g || !out) return ICC_EINVAL;
  const auto& s = g->s;
  int64_t e = 0;
  for (int64_t k = 0; k < s.nel[2]; ++k)
    for (int64_t j = 0; j < s.nel[1]; ++j)
      for (int64_t i = 0; i < s.nel[0]; ++i, ++e) {
        int64_t* o = out + e * NSH;
        for (int c = 0; c < 3; ++c)
          for (int a = 0; a < NSC; ++a) {
            int64_t ax = a % 3, ay = (a / 3) % 3, az = a / 9;
            o[c * NSC + a] = c * s.ndof_sc + (i + ax) + s.ncp[0] * ((j + ay) + s.ncp[1] * (k + az));
          }
      }
  return ICC_OK;
}

ICC_API int icc_grid_type_id(const icc_grid* g, int64_t* out) {
  if (!g || !out) return ICC_EINVAL;
  const auto& s = g->s;
  int64_t e = 0;
  for (int64_t k = 0; k < s.nel[2]; ++k)
    for (int64_t j = 0; j < s.nel[1]; ++j)
      for (int64_t i = 0; i < s.nel[0]; ++i, ++e) out[e] = reference_type(s, i, j, k);
  return ICC_OK;
}

ICC_API int icc_grid_element_matrix(const icc_grid* g, int64_t type, double* out) {
  if (!g || !out || type < 0 || type >= (int64_t)g->s.rtype_internal.size()) return ICC_EINVAL;
  std::memcpy(out, g->s.kmat[g->s.rtype_internal[type]].data(), sizeof(double) * NSH * NSH);
  return ICC_OK;
}

// ---------------------------------------------------------------------------
namespace icc {
static Work* work_of(icc_op* op);
}

ICC_API int icc_op_create(icc_grid* g, icc_op** out) {
  if (!g || !out) return ICC_EINVAL;
  auto* op = new icc_op();
  op->g = g;
  const auto& s = g->s;
  int rc;
  if ((rc = dalloc(&op->d_scale, s.nel_tot)) || (rc = dalloc(&op->d_scale32, s.nel_tot))) {
    icc_op_destroy(op);
    return rc;
  }
  hipStream_t st = g->ctx->stream;
  hipError_t e;
  if ((e = launch_fill<double>(op->d_scale, 1.0, s.nel_tot, st)) ||
      (e = launch_fill<float>(op->d_scale32, 1.0f, s.nel_tot, st))) {
    icc_op_destroy(op);
    return hip_fail(e, "icc_op_create");
  }
  g->ops.push_back(op);
  *out = op;
  return ICC_OK;
}

namespace icc {
static void free_work(Work& w) {
  hipFree(w.x);
  hipFree(w.r);
  hipFree(w.p);
  hipFree(w.ap);
  hipFree(w.dinv);
  hipFree(w.tmp);
  hipFree(w.f);
  hipFree(w.ubest);
  hipFree(w.x32);
  hipFree(w.r32);
  hipFree(w.p32);
  hipFree(w.ap32);
  hipFree(w.dinv32);
  hipFree(w.parts);
  hipFree(w.pap);
  hipFree(w.st);
  if (w.st_host) hipHostFree(w.st_host);
  w = Work();
}
// The workspace is owned by the op (created on first use, freed with it), so
// contexts driven from different host threads share no table.
static Work* work_of(icc_op* op) {
  if (op->work == nullptr) op->work = new Work();
  return static_cast<Work*>(op->work);
}
static void drop_work(icc_op* op) {
  if (op->work == nullptr) return;
  Work* w = static_cast<Work*>(op->work);
  free_work(*w);
  delete w;
  op->work = nullptr;
}

static int csr_blocks(const icc_op* op) { return op->s_nnz > 0 ? vec_blocks(op->g->s.ndof) : 0; }
// y = S x (overwrites y) in precision T, csr_blocks(op) partials into parts;
// y_zeroed: y is known to be zero already (rows without entries left alone)
template <typename T>
static hipError_t op_csr(const icc_op* op, const T* x, T* y, double* parts, hipStream_t st, bool y_zeroed = false) {
  const int64_t n = op->g->s.ndof;
  if constexpr (std::is_same<T, double>::value)
    return launch_csr_spmv<double, double>(n, op->d_s_indptr, op->d_s_indices, op->d_s_data, op->d_s_rows,
                                           op->s_nrows, csr_blocks(op), x, y, op->d_free, parts, y_zeroed, st);
  else
    return launch_csr_spmv<float, float>(n, op->d_s_indptr, op->d_s_indices, op->d_s_data32, op->d_s_rows,
                                         op->s_nrows, csr_blocks(op), x, y, op->d_free, parts, y_zeroed, st);
}
// partials of the element term: one per apply workgroup, or per gather block
// in the deterministic mode
// (the split-FP16 apply, prec = ICC_F16X, has its own tiles and one partial per
// split-tile workgroup, so the layout below depends on the precision)
static int elem_grid(const icc_op* op, int prec = ICC_F64) {
  if (op->det) return det_grid(op->g);
  return prec == ICC_F16X ? op->g->split_G : apply_grid(op->g);
}
// p.Ap partial layout of one apply: [elem_grid | csr_blocks | face_blocks]
static int npap_of(const icc_op* op, int prec = ICC_F64) {
  return elem_grid(op, prec) + csr_blocks(op) + face_blocks(op);
}

// Lazily allocates the solver workspace of op.
static int ensure_work(icc_op* op, Work*& out) {
  Work* w = work_of(op);
  const int64_t n = op->g->s.ndof;
  if (w->n != n) {
    free_work(*w);
    w->n = n;
    int rc;
    size_t npap = (size_t)APPLY_MAX_WG_PER_CU * (op->g->ctx->num_cu > 0 ? op->g->ctx->num_cu : 256) + VEC_MAX_BLOCKS +
                  FACE_MAX_BLOCKS;
    if ((rc = dalloc(&w->x, n)) || (rc = dalloc(&w->r, n)) || (rc = dalloc(&w->p, n)) || (rc = dalloc(&w->ap, n)) ||
        (rc = dalloc(&w->dinv, n)) || (rc = dalloc(&w->tmp, n)) || (rc = dalloc(&w->f, n)) || (rc = dalloc(&w->ubest, n)) ||
        (rc = dalloc(&w->x32, n)) || (rc = dalloc(&w->r32, n)) || (rc = dalloc(&w->p32, n)) ||
        (rc = dalloc(&w->ap32, n)) || (rc = dalloc(&w->dinv32, n)) || (rc = dalloc(&w->parts, 2 * VEC_MAX_BLOCKS)) ||
        (rc = dalloc(&w->pap, npap)) || (rc = dalloc(&w->st, 1))) {
      free_work(*w);
      return rc;
    }
    if (hipHostMalloc((void**)&w->st_host, 2 * sizeof(PcgState), hipHostMallocDefault) != hipSuccess) {
      free_work(*w);
      set_error("hipHostMalloc failed");
      return ICC_ENOMEM;
    }
    PcgState z;
    std::memset(&z, 0, sizeof(z));
    ICC_HIP(hipMemcpyAsync(w->st, &z, sizeof(z), hipMemcpyHostToDevice, op->g->ctx->stream));
    ICC_HIP(hipStreamSynchronize(op->g->ctx->stream));
  }
  out = w;
  return ICC_OK;
}
}  // namespace icc

ICC_API int icc_op_destroy(icc_op* op) {
  if (!op) return ICC_OK;
  if (op->g) {
    auto& v = op->g->ops;
    v.erase(std::remove(v.begin(), v.end(), op), v.end());
  }
  drop_work(op);
  hipFree(op->d_scale);
  hipFree(op->d_scale32);
  hipFree(op->d_free);
  hipFree(op->d_tmask);
  hipFree(op->d_stmask);
  hipFree(op->d_s_indptr);
  hipFree(op->d_s_indices);
  hipFree(op->d_s_data);
  hipFree(op->d_s_data32);
  hipFree(op->d_s_diag);
  hipFree(op->d_s_rows);
  hipFree(op->d_E);
  hipFree(op->d_faces);
  hipFree(op->d_f_diag);
  hipFree(op->d_ftab);
  hipFree(op->d_fblk);
  hipFree(op->d_fptr);
  delete op;
  return ICC_OK;
}

ICC_API int icc_op_set_scale_dev(icc_op* op, const double* s_dev) {
  if (!op || !s_dev) return ICC_EINVAL;
  const int64_t nel = op->g->s.nel_tot;
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(op->d_scale, s_dev, nel * 8, hipMemcpyDeviceToDevice, st));
  ICC_HIP(launch_to_f32(op->d_scale, op->d_scale32, nel, st));
  return ICC_OK;
}

ICC_API int icc_op_set_scale(icc_op* op, const double* s) {
  if (!op || !s) return ICC_EINVAL;
  const int64_t nel = op->g->s.nel_tot;
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(op->d_scale, s, nel * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(launch_to_f32(op->d_scale, op->d_scale32, nel, st));
  ICC_HIP(hipStreamSynchronize(st));
  return ICC_OK;
}

namespace icc {
// Per tile: does any staged point carry a fixed DOF?  Only such tiles read
// the mask inside the apply kernel.  Rebuilt whenever the grid is re-tiled.
// 32-bit flags (scalar loads in the kernels): 1 if the tile stages a fixed DOF
static std::vector<int32_t> tile_mask_of(const GridSetup& s, const std::vector<Tile>& tiles, const uint8_t* fm) {
  std::vector<int32_t> tm(tiles.size(), 0);
  for (size_t ti = 0; ti < tiles.size(); ++ti) {
    const Tile& t = tiles[ti];
    bool fixed = false;
    for (int c = 0; c < 3 && !fixed; ++c)
      for (int64_t k = t.k0; k < t.k0 + t.nz + 2 && !fixed; ++k)
        for (int64_t j = t.j0; j < t.j0 + t.ny + 2 && !fixed; ++j)
          for (int64_t i = t.i0; i < t.i0 + t.nx + 2; ++i)
            if (!fm[c * s.ndof_sc + i + s.ncp[0] * (j + s.ncp[1] * k)]) {
              fixed = true;
              break;
            }
    tm[ti] = fixed ? 1 : 0;
  }
  return tm;
}

static int upload_tile_mask(icc_op* op) {
  const GridSetup& s = op->g->s;
  const uint8_t* fm = op->h_free.data();
  const std::vector<int32_t> tm = tile_mask_of(s, s.tiles, fm), stm = tile_mask_of(s, s.stiles, fm);
  hipFree(op->d_tmask);
  hipFree(op->d_stmask);
  op->d_tmask = nullptr;
  op->d_stmask = nullptr;
  ICC_TRY(dalloc(&op->d_tmask, (int64_t)tm.size()));
  ICC_TRY(dalloc(&op->d_stmask, (int64_t)stm.size()));
  ICC_HIP(hipMemcpyAsync(op->d_tmask, tm.data(), tm.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                         op->g->ctx->stream));
  ICC_HIP(hipMemcpyAsync(op->d_stmask, stm.data(), stm.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                         op->g->ctx->stream));
  ICC_HIP(hipStreamSynchronize(op->g->ctx->stream));
  return ICC_OK;
}
}  // namespace icc

ICC_API int icc_op_set_free_mask(icc_op* op, const uint8_t* fm) {
  if (!op) return ICC_EINVAL;
  const int64_t n = op->g->s.ndof;
  if (!fm) {
    hipFree(op->d_free);
    hipFree(op->d_tmask);
    hipFree(op->d_stmask);
    op->d_free = nullptr;
    op->d_tmask = nullptr;
    op->d_stmask = nullptr;
    op->h_free.clear();
    return ICC_OK;
  }
  op->h_free.assign(fm, fm + n);
  if (!op->d_free) ICC_TRY(dalloc(&op->d_free, n));
  ICC_HIP(hipMemcpyAsync(op->d_free, fm, n, hipMemcpyHostToDevice, op->g->ctx->stream));
  return upload_tile_mask(op);
}

// Deterministic mode: the element term of every FP64 / FP32 apply of op (and
// of its PCG iterations) runs as per-element outputs plus a fixed-order
// gather instead of the atomic tile scatter, so applies and solves repeat
// bit for bit.  Needs nel * 81 doubles of device memory; the face
// stabilisation (atomic) is not covered.
ICC_API int icc_op_set_deterministic(icc_op* op, int on) {
  if (!op || (on != 0 && on != 1)) return ICC_EINVAL;
  if (on && op->f_n > 0) {
    set_error("icc_op_set_deterministic: the face stabilisation has no deterministic mode");
    return ICC_ENOTSUP;
  }
  if (on && op->d_E == nullptr) {
    double* e = nullptr;
    ICC_TRY(dalloc(&e, (size_t)op->g->s.nel_tot * NSH));
    op->d_E = e;
  }
  if (!on) {
    (void)hipFree(op->d_E);
    op->d_E = nullptr;
  }
  op->det = on;
  return ICC_OK;
}

ICC_API int icc_grid_set_tiling(icc_grid* g, int variant, int wg_per_cu) {
  if (!g || variant < 0 || variant >= n_tile_variants() || (wg_per_cu != 1 && wg_per_cu != 2 && wg_per_cu != 4))
    return ICC_EINVAL;
  ICC_HIP(hipStreamSynchronize(g->ctx->stream));
  build_tiles(g->s, variant);
  g->wg_per_cu = wg_per_cu;
  ICC_TRY(upload_tiles(g, variant < 6));
  g->tile_variant = variant;
  for (icc_op* op : g->ops)
    if (op->d_free) ICC_TRY(upload_tile_mask(op));
  return ICC_OK;
}

// Element groups (x-lines of <= 16 elements, or 16-element groups) per tile
// of the split-FP16 apply: fewer groups give more, smaller tiles per CU (a
// deeper pipeline) at more staged halo points per element.
ICC_API int icc_grid_set_split_tiling(icc_grid* g, int max_groups) {
  if (!g || max_groups < 1 || max_groups > SP_MAX_GROUPS) return ICC_EINVAL;
  ICC_HIP(hipStreamSynchronize(g->ctx->stream));
  g->s.split_max_groups = max_groups;
  build_split_tiles(g->s);
  ICC_TRY(upload_split_tiles(g));
  for (icc_op* op : g->ops)
    if (op->d_free) ICC_TRY(upload_tile_mask(op));
  return ICC_OK;
}

ICC_API int icc_grid_tiling(const icc_grid* g, int64_t out[3]) {
  if (!g || !out) return ICC_EINVAL;
  out[0] = g->tile_variant;
  out[1] = g->wg_per_cu;
  out[2] = g->ntiles;
  return ICC_OK;
}

ICC_API int icc_op_set_sparse(icc_op* op, int64_t nrows, const int64_t* indptr, const int64_t* indices,
                              const double* data) {
  if (!op) return ICC_EINVAL;
  const int64_t n = op->g->s.ndof;
  hipFree(op->d_s_indptr);
  hipFree(op->d_s_indices);
  hipFree(op->d_s_data);
  hipFree(op->d_s_data32);
  hipFree(op->d_s_diag);
  hipFree(op->d_s_rows);
  op->d_s_indptr = nullptr;
  op->d_s_indices = nullptr;
  op->d_s_data = nullptr;
  op->d_s_data32 = nullptr;
  op->d_s_diag = nullptr;
  op->d_s_rows = nullptr;
  op->s_nnz = 0;
  op->s_nrows = 0;
  op->h_s_indptr.clear();
  op->h_s_indices.clear();
  op->h_s_data.clear();
  if (nrows == 0 || !indptr) return ICC_OK;
  if (nrows != n) {
    set_error("icc_op_set_sparse: nrows must equal ndof");
    return ICC_EINVAL;
  }
  const int64_t nnz = indptr[n];
  std::vector<int32_t> ip(n + 1), ix(nnz);
  std::vector<float> d32(nnz);
  std::vector<double> diag(n, 0.0);
  for (int64_t i = 0; i <= n; ++i) ip[i] = (int32_t)indptr[i];
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = indptr[i]; k < indptr[i + 1]; ++k) {
      if (indices[k] < 0 || indices[k] >= n) {
        set_error("icc_op_set_sparse: column index out of range");
        return ICC_EINVAL;
      }
      ix[k] = (int32_t)indices[k];
      d32[k] = (float)data[k];
      if (indices[k] == i) diag[i] += data[k];
    }
  // {row, first entry, end, 0} of every row with entries
  std::vector<int32_t> rows;
  for (int64_t i = 0; i < n; ++i)
    if (indptr[i + 1] > indptr[i]) rows.insert(rows.end(), {(int32_t)i, ip[i], ip[i + 1], 0});
  ICC_TRY(dalloc(&op->d_s_rows, std::max<int64_t>((int64_t)rows.size(), 4)));
  ICC_TRY(dalloc(&op->d_s_indptr, n + 1));
  ICC_TRY(dalloc(&op->d_s_indices, nnz));
  ICC_TRY(dalloc(&op->d_s_data, nnz));
  ICC_TRY(dalloc(&op->d_s_data32, nnz));
  ICC_TRY(dalloc(&op->d_s_diag, n));
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(op->d_s_indptr, ip.data(), (n + 1) * 4, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(op->d_s_indices, ix.data(), nnz * 4, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(op->d_s_data, data, nnz * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(op->d_s_data32, d32.data(), nnz * 4, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(op->d_s_diag, diag.data(), n * 8, hipMemcpyHostToDevice, st));
  if (!rows.empty())
    ICC_HIP(hipMemcpyAsync(op->d_s_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, st));
  ICC_HIP(hipStreamSynchronize(st));
  op->s_nnz = nnz;
  op->s_nrows = (int64_t)rows.size() / 4;
  op->h_s_indptr.assign(indptr, indptr + n + 1);
  op->h_s_indices.assign(indices, indices + nnz);
  op->h_s_data.assign(data, data + nnz);
  return ICC_OK;
}

namespace icc {
static void clear_faces(icc_op* op) {
  hipFree(op->d_faces);
  hipFree(op->d_f_diag);
  hipFree(op->d_ftab);
  hipFree(op->d_fblk);
  hipFree(op->d_fptr);
  op->d_faces = nullptr;
  op->d_f_diag = nullptr;
  op->d_ftab = nullptr;
  op->d_fblk = nullptr;
  op->d_fptr = nullptr;
  op->f_n = 0;
  op->f_nblk = 0;
  op->h_faces.clear();
  op->h_ftab.clear();
  op->f_kind = 0;
  op->f_alpha = 0.0;
}

// Faces between face-neighbouring cells that are both active (INSIDE or CUT)
// with at least one CUT (immersed.rs:175-187), as e1 * 4 + dir in increasing
// e1, so neighbouring threads gather neighbouring control points.
static int stab_faces(const GridSetup& s, const int8_t* status, std::vector<int32_t>& faces) {
  const int64_t n0 = s.nel[0], n1 = s.nel[1], n2 = s.nel[2];
  if (s.nel_tot * 4 > (int64_t)INT32_MAX) {
    set_error("face stabilisation: grid too large for the face encoding");
    return ICC_EINVAL;
  }
  auto act = [&](int64_t e) { return status[e] == 1 || status[e] == -1; };
  faces.clear();
  for (int64_t k = 0; k < n2; ++k)
    for (int64_t j = 0; j < n1; ++j)
      for (int64_t i = 0; i < n0; ++i) {
        const int64_t e1 = i + n0 * (j + n1 * k);
        if (!act(e1)) continue;
        const bool in_range[3] = {i + 1 < n0, j + 1 < n1, k + 1 < n2};
        const int64_t step[3] = {1, n0, n0 * n1};
        for (int d = 0; d < 3; ++d) {
          if (!in_range[d]) continue;
          const int64_t e2 = e1 + step[d];
          if (act(e2) && (status[e1] == -1 || status[e2] == -1)) faces.push_back((int32_t)(e1 * 4 + d));
        }
      }
  if ((int64_t)faces.size() * 27 > (int64_t)INT32_MAX) {
    set_error("face stabilisation: too many faces");
    return ICC_EINVAL;
  }
  return ICC_OK;
}

// Uploads the faces grouped by FACE_BLK^3 element block of e1 (block index x
// fastest; inside a block by direction, then e1) with the block list and
// the face range of each block.
static int upload_faces(icc_op* op, int kind, double coef, const std::vector<int32_t>& faces_in,
                        const std::vector<double>& dg) {
  if (op->det && !faces_in.empty()) {
    set_error("the face stabilisation has no deterministic mode (icc_op_set_deterministic)");
    return ICC_ENOTSUP;
  }
  const GridSetup& s = op->g->s;
  const int64_t nbx = (s.nel[0] + FACE_BLK - 1) / FACE_BLK, nby = (s.nel[1] + FACE_BLK - 1) / FACE_BLK;
  auto block_of = [&](int32_t code) {
    const int64_t e1 = code >> 2;
    const int64_t i = e1 % s.nel[0], j = (e1 / s.nel[0]) % s.nel[1], k = e1 / (s.nel[0] * s.nel[1]);
    return (i / FACE_BLK) + nbx * ((j / FACE_BLK) + nby * (k / FACE_BLK));
  };
  std::vector<int32_t> faces(faces_in);
  std::stable_sort(faces.begin(), faces.end(), [&](int32_t a, int32_t b) {
    const int64_t ba = block_of(a), bb = block_of(b);
    if (ba != bb) return ba < bb;
    if ((a & 3) != (b & 3)) return (a & 3) < (b & 3);
    return a < b;
  });
  // the device gets the faces as packed element coordinates and direction
  // (ex | ey << 10 | ez << 20 | dir << 30) and the blocks as packed block
  // coordinates, so that the kernel decodes them without divisions
  if (s.nel[0] > 1024 || s.nel[1] > 1024 || s.nel[2] > 1024) {
    set_error("face stabilisation: at most 1024 elements per direction");
    return ICC_ENOTSUP;
  }
  std::vector<int32_t> blk, ptr, dev(faces.size());
  int64_t last_b = -1;
  for (size_t f = 0; f < faces.size(); ++f) {
    const int64_t b = block_of(faces[f]);
    const int64_t e1 = faces[f] >> 2;
    const int64_t i = e1 % s.nel[0], j = (e1 / s.nel[0]) % s.nel[1], k = e1 / (s.nel[0] * s.nel[1]);
    dev[f] = (int32_t)((uint32_t)i | ((uint32_t)j << 10) | ((uint32_t)k << 20) | ((uint32_t)(faces[f] & 3) << 30));
    if (b != last_b) {
      last_b = b;
      const int64_t bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
      blk.push_back((int32_t)(bx | (by << 10) | (bz << 20)));
      ptr.push_back((int32_t)f);
    }
  }
  ptr.push_back((int32_t)faces.size());
  ICC_TRY(dalloc(&op->d_faces, (int64_t)faces.size()));
  ICC_TRY(dalloc(&op->d_fblk, (int64_t)blk.size()));
  ICC_TRY(dalloc(&op->d_fptr, (int64_t)ptr.size()));
  ICC_TRY(dalloc(&op->d_f_diag, s.ndof));
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(op->d_faces, dev.data(), dev.size() * 4, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(op->d_fblk, blk.data(), blk.size() * 4, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(op->d_fptr, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(op->d_f_diag, dg.data(), s.ndof * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipStreamSynchronize(st));
  op->f_n = (int64_t)faces.size();
  op->f_nblk = (int64_t)blk.size();
  op->h_faces = faces_in;
  op->f_kind = kind;
  op->f_alpha = coef;
  return ICC_OK;
}

// Strides of a face's 4 x 3 x 3 slab: along the normal, the first and the
// second tangential direction (x < y < z order).
static void slab_strides(const GridSetup& s, int dir, int64_t& st, int64_t& sa, int64_t& sb) {
  const int64_t sy = s.ncp[0], sz = s.ncp[0] * s.ncp[1];
  st = dir == 0 ? 1 : (dir == 1 ? sy : sz);
  sa = dir == 0 ? sy : 1;
  sb = dir == 2 ? sy : sz;
}
}  // namespace icc

// Legacy cut-cell stabilisation (rust/src/immersed.rs:168-222) as a
// matrix-free face term: alpha = gamma * mean(h) and, per face and
// component, the 9 springs between the control points owned by one cell only
// (the reference's shared-point block sums to zero entry by entry and is
// left out).  The diagonal (alpha at both ends of every spring) is
// accumulated on the host.
ICC_API int icc_op_set_legacy_stabilization(icc_op* op, const int8_t* status, double gamma, int64_t* nfaces) {
  if (!op) return ICC_EINVAL;
  clear_faces(op);
  if (nfaces) *nfaces = 0;
  if (!status) return ICC_OK;
  const GridSetup& s = op->g->s;
  std::vector<int32_t> faces;
  ICC_TRY(stab_faces(s, status, faces));
  const double h0 = s.len[0] / (double)s.nel[0], h1 = s.len[1] / (double)s.nel[1], h2 = s.len[2] / (double)s.nel[2];
  const double alpha = gamma * ((h0 + h1 + h2) / 3.0);
  if (faces.empty() || alpha == 0.0) return ICC_OK;
  std::vector<double> dg(s.ndof, 0.0);
  for (int32_t code : faces) {
    const int64_t e1 = code >> 2;
    const int dir = code & 3;
    const int64_t i = e1 % s.nel[0], j = (e1 / s.nel[0]) % s.nel[1], k = e1 / (s.nel[0] * s.nel[1]);
    int64_t st, sa, sb;
    slab_strides(s, dir, st, sa, sb);
    for (int c = 0; c < 3; ++c) {
      const int64_t base = c * s.ndof_sc + i + s.ncp[0] * (j + s.ncp[1] * k);
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) {
          dg[base + a * sa + b * sb] += alpha;
          dg[base + 3 * st + a * sa + b * sb] += alpha;
        }
    }
  }
  ICC_TRY(upload_faces(op, FACE_LEGACY, alpha, faces, dg));
  if (nfaces) *nfaces = op->f_n;
  return ICC_OK;
}

// Consistent ghost penalty (rust/src/immersed.rs:234-370) as a matrix-free
// face term: coef = gamma * young * mean(h)^3, per direction the jump tables
// of the second derivative and the 1D element mass matrices (k_faces.hip
// applies coef (j j^T) (x) M_a (x) M_b per face and component); the diagonal
// coef j_t^2 M_a[a][a] M_b[b][b] is accumulated on the host.
ICC_API int icc_op_set_ghost_penalty(icc_op* op, const int8_t* status, double gamma, double young, int64_t* nfaces) {
  if (!op) return ICC_EINVAL;
  clear_faces(op);
  if (nfaces) *nfaces = 0;
  if (!status) return ICC_OK;
  const GridSetup& s = op->g->s;
  std::vector<int32_t> faces;
  ICC_TRY(stab_faces(s, status, faces));
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = s.len[d] / (double)s.nel[d];
  const double hc = (h[0] + h[1] + h[2]) / 3.0;
  const double coef = gamma * young * (hc * hc * hc);
  if (faces.empty() || coef == 0.0) return ICC_OK;
  std::vector<double> jt[3], mt[3], tab;
  int joff[3], moff[3];
  for (int d = 0; d < 3; ++d) {
    ghost_tables_1d(s.nel[d], s.len[d], jt[d], mt[d]);
    joff[d] = (int)tab.size();
    tab.insert(tab.end(), jt[d].begin(), jt[d].end());
  }
  const int mass0 = (int)tab.size();
  for (int d = 0; d < 3; ++d) {
    moff[d] = (int)tab.size() - mass0;
    tab.insert(tab.end(), mt[d].begin(), mt[d].end());
  }
  std::vector<double> dg(s.ndof, 0.0);
  for (int32_t code : faces) {
    const int64_t e1 = code >> 2;
    const int dir = code & 3;
    const int64_t e[3] = {e1 % s.nel[0], (e1 / s.nel[0]) % s.nel[1], e1 / (s.nel[0] * s.nel[1])};
    const int ta = dir == 0 ? 1 : 0, tb = dir == 2 ? 1 : 2;
    const double* jv = &jt[dir][4 * (e[dir] + 1)];
    const double* Ma = &mt[ta][9 * e[ta]];
    const double* Mb = &mt[tb][9 * e[tb]];
    int64_t st, sa, sb;
    slab_strides(s, dir, st, sa, sb);
    for (int c = 0; c < 3; ++c) {
      const int64_t base = c * s.ndof_sc + e[0] + s.ncp[0] * (e[1] + s.ncp[1] * e[2]);
      for (int t = 0; t < 4; ++t)
        for (int b = 0; b < 3; ++b)
          for (int a = 0; a < 3; ++a)
            dg[base + t * st + a * sa + b * sb] += coef * jv[t] * jv[t] * Ma[4 * a] * Mb[4 * b];
    }
  }
  ICC_TRY(dalloc(&op->d_ftab, (int64_t)tab.size()));
  ICC_HIP(hipMemcpy(op->d_ftab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
  for (int d = 0; d < 3; ++d) {
    op->ftab_joff[d] = joff[d];
    op->ftab_moff[d] = moff[d];
  }
  op->ftab_mass = mass0;
  op->h_ftab = tab;
  ICC_TRY(upload_faces(op, FACE_GHOST, coef, faces, dg));
  if (nfaces) *nfaces = op->f_n;
  return ICC_OK;
}

// Frobenius norm of the assembled face term (the reference reports the
// norm of its stabilisation matrix, immersed.rs:497): the face blocks are
// summed entry by entry on the host, then squared.
ICC_API int icc_op_face_frobenius(const icc_op* op, double* fro) {
  if (!op || !fro) return ICC_EINVAL;
  *fro = 0.0;
  if (op->f_n <= 0) return ICC_OK;
  const GridSetup& s = op->g->s;
  const int64_t n0 = s.nel[0], n1 = s.nel[1], n = s.ndof;
  std::unordered_map<int64_t, double> acc;
  acc.reserve((size_t)op->f_n * 64);
  auto add = [&](int64_t r, int64_t c, double v) { acc[r * n + c] += v; };
  for (int32_t code : op->h_faces) {
    const int64_t e1 = code >> 2;
    const int dir = code & 3;
    const int64_t e[3] = {e1 % n0, (e1 / n0) % n1, e1 / (n0 * n1)};
    int64_t st, sa, sb;
    slab_strides(s, dir, st, sa, sb);
    for (int c = 0; c < 3; ++c) {
      const int64_t base = c * s.ndof_sc + e[0] + s.ncp[0] * (e[1] + s.ncp[1] * e[2]);
      if (op->f_kind == FACE_LEGACY) {
        for (int b = 0; b < 3; ++b)
          for (int a = 0; a < 3; ++a) {
            const int64_t p = base + a * sa + b * sb, q = p + 3 * st;
            add(p, p, op->f_alpha);
            add(q, q, op->f_alpha);
            add(p, q, -op->f_alpha);
            add(q, p, -op->f_alpha);
          }
      } else {
        const int ta = dir == 0 ? 1 : 0, tb = dir == 2 ? 1 : 2;
        const double* jv = &op->h_ftab[op->ftab_joff[dir] + 4 * (e[dir] + 1)];
        const double* Ma = &op->h_ftab[op->ftab_mass + op->ftab_moff[ta] + 9 * e[ta]];
        const double* Mb = &op->h_ftab[op->ftab_mass + op->ftab_moff[tb] + 9 * e[tb]];
        for (int t1 = 0; t1 < 4; ++t1)
          for (int b1 = 0; b1 < 3; ++b1)
            for (int a1 = 0; a1 < 3; ++a1)
              for (int t2 = 0; t2 < 4; ++t2)
                for (int b2 = 0; b2 < 3; ++b2)
                  for (int a2 = 0; a2 < 3; ++a2)
                    add(base + t1 * st + a1 * sa + b1 * sb, base + t2 * st + a2 * sa + b2 * sb,
                        op->f_alpha * jv[t1] * jv[t2] * Ma[3 * a1 + a2] * Mb[3 * b1 + b2]);
      }
    }
  }
  double ss = 0.0;
  for (auto& kv : acc) ss += kv.second * kv.second;
  *fro = std::sqrt(ss);
  return ICC_OK;
}

// ---------------------------------------------------------------------------
namespace icc {
// y = A x on the op stream; y zero-initialised here unless the sparse term
// overwrites it first.  pap != null: per-tile / per-row partials of x.Ax.
template <typename T>
static hipError_t apply_t(icc_op* op, const T* x, T* y, double* pap) {
  icc_grid* g = op->g;
  hipStream_t st = g->ctx->stream;
  hipError_t e;
  const int64_t n = g->s.ndof;
  if (op->s_nnz > 0) {
    if constexpr (std::is_same<T, double>::value)
      e = op_csr<double>(op, x, y, pap ? pap + elem_grid(op) : nullptr, st);
    else
      e = op_csr<float>(op, x, y, pap ? pap + elem_grid(op) : nullptr, st);
  } else {
    e = launch_fill<T>(y, T(0), n, st);
  }
  if (e == hipSuccess) e = launch_faces<T>(op, x, y, pap ? pap + elem_grid(op) + csr_blocks(op) : nullptr, nullptr, st);
  if (e != hipSuccess) return e;
  if (op->det) return launch_apply_det<T>(g, op, x, y, pap, nullptr, (T*)op->d_E, st);
  if constexpr (std::is_same<T, double>::value)
    return launch_apply_f64(g, op, x, y, pap, nullptr, st);
  else
    return launch_apply_f32(g, op, x, y, pap, nullptr, st);
}
// Element contraction of a reduced-precision mode (BF16 or split FP16) on
// FP32 vectors, accumulating into y.
static hipError_t launch_apply_lowp(int prec, const icc_grid* g, const icc_op* op, const float* x, float* y,
                                    double* pap, const PcgState* pst, hipStream_t st) {
  return prec == ICC_F16X ? launch_apply_split(g, op, x, y, pap, pst, st) : launch_apply_bf16(g, op, x, y, pap, pst, st);
}
// y = A x with the element contraction in BF16 or split FP16 (FP32 vectors;
// sparse and face terms in FP32)
static hipError_t apply_lowp(icc_op* op, int prec, const float* x, float* y) {
  icc_grid* g = op->g;
  hipStream_t st = g->ctx->stream;
  hipError_t e;
  const int64_t n = g->s.ndof;
  if (op->s_nnz > 0)
    e = op_csr<float>(op, x, y, nullptr, st);
  else
    e = launch_fill<float>(y, 0.0f, n, st);
  if (e == hipSuccess) e = launch_faces<float>(op, x, y, nullptr, nullptr, st);
  if (e != hipSuccess) return e;
  return launch_apply_lowp(prec, g, op, x, y, nullptr, nullptr, st);
}
}  // namespace icc

ICC_API int icc_apply_dev(icc_op* op, int prec, const void* x, void* y) {
  if (!op || !x || !y) return ICC_EINVAL;
  hipError_t e;
  if (prec == ICC_F64)
    e = apply_t<double>(op, (const double*)x, (double*)y, nullptr);
  else if (prec == ICC_F32)
    e = apply_t<float>(op, (const float*)x, (float*)y, nullptr);
  else if (prec == ICC_BF16 || prec == ICC_F16X)
    e = apply_lowp(op, prec, (const float*)x, (float*)y);
  else {
    set_error("icc_apply: unknown precision");
    return ICC_EINVAL;
  }
  if (e != hipSuccess) return hip_fail(e, "icc_apply_dev");
  return ICC_OK;
}

ICC_API int icc_apply_accumulate_dev(icc_op* op, int prec, const void* x, void* y) {
  if (!op || !x || !y) return ICC_EINVAL;
  if (op->s_nnz > 0 || op->f_n > 0) {
    set_error("icc_apply_accumulate_dev: operators with a sparse term are not supported");
    return ICC_ENOTSUP;
  }
  hipError_t e;
  if (prec == ICC_F64)
    e = launch_apply_f64(op->g, op, (const double*)x, (double*)y, nullptr, nullptr, op->g->ctx->stream);
  else if (prec == ICC_F32)
    e = launch_apply_f32(op->g, op, (const float*)x, (float*)y, nullptr, nullptr, op->g->ctx->stream);
  else if (prec == ICC_BF16 || prec == ICC_F16X)
    e = launch_apply_lowp(prec, op->g, op, (const float*)x, (float*)y, nullptr, nullptr, op->g->ctx->stream);
  else {
    set_error("icc_apply_accumulate_dev: unknown precision");
    return ICC_EINVAL;
  }
  if (e != hipSuccess) return hip_fail(e, "icc_apply_accumulate_dev");
  return ICC_OK;
}

ICC_API int icc_apply(icc_op* op, int prec, const double* x, double* y) {
  if (!op || !x || !y) return ICC_EINVAL;
  Work* w;
  ICC_TRY(ensure_work(op, w));
  const int64_t n = op->g->s.ndof;
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(w->tmp, x, n * 8, hipMemcpyHostToDevice, st));
  if (prec == ICC_F64) {
    ICC_HIP(apply_t<double>(op, w->tmp, w->ap, nullptr));
    ICC_HIP(hipMemcpyAsync(y, w->ap, n * 8, hipMemcpyDeviceToHost, st));
  } else if (prec == ICC_F32 || prec == ICC_BF16 || prec == ICC_F16X) {
    ICC_HIP(launch_to_f32(w->tmp, w->x32, n, st));
    if (prec == ICC_F32)
      ICC_HIP(apply_t<float>(op, w->x32, w->ap32, nullptr));
    else
      ICC_HIP(apply_lowp(op, prec, w->x32, w->ap32));
    std::vector<float> h(n);
    ICC_HIP(hipMemcpyAsync(h.data(), w->ap32, n * 4, hipMemcpyDeviceToHost, st));
    ICC_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i) y[i] = h[i];
    return ICC_OK;
  } else {
    set_error("icc_apply: unknown precision");
    return ICC_EINVAL;
  }
  ICC_HIP(hipStreamSynchronize(st));
  return ICC_OK;
}

ICC_API int icc_diag_dev(icc_op* op, double* d) {
  if (!op || !d) return ICC_EINVAL;
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(launch_diag_f64(op->g, op->d_scale, d, st));
  if (op->s_nnz > 0) {
    // d += diag(S): reuse the fill-free path via a tiny axpy through ir_update (scale 1)
    ICC_HIP(launch_ir_update<double>(op->g->s.ndof, d, op->d_s_diag, 1.0, st));
  }
  if (op->f_n > 0) ICC_HIP(launch_ir_update<double>(op->g->s.ndof, d, op->d_f_diag, 1.0, st));
  return ICC_OK;
}

ICC_API int icc_diag(icc_op* op, double* d) {
  if (!op || !d) return ICC_EINVAL;
  Work* w;
  ICC_TRY(ensure_work(op, w));
  hipStream_t st = op->g->ctx->stream;
  ICC_TRY(icc_diag_dev(op, w->tmp));
  ICC_HIP(hipMemcpyAsync(d, w->tmp, op->g->s.ndof * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  return ICC_OK;
}

// ---------------------------------------------------------------------------
namespace icc {
static int read_state(icc_op* op, Work* w, PcgState& out) {
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(w->st_host, w->st, sizeof(PcgState), hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  out = *w->st_host;
  return ICC_OK;
}

static int write_tol(icc_op* op, Work* w, double tol) {
  hipStream_t st = op->g->ctx->stream;
  // the state is reset by the init kernel; tol is written beforehand
  ICC_HIP(hipMemcpyAsync(&w->st->tol, &tol, sizeof(double), hipMemcpyHostToDevice, st));
  return ICC_OK;
}

// One PCG iteration in precision T enqueued on the op stream (sparse term
// first, then the tiles, then the two vector kernels).  Ap arrives zeroed
// (pcg_init1, the previous pcg_update2, or pdir_update after a reliable
// update), so the terms only add to it and the sparse term skips zeroing.
template <typename T>
static hipError_t enqueue_iteration(icc_op* op, Work* w, T* x, T* r, T* p, T* ap, const T* dinv, hipStream_t st,
                                    int prec = ICC_F32) {
  icc_grid* g = op->g;
  const int64_t n = g->s.ndof;
  const int npap = npap_of(op, std::is_same<T, double>::value ? ICC_F64 : prec);
  hipError_t e = hipSuccess;
  if constexpr (std::is_same<T, double>::value) {
    if (op->s_nnz > 0)
      e = op_csr<double>(op, p, ap, w->pap + elem_grid(op), st, true);
    if (e == hipSuccess) e = launch_faces<double>(op, p, ap, w->pap + elem_grid(op) + csr_blocks(op), w->st, st);
    if (e == hipSuccess)
      e = op->det ? launch_apply_det<double>(g, op, p, ap, w->pap, w->st, (double*)op->d_E, st)
                  : launch_apply_f64(g, op, p, ap, w->pap, w->st, st);
  } else {
    if (op->s_nnz > 0)
      e = op_csr<float>(op, p, ap, w->pap + elem_grid(op, prec), st, true);
    if (e == hipSuccess)
      e = launch_faces<float>(op, p, ap, w->pap + elem_grid(op, prec) + csr_blocks(op), w->st, st);
    if (e == hipSuccess)
      e = prec == ICC_F32 ? (op->det ? launch_apply_det<float>(g, op, p, ap, w->pap, w->st, (float*)op->d_E, st)
                                     : launch_apply_f32(g, op, p, ap, w->pap, w->st, st))
                          : launch_apply_lowp(prec, g, op, p, ap, w->pap, w->st, st);
  }
  if (e == hipSuccess) e = launch_pcg_update<T>(n, w->st, w->pap, npap, p, ap, x, r, dinv, w->parts, st);
  return e;
}

// A batch of PCG iterations captured once as a HIP graph.
struct IterGraph {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  ~IterGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

template <typename T>
static int capture_batch(icc_op* op, Work* w, T* x, T* r, T* p, T* ap, const T* dinv, int batch, IterGraph& G,
                         int prec = ICC_F32) {
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int i = 0; i < batch && e == hipSuccess; ++i) e = enqueue_iteration<T>(op, w, x, r, p, ap, dinv, st, prec);
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(st, &graph);
  if (e != hipSuccess) return hip_fail(e, "PCG batch capture");
  if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
  G.graph = graph;
  ICC_HIP(hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0));
  return ICC_OK;
}

// Replays the batch graph until the device state says done; one batch of
// lookahead keeps the GPU busy while the host reads the previous state.
static int run_until_done(icc_op* op, Work* w, IterGraph& G, PcgState& out) {
  hipStream_t st = op->g->ctx->stream;
  hipEvent_t ev[2];
  ICC_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  ICC_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int rc = ICC_OK;
  hipError_t e = hipSuccess;
  for (long k = 0;; ++k) {
    if ((e = hipGraphLaunch(G.exec, st)) != hipSuccess) break;
    if ((e = hipMemcpyAsync(&w->st_host[k & 1], w->st, sizeof(PcgState), hipMemcpyDeviceToHost, st)) != hipSuccess)
      break;
    if ((e = hipEventRecord(ev[k & 1], st)) != hipSuccess) break;
    if (k == 0) continue;
    if ((e = hipEventSynchronize(ev[(k - 1) & 1])) != hipSuccess) break;
    if (w->st_host[(k - 1) & 1].done) break;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipMemcpy(&out, w->st, sizeof(PcgState), hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = hip_fail(e, "PCG batch replay");
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return rc;
}

static int set_limits(icc_op* op, Work* w, double tol, int64_t maxit) {
  hipStream_t st = op->g->ctx->stream;
  int32_t mi = (int32_t)std::min<int64_t>(maxit, 2000000000);
  ICC_HIP(hipMemcpyAsync(&w->st->tol, &tol, sizeof(double), hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(&w->st->maxit, &mi, sizeof(int32_t), hipMemcpyHostToDevice, st));
  return ICC_OK;
}

constexpr int PCG_BATCH = 16;

// Jacobi PCG in precision T on device vectors: solves A x = bscale * b from
// x = 0 with dinv already formed.  status: 1 converged, 2 breakdown, 3 limit.
// prec (T = float only): the element contraction of A in FP32, BF16 or split FP16.
template <typename T, typename TB>
static int pcg_core(icc_op* op, Work* w, const TB* b, double bscale, T* x, T* r, T* p, T* ap, const T* dinv,
                    double tol, int64_t maxit, int64_t* iters, double* res, int* status, int prec = ICC_F32) {
  icc_grid* g = op->g;
  hipStream_t st = g->ctx->stream;
  const int64_t n = g->s.ndof;
  ICC_TRY(set_limits(op, w, tol, maxit));
  ICC_HIP((launch_pcg_init<T, TB>(n, b, bscale, dinv, x, r, p, ap, w->st, w->parts, st)));
  IterGraph G;
  ICC_TRY(capture_batch<T>(op, w, x, r, p, ap, dinv, PCG_BATCH, G, prec));
  PcgState hs;
  if (maxit > 0) {
    ICC_TRY(run_until_done(op, w, G, hs));
  } else {
    ICC_HIP(hipMemcpy(&hs, w->st, sizeof(PcgState), hipMemcpyDeviceToHost));
    hs.done = 3;
  }
  *iters = hs.iters;
  *res = hs.res;
  *status = hs.done;
  return ICC_OK;
}
}  // namespace icc

ICC_API int icc_pcg_dev(icc_op* op, const double* b, const double* diag, double tol, int64_t maxit, double* x,
                        int64_t* iters, double* res) {
  if (!op || !b || !diag || !x || !iters || !res) return ICC_EINVAL;
  Work* w;
  ICC_TRY(ensure_work(op, w));
  const int64_t n = op->g->s.ndof;
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(launch_inv_diag<double>(diag, w->dinv, n, 0.0, 0, st));
  int status = 0;
  // the loop accumulates into workspace vectors; x is the caller's output
  ICC_TRY((pcg_core<double, double>(op, w, b, 1.0, x, w->r, w->p, w->ap, w->dinv, tol, maxit, iters, res, &status)));
  if (status == 1) return ICC_OK;
  if (status == 2) {
    set_error("icc_pcg: breakdown (p.Ap <= 0)");
    return ICC_EBREAKDOWN;
  }
  set_error("icc_pcg: iteration limit reached");
  return ICC_ENOCONV;  // status 3
}

// icc_pcg_dev without allocation: the vectors, partials and state live in
// the library's fixed per-device workspace (k_noalloc.hip), the iterations
// are launched directly (no graph: instantiating one may allocate) and the
// state is polled every PCG_BATCH iterations into a host variable.  Calls
// that overlap on one GPU share that workspace and race.
ICC_API int icc_pcg_dev_noalloc(icc_op* op, const double* b, const double* diag, double tol, int64_t maxit,
                                double* x, int64_t* iters, double* res) {
  if (!op || !b || !diag || !x || !iters || !res) return ICC_EINVAL;
  const int64_t n = op->g->s.ndof;
  if (n > NA_MAX_NDOF) {
    set_error("icc_pcg_dev_noalloc: " + std::to_string(n) + " unknowns exceed the fixed workspace's " +
              std::to_string(NA_MAX_NDOF));
    return ICC_EINVAL;
  }
  if (npap_of(op) > NA_PAP) {
    set_error("icc_pcg_dev_noalloc: the operator's p.Ap partials exceed the fixed workspace");
    return ICC_ENOTSUP;
  }
  double* base = noalloc_scratch(op->g->ctx->device);
  if (base == nullptr) {
    set_error("icc_pcg_dev_noalloc: the device workspace's address could not be looked up");
    return ICC_EHIP;
  }
  hipStream_t st = op->g->ctx->stream;
  Work w;  // views into the workspace; nothing here owns or frees memory
  w.n = n;
  w.st = reinterpret_cast<PcgState*>(base);
  w.parts = base + NA_STATE;
  w.pap = w.parts + 2 * VEC_MAX_BLOCKS;
  w.r = w.pap + NA_PAP;
  w.p = w.r + NA_MAX_NDOF;
  w.ap = w.p + NA_MAX_NDOF;
  w.dinv = w.ap + NA_MAX_NDOF;
  // the state keeps the previous call's contents: clear it (a fill kernel)
  ICC_HIP(launch_fill<double>(base, 0.0, NA_STATE, st));
  ICC_HIP(launch_inv_diag<double>(diag, w.dinv, n, 0.0, 0, st));
  ICC_TRY(set_limits(op, &w, tol, maxit));
  ICC_HIP((launch_pcg_init<double, double>(n, b, 1.0, w.dinv, x, w.r, w.p, w.ap, w.st, w.parts, st)));
  PcgState hs;
  for (;;) {
    if (maxit > 0)
      for (int i = 0; i < PCG_BATCH; ++i) ICC_HIP(enqueue_iteration<double>(op, &w, x, w.r, w.p, w.ap, w.dinv, st));
    ICC_HIP(hipMemcpyAsync(&hs, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    ICC_HIP(hipStreamSynchronize(st));
    if (maxit <= 0) hs.done = 3;
    if (hs.done) break;
  }
  *iters = hs.iters;
  *res = hs.res;
  if (hs.done == 1) return ICC_OK;
  if (hs.done == 2) {
    set_error("icc_pcg_dev_noalloc: breakdown (p.Ap <= 0)");
    return ICC_EBREAKDOWN;
  }
  set_error("icc_pcg_dev_noalloc: iteration limit reached");
  return ICC_ENOCONV;
}

ICC_API int icc_pcg(icc_op* op, const double* b, const double* diag, double tol, int64_t maxit, double* x,
                    int64_t* iters, double* res) {
  if (!op || !b || !diag || !x || !iters || !res) return ICC_EINVAL;
  Work* w;
  ICC_TRY(ensure_work(op, w));
  const int64_t n = op->g->s.ndof;
  hipStream_t st = op->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(w->f, b, n * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(w->tmp, diag, n * 8, hipMemcpyHostToDevice, st));
  int rc = icc_pcg_dev(op, w->f, w->tmp, tol, maxit, w->x, iters, res);
  if (rc != ICC_OK && rc != ICC_EBREAKDOWN && rc != ICC_ENOCONV) return rc;
  ICC_HIP(hipMemcpyAsync(x, w->x, n * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  return rc;
}

// ---------------------------------------------------------------------------
ICC_API int icc_solve_ir_dev(icc_op* op, int inner, const double* f, double tol, double inner_tol, int64_t max_outer,
                             int64_t max_inner, double* u, double* stats) {
  if (!op || !f || !u) return ICC_EINVAL;
  if (inner != ICC_F64 && inner != ICC_F32 && inner != ICC_BF16 && inner != ICC_F16X) {
    set_error("icc_solve_ir: unknown inner precision");
    return ICC_EINVAL;
  }
  if (op->det && inner != ICC_F64 && inner != ICC_F32) {
    set_error("icc_solve_ir: the deterministic mode covers the FP64 and FP32 operators only");
    return ICC_EINVAL;
  }
  Work* w;
  ICC_TRY(ensure_work(op, w));
  icc_grid* g = op->g;
  const int64_t n = g->s.ndof;
  hipStream_t st = g->ctx->stream;
  // Jacobi from the FP64 diagonal, clamped like solve_state (topopt_iga.rs:147)
  ICC_TRY(icc_diag_dev(op, w->tmp));
  ICC_HIP(launch_inv_diag<double>(w->tmp, w->dinv, n, 1e-12, 1, st));
  ICC_HIP(launch_inv_diag<float>(w->tmp, w->dinv32, n, 1e-12, 1, st));
  double st_outer = 0, st_inner = 0, st_res = 1, st_app64 = 0, st_appin = 0;
  int rc_final = ICC_ENOCONV;
  PcgState hs;
  // The true residual of every outer step is kept: the best iterate is what
  // a failed solve returns, and a step that does not cut the residual by
  // IR_STALL ends the refinement (an inner operator too coarse for the
  // system's conditioning makes no progress, and running every outer step
  // would only repeat full inner solves).
  constexpr double IR_STALL = 0.5;
  double res_prev = 0.0, res_best = 0.0;
  for (int64_t outer = 0; outer <= max_outer; ++outer) {
    // r = M (f - A u) in FP64, compensated norms
    ICC_HIP(apply_t<double>(op, u, w->ap, nullptr));
    st_app64 += 1;
    ICC_HIP(launch_residual(n, f, w->ap, op->d_free, w->r, w->st, w->parts, st));
    ICC_TRY(read_state(op, w, hs));
    const double nr = std::sqrt(hs.rr), nf = std::sqrt(hs.norm_b);
    st_res = nf > 0 ? nr / nf : nr;
    if (st_res <= tol) {
      rc_final = ICC_OK;
      break;
    }
    if (outer == 0 || st_res < res_best) {
      res_best = st_res;
      ICC_HIP(hipMemcpyAsync(w->ubest, u, n * 8, hipMemcpyDeviceToDevice, st));
    }
    if (outer == max_outer || (outer > 0 && !(st_res <= IR_STALL * res_prev))) {
      if (outer < max_outer) set_error("icc_solve_ir: the refinement stagnated (the inner operator is too coarse)");
      if (st_res > res_best) {  // return the best iterate seen
        ICC_HIP(hipMemcpyAsync(u, w->ubest, n * 8, hipMemcpyDeviceToDevice, st));
        st_res = res_best;
      }
      rc_final = ICC_ENOCONV;
      break;
    }
    res_prev = st_res;
    st_outer += 1;
    int64_t its = 0;
    double res = 0;
    int status = 0;
    if (inner != ICC_F64) {
      ICC_TRY((pcg_core<float, double>(op, w, w->r, 1.0 / nr, w->x32, w->r32, w->p32, w->ap32, w->dinv32, inner_tol,
                                       max_inner, &its, &res, &status, inner)));
      ICC_HIP(launch_ir_update<float>(n, u, w->x32, nr, st));
    } else {
      ICC_TRY((pcg_core<double, double>(op, w, w->r, 1.0 / nr, w->x, w->tmp, w->p, w->ap, w->dinv, inner_tol,
                                        max_inner, &its, &res, &status)));
      ICC_HIP(launch_ir_update<double>(n, u, w->x, nr, st));
    }
    st_inner += (double)its;
    st_appin += (double)its;
    if (status == 2 && its <= 1) {
      set_error("icc_solve_ir: inner breakdown");
      rc_final = ICC_EBREAKDOWN;
      break;
    }
  }
  if (stats) {
    stats[0] = st_outer;
    stats[1] = st_inner;
    stats[2] = st_res;
    stats[3] = st_app64;
    stats[4] = st_appin;
    stats[5] = stats[6] = stats[7] = 0.0;
  }
  if (rc_final == ICC_ENOCONV && st_outer >= (double)max_outer)
    set_error("icc_solve_ir: outer iteration limit reached");
  return rc_final;
}

// Mixed precision with reliable updates: Jacobi PCG iterates in FP32 (FP32
// MFMA apply, FP32 vectors, FP64 reductions); whenever the FP32 residual has
// dropped by `delta` since the last update, the FP32 correction is added to
// the FP64 solution, the true residual f - A u is recomputed with the FP64
// operator and compensated norms, and the FP32 residual is replaced by it
// while the search direction is kept (no restart).  Stops when the true
// relative residual <= tol.  stats: updates, FP32 iterations, true residual,
// FP64 applies, FP32 applies.
ICC_API int icc_solve_mixed_prec_dev(icc_op* op, int inner, const double* f, double tol, double delta, int64_t maxit,
                                     double* u, double* stats) {
  if (!op || !f || !u || !(delta > 0.0 && delta < 1.0)) return ICC_EINVAL;
  if (inner != ICC_F32 && inner != ICC_BF16 && inner != ICC_F16X) {
    set_error("icc_solve_mixed: the inner precision must be ICC_F32, ICC_BF16 or ICC_F16X");
    return ICC_EINVAL;
  }
  if (op->det && inner != ICC_F32) {
    set_error("icc_solve_mixed: the deterministic mode covers the FP64 and FP32 operators only");
    return ICC_EINVAL;
  }
  Work* w;
  ICC_TRY(ensure_work(op, w));
  icc_grid* g = op->g;
  const int64_t n = g->s.ndof;
  hipStream_t st = g->ctx->stream;
  ICC_TRY(icc_diag_dev(op, w->tmp));
  ICC_HIP(launch_inv_diag<float>(w->tmp, w->dinv32, n, 1e-12, 1, st));
  double n_upd = 0, n_app64 = 0, res_true = 1.0;
  int rc_final = ICC_ENOCONV;
  PcgState hs;
  // true residual of the initial guess
  ICC_HIP(apply_t<double>(op, u, w->ap, nullptr));
  n_app64 += 1;
  ICC_HIP(launch_residual(n, f, w->ap, op->d_free, w->r, w->st, w->parts, st));
  ICC_TRY(read_state(op, w, hs));
  const double nf = std::sqrt(hs.norm_b);
  res_true = nf > 0 ? std::sqrt(hs.rr) / nf : std::sqrt(hs.rr);
  int64_t iters = 0;
  if (res_true <= tol) {
    rc_final = ICC_OK;
  } else {
    // FP32 recurrence from r (x32 = 0, p = D^-1 r), stopping at |r| < delta |r_0|
    ICC_TRY(set_limits(op, w, delta, maxit));
    ICC_HIP((launch_pcg_init<float, double>(n, w->r, 1.0, w->dinv32, w->x32, w->r32, w->p32, w->ap32, w->st, w->parts,
                                             st)));
    IterGraph G;
    ICC_TRY(capture_batch<float>(op, w, w->x32, w->r32, w->p32, w->ap32, w->dinv32, PCG_BATCH, G, inner));
    for (;;) {
      ICC_TRY(run_until_done(op, w, G, hs));
      iters = hs.iters;
      // fold the FP32 correction into u and recompute the true residual
      ICC_HIP(launch_ir_update<float>(n, u, w->x32, 1.0, st));
      ICC_HIP(apply_t<double>(op, u, w->ap, nullptr));
      n_app64 += 1;
      n_upd += 1;
      ICC_HIP(launch_residual(n, f, w->ap, op->d_free, w->r, w->st, w->parts, st));
      PcgState hr;
      ICC_TRY(read_state(op, w, hr));
      res_true = nf > 0 ? std::sqrt(hr.rr) / nf : std::sqrt(hr.rr);
      if (res_true <= tol) {
        rc_final = ICC_OK;
        break;
      }
      if (hs.done == 3) break;  // iteration limit
      if (hs.done == 2) {       // breakdown of the FP32 recurrence
        rc_final = ICC_EBREAKDOWN;
        set_error("icc_solve_mixed: breakdown of the FP32 recurrence");
        break;
      }
      // replace r32 by the true residual, x32 = 0, and finish the pending
      // direction update p = D^-1 r + beta p with the replaced residual
      ICC_HIP(launch_reliable_replace(n, w->r, w->dinv32, w->r32, w->x32, w->st, w->parts, st));
      ICC_HIP(launch_pdir_update(n, w->st, w->p32, w->r32, w->dinv32, w->ap32, st));
    }
  }
  if (stats) {
    stats[0] = n_upd;
    stats[1] = (double)iters;
    stats[2] = res_true;
    stats[3] = n_app64;
    stats[4] = (double)iters;
    stats[5] = stats[6] = stats[7] = 0.0;
  }
  if (rc_final == ICC_ENOCONV) set_error("icc_solve_mixed: iteration limit reached");
  return rc_final;
}

ICC_API int icc_solve_mixed_dev(icc_op* op, const double* f, double tol, double delta, int64_t maxit, double* u,
                                double* stats) {
  return icc_solve_mixed_prec_dev(op, ICC_F32, f, tol, delta, maxit, u, stats);
}

namespace icc {
// Host-pointer wrapper of a device solve: copies f and the initial guess u to
// temporaries, runs `solve` on them and copies u back when the solve ended
// with a result (converged, iteration limit or breakdown).
template <typename F>
static int host_solve(icc_op* op, const double* f, double* u, F&& solve) {
  const int64_t n = op->g->s.ndof;
  hipStream_t st = op->g->ctx->stream;
  double *df = nullptr, *du = nullptr;
  ICC_TRY(dalloc(&df, n));
  if (dalloc(&du, n) != ICC_OK) {
    hipFree(df);
    return ICC_ENOMEM;
  }
  int rc = ICC_OK;
  if (hipMemcpyAsync(df, f, n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(du, u, n * 8, hipMemcpyHostToDevice, st) != hipSuccess) {
    set_error("host_solve: copy to device failed");
    rc = ICC_EHIP;
  } else {
    rc = solve(df, du);
    if (rc == ICC_OK || rc == ICC_ENOCONV || rc == ICC_EBREAKDOWN) {
      if (hipMemcpyAsync(u, du, n * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        set_error("host_solve: copy to host failed");
        rc = ICC_EHIP;
      }
    }
  }
  (void)hipStreamSynchronize(st);
  hipFree(df);
  hipFree(du);
  return rc;
}
}  // namespace icc

ICC_API int icc_solve_ir(icc_op* op, int inner, const double* f, double tol, double inner_tol, int64_t max_outer,
                         int64_t max_inner, double* u, double* stats) {
  if (!op || !f || !u) return ICC_EINVAL;
  return host_solve(op, f, u, [&](const double* df, double* du) {
    return icc_solve_ir_dev(op, inner, df, tol, inner_tol, max_outer, max_inner, du, stats);
  });
}

ICC_API int icc_solve_mixed_prec(icc_op* op, int inner, const double* f, double tol, double delta, int64_t maxit,
                                 double* u, double* stats) {
  if (!op || !f || !u) return ICC_EINVAL;
  return host_solve(op, f, u, [&](const double* df, double* du) {
    return icc_solve_mixed_prec_dev(op, inner, df, tol, delta, maxit, du, stats);
  });
}

ICC_API int icc_solve_mixed(icc_op* op, const double* f, double tol, double delta, int64_t maxit, double* u,
                            double* stats) {
  if (!op || !f || !u) return ICC_EINVAL;
  return host_solve(op, f, u, [&](const double* df, double* du) {
    return icc_solve_mixed_dev(op, df, tol, delta, maxit, du, stats);
  });
}

// ---------------------------------------------------------------------------
ICC_API int icc_energies_dev(icc_op* op, const double* u, double* ee) {
  if (!op || !u || !ee) return ICC_EINVAL;
  ICC_HIP(launch_energies(op->g, u, ee, nullptr, 0.0, 0.0, nullptr, nullptr, op->g->ctx->stream));
  return ICC_OK;
}

ICC_API int icc_sens_dev(icc_op* op, const double* u, const double* x, double penal, double emin, const double* w,
                         double* dc) {
  if (!op || !u || !x || !dc) return ICC_EINVAL;
  ICC_HIP(launch_energies(op->g, u, nullptr, x, penal, emin, w, dc, op->g->ctx->stream));
  return ICC_OK;
}

namespace icc {
// host wrapper: uploads `nin` host arrays of given lengths, runs fn, downloads one output
static int host_call(icc_ctx* ctx, std::vector<std::pair<const double*, int64_t>> ins, int64_t nout, double* out,
                     const std::function<int(std::vector<double*>&, double*)>& fn) {
  hipStream_t st = ctx->stream;
  std::vector<double*> dins;
  double* dout = nullptr;
  int rc = ICC_OK;
  for (auto& in : ins) {
    double* d = nullptr;
    if (in.first) {
      if ((rc = dalloc(&d, in.second)) != ICC_OK) break;
      if (hipMemcpyAsync(d, in.first, in.second * 8, hipMemcpyHostToDevice, st) != hipSuccess) rc = ICC_EHIP;
    }
    dins.push_back(d);
  }
  if (rc == ICC_OK) rc = dalloc(&dout, nout);
  if (rc == ICC_OK) rc = fn(dins, dout);
  if (rc == ICC_OK) {
    if (hipMemcpyAsync(out, dout, nout * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = ICC_EHIP;
  } else {
    hipStreamSynchronize(st);
  }
  for (double* d : dins) hipFree(d);
  hipFree(dout);
  return rc;
}
}  // namespace icc

ICC_API int icc_energies(icc_op* op, const double* u, double* ee) {
  if (!op || !u || !ee) return ICC_EINVAL;
  const auto& s = op->g->s;
  return host_call(op->g->ctx, {{u, s.ndof}}, s.nel_tot, ee,
                   [&](std::vector<double*>& in, double* out) { return icc_energies_dev(op, in[0], out); });
}

ICC_API int icc_sens(icc_op* op, const double* u, const double* x, double penal, double emin, const double* w,
                     double* dc) {
  if (!op || !u || !x || !dc) return ICC_EINVAL;
  const auto& s = op->g->s;
  return host_call(op->g->ctx, {{u, s.ndof}, {x, s.nel_tot}, {w, s.nel_tot}}, s.nel_tot, dc,
                   [&](std::vector<double*>& in, double* out) {
                     return icc_sens_dev(op, in[0], in[1], penal, emin, in[2], out);
                   });
}

// ---------------------------------------------------------------------------
ICC_API int icc_filter_create(icc_grid* g, double rmin, icc_filter** out) {
  if (!g || !out || !(rmin > 0)) return ICC_EINVAL;
  auto* f = new icc_filter();
  f->g = g;
  const auto& s = g->s;
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = s.len[d] / (double)s.nel[d];
  f->r = rmin * (h[0] + h[1] + h[2]) / 3.0;
  for (int d = 0; d < 3; ++d) f->rad[d] = (int)std::ceil(f->r / h[d]) + 1;
  int rc = dalloc(&f->d_hs, s.nel_tot);
  if (rc != ICC_OK) {
    delete f;
    return rc;
  }
  hipError_t e = launch_filter(f, 3, nullptr, nullptr, f->d_hs, g->ctx->stream);
  if (e != hipSuccess) {
    icc_filter_destroy(f);
    return hip_fail(e, "icc_filter_create");
  }
  *out = f;
  return ICC_OK;
}

ICC_API int icc_filter_destroy(icc_filter* f) {
  if (!f) return ICC_OK;
  hipFree(f->d_hs);
  hipFree(f->d_tmp);
  hipFree(f->d_act);
  hipFree(f->d_off);
  hipFree(f->d_w);
  delete f;
  return ICC_OK;
}

// Filter on a design domain (topopt_immersed.rs:117-151): offsets within
// r = rmin * mean(h) measured from the offset itself, weight r - dist, only
// active in-range neighbours of active cells; inactive rows are empty.
ICC_API int icc_filter_create_active(icc_grid* g, double rmin, const uint8_t* active, icc_filter** out) {
  if (!g || !out || !active || !(rmin > 0)) return ICC_EINVAL;
  auto* f = new icc_filter();
  f->g = g;
  const auto& s = g->s;
  std::vector<int32_t> offs;
  std::vector<double> ows;
  f->r = design_filter_offsets(s, rmin, offs, ows);
  f->noff = (int)ows.size();
  const int64_t ne = s.nel_tot;
  int rc;
  if ((rc = dalloc(&f->d_hs, ne)) || (rc = dalloc(&f->d_act, ne)) || (rc = dalloc(&f->d_off, offs.size())) ||
      (rc = dalloc(&f->d_w, ows.size()))) {
    icc_filter_destroy(f);
    return rc;
  }
  hipStream_t st = g->ctx->stream;
  hipError_t e;
  if ((e = hipMemcpyAsync(f->d_act, active, ne, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(f->d_off, offs.data(), offs.size() * 4, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(f->d_w, ows.data(), ows.size() * 8, hipMemcpyHostToDevice, st)) ||
      (e = launch_imm_filter(1, s.nel, f->d_off, f->d_w, f->noff, f->d_act, nullptr, nullptr, nullptr, f->d_hs, st)) ||
      (e = hipStreamSynchronize(st))) {
    icc_filter_destroy(f);
    return hip_fail(e, "icc_filter_create_active");
  }
  *out = f;
  return ICC_OK;
}

namespace icc {
// mode 0 sensitivity, 1 density forward, 2 density backward (structured
// numbering; the design-domain kernel numbers them 0, 2, 3)
static hipError_t run_filter(icc_filter* f, int mode, const double* x, const double* v, double* out) {
  hipStream_t st = f->g->ctx->stream;
  if (f->d_act == nullptr) return launch_filter(f, mode, x, v, out, st);
  static const int imm_mode[3] = {0, 2, 3};
  return launch_imm_filter(imm_mode[mode], f->g->s.nel, f->d_off, f->d_w, f->noff, f->d_act, x, v, f->d_hs, out, st);
}
}  // namespace icc

ICC_API int icc_filter_sens_dev(icc_filter* f, const double* x, const double* dc, double* out) {
  if (!f || !x || !dc || !out) return ICC_EINVAL;
  ICC_HIP(run_filter(f, 0, x, dc, out));
  return ICC_OK;
}
ICC_API int icc_filter_fwd_dev(icc_filter* f, const double* x, double* out) {
  if (!f || !x || !out) return ICC_EINVAL;
  ICC_HIP(run_filter(f, 1, nullptr, x, out));
  return ICC_OK;
}
ICC_API int icc_filter_bwd_dev(icc_filter* f, const double* gr, double* out) {
  if (!f || !gr || !out) return ICC_EINVAL;
  ICC_HIP(run_filter(f, 2, nullptr, gr, out));
  return ICC_OK;
}
ICC_API int icc_filter_sens(icc_filter* f, const double* x, const double* dc, double* out) {
  if (!f || !x || !dc || !out) return ICC_EINVAL;
  int64_t n = f->g->s.nel_tot;
  return host_call(f->g->ctx, {{x, n}, {dc, n}}, n, out,
                   [&](std::vector<double*>& in, double* o) { return icc_filter_sens_dev(f, in[0], in[1], o); });
}
ICC_API int icc_filter_fwd(icc_filter* f, const double* x, double* out) {
  if (!f || !x || !out) return ICC_EINVAL;
  int64_t n = f->g->s.nel_tot;
  return host_call(f->g->ctx, {{x, n}}, n, out,
                   [&](std::vector<double*>& in, double* o) { return icc_filter_fwd_dev(f, in[0], o); });
}
ICC_API int icc_filter_bwd(icc_filter* f, const double* gr, double* out) {
  if (!f || !gr || !out) return ICC_EINVAL;
  int64_t n = f->g->s.nel_tot;
  return host_call(f->g->ctx, {{gr, n}}, n, out,
                   [&](std::vector<double*>& in, double* o) { return icc_filter_bwd_dev(f, in[0], o); });
}

// ---------------------------------------------------------------------------
ICC_API int icc_wq_diag_dev(icc_grid* g, const double* x, double penal, double emin, double* d) {
  if (!g || !d) return ICC_EINVAL;
  hipStream_t st = g->ctx->stream;
  if (!g->d_wq) {
    std::vector<double> tab;
    for (int dd = 0; dd < 3; ++dd) {
      std::vector<double> a0, a1;
      wq_tables_1d(g->s.nel[dd], a0, a1);
      tab.insert(tab.end(), a0.begin(), a0.end());
      tab.insert(tab.end(), a1.begin(), a1.end());
    }
    ICC_TRY(dalloc(&g->d_wq, tab.size()));
    ICC_HIP(hipMemcpyAsync(g->d_wq, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, st));
    ICC_HIP(hipStreamSynchronize(st));
  }
  ICC_HIP(launch_wq_diag(g, g->d_wq, x, penal, emin, d, st));
  return ICC_OK;
}

ICC_API int icc_wq_diag(icc_grid* g, const double* x, double penal, double emin, double* d) {
  if (!g || !d) return ICC_EINVAL;
  const auto& s = g->s;
  return host_call(g->ctx, {{x, s.nel_tot}}, s.ndof, d,
                   [&](std::vector<double*>& in, double* out) { return icc_wq_diag_dev(g, in[0], penal, emin, out); });
}

ICC_API int icc_classify_cylinders(icc_ctx* ctx, const double box_lo[3], const double box_hi[3], const double* cyl,
                                   int64_t ncyl, const double grid_bounds[6], const int64_t res[3],
                                   const int64_t sub[3], int8_t* status, double* weights) {
  if (!ctx || !box_lo || !box_hi || (ncyl > 0 && !cyl) || !grid_bounds || !res || !sub || !status || !weights)
    return ICC_EINVAL;
  for (int d = 0; d < 3; ++d)
    if (res[d] < 1 || sub[d] < 1 || !(grid_bounds[2 * d + 1] > grid_bounds[2 * d])) {
      set_error("icc_classify_cylinders: bad resolution or bounds");
      return ICC_EINVAL;
    }
  hipStream_t st = ctx->stream;
  const int64_t nc = res[0] * res[1] * res[2], nv = (res[0] + 1) * (res[1] + 1) * (res[2] + 1);
  double *d_cyl = nullptr, *d_w = nullptr;
  uint8_t* d_in = nullptr;
  int8_t* d_st = nullptr;
  int rc = ICC_OK;
  if ((rc = dalloc(&d_cyl, 3 * std::max<int64_t>(ncyl, 1))) || (rc = dalloc(&d_w, nc)) || (rc = dalloc(&d_in, nv)) ||
      (rc = dalloc(&d_st, nc))) {
    hipFree(d_cyl);
    hipFree(d_w);
    hipFree(d_in);
    hipFree(d_st);
    return rc;
  }
  hipError_t e = hipSuccess;
  if (ncyl > 0) e = hipMemcpyAsync(d_cyl, cyl, 3 * ncyl * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = launch_classify_cylinders(box_lo, box_hi, d_cyl, ncyl, grid_bounds, res, sub, d_in, d_st, d_w, st);
  if (e == hipSuccess) e = hipMemcpyAsync(status, d_st, nc, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(weights, d_w, nc * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(d_cyl);
  hipFree(d_w);
  hipFree(d_in);
  hipFree(d_st);
  if (e != hipSuccess) return hip_fail(e, "icc_classify_cylinders");
  return ICC_OK;
}

ICC_API int icc_debug_mfma_layout_bf16(icc_ctx* ctx, double* out256) {
  if (!ctx || !out256) return ICC_EINVAL;
  double* d = nullptr;
  ICC_TRY(dalloc(&d, 256));
  hipError_t e = launch_layout_probe_bf16(d, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out256, d, 256 * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "icc_debug_mfma_layout_bf16");
  return ICC_OK;
}

ICC_API int icc_debug_mfma_layout_f16(icc_ctx* ctx, double* out512) {
  if (!ctx || !out512) return ICC_EINVAL;
  double* d = nullptr;
  ICC_TRY(dalloc(&d, 512));
  hipError_t e = launch_layout_probe_f16(d, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out512, d, 512 * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "icc_debug_mfma_layout_f16");
  return ICC_OK;
}

ICC_API int icc_debug_mfma_layout(icc_ctx* ctx, double* out512) {
  if (!ctx || !out512) return ICC_EINVAL;
  double* d = nullptr;
  ICC_TRY(dalloc(&d, 512));
  hipError_t e = launch_layout_probe(d, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out512, d, 512 * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "icc_debug_mfma_layout");
  return ICC_OK;
}

// ---------------------------------------------------------------------------
// Tiling autotuner: times `reps` accumulate-only applies of op in precision
// prec for every tiling variant and workgroups-per-CU choice, and keeps the
// fastest (ties to the lower variant).  The operator's vectors are scratch.
ICC_API int icc_grid_autotune(icc_op* op, int prec, int reps, int64_t* best_variant, int64_t* best_wg,
                              double* best_us) {
  if (!op || (prec != ICC_F64 && prec != ICC_F32 && prec != ICC_BF16) || reps < 1) return ICC_EINVAL;
  if (op->s_nnz > 0 || op->f_n > 0) {
    set_error("icc_grid_autotune: operators with a sparse term are not supported");
    return ICC_ENOTSUP;
  }
  Work* w;
  ICC_TRY(ensure_work(op, w));
  icc_grid* g = op->g;
  hipStream_t st = g->ctx->stream;
  const int64_t n = g->s.ndof;
  ICC_HIP(launch_fill<double>(w->tmp, 1.0, n, st));
  ICC_HIP(launch_fill<double>(w->ap, 0.0, n, st));
  ICC_HIP(launch_fill<float>(w->x32, 1.0f, n, st));
  ICC_HIP(launch_fill<float>(w->ap32, 0.0f, n, st));
  hipEvent_t e0, e1;
  ICC_HIP(hipEventCreate(&e0));
  ICC_HIP(hipEventCreate(&e1));
  int bv = 0, bw = 2;
  float bt = 1e30f;
  int rc = ICC_OK;
  for (int v = 0; v < n_tile_variants() && rc == ICC_OK; ++v)
    for (int wg = 1; wg <= APPLY_MAX_WG_PER_CU && rc == ICC_OK; wg *= 2) {
      if ((rc = icc_grid_set_tiling(g, v, wg)) != ICC_OK) break;
      hipError_t e = hipSuccess;
      for (int r = -2; r < reps && e == hipSuccess; ++r) {
        if (r == 0) e = hipEventRecord(e0, st);
        if (e != hipSuccess) break;
        if (prec == ICC_F64)
          e = launch_apply_f64(g, op, w->tmp, w->ap, nullptr, nullptr, st);
        else if (prec == ICC_F32)
          e = launch_apply_f32(g, op, w->x32, w->ap32, nullptr, nullptr, st);
        else
          e = launch_apply_bf16(g, op, w->x32, w->ap32, nullptr, nullptr, st);
      }
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = hip_fail(e, "icc_grid_autotune");
        break;
      }
      if (ms < bt) {
        bt = ms;
        bv = v;
        bw = wg;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != ICC_OK) return rc;
  ICC_TRY(icc_grid_set_tiling(g, bv, bw));
  if (best_variant) *best_variant = bv;
  if (best_wg) *best_wg = bw;
  if (best_us) *best_us = 1e3 * (double)bt / reps;
  return ICC_OK;
}

// ---------------------------------------------------------------------------
// Batched instances: the PCG iterations of every unfinished instance are
// captured, one instance after another, into one HIP graph of PCG_BATCH
// iterations per instance on the context stream, replayed until every
// instance has stopped; each instance's kernels return at once when its own
// device state is done.  Mixed precision: the reliable updates of the
// instances whose FP32 residual dropped by delta run on the context stream
// between graph replays.  (Capturing the instances as concurrent graph
// branches on forked streams gave wrong iteration counts on this platform,
// so the chains are serialised.)
struct icc_batch {
  icc_ctx* ctx = nullptr;
  std::vector<icc_grid*> g;
  std::vector<icc_op*> op;
  std::vector<double*> d_f, d_u;
  int wg_per_cu[2] = {1, 2};  // batched apply workgroups per CU: FP64, FP32
};

ICC_API int icc_batch_destroy(icc_batch* b) {
  if (!b) return ICC_OK;
  for (size_t k = 0; k < b->op.size(); ++k) {
    icc_op_destroy(b->op[k]);
    icc_grid_destroy(b->g[k]);
  }
  for (double* p : b->d_f) hipFree(p);
  for (double* p : b->d_u) hipFree(p);
  delete b;
  return ICC_OK;
}

ICC_API int icc_batch_create(icc_ctx* ctx, int64_t n, const int64_t* nel, const double* size, double young,
                             double poisson, icc_batch** out) {
  if (!ctx || n < 1 || !nel || !out) return ICC_EINVAL;
  auto* b = new icc_batch();
  b->ctx = ctx;
  int rc = ICC_OK;
  for (int64_t k = 0; k < n && rc == ICC_OK; ++k) {
    const double L[3] = {size ? size[3 * k] : 2.0, size ? size[3 * k + 1] : 1.0, size ? size[3 * k + 2] : 1.0};
    const double bounds[6] = {0.0, L[0], 0.0, L[1], 0.0, L[2]};
    icc_grid* g = nullptr;
    icc_op* op = nullptr;
    rc = icc_grid_create(ctx, nel + 3 * k, bounds, 2, young, poisson, &g);
    if (rc == ICC_OK) rc = icc_op_create(g, &op);
    b->g.push_back(g);
    b->op.push_back(op);
    double *f = nullptr, *u = nullptr;
    if (rc == ICC_OK && (rc = dalloc(&f, g->s.ndof)) == ICC_OK) rc = dalloc(&u, g->s.ndof);
    b->d_f.push_back(f);
    b->d_u.push_back(u);
    if (rc == ICC_OK) {
      // default: clamp x = 0, no load
      const GridSetup& gs = g->s;
      std::vector<uint8_t> fr(gs.ndof, 1);
      for (int64_t i = 0; i < gs.ndof_sc; ++i)
        if (i % gs.ncp[0] == 0)
          for (int c = 0; c < 3; ++c) fr[c * gs.ndof_sc + i] = 0;
      rc = icc_op_set_free_mask(op, fr.data());
      if (rc == ICC_OK && hipMemsetAsync(f, 0, gs.ndof * 8, ctx->stream) != hipSuccess) rc = ICC_EHIP;
    }
  }
  if (rc == ICC_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = ICC_EHIP;
  if (rc != ICC_OK) {
    if (rc == ICC_EHIP) set_error("icc_batch_create: HIP error");
    icc_batch_destroy(b);
    return rc;
  }
  *out = b;
  return ICC_OK;
}

ICC_API int icc_batch_sizes(const icc_batch* b, int64_t* ndof, int64_t* nel) {
  if (!b) return ICC_EINVAL;
  for (size_t k = 0; k < b->g.size(); ++k) {
    if (ndof) ndof[k] = b->g[k]->s.ndof;
    if (nel) nel[k] = b->g[k]->s.nel_tot;
  }
  return ICC_OK;
}

ICC_API int icc_batch_set(icc_batch* b, int64_t k, const double* scale, const uint8_t* free_mask, const double* force) {
  if (!b || k < 0 || k >= (int64_t)b->op.size()) return ICC_EINVAL;
  icc_op* op = b->op[k];
  const GridSetup& gs = b->g[k]->s;
  if (scale) ICC_TRY(icc_op_set_scale(op, scale));
  std::vector<uint8_t> fr;
  if (free_mask) {
    fr.assign(free_mask, free_mask + gs.ndof);
    ICC_TRY(icc_op_set_free_mask(op, free_mask));
  } else {
    fr.assign(gs.ndof, 1);
    for (int64_t i = 0; i < gs.ndof_sc; ++i)
      if (i % gs.ncp[0] == 0)
        for (int c = 0; c < 3; ++c) fr[c * gs.ndof_sc + i] = 0;
    ICC_TRY(icc_op_set_free_mask(op, fr.data()));
  }
  if (force) {
    std::vector<double> fm(force, force + gs.ndof);
    for (int64_t i = 0; i < gs.ndof; ++i)
      if (!fr[i]) fm[i] = 0.0;
    ICC_HIP(hipMemcpyAsync(b->d_f[k], fm.data(), gs.ndof * 8, hipMemcpyHostToDevice, b->ctx->stream));
    ICC_HIP(hipStreamSynchronize(b->ctx->stream));
  }
  return ICC_OK;
}

ICC_API int icc_batch_autotune(icc_batch* b, int prec, int reps) {
  if (!b) return ICC_EINVAL;
  for (size_t k = 0; k < b->op.size(); ++k) ICC_TRY(icc_grid_autotune(b->op[k], prec, reps, nullptr, nullptr, nullptr));
  return ICC_OK;
}

namespace icc {
void traction_xmax(const GridSetup& s, const double t[3], std::vector<double>& f);
}

ICC_API int icc_batch_set_traction(icc_batch* b, int64_t k, const double t[3]) {
  if (!b || !t || k < 0 || k >= (int64_t)b->op.size()) return ICC_EINVAL;
  const GridSetup& gs = b->g[k]->s;
  std::vector<double> f;
  traction_xmax(gs, t, f);
  std::vector<uint8_t> fr(gs.ndof, 1);
  if (b->op[k]->d_free) ICC_HIP(hipMemcpy(fr.data(), b->op[k]->d_free, gs.ndof, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < gs.ndof; ++i)
    if (!fr[i]) f[i] = 0.0;
  ICC_HIP(hipMemcpyAsync(b->d_f[k], f.data(), gs.ndof * 8, hipMemcpyHostToDevice, b->ctx->stream));
  ICC_HIP(hipStreamSynchronize(b->ctx->stream));
  return ICC_OK;
}

namespace icc {
// Device data of one batched solve: every instance's PCG vectors in
// precision T concatenated at 64-aligned offsets, their stiffness scales,
// free masks, A fragments (types renumbered by instance) and PCG states; and
// the plan of the unfinished instances: their tiles as one list with a
// TileInst per tile, the workgroup runs of the batched apply, the p.Ap slots
// of (run, instance) and the vector segments of the batched updates.  The
// plan is rebuilt (and the iteration graph captured again) when instances
// finish.
template <typename T>
struct BatchRun {
  std::vector<int64_t> off, eoff;
  std::vector<int32_t> toff;
  T *x = nullptr, *r = nullptr, *p = nullptr, *ap = nullptr, *dinv = nullptr, *scale = nullptr, *kfrag = nullptr;
  uint8_t* fr = nullptr;
  PcgState* st = nullptr;
  PcgState* st_host = nullptr;
  double *pap = nullptr, *parts = nullptr;
  int pap_cap = 0;
  VecSeg* segs = nullptr;
  // plan
  Tile* tiles = nullptr;
  int32_t *first = nullptr, *tmask = nullptr, *seg_of = nullptr;
  TileInst* tinst = nullptr;
  int G = 0, nblk = 0;
  void free_plan() {
    (void)hipFree(tiles);
    (void)hipFree(first);
    (void)hipFree(tmask);
    (void)hipFree(seg_of);
    (void)hipFree(tinst);
    tiles = nullptr;
    first = tmask = seg_of = nullptr;
    tinst = nullptr;
  }
  ~BatchRun() {
    free_plan();
    for (void* q : {(void*)x, (void*)r, (void*)p, (void*)ap, (void*)dinv, (void*)scale, (void*)kfrag, (void*)fr,
                    (void*)st, (void*)pap, (void*)parts, (void*)segs})
      (void)hipFree(q);
    if (st_host) (void)hipHostFree(st_host);
  }
};

template <typename T>
static const T* kfrag_of(const icc_grid* g);
template <>
const double* kfrag_of<double>(const icc_grid* g) {
  return g->d_kfrag64;
}
template <>
const float* kfrag_of<float>(const icc_grid* g) {
  return g->d_kfrag32;
}
template <typename T>
static const T* scale_of(const icc_op* op);
template <>
const double* scale_of<double>(const icc_op* op) {
  return op->d_scale;
}
template <>
const float* scale_of<float>(const icc_op* op) {
  return op->d_scale32;
}

// Allocates the concatenated data of all instances and copies the scales,
// free masks and A fragments in.
template <typename T>
static int batch_setup(icc_batch* b, BatchRun<T>& R) {
  const int nb = (int)b->op.size();
  hipStream_t m = b->ctx->stream;
  int64_t n_tot = 0, e_tot = 0, t_tot = 0, blk_tot = 0;
  for (int k = 0; k < nb; ++k) {
    const GridSetup& s = b->g[k]->s;
    R.off.push_back(n_tot);
    R.eoff.push_back(e_tot);
    R.toff.push_back((int32_t)t_tot);
    n_tot += (s.ndof + 63) / 64 * 64;
    e_tot += s.nel_tot;
    t_tot += (int64_t)s.kmat.size();
    blk_tot += vec_blocks(s.ndof);
  }
  if (n_tot >= ((int64_t)1 << 31)) {
    set_error("icc_batch_solve: more than 2^31 unknowns in one batch");
    return ICC_EINVAL;
  }
  const size_t fsz = (size_t)MT * KS * 64;  // fragment entries per type
  const int gmax = APPLY_MAX_WG_PER_CU * (b->ctx->num_cu > 0 ? b->ctx->num_cu : 256);
  R.pap_cap = gmax + nb;
  int rc;
  if ((rc = dalloc(&R.x, n_tot)) || (rc = dalloc(&R.r, n_tot)) || (rc = dalloc(&R.p, n_tot)) ||
      (rc = dalloc(&R.ap, n_tot)) || (rc = dalloc(&R.dinv, n_tot)) || (rc = dalloc(&R.scale, e_tot)) ||
      (rc = dalloc(&R.kfrag, t_tot * fsz)) || (rc = dalloc(&R.fr, n_tot)) || (rc = dalloc(&R.st, nb)) ||
      (rc = dalloc(&R.pap, R.pap_cap)) || (rc = dalloc(&R.parts, 2 * blk_tot)) || (rc = dalloc(&R.segs, nb)))
    return rc;
  if (hipHostMalloc((void**)&R.st_host, nb * sizeof(PcgState), hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    return ICC_ENOMEM;
  }
  std::vector<PcgState> z(nb);
  std::memset(z.data(), 0, nb * sizeof(PcgState));
  ICC_HIP(hipMemcpyAsync(R.st, z.data(), nb * sizeof(PcgState), hipMemcpyHostToDevice, m));
  ICC_HIP(launch_fill<T>(R.x, T(0), n_tot, m));
  ICC_HIP(launch_fill<T>(R.p, T(0), n_tot, m));
  ICC_HIP(launch_fill<T>(R.ap, T(0), n_tot, m));
  for (int k = 0; k < nb; ++k) {
    const GridSetup& s = b->g[k]->s;
    const icc_op* op = b->op[k];
    ICC_HIP(hipMemcpyAsync(R.scale + R.eoff[k], scale_of<T>(op), s.nel_tot * sizeof(T), hipMemcpyDeviceToDevice, m));
    ICC_HIP(hipMemcpyAsync(R.kfrag + (size_t)R.toff[k] * fsz, kfrag_of<T>(b->g[k]), s.kmat.size() * fsz * sizeof(T),
                           hipMemcpyDeviceToDevice, m));
    if (op->d_free)
      ICC_HIP(hipMemcpyAsync(R.fr + R.off[k], op->d_free, s.ndof, hipMemcpyDeviceToDevice, m));
    else
      ICC_HIP(hipMemsetAsync(R.fr + R.off[k], 1, s.ndof, m));
  }
  ICC_HIP(hipStreamSynchronize(m));
  return ICC_OK;
}

// Plan of the instances `live` (in order): their tiles concatenated, split
// into G contiguous runs of balanced work (tile_cost), p.Ap slots numbered
// instance by instance (an instance's runs are contiguous), and each
// instance's vector blocks.
template <typename T>
static int batch_plan(icc_batch* b, BatchRun<T>& R, const std::vector<int>& live, int wg_per_cu) {
  R.free_plan();
  std::vector<Tile> tl;
  std::vector<int32_t> inst, tm;
  for (int k : live) {
    const GridSetup& s = b->g[k]->s;
    const icc_op* op = b->op[k];
    const std::vector<int32_t> m =
        op->h_free.empty() ? std::vector<int32_t>(s.tiles.size(), 0) : tile_mask_of(s, s.tiles, op->h_free.data());
    for (size_t i = 0; i < s.tiles.size(); ++i) {
      Tile t = s.tiles[i];
      t.type += R.toff[k];
      tl.push_back(t);
      inst.push_back(k);
      tm.push_back(m[i]);
    }
  }
  const int64_t nt = (int64_t)tl.size();
  int G = wg_per_cu * (b->ctx->num_cu > 0 ? b->ctx->num_cu : 256);
  if (G > nt) G = (int)nt;
  if (G >= 8) G -= G % 8;
  if (G < 1) G = 1;
  // contiguous runs ending as close as possible to the cumulative targets
  std::vector<double> cum(nt + 1, 0.0);
  for (int64_t i = 0; i < nt; ++i) cum[i + 1] = cum[i] + apply_tile_cost(tl[i]);
  std::vector<int32_t> first(G + 1, 0);
  first[G] = (int32_t)nt;
  int64_t i = 0;
  for (int c = 1; c < G; ++c) {
    const double target = cum[nt] * c / G;
    while (i < nt && cum[i + 1] <= target) ++i;
    if (i < nt && (target - cum[i]) > (cum[i + 1] - target)) ++i;
    if (i < first[c - 1] + 1 && first[c - 1] < nt) i = first[c - 1] + 1;  // no empty runs while tiles remain
    first[c] = (int32_t)std::min<int64_t>(i, nt);
  }
  // runs of each instance and its slots
  std::vector<int> c_first(b->op.size(), -1), c_last(b->op.size(), -1);
  for (int c = 0; c < G; ++c)
    for (int64_t t = first[c]; t < first[c + 1]; ++t) {
      const int k = inst[t];
      if (c_first[k] < 0) c_first[k] = c;
      c_last[k] = c;
    }
  std::vector<VecSeg> segs(b->op.size());
  std::vector<int32_t> seg_of;
  int32_t pap0 = 0, blk0 = 0;
  for (int k : live) {
    VecSeg& sg = segs[k];
    sg.off = R.off[k];
    sg.n = b->g[k]->s.ndof;
    sg.blk0 = blk0;
    sg.nblk = vec_blocks(sg.n);
    sg.pap0 = pap0;
    sg.npap = c_first[k] < 0 ? 0 : c_last[k] - c_first[k] + 1;
    pap0 += sg.npap;
    blk0 += sg.nblk;
    for (int q = 0; q < sg.nblk; ++q) seg_of.push_back(k);
  }
  if (pap0 > R.pap_cap) {
    set_error("icc_batch_solve: too many p.Ap slots");
    return ICC_EINVAL;
  }
  std::vector<TileInst> ti(nt);
  for (int c = 0; c < G; ++c)
    for (int64_t t = first[c]; t < first[c + 1]; ++t) {
      const int k = inst[t];
      const GridSetup& s = b->g[k]->s;
      TileInst& q = ti[t];
      q.dbase = (int32_t)R.off[k];
      q.ebase = (int32_t)R.eoff[k];
      q.ncp0 = (int32_t)s.ncp[0];
      q.ncp1 = (int32_t)s.ncp[1];
      q.ndof_sc = (int32_t)s.ndof_sc;
      q.nel0 = (int32_t)s.nel[0];
      q.nel1 = (int32_t)s.nel[1];
      q.slot = segs[k].pap0 + (c - c_first[k]);
    }
  int rc;
  if ((rc = dalloc(&R.tiles, nt)) || (rc = dalloc(&R.first, G + 1)) || (rc = dalloc(&R.tmask, nt)) ||
      (rc = dalloc(&R.seg_of, seg_of.size())) || (rc = dalloc(&R.tinst, nt)))
    return rc;
  hipStream_t m = b->ctx->stream;
  ICC_HIP(hipMemcpyAsync(R.tiles, tl.data(), nt * sizeof(Tile), hipMemcpyHostToDevice, m));
  ICC_HIP(hipMemcpyAsync(R.first, first.data(), (G + 1) * sizeof(int32_t), hipMemcpyHostToDevice, m));
  ICC_HIP(hipMemcpyAsync(R.tmask, tm.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice, m));
  ICC_HIP(hipMemcpyAsync(R.seg_of, seg_of.data(), seg_of.size() * sizeof(int32_t), hipMemcpyHostToDevice, m));
  ICC_HIP(hipMemcpyAsync(R.tinst, ti.data(), nt * sizeof(TileInst), hipMemcpyHostToDevice, m));
  ICC_HIP(hipMemcpyAsync(R.segs, segs.data(), segs.size() * sizeof(VecSeg), hipMemcpyHostToDevice, m));
  ICC_HIP(hipStreamSynchronize(m));
  R.G = G;
  R.nblk = (int)seg_of.size();
  return ICC_OK;
}

// One batched PCG iteration: the apply of every planned instance in one
// launch, then the two update launches over their segments.
template <typename T>
static hipError_t batch_iteration(BatchRun<T>& R, hipStream_t m) {
  BatchApply ba;
  ba.tiles = R.tiles;
  ba.wg_first = R.first;
  ba.tinst = R.tinst;
  ba.tmask = R.tmask;
  ba.G = R.G;
  hipError_t e;
  if constexpr (std::is_same<T, double>::value) {
    ba.kfrag64 = R.kfrag;
    e = launch_apply_batch_f64(ba, R.p, R.ap, R.scale, R.fr, R.pap, m);
  } else {
    ba.kfrag32 = R.kfrag;
    e = launch_apply_batch_f32(ba, R.p, R.ap, R.scale, R.fr, R.pap, m);
  }
  if (e == hipSuccess)
    e = launch_pcg_update_batch<T>(R.nblk, R.segs, R.seg_of, R.st, R.pap, R.p, R.ap, R.x, R.r, R.dinv, R.parts, m);
  return e;
}

template <typename T>
static int batch_capture(BatchRun<T>& R, hipStream_t m, IterGraph& G) {
  ICC_HIP(hipStreamBeginCapture(m, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int i = 0; i < PCG_BATCH && e == hipSuccess; ++i) e = batch_iteration<T>(R, m);
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(m, &graph);
  if (e != hipSuccess) return hip_fail(e, "batched capture");
  if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
  G.graph = graph;
  ICC_HIP(hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0));
  return ICC_OK;
}

// FP64 true residual r = f - A u (masked) of instance k into state st.
static hipError_t batch_residual(icc_batch* b, int k, Work* w, PcgState* st) {
  hipError_t e = apply_t<double>(b->op[k], b->d_u[k], w->ap, nullptr);
  if (e == hipSuccess)
    e = launch_residual(b->g[k]->s.ndof, b->d_f[k], w->ap, b->op[k]->d_free, w->r, st, w->parts, b->ctx->stream);
  return e;
}

template <typename T>
static int batch_solve_t(icc_batch* b, double tol, double delta, int64_t maxit, double* u, double* stats) {
  constexpr bool MIXED = std::is_same<T, float>::value;
  const int nb = (int)b->op.size();
  hipStream_t m = b->ctx->stream;
  BatchRun<T> R;
  ICC_TRY(batch_setup<T>(b, R));
  std::vector<Work*> w(nb, nullptr);
  std::vector<int> status(nb, ICC_ENOCONV), finished(nb, 0);
  std::vector<double> nf(nb, 1.0), res_true(nb, 1.0), updates(nb, 0.0);
  std::vector<int64_t> iters(nb, 0);
  const int32_t mi = (int32_t)std::min<int64_t>(maxit > 0 ? maxit : (int64_t)2000000000, 2000000000);
  const double inner_tol = MIXED ? delta : tol;
  for (int k = 0; k < nb; ++k) {
    ICC_TRY(ensure_work(b->op[k], w[k]));
    const int64_t n = b->g[k]->s.ndof, o = R.off[k];
    ICC_TRY(icc_diag_dev(b->op[k], w[k]->tmp));
    ICC_HIP(launch_inv_diag<T>(w[k]->tmp, R.dinv + o, n, 1e-12, 1, m));
    ICC_HIP(launch_fill<double>(b->d_u[k], 0.0, n, m));
    ICC_HIP(hipMemcpyAsync(&R.st[k].tol, &inner_tol, sizeof(double), hipMemcpyHostToDevice, m));
    ICC_HIP(hipMemcpyAsync(&R.st[k].maxit, &mi, sizeof(int32_t), hipMemcpyHostToDevice, m));
    if (MIXED) ICC_HIP(batch_residual(b, k, w[k], R.st + k));
  }
  std::vector<int> live;
  if (MIXED) {
    ICC_HIP(hipMemcpyAsync(R.st_host, R.st, nb * sizeof(PcgState), hipMemcpyDeviceToHost, m));
    ICC_HIP(hipStreamSynchronize(m));
    for (int k = 0; k < nb; ++k) {
      const PcgState& hs = R.st_host[k];
      nf[k] = std::sqrt(hs.norm_b);
      res_true[k] = nf[k] > 0 ? std::sqrt(hs.rr) / nf[k] : std::sqrt(hs.rr);
      if (res_true[k] <= tol) {
        finished[k] = 1;
        status[k] = ICC_OK;
      }
    }
  }
  for (int k = 0; k < nb; ++k) {
    if (finished[k]) continue;
    const int64_t n = b->g[k]->s.ndof, o = R.off[k];
    if constexpr (MIXED)
      ICC_HIP((launch_pcg_init<float, double>(n, w[k]->r, 1.0, R.dinv + o, R.x + o, R.r + o, R.p + o, R.ap + o,
                                              R.st + k, w[k]->parts, m)));
    else
      ICC_HIP((launch_pcg_init<double, double>(n, b->d_f[k], 1.0, R.dinv + o, R.x + o, R.r + o, R.p + o, R.ap + o,
                                               R.st + k, w[k]->parts, m)));
    live.push_back(k);
  }
  // replays are bounded on the host as well: every replay advances each
  // running instance by PCG_BATCH iterations unless it stopped, and a stop
  // ends the instance or resets it through a reliable update
  int64_t replays = 0;
  const int64_t max_replays = 4 * ((int64_t)mi / PCG_BATCH + 2) * (MIXED ? 2 : 1) + 64;
  std::unique_ptr<IterGraph> G;
  std::vector<int> in_graph;
  const int wg = std::max(1, std::min(APPLY_MAX_WG_PER_CU, b->wg_per_cu[MIXED ? 1 : 0]));
  while (true) {
    std::vector<int> running;
    for (int k : live)
      if (!finished[k]) running.push_back(k);
    if (running.empty()) break;
    if (running != in_graph) {
      G.reset();
      ICC_TRY(batch_plan<T>(b, R, running, wg));
      G.reset(new IterGraph());
      ICC_TRY(batch_capture<T>(R, m, *G));
      in_graph = running;
    }
    if (++replays > max_replays) {
      set_error("icc_batch_solve: replay limit reached");
      break;
    }
    ICC_HIP(hipGraphLaunch(G->exec, m));
    ICC_HIP(hipMemcpyAsync(R.st_host, R.st, nb * sizeof(PcgState), hipMemcpyDeviceToHost, m));
    ICC_HIP(hipStreamSynchronize(m));
    std::vector<int> pend;
    for (int k : running) {
      const PcgState& hs = R.st_host[k];
      iters[k] = hs.iters;
      if (hs.done == 0) continue;
      if (!MIXED) {
        finished[k] = 1;
        status[k] = hs.done == 1 ? ICC_OK : hs.done == 2 ? ICC_EBREAKDOWN : ICC_ENOCONV;
      } else {
        pend.push_back(k);
      }
    }
    if constexpr (MIXED) {
      // reliable updates: u += x32, FP64 true residual
      std::vector<int> done_code(nb, 0);
      for (int k : pend) {
        done_code[k] = R.st_host[k].done;
        const int64_t n = b->g[k]->s.ndof;
        ICC_HIP(launch_ir_update<float>(n, b->d_u[k], R.x + R.off[k], 1.0, m));
        ICC_HIP(batch_residual(b, k, w[k], R.st + k));
        updates[k] += 1.0;
      }
      if (!pend.empty()) {
        ICC_HIP(hipMemcpyAsync(R.st_host, R.st, nb * sizeof(PcgState), hipMemcpyDeviceToHost, m));
        ICC_HIP(hipStreamSynchronize(m));
      }
      for (int k : pend) {
        const PcgState& hr = R.st_host[k];
        res_true[k] = nf[k] > 0 ? std::sqrt(hr.rr) / nf[k] : std::sqrt(hr.rr);
        if (res_true[k] <= tol || done_code[k] == 3 || done_code[k] == 2) {
          finished[k] = 1;
          status[k] = res_true[k] <= tol ? ICC_OK : done_code[k] == 2 ? ICC_EBREAKDOWN : ICC_ENOCONV;
        } else {
          const int64_t n = b->g[k]->s.ndof, o = R.off[k];
          ICC_HIP(launch_reliable_replace(n, w[k]->r, R.dinv + o, R.r + o, R.x + o, R.st + k, w[k]->parts, m));
          ICC_HIP(launch_pdir_update(n, R.st + k, R.p + o, R.r + o, R.dinv + o, R.ap + o, m));
        }
      }
    }
  }
  ICC_HIP(hipStreamSynchronize(m));
  G.reset();
  // FP64: the solution is the recurrence's x; true residuals; copy out
  int64_t off = 0;
  for (int k = 0; k < nb; ++k) {
    const int64_t n = b->g[k]->s.ndof;
    if (!MIXED) {
      ICC_HIP(hipMemcpyAsync(b->d_u[k], R.x + R.off[k], n * 8, hipMemcpyDeviceToDevice, m));
      ICC_HIP(batch_residual(b, k, w[k], R.st + k));
      ICC_HIP(hipMemcpyAsync(R.st_host + k, R.st + k, sizeof(PcgState), hipMemcpyDeviceToHost, m));
      ICC_HIP(hipStreamSynchronize(m));
      const PcgState& hr = R.st_host[k];
      const double nfk = std::sqrt(hr.norm_b);
      res_true[k] = nfk > 0 ? std::sqrt(hr.rr) / nfk : std::sqrt(hr.rr);
    }
    if (u) ICC_HIP(hipMemcpyAsync(u + off, b->d_u[k], n * 8, hipMemcpyDeviceToHost, m));
    if (stats) {
      stats[4 * k] = (double)iters[k];
      stats[4 * k + 1] = updates[k];
      stats[4 * k + 2] = res_true[k];
      stats[4 * k + 3] = (double)status[k];
    }
    off += n;
  }
  ICC_HIP(hipStreamSynchronize(m));
  for (int k = 0; k < nb; ++k)
    if (status[k] != ICC_OK) {
      set_error("icc_batch_solve: an instance did not converge (see stats)");
      return status[k];
    }
  return ICC_OK;
}
}  // namespace icc

ICC_API int icc_batch_solve(icc_batch* b, int prec, double tol, double delta, int64_t maxit, double* u,
                            double* stats) {
  if (!b || (prec != ICC_F64 && prec != ICC_F32) || !(tol > 0.0)) return ICC_EINVAL;
  if (prec == ICC_F32 && !(delta > 0.0 && delta < 1.0)) return ICC_EINVAL;
  return prec == ICC_F32 ? batch_solve_t<float>(b, tol, delta, maxit, u, stats)
                         : batch_solve_t<double>(b, tol, delta, maxit, u, stats);
}
