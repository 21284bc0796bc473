// Weighted-quadrature stiffness formation above the C ABI: the GPU
// counterpart of wq_stiffness / wq_stiffness_layout (rust/src/wq.rs:324-514,
// interior or Calabro point layout, p = 2).  The per-direction 1D rules and formation tables
// are built on the host once per handle (wq.cpp); forming the matrix for a
// density field runs three kernels (k_wq.hip): the SIMP factor at the WQ
// points, the sum-factorised rows, and the in-place symmetrisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <vector>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

#define ICC_API extern "C" __attribute__((visibility("default")))

using namespace icc;

struct icc_wq {
  icc_grid* g = nullptr;
  double* d_S = nullptr;      // [9 * 125][ndof_sc] stencil matrix
  int np[3] = {0, 0, 0};      // WQ points per direction
  double* d_simp = nullptr;   // [np0 np1 np2] SIMP factor at the WQ points
  double* d_tabs = nullptr;   // formation tables, directions 0, 1, 2
  int* d_pstart = nullptr;    // first WQ point of each row, per direction
  int* d_pbin = nullptr;      // element of each WQ point, per direction
  int* d_pfirst = nullptr;    // first nonzero basis index at each WQ point, per direction
  double* d_pval = nullptr;   // the 3 nonzero basis values there
  double* d_x = nullptr;      // density upload buffer
  double* d_v = nullptr;      // apply buffers
  double* d_w = nullptr;
  int formed = 0;
};

namespace {

template <typename T>
int dalloc(T** p, size_t n) {
  if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    set_error("hipMalloc failed");
    return ICC_ENOMEM;
  }
  return ICC_OK;
}

#define ICC_TRY(call)                  \
  do {                                 \
    int rc__ = (call);                 \
    if (rc__ != ICC_OK) return rc__;   \
  } while (0)

}  // namespace

ICC_API int icc_wq_destroy(icc_wq* m) {
  if (!m) return ICC_OK;
  hipFree(m->d_S);
  hipFree(m->d_simp);
  hipFree(m->d_tabs);
  hipFree(m->d_pstart);
  hipFree(m->d_pbin);
  hipFree(m->d_pfirst);
  hipFree(m->d_pval);
  hipFree(m->d_x);
  hipFree(m->d_v);
  hipFree(m->d_w);
  delete m;
  return ICC_OK;
}

ICC_API int icc_wq_create_layout(icc_grid* g, int layout, icc_wq** out) {
  if (!g || !out || (layout != ICC_WQ_INTERIOR && layout != ICC_WQ_CALABRO)) return ICC_EINVAL;
  const GridSetup& s = g->s;
  if (layout == ICC_WQ_CALABRO && (s.nel[0] < 2 || s.nel[1] < 2 || s.nel[2] < 2)) {
    set_error("icc_wq_create_layout: the Calabro layout needs at least 2 elements per direction (one element "
              "gives a singular system for the middle function's rule)");
    return ICC_ENOTSUP;
  }
  auto* m = new icc_wq();
  m->g = g;
  std::vector<double> tabs, pval;
  std::vector<int> pstart, pbin, pfirst;
  for (int d = 0; d < 3; ++d) {
    WqRulesP2 r;
    if (layout == ICC_WQ_CALABRO)
      wq_rules_calabro_p2(s.nel[d], r);
    else
      wq_rules_p2(s.nel[d], r);
    std::vector<double> t;
    std::vector<int> ps;
    try {
      wq_form_tables(r, t, ps);
    } catch (const std::exception& ex) {
      delete m;
      set_error(ex.what());
      return ICC_ENOTSUP;
    }
    tabs.insert(tabs.end(), t.begin(), t.end());
    pstart.insert(pstart.end(), ps.begin(), ps.end());
    m->np[d] = r.npt;
    // the point's element holds every nonzero function there (a knot belongs
    // to its right element, where the left neighbour's last function is zero)
    for (int q = 0; q < r.npt; ++q) {
      const int e = r.bin[q];
      pbin.push_back(e);
      pfirst.push_back(e);
      for (int a = 0; a < 3; ++a) pval.push_back(r.b[(size_t)q * r.ndof + e + a]);
    }
  }
  const int64_t nsc = s.ndof_sc, npt = (int64_t)m->np[0] * m->np[1] * m->np[2];
  int rc;
  if ((rc = dalloc(&m->d_S, (size_t)9 * 125 * nsc)) || (rc = dalloc(&m->d_simp, (size_t)npt)) ||
      (rc = dalloc(&m->d_tabs, tabs.size())) || (rc = dalloc(&m->d_pstart, pstart.size())) ||
      (rc = dalloc(&m->d_pbin, pbin.size())) || (rc = dalloc(&m->d_pfirst, pfirst.size())) ||
      (rc = dalloc(&m->d_pval, pval.size())) || (rc = dalloc(&m->d_x, std::max<int64_t>(s.nel_tot, nsc))) ||
      (rc = dalloc(&m->d_v, s.ndof)) || (rc = dalloc(&m->d_w, s.ndof))) {
    icc_wq_destroy(m);
    return rc;
  }
  hipStream_t st = g->ctx->stream;
  hipError_t e;
  if ((e = hipMemcpyAsync(m->d_tabs, tabs.data(), tabs.size() * 8, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(m->d_pstart, pstart.data(), pstart.size() * 4, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(m->d_pbin, pbin.data(), pbin.size() * 4, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(m->d_pfirst, pfirst.data(), pfirst.size() * 4, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(m->d_pval, pval.data(), pval.size() * 8, hipMemcpyHostToDevice, st)) ||
      (e = hipStreamSynchronize(st))) {
    icc_wq_destroy(m);
    return hip_fail(e, "icc_wq_create");
  }
  *out = m;
  return ICC_OK;
}

ICC_API int icc_wq_create(icc_grid* g, icc_wq** out) { return icc_wq_create_layout(g, ICC_WQ_INTERIOR, out); }

ICC_API double icc_debug_quantize_kmat(const double* k, int bits, double* out) {
  if (!k || !out || bits < 2 || bits > 50) return -1.0;
  return icc::quantize_kmat(k, bits, out);
}

ICC_API int64_t icc_debug_wq_rule(int64_t nel, int layout, int which, int64_t i, double* out) {
  if (nel < 1 || (layout != ICC_WQ_INTERIOR && layout != ICC_WQ_CALABRO) || which < 0 || which > 4)
    return -(int64_t)ICC_EINVAL;
  if (layout == ICC_WQ_CALABRO && nel < 2) return -(int64_t)ICC_ENOTSUP;
  WqRulesP2 r;
  if (layout == ICC_WQ_CALABRO)
    wq_rules_calabro_p2(nel, r);
  else
    wq_rules_p2(nel, r);
  if (which > 0 && (i < 0 || i >= r.ndof)) return -(int64_t)ICC_EINVAL;
  const std::vector<double>* w[4] = {&r.w0[0], &r.w1[0], &r.w2[0], &r.w3[0]};
  const std::vector<double>& v = which == 0 ? r.pts : w[which - 1][i];
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(double));
  return (int64_t)v.size();
}

ICC_API int icc_wq_form_dev(icc_wq* m, int density, const double* x_dev, double penal, double emin) {
  if (!m || density < ICC_DENSITY_NONE || density > ICC_DENSITY_SPLINE || (density != ICC_DENSITY_NONE && !x_dev))
    return ICC_EINVAL;
  hipStream_t st = m->g->ctx->stream;
  ICC_HIP(launch_wq_simp(m->g, density, m->np, x_dev, penal, emin, m->d_pbin, m->d_pfirst, m->d_pval, m->d_simp,
                         st));
  ICC_HIP(launch_wq_form(m->g, m->np, m->d_pstart, m->d_tabs, m->d_simp, m->d_S, st));
  m->formed = 1;
  return ICC_OK;
}

ICC_API int icc_wq_form(icc_wq* m, int density, const double* x, double penal, double emin) {
  if (!m || density < ICC_DENSITY_NONE || density > ICC_DENSITY_SPLINE || (density != ICC_DENSITY_NONE && !x))
    return ICC_EINVAL;
  const GridSetup& s = m->g->s;
  hipStream_t st = m->g->ctx->stream;
  if (density != ICC_DENSITY_NONE) {
    const int64_t n = density == ICC_DENSITY_ELEMENT ? s.nel_tot : s.ndof_sc;
    ICC_HIP(hipMemcpyAsync(m->d_x, x, n * 8, hipMemcpyHostToDevice, st));
  }
  ICC_TRY(icc_wq_form_dev(m, density, density == ICC_DENSITY_NONE ? nullptr : m->d_x, penal, emin));
  ICC_HIP(hipStreamSynchronize(st));
  return ICC_OK;
}

ICC_API int icc_wq_apply_dev(icc_wq* m, const double* x_dev, double* y_dev) {
  if (!m || !x_dev || !y_dev) return ICC_EINVAL;
  if (!m->formed) {
    set_error("icc_wq_apply: the matrix has not been formed");
    return ICC_EINVAL;
  }
  ICC_HIP(launch_wq_spmv(m->g, m->d_S, x_dev, y_dev, m->g->ctx->stream));
  return ICC_OK;
}

ICC_API int icc_wq_apply(icc_wq* m, const double* x, double* y) {
  if (!m || !x || !y) return ICC_EINVAL;
  const int64_t n = m->g->s.ndof;
  hipStream_t st = m->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(m->d_v, x, n * 8, hipMemcpyHostToDevice, st));
  ICC_TRY(icc_wq_apply_dev(m, m->d_v, m->d_w));
  ICC_HIP(hipMemcpyAsync(y, m->d_w, n * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  return ICC_OK;
}

namespace {
// stencil entry o of control point P -> neighbour index, or -1 outside the grid
inline int64_t neighbour(const GridSetup& s, int64_t P, int o) {
  const int64_t x = P % s.ncp[0] + o % 5 - 2, y = (P / s.ncp[0]) % s.ncp[1] + (o / 5) % 5 - 2,
                z = P / (s.ncp[0] * s.ncp[1]) + o / 25 - 2;
  if (x < 0 || y < 0 || z < 0 || x >= s.ncp[0] || y >= s.ncp[1] || z >= s.ncp[2]) return -1;
  return x + s.ncp[0] * (y + s.ncp[1] * z);
}
}  // namespace

// CSR of the formed matrix: rows and columns component-blocked (c * ndof_sc
// + control point), columns ascending within a row, every in-grid stencil
// entry stored (the reference's pattern: all trial functions whose support
// meets the test function's WQ points).  indptr / indices / data may be NULL
// to query *nnz only.
ICC_API int icc_wq_csr(icc_wq* m, int64_t* nnz, int64_t* indptr, int64_t* indices, double* data) {
  if (!m || !nnz) return ICC_EINVAL;
  if (!m->formed) {
    set_error("icc_wq_csr: the matrix has not been formed");
    return ICC_EINVAL;
  }
  const GridSetup& s = m->g->s;
  const int64_t nsc = s.ndof_sc;
  int64_t per_sc = 0;
  std::vector<int> cnt(nsc);
  for (int64_t P = 0; P < nsc; ++P) {
    int c = 0;
    for (int o = 0; o < 125; ++o) c += neighbour(s, P, o) >= 0;
    cnt[P] = c;
    per_sc += c;
  }
  *nnz = 9 * per_sc;
  if (!indptr || !indices || !data) return ICC_OK;
  std::vector<double> S((size_t)9 * 125 * nsc);
  hipStream_t st = m->g->ctx->stream;
  ICC_HIP(hipMemcpyAsync(S.data(), m->d_S, S.size() * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  int64_t k = 0;
  indptr[0] = 0;
  for (int i = 0; i < 3; ++i)
    for (int64_t P = 0; P < nsc; ++P) {
      for (int j = 0; j < 3; ++j)
        for (int o = 0; o < 125; ++o) {
          const int64_t Q = neighbour(s, P, o);
          if (Q < 0) continue;
          indices[k] = j * nsc + Q;
          data[k] = S[((size_t)(3 * i + j) * 125 + o) * nsc + P];
          ++k;
        }
      indptr[i * nsc + P + 1] = k;
    }
  return ICC_OK;
}
