// Immersed solve with an analytic geometry above the C ABI: the host side of
// build_immersed_problem / solve_immersed_problem (rust/src/immersed.rs:466-570)
// for a box minus cylinders parallel to z, on a p = 2 background grid aligned
// to the box.  Classification and volume fractions run on the GPU; the scale,
// support mask, penalty face term (CSR) and traction are built on the host
// once; the solve is the masked operator sum_e s_e K_e + P through the MFMA
// apply plus the sparse face term.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

#define ICC_API extern "C" __attribute__((visibility("default")))

using namespace icc;

namespace icc {
void traction_xmax(const GridSetup& s, const double t[3], std::vector<double>& f);
void penalty_xmin(const GridSetup& s, double beta, const std::vector<uint8_t>& face_on, std::vector<int64_t>& rows,
                  std::vector<int64_t>& cols, std::vector<double>& vals);
}

struct icc_immersed {
  icc_ctx* ctx = nullptr;
  icc_immersed_config cfg;
  icc_grid* g = nullptr;
  icc_op* op = nullptr;
  std::vector<int8_t> status;
  std::vector<double> w, scale, force;
  std::vector<uint8_t> active, fr;
  double beta = 0.0;
  int64_t n_active = 0, n_free = 0, nnz_pen = 0;
  double *d_f = nullptr, *d_u = nullptr, *d_diag = nullptr, *d_tmp = nullptr;
};

namespace {

template <typename T>
int alloc_dev(T** p, size_t n) {
  if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    set_error("hipMalloc failed");
    return ICC_ENOMEM;
  }
  return ICC_OK;
}

// Sorted CSR of the triplets (duplicates summed in triplet order).
void triplets_to_csr(int64_t n, const std::vector<int64_t>& r, const std::vector<int64_t>& c,
                     const std::vector<double>& v, std::vector<int64_t>& indptr, std::vector<int64_t>& indices,
                     std::vector<double>& data) {
  std::vector<int64_t> order(r.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return r[a] != r[b] ? r[a] < r[b] : c[a] < c[b]; });
  indptr.assign(n + 1, 0);
  indices.clear();
  data.clear();
  for (size_t q = 0; q < order.size(); ++q) {
    const int64_t k = order[q];
    if (!indices.empty() && q > 0 && r[order[q - 1]] == r[k] && c[order[q - 1]] == c[k]) {
      data.back() += v[k];
      continue;
    }
    indices.push_back(c[k]);
    data.push_back(v[k]);
    indptr[r[k] + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) indptr[i + 1] += indptr[i];
}

}  // namespace

ICC_API void icc_immersed_config_default(icc_immersed_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->box_hi[0] = 2.0;
  c->box_hi[1] = 1.0;
  c->box_hi[2] = 1.0;
  c->res[0] = 32;
  c->res[1] = 16;
  c->res[2] = 16;
  c->sub[0] = c->sub[1] = c->sub[2] = 4;
  c->young = 1.0;
  c->poisson = 0.3;
  c->emin = 1e-4;
  c->active_min = 0.05;
  c->penalty_factor = 1e3;
  c->traction[1] = -1.0;
  c->pcg_tol = 1e-8;
  c->pcg_maxit = 0;
  c->mode = 1;
  c->solver = 0;
}

ICC_API int icc_immersed_destroy(icc_immersed* h) {
  if (!h) return ICC_OK;
  icc_op_destroy(h->op);
  icc_grid_destroy(h->g);
  hipFree(h->d_f);
  hipFree(h->d_u);
  hipFree(h->d_diag);
  hipFree(h->d_tmp);
  delete h;
  return ICC_OK;
}

ICC_API int icc_immersed_create(icc_ctx* ctx, const icc_immersed_config* cfg, const double* cyl, int64_t ncyl,
                                icc_immersed** out) {
  if (!ctx || !cfg || !out || (ncyl > 0 && !cyl) || ncyl < 0) return ICC_EINVAL;
  if (cfg->mode != 0 && cfg->mode != 1) {
    set_error("icc_immersed_create: mode must be 0 or 1");
    return ICC_EINVAL;
  }
  if (cfg->solver < 0 || cfg->solver > 2) {
    set_error("icc_immersed_create: solver must be 0, 1 or 2");
    return ICC_EINVAL;
  }
  for (int d = 0; d < 3; ++d)
    if (cfg->sub[d] < 1 || cfg->res[d] < 1) {
      set_error("icc_immersed_create: res and sub must be >= 1");
      return ICC_EINVAL;
    }
  auto* h = new icc_immersed();
  h->ctx = ctx;
  h->cfg = *cfg;
  const double gb[6] = {cfg->box_lo[0], cfg->box_hi[0], cfg->box_lo[1], cfg->box_hi[1], cfg->box_lo[2], cfg->box_hi[2]};
  int rc = icc_grid_create(ctx, cfg->res, gb, 2, cfg->young, cfg->poisson, &h->g);
  if (rc == ICC_OK) rc = icc_op_create(h->g, &h->op);
  if (rc != ICC_OK) {
    icc_immersed_destroy(h);
    return rc;
  }
  const GridSetup& s = h->g->s;
  const int64_t ne = s.nel_tot, n = s.ndof;
  h->status.resize(ne);
  h->w.resize(ne);
  rc = icc_classify_cylinders(ctx, cfg->box_lo, cfg->box_hi, cyl, ncyl, gb, cfg->res, cfg->sub, h->status.data(),
                              h->w.data());
  if (rc != ICC_OK) {
    icc_immersed_destroy(h);
    return rc;
  }
  // scale and active set (immersed.rs:484; topopt_immersed.rs:84)
  h->scale.assign(ne, 0.0);
  h->active.assign(ne, 0);
  for (int64_t e = 0; e < ne; ++e) {
    if (cfg->mode == 0) {
      h->scale[e] = cfg->emin + (1.0 - cfg->emin) * h->w[e];
      h->active[e] = 1;
    } else if (h->w[e] >= cfg->active_min) {
      h->scale[e] = h->w[e];
      h->active[e] = 1;
    }
    h->n_active += h->active[e];
  }
  // DOFs supported by a cell of nonzero scale stay free
  h->fr.assign(n, 0);
  for (int64_t k = 0; k < s.nel[2]; ++k)
    for (int64_t j = 0; j < s.nel[1]; ++j)
      for (int64_t i = 0; i < s.nel[0]; ++i) {
        if (!(h->scale[i + s.nel[0] * (j + s.nel[1] * k)] > 0.0)) continue;
        for (int az = 0; az < 3; ++az)
          for (int ay = 0; ay < 3; ++ay)
            for (int ax = 0; ax < 3; ++ax) {
              const int64_t g = (i + ax) + s.ncp[0] * ((j + ay) + s.ncp[1] * (k + az));
              for (int c = 0; c < 3; ++c) h->fr[c * s.ndof_sc + g] = 1;
            }
      }
  for (int64_t i = 0; i < n; ++i) h->n_free += h->fr[i];
  // penalty u = 0 on x = box_lo[0] (immersed_bc.rs:161-190 with g = 0)
  double hmin = 1e300;
  for (int d = 0; d < 3; ++d) hmin = std::fmin(hmin, (cfg->box_hi[d] - cfg->box_lo[d]) / (double)cfg->res[d]);
  h->beta = cfg->penalty_factor * cfg->young / hmin;
  std::vector<uint8_t> face_on(s.nel[1] * s.nel[2], 0);
  for (int64_t k = 0; k < s.nel[2]; ++k)
    for (int64_t j = 0; j < s.nel[1]; ++j) face_on[j + s.nel[1] * k] = h->scale[s.nel[0] * (j + s.nel[1] * k)] > 0.0;
  std::vector<int64_t> pr, pc, ip, ix;
  std::vector<double> pv, pd;
  penalty_xmin(s, h->beta, face_on, pr, pc, pv);
  triplets_to_csr(n, pr, pc, pv, ip, ix, pd);
  h->nnz_pen = (int64_t)pd.size();
  // traction on x = box_hi[0], masked
  traction_xmax(s, cfg->traction, h->force);
  for (int64_t i = 0; i < n; ++i)
    if (!h->fr[i]) h->force[i] = 0.0;
  if ((rc = icc_op_set_scale(h->op, h->scale.data())) || (rc = icc_op_set_free_mask(h->op, h->fr.data())) ||
      (rc = icc_op_set_sparse(h->op, n, ip.data(), ix.data(), pd.data())) || (rc = alloc_dev(&h->d_f, n)) ||
      (rc = alloc_dev(&h->d_u, n)) || (rc = alloc_dev(&h->d_diag, n)) || (rc = alloc_dev(&h->d_tmp, n))) {
    icc_immersed_destroy(h);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(h->d_f, h->force.data(), n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    icc_immersed_destroy(h);
    return hip_fail(e, "icc_immersed_create");
  }
  *out = h;
  return ICC_OK;
}

ICC_API int icc_immersed_sizes(const icc_immersed* h, int64_t out[5], double* beta) {
  if (!h || !out) return ICC_EINVAL;
  out[0] = h->g->s.nel_tot;
  out[1] = h->g->s.ndof;
  out[2] = h->n_active;
  out[3] = h->n_free;
  out[4] = h->nnz_pen;
  if (beta) *beta = h->beta;
  return ICC_OK;
}

ICC_API int icc_immersed_problem(const icc_immersed* h, int8_t* status, double* weights, uint8_t* active,
                                 double* scale, uint8_t* free_mask, double* force) {
  if (!h) return ICC_EINVAL;
  const int64_t ne = h->g->s.nel_tot, n = h->g->s.ndof;
  if (status) std::memcpy(status, h->status.data(), ne);
  if (weights) std::memcpy(weights, h->w.data(), ne * 8);
  if (active) std::memcpy(active, h->active.data(), ne);
  if (scale) std::memcpy(scale, h->scale.data(), ne * 8);
  if (free_mask) std::memcpy(free_mask, h->fr.data(), n);
  if (force) std::memcpy(force, h->force.data(), n * 8);
  return ICC_OK;
}

ICC_API int icc_immersed_apply(icc_immersed* h, const double* x, double* y) {
  if (!h || !x || !y) return ICC_EINVAL;
  return icc_apply(h->op, ICC_F64, x, y);
}

ICC_API int icc_immersed_solve_dev(icc_immersed* h, double* u, double* stats) {
  if (!h || !u) return ICC_EINVAL;
  const GridSetup& s = h->g->s;
  const int64_t n = s.ndof;
  hipStream_t st = h->ctx->stream;
  const int64_t maxit = h->cfg.pcg_maxit > 0 ? h->cfg.pcg_maxit : 20 * n;
  double st8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int rc;
  if (h->cfg.solver == 0) {
    // diag = (K + P).diagonal().max(1e-6) (immersed.rs:555)
    if ((rc = icc_diag_dev(h->op, h->d_tmp)) != ICC_OK) return rc;
    ICC_HIP(launch_clamp_min(n, h->d_tmp, 1e-6, h->d_diag, st));
    int64_t iters = 0;
    double res = 0.0;
    rc = icc_pcg_dev(h->op, h->d_f, h->d_diag, h->cfg.pcg_tol, maxit, u, &iters, &res);
    if (rc != ICC_OK && rc != ICC_ENOCONV && rc != ICC_EBREAKDOWN) return rc;
    st8[0] = (double)iters;
    st8[1] = res;
    st8[4] = (double)iters;
  } else {
    ICC_HIP(launch_fill<double>(u, 0.0, n, st));
    rc = icc_solve_mixed_prec_dev(h->op, h->cfg.solver == 2 ? ICC_F16X : ICC_F32, h->d_f, h->cfg.pcg_tol, 0.1, maxit,
                                  u, st8);
    if (rc != ICC_OK && rc != ICC_ENOCONV && rc != ICC_EBREAKDOWN) return rc;
    const double upd = st8[0], its = st8[1], f64 = st8[3];
    st8[0] = its;
    st8[1] = upd;
    st8[4] = f64;
  }
  // true relative residual and compliance on the host
  std::vector<double> uh(n), au(n);
  ICC_HIP(hipMemcpyAsync(uh.data(), u, n * 8, hipMemcpyDeviceToHost, st));
  int rc2 = icc_apply_dev(h->op, ICC_F64, u, h->d_tmp);
  if (rc2 != ICC_OK) return rc2;
  ICC_HIP(hipMemcpyAsync(au.data(), h->d_tmp, n * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  double rr = 0.0, ff = 0.0, c = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    if (!h->fr[i]) continue;
    const double r = h->force[i] - au[i];
    rr += r * r;
    ff += h->force[i] * h->force[i];
    c += h->force[i] * uh[i];
  }
  st8[2] = ff > 0.0 ? std::sqrt(rr / ff) : std::sqrt(rr);
  st8[3] = c;
  if (stats) std::memcpy(stats, st8, sizeof(st8));
  return rc;
}

ICC_API int icc_immersed_solve(icc_immersed* h, double* u, double* stats) {
  if (!h || !u) return ICC_EINVAL;
  const int64_t n = h->g->s.ndof;
  int rc = icc_immersed_solve_dev(h, h->d_u, stats);
  if (rc != ICC_OK && rc != ICC_ENOCONV && rc != ICC_EBREAKDOWN) return rc;
  ICC_HIP(hipMemcpyAsync(u, h->d_u, n * 8, hipMemcpyDeviceToHost, h->ctx->stream));
  ICC_HIP(hipStreamSynchronize(h->ctx->stream));
  return rc;
}
