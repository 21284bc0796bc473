// Structured SIMP cantilever driver above the C ABI: the host side of
// CantileverTopOpt (rust/src/topopt_iga.rs) with the state solve, energies,
// sensitivities, filter and optimality-criteria bisection on the GPU, in the
// reference's loop order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

#define ICC_API extern "C" __attribute__((visibility("default")))

using namespace icc;

namespace icc {
// Consistent traction F(A, c) = t_c * int_{x = x_max} N_A dA (3-point Gauss per face element).
void traction_xmax(const GridSetup& s, const double t[3], std::vector<double>& f);
}

struct icc_topopt {
  icc_ctx* ctx = nullptr;
  icc_topopt_config cfg;
  icc_grid* g = nullptr;
  icc_op* op = nullptr;
  icc_filter* filt = nullptr;
  int64_t nel = 0, ndof = 0;
  std::vector<double> force;
  std::vector<uint8_t> fr;
  // device buffers
  double *x = nullptr, *xt = nullptr, *s = nullptr, *u = nullptr, *u0 = nullptr, *b = nullptr, *du = nullptr,
         *dg = nullptr, *f = nullptr, *dc = nullptr, *tmp = nullptr, *ones = nullptr, *dv = nullptr;
  uint8_t* fr_d = nullptr;
  PcgState* st = nullptr;
  PcgState* st_h = nullptr;
  double* parts = nullptr;
  // optimality-criteria update: previous and candidate designs, bisection partials and state
  double *xold = nullptr, *xnew = nullptr, *pmax = nullptr, *ocparts = nullptr, *sparts = nullptr;
  OcState* ocs = nullptr;
  OcState* ocs_h = nullptr;
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

#define TRY(call)                      \
  do {                                 \
    int rc__ = (call);                 \
    if (rc__ != ICC_OK) return rc__;   \
  } while (0)

int host_sync_state(icc_topopt* t, PcgState& out) {
  hipStream_t st = t->ctx->stream;
  ICC_HIP(hipMemcpyAsync(t->st_h, t->st, sizeof(PcgState), hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  out = *t->st_h;
  return ICC_OK;
}

// solve_state on device: x_dev (filtered densities) -> u_dev, warm start u0_dev.
int solve_state_dev(icc_topopt* t, const double* x_dev, const double* u0_dev, double* u_dev, int64_t* iters,
                    double* res_out) {
  const icc_topopt_config& c = t->cfg;
  hipStream_t st = t->ctx->stream;
  const int64_t n = t->ndof;
  // scale = emin + (1 - emin) x^penal (topopt_iga.rs:132)
  ICC_HIP(launch_simp_scale(t->nel, x_dev, c.penal, c.emin, nullptr, nullptr, t->s, st));
  TRY(icc_op_set_scale_dev(t->op, t->s));
  // b = free ? F - K u0 : 0 ; rb = |b|, nb = |F_free| (topopt_iga.rs:142-146)
  TRY(icc_apply_dev(t->op, ICC_F64, u0_dev, t->tmp));
  ICC_HIP(launch_residual(n, t->f, t->tmp, t->fr_d, t->b, t->st, t->parts, st));
  PcgState hs;
  TRY(host_sync_state(t, hs));
  const double rb = std::sqrt(hs.rr), nb = std::sqrt(hs.norm_b);
  // diag = K.diagonal().max(1e-12) (topopt_iga.rs:147)
  TRY(icc_diag_dev(t->op, t->tmp));
  ICC_HIP(launch_clamp_min(n, t->tmp, 1e-12, t->dg, st));
  const double tol = rb > 0.0 ? std::fmin(c.pcg_tol * nb / rb, 0.5) : 1.0;
  int64_t its = 0;
  double res = 0.0;
  if (c.solver == 0) {
    int rc = icc_pcg_dev(t->op, t->b, t->dg, tol, 20 * n, t->du, &its, &res);
    if (rc != ICC_OK && rc != ICC_EBREAKDOWN && rc != ICC_ENOCONV) return rc;
  } else {
    ICC_HIP(launch_fill<double>(t->du, 0.0, n, st));
    double stats[8];
    int rc = icc_solve_mixed_prec_dev(t->op, c.solver == 2 ? ICC_F16X : ICC_F32, t->b, tol, 0.1, 20 * n, t->du,
                                      stats);
    if (rc != ICC_OK && rc != ICC_ENOCONV && rc != ICC_EBREAKDOWN) return rc;
    its = (int64_t)stats[1];
    res = stats[2];
  }
  // u = free ? u0 + du : 0
  ICC_HIP(launch_masked_add(n, u0_dev, t->du, t->fr_d, u_dev, st));
  *iters = its;
  *res_out = res * rb / std::fmax(nb, 1e-300);
  return ICC_OK;
}

// t->x (device, design) -> t->xt (device, physical densities)
int physical_densities_dev(icc_topopt* t) {
  hipStream_t st = t->ctx->stream;
  if (t->cfg.filter == 1)
    TRY(icc_filter_fwd_dev(t->filt, t->x, t->xt));
  else
    ICC_HIP(hipMemcpyAsync(t->xt, t->x, t->nel * 8, hipMemcpyDeviceToDevice, st));
  return ICC_OK;
}

// x (host, design) -> t->xt (device, physical densities)
int physical_densities(icc_topopt* t, const double* x_host) {
  ICC_HIP(hipMemcpyAsync(t->x, x_host, t->nel * 8, hipMemcpyHostToDevice, t->ctx->stream));
  return physical_densities_dev(t);
}

// OC bisection (topopt_iga.rs:206-222) on the device: t->xnew = the candidate
// of the last bisection step from t->xold and the filtered sensitivities dc.
// The sensitivity filter bisects on sum x = volfrac nel. The density filter
// bisects on sum H x / Hs = sum dv x with dv = H^T (1 / Hs), and divides dc by
// dv under the square root.
int oc_update_dev(icc_topopt* t, const double* dc) {
  const icc_topopt_config& cf = t->cfg;
  hipStream_t st = t->ctx->stream;
  const int64_t ne = t->nel;
  const double* dv = cf.filter == 1 ? t->dv : nullptr;
  const double target = cf.volfrac * (double)ne;
  ICC_HIP(launch_oc_init(ne, nullptr, dc, t->xold, t->xnew, t->pmax, st));
  constexpr int CHUNK = 8, MAX_STEPS = 4096;
  int k = 0;
  for (;;) {
    for (int q = 0; q < CHUNK; ++q, ++k)
      ICC_HIP(launch_oc_step(k, ne, nullptr, dv, dv, dc, t->xold, t->xnew, target, cf.move_limit, t->pmax,
                             t->ocparts, t->ocs, st));
    ICC_HIP(hipMemcpyAsync(t->ocs_h, t->ocs + (k & 1), sizeof(OcState), hipMemcpyDeviceToHost, st));
    ICC_HIP(hipStreamSynchronize(st));
    if (t->ocs_h->done) return ICC_OK;
    if (k >= MAX_STEPS) {
      set_error("icc_topopt_run: OC bisection did not terminate");
      return ICC_ENOCONV;
    }
  }
}

double dot_host(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

}  // namespace

ICC_API void icc_topopt_config_default(icc_topopt_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->nel[0] = 16;
  c->nel[1] = 8;
  c->nel[2] = 4;
  c->size[0] = 1.2;
  c->size[1] = 0.6;
  c->size[2] = 0.3;
  c->volfrac = 0.3;
  c->penal = 3.0;
  c->rmin = 1.5;
  c->max_iter = 40;
  c->young = 1.0;
  c->poisson = 0.3;
  c->emin = 1e-3;
  c->move_limit = 0.2;
  c->pcg_tol = 1e-10;
}

ICC_API int icc_topopt_destroy(icc_topopt* t) {
  if (!t) return ICC_OK;
  icc_filter_destroy(t->filt);
  icc_op_destroy(t->op);
  icc_grid_destroy(t->g);
  double* bufs[] = {t->x,  t->xt,   t->s,    t->u,    t->u0,   t->b,       t->du,      t->dg,
                    t->f,  t->dc,   t->tmp,  t->ones, t->dv,   t->parts,   t->xold,    t->xnew,
                    t->pmax, t->ocparts, t->sparts};
  for (double* p : bufs) hipFree(p);
  hipFree(t->fr_d);
  hipFree(t->st);
  hipFree(t->ocs);
  if (t->st_h) hipHostFree(t->st_h);
  if (t->ocs_h) hipHostFree(t->ocs_h);
  delete t;
  return ICC_OK;
}

ICC_API int icc_topopt_create(icc_ctx* ctx, const icc_topopt_config* cfg, icc_topopt** out) {
  if (!ctx || !cfg || !out) return ICC_EINVAL;
  if (cfg->solver < 0 || cfg->solver > 2) {
    set_error("icc_topopt_create: solver must be 0, 1 or 2");
    return ICC_EINVAL;
  }
  auto* t = new icc_topopt();
  t->ctx = ctx;
  t->cfg = *cfg;
  const double bounds[6] = {0.0, cfg->size[0], 0.0, cfg->size[1], 0.0, cfg->size[2]};
  int rc = icc_grid_create(ctx, cfg->nel, bounds, 2, cfg->young, cfg->poisson, &t->g);
  if (rc == ICC_OK) rc = icc_op_create(t->g, &t->op);
  if (rc == ICC_OK) rc = icc_filter_create(t->g, cfg->rmin, &t->filt);
  if (rc != ICC_OK) {
    icc_topopt_destroy(t);
    return rc;
  }
  const GridSetup& s = t->g->s;
  t->nel = s.nel_tot;
  t->ndof = s.ndof;
  // clamp x = 0 and load (topopt_iga.rs:79-98)
  t->fr.assign(s.ndof, 1);
  t->force.assign(s.ndof, 0.0);
  std::vector<int64_t> load;
  for (int64_t i = 0; i < s.ndof_sc; ++i) {
    int64_t s0 = i % s.ncp[0], s1 = (i / s.ncp[0]) % s.ncp[1], s2 = i / (s.ncp[0] * s.ncp[1]);
    if (s0 == 0)
      for (int c = 0; c < 3; ++c) t->fr[c * s.ndof_sc + i] = 0;
    int64_t i1 = s0 + 1;
    double i2 = (double)s1 + 1.0, i3 = (double)s2 + 1.0;
    if (i1 == s.ncp[0] && std::fabs(i2 - 1.0) <= 1.0 && std::fabs(i3 - (double)s.ncp[2] / 2.0) <= 1.0) load.push_back(i);
  }
  if (cfg->load == 0) {
    for (int64_t i : load) t->force[s.ndof_sc + i] = -1.0 / (double)load.size();
  } else {
    const double tr[3] = {0.0, -1.0, 0.0};
    traction_xmax(s, tr, t->force);
  }
  const int64_t n = s.ndof, ne = s.nel_tot;
  if ((rc = dalloc(&t->x, ne)) || (rc = dalloc(&t->xt, ne)) || (rc = dalloc(&t->s, ne)) || (rc = dalloc(&t->u, n)) ||
      (rc = dalloc(&t->u0, n)) || (rc = dalloc(&t->b, n)) || (rc = dalloc(&t->du, n)) || (rc = dalloc(&t->dg, n)) ||
      (rc = dalloc(&t->f, n)) || (rc = dalloc(&t->dc, ne)) || (rc = dalloc(&t->tmp, n)) ||
      (rc = dalloc(&t->ones, ne)) || (rc = dalloc(&t->dv, ne)) || (rc = dalloc(&t->parts, 2 * VEC_MAX_BLOCKS)) ||
      (rc = dalloc(&t->fr_d, n)) || (rc = dalloc(&t->st, 1)) || (rc = dalloc(&t->xold, ne)) ||
      (rc = dalloc(&t->xnew, ne)) || (rc = dalloc(&t->pmax, VEC_MAX_BLOCKS)) ||
      (rc = dalloc(&t->ocparts, 2 * VEC_MAX_BLOCKS)) || (rc = dalloc(&t->sparts, 3 * VEC_MAX_BLOCKS)) ||
      (rc = dalloc(&t->ocs, 2))) {
    icc_topopt_destroy(t);
    return rc;
  }
  if (hipHostMalloc((void**)&t->st_h, sizeof(PcgState), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&t->ocs_h, sizeof(OcState), hipHostMallocDefault) != hipSuccess) {
    icc_topopt_destroy(t);
    return ICC_ENOMEM;
  }
  hipStream_t st = ctx->stream;
  PcgState z;
  std::memset(&z, 0, sizeof(z));
  hipError_t e;
  if ((e = hipMemcpyAsync(t->st, &z, sizeof(z), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(t->f, t->force.data(), n * 8, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(t->fr_d, t->fr.data(), n, hipMemcpyHostToDevice, st)) ||
      (e = launch_fill<double>(t->ones, 1.0, ne, st)) || (e = hipStreamSynchronize(st))) {
    icc_topopt_destroy(t);
    return hip_fail(e, "icc_topopt_create");
  }
  rc = icc_op_set_free_mask(t->op, t->fr.data());
  if (rc == ICC_OK && cfg->filter == 1) rc = icc_filter_bwd_dev(t->filt, t->ones, t->dv);  // dv = H^T (1 / Hs)
  if (rc != ICC_OK) {
    icc_topopt_destroy(t);
    return rc;
  }
  *out = t;
  return ICC_OK;
}

ICC_API int icc_topopt_sizes(const icc_topopt* t, int64_t out[2]) {
  if (!t || !out) return ICC_EINVAL;
  out[0] = t->nel;
  out[1] = t->ndof;
  return ICC_OK;
}

ICC_API int icc_topopt_force(const icc_topopt* t, double* force, uint8_t* fm) {
  if (!t) return ICC_EINVAL;
  if (force) std::memcpy(force, t->force.data(), t->ndof * 8);
  if (fm) std::memcpy(fm, t->fr.data(), t->ndof);
  return ICC_OK;
}

ICC_API int icc_topopt_solve_state(icc_topopt* t, const double* x, const double* u0, double* u, int64_t* iters,
                                   double* res) {
  if (!t || !x || !u0 || !u || !iters || !res) return ICC_EINVAL;
  hipStream_t st = t->ctx->stream;
  ICC_HIP(hipMemcpyAsync(t->x, x, t->nel * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(t->u0, u0, t->ndof * 8, hipMemcpyHostToDevice, st));
  TRY(solve_state_dev(t, t->x, t->u0, t->u, iters, res));
  ICC_HIP(hipMemcpyAsync(u, t->u, t->ndof * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  return ICC_OK;
}

ICC_API int icc_topopt_element_energies(icc_topopt* t, const double* u, double* ee) {
  if (!t || !u || !ee) return ICC_EINVAL;
  return icc_energies(t->op, u, ee);
}

ICC_API int icc_topopt_compliance(icc_topopt* t, const double* x, double* u_inout, double* c, double* dcdx) {
  if (!t || !x || !u_inout || !c) return ICC_EINVAL;
  hipStream_t st = t->ctx->stream;
  const icc_topopt_config& cf = t->cfg;
  TRY(physical_densities(t, x));
  ICC_HIP(hipMemcpyAsync(t->u0, u_inout, t->ndof * 8, hipMemcpyHostToDevice, st));
  int64_t its;
  double res;
  TRY(solve_state_dev(t, t->xt, t->u0, t->u, &its, &res));
  std::vector<double> uh(t->ndof);
  ICC_HIP(hipMemcpyAsync(uh.data(), t->u, t->ndof * 8, hipMemcpyDeviceToHost, st));
  if (dcdx) {
    // dc/dx~ = -penal (1 - emin) x~^(penal-1) u_e^T K_e u_e, then the filter adjoint
    TRY(icc_sens_dev(t->op, t->u, t->xt, cf.penal, cf.emin, nullptr, t->dc));
    if (cf.filter == 1) {
      TRY(icc_filter_bwd_dev(t->filt, t->dc, t->tmp));
      ICC_HIP(hipMemcpyAsync(dcdx, t->tmp, t->nel * 8, hipMemcpyDeviceToHost, st));
    } else {
      ICC_HIP(hipMemcpyAsync(dcdx, t->dc, t->nel * 8, hipMemcpyDeviceToHost, st));
    }
  }
  ICC_HIP(hipStreamSynchronize(st));
  std::memcpy(u_inout, uh.data(), t->ndof * 8);
  *c = dot_host(t->force, uh);
  return ICC_OK;
}

ICC_API int icc_topopt_run(icc_topopt* t, icc_report_fn report, void* user, double* compliance, double* change,
                           int64_t* pcg_iterations, double* density, int64_t* n_iter) {
  if (!t || !density) return ICC_EINVAL;
  const icc_topopt_config& cf = t->cfg;
  hipStream_t st = t->ctx->stream;
  const int64_t nel = t->nel, n = t->ndof;
  const int G = topopt_blocks(nel);
  std::vector<double> sp(3 * (size_t)G);
  ICC_HIP(launch_fill<double>(t->x, cf.volfrac, nel, st));
  ICC_HIP(launch_fill<double>(t->u, 0.0, n, st));
  int64_t it_done = 0;
  for (int64_t iter = 1; iter <= cf.max_iter; ++iter) {
    ICC_HIP(hipMemcpyAsync(t->xold, t->x, nel * 8, hipMemcpyDeviceToDevice, st));
    TRY(physical_densities_dev(t));
    ICC_HIP(hipMemcpyAsync(t->u0, t->u, n * 8, hipMemcpyDeviceToDevice, st));
    int64_t its;
    double res;
    TRY(solve_state_dev(t, t->xt, t->u0, t->u, &its, &res));
    // dc_raw (topopt_iga.rs:199), fused with the element energies
    TRY(icc_sens_dev(t->op, t->u, t->xt, cf.penal, cf.emin, nullptr, t->dc));
    if (cf.filter == 1)
      TRY(icc_filter_bwd_dev(t->filt, t->dc, t->tmp));
    else
      TRY(icc_filter_sens_dev(t->filt, t->x, t->dc, t->tmp));  // topopt_iga.rs:200-205
    TRY(oc_update_dev(t, t->tmp));
    std::swap(t->x, t->xnew);
    // c = F.u (topopt_iga.rs:197), change = max |x - x_old|, volume = mean x
    ICC_HIP(launch_topopt_stats(n, t->f, t->u, nel, nullptr, nullptr, t->x, t->xold, t->sparts, st));
    ICC_HIP(hipMemcpyAsync(sp.data(), t->sparts, sp.size() * 8, hipMemcpyDeviceToHost, st));
    ICC_HIP(hipStreamSynchronize(st));
    double c = 0.0, ch = 0.0, vol = 0.0;
    for (int i = 0; i < G; ++i) {
      c += sp[i];
      ch = std::fmax(ch, sp[G + i]);
      vol += sp[2 * G + i];
    }
    vol /= (double)nel;
    if (compliance) compliance[iter - 1] = std::fabs(c);
    if (change) change[iter - 1] = ch;
    if (pcg_iterations) pcg_iterations[iter - 1] = its;
    it_done = iter;
    if (report) report(iter, std::fabs(c), vol, ch, user);
    if (ch < 0.01 && iter >= 15) break;
  }
  ICC_HIP(hipMemcpyAsync(density, t->x, nel * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  if (n_iter) *n_iter = it_done;
  return ICC_OK;
}
