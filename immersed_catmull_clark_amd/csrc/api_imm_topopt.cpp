// Immersed geometry from a triangle B-rep and the immersed SIMP topology
// optimisation driver above the C ABI: the host side of ImmersedTopOpt
// (rust/src/topopt_immersed.rs:78-247).
//
// Set-up: padded bounds, ray-parity classification and volume fractions on
// the GPU, design domain w >= 0.05, the face stabilisation (consistent
// ghost penalty or the legacy springs) as a matrix-free operator term, bottom clamp, top-centre load, and the
// design-domain filter table.  Every design iteration then runs on the
// device: SIMP scale, warm-started Jacobi PCG (FP64, or FP32 with FP64
// reliable updates) with the reference's rescaled tolerance, element
// energies with the w-weighted sensitivity, the filter, the OC bisection
// and the history scalars; the host reads a handful of scalars per
// iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

#define ICC_API extern "C" __attribute__((visibility("default")))

using namespace icc;


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

#define TRY(call)                    \
  do {                               \
    int rc__ = (call);               \
    if (rc__ != ICC_OK) return rc__; \
  } while (0)

bool mesh_ok(const double* v, int64_t nv, const int64_t* t, int64_t nt) {
  if (!v || !t || nv < 1 || nt < 1) return false;
  for (int64_t i = 0; i < 3 * nt; ++i)
    if (t[i] < 0 || t[i] >= nv) return false;
  return true;
}

// Triangle records of the ray-parity predicate (immersed.rs:35-46).
std::vector<double> triangle_records(const double* v, const int64_t* t, int64_t nt) {
  std::vector<double> r((size_t)nt * MESH_TRI_STRIDE);
  for (int64_t q = 0; q < nt; ++q) {
    double* R = r.data() + q * MESH_TRI_STRIDE;
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) R[3 * c + d] = v[3 * t[3 * q + c] + d];
    R[9] = std::fmin(std::fmin(R[1], R[4]), R[7]);
    R[10] = std::fmax(std::fmax(R[1], R[4]), R[7]);
    R[11] = std::fmin(std::fmin(R[2], R[5]), R[8]);
    R[12] = std::fmax(std::fmax(R[2], R[5]), R[8]);
  }
  return r;
}

}  // namespace

// padded_bounds (rust/src/immersed.rs:450-463)
ICC_API int icc_mesh_padded_bounds(const double* vertices, int64_t nv, double padding, double out[6]) {
  if (!vertices || nv < 1 || !out) return ICC_EINVAL;
  for (int d = 0; d < 3; ++d) {
    double lo = std::numeric_limits<double>::max(), hi = std::numeric_limits<double>::lowest();
    for (int64_t i = 0; i < nv; ++i) {
      lo = std::fmin(lo, vertices[3 * i + d]);
      hi = std::fmax(hi, vertices[3 * i + d]);
    }
    const double pad = padding * (hi - lo);
    out[2 * d] = lo - pad;
    out[2 * d + 1] = hi + pad;
  }
  return ICC_OK;
}

ICC_API int icc_classify_mesh(icc_ctx* ctx, const double* vertices, int64_t nv, const int64_t* tris, int64_t ntri,
                              const double grid_bounds[6], const int64_t res[3], const int64_t sub[3], int8_t* status,
                              double* weights) {
  if (!ctx || !grid_bounds || !res || !sub || !status || !weights) return ICC_EINVAL;
  if (!mesh_ok(vertices, nv, tris, ntri)) {
    set_error("icc_classify_mesh: empty mesh or triangle index out of range");
    return ICC_EINVAL;
  }
  for (int d = 0; d < 3; ++d)
    if (res[d] < 1 || sub[d] < 1 || !(grid_bounds[2 * d + 1] > grid_bounds[2 * d])) {
      set_error("icc_classify_mesh: bad resolution or bounds");
      return ICC_EINVAL;
    }
  const std::vector<double> rec = triangle_records(vertices, tris, ntri);
  hipStream_t st = ctx->stream;
  const int64_t nc = res[0] * res[1] * res[2], nvx = (res[0] + 1) * (res[1] + 1) * (res[2] + 1);
  double *d_tri = nullptr, *d_w = nullptr;
  uint8_t* d_in = nullptr;
  int8_t* d_st = nullptr;
  int rc = ICC_OK;
  if ((rc = dalloc(&d_tri, rec.size())) || (rc = dalloc(&d_w, nc)) || (rc = dalloc(&d_in, nvx)) ||
      (rc = dalloc(&d_st, nc))) {
    hipFree(d_tri);
    hipFree(d_w);
    hipFree(d_in);
    hipFree(d_st);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(d_tri, rec.data(), rec.size() * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_classify_mesh(d_tri, ntri, grid_bounds, res, sub, d_in, d_st, d_w, st);
  if (e == hipSuccess) e = hipMemcpyAsync(status, d_st, nc, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(weights, d_w, nc * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  hipFree(d_tri);
  hipFree(d_w);
  hipFree(d_in);
  hipFree(d_st);
  if (e != hipSuccess) return hip_fail(e, "icc_classify_mesh");
  return ICC_OK;
}

// ---------------------------------------------------------------------------
struct icc_imm_topopt {
  icc_ctx* ctx = nullptr;
  icc_imm_topopt_config cfg;
  icc_grid* g = nullptr;
  icc_op* op = nullptr;
  double gb[6];
  int64_t nel = 0, ndof = 0, n_active = 0, n_faces = 0, n_free = 0;
  double cad_volume = 0.0, nf = 0.0;
  std::vector<int8_t> status;
  std::vector<double> w, force;
  std::vector<uint8_t> active, fr;
  int noff = 0;
  // device
  double *x = nullptr, *xold = nullptr, *xnew = nullptr, *wd = nullptr, *s = nullptr, *u = nullptr, *b = nullptr,
         *du = nullptr, *dg = nullptr, *f = nullptr, *dcraw = nullptr, *dc = nullptr, *tmp = nullptr, *hs = nullptr,
         *ow = nullptr, *parts = nullptr, *pmax = nullptr, *sparts = nullptr;
  int32_t* off = nullptr;
  uint8_t *act_d = nullptr, *fr_d = nullptr;
  PcgState* pst = nullptr;
  OcState* ocs = nullptr;
  PcgState* pst_h = nullptr;
  OcState* ocs_h = nullptr;
};

ICC_API void icc_imm_topopt_config_default(icc_imm_topopt_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->grid_res[0] = 28;
  c->grid_res[1] = 20;
  c->grid_res[2] = 16;
  c->volfrac = 0.35;
  c->penal = 3.0;
  c->rmin = 1.8;
  c->max_iter = 30;
  c->young = 1.0;
  c->poisson = 0.3;
  c->emin = 1e-4;
  c->stabilization = ICC_STAB_GHOST;
  c->gamma = 1e-2;
  c->pcg_tol = 1e-8;
  c->pcg_maxit = 20000;
  c->solver = 0;
}

ICC_API int icc_imm_topopt_destroy(icc_imm_topopt* t) {
  if (!t) return ICC_OK;
  icc_op_destroy(t->op);
  icc_grid_destroy(t->g);
  double* bufs[] = {t->x,  t->xold, t->xnew, t->wd, t->s,  t->u,  t->b,     t->du,    t->dg,
                    t->f,  t->dcraw, t->dc,  t->tmp, t->hs, t->ow, t->parts, t->pmax, t->sparts};
  for (double* p : bufs) hipFree(p);
  hipFree(t->off);
  hipFree(t->act_d);
  hipFree(t->fr_d);
  hipFree(t->pst);
  hipFree(t->ocs);
  if (t->pst_h) hipHostFree(t->pst_h);
  if (t->ocs_h) hipHostFree(t->ocs_h);
  delete t;
  return ICC_OK;
}

ICC_API int icc_imm_topopt_create(icc_ctx* ctx, const icc_imm_topopt_config* cfg, const double* vertices, int64_t nv,
                                  const int64_t* tris, int64_t ntri, icc_imm_topopt** out) {
  if (!ctx || !cfg || !out) return ICC_EINVAL;
  if (!mesh_ok(vertices, nv, tris, ntri)) {
    set_error("icc_imm_topopt_create: empty mesh or triangle index out of range");
    return ICC_EINVAL;
  }
  for (int d = 0; d < 3; ++d)
    if (cfg->grid_res[d] < 1) {
      set_error("icc_imm_topopt_create: grid_res must be >= 1");
      return ICC_EINVAL;
    }
  if (cfg->stabilization != ICC_STAB_NONE && cfg->stabilization != ICC_STAB_LEGACY &&
      cfg->stabilization != ICC_STAB_GHOST) {
    set_error("icc_imm_topopt_create: unknown stabilization");
    return ICC_EINVAL;
  }
  if (cfg->solver < 0 || cfg->solver > 2) return ICC_EINVAL;
  auto* t = new icc_imm_topopt();
  t->ctx = ctx;
  t->cfg = *cfg;
  int rc = icc_mesh_padded_bounds(vertices, nv, 0.05, t->gb);
  if (rc == ICC_OK) rc = icc_grid_create(ctx, cfg->grid_res, t->gb, 2, cfg->young, cfg->poisson, &t->g);
  if (rc == ICC_OK) rc = icc_op_create(t->g, &t->op);
  if (rc != ICC_OK) {
    icc_imm_topopt_destroy(t);
    return rc;
  }
  const GridSetup& s = t->g->s;
  const int64_t ne = s.nel_tot, n = s.ndof;
  t->nel = ne;
  t->ndof = n;
  t->status.resize(ne);
  t->w.resize(ne);
  const int64_t sub[3] = {4, 4, 4};
  rc = icc_classify_mesh(ctx, vertices, nv, tris, ntri, t->gb, cfg->grid_res, sub, t->status.data(), t->w.data());
  if (rc != ICC_OK) {
    icc_imm_topopt_destroy(t);
    return rc;
  }
  // design domain and CAD volume (topopt_immersed.rs:84-85)
  t->active.assign(ne, 0);
  for (int64_t e = 0; e < ne; ++e) {
    t->active[e] = t->w[e] >= 0.05 ? 1 : 0;
    if (t->active[e]) {
      t->cad_volume += t->w[e];
      t->n_active++;
    }
  }
  if (cfg->stabilization != ICC_STAB_NONE) {
    rc = cfg->stabilization == ICC_STAB_LEGACY
             ? icc_op_set_legacy_stabilization(t->op, t->status.data(), cfg->gamma, &t->n_faces)
             : icc_op_set_ghost_penalty(t->op, t->status.data(), cfg->gamma, cfg->young, &t->n_faces);
    if (rc != ICC_OK) {
      icc_imm_topopt_destroy(t);
      return rc;
    }
  }
  // bottom clamp, load on the top face centre (topopt_immersed.rs:95-114)
  t->fr.assign(n, 1);
  t->force.assign(n, 0.0);
  const int64_t nd0 = s.ncp[0], nd1 = s.ncp[1], nd2 = s.ncp[2];
  const int64_t mid_x = (int64_t)std::round((double)nd0 / 2.0), mid_y = (int64_t)std::round((double)nd1 / 2.0);
  std::vector<int64_t> load;
  for (int64_t i = 0; i < s.ndof_sc; ++i) {
    const int64_t s0 = i % nd0, s1 = (i / nd0) % nd1, s2 = i / (nd0 * nd1);
    if (s2 == 0)
      for (int c = 0; c < 3; ++c) t->fr[c * s.ndof_sc + i] = 0;
    if (s2 == nd2 - 1 && std::llabs(s0 + 1 - mid_x) <= 2 && std::llabs(s1 + 1 - mid_y) <= 2) load.push_back(i);
  }
  for (int64_t i : load) t->force[2 * s.ndof_sc + i] = -1.0 / (double)load.size();
  for (int64_t i = 0; i < n; ++i) {
    t->n_free += t->fr[i];
    if (t->fr[i]) t->nf += t->force[i] * t->force[i];
  }
  t->nf = std::sqrt(t->nf);
  // filter offsets on the design domain (topopt_immersed.rs:117-131)
  std::vector<int32_t> offs;
  std::vector<double> ows;
  design_filter_offsets(s, cfg->rmin, offs, ows);
  t->noff = (int)ows.size();
  const int G = topopt_blocks(ne);
  if ((rc = dalloc(&t->x, ne)) || (rc = dalloc(&t->xold, ne)) || (rc = dalloc(&t->xnew, ne)) ||
      (rc = dalloc(&t->wd, ne)) || (rc = dalloc(&t->s, ne)) || (rc = dalloc(&t->u, n)) || (rc = dalloc(&t->b, n)) ||
      (rc = dalloc(&t->du, n)) || (rc = dalloc(&t->dg, n)) || (rc = dalloc(&t->f, n)) || (rc = dalloc(&t->dcraw, ne)) ||
      (rc = dalloc(&t->dc, ne)) || (rc = dalloc(&t->tmp, n)) || (rc = dalloc(&t->hs, ne)) ||
      (rc = dalloc(&t->ow, ows.size())) || (rc = dalloc(&t->off, offs.size())) ||
      (rc = dalloc(&t->parts, 2 * (size_t)VEC_MAX_BLOCKS)) || (rc = dalloc(&t->pmax, G)) ||
      (rc = dalloc(&t->sparts, 3 * (size_t)G)) || (rc = dalloc(&t->act_d, ne)) || (rc = dalloc(&t->fr_d, n)) ||
      (rc = dalloc(&t->pst, 1)) || (rc = dalloc(&t->ocs, 2))) {
    icc_imm_topopt_destroy(t);
    return rc;
  }
  if (hipHostMalloc((void**)&t->pst_h, sizeof(PcgState), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&t->ocs_h, sizeof(OcState), hipHostMallocDefault) != hipSuccess) {
    icc_imm_topopt_destroy(t);
    set_error("hipHostMalloc failed");
    return ICC_ENOMEM;
  }
  hipStream_t st = ctx->stream;
  PcgState z;
  std::memset(&z, 0, sizeof(z));
  hipError_t e;
  if ((e = hipMemcpyAsync(t->pst, &z, sizeof(z), hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(t->wd, t->w.data(), ne * 8, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(t->act_d, t->active.data(), ne, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(t->f, t->force.data(), n * 8, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(t->fr_d, t->fr.data(), n, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(t->ow, ows.data(), ows.size() * 8, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(t->off, offs.data(), offs.size() * 4, hipMemcpyHostToDevice, st)) ||
      (e = launch_imm_filter(1, s.nel, t->off, t->ow, t->noff, t->act_d, nullptr, nullptr, nullptr, t->hs, st)) ||
      (e = hipStreamSynchronize(st))) {
    icc_imm_topopt_destroy(t);
    return hip_fail(e, "icc_imm_topopt_create");
  }
  rc = icc_op_set_free_mask(t->op, t->fr.data());
  if (rc != ICC_OK) {
    icc_imm_topopt_destroy(t);
    return rc;
  }
  *out = t;
  return ICC_OK;
}

ICC_API int icc_imm_topopt_sizes(const icc_imm_topopt* t, int64_t out[5], double* cad_volume, double grid_bounds[6]) {
  if (!t) return ICC_EINVAL;
  if (out) {
    out[0] = t->nel;
    out[1] = t->ndof;
    out[2] = t->n_active;
    out[3] = t->n_faces;
    out[4] = t->n_free;
  }
  if (cad_volume) *cad_volume = t->cad_volume;
  if (grid_bounds) std::memcpy(grid_bounds, t->gb, sizeof(t->gb));
  return ICC_OK;
}

ICC_API int icc_imm_topopt_problem(const icc_imm_topopt* t, int8_t* status, double* weights, uint8_t* active,
                                   double* force, uint8_t* free_mask) {
  if (!t) return ICC_EINVAL;
  if (status) std::memcpy(status, t->status.data(), t->nel);
  if (weights) std::memcpy(weights, t->w.data(), t->nel * 8);
  if (active) std::memcpy(active, t->active.data(), t->nel);
  if (force) std::memcpy(force, t->force.data(), t->ndof * 8);
  if (free_mask) std::memcpy(free_mask, t->fr.data(), t->ndof);
  return ICC_OK;
}

namespace {

// One design iteration (topopt_immersed.rs:163-243) on the device.
int imm_iteration(icc_imm_topopt* t, double* c_out, double* vol_out, double* ch_out, int64_t* its_out) {
  const icc_imm_topopt_config& cf = t->cfg;
  hipStream_t st = t->ctx->stream;
  const int64_t n = t->ndof, ne = t->nel;
  const GridSetup& s = t->g->s;
  ICC_HIP(hipMemcpyAsync(t->xold, t->x, ne * 8, hipMemcpyDeviceToDevice, st));
  // scale = active ? w (emin + (1 - emin) x^p) : emin
  ICC_HIP(launch_simp_scale(ne, t->x, cf.penal, cf.emin, t->wd, t->act_d, t->s, st));
  TRY(icc_op_set_scale_dev(t->op, t->s));
  // b = free ? F - (K + S) u : 0 and |b|
  TRY(icc_apply_dev(t->op, ICC_F64, t->u, t->tmp));
  ICC_HIP(launch_residual(n, t->f, t->tmp, t->fr_d, t->b, t->pst, t->parts, st));
  ICC_HIP(hipMemcpyAsync(t->pst_h, t->pst, sizeof(PcgState), hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  const double rb = std::sqrt(t->pst_h->rr);
  // diag = max(diag K + diag S, 1e-6)
  TRY(icc_diag_dev(t->op, t->tmp));
  ICC_HIP(launch_clamp_min(n, t->tmp, 1e-6, t->dg, st));
  const double tol = rb > 0.0 ? std::fmin(cf.pcg_tol * t->nf / rb, 0.5) : 1.0;
  int64_t its = 0;
  if (cf.solver == 0) {
    double res = 0.0;
    int rc = icc_pcg_dev(t->op, t->b, t->dg, tol, cf.pcg_maxit, t->du, &its, &res);
    if (rc != ICC_OK && rc != ICC_EBREAKDOWN && rc != ICC_ENOCONV) return rc;
  } else {
    ICC_HIP(launch_fill<double>(t->du, 0.0, n, st));
    double stats[8];
    int rc = icc_solve_mixed_prec_dev(t->op, cf.solver == 2 ? ICC_F16X : ICC_F32, t->b, tol, 0.1, cf.pcg_maxit, t->du,
                                      stats);
    if (rc != ICC_OK && rc != ICC_ENOCONV && rc != ICC_EBREAKDOWN) return rc;
    its = (int64_t)stats[1];
  }
  ICC_HIP(launch_masked_add(n, t->u, t->du, t->fr_d, t->tmp, st));
  std::swap(t->u, t->tmp);
  // dc_raw = -w p (1 - emin) x^(p-1) ee, then the design-domain filter
  TRY(icc_sens_dev(t->op, t->u, t->x, cf.penal, cf.emin, t->wd, t->dcraw));
  ICC_HIP(launch_imm_filter(0, s.nel, t->off, t->ow, t->noff, t->act_d, t->x, t->dcraw, t->hs, t->dc, st));
  // OC bisection with the CAD-weighted volume target
  const double target = cf.volfrac * t->cad_volume;
  ICC_HIP(launch_oc_init(ne, t->act_d, t->dc, t->xold, t->xnew, t->pmax, st));
  constexpr int CHUNK = 8, MAX_STEPS = 4096;
  int k = 0;
  for (;;) {
    for (int q = 0; q < CHUNK; ++q, ++k)
      ICC_HIP(launch_oc_step(k, ne, t->act_d, t->wd, nullptr, t->dc, t->xold, t->xnew, target, 0.2, t->pmax,
                             t->parts, t->ocs, st));
    ICC_HIP(hipMemcpyAsync(t->ocs_h, t->ocs + (k & 1), sizeof(OcState), hipMemcpyDeviceToHost, st));
    ICC_HIP(hipStreamSynchronize(st));
    if (t->ocs_h->done) break;
    if (k >= MAX_STEPS) {
      set_error("icc_imm_topopt_run: OC bisection did not terminate");
      return ICC_ENOCONV;
    }
  }
  std::swap(t->x, t->xnew);
  // compliance F.u, change and volume
  const int G = topopt_blocks(ne);
  ICC_HIP(launch_topopt_stats(n, t->f, t->u, ne, t->act_d, t->wd, t->x, t->xold, t->sparts, st));
  std::vector<double> sp(3 * (size_t)G);
  ICC_HIP(hipMemcpyAsync(sp.data(), t->sparts, sp.size() * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  double c = 0.0, ch = 0.0, vol = 0.0;
  for (int i = 0; i < G; ++i) {
    c += sp[i];
    ch = std::fmax(ch, sp[G + i]);
    vol += sp[2 * G + i];
  }
  *c_out = c;
  *vol_out = vol / t->cad_volume;
  *ch_out = ch;
  *its_out = its;
  return ICC_OK;
}

}  // namespace

ICC_API int icc_imm_topopt_run(icc_imm_topopt* t, icc_imm_report_fn report, void* user, double* compliance,
                               double* volume, double* change, int64_t* pcg_iterations, double* density) {
  if (!t) return ICC_EINVAL;
  const icc_imm_topopt_config& cf = t->cfg;
  hipStream_t st = t->ctx->stream;
  if (t->n_active == 0) {
    set_error("icc_imm_topopt_run: empty design domain");
    return ICC_EINVAL;
  }
  ICC_HIP(launch_fill_active(t->nel, t->act_d, cf.volfrac, t->x, st));
  ICC_HIP(launch_fill<double>(t->u, 0.0, t->ndof, st));
  for (int64_t iter = 1; iter <= cf.max_iter; ++iter) {
    double c, vol, ch;
    int64_t its;
    TRY(imm_iteration(t, &c, &vol, &ch, &its));
    if (compliance) compliance[iter - 1] = c;
    if (volume) volume[iter - 1] = vol;
    if (change) change[iter - 1] = ch;
    if (pcg_iterations) pcg_iterations[iter - 1] = its;
    if (report) report(iter, c, vol, ch, its, user);
  }
  if (density) {
    ICC_HIP(hipMemcpyAsync(density, t->x, t->nel * 8, hipMemcpyDeviceToHost, st));
    ICC_HIP(hipStreamSynchronize(st));
  }
  return ICC_OK;
}

ICC_API int icc_imm_topopt_state(icc_imm_topopt* t, double* u) {
  if (!t || !u) return ICC_EINVAL;
  ICC_HIP(hipMemcpyAsync(u, t->u, t->ndof * 8, hipMemcpyDeviceToHost, t->ctx->stream));
  ICC_HIP(hipStreamSynchronize(t->ctx->stream));
  return ICC_OK;
}

// ---------------------------------------------------------------------------
// build_immersed_problem / solve_immersed_problem for a triangle B-rep
// (rust/src/immersed.rs:392-570): grid on padded_bounds(mesh, padding), ray-
// parity classification with subcell_res^3 samples, stiffness scale
// emin + (1 - emin) w on every cell, the selected stabilisation, a clamped
// background face, the caller's force or the default top-centre load, and
// Jacobi PCG with the diagonal clamped at 1e-6 (FP64), or the mixed solver.
struct icc_immersed_mesh {
  icc_ctx* ctx = nullptr;
  icc_immersed_mesh_config cfg;
  icc_grid* g = nullptr;
  icc_op* op = nullptr;
  double gb[6];
  int64_t nel = 0, ndof = 0, n_faces = 0, n_free = 0;
  std::vector<int8_t> status;
  std::vector<double> w, force;
  std::vector<uint8_t> fr;
  double *d_f = nullptr, *d_u = nullptr, *d_diag = nullptr, *d_tmp = nullptr;
};

ICC_API void icc_immersed_mesh_config_default(icc_immersed_mesh_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->grid_res[0] = 24;
  c->grid_res[1] = 16;
  c->grid_res[2] = 12;
  c->padding = 0.05;
  c->young = 1.0;
  c->poisson = 0.3;
  c->stabilization = ICC_STAB_GHOST;
  c->gamma = 1e-2;
  c->emin = 1e-4;
  c->subcell_res[0] = c->subcell_res[1] = c->subcell_res[2] = 4;
  c->clamped_face = ICC_CLAMP_BOTTOM;
  c->pcg_tol = 1e-8;
  c->pcg_maxit = 5000;
  c->solver = 0;
}

ICC_API int icc_immersed_mesh_destroy(icc_immersed_mesh* h) {
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

ICC_API int icc_immersed_mesh_create(icc_ctx* ctx, const icc_immersed_mesh_config* cfg, const double* vertices,
                                     int64_t nv, const int64_t* tris, int64_t ntri, const double* force,
                                     icc_immersed_mesh** out) {
  if (!ctx || !cfg || !out) return ICC_EINVAL;
  if (!mesh_ok(vertices, nv, tris, ntri)) {
    set_error("icc_immersed_mesh_create: empty mesh or triangle index out of range");
    return ICC_EINVAL;
  }
  for (int d = 0; d < 3; ++d)
    if (cfg->grid_res[d] < 1 || cfg->subcell_res[d] < 1) {
      set_error("icc_immersed_mesh_create: grid_res and subcell_res must be >= 1");
      return ICC_EINVAL;
    }
  if ((cfg->stabilization != ICC_STAB_NONE && cfg->stabilization != ICC_STAB_LEGACY &&
       cfg->stabilization != ICC_STAB_GHOST) ||
      cfg->clamped_face < ICC_CLAMP_NONE || cfg->clamped_face > ICC_CLAMP_LEFT || cfg->solver < 0 || cfg->solver > 2) {
    set_error("icc_immersed_mesh_create: unknown stabilization, clamped face or solver");
    return ICC_EINVAL;
  }
  auto* h = new icc_immersed_mesh();
  h->ctx = ctx;
  h->cfg = *cfg;
  int rc = icc_mesh_padded_bounds(vertices, nv, cfg->padding, h->gb);
  if (rc == ICC_OK) rc = icc_grid_create(ctx, cfg->grid_res, h->gb, 2, cfg->young, cfg->poisson, &h->g);
  if (rc == ICC_OK) rc = icc_op_create(h->g, &h->op);
  if (rc != ICC_OK) {
    icc_immersed_mesh_destroy(h);
    return rc;
  }
  const GridSetup& s = h->g->s;
  const int64_t ne = s.nel_tot, n = s.ndof;
  h->nel = ne;
  h->ndof = n;
  h->status.resize(ne);
  h->w.resize(ne);
  rc = icc_classify_mesh(ctx, vertices, nv, tris, ntri, h->gb, cfg->grid_res, cfg->subcell_res, h->status.data(),
                         h->w.data());
  // scale emin + (1 - emin) w on every cell (immersed.rs:484)
  std::vector<double> scale(ne);
  for (int64_t e = 0; e < ne; ++e) scale[e] = cfg->emin + (1.0 - cfg->emin) * h->w[e];
  if (rc == ICC_OK) rc = icc_op_set_scale(h->op, scale.data());
  if (rc == ICC_OK && cfg->stabilization == ICC_STAB_LEGACY)
    rc = icc_op_set_legacy_stabilization(h->op, h->status.data(), cfg->gamma, &h->n_faces);
  if (rc == ICC_OK && cfg->stabilization == ICC_STAB_GHOST)
    rc = icc_op_set_ghost_penalty(h->op, h->status.data(), cfg->gamma, cfg->young, &h->n_faces);
  if (rc != ICC_OK) {
    icc_immersed_mesh_destroy(h);
    return rc;
  }
  // clamped face (immersed.rs:512-519) and the load (:520-533, or the caller's)
  const int64_t nd0 = s.ncp[0], nd1 = s.ncp[1], nd2 = s.ncp[2];
  h->fr.assign(n, 1);
  for (int64_t i = 0; i < s.ndof_sc; ++i) {
    const int64_t i1 = i % nd0, i3 = i / (nd0 * nd1);
    const bool cl = (cfg->clamped_face == ICC_CLAMP_BOTTOM && i3 == 0) ||
                    (cfg->clamped_face == ICC_CLAMP_TOP && i3 == nd2 - 1) ||
                    (cfg->clamped_face == ICC_CLAMP_LEFT && i1 == 0);
    if (cl)
      for (int c = 0; c < 3; ++c) h->fr[c * s.ndof_sc + i] = 0;
  }
  if (force) {
    h->force.assign(force, force + n);
  } else {
    h->force.assign(n, 0.0);
    const int64_t mid_x = (int64_t)std::round((double)nd0 / 2.0), mid_y = (int64_t)std::round((double)nd1 / 2.0);
    std::vector<int64_t> load;
    for (int64_t i = 0; i < s.ndof_sc; ++i) {
      const int64_t i1 = i % nd0, i2 = (i / nd0) % nd1, i3 = i / (nd0 * nd1);
      if (i3 == nd2 - 1 && std::llabs(i1 + 1 - mid_x) <= 2 && std::llabs(i2 + 1 - mid_y) <= 2) load.push_back(i);
    }
    for (int64_t i : load) h->force[2 * s.ndof_sc + i] = -1.0 / (double)load.size();
  }
  std::vector<double> b(n);
  for (int64_t i = 0; i < n; ++i) {
    h->n_free += h->fr[i];
    b[i] = h->fr[i] ? h->force[i] : 0.0;
  }
  if ((rc = icc_op_set_free_mask(h->op, h->fr.data())) || (rc = dalloc(&h->d_f, n)) || (rc = dalloc(&h->d_u, n)) ||
      (rc = dalloc(&h->d_diag, n)) || (rc = dalloc(&h->d_tmp, n))) {
    icc_immersed_mesh_destroy(h);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(h->d_f, b.data(), n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    icc_immersed_mesh_destroy(h);
    return hip_fail(e, "icc_immersed_mesh_create");
  }
  *out = h;
  return ICC_OK;
}

ICC_API int icc_immersed_mesh_sizes(const icc_immersed_mesh* h, int64_t out[4], double grid_bounds[6]) {
  if (!h) return ICC_EINVAL;
  if (out) {
    out[0] = h->nel;
    out[1] = h->ndof;
    out[2] = h->n_faces;
    out[3] = h->n_free;
  }
  if (grid_bounds) std::memcpy(grid_bounds, h->gb, sizeof(h->gb));
  return ICC_OK;
}

ICC_API int icc_immersed_mesh_problem(const icc_immersed_mesh* h, int8_t* status, double* weights, double* force,
                                      uint8_t* free_mask) {
  if (!h) return ICC_EINVAL;
  if (status) std::memcpy(status, h->status.data(), h->nel);
  if (weights) std::memcpy(weights, h->w.data(), h->nel * 8);
  if (force) std::memcpy(force, h->force.data(), h->ndof * 8);
  if (free_mask) std::memcpy(free_mask, h->fr.data(), h->ndof);
  return ICC_OK;
}

ICC_API int icc_immersed_mesh_set_solver(icc_immersed_mesh* h, double tol, int64_t maxit, int solver) {
  if (!h || !(tol > 0.0) || maxit < 1 || solver < 0 || solver > 2) return ICC_EINVAL;
  h->cfg.pcg_tol = tol;
  h->cfg.pcg_maxit = maxit;
  h->cfg.solver = solver;
  return ICC_OK;
}

ICC_API int icc_immersed_mesh_apply(icc_immersed_mesh* h, const double* x, double* y) {
  if (!h || !x || !y) return ICC_EINVAL;
  return icc_apply(h->op, ICC_F64, x, y);
}

ICC_API int icc_immersed_mesh_solve(icc_immersed_mesh* h, double* u, double* stats) {
  if (!h || !u) return ICC_EINVAL;
  const int64_t n = h->ndof;
  hipStream_t st = h->ctx->stream;
  double st8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int rc;
  if (h->cfg.solver == 0) {
    TRY(icc_diag_dev(h->op, h->d_tmp));
    ICC_HIP(launch_clamp_min(n, h->d_tmp, 1e-6, h->d_diag, st));
    int64_t iters = 0;
    double res = 0.0;
    rc = icc_pcg_dev(h->op, h->d_f, h->d_diag, h->cfg.pcg_tol, h->cfg.pcg_maxit, h->d_u, &iters, &res);
    if (rc != ICC_OK && rc != ICC_ENOCONV && rc != ICC_EBREAKDOWN) return rc;
    st8[0] = (double)iters;
    st8[1] = res;
  } else {
    ICC_HIP(launch_fill<double>(h->d_u, 0.0, n, st));
    rc = icc_solve_mixed_prec_dev(h->op, h->cfg.solver == 2 ? ICC_F16X : ICC_F32, h->d_f, h->cfg.pcg_tol, 0.1,
                                  h->cfg.pcg_maxit, h->d_u, st8);
    if (rc != ICC_OK && rc != ICC_ENOCONV && rc != ICC_EBREAKDOWN) return rc;
    const double upd = st8[0], its = st8[1];
    st8[0] = its;
    st8[1] = upd;
  }
  // u = free ? x : 0 (immersed.rs:567), true residual and compliance F.u
  std::vector<double> au(n);
  ICC_HIP(hipMemcpyAsync(u, h->d_u, n * 8, hipMemcpyDeviceToHost, st));
  int rc2 = icc_apply_dev(h->op, ICC_F64, h->d_u, h->d_tmp);
  if (rc2 != ICC_OK) return rc2;
  ICC_HIP(hipMemcpyAsync(au.data(), h->d_tmp, n * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  double rr = 0.0, ff = 0.0, c = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    if (!h->fr[i]) {
      u[i] = 0.0;
      continue;
    }
    const double r = h->force[i] - au[i];
    rr += r * r;
    ff += h->force[i] * h->force[i];
  }
  for (int64_t i = 0; i < n; ++i) c += h->force[i] * u[i];
  st8[2] = ff > 0.0 ? std::sqrt(rr / ff) : std::sqrt(rr);
  st8[3] = c;
  if (stats) std::memcpy(stats, st8, sizeof(st8));
  return rc;
}

ICC_API int icc_immersed_mesh_grid(icc_immersed_mesh* h, icc_grid** out) {
  if (!h || !out) return ICC_EINVAL;
  *out = h->g;
  return ICC_OK;
}

// ImmersedProblem::add_surface_terms (rust/src/immersed.rs:541-548): the
// surface triplets join the operator's sparse term, the load joins the force
// (and the right-hand side on the free DOFs).
ICC_API int icc_immersed_mesh_add_surface(icc_immersed_mesh* h, icc_surface* s) {
  if (!h || !s) return ICC_EINVAL;
  if (surface_grid(s) != h->g) {
    set_error("icc_immersed_mesh_add_surface: the surface terms belong to another grid");
    return ICC_EINVAL;
  }
  int64_t sz[3];
  TRY(icc_surface_sizes(s, sz));
  const int64_t n = h->ndof;
  std::vector<int64_t> r(sz[0]), c(sz[0]);
  std::vector<double> v(sz[0]), f(n);
  TRY(icc_surface_terms(s, r.data(), c.data(), v.data(), f.data()));
  TRY(icc_op_add_triplets(h->op, sz[0], r.data(), c.data(), v.data()));
  std::vector<double> b(n);
  for (int64_t i = 0; i < n; ++i) {
    h->force[i] += f[i];
    b[i] = h->fr[i] ? h->force[i] : 0.0;
  }
  ICC_HIP(hipMemcpyAsync(h->d_f, b.data(), n * 8, hipMemcpyHostToDevice, h->ctx->stream));
  ICC_HIP(hipStreamSynchronize(h->ctx->stream));
  return ICC_OK;
}

ICC_API int icc_immersed_mesh_set_force(icc_immersed_mesh* h, const double* force) {
  if (!h || !force) return ICC_EINVAL;
  const int64_t n = h->ndof;
  h->force.assign(force, force + n);
  std::vector<double> b(n);
  for (int64_t i = 0; i < n; ++i) b[i] = h->fr[i] ? h->force[i] : 0.0;
  ICC_HIP(hipMemcpyAsync(h->d_f, b.data(), n * 8, hipMemcpyHostToDevice, h->ctx->stream));
  ICC_HIP(hipStreamSynchronize(h->ctx->stream));
  return ICC_OK;
}
