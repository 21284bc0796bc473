// Oracle C API for ctypes (tests/, smoke check, bench cpu_baseline only).
// Test infrastructure: the product never links this library.
#include <algorithm>
#include <cmath>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "oracle.hpp"

using namespace orc;
using i64 = int64_t;

namespace {
struct SpaceH {
  SpaceBox sp;
  bool have_em = false;
  ElementMatrices em;
  std::unique_ptr<CsrMatrix> csr;
};
struct ImmH {
  ImmersedProblem pb;
};
struct TopoH {
  std::unique_ptr<CantileverTopOpt> t;
};
}  // namespace

#define ORC_API extern "C" __attribute__((visibility("default")))

ORC_API void* orc_space_new(i64 dim, const double* bounds, const i64* nsub, i64 degree) {
  auto* h = new SpaceH();
  std::vector<std::array<double, 2>> b;
  std::vector<usize> n;
  for (i64 d = 0; d < dim; ++d) {
    b.push_back({bounds[2 * d], bounds[2 * d + 1]});
    n.push_back((usize)nsub[d]);
  }
  h->sp = SpaceBox::make(b, n, (usize)degree);
  return h;
}
ORC_API void orc_space_free(void* p) { delete (SpaceH*)p; }

// out: dim, nel, ndof_sc, ndof, nsh, nsh_sc, ndof_dir[3], nel_dir[3]
ORC_API void orc_space_sizes(void* p, i64* out) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  out[0] = (i64)s.dim;
  out[1] = (i64)s.nel;
  out[2] = (i64)s.ndof_sc;
  out[3] = (i64)s.ndof;
  out[4] = (i64)s.nsh;
  out[5] = (i64)s.nsh_sc;
  for (int d = 0; d < 3; ++d) {
    out[6 + d] = d < (int)s.dim ? (i64)s.ndof_dir[d] : 1;
    out[9 + d] = d < (int)s.dim ? (i64)s.nel_dir[d] : 1;
  }
}

ORC_API void orc_space_connectivity(void* p, i64* out) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  for (usize i = 0; i < s.connectivity.size(); ++i) out[i] = (i64)s.connectivity[i];
}

ORC_API i64 orc_space_elem_mats(void* p, double lambda, double mu) {
  auto* h = (SpaceH*)p;
  h->em = elasticity_element_matrices(h->sp, lambda, mu);
  h->have_em = true;
  return (i64)h->em.types.size();
}

ORC_API void orc_space_types(void* p, double* types, i64* type_id) {
  auto* h = (SpaceH*)p;
  usize nn = h->em.nsh * h->em.nsh;
  if (types)
    for (usize t = 0; t < h->em.types.size(); ++t) std::memcpy(types + t * nn, h->em.types[t].data(), nn * sizeof(double));
  if (type_id)
    for (usize e = 0; e < h->em.type_id.size(); ++e) type_id[e] = (i64)h->em.type_id[e];
}

// Body force of the reference's kernel fixture (iga_vs_matlab.rs:62-68).
ORC_API void orc_space_load_fixture(void* p, double* out) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  usize dim = s.dim;
  Vec f = load_vector(s, [dim](const std::array<double, 3>& x) -> std::array<double, 3> {
    if (dim == 2) return {x[0] * x[1], std::sin(x[0]) + x[1] * x[1], 0.0};
    return {x[0] * x[1], std::sin(x[0]) + x[2], x[1] * x[1]};
  });
  std::memcpy(out, f.data(), f.size() * sizeof(double));
}

// Consistent traction t on the face x = x_max: F(A,c) = t_c * int_face N_A dA,
// by (p+1)-point Gauss on each face element (exact for the spline trace).
ORC_API void orc_space_traction_xmax(void* p, const double* t, double* out) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  std::memset(out, 0, s.ndof * sizeof(double));
  const Space1d& uy = s.univ[1];
  const Space1d& uz = s.univ[2];
  usize i_last = s.ndof_dir[0] - 1;
  for (usize ez = 0; ez < uz.nel; ++ez)
    for (usize ey = 0; ey < uy.nel; ++ey)
      for (usize qz = 0; qz < uz.nquad; ++qz)
        for (usize qy = 0; qy < uy.nquad; ++qy) {
          double w = uy.qw[ey * uy.nquad + qy] * s.lengths[1] * uz.qw[ez * uz.nquad + qz] * s.lengths[2];
          for (usize az = 0; az < uz.nsh(); ++az)
            for (usize ay = 0; ay < uy.nsh(); ++ay) {
              double n = uy.shape[(ey * uy.nquad + qy) * uy.nsh() + ay] * uz.shape[(ez * uz.nquad + qz) * uz.nsh() + az];
              usize gy = uy.connectivity[ey * uy.nsh() + ay], gz = uz.connectivity[ez * uz.nsh() + az];
              usize g = i_last + s.ndof_dir[0] * (gy + s.ndof_dir[1] * gz);
              for (usize c = 0; c < 3; ++c) out[c * s.ndof_sc + g] += w * n * t[c];
            }
        }
}

ORC_API void orc_space_greville(void* p, double* out) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  auto g = s.greville_points();
  for (usize i = 0; i < g.size(); ++i)
    for (int d = 0; d < 3; ++d) out[i * 3 + d] = g[i][d];
}

// Assemble sum_e scale_e K_e (sparse.rs:179-200; scale NULL = 1) and cache it.
ORC_API i64 orc_space_assemble(void* p, const double* scale) {
  auto* h = (SpaceH*)p;
  if (!h->have_em) return -1;
  h->csr.reset(new CsrMatrix(orc::assemble(h->sp.ndof, h->sp.connectivity, h->sp.nsh, h->sp.nel, h->em, scale)));
  return (i64)h->csr->nnz();
}

// Assemble with the reusable pattern path (AssemblyPattern, sparse.rs:128-175).
ORC_API i64 orc_space_assemble_pattern(void* p, const double* scale) {
  auto* h = (SpaceH*)p;
  if (!h->have_em) return -1;
  AssemblyPattern pat(h->sp.ndof, h->sp.connectivity, h->sp.nsh, h->sp.nel);
  Vec s(scale, scale + h->sp.nel);
  h->csr.reset(new CsrMatrix(pat.assemble(h->em, s)));
  return (i64)h->csr->nnz();
}

ORC_API void orc_csr_matvec(void* p, const double* x, double* y) { ((SpaceH*)p)->csr->matvec(x, y); }
ORC_API void orc_csr_diag(void* p, double* d) {
  Vec v = ((SpaceH*)p)->csr->diagonal();
  std::memcpy(d, v.data(), v.size() * sizeof(double));
}
ORC_API double orc_csr_frobenius(void* p) { return ((SpaceH*)p)->csr->frobenius(); }

// Masked Jacobi PCG on the cached CSR (the closure of topopt_iga.rs:151-159).
ORC_API void orc_csr_pcg(void* p, const double* b, const double* diag, const uint8_t* fr, double tol, i64 maxit,
                         double* x, i64* iters, double* res) {
  auto* h = (SpaceH*)p;
  usize n = h->sp.ndof;
  Vec bb(b, b + n), dd(diag, diag + n), pm(n), tmp(n);
  PcgResult r = pcg_solve((usize)maxit, tol, n, bb, dd, [&](const double* pv, double* ap) {
    for (usize i = 0; i < n; ++i) pm[i] = fr[i] ? pv[i] : 0.0;
    h->csr->matvec(pm.data(), tmp.data());
    for (usize i = 0; i < n; ++i) ap[i] = fr[i] ? tmp[i] : 0.0;
  });
  std::memcpy(x, r.x.data(), n * sizeof(double));
  *iters = (i64)r.iters;
  *res = r.res;
}

// Unscaled element energies u_e^T K_e u_e (topopt_iga.rs:165-184).
ORC_API void orc_space_energies(void* p, const double* u, double* out) {
  auto* h = (SpaceH*)p;
  const SpaceBox& s = h->sp;
  usize nsh = s.nsh;
#pragma omp parallel for schedule(static)
  for (long e = 0; e < (long)s.nel; ++e) {
    const double* ke = h->em.of_element((usize)e);
    const usize* ce = &s.connectivity[(usize)e * nsh];
    double acc = 0.0;
    for (usize a = 0; a < nsh; ++a) {
      double row = 0.0;
      for (usize bb = 0; bb < nsh; ++bb) row += ke[a * nsh + bb] * u[ce[bb]];
      acc += u[ce[a]] * row;
    }
    out[e] = acc;
  }
}

ORC_API void orc_fast_sens(void* p, const double* u, const double* x, i64 spline, double penal, double emin,
                           double young, double poisson, double* out) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  Vec uu(u, u + s.ndof), xx(x, x + (spline ? s.ndof_sc : s.nel));
  Vec o = fast_sensitivities(s, uu, xx, spline != 0, penal, emin, young, poisson);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// Sensitivity filter rows of topopt_iga.rs:100-126 (structured enumeration).
// mode 0: sensitivity filter dc~ = sum_f w x_f dc_f / (Hs_e max(x_e, 1e-3))
// mode 1: density filter forward  x~ = H x / Hs
// mode 2: density filter backward g_x = H^T (g / Hs)
ORC_API void orc_filter_apply(void* p, double rmin, i64 mode, const double* x, const double* v, double* out) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  auto rows = build_filter_structured(s, rmin);
  usize nel = s.nel;
  Vec hs(nel);
  for (usize e = 0; e < nel; ++e) {
    double a = 0.0;
    for (double w : rows[e].w) a += w;
    hs[e] = a;
  }
  for (usize e = 0; e < nel; ++e) {
    double acc = 0.0;
    const auto& r = rows[e];
    for (usize k = 0; k < r.f.size(); ++k) {
      usize f = r.f[k];
      if (mode == 0)
        acc += r.w[k] * x[f] * v[f];
      else if (mode == 1)
        acc += r.w[k] * v[f];
      else
        acc += r.w[k] * (v[f] / hs[f]);  // H symmetric: (H^T g)_e = sum_f H_ef g_f
    }
    if (mode == 0)
      out[e] = acc / (hs[e] * std::fmax(x[e], 1e-3));
    else if (mode == 1)
      out[e] = acc / hs[e];
    else
      out[e] = acc;
  }
}

// Design-domain filter (topopt_immersed.rs:117-151) with the modes of
// orc_filter_apply; the row sum of an empty row is 1.
ORC_API void orc_design_filter_apply(void* p, double rmin, i64 mode, const uint8_t* active, const double* x,
                                     const double* v, double* out) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  std::vector<uint8_t> act(active, active + s.nel);
  auto rows = build_filter_design(s, rmin, act);
  usize nel = s.nel;
  Vec hs(nel);
  for (usize e = 0; e < nel; ++e) {
    double a = 0.0;
    for (double w : rows[e].w) a += w;
    hs[e] = a < 1e-6 ? 1.0 : a;
  }
  for (usize e = 0; e < nel; ++e) {
    double acc = 0.0;
    const auto& r = rows[e];
    for (usize k = 0; k < r.f.size(); ++k) {
      usize f = r.f[k];
      if (mode == 0)
        acc += r.w[k] * x[f] * v[f];
      else if (mode == 1)
        acc += r.w[k] * v[f];
      else
        acc += r.w[k] * (v[f] / hs[f]);
    }
    if (mode == 0)
      out[e] = acc / (hs[e] * std::fmax(x[e], 1e-3));
    else if (mode == 1)
      out[e] = acc / hs[e];
    else
      out[e] = acc;
  }
}

// Filter row check: returns 1 if the structured rows equal the O(nel^2) scan.
ORC_API i64 orc_filter_rows_match(void* p, double rmin) {
  const SpaceBox& s = ((SpaceH*)p)->sp;
  auto a = build_filter_structured(s, rmin), b = build_filter_bruteforce(s, rmin);
  for (usize e = 0; e < s.nel; ++e) {
    if (a[e].f != b[e].f) return 0;
    if (a[e].w != b[e].w) return 0;
  }
  return 1;
}

// ---- structured SIMP cantilever -----------------------------------------
static CantileverCfg make_cfg(const i64* nel, const double* size, const double* par) {
  CantileverCfg c;
  for (int d = 0; d < 3; ++d) {
    c.nel[d] = (usize)nel[d];
    c.size[d] = size[d];
  }
  c.volfrac = par[0];
  c.penal = par[1];
  c.rmin = par[2];
  c.max_iter = (usize)par[3];
  c.young = par[4];
  c.poisson = par[5];
  c.emin = par[6];
  c.move_limit = par[7];
  c.pcg_tol = par[8];
  return c;
}

// par = volfrac, penal, rmin, max_iter, young, poisson, emin, move, pcg_tol
ORC_API i64 orc_topopt_run(const i64* nel, const double* size, const double* par, double* compliance, double* change,
                           i64* pcg_iters, double* density) {
  CantileverTopOpt t(make_cfg(nel, size, par));
  auto h = t.run();
  for (usize i = 0; i < h.compliance.size(); ++i) {
    compliance[i] = h.compliance[i];
    change[i] = h.change[i];
    pcg_iters[i] = (i64)h.pcg_iterations[i];
  }
  std::memcpy(density, h.density.data(), h.density.size() * sizeof(double));
  return (i64)h.compliance.size();
}

ORC_API void* orc_topopt_new(const i64* nel, const double* size, const double* par) {
  auto* h = new TopoH();
  h->t.reset(new CantileverTopOpt(make_cfg(nel, size, par)));
  return h;
}
ORC_API void orc_topopt_free(void* p) { delete (TopoH*)p; }
ORC_API void orc_topopt_force_free(void* p, double* force, uint8_t* fr) {
  auto& t = *((TopoH*)p)->t;
  std::memcpy(force, t.force.data(), t.force.size() * sizeof(double));
  std::memcpy(fr, t.free_.data(), t.free_.size());
}
ORC_API void orc_topopt_solve_state(void* p, const double* x, const double* u0, double* u, i64* iters, double* res) {
  auto& t = *((TopoH*)p)->t;
  Vec xx(x, x + t.sp.nel), uu(u0, u0 + t.sp.ndof);
  PcgResult r = t.solve_state(xx, uu);
  std::memcpy(u, r.x.data(), r.x.size() * sizeof(double));
  *iters = (i64)r.iters;
  *res = r.res;
}

// ---- immersed -------------------------------------------------------------
static TriangleMesh3D make_mesh(const double* nodes, i64 nv, const i64* tris, i64 nt) {
  TriangleMesh3D m;
  for (i64 i = 0; i < nv; ++i) m.vertices.push_back({nodes[3 * i], nodes[3 * i + 1], nodes[3 * i + 2]});
  for (i64 i = 0; i < nt; ++i) m.triangles.push_back({(usize)tris[3 * i], (usize)tris[3 * i + 1], (usize)tris[3 * i + 2]});
  return m;
}

// par = young, poisson, gamma, emin, padding; ipar = res[3], degree, stab_kind, clamp, sub[3]
ORC_API void* orc_immersed_new(const double* nodes, i64 nv, const i64* tris, i64 nt, const i64* ipar, const double* par) {
  auto* h = new ImmH();
  TriangleMesh3D m = make_mesh(nodes, nv, tris, nt);
  h->pb = build_immersed_problem(m, {(usize)ipar[0], (usize)ipar[1], (usize)ipar[2]}, (usize)ipar[3], par[0], par[1],
                                 (int)ipar[4], par[2], par[3], {(usize)ipar[6], (usize)ipar[7], (usize)ipar[8]},
                                 (int)ipar[5], par[4]);
  return h;
}
ORC_API void orc_immersed_free(void* p) { delete (ImmH*)p; }
ORC_API i64 orc_immersed_ndof(void* p) { return (i64)((ImmH*)p)->pb.sp.ndof; }
// out: bounds[6], stab_frobenius
ORC_API void orc_immersed_info(void* p, double* bounds, double* stab_fro, int8_t* status, double* weights,
                               double* force, uint8_t* fr) {
  auto& pb = ((ImmH*)p)->pb;
  for (int d = 0; d < 3; ++d) {
    bounds[2 * d] = pb.grid_bounds[d][0];
    bounds[2 * d + 1] = pb.grid_bounds[d][1];
  }
  *stab_fro = pb.stabilization_frobenius;
  std::memcpy(status, pb.status.data(), pb.status.size());
  std::memcpy(weights, pb.weights.data(), pb.weights.size() * sizeof(double));
  std::memcpy(force, pb.force.data(), pb.force.size() * sizeof(double));
  std::memcpy(fr, pb.free_.data(), pb.free_.size());
}
ORC_API void orc_immersed_matvec(void* p, const double* x, double* y) { ((ImmH*)p)->pb.stiffness.matvec(x, y); }
ORC_API void orc_immersed_solve(void* p, double tol, i64 maxit, double* u, double* compliance, i64* iters, double* res) {
  auto& pb = ((ImmH*)p)->pb;
  ImmersedSolution s = solve_immersed_problem(pb, tol, (usize)maxit);
  std::memcpy(u, s.u.data(), s.u.size() * sizeof(double));
  *compliance = s.compliance;
  *iters = (i64)s.iterations;
  *res = s.residual;
}

// ipar = res[3], max_iter, stab_kind ; par = volfrac, penal, rmin, young, poisson, emin, gamma, pcg_tol
ORC_API i64 orc_immersed_topopt_run(const double* nodes, i64 nv, const i64* tris, i64 nt, const i64* ipar,
                                    const double* par, double* compliance, double* volume, double* density,
                                    double* change, i64* pcg_iters) {
  ImmersedTopOptCfg c;
  for (int d = 0; d < 3; ++d) c.grid_res[d] = (usize)ipar[d];
  c.max_iter = (usize)ipar[3];
  c.stab_kind = (int)ipar[4];
  c.volfrac = par[0];
  c.penal = par[1];
  c.rmin = par[2];
  c.young = par[3];
  c.poisson = par[4];
  c.emin = par[5];
  c.gamma = par[6];
  c.pcg_tol = par[7];
  auto h = immersed_topopt_run(make_mesh(nodes, nv, tris, nt), c);
  for (usize i = 0; i < h.compliance.size(); ++i) {
    compliance[i] = h.compliance[i];
    volume[i] = h.volume[i];
    if (change) change[i] = h.change[i];
    if (pcg_iters) pcg_iters[i] = (i64)h.pcg_iterations[i];
  }
  std::memcpy(density, h.density.data(), h.density.size() * sizeof(double));
  return (i64)h.compliance.size();
}

// Classification with the cylinder-hole SDF predicate (a box minus cylinders
// parallel to z): inside iff inside the box and outside every cylinder, with a
// strict sdf < 0.  cyl = [cx, cy, r] * ncyl.  Uses the reference's corner and
// sub-cell logic (immersed.rs:88-155) with this predicate in place of the ray
// caster.
ORC_API void orc_classify_cylinders(const double* box_lo, const double* box_hi, const double* cyl, i64 ncyl,
                                    const double* grid_bounds, const i64* res, const i64* sub, int8_t* status,
                                    double* weights) {
  InsideFn inside = box_minus_cylinders(box_lo, box_hi, std::vector<double>(cyl, cyl + 3 * ncyl));
  std::array<std::array<double, 2>, 3> gb;
  for (int d = 0; d < 3; ++d) gb[d] = {grid_bounds[2 * d], grid_bounds[2 * d + 1]};
  Vec w;
  std::vector<int8_t> st;
  element_weights(inside, gb, {(usize)res[0], (usize)res[1], (usize)res[2]}, {(usize)sub[0], (usize)sub[1], (usize)sub[2]},
                  w, st);
  std::memcpy(status, st.data(), st.size());
  std::memcpy(weights, w.data(), w.size() * sizeof(double));
}

// element_weights with the ParityRayCaster predicate (immersed.rs:121-155),
// and padded_bounds (immersed.rs:450-463)
ORC_API void orc_classify_mesh(const double* nodes, i64 nv, const i64* tris, i64 nt, const double* grid_bounds,
                               const i64* res, const i64* sub, int8_t* status, double* weights) {
  TriangleMesh3D m = make_mesh(nodes, nv, tris, nt);
  ParityRayCaster rc(m);
  std::array<std::array<double, 2>, 3> gb;
  for (int d = 0; d < 3; ++d) gb[d] = {grid_bounds[2 * d], grid_bounds[2 * d + 1]};
  Vec w;
  std::vector<int8_t> st;
  element_weights([&](const std::array<double, 3>& q) { return rc.is_inside(q); }, gb,
                  {(usize)res[0], (usize)res[1], (usize)res[2]}, {(usize)sub[0], (usize)sub[1], (usize)sub[2]}, w, st);
  std::memcpy(status, st.data(), st.size());
  std::memcpy(weights, w.data(), w.size() * sizeof(double));
}
ORC_API void orc_padded_bounds(const double* nodes, i64 nv, double padding, double* out) {
  TriangleMesh3D m;
  for (i64 i = 0; i < nv; ++i) m.vertices.push_back({nodes[3 * i], nodes[3 * i + 1], nodes[3 * i + 2]});
  auto b = padded_bounds(m, padding);
  for (int d = 0; d < 3; ++d) {
    out[2 * d] = b[d][0];
    out[2 * d + 1] = b[d][1];
  }
}

// y = G x with G the ghost-penalty matrix (immersed.rs:234-370) of the grid
// (bounds[6], res[3], degree) and cell status; returns |G|_F.
ORC_API double orc_ghost_penalty_matvec(const double* bounds, const i64* res, i64 degree, const int8_t* status,
                                        double gamma, double young, const double* x, double* y) {
  SpaceBox sp = SpaceBox::make({{bounds[0], bounds[1]}, {bounds[2], bounds[3]}, {bounds[4], bounds[5]}},
                               {(usize)res[0], (usize)res[1], (usize)res[2]}, (usize)degree);
  std::vector<int8_t> st(status, status + sp.nel);
  std::vector<usize> r, c;
  Vec v;
  ghost_penalty(sp, st, gamma, young, r, c, v);
  CsrMatrix g = CsrMatrix::from_triplets(sp.ndof, sp.ndof, r, c, v);
  g.matvec(x, y);
  return g.frobenius();
}

// ---- analytic immersed problem (handle) ----
struct SdfH {
  SdfImmersed pb;
};
// par: young, poisson, emin, active_min, penalty_factor, tx, ty, tz
ORC_API void* orc_sdf_new(const double* box_lo, const double* box_hi, const double* cyl, i64 ncyl, const i64* res,
                          const i64* sub, const double* par, int mode) {
  auto* h = new SdfH();
  h->pb = build_sdf_immersed(box_lo, box_hi, std::vector<double>(cyl, cyl + 3 * ncyl),
                             {(usize)res[0], (usize)res[1], (usize)res[2]}, {(usize)sub[0], (usize)sub[1], (usize)sub[2]},
                             par[0], par[1], par[2], par[3], par[4], par + 5, mode);
  return h;
}
ORC_API void orc_sdf_free(void* p) { delete (SdfH*)p; }
// out: nel, ndof, nnz, beta bits are returned separately
ORC_API double orc_sdf_sizes(void* p, i64* out) {
  const SdfImmersed& pb = ((SdfH*)p)->pb;
  out[0] = (i64)pb.sp.nel;
  out[1] = (i64)pb.sp.ndof;
  out[2] = (i64)pb.op.nnz();
  return pb.beta;
}
ORC_API void orc_sdf_get(void* p, int8_t* status, double* w, uint8_t* active, double* scale, uint8_t* fr,
                         double* force) {
  const SdfImmersed& pb = ((SdfH*)p)->pb;
  std::memcpy(status, pb.status.data(), pb.status.size());
  std::memcpy(w, pb.weights.data(), pb.weights.size() * 8);
  std::memcpy(active, pb.active.data(), pb.active.size());
  std::memcpy(scale, pb.scale.data(), pb.scale.size() * 8);
  std::memcpy(fr, pb.free_.data(), pb.free_.size());
  std::memcpy(force, pb.force.data(), pb.force.size() * 8);
}
// y = M op M x
ORC_API void orc_sdf_matvec(void* p, const double* x, double* y) {
  const SdfImmersed& pb = ((SdfH*)p)->pb;
  const usize n = pb.sp.ndof;
  Vec xm(n);
  for (usize i = 0; i < n; ++i) xm[i] = pb.free_[i] ? x[i] : 0.0;
  pb.op.matvec(xm.data(), y);
  for (usize i = 0; i < n; ++i)
    if (!pb.free_[i]) y[i] = 0.0;
}
// stats: iterations, residual, compliance
ORC_API void orc_sdf_solve(void* p, double tol, i64 maxit, double* u, double* stats) {
  const SdfImmersed& pb = ((SdfH*)p)->pb;
  ImmersedSolution sol = solve_sdf_immersed(pb, tol, (usize)maxit);
  std::memcpy(u, sol.u.data(), sol.u.size() * 8);
  stats[0] = (double)sol.iterations;
  stats[1] = sol.residual;
  stats[2] = sol.compliance;
}

// ---- weighted quadrature ---------------------------------------------------
// which: 0 points, 1.. = weights[t][i] with t = which - 1
ORC_API i64 orc_wq_rule(i64 nel, i64 p, i64 layout, i64 which, i64 i, double* out) {
  Vec knots = open_knots((usize)nel, (usize)p, p - 1);
  WqRules1d r = wq_rules_1d(knots, (usize)p, (int)layout);
  const Vec& v = which == 0 ? r.points : r.weights[which - 1][(usize)i];
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(double));
  return (i64)v.size();
}

// Assemble the WQ stiffness into the space handle's cached CSR.
ORC_API i64 orc_space_assemble_wq(void* p, double young, double poisson, i64 density_kind, const double* x,
                                  double penal, double emin, i64 layout) {
  auto* h = (SpaceH*)p;
  h->csr.reset(new CsrMatrix(
      wq_stiffness_layout(h->sp, young, poisson, (int)density_kind, x, penal, emin, (int)layout)));
  return (i64)h->csr->nnz();
}

// ---- CPU baseline support -------------------------------------------------------
// Builds the same CSR matrix as AssemblyPattern::new + assemble
// (sparse.rs:128-175) for 3D p = 2 grids without the nel*81^2 position map:
// rows enumerate the 3 x 5^3 neighbour columns in increasing index, and each
// entry sums s_e K_e[a][b] over the elements holding both points in
// increasing e, starting from 0.0 -- the order assemble() uses, so the values
// are bitwise equal.  Rows are built in parallel.
ORC_API i64 orc_space_assemble_structured(void* p, const double* scale) {
  auto* h = (SpaceH*)p;
  const SpaceBox& s = h->sp;
  if (!h->have_em || s.dim != 3 || s.nsh != 81) return -1;
  const long n0 = (long)s.ndof_dir[0], n1 = (long)s.ndof_dir[1], n2 = (long)s.ndof_dir[2];
  const long e0 = (long)s.nel_dir[0], e1 = (long)s.nel_dir[1], e2 = (long)s.nel_dir[2];
  const long nsc = (long)s.ndof_sc, n = 3 * nsc;
  auto* m = new CsrMatrix();
  m->nrows = m->ncols = (usize)n;
  m->indptr.assign((usize)n + 1, 0);
  auto span = [](long p_, long np) { return std::min(p_ + 2, np - 1) - std::max(p_ - 2, 0L) + 1; };
  for (long r = 0; r < n; ++r) {
    long P = r % nsc;
    long pi = P % n0, pj = (P / n0) % n1, pk = P / (n0 * n1);
    m->indptr[(usize)r + 1] = (usize)(3 * span(pi, n0) * span(pj, n1) * span(pk, n2));
  }
  for (long r = 0; r < n; ++r) m->indptr[(usize)r + 1] += m->indptr[(usize)r];
  m->indices.resize(m->indptr[(usize)n]);
  m->data.resize(m->indptr[(usize)n]);
#pragma omp parallel for schedule(dynamic, 256)
  for (long r = 0; r < n; ++r) {
    const long c = r / nsc, P = r % nsc;
    const long pi = P % n0, pj = (P / n0) % n1, pk = P / (n0 * n1);
    usize k = m->indptr[(usize)r];
    for (long c2 = 0; c2 < 3; ++c2)
      for (long qk = std::max(pk - 2, 0L); qk <= std::min(pk + 2, n2 - 1); ++qk)
        for (long qj = std::max(pj - 2, 0L); qj <= std::min(pj + 2, n1 - 1); ++qj)
          for (long qi = std::max(pi - 2, 0L); qi <= std::min(pi + 2, n0 - 1); ++qi, ++k) {
            m->indices[k] = (usize)(c2 * nsc + qi + n0 * (qj + n1 * qk));
            double v = 0.0;
            for (long ek = std::max(std::max(pk, qk) - 2, 0L); ek <= std::min(std::min(pk, qk), e2 - 1); ++ek)
              for (long ej = std::max(std::max(pj, qj) - 2, 0L); ej <= std::min(std::min(pj, qj), e1 - 1); ++ej)
                for (long ei = std::max(std::max(pi, qi) - 2, 0L); ei <= std::min(std::min(pi, qi), e0 - 1); ++ei) {
                  const long e = ei + e0 * (ej + e1 * ek);
                  const long a = (pi - ei) + 3 * ((pj - ej) + 3 * (pk - ek));
                  const long b = (qi - ei) + 3 * ((qj - ej) + 3 * (qk - ek));
                  const double* ke = h->em.of_element((usize)e);
                  v += (scale ? scale[e] : 1.0) * ke[(c * 27 + a) * 81 + (c2 * 27 + b)];
                }
            m->data[k] = v;
          }
  }
  h->csr.reset(m);
  return (i64)m->nnz();
}

// Compares the cached CSR with a fresh AssemblyPattern assembly: returns the
// number of differing entries (pattern or bits); -1 on pattern mismatch.
ORC_API i64 orc_compare_with_pattern(void* p, const double* scale) {
  auto* h = (SpaceH*)p;
  AssemblyPattern pat(h->sp.ndof, h->sp.connectivity, h->sp.nsh, h->sp.nel);
  Vec s(scale, scale + h->sp.nel);
  CsrMatrix ref = pat.assemble(h->em, s);
  if (ref.indptr != h->csr->indptr || ref.indices != h->csr->indices) return -1;
  i64 diff = 0;
  for (usize k = 0; k < ref.data.size(); ++k) diff += (ref.data[k] != h->csr->data[k]);
  return diff;
}

ORC_API i64 orc_num_threads(void) {
#ifdef _OPENMP
  return (i64)omp_get_max_threads();
#else
  return 1;
#endif
}
ORC_API void orc_set_threads(i64 n) {
#ifdef _OPENMP
  omp_set_num_threads((int)n);
#else
  (void)n;
#endif
}

// ---- surface boundary terms (immersed_bc.rs) --------------------------------
namespace {
struct SurfH {
  const SpaceBox* sp = nullptr;
  SurfaceAccumulator acc;
  SurfaceTerms st;
  bool done = false;
};
std::vector<usize> tri_list(const i64* sel, i64 nsel) {
  std::vector<usize> t(nsel);
  for (i64 i = 0; i < nsel; ++i) t[i] = (usize)sel[i];
  return t;
}
}  // namespace

// (x, w, n) per point as 7 doubles; returns the count (out may be NULL)
ORC_API i64 orc_surface_quadrature(const double* nodes, i64 nv, const i64* tris, i64 nt, const i64* sel, i64 nsel,
                                   double max_edge, double* out) {
  TriangleMesh3D m = make_mesh(nodes, nv, tris, nt);
  auto q = surface_quadrature(m, tri_list(sel, nsel), max_edge);
  if (out)
    for (usize i = 0; i < q.size(); ++i) {
      for (int d = 0; d < 3; ++d) out[7 * i + d] = q[i].x[d];
      out[7 * i + 3] = q[i].w;
      for (int d = 0; d < 3; ++d) out[7 * i + 4 + d] = q[i].n[d];
    }
  return (i64)q.size();
}
// space: an orc_space handle (imm == 0) or an orc_immersed handle (imm == 1)
static const SpaceBox* space_of(void* p, i64 imm) { return imm ? &((ImmH*)p)->pb.sp : &((SpaceH*)p)->sp; }
ORC_API i64 orc_scalar_basis_at(void* p, i64 imm, const double* x, i64* idx, double* val) {
  auto nz = scalar_basis_at(*space_of(p, imm), {x[0], x[1], x[2]});
  for (usize a = 0; a < nz.size(); ++a) {
    idx[a] = (i64)nz[a].first;
    val[a] = nz[a].second;
  }
  return (i64)nz.size();
}
ORC_API void* orc_surface_new(void* p, i64 imm) {
  auto* h = new SurfH();
  h->sp = space_of(p, imm);
  h->st.force.assign(h->sp->ndof, 0.0);
  return h;
}
ORC_API void orc_surface_free(void* h) { delete (SurfH*)h; }
// one condition of surface_boundary_terms: kind 0 Dirichlet (par = comps[3], values[3]),
// 1 traction (par[3..6] = t), 2 pressure (par[6]), 3 Robin (par[7], par[8] = kn, kt);
// beta = the Dirichlet penalty
ORC_API void orc_surface_add(void* hp, i64 kind, const double* nodes, i64 nv, const i64* tris, i64 nt, const i64* sel,
                             i64 nsel, const double* par, double beta) {
  auto* h = (SurfH*)hp;
  TriangleMesh3D m = make_mesh(nodes, nv, tris, nt);
  SurfaceBc bc;
  bc.kind = (SurfaceBc::Kind)kind;
  bc.tris = tri_list(sel, nsel);
  for (int c = 0; c < 3; ++c) {
    bc.components[c] = par[c] != 0.0;
    bc.values[c] = par[3 + c];
  }
  bc.pressure = par[6];
  bc.kn = par[7];
  bc.kt = par[8];
  surface_add_condition(*h->sp, m, bc, beta, h->acc, h->st);
}
// dirichlet_penalty_fn with the affine g(x) = g0 + G x (G row-major 3x3)
ORC_API void orc_surface_dirichlet_affine(void* hp, const double* nodes, i64 nv, const i64* tris, i64 nt,
                                          const i64* sel, i64 nsel, double beta, const double* g0, const double* G) {
  auto* h = (SurfH*)hp;
  TriangleMesh3D m = make_mesh(nodes, nv, tris, nt);
  const std::array<double, 3> a = {g0[0], g0[1], g0[2]};
  std::array<double, 9> M;
  for (int i = 0; i < 9; ++i) M[i] = G[i];
  SurfaceTerms st = dirichlet_penalty_fn(*h->sp, m, tri_list(sel, nsel), beta, [&](const std::array<double, 3>& x) {
    std::array<double, 3> r;
    for (int i = 0; i < 3; ++i) r[i] = a[i] + M[3 * i] * x[0] + M[3 * i + 1] * x[1] + M[3 * i + 2] * x[2];
    return r;
  });
  h->st = st;
  h->done = true;
}
// finalises (into_triplets) once; returns the triplet count; rows/cols/vals/force may be NULL
ORC_API i64 orc_surface_get(void* hp, i64* rows, i64* cols, double* vals, double* force) {
  auto* h = (SurfH*)hp;
  if (!h->done) {
    h->acc.into_triplets(h->sp->ndof_sc, h->st);
    h->done = true;
  }
  const usize n = h->st.vals.size();
  for (usize k = 0; k < n; ++k) {
    if (rows) rows[k] = (i64)h->st.rows[k];
    if (cols) cols[k] = (i64)h->st.cols[k];
    if (vals) vals[k] = h->st.vals[k];
  }
  if (force) std::memcpy(force, h->st.force.data(), h->st.force.size() * sizeof(double));
  return (i64)n;
}
// ImmersedProblem::add_surface_terms on an orc_immersed handle
ORC_API void orc_immersed_add_surface(void* ip, void* hp) {
  auto* h = (SurfH*)hp;
  orc_surface_get(hp, nullptr, nullptr, nullptr, nullptr);
  add_surface_terms(((ImmH*)ip)->pb, h->st);
}
// ImmersedProblem::force is zeroed by the callers of surface terms (immersed_cad_bc.rs:129)
ORC_API void orc_immersed_zero_force(void* ip) {
  auto& f = ((ImmH*)ip)->pb.force;
  std::fill(f.begin(), f.end(), 0.0);
}
// u at a physical point (fields.rs eval_vector_at)
ORC_API void orc_immersed_eval_vector_at(void* ip, const double* u, const double* x, double* out) {
  auto& pb = ((ImmH*)ip)->pb;
  Vec uu(u, u + pb.sp.ndof);
  auto r = eval_vector_at(pb.sp, uu, {x[0], x[1], x[2]});
  for (int d = 0; d < 3; ++d) out[d] = r[d];
}
