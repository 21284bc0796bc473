// Oracle: structured SIMP cantilever topology optimisation (restates
// rust/src/topopt_iga.rs), quadrature-point sensitivities (restates
// rust/src/topopt2d.rs:13-48 fast_sensitivities), and the textbook density
// filter used by the adjoint SIMP configuration (no reference counterpart:
// x~ = H x / Hs, dc/dx = H^T (dc/dx~ / Hs); Bourdin 2001, Andreassen et al.
// 2011).  Test infrastructure only (see oracle.hpp).
#include <algorithm>
#include <cmath>

#include "oracle.hpp"

namespace orc {

static inline double sq(double v) { return v * v; }  // Rust powi(2) == v * v

// topopt_iga.rs:100-126: r = rmin * mean(h), centres at (s + 0.5) h,
// d2 <= r^2, weight max(r - sqrt(d2), 0); O(nel^2) scan in increasing f.
std::vector<FilterRow> build_filter_bruteforce(const SpaceBox& sp, double rmin) {
  Vec h = sp.element_size();
  double r = rmin * (h[0] + h[1] + h[2]) / 3.0;
  std::vector<std::array<double, 3>> centres(sp.nel);
  for (usize e = 0; e < sp.nel; ++e) {
    auto s = unravel(e, sp.nel_dir);
    centres[e] = {((double)s[0] + 0.5) * h[0], ((double)s[1] + 0.5) * h[1], ((double)s[2] + 0.5) * h[2]};
  }
  std::vector<FilterRow> filter(sp.nel);
  for (usize e = 0; e < sp.nel; ++e) {
    auto ce = centres[e];
    for (usize f = 0; f < sp.nel; ++f) {
      auto cf = centres[f];
      double d2 = sq(ce[0] - cf[0]) + sq(ce[1] - cf[1]) + sq(ce[2] - cf[2]);
      if (d2 <= r * r) {
        filter[e].f.push_back(f);
        filter[e].w.push_back(std::fmax(r - std::sqrt(d2), 0.0));
      }
    }
  }
  return filter;
}

// Same rows as build_filter_bruteforce: candidates are limited to the box of
// offsets |d_i| <= ceil(r / h_i) + 1 (a superset of the ball), visited with z
// slowest and x fastest, i.e. in increasing linear index f, and tested with
// the identical d2 expression on the identical centres.
std::vector<FilterRow> build_filter_structured(const SpaceBox& sp, double rmin) {
  Vec h = sp.element_size();
  double r = rmin * (h[0] + h[1] + h[2]) / 3.0;
  long R[3];
  for (int d = 0; d < 3; ++d) R[d] = (long)std::ceil(r / h[d]) + 1;
  long n[3] = {(long)sp.nel_dir[0], (long)sp.nel_dir[1], (long)sp.nel_dir[2]};
  std::vector<FilterRow> filter(sp.nel);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < (long)sp.nel; ++e) {
    long s0 = e % n[0], s1 = (e / n[0]) % n[1], s2 = e / (n[0] * n[1]);
    double ce[3] = {((double)s0 + 0.5) * h[0], ((double)s1 + 0.5) * h[1], ((double)s2 + 0.5) * h[2]};
    for (long k = std::max(0L, s2 - R[2]); k <= std::min(n[2] - 1, s2 + R[2]); ++k)
      for (long j = std::max(0L, s1 - R[1]); j <= std::min(n[1] - 1, s1 + R[1]); ++j)
        for (long i = std::max(0L, s0 - R[0]); i <= std::min(n[0] - 1, s0 + R[0]); ++i) {
          double cf[3] = {((double)i + 0.5) * h[0], ((double)j + 0.5) * h[1], ((double)k + 0.5) * h[2]};
          double d2 = sq(ce[0] - cf[0]) + sq(ce[1] - cf[1]) + sq(ce[2] - cf[2]);
          if (d2 <= r * r) {
            filter[e].f.push_back((usize)(i + n[0] * (j + n[1] * k)));
            filter[e].w.push_back(std::fmax(r - std::sqrt(d2), 0.0));
          }
        }
  }
  return filter;
}

// topopt_iga.rs:73-129 CantileverTopOpt::new
CantileverTopOpt::CantileverTopOpt(const CantileverCfg& c, bool build_pattern) : cfg(c) {
  sp = SpaceBox::make({{0.0, cfg.size[0]}, {0.0, cfg.size[1]}, {0.0, cfg.size[2]}},
                      {cfg.nel[0], cfg.nel[1], cfg.nel[2]}, cfg.degree);
  double lambda, mu;
  lame(cfg.young, cfg.poisson, lambda, mu);
  em = elasticity_element_matrices(sp, lambda, mu);
  const auto& nd = sp.ndof_dir;
  free_.assign(sp.ndof, 1);
  std::vector<usize> load;
  for (usize i = 0; i < sp.ndof_sc; ++i) {
    auto s = unravel(i, nd);
    if (s[0] == 0)
      for (usize cc = 0; cc < 3; ++cc) free_[cc * sp.ndof_sc + i] = 0;
    usize i1 = s[0] + 1;
    double i2 = (double)s[1] + 1.0, i3 = (double)s[2] + 1.0;
    if (i1 == nd[0] && std::fabs(i2 - 1.0) <= 1.0 && std::fabs(i3 - (double)nd[2] / 2.0) <= 1.0) load.push_back(i);
  }
  force.assign(sp.ndof, 0.0);
  for (usize i : load) force[sp.ndof_sc + i] = -1.0 / (double)load.size();
  filter = build_filter_structured(sp, cfg.rmin);
  filter_sum.resize(sp.nel);
  for (usize e = 0; e < sp.nel; ++e) {
    double s = 0.0;
    for (double w : filter[e].w) s += w;
    filter_sum[e] = s;
  }
  if (build_pattern) pattern = AssemblyPattern(sp.ndof, sp.connectivity, sp.nsh, sp.nel);
}

// topopt_iga.rs:131-134 assemble: scale = emin + (1 - emin) x^penal
CsrMatrix CantileverTopOpt::assemble(const Vec& x) const {
  Vec scale(x.size());
  for (usize e = 0; e < x.size(); ++e) scale[e] = cfg.emin + (1.0 - cfg.emin) * std::pow(x[e], cfg.penal);
  return pattern.assemble(em, scale);
}

// topopt_iga.rs:137-162 solve_state (warm start, rescaled tolerance, masked closure)
PcgResult CantileverTopOpt::solve_state(const Vec& x, const Vec& u0) const {
  CsrMatrix k = assemble(x);
  usize n = sp.ndof;
  Vec ku0(n, 0.0);
  k.matvec(u0.data(), ku0.data());
  Vec b(n);
  for (usize i = 0; i < n; ++i) b[i] = free_[i] ? force[i] - ku0[i] : 0.0;
  double nb = 0.0;
  for (usize i = 0; i < n; ++i)
    if (free_[i]) nb += force[i] * force[i];
  nb = std::sqrt(nb);
  double rb = 0.0;
  for (double v : b) rb += v * v;
  rb = std::sqrt(rb);
  Vec diag = k.diagonal();
  for (double& d : diag) d = std::fmax(d, 1e-12);
  double tol = rb > 0.0 ? std::fmin(cfg.pcg_tol * nb / rb, 0.5) : 1.0;
  Vec pm(n, 0.0), tmp(n, 0.0);
  PcgResult res = pcg_solve(20 * n, tol, n, b, diag, [&](const double* p, double* ap) {
    for (usize i = 0; i < n; ++i) pm[i] = free_[i] ? p[i] : 0.0;
    k.matvec(pm.data(), tmp.data());
    for (usize i = 0; i < n; ++i) ap[i] = free_[i] ? tmp[i] : 0.0;
  });
  PcgResult out;
  out.x.resize(n);
  for (usize i = 0; i < n; ++i) out.x[i] = free_[i] ? u0[i] + res.x[i] : 0.0;
  out.iters = res.iters;
  out.res = res.res * rb / std::fmax(nb, 1e-300);
  return out;
}

// topopt_iga.rs:165-184 element_energies: u_e^T K_e u_e
Vec CantileverTopOpt::element_energies(const Vec& u) const {
  usize nsh = sp.nsh;
  Vec out(sp.nel);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < (long)sp.nel; ++e) {
    const double* ke = em.of_element((usize)e);
    const usize* ce = &sp.connectivity[(usize)e * nsh];
    double s = 0.0;
    for (usize a = 0; a < nsh; ++a) {
      double ua = u[ce[a]];
      double row = 0.0;
      for (usize b = 0; b < nsh; ++b) row += ke[a * nsh + b] * u[ce[b]];
      s += ua * row;
    }
    out[(usize)e] = s;
  }
  return out;
}

// topopt_iga.rs:187-234 run (sensitivity filter + OC bisection)
CantileverTopOpt::History CantileverTopOpt::run() const {
  usize nel = sp.nel;
  Vec x(nel, cfg.volfrac), u(sp.ndof, 0.0);
  History hist;
  for (usize iter = 1; iter <= cfg.max_iter; ++iter) {
    Vec x_old = x;
    PcgResult st = solve_state(x, u);
    u = st.x;
    double c = 0.0;
    for (usize i = 0; i < sp.ndof; ++i) c += force[i] * u[i];
    Vec ee = element_energies(u);
    Vec dc_raw(nel), dc(nel);
    for (usize e = 0; e < nel; ++e)
      dc_raw[e] = -cfg.penal * (1.0 - cfg.emin) * std::pow(x[e], cfg.penal - 1.0) * ee[e];
    for (usize e = 0; e < nel; ++e) {
      double s = 0.0;
      for (usize k = 0; k < filter[e].f.size(); ++k) s += filter[e].w[k] * x[filter[e].f[k]] * dc_raw[filter[e].f[k]];
      dc[e] = s / (filter_sum[e] * std::fmax(x[e], 1e-3));
    }
    double l1 = 0.0, mx = 0.0;
    for (double v : dc) mx = std::fmax(mx, std::fabs(v));
    double l2 = 2.0 * mx;
    Vec cand = x_old;
    while ((l2 - l1) / (l1 + l2 + 1e-10) > 1e-4) {
      double lmid = 0.5 * (l1 + l2);
      for (usize e = 0; e < nel; ++e) {
        double be = std::sqrt(-dc[e] / lmid);
        // f64::min / f64::max return the non-NaN operand, as fmin / fmax do
        cand[e] = std::fmax(
            std::fmax(std::fmin(std::fmin(x_old[e] * be, x_old[e] + cfg.move_limit), 1.0), x_old[e] - cfg.move_limit),
            0.001);
      }
      double sum = 0.0;
      for (double v : cand) sum += v;
      double mean = sum / (double)nel;
      if (mean > cfg.volfrac)
        l1 = lmid;
      else
        l2 = lmid;
    }
    x = cand;
    double change = 0.0;
    for (usize e = 0; e < nel; ++e) change = std::fmax(change, std::fabs(x[e] - x_old[e]));
    hist.compliance.push_back(std::fabs(c));
    hist.change.push_back(change);
    hist.pcg_iterations.push_back(st.iters);
    if (change < 0.01 && iter >= 15) break;
  }
  hist.density = x;
  return hist;
}

// topopt2d.rs:13-48 fast_sensitivities
Vec fast_sensitivities(const SpaceBox& sp, const Vec& u, const Vec& x, bool spline, double penal, double emin,
                       double young, double poisson) {
  double lam, mu;
  lame(young, poisson, lam, mu);
  usize dim = sp.dim;
  Vec out(spline ? sp.ndof_sc : sp.nel, 0.0);
  std::vector<std::array<double, 3>> xq;
  Vec w, nv, dn;
  for (usize e = 0; e < sp.nel; ++e) {
    element_quad_points(sp, e, xq, w);
    auto grads = element_vector_gradients(sp, u, e);
    if (spline) element_shape_functions(sp, e, nv, dn);
    double ee = 0.0;
    for (usize q = 0; q < sp.nqn; ++q) {
      const auto& g = grads[q];
      double tr = 0.0;
      for (usize i = 0; i < dim; ++i) tr += g[i][i];
      double epse = 0.0;
      for (usize i = 0; i < dim; ++i)
        for (usize j = 0; j < dim; ++j) epse += sq(0.5 * (g[i][j] + g[j][i]));
      double sed = lam * tr * tr + 2.0 * mu * epse;
      if (!spline) {
        ee += w[q] * sed;
      } else {
        double rho = 0.0;
        for (usize a = 0; a < sp.nsh_sc; ++a) rho += nv[q * sp.nsh_sc + a] * x[sp.conn_sc[e * sp.nsh_sc + a]];
        double s = penal * std::pow(rho, penal - 1.0) * (1.0 - emin) * sed;
        for (usize a = 0; a < sp.nsh_sc; ++a) out[sp.conn_sc[e * sp.nsh_sc + a]] -= w[q] * nv[q * sp.nsh_sc + a] * s;
      }
    }
    if (!spline) out[e] = -penal * std::pow(x[e], penal - 1.0) * (1.0 - emin) * ee;
  }
  return out;
}

}  // namespace orc
