// Oracle: immersed background-grid classification, volume fractions, the
// legacy cut-cell stabilisation, the consistent ghost penalty, the assembled
// immersed problem and the immersed SIMP loop.  Restates rust/src/cut_cell.rs (new_box),
// rust/src/immersed.rs and rust/src/topopt_immersed.rs of the reference.
// Test infrastructure only (see oracle.hpp).
#include <algorithm>
#include <cmath>
#include <limits>

#include "oracle.hpp"

namespace orc {

static constexpr int8_t OUTSIDE = 0, INSIDE = 1, CUT = -1;

// cut_cell.rs:27-53 TriangleMesh3D::new_box
TriangleMesh3D TriangleMesh3D::new_box(std::array<double, 3> mn, std::array<double, 3> mx) {
  TriangleMesh3D m;
  m.vertices = {{mn[0], mn[1], mn[2]}, {mx[0], mn[1], mn[2]}, {mx[0], mx[1], mn[2]}, {mn[0], mx[1], mn[2]},
                {mn[0], mn[1], mx[2]}, {mx[0], mn[1], mx[2]}, {mx[0], mx[1], mx[2]}, {mn[0], mx[1], mx[2]}};
  m.triangles = {{0, 2, 1}, {0, 3, 2}, {4, 5, 6}, {4, 6, 7}, {0, 1, 5}, {0, 5, 4},
                 {2, 3, 7}, {2, 7, 6}, {0, 4, 7}, {0, 7, 3}, {1, 2, 6}, {1, 6, 5}};
  return m;
}

// immersed.rs:35-46 ParityRayCaster::new
ParityRayCaster::ParityRayCaster(const TriangleMesh3D& mesh) {
  for (auto& t : mesh.triangles) {
    std::array<std::array<double, 3>, 3> tri{mesh.vertices[t[0]], mesh.vertices[t[1]], mesh.vertices[t[2]]};
    tris.push_back(tri);
    auto& a = tri[0];
    auto& b = tri[1];
    auto& c = tri[2];
    yz_box.push_back({std::fmin(std::fmin(a[1], b[1]), c[1]), std::fmax(std::fmax(a[1], b[1]), c[1]),
                      std::fmin(std::fmin(a[2], b[2]), c[2]), std::fmax(std::fmax(a[2], b[2]), c[2])});
  }
}

// immersed.rs:48-75 is_inside (ray parity along +x, barycentric test in y-z)
bool ParityRayCaster::is_inside(const std::array<double, 3>& q) const {
  double qx = q[0], qy = q[1], qz = q[2];
  usize count = 0;
  for (usize t = 0; t < tris.size(); ++t) {
    const auto& bb = yz_box[t];
    if (!(qy >= bb[0] && qy <= bb[1] && qz >= bb[2] && qz <= bb[3])) continue;
    const auto& a = tris[t][0];
    const auto& b = tris[t][1];
    const auto& c = tris[t][2];
    double v0[2] = {c[1] - a[1], c[2] - a[2]};
    double v1[2] = {b[1] - a[1], b[2] - a[2]};
    double v2[2] = {qy - a[1], qz - a[2]};
    double dot00 = v0[0] * v0[0] + v0[1] * v0[1];
    double dot01 = v0[0] * v1[0] + v0[1] * v1[1];
    double dot02 = v0[0] * v2[0] + v0[1] * v2[1];
    double dot11 = v1[0] * v1[0] + v1[1] * v1[1];
    double dot12 = v1[0] * v2[0] + v1[1] * v2[1];
    double inv = 1.0 / (dot00 * dot11 - dot01 * dot01 + 1e-15);
    double u = (dot11 * dot02 - dot01 * dot12) * inv;
    double v = (dot00 * dot12 - dot01 * dot02) * inv;
    if (u >= 0.0 && v >= 0.0 && u + v <= 1.0) {
      double x_hit = a[0] + v * (b[0] - a[0]) + u * (c[0] - a[0]);
      if (x_hit > qx) count += 1;
    }
  }
  return count % 2 == 1;
}

// immersed.rs:79-84 MATLAB-compatible linspace
Vec matlab_linspace(double a, double b, usize n) {
  double step = (b - a) / (double)(n - 1);
  Vec v(n);
  for (usize i = 0; i < n; ++i) v[i] = a + (double)i * step;
  v[n - 1] = b;
  return v;
}

// immersed.rs:88-117 classify_background_cells (8 corners)
std::vector<int8_t> classify_background_cells(const InsideFn& inside, const std::array<std::array<double, 2>, 3>& bounds,
                                              const std::array<usize, 3>& res) {
  std::vector<Vec> xs(3);
  for (int d = 0; d < 3; ++d) xs[d] = matlab_linspace(bounds[d][0], bounds[d][1], res[d] + 1);
  usize nv[3] = {res[0] + 1, res[1] + 1, res[2] + 1};
  std::vector<uint8_t> inside_v(nv[0] * nv[1] * nv[2]);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)inside_v.size(); ++i) {
    usize a = (usize)i % nv[0], b = ((usize)i / nv[0]) % nv[1], c = (usize)i / (nv[0] * nv[1]);
    inside_v[(usize)i] = inside({xs[0][a], xs[1][b], xs[2][c]}) ? 1 : 0;
  }
  std::vector<int8_t> status(res[0] * res[1] * res[2], OUTSIDE);
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  for (usize k = 0; k < res[2]; ++k)
    for (usize j = 0; j < res[1]; ++j)
      for (usize i = 0; i < res[0]; ++i) {
        int s = 0;
        for (auto& o : off)
          if (inside_v[(i + o[0]) + nv[0] * ((j + o[1]) + nv[1] * (k + o[2]))]) s += 1;
        status[i + res[0] * (j + res[1] * k)] = s == 8 ? INSIDE : (s == 0 ? OUTSIDE : CUT);
      }
  return status;
}

// immersed.rs:121-155 element_weights (sub^3 midpoint sampling of cut cells)
void element_weights(const InsideFn& inside, const std::array<std::array<double, 2>, 3>& bounds,
                     const std::array<usize, 3>& res, const std::array<usize, 3>& sub, Vec& weights,
                     std::vector<int8_t>& status) {
  status = classify_background_cells(inside, bounds, res);
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = (bounds[d][1] - bounds[d][0]) / (double)res[d];
  weights.assign(status.size(), 0.0);
#pragma omp parallel for schedule(dynamic, 64)
  for (long ee = 0; ee < (long)status.size(); ++ee) {
    usize e = (usize)ee;
    int8_t s = status[e];
    if (s == INSIDE) {
      weights[e] = 1.0;
    } else if (s == OUTSIDE) {
      weights[e] = 0.0;
    } else {
      usize idx[3] = {e % res[0], (e / res[0]) % res[1], e / (res[0] * res[1])};
      double lo[3], hi[3], dx[3];
      for (int d = 0; d < 3; ++d) {
        lo[d] = bounds[d][0] + (double)idx[d] * h[d];
        hi[d] = bounds[d][0] + (double)(idx[d] + 1) * h[d];
        dx[d] = (hi[d] - lo[d]) / (double)sub[d];
      }
      usize n_in = 0;
      for (usize c = 0; c < sub[2]; ++c)
        for (usize b = 0; b < sub[1]; ++b)
          for (usize a = 0; a < sub[0]; ++a) {
            std::array<double, 3> q{lo[0] + ((double)a + 0.5) * dx[0], lo[1] + ((double)b + 0.5) * dx[1],
                                    lo[2] + ((double)c + 0.5) * dx[2]};
            if (inside(q)) n_in += 1;
          }
      weights[e] = (double)n_in / (double)(sub[0] * sub[1] * sub[2]);
    }
  }
}

// immersed.rs:168-222 legacy_stabilization (MATLAB "ghost penalty" port)
void legacy_stabilization(const SpaceBox& sp, const std::vector<int8_t>& status, double gamma, std::vector<usize>& ri,
                          std::vector<usize>& ci, Vec& vi) {
  usize res[3] = {sp.nel_dir[0], sp.nel_dir[1], sp.nel_dir[2]};
  Vec h = sp.element_size();
  double alpha = gamma * ((h[0] + h[1] + h[2]) / 3.0);
  auto lin = [&](usize i, usize j, usize k) { return i + res[0] * (j + res[1] * k); };
  auto active = [&](usize e) { return status[e] == INSIDE || status[e] == CUT; };
  auto cut = [&](usize e) { return status[e] == CUT; };
  std::vector<std::pair<usize, usize>> faces;
  static const usize dirs[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (auto& dd : dirs)
    for (usize i = 0; i < res[0] - dd[0]; ++i)
      for (usize j = 0; j < res[1] - dd[1]; ++j)
        for (usize k = 0; k < res[2] - dd[2]; ++k) {
          usize e1 = lin(i, j, k), e2 = lin(i + dd[0], j + dd[1], k + dd[2]);
          if (active(e1) && active(e2) && (cut(e1) || cut(e2))) faces.push_back({e1, e2});
        }
  usize nsh_sc = sp.nsh_sc;
  for (auto& f : faces) {
    usize e1 = f.first, e2 = f.second;
    for (usize c = 0; c < 3; ++c) {
      std::vector<usize> d1(sp.connectivity.begin() + (long)(e1 * sp.nsh + c * nsh_sc),
                            sp.connectivity.begin() + (long)(e1 * sp.nsh + (c + 1) * nsh_sc));
      std::vector<usize> d2(sp.connectivity.begin() + (long)(e2 * sp.nsh + c * nsh_sc),
                            sp.connectivity.begin() + (long)(e2 * sp.nsh + (c + 1) * nsh_sc));
      std::sort(d1.begin(), d1.end());
      d1.erase(std::unique(d1.begin(), d1.end()), d1.end());
      std::sort(d2.begin(), d2.end());
      d2.erase(std::unique(d2.begin(), d2.end()), d2.end());
      std::vector<usize> common, only1, only2;
      for (usize x : d1) (std::binary_search(d2.begin(), d2.end(), x) ? common : only1).push_back(x);
      for (usize x : d2)
        if (!std::binary_search(d1.begin(), d1.end(), x)) only2.push_back(x);
      usize n = common.size();
      if (n > 0) {
        double nf = (double)n;
        for (usize m = 0; m < n * n; ++m) {
          usize i = m % n, j = m / n;
          double s = (alpha / nf) * ((i == j ? 1.0 : 0.0) - 1.0 / nf);
          ri.push_back(common[i]);
          ci.push_back(common[i]);
          vi.push_back(s);
        }
      }
      usize np = std::min(only1.size(), only2.size());
      for (usize k = 0; k < np; ++k) {
        usize p1 = only1[k], p2 = only2[k];
        ri.insert(ri.end(), {p1, p2, p1, p2});
        ci.insert(ci.end(), {p1, p2, p2, p1});
        vi.insert(vi.end(), {alpha, alpha, -alpha, -alpha});
      }
    }
  }
}

// immersed.rs:234-370 ghost_penalty: gamma E h^(2p-1) sum_F int_F [d_n^p u].[d_n^p v]
// over the same faces as the legacy term, factored into the 1D jump vector of
// the p-th derivative (normal direction) times the 1D element mass matrices
// (tangential directions); one block per component.
void ghost_penalty(const SpaceBox& sp, const std::vector<int8_t>& status, double gamma, double young,
                   std::vector<usize>& ri, std::vector<usize>& ci, Vec& vi) {
  const usize dim = sp.dim;
  Vec h = sp.element_size();
  double hsum = 0.0;
  for (double v : h) hsum += v;
  const double h_char = hsum / (double)dim;
  const auto& res = sp.nel_dir;
  auto active = [&](usize e) { return status[e] == INSIDE || status[e] == CUT; };
  auto cut = [&](usize e) { return status[e] == CUT; };
  // mass[d][e] (nl x nl), physical
  std::vector<std::vector<Vec>> mass(dim);
  for (usize d = 0; d < dim; ++d) {
    const Space1d& u = sp.univ[d];
    const usize nl = u.nsh();
    for (usize e = 0; e < u.nel; ++e) {
      Vec m(nl * nl, 0.0);
      for (usize q = 0; q < u.nquad; ++q) {
        const double w = u.qw[e * u.nquad + q] * sp.lengths[d];
        const usize base = (e * u.nquad + q) * nl;
        for (usize a = 0; a < nl; ++a)
          for (usize b = 0; b < nl; ++b) m[a * nl + b] += w * u.shape[base + a] * u.shape[base + b];
      }
      mass[d].push_back(m);
    }
  }
  // jumps[d][k] = (indices, values) across break k (elements k-1 | k)
  std::vector<std::vector<std::pair<std::vector<usize>, Vec>>> jumps(dim);
  for (usize d = 0; d < dim; ++d) {
    const Space1d& u = sp.univ[d];
    const usize p = u.degree, nl = u.nsh();
    const double lp = std::pow(sp.lengths[d], (double)p);
    for (usize k = 0; k < u.nel; ++k) {
      if (k == 0) {
        jumps[d].push_back({});
        continue;
      }
      const double x = u.breaks[k];
      const usize lf = u.connectivity[(k - 1) * nl], rf = u.connectivity[k * nl];
      std::vector<Vec> dl, dr;
      basis_funs_ders(u.knots, p, x, p, (long)(lf + p), dl);
      basis_funs_ders(u.knots, p, x, p, (long)(rf + p), dr);
      const usize lo = std::min(lf, rf), hi = std::max(lf + p, rf + p);
      std::vector<usize> idx;
      Vec vals;
      for (usize g = lo; g <= hi; ++g) {
        const double rv = (g >= rf && g <= rf + p) ? dr[p][g - rf] : 0.0;
        const double lv = (g >= lf && g <= lf + p) ? dl[p][g - lf] : 0.0;
        idx.push_back(g);
        vals.push_back((rv - lv) / lp);
      }
      jumps[d].push_back({idx, vals});
    }
  }
  for (usize d = 0; d < dim; ++d) {
    const usize p = sp.degree[d];
    const double coef = gamma * young * std::pow(h_char, (double)(2 * p - 1));
    for (usize e2 = 0; e2 < sp.nel; ++e2) {
      auto s2 = unravel(e2, res);
      if (s2[d] == 0) continue;
      auto s1 = s2;
      s1[d] -= 1;
      usize lin1 = 0, stride = 1;
      for (usize dd = 0; dd < dim; ++dd) {
        lin1 += s1[dd] * stride;
        stride *= res[dd];
      }
      if (!(active(lin1) && active(e2) && (cut(lin1) || cut(e2)))) continue;
      const auto& jidx = jumps[d][s2[d]].first;
      const auto& jval = jumps[d][s2[d]].second;
      std::vector<std::vector<usize>> sets(dim);
      for (usize dd = 0; dd < dim; ++dd) {
        if (dd == d) {
          sets[dd] = jidx;
        } else {
          const Space1d& u = sp.univ[dd];
          sets[dd].assign(u.connectivity.begin() + (long)(s2[dd] * u.nsh()),
                          u.connectivity.begin() + (long)((s2[dd] + 1) * u.nsh()));
        }
      }
      std::vector<usize> sizes;
      usize nloc = 1;
      for (auto& st : sets) {
        sizes.push_back(st.size());
        nloc *= st.size();
      }
      std::vector<usize> gidx(nloc);
      for (usize a = 0; a < nloc; ++a) {
        auto sa = unravel(a, sizes);
        usize g = 0, str = 1;
        for (usize dd = 0; dd < dim; ++dd) {
          g += sets[dd][sa[dd]] * str;
          str *= sp.ndof_dir[dd];
        }
        gidx[a] = g;
      }
      for (usize a = 0; a < nloc; ++a) {
        auto sa = unravel(a, sizes);
        for (usize b = 0; b < nloc; ++b) {
          auto sb = unravel(b, sizes);
          double v = coef * jval[sa[d]] * jval[sb[d]];
          for (usize t = 0; t < dim; ++t) {
            if (t == d) continue;
            const usize nl = sp.nsh_dir[t];
            v *= mass[t][s2[t]][sa[t] * nl + sb[t]];
          }
          if (v == 0.0) continue;
          for (usize c = 0; c < dim; ++c) {
            ri.push_back(c * sp.ndof_sc + gidx[a]);
            ci.push_back(c * sp.ndof_sc + gidx[b]);
            vi.push_back(v);
          }
        }
      }
    }
  }
}

// immersed.rs:450-463 padded_bounds
std::array<std::array<double, 2>, 3> padded_bounds(const TriangleMesh3D& mesh, double padding) {
  std::array<std::array<double, 2>, 3> b;
  for (int d = 0; d < 3; ++d) b[d] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::lowest()};
  for (auto& v : mesh.vertices)
    for (int d = 0; d < 3; ++d) {
      b[d][0] = std::fmin(b[d][0], v[d]);
      b[d][1] = std::fmax(b[d][1], v[d]);
    }
  for (int d = 0; d < 3; ++d) {
    double pad = padding * (b[d][1] - b[d][0]);
    b[d] = {b[d][0] - pad, b[d][1] + pad};
  }
  return b;
}

// immersed.rs:466-538 build_immersed_problem (assembled form)
ImmersedProblem build_immersed_problem(const TriangleMesh3D& mesh, const std::array<usize, 3>& res, usize degree,
                                       double young, double poisson, int stab_kind, double gamma, double emin,
                                       const std::array<usize, 3>& sub, int clamp, double padding) {
  ImmersedProblem pb;
  pb.grid_bounds = padded_bounds(mesh, padding);
  auto& gb = pb.grid_bounds;
  pb.sp = SpaceBox::make({gb[0], gb[1], gb[2]}, {res[0], res[1], res[2]}, degree);
  const SpaceBox& sp = pb.sp;
  ParityRayCaster rc(mesh);
  element_weights([&](const std::array<double, 3>& q) { return rc.is_inside(q); }, gb, res, sub, pb.weights, pb.status);
  double lambda, mu;
  lame(young, poisson, lambda, mu);
  ElementMatrices em = elasticity_element_matrices(sp, lambda, mu);
  std::vector<usize> rows, cols;
  Vec vals;
  if (stab_kind == 1) legacy_stabilization(sp, pb.status, gamma, rows, cols, vals);
  if (stab_kind == 2) ghost_penalty(sp, pb.status, gamma, young, rows, cols, vals);
  usize n_gp = vals.size();
  CsrMatrix stab = CsrMatrix::from_triplets(sp.ndof, sp.ndof, rows, cols, vals);
  pb.stabilization_frobenius = stab.frobenius();
  (void)n_gp;
  usize nsh = sp.nsh;
  rows.reserve(rows.size() + nsh * nsh * sp.nel);
  for (usize e = 0; e < sp.nel; ++e) {
    double s = emin + (1.0 - emin) * pb.weights[e];
    const double* ke = em.of_element(e);
    const usize* ce = &sp.connectivity[e * nsh];
    for (usize a = 0; a < nsh; ++a)
      for (usize b = 0; b < nsh; ++b) {
        rows.push_back(ce[a]);
        cols.push_back(ce[b]);
        vals.push_back(s * ke[a * nsh + b]);
      }
  }
  pb.stiffness = CsrMatrix::from_triplets(sp.ndof, sp.ndof, rows, cols, vals);
  const auto& nd = sp.ndof_dir;
  pb.free_.assign(sp.ndof, 1);
  for (usize i = 0; i < sp.ndof_sc; ++i) {
    usize i1 = i % nd[0], i3 = i / (nd[0] * nd[1]);
    bool cl = (clamp == 0 && i3 == 0) || (clamp == 1 && i3 == nd[2] - 1) || (clamp == 2 && i1 == 0);
    if (cl)
      for (usize c = 0; c < 3; ++c) pb.free_[c * sp.ndof_sc + i] = 0;
  }
  long mid_x = (long)std::round((double)nd[0] / 2.0), mid_y = (long)std::round((double)nd[1] / 2.0);
  std::vector<usize> load;
  for (usize i = 0; i < sp.ndof_sc; ++i) {
    usize i1 = i % nd[0], i2 = (i / nd[0]) % nd[1], i3 = i / (nd[0] * nd[1]);
    if (i3 == nd[2] - 1 && std::labs((long)i1 + 1 - mid_x) <= 2 && std::labs((long)i2 + 1 - mid_y) <= 2)
      load.push_back(i);
  }
  pb.force.assign(sp.ndof, 0.0);
  for (usize i : load) pb.force[2 * sp.ndof_sc + i] = -1.0 / (double)load.size();
  return pb;
}

// immersed.rs:551-570 solve_immersed_problem
ImmersedSolution solve_immersed_problem(const ImmersedProblem& pb, double tol, usize maxit) {
  usize n = pb.sp.ndof;
  const auto& fr = pb.free_;
  Vec b(n);
  for (usize i = 0; i < n; ++i) b[i] = fr[i] ? pb.force[i] : 0.0;
  Vec diag = pb.stiffness.diagonal();
  for (double& d : diag) d = std::fmax(d, 1e-6);
  Vec pm(n, 0.0), tmp(n, 0.0);
  PcgResult r = pcg_solve(maxit, tol, n, b, diag, [&](const double* p, double* ap) {
    for (usize i = 0; i < n; ++i) pm[i] = fr[i] ? p[i] : 0.0;
    pb.stiffness.matvec(pm.data(), tmp.data());
    for (usize i = 0; i < n; ++i) ap[i] = fr[i] ? tmp[i] : 0.0;
  });
  ImmersedSolution sol;
  sol.u.resize(n);
  for (usize i = 0; i < n; ++i) sol.u[i] = fr[i] ? r.x[i] : 0.0;
  double c = 0.0;
  for (usize i = 0; i < n; ++i) c += pb.force[i] * sol.u[i];
  sol.compliance = c;
  sol.iterations = r.iters;
  sol.residual = r.res;
  return sol;
}

InsideFn box_minus_cylinders(const double lo[3], const double hi[3], const std::vector<double>& cyl) {
  std::array<double, 3> l = {lo[0], lo[1], lo[2]}, h = {hi[0], hi[1], hi[2]};
  return [l, h, cyl](const std::array<double, 3>& q) {
    double sdf = -1e300;
    for (int d = 0; d < 3; ++d) sdf = std::fmax(sdf, std::fmax(l[d] - q[d], q[d] - h[d]));
    for (usize c = 0; 3 * c + 2 < cyl.size(); ++c) {
      double dx = q[0] - cyl[3 * c], dy = q[1] - cyl[3 * c + 1];
      sdf = std::fmax(sdf, cyl[3 * c + 2] - std::sqrt(dx * dx + dy * dy));  // > 0 inside a hole
    }
    return sdf < 0.0;
  };
}

// The immersed problem of build_immersed_problem (immersed.rs:466-538) for the
// analytic geometry, on a background grid aligned to the box:
// * weights / status: element_weights with the box-minus-cylinders predicate;
// * scale per mode (see oracle.hpp);
// * fixed DOFs: those whose support holds no cell with s_e > 0 (the clamp is
//   replaced by the penalty below, so no face is clamped);
// * penalty Dirichlet u = 0 on the face x = lo: beta int u.v dA over the face
//   cells whose volume cell has s_e > 0 (dirichlet_penalty_fn,
//   immersed_bc.rs:161-183, with g = 0; beta = factor * E / h_min as
//   surface_boundary_terms, immersed_bc.rs:186-190);
// * consistent traction t on the face x = hi.
SdfImmersed build_sdf_immersed(const double lo[3], const double hi[3], const std::vector<double>& cyl,
                               const std::array<usize, 3>& res, const std::array<usize, 3>& sub, double young,
                               double poisson, double emin, double active_min, double penalty_factor,
                               const double traction[3], int mode) {
  SdfImmersed pb;
  std::array<std::array<double, 2>, 3> gb;
  for (int d = 0; d < 3; ++d) gb[d] = {lo[d], hi[d]};
  pb.sp = SpaceBox::make({gb[0], gb[1], gb[2]}, {res[0], res[1], res[2]}, 2);
  const SpaceBox& sp = pb.sp;
  element_weights(box_minus_cylinders(lo, hi, cyl), gb, res, sub, pb.weights, pb.status);
  pb.scale.assign(sp.nel, 0.0);
  pb.active.assign(sp.nel, 0);
  for (usize e = 0; e < sp.nel; ++e) {
    if (mode == 0) {
      pb.scale[e] = emin + (1.0 - emin) * pb.weights[e];
      pb.active[e] = 1;
    } else if (pb.weights[e] >= active_min) {
      pb.scale[e] = pb.weights[e];
      pb.active[e] = 1;
    }
  }
  double lambda, mu;
  lame(young, poisson, lambda, mu);
  ElementMatrices em = elasticity_element_matrices(sp, lambda, mu);
  std::vector<usize> rows, cols;
  Vec vals;
  const usize nsh = sp.nsh, nsc = sp.ndof_sc;
  pb.free_.assign(sp.ndof, 0);
  for (usize e = 0; e < sp.nel; ++e) {
    if (!(pb.scale[e] > 0.0)) continue;
    const double* ke = em.of_element(e);
    const usize* ce = &sp.connectivity[e * nsh];
    for (usize a = 0; a < nsh; ++a) {
      pb.free_[ce[a]] = 1;
      for (usize b = 0; b < nsh; ++b) {
        rows.push_back(ce[a]);
        cols.push_back(ce[b]);
        vals.push_back(pb.scale[e] * ke[a * nsh + b]);
      }
    }
  }
  double hmin = 1e300;
  for (int d = 0; d < 3; ++d) hmin = std::fmin(hmin, (hi[d] - lo[d]) / (double)res[d]);
  pb.beta = penalty_factor * young / hmin;
  const Space1d& uy = sp.univ[1];
  const Space1d& uz = sp.univ[2];
  const usize nqy = uy.nquad, nqz = uz.nquad, ny = uy.nsh(), nz = uz.nsh();
  for (usize ez = 0; ez < uz.nel; ++ez)
    for (usize ey = 0; ey < uy.nel; ++ey) {
      const usize e0 = sp.nel_dir[0] * (ey + sp.nel_dir[1] * ez);
      if (!(pb.scale[e0] > 0.0)) continue;
      for (usize qz = 0; qz < nqz; ++qz)
        for (usize qy = 0; qy < nqy; ++qy) {
          const double w = uy.qw[ey * nqy + qy] * sp.lengths[1] * uz.qw[ez * nqz + qz] * sp.lengths[2];
          const double* Ny = &uy.shape[(ey * nqy + qy) * ny];
          const double* Nz = &uz.shape[(ez * nqz + qz) * nz];
          for (usize az = 0; az < nz; ++az)
            for (usize ay = 0; ay < ny; ++ay) {
              const usize ga = sp.ndof_dir[0] * (uy.connectivity[ey * ny + ay] + sp.ndof_dir[1] * uz.connectivity[ez * nz + az]);
              for (usize bz = 0; bz < nz; ++bz)
                for (usize by = 0; by < ny; ++by) {
                  const usize gbb = sp.ndof_dir[0] * (uy.connectivity[ey * ny + by] + sp.ndof_dir[1] * uz.connectivity[ez * nz + bz]);
                  const double v = pb.beta * w * (Ny[ay] * Nz[az]) * (Ny[by] * Nz[bz]);
                  for (usize c = 0; c < 3; ++c) {
                    rows.push_back(c * nsc + ga);
                    cols.push_back(c * nsc + gbb);
                    vals.push_back(v);
                  }
                }
            }
        }
    }
  pb.op = CsrMatrix::from_triplets(sp.ndof, sp.ndof, rows, cols, vals);
  pb.force.assign(sp.ndof, 0.0);
  const usize il = sp.ndof_dir[0] - 1;
  for (usize ez = 0; ez < uz.nel; ++ez)
    for (usize ey = 0; ey < uy.nel; ++ey)
      for (usize qz = 0; qz < nqz; ++qz)
        for (usize qy = 0; qy < nqy; ++qy) {
          const double w = uy.qw[ey * nqy + qy] * sp.lengths[1] * uz.qw[ez * nqz + qz] * sp.lengths[2];
          for (usize az = 0; az < nz; ++az)
            for (usize ay = 0; ay < ny; ++ay) {
              const double n = uy.shape[(ey * nqy + qy) * ny + ay] * uz.shape[(ez * nqz + qz) * nz + az];
              const usize g = il + sp.ndof_dir[0] * (uy.connectivity[ey * ny + ay] + sp.ndof_dir[1] * uz.connectivity[ez * nz + az]);
              for (usize c = 0; c < 3; ++c) pb.force[c * nsc + g] += w * n * traction[c];
            }
        }
  for (usize i = 0; i < sp.ndof; ++i)
    if (!pb.free_[i]) pb.force[i] = 0.0;
  return pb;
}

// solve_immersed_problem (immersed.rs:551-570) on the analytic problem
ImmersedSolution solve_sdf_immersed(const SdfImmersed& pb, double tol, usize maxit) {
  const usize n = pb.sp.ndof;
  const auto& fr = pb.free_;
  Vec b(n);
  for (usize i = 0; i < n; ++i) b[i] = fr[i] ? pb.force[i] : 0.0;
  Vec diag = pb.op.diagonal();
  for (double& d : diag) d = std::fmax(d, 1e-6);
  Vec pm(n, 0.0), tmp(n, 0.0);
  PcgResult r = pcg_solve(maxit, tol, n, b, diag, [&](const double* p, double* ap) {
    for (usize i = 0; i < n; ++i) pm[i] = fr[i] ? p[i] : 0.0;
    pb.op.matvec(pm.data(), tmp.data());
    for (usize i = 0; i < n; ++i) ap[i] = fr[i] ? tmp[i] : 0.0;
  });
  ImmersedSolution sol;
  sol.u.resize(n);
  for (usize i = 0; i < n; ++i) sol.u[i] = fr[i] ? r.x[i] : 0.0;
  double c = 0.0;
  for (usize i = 0; i < n; ++i) c += pb.force[i] * sol.u[i];
  sol.compliance = c;
  sol.iterations = r.iters;
  sol.residual = r.res;
  return sol;
}

// topopt_immersed.rs:78-153 (new) and 155-247 (run)
// Filter rows on a design domain (topopt_immersed.rs:117-151): offsets within
// r = rmin * mean(h), taken dk outermost and di innermost, weight r - dist;
// only active in-range neighbours of active cells, inactive rows empty.
std::vector<FilterRow> build_filter_design(const SpaceBox& sp, double rmin, const std::vector<uint8_t>& active) {
  Vec h = sp.element_size();
  const auto& res = sp.nel_dir;
  double r = rmin * (h[0] + h[1] + h[2]) / 3.0;
  long rmax[3];
  for (int d = 0; d < 3; ++d) rmax[d] = (long)std::ceil(r / h[d]);
  struct Off {
    long di, dj, dk;
    double w;
  };
  std::vector<Off> offsets;
  for (long dk = -rmax[2]; dk <= rmax[2]; ++dk)
    for (long dj = -rmax[1]; dj <= rmax[1]; ++dj)
      for (long di = -rmax[0]; di <= rmax[0]; ++di) {
        double dist = std::sqrt(((double)di * h[0]) * ((double)di * h[0]) + ((double)dj * h[1]) * ((double)dj * h[1]) +
                                ((double)dk * h[2]) * ((double)dk * h[2]));
        if (dist <= r) offsets.push_back({di, dj, dk, r - dist});
      }
  std::vector<FilterRow> filter(sp.nel);
  for (usize e = 0; e < sp.nel; ++e) {
    if (!active[e]) continue;
    auto s = unravel(e, res);
    for (auto& o : offsets) {
      long ni = (long)s[0] + o.di, nj = (long)s[1] + o.dj, nk = (long)s[2] + o.dk;
      if (ni < 0 || nj < 0 || nk < 0 || ni >= (long)res[0] || nj >= (long)res[1] || nk >= (long)res[2]) continue;
      usize f = (usize)ni + res[0] * ((usize)nj + res[1] * (usize)nk);
      if (active[f]) {
        filter[e].f.push_back(f);
        filter[e].w.push_back(o.w);
      }
    }
  }
  return filter;
}

ImmersedTopOptHistory immersed_topopt_run(const TriangleMesh3D& brep, const ImmersedTopOptCfg& cfg) {
  auto gb = padded_bounds(brep, 0.05);
  std::array<usize, 3> res{cfg.grid_res[0], cfg.grid_res[1], cfg.grid_res[2]};
  SpaceBox sp = SpaceBox::make({gb[0], gb[1], gb[2]}, {res[0], res[1], res[2]}, 2);
  ParityRayCaster rc(brep);
  Vec weights;
  std::vector<int8_t> status;
  element_weights([&](const std::array<double, 3>& q) { return rc.is_inside(q); }, gb, res, {4, 4, 4}, weights, status);
  usize nel = sp.nel, n = sp.ndof;
  std::vector<uint8_t> active(nel);
  double cad_volume = 0.0;
  for (usize e = 0; e < nel; ++e) {
    active[e] = weights[e] >= 0.05;
    if (active[e]) cad_volume += weights[e];
  }
  double lambda, mu;
  lame(cfg.young, cfg.poisson, lambda, mu);
  ElementMatrices em = elasticity_element_matrices(sp, lambda, mu);
  AssemblyPattern pattern(sp.ndof, sp.connectivity, sp.nsh, sp.nel);
  std::vector<usize> r_, c_;
  Vec v_;
  if (cfg.stab_kind == 1) legacy_stabilization(sp, status, cfg.gamma, r_, c_, v_);
  if (cfg.stab_kind == 2) ghost_penalty(sp, status, cfg.gamma, cfg.young, r_, c_, v_);
  CsrMatrix stab = CsrMatrix::from_triplets(sp.ndof, sp.ndof, r_, c_, v_);
  auto nd = sp.ndof_dir;
  std::vector<uint8_t> fr(n, 1);
  long mid_x = (long)std::round((double)nd[0] / 2.0), mid_y = (long)std::round((double)nd[1] / 2.0);
  std::vector<usize> load;
  for (usize i = 0; i < sp.ndof_sc; ++i) {
    auto s = unravel(i, nd);
    if (s[2] == 0)
      for (usize cc = 0; cc < 3; ++cc) fr[cc * sp.ndof_sc + i] = 0;
    if (s[2] == nd[2] - 1 && std::labs((long)s[0] + 1 - mid_x) <= 2 && std::labs((long)s[1] + 1 - mid_y) <= 2)
      load.push_back(i);
  }
  Vec force(n, 0.0);
  for (usize i : load) force[2 * sp.ndof_sc + i] = -1.0 / (double)load.size();
  // filter on the design domain (offset stencil, topopt_immersed.rs:117-151)
  std::vector<FilterRow> filter = build_filter_design(sp, cfg.rmin, active);
  Vec filter_sum(nel);
  for (usize e = 0; e < nel; ++e) {
    double s = 0.0;
    for (double w : filter[e].w) s += w;
    filter_sum[e] = s < 1e-6 ? 1.0 : s;
  }
  // run
  Vec x(nel), u(n, 0.0);
  for (usize e = 0; e < nel; ++e) x[e] = active[e] ? cfg.volfrac : 0.0;
  ImmersedTopOptHistory hist;
  double nf = 0.0;
  for (usize i = 0; i < n; ++i)
    if (fr[i]) nf += force[i] * force[i];
  nf = std::sqrt(nf);
  for (usize iter = 1; iter <= cfg.max_iter; ++iter) {
    Vec x_old = x;
    Vec scale(nel);
    for (usize e = 0; e < nel; ++e)
      scale[e] = active[e] ? weights[e] * (cfg.emin + (1.0 - cfg.emin) * std::pow(x[e], cfg.penal)) : cfg.emin;
    CsrMatrix k = pattern.assemble(em, scale);
    Vec dk = k.diagonal(), ds = stab.diagonal(), diag(n);
    for (usize i = 0; i < n; ++i) diag[i] = std::fmax(dk[i] + ds[i], 1e-6);
    Vec tmp(n, 0.0), tmp2(n, 0.0), pm(n, 0.0);
    auto op = [&](const double* p, double* ap) {
      for (usize i = 0; i < n; ++i) pm[i] = fr[i] ? p[i] : 0.0;
      k.matvec(pm.data(), tmp.data());
      stab.matvec(pm.data(), tmp2.data());
      for (usize i = 0; i < n; ++i) ap[i] = fr[i] ? tmp[i] + tmp2[i] : 0.0;
    };
    Vec ku(n, 0.0);
    op(u.data(), ku.data());
    Vec b(n);
    for (usize i = 0; i < n; ++i) b[i] = fr[i] ? force[i] - ku[i] : 0.0;
    double rb = 0.0;
    for (double v : b) rb += v * v;
    rb = std::sqrt(rb);
    double tol = rb > 0.0 ? std::fmin(cfg.pcg_tol * nf / rb, 0.5) : 1.0;
    PcgResult pr = pcg_solve(cfg.pcg_maxit, tol, n, b, diag, op);
    for (usize i = 0; i < n; ++i) u[i] = fr[i] ? u[i] + pr.x[i] : 0.0;
    double c = 0.0;
    for (usize i = 0; i < n; ++i) c += force[i] * u[i];
    usize nsh = sp.nsh;
    Vec ee(nel);
    for (usize e = 0; e < nel; ++e) {
      const double* ke = em.of_element(e);
      const usize* ce = &sp.connectivity[e * nsh];
      double s = 0.0;
      for (usize a = 0; a < nsh; ++a) {
        double row = 0.0;
        for (usize bb = 0; bb < nsh; ++bb) row += ke[a * nsh + bb] * u[ce[bb]];
        s += u[ce[a]] * row;
      }
      ee[e] = s;
    }
    Vec dc_raw(nel), dc(nel);
    for (usize e = 0; e < nel; ++e)
      dc_raw[e] = -weights[e] * (cfg.penal * (1.0 - cfg.emin) * std::pow(x[e], cfg.penal - 1.0) * ee[e]);
    for (usize e = 0; e < nel; ++e) {
      double s = 0.0;
      for (usize kk = 0; kk < filter[e].f.size(); ++kk) s += filter[e].w[kk] * x[filter[e].f[kk]] * dc_raw[filter[e].f[kk]];
      dc[e] = s / (filter_sum[e] * std::fmax(x[e], 1e-3));
    }
    std::vector<usize> act;
    for (usize e = 0; e < nel; ++e)
      if (active[e]) act.push_back(e);
    double l1 = 0.0, mx = 0.0;
    for (usize e : act) mx = std::fmax(mx, std::fabs(dc[e]));
    double l2 = 2.0 * mx;
    double target = cfg.volfrac * cad_volume;
    Vec cand(act.size());
    for (usize k2 = 0; k2 < act.size(); ++k2) cand[k2] = x_old[act[k2]];
    while ((l2 - l1) / (l1 + l2 + 1e-10) > 1e-4) {
      double lmid = 0.5 * (l1 + l2);
      for (usize k2 = 0; k2 < act.size(); ++k2) {
        usize e = act[k2];
        double be = std::sqrt(-dc[e] / lmid);
        cand[k2] = std::fmax(std::fmax(std::fmin(std::fmin(x_old[e] * be, x_old[e] + 0.2), 1.0), x_old[e] - 0.2), 0.001);
      }
      double vol = 0.0;
      for (usize k2 = 0; k2 < act.size(); ++k2) vol += weights[act[k2]] * cand[k2];
      if (vol > target)
        l1 = lmid;
      else
        l2 = lmid;
    }
    x.assign(nel, 0.0);
    for (usize k2 = 0; k2 < act.size(); ++k2) x[act[k2]] = cand[k2];
    double change = 0.0;
    for (usize e : act) change = std::fmax(change, std::fabs(x[e] - x_old[e]));
    double vol = 0.0;
    for (usize e : act) vol += weights[e] * x[e];
    vol /= cad_volume;
    hist.compliance.push_back(c);
    hist.volume.push_back(vol);
    hist.change.push_back(change);
    hist.pcg_iterations.push_back(pr.iters);
  }
  hist.density = x;
  return hist;
}

}  // namespace orc
