// Oracle restatement of the surface boundary conditions on triangle
// surfaces (rust/src/immersed_bc.rs) and ImmersedProblem::add_surface_terms
// (rust/src/immersed.rs:541-548).  TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// One deliberate difference: the reference keeps the per-element blocks in a
// HashMap, so into_triplets emits them in an unspecified order; here the
// blocks are kept in a std::map and emitted in increasing key order.  Only
// the order of duplicate summation in the assembled CSR depends on it (the
// reference's from_triplets sorts each row unstably, so its order is
// unspecified as well).
#include <cmath>
#include <map>

#include "oracle.hpp"

namespace orc {

// immersed_bc.rs:15-22: Dunavant degree-4 rule, barycentric points and weights (sum 1)
static const double kTriRule[6][4] = {
    {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.223381589678011},
    {0.445948490915965, 0.108103018168070, 0.445948490915965, 0.223381589678011},
    {0.445948490915965, 0.445948490915965, 0.108103018168070, 0.223381589678011},
    {0.816847572980459, 0.091576213509771, 0.091576213509771, 0.109951743655322},
    {0.091576213509771, 0.816847572980459, 0.091576213509771, 0.109951743655322},
    {0.091576213509771, 0.091576213509771, 0.816847572980459, 0.109951743655322},
};

// immersed_bc.rs:25-50 scalar_basis_at: (global index, value) of the nsh_sc
// nonzero scalar basis functions at x, the parameter clamped to [0, 1].
std::vector<std::pair<usize, double>> scalar_basis_at(const SpaceBox& sp, const std::array<double, 3>& x) {
  std::vector<usize> first(sp.dim);
  std::vector<Vec> vals(sp.dim);
  for (usize d = 0; d < sp.dim; ++d) {
    const Space1d& u = sp.univ[d];
    double xi = (x[d] - sp.lo[d]) / sp.lengths[d];
    xi = std::min(std::max(xi, 0.0), 1.0);
    Vec db;
    first[d] = basis_funs(u.knots, u.degree, xi, vals[d], db);
  }
  std::vector<std::pair<usize, double>> out;
  out.reserve(sp.nsh_sc);
  for (usize a = 0; a < sp.nsh_sc; ++a) {
    auto s = unravel(a, sp.nsh_dir);
    double w = 1.0;
    usize g = 0, stride = 1;
    for (usize d = 0; d < sp.dim; ++d) {
      w *= vals[d][s[d]];
      g += (first[d] + s[d]) * stride;
      stride *= sp.ndof_dir[d];
    }
    out.emplace_back(g, w);
  }
  return out;
}

// immersed_bc.rs:128-148 push_triangle_rule
static void push_triangle_rule(std::vector<SurfacePoint>& out, const std::array<double, 3>& a,
                               const std::array<double, 3>& b, const std::array<double, 3>& c) {
  const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (nn == 0.0) return;
  const double area = 0.5 * nn;
  const std::array<double, 3> unit = {n[0] / nn, n[1] / nn, n[2] / nn};
  for (const auto& r : kTriRule) {
    SurfacePoint q;
    for (int d = 0; d < 3; ++d) q.x[d] = r[0] * a[d] + r[1] * b[d] + r[2] * c[d];
    q.w = r[3] * area;
    q.n = unit;
    out.push_back(q);
  }
}

// immersed_bc.rs:109-126 surface_quadrature: midpoint refinement until every
// edge is at most max_edge, depth first with the reference's stack order.
std::vector<SurfacePoint> surface_quadrature(const TriangleMesh3D& mesh, const std::vector<usize>& tris,
                                             double max_edge) {
  using P = std::array<double, 3>;
  std::vector<SurfacePoint> out;
  out.reserve(tris.size() * 6);
  auto len = [](const P& p, const P& q) {
    return std::sqrt(std::pow(p[0] - q[0], 2) + std::pow(p[1] - q[1], 2) + std::pow(p[2] - q[2], 2));
  };
  auto mid = [](const P& p, const P& q) {
    return P{0.5 * (p[0] + q[0]), 0.5 * (p[1] + q[1]), 0.5 * (p[2] + q[2])};
  };
  for (usize t : tris) {
    const auto& tri = mesh.triangles[t];
    std::vector<std::array<P, 3>> stack{{mesh.vertices[tri[0]], mesh.vertices[tri[1]], mesh.vertices[tri[2]]}};
    while (!stack.empty()) {
      const auto abc = stack.back();
      stack.pop_back();
      const P &a = abc[0], &b = abc[1], &c = abc[2];
      if (std::max(std::max(len(a, b), len(b, c)), len(c, a)) > max_edge) {
        const P ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
        stack.push_back({a, ab, ca});
        stack.push_back({ab, b, bc});
        stack.push_back({ca, bc, c});
        stack.push_back({ab, bc, ca});
        continue;
      }
      push_triangle_rule(out, a, b, c);
    }
  }
  return out;
}

// immersed_bc.rs:55-100 SurfaceAccumulator: one dense (3n x 3n) block per
// background element, keyed by the first basis function of the point.
void SurfaceAccumulator::add(const std::vector<std::pair<usize, double>>& nz, const double ks[3][3], double w) {
  const usize n = nz.size();
  const usize key = nz[0].first;
  auto it = blocks.find(key);
  if (it == blocks.end()) {
    std::vector<usize> idx(n);
    for (usize a = 0; a < n; ++a) idx[a] = nz[a].first;
    it = blocks.emplace(key, std::make_pair(idx, Vec(9 * n * n, 0.0))).first;
  }
  Vec& m = it->second.second;
  for (int ci = 0; ci < 3; ++ci)
    for (int cj = 0; cj < 3; ++cj) {
      const double k = ks[ci][cj] * w;
      if (k == 0.0) continue;
      for (usize a = 0; a < n; ++a) {
        const double ka = k * nz[a].second;
        const usize row = (ci * n + a) * 3 * n + cj * n;
        for (usize b = 0; b < n; ++b) m[row + b] += ka * nz[b].second;
      }
    }
}

void SurfaceAccumulator::into_triplets(usize nsc, SurfaceTerms& st) const {
  for (const auto& kv : blocks) {
    const auto& idx = kv.second.first;
    const Vec& m = kv.second.second;
    const usize n = idx.size();
    for (usize ci = 0; ci < 3; ++ci)
      for (usize a = 0; a < n; ++a)
        for (usize cj = 0; cj < 3; ++cj)
          for (usize b = 0; b < n; ++b) {
            const double v = m[(ci * n + a) * 3 * n + cj * n + b];
            if (v != 0.0) {
              st.rows.push_back(ci * nsc + idx[a]);
              st.cols.push_back(cj * nsc + idx[b]);
              st.vals.push_back(v);
            }
          }
  }
}

static double h_min_of(const SpaceBox& sp) {
  double h = 1e300;
  for (double v : sp.element_size()) h = std::min(h, v);
  return h;
}

// The body of immersed_bc.rs:186-247 surface_boundary_terms for one labelled
// condition on the triangles bc.tris (the caller's face), into acc and st.force;
// beta is the Dirichlet penalty factor * E / h_min.
void surface_add_condition(const SpaceBox& sp, const TriangleMesh3D& mesh, const SurfaceBc& bc, double beta,
                           SurfaceAccumulator& acc, SurfaceTerms& st) {
  const usize nsc = sp.ndof_sc;
  const double hmin = h_min_of(sp);
  for (const auto& q : surface_quadrature(mesh, bc.tris, 0.25 * hmin)) {
    const auto nz = scalar_basis_at(sp, q.x);
    const double w = q.w;
    switch (bc.kind) {
      case SurfaceBc::DIRICHLET: {
        double ks[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int c = 0; c < 3; ++c) {
          if (!bc.components[c]) continue;
          ks[c][c] = beta;
          for (const auto& p : nz) st.force[c * nsc + p.first] += beta * w * p.second * bc.values[c];
        }
        acc.add(nz, ks, w);
        break;
      }
      case SurfaceBc::TRACTION:
        for (const auto& p : nz)
          for (int c = 0; c < 3; ++c) st.force[c * nsc + p.first] += w * p.second * bc.values[c];
        break;
      case SurfaceBc::PRESSURE:
        for (const auto& p : nz)
          for (int c = 0; c < 3; ++c) st.force[c * nsc + p.first] -= w * p.second * bc.pressure * q.n[c];
        break;
      case SurfaceBc::ROBIN: {
        double ks[3][3];
        for (int ci = 0; ci < 3; ++ci)
          for (int cj = 0; cj < 3; ++cj) {
            const double delta = ci == cj ? 1.0 : 0.0;
            ks[ci][cj] = bc.kn * q.n[ci] * q.n[cj] + bc.kt * (delta - q.n[ci] * q.n[cj]);
          }
        acc.add(nz, ks, w);
        break;
      }
    }
  }
}

// immersed_bc.rs:186-247 surface_boundary_terms over an explicit list of
// labelled conditions (the face lookup by name of CadAssembly is the caller's).
SurfaceTerms surface_boundary_terms(const SpaceBox& sp, const TriangleMesh3D& mesh, const std::vector<SurfaceBc>& bcs,
                                    double penalty_factor, double young) {
  SurfaceTerms st;
  st.force.assign(sp.ndof, 0.0);
  const double beta = penalty_factor * young / h_min_of(sp);
  SurfaceAccumulator acc;
  for (const auto& bc : bcs) surface_add_condition(sp, mesh, bc, beta, acc, st);
  acc.into_triplets(sp.ndof_sc, st);
  return st;
}

// immersed_bc.rs:161-182 dirichlet_penalty_fn: beta int (u - g(x)).v on tris.
SurfaceTerms dirichlet_penalty_fn(const SpaceBox& sp, const TriangleMesh3D& mesh, const std::vector<usize>& tris,
                                  double beta, const std::function<std::array<double, 3>(const std::array<double, 3>&)>& g) {
  const usize nsc = sp.ndof_sc;
  SurfaceTerms st;
  st.force.assign(sp.ndof, 0.0);
  SurfaceAccumulator acc;
  const double ks[3][3] = {{beta, 0.0, 0.0}, {0.0, beta, 0.0}, {0.0, 0.0, beta}};
  for (const auto& q : surface_quadrature(mesh, tris, 0.25 * h_min_of(sp))) {
    const auto nz = scalar_basis_at(sp, q.x);
    const auto gx = g(q.x);
    for (int c = 0; c < 3; ++c)
      for (const auto& p : nz) st.force[c * nsc + p.first] += beta * q.w * p.second * gx[c];
    acc.add(nz, ks, q.w);
  }
  acc.into_triplets(nsc, st);
  return st;
}

// immersed.rs:541-548 ImmersedProblem::add_surface_terms
void add_surface_terms(ImmersedProblem& pb, const SurfaceTerms& st) {
  pb.stiffness = pb.stiffness.plus_triplets(st.rows, st.cols, st.vals);
  for (usize i = 0; i < pb.force.size(); ++i) pb.force[i] += st.force[i];
}

}  // namespace orc
