// Oracle: B-spline spaces, quadrature, exact elasticity element matrices and
// field evaluation.  Restates rust/src/iga/{quadrature,basis,space,elasticity,
// fields}.rs of the reference; each function cites the lines it follows.
// Test infrastructure only (see oracle.hpp).
#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// quadrature.rs:35-51 legendre_with_derivative
static void legendre_with_derivative(usize n, double z, double& p_out, double& dp_out) {
  double p0 = 1.0, p1 = z;
  if (n == 0) {
    p_out = 1.0;
    dp_out = 0.0;
    return;
  }
  for (usize k = 2; k <= n; ++k) {
    double kf = (double)k;
    double p2 = ((2.0 * kf - 1.0) * z * p1 - (kf - 1.0) * p0) / kf;
    p0 = p1;
    p1 = p2;
  }
  double dp = (double)n * (z * p1 - p0) / (z * z - 1.0);
  p_out = p1;
  dp_out = dp;
}

// quadrature.rs:5-31 gauss_legendre (Tricomi guess + Newton, ascending nodes)
void gauss_legendre(usize n, Vec& x, Vec& w) {
  if (n < 1) throw std::invalid_argument("gauss_legendre: n must be >= 1");
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (usize i = 0; i < n; ++i) {
    double z = std::cos(M_PI * ((double)i + 0.75) / ((double)n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p, d;
      legendre_with_derivative(n, z, p, d);
      dp = d;
      double dz = p / d;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p, d;
    legendre_with_derivative(n, z, p, d);
    if (d != 0.0) dp = d;
    x[i] = -z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// quadrature.rs:51-64 open_knots
Vec open_knots(usize nel, usize degree, long regularity) {
  if (nel < 1) throw std::invalid_argument("open_knots: nel must be >= 1");
  if (!(regularity >= -1 && regularity < (long)degree)) throw std::invalid_argument("open_knots: regularity");
  usize mult = (usize)((long)degree - regularity);
  Vec knots(degree + 1, 0.0);
  for (usize i = 1; i < nel; ++i) {
    double b = (double)i / (double)nel;
    for (usize m = 0; m < mult; ++m) knots.push_back(b);
  }
  for (usize m = 0; m < degree + 1; ++m) knots.push_back(1.0);
  return knots;
}

// basis.rs:5-7
usize num_basis(const Vec& knots, usize degree) { return knots.size() - degree - 1; }

// basis.rs:11-15 find_span (partition_point + clamp)
usize find_span(const Vec& knots, usize degree, double x) {
  usize ndof = num_basis(knots, degree);
  usize count = (usize)(std::upper_bound(knots.begin(), knots.end(), x) - knots.begin());
  usize s = count == 0 ? 0 : count - 1;
  return std::min(std::max(s, degree), ndof - 1);
}

// basis.rs:20-70 basis_funs (Cox-de Boor + first derivative)
usize basis_funs(const Vec& knots, usize degree, double x, Vec& b, Vec& db) {
  usize p = degree;
  usize s = find_span(knots, p, x);
  b.assign(1, 1.0);
  Vec b_prev;
  for (usize k = 1; k <= p; ++k) {
    if (k == p) b_prev = b;
    Vec bn(k + 1, 0.0);
    for (usize r = 0; r <= k; ++r) {
      usize i = s + r - k;
      double val = 0.0;
      if (r > 0) {
        double d = knots[i + k] - knots[i];
        if (d > 0.0) val += (x - knots[i]) / d * b[r - 1];
      }
      if (r < k) {
        double d = knots[i + k + 1] - knots[i + 1];
        if (d > 0.0) val += (knots[i + k + 1] - x) / d * b[r];
      }
      bn[r] = val;
    }
    b = bn;
  }
  db.assign(p + 1, 0.0);
  if (p > 0) {
    for (usize r = 0; r <= p; ++r) {
      usize i = s + r - p;
      double val = 0.0;
      if (r > 0) {
        double d = knots[i + p] - knots[i];
        if (d > 0.0) val += b_prev[r - 1] / d;
      }
      if (r < p) {
        double d = knots[i + p + 1] - knots[i + 1];
        if (d > 0.0) val -= b_prev[r] / d;
      }
      db[r] = (double)p * val;
    }
  }
  return s - p;
}

// basis.rs:76-136 basis_funs_ders (NURBS Book A2.3)
usize basis_funs_ders(const Vec& knots, usize degree, double x, usize n, long span, std::vector<Vec>& ders) {
  usize p = degree;
  usize s = span >= 0 ? (usize)span : find_span(knots, p, x);
  std::vector<Vec> ndu(p + 1, Vec(p + 1, 0.0));
  Vec left(p + 1, 0.0), right(p + 1, 0.0);
  ndu[0][0] = 1.0;
  for (usize j = 1; j <= p; ++j) {
    left[j] = x - knots[s + 1 - j];
    right[j] = knots[s + j] - x;
    double saved = 0.0;
    for (usize r = 0; r < j; ++r) {
      ndu[j][r] = right[r + 1] + left[j - r];
      double temp = ndu[r][j - 1] / ndu[j][r];
      ndu[r][j] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    ndu[j][j] = saved;
  }
  ders.assign(n + 1, Vec(p + 1, 0.0));
  for (usize j = 0; j <= p; ++j) ders[0][j] = ndu[j][p];
  std::vector<Vec> a(2, Vec(p + 1, 0.0));
  for (usize r = 0; r <= p; ++r) {
    usize s1 = 0, s2 = 1;
    a[0][0] = 1.0;
    for (usize k = 1; k <= std::min(n, p); ++k) {
      double d = 0.0;
      long rk = (long)r - (long)k;
      usize pk = p - k;
      if (r >= k) {
        a[s2][0] = a[s1][0] / ndu[pk + 1][(usize)rk];
        d = a[s2][0] * ndu[(usize)rk][pk];
      }
      usize j1 = rk >= -1 ? 1 : (usize)(-rk);
      usize j2 = ((long)r - 1) <= (long)pk ? k - 1 : p - r;
      for (usize j = j1; j <= j2; ++j) {
        usize idx = (usize)(rk + (long)j);
        a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][idx];
        d += a[s2][j] * ndu[idx][pk];
      }
      if (r <= pk) {
        a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
        d += a[s2][k] * ndu[r][pk];
      }
      ders[k][r] = d;
      std::swap(s1, s2);
    }
  }
  double fac = (double)p;
  for (usize k = 1; k <= std::min(n, p); ++k) {
    for (usize j = 0; j <= p; ++j) ders[k][j] *= fac;
    fac *= (double)(p - k);
  }
  return s - p;
}

// basis.rs:140-152 basis_dense
void basis_dense(const Vec& knots, usize degree, const Vec& xs, Vec& n, Vec& dn) {
  usize ndof = num_basis(knots, degree);
  n.assign(xs.size() * ndof, 0.0);
  dn.assign(xs.size() * ndof, 0.0);
  Vec v, d;
  for (usize q = 0; q < xs.size(); ++q) {
    usize first = basis_funs(knots, degree, xs[q], v, d);
    for (usize r = 0; r <= degree; ++r) {
      n[q * ndof + first + r] = v[r];
      dn[q * ndof + first + r] = d[r];
    }
  }
}

// space.rs:36-72 Space1d::new
Space1d::Space1d(Vec knots_, usize degree_, usize nquad_) : knots(std::move(knots_)), degree(degree_), nquad(nquad_) {
  usize p = degree;
  for (double k : knots)
    if (breaks.empty() || k != breaks.back()) breaks.push_back(k);
  nel = breaks.size() - 1;
  ndof = num_basis(knots, p);
  supp.assign(ndof, {});
  Vec xg, wg;
  gauss_legendre(nquad, xg, wg);
  Vec v, d;
  for (usize e = 0; e < nel; ++e) {
    double a = breaks[e], b = breaks[e + 1];
    usize first = basis_funs(knots, p, 0.5 * (a + b), v, d);
    for (usize r = 0; r <= p; ++r) {
      connectivity.push_back(first + r);
      supp[first + r].push_back(e);
    }
    double h = b - a;
    for (usize q = 0; q < nquad; ++q) {
      double x = a + 0.5 * (xg[q] + 1.0) * h;
      qn.push_back(x);
      qw.push_back(0.5 * wg[q] * h);
      basis_funs(knots, p, x, v, d);
      shape.insert(shape.end(), v.begin(), v.end());
      dshape.insert(dshape.end(), d.begin(), d.end());
    }
  }
}

// space.rs:204-211 unravel (direction 0 fastest)
std::vector<usize> unravel(usize idx, const std::vector<usize>& dims) {
  std::vector<usize> out;
  out.reserve(dims.size());
  for (usize n : dims) {
    out.push_back(idx % n);
    idx /= n;
  }
  return out;
}

// space.rs:107-160 SpaceBox::new / with_params (maximal regularity, p+1 Gauss points)
SpaceBox SpaceBox::make(const std::vector<std::array<double, 2>>& bounds, const std::vector<usize>& nsub, usize deg) {
  SpaceBox sb;
  usize dim = bounds.size();
  if (!(dim == 2 || dim == 3) || nsub.size() != dim) throw std::invalid_argument("SpaceBox: dim must be 2 or 3");
  sb.dim = dim;
  for (auto& b : bounds) {
    sb.lo.push_back(b[0]);
    sb.lengths.push_back(b[1] - b[0]);
  }
  sb.degree.assign(dim, deg);
  sb.nquad.assign(dim, deg + 1);
  for (usize d = 0; d < dim; ++d) sb.univ.emplace_back(open_knots(nsub[d], deg, (long)deg - 1), deg, deg + 1);
  sb.nel = 1;
  sb.ndof_sc = 1;
  sb.nsh_sc = 1;
  sb.nqn = 1;
  for (usize d = 0; d < dim; ++d) {
    sb.nel_dir.push_back(sb.univ[d].nel);
    sb.nel *= sb.univ[d].nel;
    sb.ndof_dir.push_back(sb.univ[d].ndof);
    sb.ndof_sc *= sb.univ[d].ndof;
    sb.nsh_dir.push_back(deg + 1);
    sb.nsh_sc *= deg + 1;
    sb.nqn *= deg + 1;
  }
  sb.nsh = dim * sb.nsh_sc;
  sb.ndof = dim * sb.ndof_sc;
  sb.conn_sc.reserve(sb.nel * sb.nsh_sc);
  for (usize e = 0; e < sb.nel; ++e) {
    auto esub = unravel(e, sb.nel_dir);
    for (usize a = 0; a < sb.nsh_sc; ++a) {
      auto asub = unravel(a, sb.nsh_dir);
      usize g = 0, stride = 1;
      for (usize d = 0; d < dim; ++d) {
        const Space1d& u = sb.univ[d];
        g += u.connectivity[esub[d] * u.nsh() + asub[d]] * stride;
        stride *= sb.ndof_dir[d];
      }
      sb.conn_sc.push_back(g);
    }
  }
  sb.connectivity.reserve(sb.nel * sb.nsh);
  for (usize e = 0; e < sb.nel; ++e)
    for (usize c = 0; c < dim; ++c)
      for (usize a = 0; a < sb.nsh_sc; ++a) sb.connectivity.push_back(c * sb.ndof_sc + sb.conn_sc[e * sb.nsh_sc + a]);
  return sb;
}

// space.rs:163-165
Vec SpaceBox::element_size() const {
  Vec h;
  for (usize d = 0; d < dim; ++d) h.push_back(lengths[d] / (double)nel_dir[d]);
  return h;
}

// space.rs:168-172
std::vector<usize> SpaceBox::boundary_dofs_sc(usize side) const {
  usize d = side / 2;
  usize target = side % 2 == 0 ? 0 : ndof_dir[d] - 1;
  std::vector<usize> out;
  for (usize i = 0; i < ndof_sc; ++i)
    if (unravel(i, ndof_dir)[d] == target) out.push_back(i);
  return out;
}

// space.rs:175-178
std::vector<usize> SpaceBox::boundary_dofs(usize side) const {
  auto sc = boundary_dofs_sc(side);
  std::vector<usize> out;
  for (usize c = 0; c < dim; ++c)
    for (usize i : sc) out.push_back(c * ndof_sc + i);
  return out;
}

// space.rs:181-200 greville_points
std::vector<std::array<double, 3>> SpaceBox::greville_points() const {
  std::vector<Vec> g1(dim);
  for (usize d = 0; d < dim; ++d) {
    const Space1d& u = univ[d];
    usize p = u.degree;
    for (usize i = 0; i < u.ndof; ++i) {
      double m;
      if (p == 0) {
        m = 0.5 * (u.knots[i] + u.knots[i + 1]);
      } else {
        double s = 0.0;
        for (usize k = i + 1; k <= i + p; ++k) s += u.knots[k];
        m = s / (double)p;
      }
      g1[d].push_back(lo[d] + lengths[d] * m);
    }
  }
  std::vector<std::array<double, 3>> out(ndof_sc, {0.0, 0.0, 0.0});
  for (usize i = 0; i < ndof_sc; ++i) {
    auto s = unravel(i, ndof_dir);
    for (usize d = 0; d < dim; ++d) out[i][d] = g1[d][s[d]];
  }
  return out;
}

// elasticity.rs:153-170 kron (index of b fastest)
static Vec kron(const Vec& a, usize na, const Vec& b, usize nb) {
  usize n = na * nb;
  Vec out(n * n, 0.0);
  for (usize ia = 0; ia < na; ++ia)
    for (usize ja = 0; ja < na; ++ja) {
      double av = a[ia * na + ja];
      if (av == 0.0) continue;
      for (usize ib = 0; ib < nb; ++ib)
        for (usize jb = 0; jb < nb; ++jb) out[(ia * nb + ib) * n + ja * nb + jb] = av * b[ib * nb + jb];
    }
  return out;
}

// elasticity.rs:30-148 elasticity_element_matrices
ElementMatrices elasticity_element_matrices(const SpaceBox& sp, double lambda, double mu) {
  usize dim = sp.dim, nsh_sc = sp.nsh_sc;
  // mats[d][t][k], k = 00, 10, 01, 11 (first index = test)
  std::vector<std::vector<std::array<Vec, 4>>> mats(dim);
  std::vector<std::vector<usize>> tid_dir(dim);
  for (usize d = 0; d < dim; ++d) {
    const Space1d& u = sp.univ[d];
    double ld = sp.lengths[d];
    usize nl = u.nsh();
    std::map<std::vector<long long>, usize> groups;
    std::vector<std::array<Vec, 4>> per_elem;
    double scale = 0.0;
    for (usize e = 0; e < u.nel; ++e) {
      std::array<Vec, 4> m;
      for (auto& mk : m) mk.assign(nl * nl, 0.0);
      for (usize q = 0; q < u.nquad; ++q) {
        double w = u.qw[e * u.nquad + q] * ld;
        usize base = (e * u.nquad + q) * nl;
        for (usize a = 0; a < nl; ++a) {
          double na = u.shape[base + a];
          double da = u.dshape[base + a] / ld;
          for (usize b = 0; b < nl; ++b) {
            double nb = u.shape[base + b];
            double db = u.dshape[base + b] / ld;
            m[0][a * nl + b] += w * na * nb;
            m[1][a * nl + b] += w * da * nb;
            m[2][a * nl + b] += w * na * db;
            m[3][a * nl + b] += w * da * db;
          }
        }
      }
      for (auto& mk : m)
        for (double v : mk) scale = std::max(scale, std::fabs(v));
      per_elem.push_back(m);
    }
    std::vector<usize> tids;
    for (auto& m : per_elem) {
      std::vector<long long> key;
      for (auto& mk : m)
        for (double v : mk) key.push_back((long long)std::round(v / scale * 1e12));
      usize next = mats[d].size();
      auto it = groups.find(key);
      usize t;
      if (it == groups.end()) {
        groups.emplace(key, next);
        t = next;
        mats[d].push_back(m);
      } else {
        t = it->second;
      }
      tids.push_back(t);
    }
    tid_dir[d] = tids;
  }
  std::map<std::vector<usize>, usize> combos;
  std::vector<std::vector<usize>> type_list;
  ElementMatrices em;
  em.type_id.reserve(sp.nel);
  for (usize e = 0; e < sp.nel; ++e) {
    auto esub = unravel(e, sp.nel_dir);
    std::vector<usize> key;
    for (usize d = 0; d < dim; ++d) key.push_back(tid_dir[d][esub[d]]);
    usize next = type_list.size();
    auto it = combos.find(key);
    usize t;
    if (it == combos.end()) {
      combos.emplace(key, next);
      type_list.push_back(key);
      t = next;
    } else {
      t = it->second;
    }
    em.type_id.push_back(t);
  }
  usize nsh = sp.nsh;
  em.nsh = nsh;
  for (auto& key : type_list) {
    std::vector<std::vector<Vec>> g(dim, std::vector<Vec>(dim));
    for (usize i = 0; i < dim; ++i)
      for (usize j = 0; j < dim; ++j) {
        Vec gij(1, 1.0);
        usize n = 1;
        for (usize d = 0; d < dim; ++d) {
          int k = (d == i && d == j) ? 3 : (d == i) ? 1 : (d == j) ? 2 : 0;
          const Vec& m = mats[d][key[d]][k];
          usize nl = sp.nsh_dir[d];
          gij = kron(m, nl, gij, n);
          n *= nl;
        }
        g[i][j] = gij;
      }
    Vec lap(nsh_sc * nsh_sc, 0.0);
    for (usize k = 0; k < dim; ++k)
      for (usize l = 0; l < lap.size(); ++l) lap[l] += g[k][k][l];
    Vec ke(nsh * nsh, 0.0);
    for (usize ci = 0; ci < dim; ++ci)
      for (usize cj = 0; cj < dim; ++cj)
        for (usize a = 0; a < nsh_sc; ++a)
          for (usize b = 0; b < nsh_sc; ++b) {
            double v = lambda * g[ci][cj][a * nsh_sc + b] + mu * g[cj][ci][a * nsh_sc + b];
            if (ci == cj) v += mu * lap[a * nsh_sc + b];
            ke[(ci * nsh_sc + a) * nsh + cj * nsh_sc + b] = v;
          }
    em.types.push_back(std::move(ke));
  }
  return em;
}

// elasticity.rs:173-177 lame
void lame(double young, double poisson, double& lambda, double& mu) {
  lambda = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
  mu = young / (2.0 * (1.0 + poisson));
}

// fields.rs:9-27 element_quad_points
void element_quad_points(const SpaceBox& sp, usize e, std::vector<std::array<double, 3>>& x, Vec& w) {
  auto esub = unravel(e, sp.nel_dir);
  x.clear();
  w.clear();
  for (usize q = 0; q < sp.nqn; ++q) {
    auto qsub = unravel(q, sp.nquad);
    std::array<double, 3> xq{0.0, 0.0, 0.0};
    double wq = 1.0;
    for (usize d = 0; d < sp.dim; ++d) {
      const Space1d& u = sp.univ[d];
      usize k = esub[d] * u.nquad + qsub[d];
      xq[d] = sp.lo[d] + sp.lengths[d] * u.qn[k];
      wq *= sp.lengths[d] * u.qw[k];
    }
    x.push_back(xq);
    w.push_back(wq);
  }
}

// fields.rs:31-53 element_shape_functions
void element_shape_functions(const SpaceBox& sp, usize e, Vec& n, Vec& dn) {
  usize dim = sp.dim;
  auto esub = unravel(e, sp.nel_dir);
  n.assign(sp.nqn * sp.nsh_sc, 1.0);
  dn.assign(sp.nqn * sp.nsh_sc * dim, 1.0);
  for (usize q = 0; q < sp.nqn; ++q) {
    auto qsub = unravel(q, sp.nquad);
    for (usize a = 0; a < sp.nsh_sc; ++a) {
      auto asub = unravel(a, sp.nsh_dir);
      for (usize d = 0; d < dim; ++d) {
        const Space1d& u = sp.univ[d];
        usize idx = (esub[d] * u.nquad + qsub[d]) * u.nsh() + asub[d];
        double v = u.shape[idx];
        double dv = u.dshape[idx] / sp.lengths[d];
        n[q * sp.nsh_sc + a] *= v;
        for (usize k = 0; k < dim; ++k) dn[(q * sp.nsh_sc + a) * dim + k] *= (k == d ? dv : v);
      }
    }
  }
}

// fields.rs:56-76 load_vector
Vec load_vector(const SpaceBox& sp, const std::function<std::array<double, 3>(const std::array<double, 3>&)>& f) {
  Vec out(sp.ndof, 0.0);
  std::vector<std::array<double, 3>> x;
  Vec w, n, dn;
  for (usize e = 0; e < sp.nel; ++e) {
    element_quad_points(sp, e, x, w);
    element_shape_functions(sp, e, n, dn);
    for (usize q = 0; q < sp.nqn; ++q) {
      auto fq = f(x[q]);
      for (usize a = 0; a < sp.nsh_sc; ++a) {
        usize g = sp.conn_sc[e * sp.nsh_sc + a];
        double wn = w[q] * n[q * sp.nsh_sc + a];
        for (usize c = 0; c < sp.dim; ++c) out[c * sp.ndof_sc + g] += wn * fq[c];
      }
    }
  }
  return out;
}

// fields.rs:97-115 element_vector_gradients
std::vector<std::array<std::array<double, 3>, 3>> element_vector_gradients(const SpaceBox& sp, const Vec& u, usize e) {
  usize dim = sp.dim;
  Vec n, dn;
  element_shape_functions(sp, e, n, dn);
  std::vector<std::array<std::array<double, 3>, 3>> out(sp.nqn);
  for (usize q = 0; q < sp.nqn; ++q) {
    std::array<std::array<double, 3>, 3> g{};
    for (usize a = 0; a < sp.nsh_sc; ++a) {
      usize gi = sp.conn_sc[e * sp.nsh_sc + a];
      for (usize c = 0; c < dim; ++c) {
        double uc = u[c * sp.ndof_sc + gi];
        for (usize k = 0; k < dim; ++k) g[c][k] += dn[(q * sp.nsh_sc + a) * dim + k] * uc;
      }
    }
    out[q] = g;
  }
  return out;
}

// fields.rs:118-145 eval_vector_at
std::array<double, 3> eval_vector_at(const SpaceBox& sp, const Vec& u, const std::array<double, 3>& x) {
  usize dim = sp.dim;
  std::vector<usize> first(dim);
  std::vector<Vec> vals(dim);
  Vec v, d;
  for (usize dd = 0; dd < dim; ++dd) {
    const Space1d& u1 = sp.univ[dd];
    double xi = std::min(std::max((x[dd] - sp.lo[dd]) / sp.lengths[dd], 0.0), 1.0);
    first[dd] = basis_funs(u1.knots, u1.degree, xi, v, d);
    vals[dd] = v;
  }
  std::array<double, 3> out{0.0, 0.0, 0.0};
  for (usize a = 0; a < sp.nsh_sc; ++a) {
    auto asub = unravel(a, sp.nsh_dir);
    double w = 1.0;
    usize g = 0, stride = 1;
    for (usize dd = 0; dd < dim; ++dd) {
      w *= vals[dd][asub[dd]];
      g += (first[dd] + asub[dd]) * stride;
      stride *= sp.ndof_dir[dd];
    }
    for (usize c = 0; c < dim; ++c) out[c] += w * u[c * sp.ndof_sc + g];
  }
  return out;
}

}  // namespace orc
