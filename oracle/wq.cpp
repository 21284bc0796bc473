// Oracle: weighted-quadrature (WQ) rules and WQ stiffness formation with sum
// factorisation.  Restates rust/src/wq.rs of the reference.  The WQ diagonal
// the GPU forms is the (c N + ii, c N + ii) entries of wq_stiffness_layout.
// Test infrastructure only (see oracle.hpp).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// wq.rs:26-31 linspace
static Vec wq_linspace(double a, double b, usize n) {
  Vec v(n);
  for (usize i = 0; i < n; ++i) v[i] = a + (double)i * (b - a) / (double)(n - 1);
  v[0] = a;
  v[n - 1] = b;
  return v;
}

// wq.rs:34-64 min_norm (normal equations, Gaussian elimination with partial pivoting)
static Vec min_norm(const Vec& a, usize m, usize n, const Vec& rhs) {
  Vec g(m * m);
  for (usize i = 0; i < m; ++i)
    for (usize j = 0; j < m; ++j) {
      double s = 0.0;
      for (usize k = 0; k < n; ++k) s += a[i * n + k] * a[j * n + k];
      g[i * m + j] = s;
    }
  Vec y(rhs);
  for (usize col = 0; col < m; ++col) {
    // max_by returns the last maximal element on ties
    usize piv = col;
    for (usize i = col; i < m; ++i)
      if (std::fabs(g[i * m + col]) >= std::fabs(g[piv * m + col])) piv = i;
    if (piv != col) {
      for (usize j = 0; j < m; ++j) std::swap(g[col * m + j], g[piv * m + j]);
      std::swap(y[col], y[piv]);
    }
    for (usize i = col + 1; i < m; ++i) {
      double f = g[i * m + col] / g[col * m + col];
      for (usize j = col; j < m; ++j) g[i * m + j] -= f * g[col * m + j];
      y[i] -= f * y[col];
    }
  }
  for (usize ii = m; ii-- > 0;) {
    double s = 0.0;
    for (usize j = ii + 1; j < m; ++j) s += g[ii * m + j] * y[j];
    y[ii] = (y[ii] - s) / g[ii * m + ii];
  }
  Vec out(n);
  for (usize k = 0; k < n; ++k) {
    double s = 0.0;
    for (usize i = 0; i < m; ++i) s += a[i * n + k] * y[i];
    out[k] = s;
  }
  return out;
}

// wq.rs:68-114 min_norm_qr (Householder QR of A^T)
static Vec min_norm_qr(const Vec& a, usize m, usize n, const Vec& rhs) {
  Vec r(n * m, 0.0);
  for (usize i = 0; i < m; ++i)
    for (usize k = 0; k < n; ++k) r[i * n + k] = a[i * n + k];
  std::vector<Vec> vs;
  for (usize j = 0; j < m; ++j) {
    const double* col = &r[j * n];
    double nrm = 0.0;
    for (usize k = j; k < n; ++k) nrm += col[k] * col[k];
    nrm = std::sqrt(nrm);
    Vec v(n, 0.0);
    for (usize k = j; k < n; ++k) v[k] = col[k];
    double alpha = col[j] >= 0.0 ? -nrm : nrm;
    v[j] -= alpha;
    double vn = 0.0;
    for (usize k = j; k < n; ++k) vn += v[k] * v[k];
    if (vn > 0.0) {
      for (usize c = j; c < m; ++c) {
        double* cc = &r[c * n];
        double s = 0.0;
        for (usize k = j; k < n; ++k) s += v[k] * cc[k];
        s = s * 2.0 / vn;
        for (usize k = j; k < n; ++k) cc[k] -= s * v[k];
      }
    }
    vs.push_back(v);
  }
  Vec y(n, 0.0);
  for (usize i = 0; i < m; ++i) {
    double s = 0.0;
    for (usize c = 0; c < i; ++c) s += r[i * n + c] * y[c];
    y[i] = (rhs[i] - s) / r[i * n + i];
  }
  for (usize j = m; j-- > 0;) {
    const Vec& v = vs[j];
    double vn = 0.0;
    for (usize k = j; k < n; ++k) vn += v[k] * v[k];
    if (vn > 0.0) {
      double s = 0.0;
      for (usize k = j; k < n; ++k) s += v[k] * y[k];
      s = s * 2.0 / vn;
      for (usize k = j; k < n; ++k) y[k] -= s * v[k];
    }
  }
  return y;
}

static Vec breaks_of(const Vec& knots) {
  Vec brk;
  for (double k : knots)
    if (brk.empty() || brk.back() != k) brk.push_back(k);
  return brk;
}

// wq.rs:141-240 rules_interior
static WqRules1d rules_interior(const Vec& knots, usize p) {
  if (p < 1) throw std::invalid_argument("wq: p >= 1");
  Vec brk = breaks_of(knots);
  usize nel = brk.size() - 1;
  Vec pts;
  std::vector<usize> el;
  for (usize k = 0; k < nel; ++k) {
    usize q = (k == 0 || k == nel - 1) ? std::max<usize>(3, p + 1) : 3;
    double h = brk[k + 1] - brk[k];
    for (usize m = 1; m <= q; ++m) {
      pts.push_back(brk[k] + (double)(2 * m - 1) / (double)(2 * q) * h);
      el.push_back(k);
    }
  }
  usize npt = pts.size();
  Vec hq(npt);
  for (usize q = 0; q < npt; ++q) hq[q] = brk[el[q] + 1] - brk[el[q]];
  usize ndof = knots.size() - p - 1;
  WqRules1d out;
  basis_dense(knots, p, pts, out.b, out.db);
  Space1d sp(knots, p, p + 1);
  Vec kt(p + 1, brk[0]);
  for (usize i = 1; i < nel; ++i) {
    kt.push_back(brk[i]);
    kt.push_back(brk[i]);
  }
  for (usize i = 0; i < p + 1; ++i) kt.push_back(brk[nel]);
  usize nt = kt.size() - p - 1;
  Vec bt, dummy;
  basis_dense(kt, p, pts, bt, dummy);
  Vec bm;
  basis_dense(knots, p - 1, pts, bm, dummy);
  usize nm = knots.size() - p;
  for (usize i = 0; i < ndof; ++i) {
    const auto& els = sp.supp[i];
    std::vector<usize> ind;
    for (usize q = 0; q < npt; ++q)
      if (std::find(els.begin(), els.end(), el[q]) != els.end()) ind.push_back(q);
    std::vector<usize> nb_p;
    for (usize e : els)
      for (usize r = 0; r < p + 1; ++r) nb_p.push_back(sp.connectivity[e * (p + 1) + r]);
    std::sort(nb_p.begin(), nb_p.end());
    nb_p.erase(std::unique(nb_p.begin(), nb_p.end()), nb_p.end());
    Vec xg, wg;
    for (usize e : els)
      for (usize q = 0; q < sp.nquad; ++q) {
        xg.push_back(sp.qn[e * sp.nquad + q]);
        wg.push_back(sp.qw[e * sp.nquad + q]);
      }
    Vec tg;
    basis_dense(kt, p, xg, tg, dummy);
    auto rule = [&](const std::function<double(usize)>& mvals, const std::function<double(usize)>& mg) {
      Vec z(ind.size());
      for (usize c = 0; c < ind.size(); ++c) z[c] = mvals(ind[c]) * hq[ind[c]];
      std::vector<usize> act;
      for (usize c = 0; c < ind.size(); ++c)
        if (z[c] > 0.0) act.push_back(c);
      Vec o(ind.size(), 0.0);
      if (act.empty()) return o;
      std::vector<usize> trial;
      for (usize j = 0; j < nt; ++j) {
        bool any = false;
        for (usize c : act)
          if (bt[ind[c] * nt + j] != 0.0) any = true;
        if (any) trial.push_back(j);
      }
      usize m = trial.size(), n = act.size();
      Vec a(m * n, 0.0);
      for (usize rr = 0; rr < m; ++rr)
        for (usize c = 0; c < n; ++c) a[rr * n + c] = bt[ind[act[c]] * nt + trial[rr]] * z[act[c]];
      Vec rhs(m);
      for (usize rr = 0; rr < m; ++rr) {
        double s = 0.0;
        for (usize g = 0; g < xg.size(); ++g) s += tg[g * nt + trial[rr]] * wg[g] * mg(g);
        rhs[rr] = s;
      }
      Vec v = min_norm_qr(a, m, n, rhs);
      for (usize c = 0; c < n; ++c) o[act[c]] = z[act[c]] * v[c];
      return o;
    };
    Vec bg, bmg;
    basis_dense(knots, p, xg, bg, dummy);
    basis_dense(knots, p - 1, xg, bmg, dummy);
    Vec w0 = rule([&](usize q) { return out.b[q * ndof + i]; }, [&](usize g) { return bg[g * ndof + i]; });
    double da = knots[i + p] - knots[i];
    double dbk = knots[i + p + 1] - knots[i + 1];
    Vec w1(ind.size(), 0.0);
    if (da > 0.0) {
      Vec r = rule([&](usize q) { return bm[q * nm + i]; }, [&](usize g) { return bmg[g * nm + i]; });
      for (usize c = 0; c < w1.size(); ++c) w1[c] += (double)p / da * r[c];
    }
    if (dbk > 0.0) {
      Vec r = rule([&](usize q) { return bm[q * nm + i + 1]; }, [&](usize g) { return bmg[g * nm + i + 1]; });
      for (usize c = 0; c < w1.size(); ++c) w1[c] -= (double)p / dbk * r[c];
    }
    out.weights[0].push_back(w0);
    out.weights[1].push_back(w1);
    out.weights[2].push_back(w0);
    out.weights[3].push_back(w1);
    out.ind_points.push_back(ind);
    out.neighbors.push_back(nb_p);
  }
  out.points = pts;
  out.ndof = ndof;
  return out;
}

// wq.rs:242-310 rules_calabro
static WqRules1d rules_calabro(const Vec& knots, usize p) {
  if (p < 1) throw std::invalid_argument("wq: p >= 1");
  Vec brk = breaks_of(knots);
  usize nb = brk.size();
  Vec pts = wq_linspace(brk[0], brk[1], p + 2);
  for (usize k = 2; k < (nb >= 2 ? nb - 2 : 0); ++k) pts.push_back(brk[k]);
  for (usize k = 1; k < (nb >= 2 ? nb - 2 : 0); ++k) pts.push_back(0.5 * (brk[k] + brk[k + 1]));
  Vec tail = wq_linspace(brk[nb - 2], brk[nb - 1], p + 2);
  pts.insert(pts.end(), tail.begin(), tail.end());
  std::sort(pts.begin(), pts.end());
  pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
  usize npt = pts.size();
  usize ndof = knots.size() - p - 1;
  WqRules1d out;
  basis_dense(knots, p, pts, out.b, out.db);
  Space1d sp(knots, p, p + 1);
  Vec kd(knots.begin() + 1, knots.end() - 1);
  Space1d spd(kd, p - 1, p + 1);
  usize ndof_d = kd.size() - p;
  Vec bd, dummy;
  basis_dense(kd, p - 1, pts, bd, dummy);
  for (usize i = 0; i < ndof; ++i) {
    std::vector<usize> ind;
    for (usize q = 0; q < npt; ++q)
      if (out.b[q * ndof + i] != 0.0) ind.push_back(q);
    const auto& els = sp.supp[i];
    std::vector<usize> nb_p, nb_d;
    for (usize e : els)
      for (usize r = 0; r < p + 1; ++r) nb_p.push_back(sp.connectivity[e * (p + 1) + r]);
    std::sort(nb_p.begin(), nb_p.end());
    nb_p.erase(std::unique(nb_p.begin(), nb_p.end()), nb_p.end());
    for (usize e : els)
      for (usize r = 0; r < p; ++r) nb_d.push_back(spd.connectivity[e * p + r]);
    std::sort(nb_d.begin(), nb_d.end());
    nb_d.erase(std::unique(nb_d.begin(), nb_d.end()), nb_d.end());
    Vec xg, wg;
    for (usize e : els)
      for (usize q = 0; q < sp.nquad; ++q) {
        xg.push_back(sp.qn[e * sp.nquad + q]);
        wg.push_back(sp.qw[e * sp.nquad + q]);
      }
    Vec bg, dbg, bdg;
    basis_dense(knots, p, xg, bg, dbg);
    basis_dense(kd, p - 1, xg, bdg, dummy);
    usize nq = ind.size();
    auto mat = [&](const Vec& vals, usize nd, const std::vector<usize>& nbs) {
      Vec a(nbs.size() * nq, 0.0);
      for (usize rr = 0; rr < nbs.size(); ++rr)
        for (usize c = 0; c < nq; ++c) a[rr * nq + c] = vals[ind[c] * nd + nbs[rr]];
      return a;
    };
    Vec ap = mat(out.b, ndof, nb_p);
    Vec ad = mat(bd, ndof_d, nb_d);
    auto rhs = [&](const Vec& trial, usize nd, const std::vector<usize>& nbs, const Vec& test) {
      Vec o(nbs.size());
      for (usize rr = 0; rr < nbs.size(); ++rr) {
        double s = 0.0;
        for (usize g = 0; g < xg.size(); ++g) s += trial[g * nd + nbs[rr]] * wg[g] * test[g * ndof + i];
        o[rr] = s;
      }
      return o;
    };
    out.weights[0].push_back(min_norm(ap, nb_p.size(), nq, rhs(bg, ndof, nb_p, bg)));
    out.weights[1].push_back(min_norm(ap, nb_p.size(), nq, rhs(bg, ndof, nb_p, dbg)));
    out.weights[2].push_back(min_norm(ad, nb_d.size(), nq, rhs(bdg, ndof_d, nb_d, bg)));
    out.weights[3].push_back(min_norm(ad, nb_d.size(), nq, rhs(bdg, ndof_d, nb_d, dbg)));
    out.ind_points.push_back(ind);
    out.neighbors.push_back(nb_p);
  }
  out.points = pts;
  out.ndof = ndof;
  return out;
}

WqRules1d wq_rules_1d(const Vec& knots, usize p, int layout) {
  return layout == 0 ? rules_interior(knots, p) : rules_calabro(knots, p);
}

// wq.rs:329-514 wq_stiffness_layout
CsrMatrix wq_stiffness_layout(const SpaceBox& sp, double young, double poisson, int density_kind, const double* x,
                              double penal, double emin, int layout) {
  usize dim = sp.dim;
  std::vector<WqRules1d> rules;
  for (usize d = 0; d < dim; ++d) rules.push_back(wq_rules_1d(sp.univ[d].knots, sp.degree[d], layout));
  std::vector<usize> npts;
  usize ntot = 1;
  for (auto& r : rules) {
    npts.push_back(r.points.size());
    ntot *= r.points.size();
  }
  Vec simp;
  if (density_kind == 0) {
    simp.assign(ntot, 1.0);
  } else if (density_kind == 1) {
    std::vector<std::vector<usize>> bins(dim);
    for (usize d = 0; d < dim; ++d) {
      const Vec& br = sp.univ[d].breaks;
      usize ne = br.size() - 1;
      for (double q : rules[d].points) {
        usize pp = (usize)(std::upper_bound(br.begin(), br.end(), q) - br.begin());
        usize s = pp == 0 ? 0 : pp - 1;
        bins[d].push_back(std::min(s, ne - 1));
      }
    }
    simp.resize(ntot);
    for (usize g = 0; g < ntot; ++g) {
      auto s = unravel(g, npts);
      usize e = 0, stride = 1;
      for (usize d = 0; d < dim; ++d) {
        e += bins[d][s[d]] * stride;
        stride *= sp.nel_dir[d];
      }
      simp[g] = emin + std::pow(x[e], penal) * (1.0 - emin);
    }
  } else {
    Vec cur(x, x + sp.ndof_sc);
    std::vector<usize> sizes = sp.ndof_dir;
    for (usize d = 0; d < dim; ++d) {
      const WqRules1d& r = rules[d];
      usize nd = r.ndof, nq = r.points.size();
      std::vector<usize> next_sizes = sizes;
      next_sizes[d] = nq;
      usize total = 1;
      for (usize v : next_sizes) total *= v;
      Vec next(total, 0.0);
      for (usize idx = 0; idx < total; ++idx) {
        auto s = unravel(idx, next_sizes);
        double acc = 0.0;
        for (usize a = 0; a < nd; ++a) {
          auto src = s;
          src[d] = a;
          usize li = 0, stride = 1;
          for (usize dd = 0; dd < dim; ++dd) {
            li += src[dd] * stride;
            stride *= sizes[dd];
          }
          acc += r.b[s[d] * nd + a] * cur[li];
        }
        next[idx] = acc;
      }
      cur = next;
      sizes = next_sizes;
    }
    simp.resize(cur.size());
    for (usize i = 0; i < cur.size(); ++i) simp[i] = emin + std::pow(cur[i], penal) * (1.0 - emin);
  }
  double lam = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
  double mu = young / (2.0 * (1.0 + poisson));
  double det = 1.0;
  for (double l : sp.lengths) det *= l;
  auto coef = [&](usize i, usize j, usize k1, usize k2) {
    auto dl = [](usize a, usize b) { return a == b ? 1.0 : 0.0; };
    return det / (sp.lengths[k1] * sp.lengths[k2]) *
           (lam * dl(i, k1) * dl(j, k2) + mu * (dl(i, j) * dl(k1, k2) + dl(i, k2) * dl(j, k1)));
  };
  usize nsc = sp.ndof_sc;
  std::vector<usize> rr, cc;
  Vec vv;
  for (usize ii = 0; ii < nsc; ++ii) {
    auto ind = unravel(ii, sp.ndof_dir);
    std::vector<const std::vector<usize>*> pts(dim), nbs(dim);
    std::vector<usize> nqd(dim), nnd(dim);
    usize nloc = 1, nql = 1;
    for (usize d = 0; d < dim; ++d) {
      pts[d] = &rules[d].ind_points[ind[d]];
      nbs[d] = &rules[d].neighbors[ind[d]];
      nqd[d] = pts[d]->size();
      nnd[d] = nbs[d]->size();
      nloc *= nnd[d];
      nql *= nqd[d];
    }
    Vec s_loc(nql);
    for (usize g = 0; g < nql; ++g) {
      auto s = unravel(g, nqd);
      usize gi = 0, stride = 1;
      for (usize d = 0; d < dim; ++d) {
        gi += (*pts[d])[s[d]] * stride;
        stride *= npts[d];
      }
      s_loc[g] = simp[gi];
    }
    std::vector<Vec> blocks(dim * dim, Vec(nloc, 0.0));
    for (usize k1 = 0; k1 < dim; ++k1)
      for (usize k2 = 0; k2 < dim; ++k2) {
        std::vector<Vec> bm(dim);
        for (usize d = 0; d < dim; ++d) {
          const WqRules1d& r = rules[d];
          int t;
          bool deriv;
          if (k1 == k2 && d == k1) {
            t = 3;
            deriv = true;
          } else if (k1 != k2 && d == k1) {
            t = 1;
            deriv = false;
          } else if (k1 != k2 && d == k2) {
            t = 2;
            deriv = true;
          } else {
            t = 0;
            deriv = false;
          }
          const Vec& wts = r.weights[t][ind[d]];
          const Vec& vals = deriv ? r.db : r.b;
          Vec m(nnd[d] * nqd[d], 0.0);
          for (usize jr = 0; jr < nnd[d]; ++jr)
            for (usize qc = 0; qc < nqd[d]; ++qc) m[jr * nqd[d] + qc] = wts[qc] * vals[(*pts[d])[qc] * r.ndof + (*nbs[d])[jr]];
          bm[d] = m;
        }
        Vec cur = s_loc;
        std::vector<usize> sizes = nqd;
        for (usize d = 0; d < dim; ++d) {
          std::vector<usize> next_sizes = sizes;
          next_sizes[d] = nnd[d];
          usize total = 1;
          for (usize v : next_sizes) total *= v;
          Vec next(total, 0.0);
          for (usize idx = 0; idx < total; ++idx) {
            auto s = unravel(idx, next_sizes);
            double acc = 0.0;
            for (usize q = 0; q < nqd[d]; ++q) {
              auto src = s;
              src[d] = q;
              usize li = 0, stride = 1;
              for (usize dd = 0; dd < dim; ++dd) {
                li += src[dd] * stride;
                stride *= sizes[dd];
              }
              acc += bm[d][s[d] * nqd[d] + q] * cur[li];
            }
            next[idx] = acc;
          }
          cur = next;
          sizes = next_sizes;
        }
        for (usize i = 0; i < dim; ++i)
          for (usize j = 0; j < dim; ++j) {
            double c = coef(i, j, k1, k2);
            if (c != 0.0)
              for (usize l = 0; l < nloc; ++l) blocks[i * dim + j][l] += c * cur[l];
          }
      }
    for (usize loc = 0; loc < nloc; ++loc) {
      auto s = unravel(loc, nnd);
      usize col = 0, stride = 1;
      for (usize d = 0; d < dim; ++d) {
        col += (*nbs[d])[s[d]] * stride;
        stride *= sp.ndof_dir[d];
      }
      for (usize i = 0; i < dim; ++i)
        for (usize j = 0; j < dim; ++j) {
          double v = blocks[i * dim + j][loc];
          rr.push_back(i * nsc + ii);
          cc.push_back(j * nsc + col);
          vv.push_back(0.5 * v);
          rr.push_back(j * nsc + col);
          cc.push_back(i * nsc + ii);
          vv.push_back(0.5 * v);
        }
    }
  }
  return CsrMatrix::from_triplets(sp.ndof, sp.ndof, rr, cc, vv);
}

}  // namespace orc
