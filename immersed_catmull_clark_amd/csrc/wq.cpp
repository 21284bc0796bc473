// Weighted-quadrature (WQ) diagonal: host-side 1D rules and tables.
//
// Semantics of rust/src/wq.rs (interior layout, rules_interior :141-240, and
// the (c N + ii, c N + ii) entries of wq_stiffness_layout :329-514): for a
// scalar control point ii and component c,
//   d(c, ii) = sum_k coef(c, c, k, k) * sum_q prod_d w_{k,d}(ii_d, q_d) * simp(q),
//   coef(c, c, k, k) = det / L_k^2 * (mu + (lambda + mu) [c == k]),
//   w_{k,d}(i, q) = W11_i(q) B_i'(q) if d == k, else W00_i(q) B_i(q),
// only the k1 = k2 derivative pairs reaching the diagonal (wq.rs:397-400).
// With element-wise densities simp(q) is constant on each element, so the
// sum over WQ points groups per element: A_{k,d}(i, e) = sum_{q in e} w_{k,d}(i, q),
// and d(c, ii) = sum_k coef_k sum_{e in supp(ii)} simp_e prod_d A_{k,d}(ii_d, e_d).
// This file builds A0 = sum W00 B and A1 = sum W11 B' per direction; the
// GPU kernel does the 27-element gather per control point.  It also builds
// the per-direction tables of the full sum-factorised formation (k_wq.hip).
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "icc_internal.hpp"

namespace icc {

namespace {

// Cox-de Boor values of all degree-p functions of `knots` at x (general knot
// vector, span clamped to [p, smax], smax = ndof-1 by default, like the
// reference's find_span).
void basis_all(const std::vector<double>& knots, int p, double x, std::vector<double>& out, int smax = -1) {
  const int ndof = (int)knots.size() - p - 1;
  out.assign(ndof, 0.0);
  int cnt = (int)(std::upper_bound(knots.begin(), knots.end(), x) - knots.begin());
  int s = std::min(std::max(cnt - 1, p), smax >= 0 ? smax : ndof - 1);
  std::vector<double> b(1, 1.0);
  for (int k = 1; k <= p; ++k) {
    std::vector<double> bn(k + 1, 0.0);
    for (int r = 0; r <= k; ++r) {
      int i = s + r - k;
      double v = 0.0;
      if (r > 0) {
        double d = knots[i + k] - knots[i];
        if (d > 0.0) v += (x - knots[i]) / d * b[r - 1];
      }
      if (r < k) {
        double d = knots[i + k + 1] - knots[i + 1];
        if (d > 0.0) v += (knots[i + k + 1] - x) / d * b[r];
      }
      bn[r] = v;
    }
    b.swap(bn);
  }
  for (int r = 0; r <= p; ++r) out[s - p + r] = b[r];
}

// derivative of all degree-p functions at x
void dbasis_all(const std::vector<double>& knots, int p, double x, std::vector<double>& out) {
  const int ndof = (int)knots.size() - p - 1;
  std::vector<double> lower;
  // degree p-1 on the same knots (ndof + 1 functions), on the span of the
  // degree-p evaluation: at the last knot, the degree-(p-1) clamp would pick
  // the empty span between repeated end knots, where every value is 0
  basis_all(knots, p - 1, x, lower, ndof - 1);
  out.assign(ndof, 0.0);
  for (int i = 0; i < ndof; ++i) {
    double v = 0.0;
    double da = knots[i + p] - knots[i], db = knots[i + p + 1] - knots[i + 1];
    if (da > 0.0) v += lower[i] / da;
    if (db > 0.0) v -= lower[i + 1] / db;
    out[i] = p * v;
  }
}

// Minimum-norm solution of A w = rhs (A: m x n row-major, m <= n, full row
// rank) by Householder QR of A^T: A^T = Q R, w = Q R^-T rhs.
std::vector<double> min_norm(const std::vector<double>& A, int m, int n, const std::vector<double>& rhs) {
  std::vector<double> R(n * m);  // column j of A^T (= row j of A), stored [j][n]
  for (int j = 0; j < m; ++j)
    for (int k = 0; k < n; ++k) R[j * n + k] = A[j * n + k];
  std::vector<std::vector<double>> V;
  for (int j = 0; j < m; ++j) {
    double* col = &R[j * n];
    double nrm = 0.0;
    for (int k = j; k < n; ++k) nrm += col[k] * col[k];
    nrm = std::sqrt(nrm);
    std::vector<double> v(n, 0.0);
    for (int k = j; k < n; ++k) v[k] = col[k];
    v[j] += col[j] >= 0.0 ? nrm : -nrm;
    double vn = 0.0;
    for (int k = j; k < n; ++k) vn += v[k] * v[k];
    if (vn > 0.0)
      for (int c = j; c < m; ++c) {
        double* cc = &R[c * n];
        double s = 0.0;
        for (int k = j; k < n; ++k) s += v[k] * cc[k];
        s *= 2.0 / vn;
        for (int k = j; k < n; ++k) cc[k] -= s * v[k];
      }
    V.push_back(v);
  }
  std::vector<double> y(n, 0.0);  // R^T y = rhs
  for (int i = 0; i < m; ++i) {
    double s = 0.0;
    for (int c = 0; c < i; ++c) s += R[i * n + c] * y[c];
    y[i] = (rhs[i] - s) / R[i * n + i];
  }
  for (int j = m - 1; j >= 0; --j) {  // w = H_0 ... H_{m-1} y
    const auto& v = V[j];
    double vn = 0.0, s = 0.0;
    for (int k = j; k < n; ++k) {
      vn += v[k] * v[k];
      s += v[k] * y[k];
    }
    if (vn > 0.0) {
      s *= 2.0 / vn;
      for (int k = j; k < n; ++k) y[k] -= s * v[k];
    }
  }
  return y;
}

// Minimum-norm solution through the normal equations: (A A^T) y = rhs by
// Gaussian elimination with partial pivoting (the last row of largest
// magnitude on ties), w = A^T y.  The Calabro rules are defined with this
// solver (wq.rs:33-62), which also fixes their value where A A^T is singular
// (a single-element direction).
std::vector<double> min_norm_normal(const std::vector<double>& A, int m, int n, const std::vector<double>& rhs) {
  std::vector<double> G((size_t)m * m), y = rhs;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += A[(size_t)i * n + k] * A[(size_t)j * n + k];
      G[(size_t)i * m + j] = s;
    }
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int i = c + 1; i < m; ++i)
      if (std::fabs(G[(size_t)i * m + c]) >= std::fabs(G[(size_t)piv * m + c])) piv = i;
    if (piv != c) {
      for (int j = 0; j < m; ++j) std::swap(G[(size_t)c * m + j], G[(size_t)piv * m + j]);
      std::swap(y[c], y[piv]);
    }
    for (int i = c + 1; i < m; ++i) {
      const double f = G[(size_t)i * m + c] / G[(size_t)c * m + c];
      for (int j = c; j < m; ++j) G[(size_t)i * m + j] -= f * G[(size_t)c * m + j];
      y[i] -= f * y[c];
    }
  }
  for (int i = m - 1; i >= 0; --i) {
    double s = 0.0;
    for (int j = i + 1; j < m; ++j) s += G[(size_t)i * m + j] * y[j];
    y[i] = (y[i] - s) / G[(size_t)i * m + i];
  }
  std::vector<double> w(n);
  for (int k = 0; k < n; ++k) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += A[(size_t)i * n + k] * y[i];
    w[k] = s;
  }
  return w;
}

}  // namespace

// WQ rules of rules_interior (wq.rs:141-240) for p = 2 on the open uniform
// knot vector of `nel` elements (parametric [0, 1]): 3 points per element at
// (2m - 1)/6 of it; for each test function i the points of the elements of
// its support (element first(i) + t, t = 0..2, first(i) = max(0, i - 2)),
// weights w0 (test value) and w1 (test derivative, by the recurrence
// B_i' = p/da N_{i,p-1} - p/db N_{i+1,p-1}), each the weighted minimum-norm
// rule exact on the C^0 quadratic trial space; W01 = W00 and W11 = W10 in
// this layout.  b/db: all basis values / parametric derivatives at the points.
void wq_rules_p2(int64_t nel, WqRulesP2& r) {
  const int p = 2;
  std::vector<double> knots(3, 0.0);
  for (int64_t i = 1; i < nel; ++i) knots.push_back((double)i / (double)nel);
  for (int i = 0; i < 3; ++i) knots.push_back(1.0);
  std::vector<double> brk;
  for (double k : knots)
    if (brk.empty() || brk.back() != k) brk.push_back(k);
  const int ne = (int)brk.size() - 1;
  const int ndof = (int)knots.size() - p - 1;
  std::vector<double> pts, hq;
  std::vector<int> el;
  for (int k = 0; k < ne; ++k) {
    const int q = 3;
    double h = brk[k + 1] - brk[k];
    for (int m = 1; m <= q; ++m) {
      pts.push_back(brk[k] + (double)(2 * m - 1) / (double)(2 * q) * h);
      el.push_back(k);
      hq.push_back(h);
    }
  }
  const int npt = (int)pts.size();
  // C^0 quadratic trial space (interior breaks doubled) and Gauss data
  std::vector<double> kt(p + 1, brk[0]);
  for (int i = 1; i < ne; ++i) {
    kt.push_back(brk[i]);
    kt.push_back(brk[i]);
  }
  for (int i = 0; i <= p; ++i) kt.push_back(brk[ne]);
  const int nt = (int)kt.size() - p - 1;
  const double gx[3] = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)}, gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  std::vector<std::vector<double>> Bq(npt), dBq(npt), Tq(npt), Mq(npt);
  for (int q = 0; q < npt; ++q) {
    basis_all(knots, p, pts[q], Bq[q]);
    dbasis_all(knots, p, pts[q], dBq[q]);
    basis_all(kt, p, pts[q], Tq[q]);
    basis_all(knots, p - 1, pts[q], Mq[q]);
  }
  r.nel = ne;
  r.ndof = ndof;
  r.npt = npt;
  r.pts = pts;
  r.bin = el;
  r.b.assign((size_t)npt * ndof, 0.0);
  r.db.assign((size_t)npt * ndof, 0.0);
  for (int q = 0; q < npt; ++q)
    for (int j = 0; j < ndof; ++j) {
      r.b[(size_t)q * ndof + j] = Bq[q][j];
      r.db[(size_t)q * ndof + j] = dBq[q][j];
    }
  r.ind.assign(ndof, {});
  r.w0.assign(ndof, {});
  r.w1.assign(ndof, {});
  for (int i = 0; i < ndof; ++i) {
    const int e_lo = std::max(0, i - p), e_hi = std::min(i, ne - 1);
    std::vector<int> ind;
    for (int q = 0; q < npt; ++q)
      if (el[q] >= e_lo && el[q] <= e_hi) ind.push_back(q);
    // Gauss points on supp(B_i)
    std::vector<double> xg, wg;
    for (int e = e_lo; e <= e_hi; ++e) {
      double a = brk[e], h = brk[e + 1] - brk[e];
      for (int g = 0; g < 3; ++g) {
        xg.push_back(a + 0.5 * (gx[g] + 1.0) * h);
        wg.push_back(0.5 * gw[g] * h);
      }
    }
    std::vector<std::vector<double>> Tg(xg.size()), Bg(xg.size()), Mg(xg.size());
    for (size_t g = 0; g < xg.size(); ++g) {
      basis_all(kt, p, xg[g], Tg[g]);
      basis_all(knots, p, xg[g], Bg[g]);
      basis_all(knots, p - 1, xg[g], Mg[g]);
    }
    // weighted min-norm rule on ind, exact on the trial space
    auto rule = [&](auto mval, auto mgauss) {
      std::vector<double> z(ind.size()), out(ind.size(), 0.0);
      std::vector<int> act;
      for (size_t c = 0; c < ind.size(); ++c) {
        z[c] = mval(ind[c]) * hq[ind[c]];
        if (z[c] > 0.0) act.push_back((int)c);
      }
      if (act.empty()) return out;
      std::vector<int> trial;
      for (int j = 0; j < nt; ++j) {
        bool any = false;
        for (int c : act) any = any || Tq[ind[c]][j] != 0.0;
        if (any) trial.push_back(j);
      }
      const int m = (int)trial.size(), n = (int)act.size();
      std::vector<double> A((size_t)m * n), rhs(m);
      for (int rr = 0; rr < m; ++rr) {
        for (int c = 0; c < n; ++c) A[(size_t)rr * n + c] = Tq[ind[act[c]]][trial[rr]] * z[act[c]];
        double s = 0.0;
        for (size_t g = 0; g < xg.size(); ++g) s += Tg[g][trial[rr]] * wg[g] * mgauss(g);
        rhs[rr] = s;
      }
      std::vector<double> v = min_norm(A, m, n, rhs);
      for (int c = 0; c < n; ++c) out[act[c]] = z[act[c]] * v[c];
      return out;
    };
    std::vector<double> w0 = rule([&](int q) { return Bq[q][i]; }, [&](size_t g) { return Bg[g][i]; });
    std::vector<double> w1(ind.size(), 0.0);
    const double da = knots[i + p] - knots[i], db = knots[i + p + 1] - knots[i + 1];
    if (da > 0.0) {
      auto rr = rule([&](int q) { return Mq[q][i]; }, [&](size_t g) { return Mg[g][i]; });
      for (size_t c = 0; c < w1.size(); ++c) w1[c] += (double)p / da * rr[c];
    }
    if (db > 0.0) {
      auto rr = rule([&](int q) { return Mq[q][i + 1]; }, [&](size_t g) { return Mg[g][i + 1]; });
      for (size_t c = 0; c < w1.size(); ++c) w1[c] -= (double)p / db * rr[c];
    }
    r.ind[i] = ind;
    r.w0[i] = w0;
    r.w1[i] = w1;
  }
  r.w2 = r.w0;
  r.w3 = r.w1;
}

namespace {
// open uniform knot vector of degree 2 on [0, 1] with nel elements
std::vector<double> open_knots_p2(int64_t nel) {
  std::vector<double> knots(3, 0.0);
  for (int64_t i = 1; i < nel; ++i) knots.push_back((double)i / (double)nel);
  for (int i = 0; i < 3; ++i) knots.push_back(1.0);
  return knots;
}

// n equispaced values from a to b with both ends exact (wq.rs:26-31)
void push_linspace(double a, double b, int n, std::vector<double>& out) {
  for (int i = 0; i < n; ++i) out.push_back(i == 0 ? a : (i == n - 1 ? b : a + (double)i * (b - a) / (double)(n - 1)));
}
}  // namespace

// WQ rules of rules_calabro (wq.rs:242-310) for p = 2: p + 2 equispaced points
// in each end element, the interior knots and the midpoints of the interior
// elements; for each test function i its points are those where B_i is
// nonzero, and the four rules are unweighted minimum-norm solutions exact on
// the spline space (W00, W10: trials B_j, j in the support's elements) and on
// its derivative space (W01, W11: degree-1 functions of the knot vector
// without its end knots), against 3-point Gauss integrals over supp(B_i).
void wq_rules_calabro_p2(int64_t nel, WqRulesP2& r) {
  const int p = 2;
  const std::vector<double> knots = open_knots_p2(nel);
  std::vector<double> brk;
  for (double k : knots)
    if (brk.empty() || brk.back() != k) brk.push_back(k);
  const int nb = (int)brk.size(), ne = nb - 1;
  std::vector<double> pts;
  push_linspace(brk[0], brk[1], p + 2, pts);
  for (int k = 2; k < nb - 2; ++k) pts.push_back(brk[k]);
  for (int k = 1; k < nb - 2; ++k) pts.push_back(0.5 * (brk[k] + brk[k + 1]));
  push_linspace(brk[nb - 2], brk[nb - 1], p + 2, pts);
  std::sort(pts.begin(), pts.end());
  pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
  const int npt = (int)pts.size(), ndof = (int)knots.size() - p - 1;
  // derivative space: degree 1 on the knot vector without its first and last knot
  const std::vector<double> kd(knots.begin() + 1, knots.end() - 1);
  r.nel = ne;
  r.ndof = ndof;
  r.npt = npt;
  r.pts = pts;
  r.bin.resize(npt);
  r.b.assign((size_t)npt * ndof, 0.0);
  r.db.assign((size_t)npt * ndof, 0.0);
  std::vector<std::vector<double>> Bd(npt);
  std::vector<double> tmp;
  for (int q = 0; q < npt; ++q) {
    const int e = (int)(std::upper_bound(brk.begin(), brk.end(), pts[q]) - brk.begin()) - 1;
    r.bin[q] = std::min(std::max(e, 0), ne - 1);
    basis_all(knots, p, pts[q], tmp);
    for (int j = 0; j < ndof; ++j) r.b[(size_t)q * ndof + j] = tmp[j];
    dbasis_all(knots, p, pts[q], tmp);
    for (int j = 0; j < ndof; ++j) r.db[(size_t)q * ndof + j] = tmp[j];
    basis_all(kd, p - 1, pts[q], Bd[q]);
  }
  const double gx[3] = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)}, gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  r.ind.assign(ndof, {});
  r.w0.assign(ndof, {});
  r.w1.assign(ndof, {});
  r.w2.assign(ndof, {});
  r.w3.assign(ndof, {});
  for (int i = 0; i < ndof; ++i) {
    std::vector<int> ind;
    for (int q = 0; q < npt; ++q)
      if (r.b[(size_t)q * ndof + i] != 0.0) ind.push_back(q);
    // support elements of B_i; trials: functions e..e+2 (space) and e..e+1 (derivative space) of each
    const int e_lo = std::max(0, i - p), e_hi = std::min(i, ne - 1);
    const int jp_lo = e_lo, jp_hi = e_hi + p, jd_lo = e_lo, jd_hi = e_hi + p - 1;
    std::vector<double> xg, wg;
    for (int e = e_lo; e <= e_hi; ++e) {
      const double a = brk[e], h = brk[e + 1] - brk[e];
      for (int g = 0; g < 3; ++g) {
        xg.push_back(a + 0.5 * (gx[g] + 1.0) * h);
        wg.push_back(0.5 * gw[g] * h);
      }
    }
    const int nq = (int)ind.size(), ng = (int)xg.size();
    std::vector<std::vector<double>> Bg(ng), dBg(ng), Bdg(ng);
    for (int g = 0; g < ng; ++g) {
      basis_all(knots, p, xg[g], Bg[g]);
      dbasis_all(knots, p, xg[g], dBg[g]);
      basis_all(kd, p - 1, xg[g], Bdg[g]);
    }
    auto rule = [&](bool deriv_trial, bool deriv_test) {
      const int lo = deriv_trial ? jd_lo : jp_lo, m = (deriv_trial ? jd_hi : jp_hi) - lo + 1;
      std::vector<double> A((size_t)m * nq), rhs(m);
      for (int rr = 0; rr < m; ++rr) {
        const int j = lo + rr;
        for (int c = 0; c < nq; ++c) A[(size_t)rr * nq + c] = deriv_trial ? Bd[ind[c]][j] : r.b[(size_t)ind[c] * ndof + j];
        double s = 0.0;
        for (int g = 0; g < ng; ++g) s += (deriv_trial ? Bdg[g][j] : Bg[g][j]) * wg[g] * (deriv_test ? dBg[g][i] : Bg[g][i]);
        rhs[rr] = s;
      }
      return min_norm_normal(A, m, nq, rhs);
    };
    r.ind[i] = ind;
    r.w0[i] = rule(false, false);
    r.w1[i] = rule(false, true);
    r.w2[i] = rule(true, false);
    r.w3[i] = rule(true, true);
  }
}

// A0[(i*3 + t)], A1[(i*3 + t)]: sums over the WQ points of element
// first(i) + t for the 1D function i, first(i) = max(0, i - 2) (interior layout).
void wq_tables_1d(int64_t nel, std::vector<double>& A0, std::vector<double>& A1) {
  WqRulesP2 r;
  wq_rules_p2(nel, r);
  A0.assign((size_t)r.ndof * 3, 0.0);
  A1.assign((size_t)r.ndof * 3, 0.0);
  for (int i = 0; i < r.ndof; ++i)
    for (size_t c = 0; c < r.ind[i].size(); ++c) {
      const int q = r.ind[i][c], t = q / 3 - std::max(0, i - 2);
      A0[(size_t)i * 3 + t] += r.w0[i][c] * r.b[(size_t)q * r.ndof + i];
      A1[(size_t)i * 3 + t] += r.w1[i][c] * r.db[(size_t)q * r.ndof + i];
    }
}

// Formation tables of one direction for the sum-factorised WQ stiffness:
// M[((i * 4 + t) * 5 + jr) * 9 + qc] = W_t(i, qc) * val_t(point pstart[i] + qc, trial i - 2 + jr)
// with (W, val) = (W00, B), (W10, B), (W01, B'), (W11, B') for t = 0..3 (the
// contractions of wq.rs:418-438; B' parametric), zero for missing trials or
// points; pstart[i] is the first of the row's (contiguous) points.
void wq_form_tables(const WqRulesP2& r, std::vector<double>& M, std::vector<int>& pstart) {
  M.assign((size_t)r.ndof * 4 * 5 * 9, 0.0);
  pstart.assign(r.ndof, 0);
  for (int i = 0; i < r.ndof; ++i) {
    if (r.ind[i].empty()) continue;
    pstart[i] = r.ind[i][0];
    for (size_t c = 0; c < r.ind[i].size(); ++c) {
      const int q = r.ind[i][c];
      const int qc = q - pstart[i];
      if (qc >= 9) throw std::logic_error("wq_form_tables: more than 9 points in a row");
      for (int jr = 0; jr < 5; ++jr) {
        const int j = i - 2 + jr;
        if (j < 0 || j >= r.ndof) continue;
        const double bv = r.b[(size_t)q * r.ndof + j], dv = r.db[(size_t)q * r.ndof + j];
        const double wt[4] = {r.w0[i][c] * bv, r.w1[i][c] * bv, r.w2[i][c] * dv, r.w3[i][c] * dv};
        for (int t = 0; t < 4; ++t) M[(((size_t)i * 4 + t) * 5 + jr) * 9 + qc] = wt[t];
      }
    }
  }
}

void wq_form_tables_1d(int64_t nel, std::vector<double>& M) {
  WqRulesP2 r;
  wq_rules_p2(nel, r);
  std::vector<int> pstart;
  wq_form_tables(r, M, pstart);
}

}  // namespace icc
