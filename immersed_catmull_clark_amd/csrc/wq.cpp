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
#include <vector>

#include "icc_internal.hpp"

namespace icc {

namespace {

// Cox-de Boor values of all degree-p functions of `knots` at x (general knot
// vector, span clamped to [p, ndof-1] like the reference's find_span).
void basis_all(const std::vector<double>& knots, int p, double x, std::vector<double>& out) {
  const int ndof = (int)knots.size() - p - 1;
  out.assign(ndof, 0.0);
  int cnt = (int)(std::upper_bound(knots.begin(), knots.end(), x) - knots.begin());
  int s = std::min(std::max(cnt - 1, p), ndof - 1);
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
  basis_all(knots, p - 1, x, lower);  // degree p-1 on the same knots: ndof + 1 functions
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
}

// A0[(i*3 + t)], A1[(i*3 + t)]: sums over the WQ points of element
// first(i) + t for the 1D function i, first(i) = max(0, i - 2).
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
// M[((i * 4 + t) * 5 + jr) * 9 + qc] = w_t(i, qc) * val_t(point ind_i[qc], trial i - 2 + jr)
// with (w, val) = (w0, B), (w1, B), (w0, B'), (w1, B') for t = 0..3 (the
// W00, W10, W01, W11 contractions of wq.rs:418-438; B' parametric), zero for
// missing trials or points; the row's points are those of elements
// first(i) + qc / 3, point qc % 3 of each.
void wq_form_tables_1d(int64_t nel, std::vector<double>& M) {
  WqRulesP2 r;
  wq_rules_p2(nel, r);
  M.assign((size_t)r.ndof * 4 * 5 * 9, 0.0);
  for (int i = 0; i < r.ndof; ++i)
    for (size_t c = 0; c < r.ind[i].size(); ++c) {
      const int q = r.ind[i][c];
      const int qc = q - 3 * std::max(0, i - 2);
      for (int jr = 0; jr < 5; ++jr) {
        const int j = i - 2 + jr;
        if (j < 0 || j >= r.ndof) continue;
        const double bv = r.b[(size_t)q * r.ndof + j], dv = r.db[(size_t)q * r.ndof + j];
        const double wt[4] = {r.w0[i][c] * bv, r.w1[i][c] * bv, r.w0[i][c] * dv, r.w1[i][c] * dv};
        for (int t = 0; t < 4; ++t) M[(((size_t)i * 4 + t) * 5 + jr) * 9 + qc] = wt[t];
      }
    }
}

}  // namespace icc
