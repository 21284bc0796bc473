// Oracle: CSR matrices, element-wise assembly and Jacobi PCG.  Restates
// rust/src/sparse.rs and rust/src/solver.rs of the reference.  The row loop of
// the SpMV runs under OpenMP (schedule static) in place of rayon's
// par_iter_mut; rows are independent, so results do not depend on threads.
// Test infrastructure only (see oracle.hpp).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// sparse.rs:16-57 CsrMatrix::from_triplets (counting sort by row, stable
// sort by column, merge duplicates in insertion order)
CsrMatrix CsrMatrix::from_triplets(usize nrows, usize ncols, const std::vector<usize>& rows,
                                   const std::vector<usize>& cols, const Vec& vals) {
  if (!(rows.size() == cols.size() && rows.size() == vals.size())) throw std::invalid_argument("from_triplets");
  std::vector<usize> count(nrows + 1, 0);
  for (usize r : rows) count[r + 1] += 1;
  for (usize i = 0; i < nrows; ++i) count[i + 1] += count[i];
  std::vector<usize> next(count);
  std::vector<usize> ci(rows.size());
  Vec cv(rows.size());
  for (usize k = 0; k < rows.size(); ++k) {
    usize r = rows[k];
    ci[next[r]] = cols[k];
    cv[next[r]] = vals[k];
    next[r] += 1;
  }
  CsrMatrix m;
  m.nrows = nrows;
  m.ncols = ncols;
  m.indptr.assign(nrows + 1, 0);
  m.indices.reserve(rows.size());
  m.data.reserve(rows.size());
  std::vector<usize> perm;
  for (usize r = 0; r < nrows; ++r) {
    usize s = count[r], e = count[r + 1];
    perm.resize(e - s);
    std::iota(perm.begin(), perm.end(), s);
    // sort_unstable_by_key: the merge below sums duplicates, whose order only
    // affects rounding; a stable sort keeps insertion order (the order the
    // reference's pdqsort leaves for small runs of equal keys is not
    // specified, so duplicate sums are compared to round-off, not bits).
    std::stable_sort(perm.begin(), perm.end(), [&](usize a, usize b) { return ci[a] < ci[b]; });
    bool has_last = false;
    usize last = 0;
    for (usize k : perm) {
      if (has_last && last == ci[k]) {
        m.data.back() += cv[k];
      } else {
        m.indices.push_back(ci[k]);
        m.data.push_back(cv[k]);
        last = ci[k];
        has_last = true;
      }
    }
    m.indptr[r + 1] = m.indices.size();
  }
  return m;
}

// sparse.rs:60-68 matvec (parallel over rows)
void CsrMatrix::matvec(const double* x, double* y) const {
  const long n = (long)nrows;
#pragma omp parallel for schedule(static)
  for (long r = 0; r < n; ++r) {
    double s = 0.0;
    for (usize k = indptr[r]; k < indptr[r + 1]; ++k) s += data[k] * x[indices[k]];
    y[r] = s;
  }
}

// sparse.rs:70-78 diagonal
Vec CsrMatrix::diagonal() const {
  Vec d(nrows, 0.0);
  for (usize r = 0; r < nrows; ++r)
    for (usize k = indptr[r]; k < indptr[r + 1]; ++k)
      if (indices[k] == r) {
        d[r] = data[k];
        break;
      }
  return d;
}

// sparse.rs:80-86 get
double CsrMatrix::get(usize r, usize c) const {
  auto b = indices.begin() + (long)indptr[r], e = indices.begin() + (long)indptr[r + 1];
  auto it = std::lower_bound(b, e, c);
  if (it != e && *it == c) return data[(usize)(it - indices.begin())];
  return 0.0;
}

// sparse.rs:89-91 frobenius
double CsrMatrix::frobenius() const {
  double s = 0.0;
  for (double v : data) s += v * v;
  return std::sqrt(s);
}

// sparse.rs:98-113 plus_triplets
CsrMatrix CsrMatrix::plus_triplets(const std::vector<usize>& rows, const std::vector<usize>& cols,
                                   const Vec& vals) const {
  std::vector<usize> r, c;
  Vec v;
  r.reserve(nnz() + rows.size());
  c.reserve(nnz() + rows.size());
  v.reserve(nnz() + rows.size());
  for (usize i = 0; i < nrows; ++i)
    for (usize k = indptr[i]; k < indptr[i + 1]; ++k) {
      r.push_back(i);
      c.push_back(indices[k]);
      v.push_back(data[k]);
    }
  r.insert(r.end(), rows.begin(), rows.end());
  c.insert(c.end(), cols.begin(), cols.end());
  v.insert(v.end(), vals.begin(), vals.end());
  return from_triplets(nrows, ncols, r, c, v);
}

// sparse.rs:128-157 AssemblyPattern::new (row -> sorted unique columns; pos map)
AssemblyPattern::AssemblyPattern(usize ndof_, const std::vector<usize>& conn, usize nsh_, usize nel)
    : ndof(ndof_), nsh(nsh_) {
  std::vector<std::vector<usize>> rows(ndof);
  for (usize e = 0; e < nel; ++e) {
    const usize* ce = &conn[e * nsh];
    for (usize a = 0; a < nsh; ++a) rows[ce[a]].insert(rows[ce[a]].end(), ce, ce + nsh);
  }
  indptr.assign(ndof + 1, 0);
  for (usize r = 0; r < ndof; ++r) {
    auto& row = rows[r];
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    indices.insert(indices.end(), row.begin(), row.end());
    indptr[r + 1] = indices.size();
    std::vector<usize>().swap(row);
  }
  pos.reserve(nel * nsh * nsh);
  for (usize e = 0; e < nel; ++e) {
    const usize* ce = &conn[e * nsh];
    for (usize a = 0; a < nsh; ++a) {
      auto b = indices.begin() + (long)indptr[ce[a]], en = indices.begin() + (long)indptr[ce[a] + 1];
      for (usize bb = 0; bb < nsh; ++bb) {
        auto it = std::lower_bound(b, en, ce[bb]);
        pos.push_back(indptr[ce[a]] + (usize)(it - b));
      }
    }
  }
}

// sparse.rs:160-175 AssemblyPattern::assemble (serial, element order)
CsrMatrix AssemblyPattern::assemble(const ElementMatrices& em, const Vec& scale) const {
  usize nn = nsh * nsh;
  CsrMatrix m;
  m.nrows = ndof;
  m.ncols = ndof;
  m.indptr = indptr;
  m.indices = indices;
  m.data.assign(indices.size(), 0.0);
  for (usize e = 0; e < scale.size(); ++e) {
    double s = scale[e];
    const double* ke = em.of_element(e);
    const usize* p = &pos[e * nn];
    for (usize k = 0; k < nn; ++k) m.data[p[k]] += s * ke[k];
  }
  return m;
}

// sparse.rs:179-200 assemble (triplets)
CsrMatrix assemble(usize ndof, const std::vector<usize>& conn, usize nsh, usize nel, const ElementMatrices& em,
                   const double* scale) {
  std::vector<usize> rows, cols;
  Vec vals;
  usize nn = nsh * nsh * nel;
  rows.reserve(nn);
  cols.reserve(nn);
  vals.reserve(nn);
  for (usize e = 0; e < nel; ++e) {
    const double* ke = em.of_element(e);
    double s = scale ? scale[e] : 1.0;
    const usize* ce = &conn[e * nsh];
    for (usize a = 0; a < nsh; ++a)
      for (usize b = 0; b < nsh; ++b) {
        rows.push_back(ce[a]);
        cols.push_back(ce[b]);
        vals.push_back(s * ke[a * nsh + b]);
      }
  }
  return CsrMatrix::from_triplets(ndof, ndof, rows, cols, vals);
}

// solver.rs:14-76 PcgSolver::solve (serial vector loops in the reference's order)
PcgResult pcg_solve(usize max_iter, double tol, usize n, const Vec& b, const Vec& diag_a,
                    const std::function<void(const double*, double*)>& matvec) {
  PcgResult out;
  Vec& x = out.x;
  x.assign(n, 0.0);
  Vec r(b), z(n, 0.0), p(n, 0.0), ap(n, 0.0), inv_diag(n);
  for (usize i = 0; i < n; ++i) inv_diag[i] = std::fabs(diag_a[i]) > 1e-12 ? 1.0 / diag_a[i] : 1.0;
  for (usize i = 0; i < n; ++i) {
    z[i] = inv_diag[i] * r[i];
    p[i] = z[i];
  }
  double rz_old = 0.0;
  for (usize i = 0; i < n; ++i) rz_old += r[i] * z[i];
  double nb = 0.0;
  for (usize i = 0; i < n; ++i) nb += b[i] * b[i];
  double norm_b = std::max(std::sqrt(nb), 1e-16);
  double final_res = 1.0;
  usize final_iter = 0;
  for (usize iter = 0; iter < max_iter; ++iter) {
    final_iter = iter + 1;
    matvec(p.data(), ap.data());
    double p_ap = 0.0;
    for (usize i = 0; i < n; ++i) p_ap += p[i] * ap[i];
    if (!(p_ap > 0.0)) break;
    double alpha = rz_old / p_ap;
    for (usize i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * ap[i];
    }
    double nr = 0.0;
    for (usize i = 0; i < n; ++i) nr += r[i] * r[i];
    final_res = std::sqrt(nr) / norm_b;
    if (final_res < tol) break;
    for (usize i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
    double rz_new = 0.0;
    for (usize i = 0; i < n; ++i) rz_new += r[i] * z[i];
    double beta = rz_new / rz_old;
    for (usize i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz_old = rz_new;
  }
  out.iters = final_iter;
  out.res = final_res;
  return out;
}

}  // namespace orc
