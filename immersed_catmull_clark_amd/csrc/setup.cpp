// Host-side setup of the p = 2 grid: open uniform knots, Gauss data, the 1D
// element matrices per direction, the element types and their 81x81 matrices,
// and the constant-type tiling the apply kernels walk.
//
// Semantics follow the reference (rust/src/iga/{quadrature,space,elasticity}.rs):
// types are the distinct combinations of per-direction 1D types, numbered in
// order of first appearance (elasticity.rs:85-98); the kernels group 1D types
// by a tolerance, while the reported numbering uses the reference's rounded
// key (see build_setup); K = lambda G_ij + mu G_ji + mu delta_ij sum_k G_kk
// (elasticity.rs:125-145).
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <string>

#include "icc_internal.hpp"

namespace icc {

namespace {

// Gauss-Legendre nodes/weights on [-1, 1] by Newton on P_n (ascending).
void gauss(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    double d = n * (z * p1 - p0) / (z * z - 1.0);
    if (d != 0.0) dp = d;
    x[i] = -z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// Quadratic B-spline values / derivatives on the open uniform knot vector of
// `nel` elements at parametric xi inside element e (local functions 0..2).
void quad_basis(const std::vector<double>& knots, int64_t e, double xi, double v[3], double dv[3]) {
  // span s = e + 2 (knots has 3 repeated zeros)
  const int64_t s = e + 2;
  // degree 1 functions s-1 and s on the span [knots[s], knots[s+1])
  double b1[2] = {0.0, 0.0};
  const double d1 = knots[s + 1] - knots[s];
  if (d1 > 0.0) {
    b1[0] = (knots[s + 1] - xi) / d1 * 1.0;
    b1[1] = (xi - knots[s]) / d1 * 1.0;
  }
  // degree 2: functions s-2, s-1, s
  for (int r = 0; r < 3; ++r) {
    int64_t i = s + r - 2;
    double val = 0.0;
    if (r > 0) {
      double d = knots[i + 2] - knots[i];
      if (d > 0.0) val += (xi - knots[i]) / d * b1[r - 1];
    }
    if (r < 2) {
      double d = knots[i + 3] - knots[i + 1];
      if (d > 0.0) val += (knots[i + 3] - xi) / d * b1[r];
    }
    v[r] = val;
  }
  for (int r = 0; r < 3; ++r) {
    int64_t i = s + r - 2;
    double val = 0.0;
    if (r > 0) {
      double d = knots[i + 2] - knots[i];
      if (d > 0.0) val += b1[r - 1] / d;
    }
    if (r < 2) {
      double d = knots[i + 3] - knots[i + 1];
      if (d > 0.0) val -= b1[r] / d;
    }
    dv[r] = 2.0 * val;
  }
}

std::vector<double> open_knots_p2(int64_t nel) {
  std::vector<double> k(3, 0.0);
  for (int64_t i = 1; i < nel; ++i) {
    double b = (double)i / (double)nel;
    k.push_back(b);
  }
  for (int i = 0; i < 3; ++i) k.push_back(1.0);
  return k;
}

// 1D element matrices of one direction: m[k][a][b], k = 00, 10, 01, 11,
// first index = test, "1" = physical derivative.
void element_mats_1d(int64_t nel, double length, std::vector<std::array<double, 36>>& per_elem) {
  std::vector<double> knots = open_knots_p2(nel);
  double xg[3], wg[3];
  gauss(3, xg, wg);
  per_elem.assign(nel, {});
  for (int64_t e = 0; e < nel; ++e) {
    double a = knots[e + 2], b = knots[e + 3];
    double h = b - a;
    auto& m = per_elem[e];
    m.fill(0.0);
    for (int q = 0; q < 3; ++q) {
      double xi = a + 0.5 * (xg[q] + 1.0) * h;
      double w = 0.5 * wg[q] * h * length;
      double v[3], dv[3];
      quad_basis(knots, e, xi, v, dv);
      for (int i = 0; i < 3; ++i) {
        double ni = v[i], di = dv[i] / length;
        for (int j = 0; j < 3; ++j) {
          double nj = v[j], dj = dv[j] / length;
          m[0 * 9 + i * 3 + j] += w * ni * nj;
          m[1 * 9 + i * 3 + j] += w * di * nj;
          m[2 * 9 + i * 3 + j] += w * ni * dj;
          m[3 * 9 + i * 3 + j] += w * di * dj;
        }
      }
    }
  }
}

// Picks a tile shape for a constant-type sub-box of extents (ex, ey, ez):
// as many elements as possible (<= TILE_MAX_EL), staged points within
// TILE_MAX_PTS, ties broken by fewer staged points per element.
void tile_shape(int64_t ex, int64_t ey, int64_t ez, int& tx, int& ty, int& tz) {
  static const int cand[] = {1, 2, 4, 8, 16, 32};
  double best_score = -1.0;
  tx = ty = tz = 1;
  for (int a : cand) {
    if (a > 2 * ex && a != 1) continue;
    for (int b : cand) {
      if (b > 2 * ey && b != 1) continue;
      for (int c : cand) {
        if (c > 2 * ez && c != 1) continue;
        int ax = (int)std::min<int64_t>(a, ex), by = (int)std::min<int64_t>(b, ey), cz = (int)std::min<int64_t>(c, ez);
        int nelem = ax * by * cz;
        int pts = (ax + 2) * (by + 2) * (cz + 2);
        if (nelem > TILE_MAX_EL || pts > TILE_MAX_PTS) continue;
        // number of tiles and total staged points for the whole sub-box
        int64_t ntx = (ex + ax - 1) / ax, nty = (ey + by - 1) / by, ntz = (ez + cz - 1) / cz;
        double ntiles = (double)(ntx * nty * ntz);
        double fill = (double)(ex * ey * ez) / (ntiles * TILE_MAX_EL);
        double ppe = (double)pts / nelem;
        double score = fill * 100.0 - ppe;  // prefer full tiles, then few points per element
        if (score > best_score) {
          best_score = score;
          tx = ax;
          ty = by;
          tz = cz;
        }
      }
    }
  }
}

}  // namespace

int32_t element_type(const GridSetup& s, int64_t i, int64_t j, int64_t k) {
  int T0 = (int)s.mats1[0].size(), T1 = (int)s.mats1[1].size();
  return s.type_of[s.tid1[0][i] + T0 * (s.tid1[1][j] + T1 * s.tid1[2][k])];
}

int32_t reference_type(const GridSetup& s, int64_t i, int64_t j, int64_t k) {
  const int64_t R0 = 1 + *std::max_element(s.rtid1[0].begin(), s.rtid1[0].end());
  const int64_t R1 = 1 + *std::max_element(s.rtid1[1].begin(), s.rtid1[1].end());
  return s.rtype_of[s.rtid1[0][i] + R0 * (s.rtid1[1][j] + R1 * s.rtid1[2][k])];
}

int build_setup(const int64_t nel[3], const double bounds[6], double young, double poisson, GridSetup& s) {
  for (int d = 0; d < 3; ++d) {
    if (nel[d] < 1) {
      set_error("icc_grid_create: nel must be >= 1 in every direction");
      return ICC_EINVAL;
    }
    if (!(bounds[2 * d + 1] > bounds[2 * d])) {
      set_error("icc_grid_create: empty bounds");
      return ICC_EINVAL;
    }
    s.nel[d] = nel[d];
    s.ncp[d] = nel[d] + 2;
    s.lo[d] = bounds[2 * d];
    s.len[d] = bounds[2 * d + 1] - bounds[2 * d];
  }
  s.nel_tot = s.nel[0] * s.nel[1] * s.nel[2];
  s.ndof_sc = s.ncp[0] * s.ncp[1] * s.ncp[2];
  s.ndof = 3 * s.ndof_sc;
  if (s.ndof >= (int64_t)1 << 31) {
    set_error("icc_grid_create: grid too large for 32-bit DOF indices");
    return ICC_EINVAL;
  }
  s.lambda = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
  s.mu = young / (2.0 * (1.0 + poisson));

  // 1D types per direction.  The reference groups by the rounded key
  // round(v / scale * 1e12) (elasticity.rs:73); rounding noise of the knot
  // spacing can straddle that key at some sizes (nel = 32 or 160 on a length-2
  // side) and split one mathematical type in two.  Here an element joins the
  // first type whose matrices agree to 1e-9 * scale, so such noise never splits
  // a run; every element's matrix still equals the reference's to round-off.
  for (int d = 0; d < 3; ++d) {
    std::vector<std::array<double, 36>> per;
    element_mats_1d(s.nel[d], s.len[d], per);
    double scale = 0.0;
    for (auto& m : per)
      for (double v : m) scale = std::max(scale, std::fabs(v));
    // the reference's 1D types: equal keys round(v / scale * 1e12) over the
    // four matrices in order (elasticity.rs:72-82).  element_mats_1d repeats
    // the reference's arithmetic operation for operation, so the keys are
    // the reference's.
    {
      std::map<std::array<int64_t, 36>, int32_t> keys;
      s.rtid1[d].clear();
      for (auto& m : per) {
        std::array<int64_t, 36> key;
        for (int q = 0; q < 36; ++q) key[q] = (int64_t)std::round(m[q] / scale * 1e12);
        auto it = keys.emplace(key, (int32_t)keys.size()).first;
        s.rtid1[d].push_back(it->second);
      }
    }
    const double tol = 1e-9 * scale;
    s.tid1[d].clear();
    s.mats1[d].clear();
    for (auto& m : per) {
      int found = -1;
      for (size_t t = 0; t < s.mats1[d].size() && found < 0; ++t) {
        double diff = 0.0;
        for (int q = 0; q < 36; ++q) diff = std::max(diff, std::fabs(m[q] - s.mats1[d][t][q]));
        if (diff <= tol) found = (int)t;
      }
      if (found < 0) {
        found = (int)s.mats1[d].size();
        s.mats1[d].push_back(m);
      }
      s.tid1[d].push_back(found);
    }
  }
  int T0 = (int)s.mats1[0].size(), T1 = (int)s.mats1[1].size(), T2 = (int)s.mats1[2].size();
  s.type_of.assign((size_t)T0 * T1 * T2, -1);
  s.type_list.clear();
  for (int64_t k = 0; k < s.nel[2]; ++k)
    for (int64_t j = 0; j < s.nel[1]; ++j)
      for (int64_t i = 0; i < s.nel[0]; ++i) {
        int32_t t0 = s.tid1[0][i], t1 = s.tid1[1][j], t2 = s.tid1[2][k];
        int32_t& slot = s.type_of[t0 + T0 * (t1 + T1 * t2)];
        if (slot < 0) {
          slot = (int32_t)s.type_list.size();
          s.type_list.push_back({t0, t1, t2});
        }
      }

  // the reference's element types: distinct (r0, r1, r2) in element order
  {
    const int64_t R0 = 1 + *std::max_element(s.rtid1[0].begin(), s.rtid1[0].end());
    const int64_t R1 = 1 + *std::max_element(s.rtid1[1].begin(), s.rtid1[1].end());
    const int64_t R2 = 1 + *std::max_element(s.rtid1[2].begin(), s.rtid1[2].end());
    s.rtype_of.assign((size_t)(R0 * R1 * R2), -1);
    s.rtype_internal.clear();
    for (int64_t k = 0; k < s.nel[2]; ++k)
      for (int64_t j = 0; j < s.nel[1]; ++j)
        for (int64_t i = 0; i < s.nel[0]; ++i) {
          int32_t& slot = s.rtype_of[s.rtid1[0][i] + R0 * (s.rtid1[1][j] + R1 * s.rtid1[2][k])];
          if (slot < 0) {
            slot = (int32_t)s.rtype_internal.size();
            s.rtype_internal.push_back(element_type(s, i, j, k));
          }
        }
  }

  // 81x81 matrices: G_ij = (x) over d of m_d^{k(d,i,j)}, direction 0 fastest
  s.kmat.clear();
  for (auto& key : s.type_list) {
    const double* m[3] = {s.mats1[0][key[0]].data(), s.mats1[1][key[1]].data(), s.mats1[2][key[2]].data()};
    auto G = [&](int i, int j, int a, int b) {
      double v = 1.0;
      int ad[3] = {a % 3, (a / 3) % 3, a / 9}, bd[3] = {b % 3, (b / 3) % 3, b / 9};
      for (int d = 0; d < 3; ++d) {
        int kk = (d == i && d == j) ? 3 : (d == i) ? 1 : (d == j) ? 2 : 0;
        v = m[d][kk * 9 + ad[d] * 3 + bd[d]] * v;
      }
      return v;
    };
    std::vector<double> K(NSH * NSH);
    for (int ci = 0; ci < 3; ++ci)
      for (int cj = 0; cj < 3; ++cj)
        for (int a = 0; a < NSC; ++a)
          for (int b = 0; b < NSC; ++b) {
            double v = s.lambda * G(ci, cj, a, b) + s.mu * G(cj, ci, a, b);
            if (ci == cj) {
              double lap = 0.0;
              for (int kk = 0; kk < 3; ++kk) lap += G(kk, kk, a, b);
              v += s.mu * lap;
            }
            K[(ci * NSC + a) * NSH + cj * NSC + b] = v;
          }
    s.kmat.push_back(std::move(K));
  }
  s.kmat_q32.assign(s.kmat.size(), std::vector<double>(NSH * NSH));
  for (size_t t = 0; t < s.kmat.size(); ++t) quantize_kmat(s.kmat[t].data(), 24, s.kmat_q32[t].data());

  build_tiles(s, 0);
  build_split_tiles(s);
  return ICC_OK;
}

// ---- element matrix of the FP32 contraction -----------------------------------
//
// Rounding K_t entry by entry to FP32 breaks K_t t_c = 0 for the rigid
// translations t_c (every row of every 27x27 component block sums to zero) by
// the same residual in every element of the type.  That error slows the FP32
// Jacobi-PCG recurrence of the mixed solver: at 96x48x48 with the cantilever
// traction, 1.58x the FP64 iteration count with plain FP32 rounding against
// 1.38x with the block sums kept (1.15x with the FP64 matrix, the floor the
// FP32 vectors set).  quantize_kmat rounds each row i onto its own
// fixed-point grid q_i = 2^(e_i - bits), max_j |K_ij| < 2^e_i (bits = 24: every
// entry exact in FP32), then moves each block's rounding residual onto the
// row's entry at its own control point (column (c, i mod 27)), which leaves
// every block row sum exactly zero.  The result is symmetric only up to a few
// quanta.  (Symmetric zero-sum roundings on one grid per type, and the same
// on 22-bit grids for the split-FP16 pair, were measured slower: the first
// loses precision on the small entries, the second is limited by the split x.)
double quantize_kmat(const double* K, int bits, double* out) {
  double dev = 0.0;
  const double lim = std::ldexp(1.0, bits);
  for (int i = 0; i < NSH; ++i) {
    const double* row = K + (size_t)i * NSH;
    double* o = out + (size_t)i * NSH;
    double mx = 0.0;
    for (int j = 0; j < NSH; ++j) mx = std::max(mx, std::fabs(row[j]));
    if (mx == 0.0) {
      std::fill(o, o + NSH, 0.0);
      continue;
    }
    int e = 0;
    std::frexp(mx, &e);
    const double q = std::ldexp(1.0, e - bits);  // |row| / q < 2^bits
    double X[NSH];
    for (int j = 0; j < NSH; ++j) X[j] = std::nearbyint(row[j] / q);
    for (int c = 0; c < 3; ++c) {
      double sum = 0.0;
      for (int b = 0; b < NSC; ++b) sum += X[c * NSC + b];
      int j = c * NSC + i % NSC;
      if (std::fabs(X[j] - sum) >= lim) {  // no headroom there: the block's smallest entry takes it
        for (int b = 0; b < NSC; ++b)
          if (std::fabs(X[c * NSC + b]) < std::fabs(X[j])) j = c * NSC + b;
      }
      X[j] -= sum;
    }
    for (int j = 0; j < NSH; ++j) {
      dev = std::max(dev, std::fabs(X[j] - row[j] / q));
      o[j] = X[j] * q;
    }
  }
  return dev;
}

// Tiles of constant-type sub-boxes.  Sub-boxes at least 8 elements long in x
// are cut into x-line tiles: lines of <= 16 elements spread evenly, and up to
// a variant-dependent number of lines per tile (y x z candidates tried in the
// variant's order, the first one that fills the most lines within the
// staging limits wins).  Shorter sub-boxes use tile_shape.  Variants 6 .. 11
// repeat the shapes of 0 .. 5 (the caller partitions them differently).
int n_tile_variants() { return 12; }

void build_tiles(GridSetup& s, int variant) {
  static const int yz_orders[6][10][2] = {
      {{2, 4}, {4, 2}, {1, 8}, {8, 1}, {2, 2}, {1, 4}, {4, 1}, {1, 2}, {2, 1}, {1, 1}},
      {{4, 2}, {2, 4}, {8, 1}, {1, 8}, {2, 2}, {4, 1}, {1, 4}, {2, 1}, {1, 2}, {1, 1}},
      {{1, 8}, {2, 4}, {4, 2}, {8, 1}, {1, 4}, {2, 2}, {4, 1}, {1, 2}, {2, 1}, {1, 1}},
      {{8, 1}, {4, 2}, {2, 4}, {1, 8}, {4, 1}, {2, 2}, {1, 4}, {2, 1}, {1, 2}, {1, 1}},
      {{2, 2}, {1, 4}, {4, 1}, {1, 2}, {2, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}},
      {{1, 4}, {4, 1}, {2, 2}, {1, 2}, {2, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}},
  };
  if (variant < 0 || variant >= n_tile_variants()) variant = 0;
  variant %= 6;
  // Constant-type sub-boxes: runs of equal 1D type along each direction.
  std::vector<std::pair<int64_t, int64_t>> runs[3];
  for (int d = 0; d < 3; ++d) {
    int64_t st = 0;
    for (int64_t e = 1; e <= s.nel[d]; ++e)
      if (e == s.nel[d] || s.tid1[d][e] != s.tid1[d][st]) {
        runs[d].push_back({st, e});
        st = e;
      }
  }
  s.tiles.clear();
  for (auto& rz : runs[2])
    for (auto& ry : runs[1])
      for (auto& rx : runs[0]) {
        int64_t ex = rx.second - rx.first, ey = ry.second - ry.first, ez = rz.second - rz.first;
        int tx, ty, tz, lines = 0;
        if (ex >= 8) {
          // x-lines of <= 16 elements, spread evenly; up to 8 lines per tile
          const int64_t ntx = (ex + 15) / 16;
          tx = (int)((ex + ntx - 1) / ntx);
          lines = 1;
          int best = -1;
          ty = tz = 1;
          for (auto& c : yz_orders[variant]) {
            const int by = (int)std::min<int64_t>(c[0], ey), bz = (int)std::min<int64_t>(c[1], ez);
            const int pts = (tx + 2) * (by + 2) * (bz + 2);
            const int score = by * bz * 1000;
            if (pts <= TILE_MAX_PTS && by * bz * tx <= TILE_MAX_EL && score > best) {
              best = score;
              ty = by;
              tz = bz;
            }
          }
        } else {
          tile_shape(ex, ey, ez, tx, ty, tz);
        }
        int32_t type = element_type(s, rx.first, ry.first, rz.first);
        for (int64_t k0 = rz.first; k0 < rz.second; k0 += tz)
          for (int64_t j0 = ry.first; j0 < ry.second; j0 += ty)
            for (int64_t i0 = rx.first; i0 < rx.second; i0 += tx) {
              Tile t;
              t.i0 = (int32_t)i0;
              t.j0 = (int32_t)j0;
              t.k0 = (int32_t)k0;
              t.nx = (int32_t)std::min<int64_t>(tx, rx.second - i0);
              t.ny = (int32_t)std::min<int64_t>(ty, ry.second - j0);
              t.nz = (int32_t)std::min<int64_t>(tz, rz.second - k0);
              t.type = type;
              t.lines = lines;
              s.tiles.push_back(t);
            }
      }
}

// MFMA work of a tile in element-group pairs (each wave runs its element
// groups two at a time), plus the staging and flush of its points, which
// cost about a pair per 300 staged points.
static double tile_cost(const Tile& t) {
  const int nel = t.nx * t.ny * t.nz;
  const int groups = t.lines ? t.ny * t.nz : (nel + 15) / 16;
  const int pts = (t.nx + 2) * (t.ny + 2) * (t.nz + 2);
  return (double)((groups + 1) / 2) + pts / 300.0;
}

double apply_tile_cost(const Tile& t) { return tile_cost(t); }

// One pass of balance_tiles for an assumed total work `total`: workgroup b
// fills up to the cumulative target (b + 1) total / G, and the tile that
// crosses it is taken whole, left for the next workgroup, or split at the
// leading slab, whichever ends closest to the target.
static double balance_pass(const std::vector<Tile>& in, int G, double total, double (*tile_cost)(const Tile&),
                           std::vector<Tile>& out, std::vector<int32_t>& first) {
  out.clear();
  out.reserve(in.size() + (size_t)G);
  first.assign(1, 0);
  size_t i = 0;
  Tile cur{};
  bool pending = false;  // cur = the remainder of a split tile, next in line
  double cum = 0.0;
  auto advance = [&]() {
    if (pending)
      pending = false;
    else
      ++i;
  };
  for (int b = 0; b < G; ++b) {
    const double target = (double)(b + 1) * total / (double)G;
    while (pending || i < in.size()) {
      const Tile t = pending ? cur : in[i];
      const double c = tile_cost(t);
      if (b == G - 1 || cum + c <= target + 1e-9) {
        out.push_back(t);
        cum += c;
        advance();
        continue;
      }
      // t crosses the target: 0 = leave it, 1 = take it, 2 = split it
      int choice = 0;
      double best = target - cum;
      if (cum + c - target < best) {
        choice = 1;
        best = cum + c - target;
      }
      Tile a_best = t, r_best = t;
      const bool along_z = t.nz > 1;
      const int n = along_z ? t.nz : t.ny;
      for (int k = 1; k < n; ++k) {
        Tile a = t, r = t;
        if (along_z) {
          a.nz = k;
          r.k0 += k;
          r.nz -= k;
        } else {
          a.ny = k;
          r.j0 += k;
          r.ny -= k;
        }
        const double d = std::fabs(cum + tile_cost(a) - target);
        if (d < best) {
          best = d;
          choice = 2;
          a_best = a;
          r_best = r;
        }
      }
      if (choice == 1) {
        out.push_back(t);
        cum += c;
        advance();
      } else if (choice == 2) {
        out.push_back(a_best);
        cum += tile_cost(a_best);
        if (!pending) ++i;
        cur = r_best;
        pending = true;
      }
      break;
    }
    first.push_back((int32_t)out.size());
  }
  return cum;
}

void balance_tile_list(std::vector<Tile>& tiles, int G, double (*tile_cost)(const Tile&), std::vector<int32_t>& first) {
  const std::vector<Tile> in = tiles;
  double total = 0.0;
  for (const Tile& t : in) total += tile_cost(t);
  // splitting adds staging work, which lands on the last workgroup: re-run
  // with the total the previous pass found and keep the partition with the
  // smallest maximum
  std::vector<Tile> out, best_out;
  std::vector<int32_t> fst, best_first;
  double best_max = 0.0;
  for (int it = 0; it < 8; ++it) {
    const double found = balance_pass(in, G, total, tile_cost, out, fst);
    double mx = 0.0;
    for (int b = 0; b < G; ++b) {
      double w = 0.0;
      for (int32_t j = fst[b]; j < fst[b + 1]; ++j) w += tile_cost(out[j]);
      mx = std::max(mx, w);
    }
    if (it == 0 || mx < best_max) {
      best_max = mx;
      best_out = out;
      best_first = fst;
    }
    if (found <= total) break;
    total = found;
  }
  // the plain partition (equal tile counts, no splits) when it is no worse
  double plain_max = 0.0;
  std::vector<int32_t> plain(G + 1);
  for (int b = 0; b <= G; ++b) plain[b] = (int32_t)(((int64_t)b * (int64_t)in.size()) / G);
  for (int b = 0; b < G; ++b) {
    double w = 0.0;
    for (int32_t j = plain[b]; j < plain[b + 1]; ++j) w += tile_cost(in[j]);
    plain_max = std::max(plain_max, w);
  }
  if (plain_max <= best_max) {
    tiles = in;
    first = std::move(plain);
    return;
  }
  tiles = std::move(best_out);
  first = std::move(best_first);
}

void balance_tiles(GridSetup& s, int G, std::vector<int32_t>& first) {
  balance_tile_list(s.tiles, G, tile_cost, first);
}

// Min-max contiguous partition: the smallest bound L such that walking the
// tiles in order, each workgroup takes tiles while its cost stays <= L and
// splits the tile that would cross L (along z, else y; the leading slabs
// that fit stay, the rest carries over), needs at most G workgroups.  Found
// by bisection on L; every workgroup but possibly the last few gets work.
static bool minmax_walk(const std::vector<Tile>& in, int G, double L, double (*cost)(const Tile&),
                        std::vector<Tile>* out, std::vector<int32_t>* first) {
  if (out) {
    out->clear();
    first->assign(1, 0);
  }
  size_t i = 0;
  Tile cur{};
  bool pending = false;
  int used = 0;
  while (pending || i < in.size()) {
    if (used == G) return false;
    double cum = 0.0;
    bool took = false;
    while (pending || i < in.size()) {
      const Tile t = pending ? cur : in[i];
      const double c = cost(t);
      if (cum + c <= L) {
        if (out) out->push_back(t);
        cum += c;
        took = true;
        if (pending)
          pending = false;
        else
          ++i;
        continue;
      }
      // split t: the largest leading part that fits
      const bool along_z = t.nz > 1;
      const int n = along_z ? t.nz : t.ny;
      int kbest = 0;
      for (int k = 1; k < n; ++k) {
        Tile a = t;
        if (along_z)
          a.nz = k;
        else
          a.ny = k;
        if (cum + cost(a) <= L) kbest = k;
      }
      if (kbest > 0) {
        Tile a = t, r = t;
        if (along_z) {
          a.nz = kbest;
          r.k0 += kbest;
          r.nz -= kbest;
        } else {
          a.ny = kbest;
          r.j0 += kbest;
          r.ny -= kbest;
        }
        if (out) out->push_back(a);
        took = true;
        if (!pending) ++i;
        cur = r;
        pending = true;
      } else if (!took) {
        return false;  // a single slab does not fit
      }
      break;
    }
    ++used;
    if (out) first->push_back((int32_t)out->size());
  }
  if (out)
    while ((int)first->size() < G + 1) first->push_back((int32_t)out->size());
  return true;
}

static void partition_minmax(std::vector<Tile>& tiles, int G, double (*cost)(const Tile&), std::vector<int32_t>& first) {
  double lo = 0.0, hi = 0.0;
  for (const Tile& t : tiles) hi += cost(t);
  hi += 1.0;
  for (int it = 0; it < 60 && hi - lo > 1e-3 * hi; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (minmax_walk(tiles, G, mid, cost, nullptr, nullptr))
      hi = mid;
    else
      lo = mid;
  }
  std::vector<Tile> out;
  minmax_walk(tiles, G, hi, cost, &out, &first);
  tiles = std::move(out);
}

void partition_split_tiles(std::vector<Tile>& tiles, int G, std::vector<int32_t>& first) {
  partition_minmax(tiles, G, split_tile_cost, first);
}

// Tiles of the split-FP16 apply.  Constant-type sub-boxes at least 8
// elements long in x are cut into x-line tiles (lines of <= 16 elements
// spread evenly) of up to 16 lines, the (ny, nz) shape with the most lines
// and then the fewest staged points winning; shorter sub-boxes (the
// x-boundary layers) into boxes processed as 16-element groups, the shape
// with the most elements (<= 16 groups) and then the fewest points per
// element winning.  Every shape keeps within the SP_* staging limits.
static bool split_fits(int nx, int ny, int nz) {
  const int px = nx + 2, py = ny + 2, pz = nz + 2, pxp = sp_pxp(nx);
  return px * py * pz <= SP_MAX_PTS && 3 * py * pz <= SP_MAX_ROWS && 3 * py * pz * pxp <= SP_MAX_HPTS &&
         (pxp / 8) * 3 * py * pz <= 256 * SP_TASK_ROUNDS && nx * ny * nz <= SP_MAX_EL;
}

// Cost of a tile in thousands of cycles, fitted to per-tile timestamps of
// the split apply on MI355X (96x48x48, 5.5 % rms): the four MFMA waves take
// the groups in turn, so a tile lasts as long as ceil(groups / 4) group
// steps, plus the flush (per staged point), a fixed part for the set-up,
// the staging hand-over and the barrier, and an extra part for group-mode
// boxes, whose short x-rows stage slowly.
double split_tile_cost(const Tile& t) {
  const int nel = t.nx * t.ny * t.nz;
  const int groups = t.lines ? t.ny * t.nz : (nel + 15) / 16;
  const int pts = (t.nx + 2) * (t.ny + 2) * (t.nz + 2);
  return 4.2 + 3.05 * ((groups + 3) / 4) + 0.0026 * pts + (t.lines ? 0.0 : 2.9);
}

void build_split_tiles(GridSetup& s) {
  std::vector<std::pair<int64_t, int64_t>> runs[3];
  for (int d = 0; d < 3; ++d) {
    int64_t st = 0;
    for (int64_t e = 1; e <= s.nel[d]; ++e)
      if (e == s.nel[d] || s.tid1[d][e] != s.tid1[d][st]) {
        runs[d].push_back({st, e});
        st = e;
      }
  }
  s.stiles.clear();
  for (auto& rz : runs[2])
    for (auto& ry : runs[1])
      for (auto& rx : runs[0]) {
        const int64_t ex = rx.second - rx.first, ey = ry.second - ry.first, ez = rz.second - rz.first;
        int tx, ty = 1, tz = 1, lines;
        if (ex >= 8) {
          const int64_t ntx = (ex + 15) / 16;
          tx = (int)((ex + ntx - 1) / ntx);
          lines = 1;
          int best_lines = 0, best_pts = 0;
          const int mg = std::max(1, std::min(s.split_max_groups, SP_MAX_GROUPS));
          for (int by = 1; by <= std::min<int64_t>(ey, mg); ++by)
            for (int bz = 1; bz <= std::min<int64_t>(ez, mg); ++bz) {
              if (by * bz > mg || !split_fits(tx, by, bz)) continue;
              const int pts = (tx + 2) * (by + 2) * (bz + 2);
              if (by * bz > best_lines || (by * bz == best_lines && pts < best_pts)) {
                best_lines = by * bz;
                best_pts = pts;
                ty = by;
                tz = bz;
              }
            }
        } else {
          tx = (int)ex;
          lines = 0;
          int best_el = 0;
          double best_ppe = 0.0;
          for (int by = 1; by <= std::min<int64_t>(ey, 64); ++by)
            for (int bz = 1; bz <= std::min<int64_t>(ez, 64); ++bz) {
              const int nel = tx * by * bz;
              if ((nel + 15) / 16 > std::max(1, std::min(s.split_max_groups, SP_MAX_GROUPS)) || !split_fits(tx, by, bz))
                continue;
              const double ppe = (double)((tx + 2) * (by + 2) * (bz + 2)) / nel;
              if (nel > best_el || (nel == best_el && ppe < best_ppe)) {
                best_el = nel;
                best_ppe = ppe;
                ty = by;
                tz = bz;
              }
            }
        }
        const int32_t type = element_type(s, rx.first, ry.first, rz.first);
        for (int64_t k0 = rz.first; k0 < rz.second; k0 += tz)
          for (int64_t j0 = ry.first; j0 < ry.second; j0 += ty)
            for (int64_t i0 = rx.first; i0 < rx.second; i0 += tx) {
              Tile t;
              t.i0 = (int32_t)i0;
              t.j0 = (int32_t)j0;
              t.k0 = (int32_t)k0;
              t.nx = (int32_t)std::min<int64_t>(tx, rx.second - i0);
              t.ny = (int32_t)std::min<int64_t>(ty, ry.second - j0);
              t.nz = (int32_t)std::min<int64_t>(tz, rz.second - k0);
              t.type = type;
              t.lines = lines;
              s.stiles.push_back(t);
            }
      }
}

// Consistent traction on the face x = x_max: F(A, c) = t_c * int N_A dA, the
// face integral taken exactly with 3-point Gauss per face element (the trace
// of the p = 2 basis is biquadratic).
void traction_xmax(const GridSetup& s, const double t[3], std::vector<double>& f) {
  f.assign(s.ndof, 0.0);
  std::vector<double> ky = open_knots_p2(s.nel[1]), kz = open_knots_p2(s.nel[2]);
  double xg[3], wg[3];
  gauss(3, xg, wg);
  const int64_t i = s.ncp[0] - 1;
  for (int64_t ez = 0; ez < s.nel[2]; ++ez)
    for (int64_t ey = 0; ey < s.nel[1]; ++ey) {
      const double ay = ky[ey + 2], hy = ky[ey + 3] - ay, az = kz[ez + 2], hz = kz[ez + 3] - az;
      for (int qz = 0; qz < 3; ++qz)
        for (int qy = 0; qy < 3; ++qy) {
          const double yi = ay + 0.5 * (xg[qy] + 1.0) * hy, zi = az + 0.5 * (xg[qz] + 1.0) * hz;
          const double w = (0.5 * wg[qy] * hy * s.len[1]) * (0.5 * wg[qz] * hz * s.len[2]);
          double vy[3], dvy[3], vz[3], dvz[3];
          quad_basis(ky, ey, yi, vy, dvy);
          quad_basis(kz, ez, zi, vz, dvz);
          for (int bz = 0; bz < 3; ++bz)
            for (int by = 0; by < 3; ++by) {
              const int64_t g = i + s.ncp[0] * ((ey + by) + s.ncp[1] * (ez + bz));
              const double wn = w * vy[by] * vz[bz];
              for (int c = 0; c < 3; ++c) f[c * s.ndof_sc + g] += wn * t[c];
            }
        }
    }
}

// Penalty face term on x = x_min: beta * int N_A N_B dA for each component,
// over the face cells (ey, ez) with face_on[ey + nel1 * ez] set, 3-point Gauss
// per face direction.  On x = x_min only the first control-point layer has a
// nonzero trace (open knots), so rows and columns are DOFs of that layer.
void penalty_xmin(const GridSetup& s, double beta, const std::vector<uint8_t>& face_on, std::vector<int64_t>& rows,
                  std::vector<int64_t>& cols, std::vector<double>& vals) {
  std::vector<double> ky = open_knots_p2(s.nel[1]), kz = open_knots_p2(s.nel[2]);
  double xg[3], wg[3];
  gauss(3, xg, wg);
  for (int64_t ez = 0; ez < s.nel[2]; ++ez)
    for (int64_t ey = 0; ey < s.nel[1]; ++ey) {
      if (!face_on[ey + s.nel[1] * ez]) continue;
      const double ay = ky[ey + 2], hy = ky[ey + 3] - ay, az = kz[ez + 2], hz = kz[ez + 3] - az;
      for (int qz = 0; qz < 3; ++qz)
        for (int qy = 0; qy < 3; ++qy) {
          const double yi = ay + 0.5 * (xg[qy] + 1.0) * hy, zi = az + 0.5 * (xg[qz] + 1.0) * hz;
          const double w = (0.5 * wg[qy] * hy * s.len[1]) * (0.5 * wg[qz] * hz * s.len[2]);
          double vy[3], dvy[3], vz[3], dvz[3];
          quad_basis(ky, ey, yi, vy, dvy);
          quad_basis(kz, ez, zi, vz, dvz);
          for (int az2 = 0; az2 < 3; ++az2)
            for (int ay2 = 0; ay2 < 3; ++ay2) {
              const int64_t ga = s.ncp[0] * ((ey + ay2) + s.ncp[1] * (ez + az2));
              for (int bz = 0; bz < 3; ++bz)
                for (int by = 0; by < 3; ++by) {
                  const int64_t gb = s.ncp[0] * ((ey + by) + s.ncp[1] * (ez + bz));
                  const double v = beta * w * (vy[ay2] * vz[az2]) * (vy[by] * vz[bz]);
                  for (int c = 0; c < 3; ++c) {
                    rows.push_back(c * s.ndof_sc + ga);
                    cols.push_back(c * s.ndof_sc + gb);
                    vals.push_back(v);
                  }
                }
            }
        }
    }
}

// Ghost-penalty tables of one direction (p = 2 open uniform knots on a
// segment of `length`): jump[k][t] = jump of the second physical derivative
// of function k - 1 + t across break k (elements k - 1 | k), row 0 unused;
// mass[e][a][b] = int N_a N_b over element e (physical).  The second
// derivative is constant on each element: (N'(end) - N'(start)) / h.
void ghost_tables_1d(int64_t nel, double length, std::vector<double>& jump, std::vector<double>& mass) {
  std::vector<double> knots = open_knots_p2(nel);
  std::vector<std::array<double, 36>> em;
  element_mats_1d(nel, length, em);
  std::vector<double> d2(nel * 3);
  for (int64_t e = 0; e < nel; ++e) {
    const double a = knots[e + 2], b = knots[e + 3];
    double v[3], da[3], db[3];
    quad_basis(knots, e, a, v, da);
    quad_basis(knots, e, b, v, db);
    for (int r = 0; r < 3; ++r) d2[e * 3 + r] = (db[r] - da[r]) / (b - a);
  }
  const double l2 = length * length;
  jump.assign(nel * 4, 0.0);
  mass.assign(nel * 9, 0.0);
  for (int64_t k = 1; k < nel; ++k)
    for (int t = 0; t < 4; ++t) {
      const int64_t g = k - 1 + t;
      const double rv = (g >= k && g <= k + 2) ? d2[k * 3 + (g - k)] : 0.0;
      const double lv = (g >= k - 1 && g <= k + 1) ? d2[(k - 1) * 3 + (g - k + 1)] : 0.0;
      jump[k * 4 + t] = (rv - lv) / l2;
    }
  for (int64_t e = 0; e < nel; ++e)
    for (int m = 0; m < 9; ++m) mass[e * 9 + m] = em[e][m];
}

double design_filter_offsets(const GridSetup& s, double rmin, std::vector<int32_t>& offs, std::vector<double>& ows) {
  const double h[3] = {s.len[0] / (double)s.nel[0], s.len[1] / (double)s.nel[1], s.len[2] / (double)s.nel[2]};
  const double r = rmin * (h[0] + h[1] + h[2]) / 3.0;
  int64_t rmax[3];
  for (int d = 0; d < 3; ++d) rmax[d] = (int64_t)std::ceil(r / h[d]);
  offs.clear();
  ows.clear();
  for (int64_t dk = -rmax[2]; dk <= rmax[2]; ++dk)
    for (int64_t dj = -rmax[1]; dj <= rmax[1]; ++dj)
      for (int64_t di = -rmax[0]; di <= rmax[0]; ++di) {
        const double a = (double)di * h[0], b = (double)dj * h[1], c = (double)dk * h[2];
        const double dist = std::sqrt(a * a + b * b + c * c);
        if (dist <= r) {
          offs.insert(offs.end(), {(int32_t)di, (int32_t)dj, (int32_t)dk});
          ows.push_back(r - dist);
        }
      }
  return r;
}

}  // namespace icc
