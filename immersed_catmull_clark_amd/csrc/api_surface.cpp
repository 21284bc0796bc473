// Surface boundary conditions on triangle surfaces above the C ABI: the host
// side of rust/src/immersed_bc.rs (surface_quadrature, scalar_basis_at,
// surface_boundary_terms per labelled condition, dirichlet_penalty_fn) and
// CsrMatrix::plus_triplets (rust/src/sparse.rs:98-113) on an operator's
// sparse add-on.
//
// An icc_surface is one SurfaceTerms: its conditions append quadrature points
// (position, weight, the 3x3 stiffness coefficient ks and the load rule of
// the point) in call order.  The first query forms the terms on the device:
// the basis at every point (surf_basis_kernel), the points grouped by element
// (a stable sort by key on the host, so each element keeps the call order),
// one 81 x 81 block and 81 loads per element (surf_block_kernel), then the
// triplets of the blocks in key order with exact zeros dropped
// (SurfaceAccumulator::into_triplets, immersed_bc.rs:82-100).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"
#include "surface_basis.hpp"

#define ICC_API extern "C" __attribute__((visibility("default")))

using namespace icc;

struct icc_surface {
  icc_grid* g = nullptr;
  // points in call order
  std::vector<double> x, w, ks, fs, fv;
  std::vector<int8_t> fmode;
  // formed terms
  bool formed = false;
  int64_t nblocks = 0;
  std::vector<int64_t> rows, cols;
  std::vector<double> vals, force;
};

namespace icc {
const icc_grid* surface_grid(const icc_surface* s) { return s ? s->g : nullptr; }
}  // namespace icc

namespace {

// immersed_bc.rs:15-22: Dunavant degree-4 rule (barycentric point, weight)
const double kDunavant6[6][4] = {
    {0.108103018168070, 0.445948490915965, 0.445948490915965, 0.223381589678011},
    {0.445948490915965, 0.108103018168070, 0.445948490915965, 0.223381589678011},
    {0.445948490915965, 0.445948490915965, 0.108103018168070, 0.223381589678011},
    {0.816847572980459, 0.091576213509771, 0.091576213509771, 0.109951743655322},
    {0.091576213509771, 0.816847572980459, 0.091576213509771, 0.109951743655322},
    {0.091576213509771, 0.091576213509771, 0.816847572980459, 0.109951743655322},
};

struct P3 {
  double v[3];
};

double edge(const P3& p, const P3& q) {
  const double a = p.v[0] - q.v[0], b = p.v[1] - q.v[1], c = p.v[2] - q.v[2];
  return std::sqrt(a * a + b * b + c * c);
}
P3 midpoint(const P3& p, const P3& q) {
  return {{0.5 * (p.v[0] + q.v[0]), 0.5 * (p.v[1] + q.v[1]), 0.5 * (p.v[2] + q.v[2])}};
}

// surface_quadrature (immersed_bc.rs:109-148): each selected triangle is split
// at its edge midpoints, depth first (last pushed child first), until every
// edge is at most max_edge; each leaf gets the 6-point rule with weight
// w_i * area and the unit normal e1 x e2 / |e1 x e2| (degenerate leaves are
// dropped).  out: 7 doubles per point (x, w, n).
int quadrature(const double* vert, int64_t nv, const int64_t* tris, int64_t ntri, const int64_t* sel, int64_t nsel,
               double max_edge, std::vector<double>& out) {
  out.clear();
  for (int64_t i = 0; i < nsel; ++i) {
    if (sel[i] < 0 || sel[i] >= ntri) {
      set_error("surface: triangle index out of range");
      return ICC_EINVAL;
    }
    const int64_t* t = tris + 3 * sel[i];
    P3 abc[3];
    for (int k = 0; k < 3; ++k) {
      if (t[k] < 0 || t[k] >= nv) {
        set_error("surface: vertex index out of range");
        return ICC_EINVAL;
      }
      for (int d = 0; d < 3; ++d) abc[k].v[d] = vert[3 * t[k] + d];
    }
    std::vector<std::array<P3, 3>> stack{{abc[0], abc[1], abc[2]}};
    while (!stack.empty()) {
      const std::array<P3, 3> cur = stack.back();
      stack.pop_back();
      const P3 &a = cur[0], &b = cur[1], &c = cur[2];
      if (std::max(std::max(edge(a, b), edge(b, c)), edge(c, a)) > max_edge) {
        const P3 ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
        stack.push_back({a, ab, ca});
        stack.push_back({ab, b, bc});
        stack.push_back({ca, bc, c});
        stack.push_back({ab, bc, ca});
        continue;
      }
      const double e1[3] = {b.v[0] - a.v[0], b.v[1] - a.v[1], b.v[2] - a.v[2]};
      const double e2[3] = {c.v[0] - a.v[0], c.v[1] - a.v[1], c.v[2] - a.v[2]};
      const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
      const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (nn == 0.0) continue;
      const double area = 0.5 * nn;
      for (const auto& r : kDunavant6) {
        for (int d = 0; d < 3; ++d) out.push_back(r[0] * a.v[d] + r[1] * b.v[d] + r[2] * c.v[d]);
        out.push_back(r[3] * area);
        for (int d = 0; d < 3; ++d) out.push_back(n[d] / nn);
      }
    }
  }
  return ICC_OK;
}

double h_min(const GridSetup& s) {
  double h = 1e300;
  for (int d = 0; d < 3; ++d) h = std::min(h, s.len[d] / (double)s.nel[d]);
  return h;
}

// open_knots (quadrature.rs:51-64) for p = 2
std::vector<double> knots_p2(int64_t nel) {
  std::vector<double> k(3, 0.0);
  for (int64_t i = 1; i < nel; ++i) k.push_back((double)i / (double)nel);
  for (int i = 0; i < 3; ++i) k.push_back(1.0);
  return k;
}

template <typename T>
int dalloc(T** p, size_t n) {
  if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    set_error("hipMalloc failed");
    return ICC_ENOMEM;
  }
  return ICC_OK;
}

struct DevBufs {
  std::vector<void*> v;
  template <typename T>
  int alloc(T** p, size_t n) {
    int rc = dalloc(p, n);
    if (rc == ICC_OK) v.push_back(*p);
    return rc;
  }
  ~DevBufs() {
    for (void* p : v) (void)hipFree(p);
  }
};

#define S_TRY(call)                  \
  do {                               \
    int rc__ = (call);               \
    if (rc__ != ICC_OK) return rc__; \
  } while (0)

// Blocks formed per launch (8 B x 6561 each: 215 MB at most).
constexpr int64_t SURF_CHUNK = 4096;

int form(icc_surface* s) {
  if (s->formed) return ICC_OK;
  const GridSetup& gs = s->g->s;
  const int64_t np = (int64_t)s->w.size();
  const int64_t nsc = gs.ndof_sc;
  s->rows.clear();
  s->cols.clear();
  s->vals.clear();
  s->force.assign(gs.ndof, 0.0);
  s->nblocks = 0;
  if (np == 0) {
    s->formed = true;
    return ICC_OK;
  }
  hipStream_t st = s->g->ctx->stream;
  DevBufs B;
  double *d_x, *d_w, *d_ks, *d_fs, *d_fv, *d_n, *d_k[3];
  int8_t* d_fm;
  int64_t* d_key;
  std::vector<double> kn[3];
  int nk[3];
  for (int d = 0; d < 3; ++d) {
    kn[d] = knots_p2(gs.nel[d]);
    nk[d] = (int)kn[d].size();
    S_TRY(B.alloc(&d_k[d], kn[d].size()));
  }
  S_TRY(B.alloc(&d_x, 3 * np));
  S_TRY(B.alloc(&d_w, np));
  S_TRY(B.alloc(&d_ks, 9 * np));
  S_TRY(B.alloc(&d_fs, np));
  S_TRY(B.alloc(&d_fv, 3 * np));
  S_TRY(B.alloc(&d_fm, np));
  S_TRY(B.alloc(&d_n, 27 * np));
  S_TRY(B.alloc(&d_key, np));
  for (int d = 0; d < 3; ++d)
    ICC_HIP(hipMemcpyAsync(d_k[d], kn[d].data(), kn[d].size() * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(d_x, s->x.data(), 3 * np * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(d_w, s->w.data(), np * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(d_ks, s->ks.data(), 9 * np * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(d_fs, s->fs.data(), np * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(d_fv, s->fv.data(), 3 * np * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(d_fm, s->fmode.data(), np, hipMemcpyHostToDevice, st));
  const double* kp[3] = {d_k[0], d_k[1], d_k[2]};
  ICC_HIP(launch_surface_basis(np, d_x, kp, nk, gs.lo, gs.len, gs.ncp[0], gs.ncp[1], d_key, d_n, st));
  std::vector<int64_t> key(np);
  ICC_HIP(hipMemcpyAsync(key.data(), d_key, np * 8, hipMemcpyDeviceToHost, st));
  ICC_HIP(hipStreamSynchronize(st));
  // points grouped by element block, call order kept inside a block
  std::vector<int64_t> perm(np);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  std::vector<int64_t> seg{0};
  for (int64_t i = 1; i <= np; ++i)
    if (i == np || key[perm[i]] != key[perm[i - 1]]) seg.push_back(i);
  const int64_t nseg = (int64_t)seg.size() - 1;
  s->nblocks = nseg;
  int64_t *d_perm, *d_seg;
  double *d_blk, *d_fe;
  const int64_t chunk = std::min(nseg, SURF_CHUNK);
  S_TRY(B.alloc(&d_perm, np));
  S_TRY(B.alloc(&d_seg, seg.size()));
  S_TRY(B.alloc(&d_blk, chunk * 6561));
  S_TRY(B.alloc(&d_fe, chunk * 81));
  ICC_HIP(hipMemcpyAsync(d_perm, perm.data(), np * 8, hipMemcpyHostToDevice, st));
  ICC_HIP(hipMemcpyAsync(d_seg, seg.data(), seg.size() * 8, hipMemcpyHostToDevice, st));
  std::vector<double> blk(chunk * 6561), fe(chunk * 81);
  for (int64_t c0 = 0; c0 < nseg; c0 += chunk) {
    const int64_t nb = std::min(chunk, nseg - c0);
    ICC_HIP(launch_surface_blocks(nb, d_seg + c0, d_perm, d_n, d_w, d_ks, d_fm, d_fs, d_fv, d_blk, d_fe, st));
    ICC_HIP(hipMemcpyAsync(blk.data(), d_blk, nb * 6561 * 8, hipMemcpyDeviceToHost, st));
    ICC_HIP(hipMemcpyAsync(fe.data(), d_fe, nb * 81 * 8, hipMemcpyDeviceToHost, st));
    ICC_HIP(hipStreamSynchronize(st));
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t k0 = key[perm[seg[c0 + b]]];
      const int64_t fx = k0 % gs.ncp[0], fy = (k0 / gs.ncp[0]) % gs.ncp[1], fz = k0 / (gs.ncp[0] * gs.ncp[1]);
      int64_t idx[27];
      for (int a = 0; a < 27; ++a)
        idx[a] = (fx + a % 3) + gs.ncp[0] * ((fy + (a / 3) % 3) + gs.ncp[1] * (fz + a / 9));
      const double* m = blk.data() + b * 6561;
      for (int ci = 0; ci < 3; ++ci)
        for (int a = 0; a < 27; ++a)
          for (int cj = 0; cj < 3; ++cj)
            for (int bb = 0; bb < 27; ++bb) {
              const double v = m[(ci * 27 + a) * 81 + cj * 27 + bb];
              if (v != 0.0) {
                s->rows.push_back(ci * nsc + idx[a]);
                s->cols.push_back(cj * nsc + idx[bb]);
                s->vals.push_back(v);
              }
            }
      for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 27; ++a) s->force[c * nsc + idx[a]] += fe[b * 81 + c * 27 + a];
    }
  }
  s->formed = true;
  return ICC_OK;
}

bool mesh_ok(const double* v, int64_t nv, const int64_t* t, int64_t nt, const int64_t* sel, int64_t nsel) {
  return v && t && nv > 0 && nt > 0 && (sel || nsel == 0) && nsel >= 0;
}

// Appends the points of one condition; ks_of(n, ks) and load(n, x, fs, fv, mode) per point.
template <typename KS, typename LOAD>
int append(icc_surface* s, const double* vert, int64_t nv, const int64_t* tris, int64_t ntri, const int64_t* sel,
           int64_t nsel, double max_edge, KS ks_of, LOAD load) {
  if (!mesh_ok(vert, nv, tris, ntri, sel, nsel)) return ICC_EINVAL;
  const double me = max_edge > 0.0 ? max_edge : 0.25 * h_min(s->g->s);
  std::vector<double> q;
  S_TRY(quadrature(vert, nv, tris, ntri, sel, nsel, me, q));
  const size_t n = q.size() / 7;
  for (size_t i = 0; i < n; ++i) {
    const double* p = q.data() + 7 * i;
    double k[9], fv[3], fs = 0.0;
    int8_t mode = 2;
    ks_of(p + 4, k);
    load(p + 4, p, fs, fv, mode);
    s->x.insert(s->x.end(), p, p + 3);
    s->w.push_back(p[3]);
    s->ks.insert(s->ks.end(), k, k + 9);
    s->fs.push_back(fs);
    s->fv.insert(s->fv.end(), fv, fv + 3);
    s->fmode.push_back(mode);
  }
  s->formed = false;
  return ICC_OK;
}

}  // namespace

ICC_API int icc_surface_quadrature(const double* vertices, int64_t nv, const int64_t* tris, int64_t ntri,
                                   const int64_t* sel, int64_t nsel, double max_edge, int64_t cap, double* out,
                                   int64_t* npts) {
  if (!npts || !mesh_ok(vertices, nv, tris, ntri, sel, nsel) || !(max_edge > 0.0)) return ICC_EINVAL;
  std::vector<double> q;
  S_TRY(quadrature(vertices, nv, tris, ntri, sel, nsel, max_edge, q));
  *npts = (int64_t)(q.size() / 7);
  if (out) {
    if (cap < *npts) {
      set_error("icc_surface_quadrature: output holds fewer points than the rule produces");
      return ICC_EINVAL;
    }
    std::memcpy(out, q.data(), q.size() * 8);
  }
  return ICC_OK;
}

ICC_API int icc_scalar_basis_at(const icc_grid* g, const double x[3], int64_t* idx, double* val) {
  if (!g || !x || !idx || !val) return ICC_EINVAL;
  const GridSetup& gs = g->s;
  int first[3];
  double v[3][3];
  for (int d = 0; d < 3; ++d) {
    const std::vector<double> kn = knots_p2(gs.nel[d]);
    double xi = (x[d] - gs.lo[d]) / gs.len[d];
    xi = xi < 0.0 ? 0.0 : (xi > 1.0 ? 1.0 : xi);
    const int s = surf_span(kn.data(), (int)kn.size(), xi);
    surf_basis_p2(kn.data(), s, xi, v[d]);
    first[d] = s - 2;
  }
  for (int a = 0; a < 27; ++a) {
    const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
    double w = 1.0;
    w *= v[0][ax];
    w *= v[1][ay];
    w *= v[2][az];
    idx[a] = (first[0] + ax) + gs.ncp[0] * ((first[1] + ay) + gs.ncp[1] * (int64_t)(first[2] + az));
    val[a] = w;
  }
  return ICC_OK;
}

ICC_API int icc_surface_create(icc_grid* g, icc_surface** out) {
  if (!g || !out) return ICC_EINVAL;
  auto* s = new icc_surface();
  s->g = g;
  s->force.assign(g->s.ndof, 0.0);
  *out = s;
  return ICC_OK;
}

ICC_API int icc_surface_destroy(icc_surface* s) {
  delete s;
  return ICC_OK;
}

ICC_API int icc_surface_add(icc_surface* s, int kind, const double* par, const double* vertices, int64_t nv,
                            const int64_t* tris, int64_t ntri, const int64_t* sel, int64_t nsel, double max_edge) {
  if (!s || !par) return ICC_EINVAL;
  switch (kind) {
    case ICC_BC_DIRICHLET: {
      // ks = beta on the constrained diagonal; load beta w N_a value_c (immersed_bc.rs:204-215)
      const double beta = par[6];
      return append(
          s, vertices, nv, tris, ntri, sel, nsel, max_edge,
          [&](const double*, double* k) {
            for (int i = 0; i < 9; ++i) k[i] = 0.0;
            for (int c = 0; c < 3; ++c)
              if (par[c] != 0.0) k[4 * c] = beta;
          },
          [&](const double*, const double*, double& fs, double* fv, int8_t& mode) {
            fs = beta;
            for (int c = 0; c < 3; ++c) fv[c] = par[c] != 0.0 ? par[3 + c] : 0.0;
            mode = 0;
          });
    }
    case ICC_BC_TRACTION:  // load w N_a t_c (immersed_bc.rs:216-222)
      return append(
          s, vertices, nv, tris, ntri, sel, nsel, max_edge,
          [](const double*, double* k) {
            for (int i = 0; i < 9; ++i) k[i] = 0.0;
          },
          [&](const double*, const double*, double& fs, double* fv, int8_t& mode) {
            fs = 1.0;
            for (int c = 0; c < 3; ++c) fv[c] = par[c];
            mode = 0;
          });
    case ICC_BC_PRESSURE:  // load -w N_a p n_c (immersed_bc.rs:223-229)
      return append(
          s, vertices, nv, tris, ntri, sel, nsel, max_edge,
          [](const double*, double* k) {
            for (int i = 0; i < 9; ++i) k[i] = 0.0;
          },
          [&](const double* n, const double*, double& fs, double* fv, int8_t& mode) {
            fs = par[0];
            for (int c = 0; c < 3; ++c) fv[c] = n[c];
            mode = 1;
          });
    case ICC_BC_ROBIN:  // ks = kn n n^T + kt (I - n n^T) (immersed_bc.rs:230-239)
      return append(
          s, vertices, nv, tris, ntri, sel, nsel, max_edge,
          [&](const double* n, double* k) {
            for (int ci = 0; ci < 3; ++ci)
              for (int cj = 0; cj < 3; ++cj) {
                const double delta = ci == cj ? 1.0 : 0.0;
                k[3 * ci + cj] = par[0] * n[ci] * n[cj] + par[1] * (delta - n[ci] * n[cj]);
              }
          },
          [](const double*, const double*, double& fs, double* fv, int8_t& mode) {
            fs = 0.0;
            fv[0] = fv[1] = fv[2] = 0.0;
            mode = 2;
          });
    default:
      set_error("icc_surface_add: unknown condition kind");
      return ICC_EINVAL;
  }
}

ICC_API int icc_surface_add_dirichlet_fn(icc_surface* s, double beta, icc_point_fn gfn, void* user,
                                         const double* vertices, int64_t nv, const int64_t* tris, int64_t ntri,
                                         const int64_t* sel, int64_t nsel, double max_edge) {
  if (!s || !gfn) return ICC_EINVAL;
  // dirichlet_penalty_fn (immersed_bc.rs:161-182): ks = beta I, load beta w N_a g_c(x)
  return append(
      s, vertices, nv, tris, ntri, sel, nsel, max_edge,
      [&](const double*, double* k) {
        for (int i = 0; i < 9; ++i) k[i] = 0.0;
        k[0] = k[4] = k[8] = beta;
      },
      [&](const double*, const double* x, double& fs, double* fv, int8_t& mode) {
        fs = beta;
        gfn(x, fv, user);
        mode = 0;
      });
}

ICC_API int icc_surface_sizes(icc_surface* s, int64_t out[3]) {
  if (!s || !out) return ICC_EINVAL;
  S_TRY(form(s));
  out[0] = (int64_t)s->vals.size();
  out[1] = s->nblocks;
  out[2] = (int64_t)s->w.size();
  return ICC_OK;
}

ICC_API int icc_surface_terms(icc_surface* s, int64_t* rows, int64_t* cols, double* vals, double* force) {
  if (!s) return ICC_EINVAL;
  S_TRY(form(s));
  const size_t n = s->vals.size();
  if (rows) std::memcpy(rows, s->rows.data(), n * 8);
  if (cols) std::memcpy(cols, s->cols.data(), n * 8);
  if (vals) std::memcpy(vals, s->vals.data(), n * 8);
  if (force) std::memcpy(force, s->force.data(), s->force.size() * 8);
  return ICC_OK;
}

// CsrMatrix::plus_triplets (sparse.rs:98-113) then from_triplets (:15-56):
// the existing entries row by row, then the new triplets; each row sorted by
// column (stable, so duplicates are summed in that order) and merged.
ICC_API int icc_op_add_triplets(icc_op* op, int64_t n, const int64_t* rows, const int64_t* cols, const double* vals) {
  if (!op || n < 0 || (n > 0 && (!rows || !cols || !vals))) return ICC_EINVAL;
  const int64_t nr = op->g->s.ndof;
  for (int64_t k = 0; k < n; ++k)
    if (rows[k] < 0 || rows[k] >= nr || cols[k] < 0 || cols[k] >= nr) {
      set_error("icc_op_add_triplets: index out of range");
      return ICC_EINVAL;
    }
  const std::vector<int64_t>& ip0 = op->h_s_indptr;
  const bool have = !ip0.empty();
  std::vector<int64_t> count(nr + 1, 0);
  if (have)
    for (int64_t i = 0; i < nr; ++i) count[i + 1] += ip0[i + 1] - ip0[i];
  for (int64_t k = 0; k < n; ++k) count[rows[k] + 1] += 1;
  for (int64_t i = 0; i < nr; ++i) count[i + 1] += count[i];
  std::vector<int64_t> next(count.begin(), count.end() - 1), ci(count[nr]);
  std::vector<double> cv(count[nr]);
  if (have)
    for (int64_t i = 0; i < nr; ++i)
      for (int64_t k = ip0[i]; k < ip0[i + 1]; ++k) {
        ci[next[i]] = op->h_s_indices[k];
        cv[next[i]++] = op->h_s_data[k];
      }
  for (int64_t k = 0; k < n; ++k) {
    ci[next[rows[k]]] = cols[k];
    cv[next[rows[k]]++] = vals[k];
  }
  std::vector<int64_t> indptr(nr + 1, 0), indices;
  std::vector<double> data;
  indices.reserve(ci.size());
  data.reserve(ci.size());
  std::vector<int64_t> perm;
  for (int64_t r = 0; r < nr; ++r) {
    perm.resize(count[r + 1] - count[r]);
    std::iota(perm.begin(), perm.end(), count[r]);
    std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return ci[a] < ci[b]; });
    int64_t last = -1;
    for (int64_t k : perm) {
      if (last == ci[k]) {
        data.back() += cv[k];
      } else {
        indices.push_back(ci[k]);
        data.push_back(cv[k]);
        last = ci[k];
      }
    }
    indptr[r + 1] = (int64_t)indices.size();
  }
  return icc_op_set_sparse(op, nr, indptr.data(), indices.data(), data.data());
}
