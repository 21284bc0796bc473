// CPU oracle: a restatement, in C++, of the reference crate's CPU path for the
// matrix-free p = 2 elasticity hot path (B-spline space, exact element matrices,
// CSR assembly + SpMV, Jacobi PCG, SIMP topopt, immersed classification,
// weighted-quadrature formation, sensitivities).
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (libicc.so, the Python
// package) links or imports this code; only tests/, the smoke check and the
// bench's cpu_baseline leg load it, to check the product against.
//
// Loops keep the reference's order and the file is compiled with
// -ffp-contract=off (Rust never fuses FMAs), so FP64 results are comparable
// bit for bit where the reference is deterministic.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <vector>

namespace orc {

using usize = std::size_t;
using Vec = std::vector<double>;

// ---- rust/src/iga/quadrature.rs -------------------------------------------
void gauss_legendre(usize n, Vec& x, Vec& w);
Vec open_knots(usize nel, usize degree, long regularity);

// ---- rust/src/iga/basis.rs ------------------------------------------------
usize num_basis(const Vec& knots, usize degree);
usize find_span(const Vec& knots, usize degree, double x);
usize basis_funs(const Vec& knots, usize degree, double x, Vec& b, Vec& db);
usize basis_funs_ders(const Vec& knots, usize degree, double x, usize n, long span,
                      std::vector<Vec>& ders);
void basis_dense(const Vec& knots, usize degree, const Vec& xs, Vec& n, Vec& dn);

// ---- rust/src/iga/space.rs ------------------------------------------------
struct Space1d {
  Vec knots;
  usize degree = 0, ndof = 0, nel = 0;
  Vec breaks;
  std::vector<usize> connectivity;
  std::vector<std::vector<usize>> supp;
  usize nquad = 0;
  Vec qn, qw, shape, dshape;
  Space1d() = default;
  Space1d(Vec knots, usize degree, usize nquad);
  usize nsh() const { return degree + 1; }
};

std::vector<usize> unravel(usize idx, const std::vector<usize>& dims);

struct SpaceBox {
  usize dim = 0;
  Vec lo, lengths;
  std::vector<usize> nel_dir;
  usize nel = 0;
  std::vector<usize> degree, nquad;
  usize nqn = 0;
  std::vector<Space1d> univ;
  std::vector<usize> ndof_dir;
  usize ndof_sc = 0, ndof = 0;
  std::vector<usize> nsh_dir;
  usize nsh_sc = 0, nsh = 0;
  std::vector<usize> conn_sc, connectivity;

  static SpaceBox make(const std::vector<std::array<double, 2>>& bounds,
                       const std::vector<usize>& nsub, usize degree);
  Vec element_size() const;
  std::vector<usize> boundary_dofs_sc(usize side) const;
  std::vector<usize> boundary_dofs(usize side) const;
  std::vector<std::array<double, 3>> greville_points() const;
};

// ---- rust/src/iga/elasticity.rs -------------------------------------------
struct ElementMatrices {
  usize nsh = 0;
  std::vector<Vec> types;
  std::vector<usize> type_id;
  const double* of_element(usize e) const { return types[type_id[e]].data(); }
};
ElementMatrices elasticity_element_matrices(const SpaceBox& sp, double lambda, double mu);
void lame(double young, double poisson, double& lambda, double& mu);

// ---- rust/src/iga/fields.rs -----------------------------------------------
void element_quad_points(const SpaceBox& sp, usize e, std::vector<std::array<double, 3>>& x, Vec& w);
void element_shape_functions(const SpaceBox& sp, usize e, Vec& n, Vec& dn);
Vec load_vector(const SpaceBox& sp, const std::function<std::array<double, 3>(const std::array<double, 3>&)>& f);
std::vector<std::array<std::array<double, 3>, 3>> element_vector_gradients(const SpaceBox& sp, const Vec& u, usize e);
std::array<double, 3> eval_vector_at(const SpaceBox& sp, const Vec& u, const std::array<double, 3>& x);

// ---- rust/src/sparse.rs ---------------------------------------------------
struct CsrMatrix {
  usize nrows = 0, ncols = 0;
  std::vector<usize> indptr, indices;
  Vec data;
  static CsrMatrix from_triplets(usize nrows, usize ncols, const std::vector<usize>& rows,
                                 const std::vector<usize>& cols, const Vec& vals);
  void matvec(const double* x, double* y) const;  // row-parallel (OpenMP)
  Vec diagonal() const;
  double get(usize r, usize c) const;
  double frobenius() const;
  usize nnz() const { return data.size(); }
  CsrMatrix plus_triplets(const std::vector<usize>& rows, const std::vector<usize>& cols, const Vec& vals) const;
};

struct AssemblyPattern {
  usize ndof = 0, nsh = 0;
  std::vector<usize> indptr, indices, pos;
  AssemblyPattern() = default;
  AssemblyPattern(usize ndof, const std::vector<usize>& conn, usize nsh, usize nel);
  CsrMatrix assemble(const ElementMatrices& em, const Vec& scale) const;
};

CsrMatrix assemble(usize ndof, const std::vector<usize>& conn, usize nsh, usize nel,
                   const ElementMatrices& em, const double* scale);

// ---- rust/src/solver.rs ---------------------------------------------------
struct PcgResult {
  Vec x;
  usize iters = 0;
  double res = 1.0;
};
PcgResult pcg_solve(usize max_iter, double tol, usize n, const Vec& b, const Vec& diag_a,
                    const std::function<void(const double*, double*)>& matvec);

// ---- rust/src/topopt_iga.rs -----------------------------------------------
struct CantileverCfg {
  usize nel[3] = {16, 8, 4};
  double size[3] = {1.2, 0.6, 0.3};
  usize degree = 2;
  double volfrac = 0.3, penal = 3.0, rmin = 1.5;
  usize max_iter = 40;
  double young = 1.0, poisson = 0.3, emin = 1e-3, move_limit = 0.2, pcg_tol = 1e-10;
};

struct FilterRow {
  std::vector<usize> f;
  Vec w;
};

struct CantileverTopOpt {
  CantileverCfg cfg;
  SpaceBox sp;
  ElementMatrices em;
  Vec force;
  std::vector<uint8_t> free_;
  std::vector<FilterRow> filter;
  Vec filter_sum;
  AssemblyPattern pattern;
  explicit CantileverTopOpt(const CantileverCfg& cfg, bool build_pattern = true);
  CsrMatrix assemble(const Vec& x) const;
  PcgResult solve_state(const Vec& x, const Vec& u0) const;  // x here is the state u
  Vec element_energies(const Vec& u) const;
  struct History {
    Vec compliance, change;
    std::vector<usize> pcg_iterations;
    Vec density;
  };
  History run() const;
};

// Sensitivity filter on a structured element grid (topopt_iga.rs:100-126),
// enumerated by offsets but visiting neighbours in increasing linear index,
// so the rows equal the O(nel^2) scan entry for entry.
std::vector<FilterRow> build_filter_structured(const SpaceBox& sp, double rmin);
std::vector<FilterRow> build_filter_bruteforce(const SpaceBox& sp, double rmin);
// Filter rows on a design domain (topopt_immersed.rs:117-151): radius
// rmin * mean(h), only active neighbours of active cells.
std::vector<FilterRow> build_filter_design(const SpaceBox& sp, double rmin, const std::vector<uint8_t>& active);

// ---- rust/src/topopt2d.rs (fast_sensitivities) ----------------------------
Vec fast_sensitivities(const SpaceBox& sp, const Vec& u, const Vec& x, bool spline, double penal,
                       double emin, double young, double poisson);

// ---- rust/src/cut_cell.rs, rust/src/immersed.rs -----------------------------
struct TriangleMesh3D {
  std::vector<std::array<double, 3>> vertices;
  std::vector<std::array<usize, 3>> triangles;
  static TriangleMesh3D new_box(std::array<double, 3> mn, std::array<double, 3> mx);
};

struct ParityRayCaster {
  std::vector<std::array<std::array<double, 3>, 3>> tris;
  std::vector<std::array<double, 4>> yz_box;
  explicit ParityRayCaster(const TriangleMesh3D& mesh);
  bool is_inside(const std::array<double, 3>& q) const;
};

using InsideFn = std::function<bool(const std::array<double, 3>&)>;
Vec matlab_linspace(double a, double b, usize n);
std::vector<int8_t> classify_background_cells(const InsideFn& inside, const std::array<std::array<double, 2>, 3>& bounds,
                                              const std::array<usize, 3>& res);
void element_weights(const InsideFn& inside, const std::array<std::array<double, 2>, 3>& bounds,
                     const std::array<usize, 3>& res, const std::array<usize, 3>& sub, Vec& weights,
                     std::vector<int8_t>& status);
void legacy_stabilization(const SpaceBox& sp, const std::vector<int8_t>& status, double gamma,
                          std::vector<usize>& ri, std::vector<usize>& ci, Vec& vi);
void ghost_penalty(const SpaceBox& sp, const std::vector<int8_t>& status, double gamma, double young,
                   std::vector<usize>& ri, std::vector<usize>& ci, Vec& vi);
std::array<std::array<double, 2>, 3> padded_bounds(const TriangleMesh3D& mesh, double padding);

struct ImmersedProblem {
  SpaceBox sp;
  std::array<std::array<double, 2>, 3> grid_bounds;
  std::vector<int8_t> status;
  Vec weights;
  CsrMatrix stiffness;
  double stabilization_frobenius = 0.0;
  Vec force;
  std::vector<uint8_t> free_;
};
// stab_kind 0 = None, 1 = Legacy{gamma}; clamp 0 = Bottom, 1 = Top, 2 = Left, -1 = none
ImmersedProblem build_immersed_problem(const TriangleMesh3D& mesh, const std::array<usize, 3>& res, usize degree,
                                       double young, double poisson, int stab_kind, double gamma, double emin,
                                       const std::array<usize, 3>& sub, int clamp, double padding);
struct ImmersedSolution {
  Vec u;
  double compliance = 0;
  usize iterations = 0;
  double residual = 0;
};
ImmersedSolution solve_immersed_problem(const ImmersedProblem& pb, double tol, usize maxit);

// ---- rust/src/immersed_bc.rs ----------------------------------------------
struct SurfacePoint {
  std::array<double, 3> x;
  double w;
  std::array<double, 3> n;
};
struct SurfaceTerms {
  std::vector<usize> rows, cols;
  Vec vals, force;
};
struct SurfaceAccumulator {
  std::map<usize, std::pair<std::vector<usize>, Vec>> blocks;  // key: first basis function of the element
  void add(const std::vector<std::pair<usize, double>>& nz, const double ks[3][3], double w);
  void into_triplets(usize nsc, SurfaceTerms& st) const;
};
// One labelled condition on the triangles `tris` (BoundaryConditionType,
// cad_model.rs): DIRICHLET {components, values}, TRACTION {values = t},
// PRESSURE {pressure}, ROBIN {kn, kt}.
struct SurfaceBc {
  enum Kind { DIRICHLET = 0, TRACTION = 1, PRESSURE = 2, ROBIN = 3 } kind = DIRICHLET;
  std::vector<usize> tris;
  bool components[3] = {true, true, true};
  double values[3] = {0, 0, 0};
  double pressure = 0.0, kn = 0.0, kt = 0.0;
};
std::vector<std::pair<usize, double>> scalar_basis_at(const SpaceBox& sp, const std::array<double, 3>& x);
std::vector<SurfacePoint> surface_quadrature(const TriangleMesh3D& mesh, const std::vector<usize>& tris,
                                             double max_edge);
void surface_add_condition(const SpaceBox& sp, const TriangleMesh3D& mesh, const SurfaceBc& bc, double beta,
                           SurfaceAccumulator& acc, SurfaceTerms& st);
SurfaceTerms surface_boundary_terms(const SpaceBox& sp, const TriangleMesh3D& mesh, const std::vector<SurfaceBc>& bcs,
                                    double penalty_factor, double young);
SurfaceTerms dirichlet_penalty_fn(const SpaceBox& sp, const TriangleMesh3D& mesh, const std::vector<usize>& tris,
                                  double beta,
                                  const std::function<std::array<double, 3>(const std::array<double, 3>&)>& g);
void add_surface_terms(ImmersedProblem& pb, const SurfaceTerms& st);

// ---- rust/src/topopt_immersed.rs ------------------------------------------
struct ImmersedTopOptCfg {
  usize grid_res[3] = {28, 20, 16};
  double volfrac = 0.35, penal = 3.0, rmin = 1.8;
  usize max_iter = 30;
  double young = 1.0, poisson = 0.3, emin = 1e-4;
  int stab_kind = 1;
  double gamma = 0.05;
  double pcg_tol = 1e-8;
  usize pcg_maxit = 20000;
};
struct ImmersedTopOptHistory {
  Vec compliance, volume, change;
  std::vector<usize> pcg_iterations;
  Vec density;
};
ImmersedTopOptHistory immersed_topopt_run(const TriangleMesh3D& brep, const ImmersedTopOptCfg& cfg);

// ---- immersed solve with an analytic geometry (box minus z-cylinders) -----
// Box [lo, hi] minus cylinders parallel to z, cyl = (cx, cy, r) triples;
// inside iff the signed distance is < 0 (strict).
InsideFn box_minus_cylinders(const double lo[3], const double hi[3], const std::vector<double>& cyl);
struct SdfImmersed {
  SpaceBox sp;  // background grid aligned to the box
  std::vector<int8_t> status;
  Vec weights, scale;
  std::vector<uint8_t> active, free_;
  CsrMatrix op;  // sum_e s_e K_e + penalty face term
  Vec force;
  double beta = 0.0;
};
// mode 0: s_e = emin + (1 - emin) w_e on every cell (immersed.rs:484);
// mode 1: s_e = w_e on cells with w_e >= active_min, other cells dropped.
SdfImmersed build_sdf_immersed(const double lo[3], const double hi[3], const std::vector<double>& cyl,
                               const std::array<usize, 3>& res, const std::array<usize, 3>& sub, double young,
                               double poisson, double emin, double active_min, double penalty_factor,
                               const double traction[3], int mode);
ImmersedSolution solve_sdf_immersed(const SdfImmersed& pb, double tol, usize maxit);

// ---- rust/src/wq.rs -------------------------------------------------------
struct WqRules1d {
  Vec points;
  usize ndof = 0;
  std::vector<std::vector<usize>> ind_points, neighbors;
  std::vector<Vec> weights[4];
  Vec b, db;
};
WqRules1d wq_rules_1d(const Vec& knots, usize p, int layout);  // 0 interior, 1 calabro
// density_kind 0 none, 1 element, 2 spline
CsrMatrix wq_stiffness_layout(const SpaceBox& sp, double young, double poisson, int density_kind, const double* x,
                              double penal, double emin, int layout);

}  // namespace orc
