/*
 * icc.h: C ABI of libicc, the MI355X (gfx950) matrix-free elasticity hot path
 * for quadratic (p = 2) trivariate B-spline background grids.
 *
 * Conventions
 *  - Every call returns an int status (ICC_OK = 0).  icc_last_error() gives a
 *    thread-local message for the last failure.
 *  - Plain pointers and sizes only.  Functions without a _dev suffix take HOST
 *    buffers owned by the caller (copied in and out).  _dev functions take
 *    DEVICE pointers (hipMalloc'd or torch CUDA tensors) and are ordered on the
 *    context's HIP stream (icc_ctx_stream); they do not synchronise.
 *  - One context = one device + one HIP stream, driven from one host thread.
 *  - Vectors are component-blocked like the reference: [u0(:); u1(:); u2(:)],
 *    scalar control points lexicographic with direction 0 fastest.
 *  - Reference citations are file:line into dshguilherme/immersed_catmull_clark.
 */
#ifndef ICC_H
#define ICC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum icc_status {
  ICC_OK = 0,
  ICC_EINVAL = 1,      /* bad argument (sizes, null pointers, unsupported degree) */
  ICC_ENOMEM = 2,      /* device or host allocation failed */
  ICC_EHIP = 3,        /* HIP runtime error */
  ICC_EBREAKDOWN = 4,  /* PCG breakdown: p.Ap <= 0 (solver.rs:47-49 stops silently; we report it) */
  ICC_ENOCONV = 5,     /* iteration limit reached before the tolerance */
  ICC_ENOTSUP = 6      /* valid request the library does not implement */
};

/* Arithmetic of the K.u contraction and of the inner solve. */
enum icc_prec {
  ICC_F64 = 0, /* FP64 MFMA (v_mfma_f64_16x16x4), FP64 vectors */
  ICC_F32 = 1, /* FP32 MFMA (v_mfma_f32_16x16x4), FP32 vectors and accumulation */
  ICC_BF16 = 2, /* K and x rounded to BF16, FP32 accumulation
                  (v_mfma_f32_16x16x32_bf16), FP32 vectors (x, y float*);
                  the apply, and the inner solve of icc_solve_ir */
  ICC_F16X = 3  /* split FP16: K (scaled per element type) and x (scaled per
                  tile) each as an FP16 hi + lo pair, three MFMA passes
                  K_hi x_hi + K_hi x_lo + K_lo x_hi (v_mfma_f32_16x16x32_f16),
                  FP32 accumulation, about 2^-21 relative error per element
                  product; FP32 vectors (x, y float*).  The default inner
                  operator of the mixed-precision solvers. */
};

typedef struct icc_ctx icc_ctx;
typedef struct icc_grid icc_grid;
typedef struct icc_op icc_op;
typedef struct icc_filter icc_filter;

const char* icc_last_error(void);
const char* icc_version(void);

/* ---- context --------------------------------------------------------------- */
int icc_ctx_create(int device, icc_ctx** out);
int icc_ctx_destroy(icc_ctx* ctx);
int icc_ctx_sync(icc_ctx* ctx);
/* the hipStream_t every _dev call is ordered on */
void* icc_ctx_stream(icc_ctx* ctx);

/* ---- grid: B-spline space + exact element matrices ------------------------ */
/* Replaces SpaceBox::new (rust/src/iga/space.rs:107-160) for dim = 3,
 * degree = 2, plus lame + elasticity_element_matrices
 * (rust/src/iga/elasticity.rs:30-148,173-177).  bounds = [lo0,hi0,lo1,hi1,lo2,hi2]. */
int icc_grid_create(icc_ctx* ctx, const int64_t nel[3], const double bounds[6], int64_t degree, double young,
                    double poisson, icc_grid** out);
int icc_grid_destroy(icc_grid* g);
/* out[0..10] = nel, ndof_sc, ndof, ntypes, nsh, ncp0, ncp1, ncp2, nel0, nel1, nel2 */
int icc_grid_sizes(const icc_grid* g, int64_t out[11]);
/* connectivity[e*81 + c*27 + a] (space.rs:133-155), bit-exact */
int icc_grid_connectivity(const icc_grid* g, int64_t* out);
/* type_id[e] in first-appearance order, the reference's numbering
 * (elasticity_element_matrices, rust/src/iga/elasticity.rs:30-98): 1D types
 * are the distinct keys round(v / scale * 1e12) of the four 1D element
 * matrices (computed with the reference's arithmetic, so the keys are bit for
 * bit the reference's), element types the distinct combinations in element
 * order.  Where rounding noise of the knot spacing straddles a key (a
 * length-2 side with 32 or 160 elements) the reference splits one
 * mathematical type in two and so does this map; the kernels group by a
 * tolerance and use one matrix for both (icc_grid_element_matrix returns the
 * same matrix for the two). */
int icc_grid_type_id(const icc_grid* g, int64_t* out);
/* row-major 81x81 element matrix of a type (test x trial), type in the
 * numbering of icc_grid_type_id (0 <= type < ntypes of icc_grid_sizes) */
int icc_grid_element_matrix(const icc_grid* g, int64_t type, double* out);
/* The same type map without a grid or a device (host code): *ntypes and, if
 * type_id is not NULL, type_id[nel0 nel1 nel2] (elasticity.rs:30-98, the
 * types of elasticity_element_matrices; they do not depend on the Lame
 * parameters). */
int icc_element_types(const int64_t nel[3], const double bounds[6], int64_t* ntypes, int64_t* type_id);

/* ---- tiling (no counterpart in the reference) ---------------------------------- */
/* The apply walks the grid in tiles of constant element type.  variant selects
 * the x-line tile shape order (variant % 6) and the partition among the
 * persistent workgroups: 0 .. 5 balance the work per workgroup (tiles split
 * where needed), 6 .. 11 give each workgroup an equal number of whole tiles;
 * 0 is the default.  wg_per_cu: apply workgroups per CU (1, 2 or 4).
 * Operators on the grid follow the new tiling. */
int icc_grid_set_tiling(icc_grid* g, int variant, int wg_per_cu);
/* element groups per tile of the split-FP16 apply (1 .. 16, default 16):
 * fewer give more, smaller tiles per CU at more staged halo points per element */
int icc_grid_set_split_tiling(icc_grid* g, int max_groups);
/* out = variant, workgroups per CU, tile count */
int icc_grid_tiling(const icc_grid* g, int64_t out[3]);
/* Times `reps` applies of op (precision prec) for every variant and
 * workgroups-per-CU choice, keeps the fastest and reports it (any out pointer
 * may be NULL; best_us = time per apply).  Operators with a sparse term: ENOTSUP. */
int icc_grid_autotune(icc_op* op, int prec, int reps, int64_t* best_variant, int64_t* best_wg, double* best_us);

/* ---- operator: y = M * sum_e s_e C_e^T K_t(e) C_e * (M x) ------------------ */
/* The matvec closure of topopt_iga.rs:151-159 / immersed.rs:558-566 over the
 * assembled K of sparse.rs:160-175; semantic spec eval_matvec_3d
 * (src/fastformation/topopt_iga_3d.m:252-260). */
int icc_op_create(icc_grid* g, icc_op** out);
int icc_op_destroy(icc_op* op);
/* s[nel]: SIMP / volume-fraction stiffness scale (topopt_iga.rs:132, immersed.rs:484) */
int icc_op_set_scale(icc_op* op, const double* s);
int icc_op_set_scale_dev(icc_op* op, const double* s_dev);
/* free[ndof]: 1 = free DOF, 0 = Dirichlet (NULL = all free) */
int icc_op_set_free_mask(icc_op* op, const uint8_t* free_mask);
/* adds a sparse term S (CSR, ndof x ndof, FP64) to the operator: stabilisation
 * or penalty surface terms (topopt_immersed.rs:174-183, immersed.rs:542-547) */
int icc_op_set_sparse(icc_op* op, int64_t nrows, const int64_t* indptr, const int64_t* indices, const double* data);
/* on = 1: the element term of every FP64 / FP32 apply of op, and of its PCG
 * iterations, runs as per-element outputs (nel x 81, device memory held by
 * op) and a fixed-order gather instead of the atomic tile scatter, so
 * applies and solves repeat bit for bit, like the reference's row-parallel
 * CSR product (rust/src/sparse.rs:60-68); about 1.6x the time of the atomic
 * apply (149 against 94 us at 96x48x48).  ICC_ENOTSUP with the face stabilisation; the BF16 and split-FP16
 * operators are not covered (ICC_EINVAL in the solvers).  on = 0 frees it. */
int icc_op_set_deterministic(icc_op* op, int on);
/* S = S + T for triplets T (rows, cols, vals; duplicates summed): replaces
 * CsrMatrix::plus_triplets (rust/src/sparse.rs:98-113) on the sparse term */
int icc_op_add_triplets(icc_op* op, int64_t n, const int64_t* rows, const int64_t* cols, const double* vals);
/* Adds the legacy cut-cell stabilisation of legacy_stabilization
 * (rust/src/immersed.rs:168-222) as a matrix-free face term: for each face
 * between two active cells (status INSIDE 1 or CUT -1) with at least one CUT,
 * and each component, springs alpha [[1,-1],[-1,1]] between the 9 pairs of
 * control points owned by one of the two cells only, alpha = gamma * mean(h).
 * The reference's block on the 18 shared points touches diagonal entries
 * only and sums to zero in each (round-off), so it is not applied.  The
 * operator's mask applies on both sides (topopt_immersed.rs:174-183).
 * status[nel]; NULL removes the term.  *nfaces (may be NULL) receives the
 * number of stabilised faces. */
int icc_op_set_legacy_stabilization(icc_op* op, const int8_t* status, double gamma, int64_t* nfaces);
/* Adds the consistent ghost penalty of ghost_penalty
 * (rust/src/immersed.rs:234-370) as a matrix-free face term on the same
 * faces: gamma * young * mean(h)^3 * int_F [d_n^2 u].[d_n^2 v], evaluated per
 * face and component as coef (j j^T) (x) M_a (x) M_b with j the jump of the
 * second normal derivative and M_a, M_b the tangential 1D element mass
 * matrices.  Replaces any face term set before; status NULL removes it. */
int icc_op_set_ghost_penalty(icc_op* op, const int8_t* status, double gamma, double young, int64_t* nfaces);
/* Frobenius norm of the assembled face term (ImmersedProblem::
 * stabilization_frobenius, rust/src/immersed.rs:497), summed on the host;
 * 0 without a face term. */
int icc_op_face_frobenius(const icc_op* op, double* fro);

/* y = A x.  Host pointers, FP64 in and out; prec selects the contraction. */
int icc_apply(icc_op* op, int prec, const double* x, double* y);
/* Device pointers: x,y are double* for ICC_F64 and float* for ICC_F32. */
int icc_apply_dev(icc_op* op, int prec, const void* x_dev, void* y_dev);
/* y += A x without zeroing y first (one kernel launch; no sparse term). */
int icc_apply_accumulate_dev(icc_op* op, int prec, const void* x_dev, void* y_dev);

/* Jacobi diagonal d_A = sum_{e ni A} s_e K_t[a,a] (+ diag S), unclamped
 * (sparse.rs:70-78 on the assembled K; topopt_iga_3d.m:119-125). */
int icc_diag(icc_op* op, double* d);
int icc_diag_dev(icc_op* op, double* d_dev);

/* ---- solvers ------------------------------------------------------------------ */
/* PcgSolver::solve (rust/src/solver.rs:14-76) with the operator as the
 * closure: x0 = 0, Jacobi from diag (|d| > 1e-12 else 1), stop at
 * |r|/max(|b|,1e-16) < tol or maxit.  FP64.  Returns ICC_OK on convergence,
 * ICC_EBREAKDOWN / ICC_ENOCONV otherwise (x, iters, res filled in every case). */
int icc_pcg(icc_op* op, const double* b, const double* diag, double tol, int64_t maxit, double* x, int64_t* iters,
            double* res);
int icc_pcg_dev(icc_op* op, const double* b_dev, const double* diag_dev, double tol, int64_t maxit, double* x_dev,
                int64_t* iters, double* res);

/* icc_pcg_dev without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (serving loops, harnesses that
 * forbid hipMalloc).  The solver's vectors, partials and state live in a
 * fixed per-GPU workspace of the library (about 34 MB, held from the
 * library's load until the process ends) sized for at most 1,048,576
 * unknowns; larger operators return ICC_EINVAL.  Same arguments, results and
 * status codes as icc_pcg_dev.  Every call on one GPU shares that workspace:
 * calls that overlap on one GPU (two host threads, two contexts) race. */
int icc_pcg_dev_noalloc(icc_op* op, const double* b_dev, const double* diag_dev, double tol, int64_t maxit,
                        double* x_dev, int64_t* iters, double* res);

/* Mixed-precision iterative refinement: FP64 outer residual r = f - A u with
 * compensated dot products, FP64 update u += d, inner Jacobi PCG in `inner`
 * precision (ICC_F64, ICC_F32, or ICC_BF16: FP32 vectors with the BF16
 * element contraction) solving A d = r to relative `inner_tol`.  The BF16
 * inner operator is perturbed by about 2^-9 relative, so refinement converges
 * only while that perturbation times the condition number stays below 1;
 * otherwise ICC_ENOCONV after max_outer steps.  Stops when the TRUE
 * relative residual |f - A u| / |f| <= tol.  stats (may be NULL), 8 doubles:
 * outer steps, total inner iterations, final true residual, FP64 applies,
 * inner applies, 0, 0, 0.  u_dev in/out (initial guess). */
int icc_solve_ir_dev(icc_op* op, int inner, const double* f_dev, double tol, double inner_tol, int64_t max_outer,
                     int64_t max_inner, double* u_dev, double* stats);
int icc_solve_ir(icc_op* op, int inner, const double* f, double tol, double inner_tol, int64_t max_outer,
                 int64_t max_inner, double* u, double* stats);

/* Mixed precision with reliable updates: FP32 Jacobi PCG (FP32 MFMA apply,
 * FP32 vectors, FP64 reductions) whose residual is replaced by the FP64 true
 * residual f - A u (compensated norms) and whose correction is folded into the
 * FP64 solution each time the FP32 residual drops by `delta`; the search
 * direction is kept across updates.  Stops when the TRUE relative residual
 * <= tol.  stats (may be NULL), 8 doubles: reliable updates, FP32 iterations,
 * final true residual, FP64 applies, FP32 applies, 0, 0, 0. */
int icc_solve_mixed_dev(icc_op* op, const double* f_dev, double tol, double delta, int64_t maxit, double* u_dev,
                        double* stats);
int icc_solve_mixed(icc_op* op, const double* f, double tol, double delta, int64_t maxit, double* u, double* stats);
/* The same solver with the inner contraction in precision `inner`: ICC_F32,
 * ICC_F16X (split FP16, the fastest inner operator with FP32-level accuracy)
 * or ICC_BF16 (converges only where 2^-9 times the condition number is below
 * 1).  The inner vectors are FP32 in every case.  Replaces the solve closure
 * of solve_state (topopt_iga.rs:137-162, solver.rs:14-76) by an FP64-accurate
 * solve to a true residual <= tol. */
int icc_solve_mixed_prec_dev(icc_op* op, int inner, const double* f_dev, double tol, double delta, int64_t maxit,
                             double* u_dev, double* stats);
int icc_solve_mixed_prec(icc_op* op, int inner, const double* f, double tol, double delta, int64_t maxit, double* u,
                         double* stats);

/* ---- sensitivities --------------------------------------------------------- */
/* ee_e = u_e^T K_t(e) u_e, unscaled (topopt_iga.rs:165-184) */
int icc_energies(icc_op* op, const double* u, double* ee);
int icc_energies_dev(icc_op* op, const double* u_dev, double* ee_dev);
/* dc_e = -penal (1 - emin) x_e^(penal-1) ee_e [* w_e if w != NULL]
 * (topopt_iga.rs:199, topopt_immersed.rs:212), fused with the energies. */
int icc_sens_dev(icc_op* op, const double* u_dev, const double* x_dev, double penal, double emin, const double* w_dev,
                 double* dc_dev);
int icc_sens(icc_op* op, const double* u, const double* x, double penal, double emin, const double* w, double* dc);

/* ---- filters -------------------------------------------------------------- */
/* Cartesian filter on element centres, r = rmin * mean(h), weight max(r - d, 0)
 * with d^2 <= r^2 (topopt_iga.rs:100-126).  active may be NULL (all active). */
int icc_filter_create(icc_grid* g, double rmin, icc_filter** out);
/* The design-domain filter of ImmersedTopOpt::new (rust/src/topopt_immersed.rs:117-151):
 * offsets within r = rmin * mean(h), weight r - dist, only active in-range
 * neighbours of active cells (active[nel]); inactive rows are empty (sensitivity
 * and forward give 0 there).  The same sens / fwd / bwd calls apply. */
int icc_filter_create_active(icc_grid* g, double rmin, const uint8_t* active, icc_filter** out);
int icc_filter_destroy(icc_filter* f);
/* sensitivity filter (topopt_iga.rs:200-205) */
int icc_filter_sens_dev(icc_filter* f, const double* x_dev, const double* dc_dev, double* out_dev);
/* density filter forward x~ = H x / Hs and backward g_x = H^T (g / Hs) */
int icc_filter_fwd_dev(icc_filter* f, const double* x_dev, double* out_dev);
int icc_filter_bwd_dev(icc_filter* f, const double* g_dev, double* out_dev);
int icc_filter_sens(icc_filter* f, const double* x, const double* dc, double* out);
int icc_filter_fwd(icc_filter* f, const double* x, double* out);
int icc_filter_bwd(icc_filter* f, const double* g, double* out);

/* ---- weighted-quadrature diagonal ------------------------------------------ */
/* Diagonal of the WQ stiffness (rust/src/wq.rs:329-514, interior layout,
 * element densities: SIMP factor emin + x^penal (1 - emin) sampled at the WQ
 * points, wq.rs:336-358).  x = NULL: uniform density. */
int icc_wq_diag(icc_grid* g, const double* x, double penal, double emin, double* d);
int icc_wq_diag_dev(icc_grid* g, const double* x_dev, double penal, double emin, double* d_dev);

/* ---- WQ stiffness formation ------------------------------------------------ */
/* GPU formation of wq_stiffness / wq_stiffness_layout (rust/src/wq.rs:324-514,
 * p = 2): the WQ elasticity stiffness of the grid's space
 * (its young / poisson), component-blocked and symmetrised 0.5 (K + K^T),
 * with the SIMP factor emin + rho^penal (1 - emin) at the WQ points.
 * density: ICC_DENSITY_NONE (factor 1), ICC_DENSITY_ELEMENT (x[nel],
 * Density::Element) or ICC_DENSITY_SPLINE (x[ndof_sc] control-point
 * densities on the displacement space, Density::Spline).  Stored on the
 * device as a 5x5x5-stencil matrix (9 * 125 * ndof_sc doubles). */
enum { ICC_DENSITY_NONE = 0, ICC_DENSITY_ELEMENT = 1, ICC_DENSITY_SPLINE = 2 };
/* WqLayout (wq.rs:116-126): Interior (the default of wq_stiffness) or Calabro */
enum { ICC_WQ_INTERIOR = 0, ICC_WQ_CALABRO = 1 };
typedef struct icc_wq icc_wq;
/* wq_stiffness: interior layout */
int icc_wq_create(icc_grid* g, icc_wq** out);
/* wq_stiffness_layout (wq.rs:329): layout ICC_WQ_INTERIOR or ICC_WQ_CALABRO.
 * ICC_ENOTSUP for the Calabro layout on a grid with one element in a
 * direction: the middle function's rule there solves a singular system. */
int icc_wq_create_layout(icc_grid* g, int layout, icc_wq** out);
int icc_wq_destroy(icc_wq* m);
int icc_wq_form(icc_wq* m, int density, const double* x, double penal, double emin);
int icc_wq_form_dev(icc_wq* m, int density, const double* x_dev, double penal, double emin);
/* y = K_wq x (unmasked, like CsrMatrix::matvec) */
int icc_wq_apply(icc_wq* m, const double* x, double* y);
int icc_wq_apply_dev(icc_wq* m, const double* x_dev, double* y_dev);
/* CSR copy on the host (the CsrMatrix wq_stiffness returns): rows ndof,
 * columns ascending, every in-grid stencil entry; NULL arrays: *nnz only. */
int icc_wq_csr(icc_wq* m, int64_t* nnz, int64_t* indptr, int64_t* indices, double* data);

/* ---- immersed classification --------------------------------------------- */
/* classify_background_cells + element_weights (rust/src/immersed.rs:88-155)
 * with the analytic box-minus-z-cylinders predicate: inside iff sdf(q) < 0.
 * cyl = [cx, cy, r] * ncyl; grid_bounds[6]; status INSIDE 1 / OUTSIDE 0 / CUT -1. */
int icc_classify_cylinders(icc_ctx* ctx, const double box_lo[3], const double box_hi[3], const double* cyl,
                           int64_t ncyl, const double grid_bounds[6], const int64_t res[3], const int64_t sub[3],
                           int8_t* status, double* weights);
/* The same with the ray-parity predicate of ParityRayCaster
 * (rust/src/immersed.rs:29-76) for a closed triangle mesh: vertices [nv][3],
 * tris [ntri][3] (vertex indices).  Status and weights are bit-identical to
 * the scalar CPU evaluation. */
int icc_classify_mesh(icc_ctx* ctx, const double* vertices, int64_t nv, const int64_t* tris, int64_t ntri,
                      const double grid_bounds[6], const int64_t res[3], const int64_t sub[3], int8_t* status,
                      double* weights);
/* padded_bounds (rust/src/immersed.rs:450-463): the vertices' bounding box
 * grown by padding * extent on each side; out = [x0, x1, y0, y1, z0, z1]. */
int icc_mesh_padded_bounds(const double* vertices, int64_t nv, double padding, double out[6]);

/* ---- structured SIMP cantilever driver ---------------------------------------- */
/* Mirrors CantileverTopOpt (rust/src/topopt_iga.rs:19-235): clamp x = 0,
 * SIMP scale emin + (1 - emin) x^penal, warm-started Jacobi PCG with the
 * rescaled tolerance, element energies, filter, OC bisection (move 0.2),
 * stop when change < 0.01 after 15 iterations. */
typedef struct icc_topopt icc_topopt;
typedef struct icc_topopt_config {
  int64_t nel[3];
  double size[3];
  double volfrac, penal, rmin;
  int64_t max_iter;
  double young, poisson, emin, move_limit, pcg_tol;
  int32_t load;   /* 0: reference nodal load (topopt_iga.rs:79-98); 1: unit downward traction on x = L */
  int32_t filter; /* 0: sensitivity filter (topopt_iga.rs:200-205); 1: density filter with adjoint backward */
  int32_t solver; /* 0: FP64 Jacobi PCG; 1: FP32 PCG with FP64 reliable updates to the same true residual;
                     2: the same with the split-FP16 inner contraction (ICC_F16X) */
  int32_t pad;
} icc_topopt_config;
/* CantileverTopOptConfig::default (topopt_iga.rs:35-51) */
void icc_topopt_config_default(icc_topopt_config* cfg);
int icc_topopt_create(icc_ctx* ctx, const icc_topopt_config* cfg, icc_topopt** out);
int icc_topopt_destroy(icc_topopt* t);
/* out = nel, ndof */
int icc_topopt_sizes(const icc_topopt* t, int64_t out[2]);
int icc_topopt_force(const icc_topopt* t, double* force, uint8_t* free_mask);
/* solve_state (topopt_iga.rs:137-162): x = densities [nel], u0 warm start [ndof] */
int icc_topopt_solve_state(icc_topopt* t, const double* x, const double* u0, double* u, int64_t* iters, double* res);
/* element_energies (topopt_iga.rs:165-184) */
int icc_topopt_element_energies(icc_topopt* t, const double* u, double* ee);
/* compliance c = F.u of the design x and its gradient dc/dx through the
 * filter of the config (density filter: H^T (dc/dx~ / Hs)); u_inout = warm start
 * in, state out.  Tolerance: cfg.pcg_tol. */
int icc_topopt_compliance(icc_topopt* t, const double* x, double* u_inout, double* c, double* dcdx);
typedef void (*icc_report_fn)(int64_t iter, double compliance, double volume, double change, void* user);
/* run (topopt_iga.rs:187-234); arrays sized max_iter (history) and nel (density).
 * The OC bisection runs on the device; ICC_ENOCONV if it has not ended after
 * 4096 steps. */
int icc_topopt_run(icc_topopt* t, icc_report_fn report, void* user, double* compliance, double* change,
                   int64_t* pcg_iterations, double* density, int64_t* n_iter);

/* ---- immersed solve with an analytic geometry ---------------------------------- */
/* build_immersed_problem + solve_immersed_problem (rust/src/immersed.rs:466-570)
 * for a box minus cylinders parallel to z, on a p = 2 background grid aligned to
 * the box:
 *  - classification and 4^3-style sub-cell volume fractions as
 *    icc_classify_cylinders (immersed.rs:79-155);
 *  - mode 0: scale emin + (1 - emin) w on every cell (immersed.rs:484);
 *    mode 1: scale w on cells with w >= active_min, other cells dropped
 *    (active set as topopt_immersed.rs:84);
 *  - DOFs whose support holds no cell of nonzero scale are fixed;
 *  - Dirichlet u = 0 on the face x = box_lo[0] by the penalty
 *    beta int u.v dA, beta = penalty_factor * young / h_min, over the face cells
 *    whose volume cell has nonzero scale (immersed_bc.rs:161-190);
 *  - consistent traction `traction` on the face x = box_hi[0];
 *  - Jacobi PCG on the free DOFs, diagonal clamped at 1e-6 (immersed.rs:555),
 *    stop at |r| / |b| < pcg_tol (solver 0), or the mixed solver of
 *    icc_solve_mixed_prec_dev to a true relative residual <= pcg_tol (solvers 1, 2). */
typedef struct icc_immersed icc_immersed;
typedef struct icc_immersed_config {
  double box_lo[3], box_hi[3];
  int64_t res[3]; /* background elements per direction */
  int64_t sub[3]; /* sub-cells per cell and direction for the volume fractions */
  double young, poisson, emin, active_min, penalty_factor;
  double traction[3];
  double pcg_tol;
  int64_t pcg_maxit; /* 0: 20 * ndof */
  int32_t mode;      /* 0: reference scale on all cells; 1: drop cells with w < active_min */
  int32_t solver;    /* 0: FP64 Jacobi PCG; 1: FP32 PCG with FP64 reliable updates; 2: the same with the
                        split-FP16 inner contraction (ICC_F16X) */
} icc_immersed_config;
/* box [0,2]x[0,1]x[0,1], 32x16x16 elements, 4^3 sub-cells, E = 1, nu = 0.3,
 * emin = 1e-4, active_min = 0.05, penalty factor 1e3, traction (0,-1,0),
 * pcg_tol 1e-8, mode 1, solver 0 */
void icc_immersed_config_default(icc_immersed_config* cfg);
/* cyl = [cx, cy, r] * ncyl */
int icc_immersed_create(icc_ctx* ctx, const icc_immersed_config* cfg, const double* cyl, int64_t ncyl,
                        icc_immersed** out);
int icc_immersed_destroy(icc_immersed* h);
/* out = nel, ndof, active cells, free DOFs, penalty nonzeros; *beta = penalty coefficient */
int icc_immersed_sizes(const icc_immersed* h, int64_t out[5], double* beta);
/* status [nel] (1 inside, 0 outside, -1 cut), weights [nel], active [nel], scale [nel],
 * free mask [ndof], force [ndof] (masked); any pointer may be NULL */
int icc_immersed_problem(const icc_immersed* h, int8_t* status, double* weights, uint8_t* active, double* scale,
                         uint8_t* free_mask, double* force);
/* y = M (K_s + P) M x on host vectors (the solver's operator) */
int icc_immersed_apply(icc_immersed* h, const double* x, double* y);
/* u [ndof] from zero; stats (8 doubles, may be NULL): iterations, recursive
 * relative residual (solver 0) or reliable updates (solvers 1, 2), true relative
 * residual, compliance F.u, FP64 applies, 0, 0, 0.  Returns ICC_ENOCONV /
 * ICC_EBREAKDOWN like icc_pcg with u filled. */
int icc_immersed_solve(icc_immersed* h, double* u, double* stats);
int icc_immersed_solve_dev(icc_immersed* h, double* u_dev, double* stats);

/* ---- immersed problem in a triangle B-rep ----------------------------------- */
/* Mirrors ImmersedOptions / build_immersed_problem / solve_immersed_problem
 * (rust/src/immersed.rs:392-570): p = 2 grid on padded_bounds(mesh, padding)
 * with grid_res cells, ray-parity classification with subcell_res^3 samples,
 * stiffness scale emin + (1 - emin) w on every cell, the stabilisation
 * (ICC_STAB_*, gamma), the clamped background face, the caller's force (or,
 * with force = NULL, -1/count on z at the top-face control points with
 * |i - round(n/2)| <= 2), and Jacobi PCG on the free DOFs with the diagonal
 * clamped at 1e-6 (solver 0) or the mixed solver to the same true residual
 * (solvers 1, 2). */
enum { ICC_CLAMP_NONE = -1, ICC_CLAMP_BOTTOM = 0, ICC_CLAMP_TOP = 1, ICC_CLAMP_LEFT = 2 };
typedef struct icc_immersed_mesh icc_immersed_mesh;
typedef struct icc_immersed_mesh_config {
  int64_t grid_res[3];
  double padding;
  double young, poisson;
  int32_t stabilization; /* ICC_STAB_NONE, ICC_STAB_LEGACY, ICC_STAB_GHOST */
  int32_t clamped_face;  /* ICC_CLAMP_* */
  double gamma;
  double emin;
  int64_t subcell_res[3];
  double pcg_tol;
  int64_t pcg_maxit;
  int32_t solver; /* 0: FP64 Jacobi PCG; 1: FP32 PCG with FP64 reliable updates; 2: the same with the
                     split-FP16 inner contraction (ICC_F16X) */
  int32_t pad;
} icc_immersed_mesh_config;
/* ImmersedOptions::default: 24x16x12, padding 0.05, E = 1, nu = 0.3, ghost
 * penalty 1e-2, emin 1e-4, 4^3 sub-cells, bottom clamped, pcg_tol 1e-8,
 * pcg_maxit 5000, solver 0 */
void icc_immersed_mesh_config_default(icc_immersed_mesh_config* cfg);
/* force [ndof] or NULL (default load) */
int icc_immersed_mesh_create(icc_ctx* ctx, const icc_immersed_mesh_config* cfg, const double* vertices, int64_t nv,
                             const int64_t* tris, int64_t ntri, const double* force, icc_immersed_mesh** out);
int icc_immersed_mesh_destroy(icc_immersed_mesh* h);
/* out = nel, ndof, stabilised faces, free DOFs; grid bounds (either may be NULL) */
int icc_immersed_mesh_sizes(const icc_immersed_mesh* h, int64_t out[4], double grid_bounds[6]);
/* status, weights [nel]; force, free mask [ndof] (any may be NULL) */
int icc_immersed_mesh_problem(const icc_immersed_mesh* h, int8_t* status, double* weights, double* force,
                              uint8_t* free_mask);
/* tolerance, iteration limit and solver of the following solves
 * (solve_immersed_problem's tol and maxit) */
int icc_immersed_mesh_set_solver(icc_immersed_mesh* h, double tol, int64_t maxit, int solver);
/* y = M (K + S) M x (the solver's operator) */
int icc_immersed_mesh_apply(icc_immersed_mesh* h, const double* x, double* y);
/* u [ndof] (zero on fixed DOFs, like ImmersedSolution::u); stats (8, may be
 * NULL): iterations, recursive relative residual (solver 0) or reliable
 * updates (solvers 1, 2), true relative residual, compliance F.u, 0, 0, 0, 0.
 * ICC_ENOCONV / ICC_EBREAKDOWN like icc_pcg with u filled. */
int icc_immersed_mesh_solve(icc_immersed_mesh* h, double* u, double* stats);

/* ---- surface boundary conditions on triangle surfaces ------------------------ */
/* Replaces rust/src/immersed_bc.rs.  surface_quadrature (:109-148): the
 * triangles sel of the mesh, split at edge midpoints until every edge is at
 * most max_edge, with the 6-point Dunavant rule; out (may be NULL): 7 doubles
 * per point (x[3], weight, unit normal[3]), at most cap points; *npts = the
 * point count. */
int icc_surface_quadrature(const double* vertices, int64_t nv, const int64_t* tris, int64_t ntri,
                           const int64_t* sel, int64_t nsel, double max_edge, int64_t cap, double* out,
                           int64_t* npts);
/* scalar_basis_at (:25-50): global indices and values of the 27 nonzero
 * scalar basis functions at x (parameter clamped to the grid) */
int icc_scalar_basis_at(const icc_grid* g, const double x[3], int64_t idx[27], double val[27]);
/* One SurfaceTerms (:152-160) on grid g: stiffness triplets and a load.  Its
 * conditions accumulate in call order; the first icc_surface_sizes /
 * icc_surface_terms forms the per-element surface blocks on the device. */
typedef struct icc_surface icc_surface;
int icc_surface_create(icc_grid* g, icc_surface** out);
int icc_surface_destroy(icc_surface* s);
/* One labelled condition of surface_boundary_terms (:186-247) on the
 * triangles sel (the face; labelling is the caller's):
 *   ICC_BC_DIRICHLET  par = {cx, cy, cz (nonzero = constrained), vx, vy, vz, beta}:
 *                     penalty beta int (u - v).w on the constrained components
 *   ICC_BC_TRACTION   par = {tx, ty, tz}: load int t.w
 *   ICC_BC_PRESSURE   par = {p}: load -int p n.w
 *   ICC_BC_ROBIN      par = {kn, kt}: int (kn n n^T + kt (I - n n^T)) u.w
 * max_edge <= 0 takes the reference's 0.25 h_min.  The reference's beta is
 * penalty_factor * E / h_min. */
enum { ICC_BC_DIRICHLET = 0, ICC_BC_TRACTION = 1, ICC_BC_PRESSURE = 2, ICC_BC_ROBIN = 3 };
int icc_surface_add(icc_surface* s, int kind, const double* par, const double* vertices, int64_t nv,
                    const int64_t* tris, int64_t ntri, const int64_t* sel, int64_t nsel, double max_edge);
/* dirichlet_penalty_fn (:161-182): beta int (u - g(x)).w on the triangles
 * sel, g evaluated on the host at every quadrature point */
typedef void (*icc_point_fn)(const double x[3], double g[3], void* user);
int icc_surface_add_dirichlet_fn(icc_surface* s, double beta, icc_point_fn g, void* user, const double* vertices,
                                 int64_t nv, const int64_t* tris, int64_t ntri, const int64_t* sel, int64_t nsel,
                                 double max_edge);
/* out = triplets, element blocks, quadrature points */
int icc_surface_sizes(icc_surface* s, int64_t out[3]);
/* triplets in SurfaceAccumulator::into_triplets order (blocks by increasing
 * first basis index; exact zeros dropped) and the load [ndof]; any may be NULL */
int icc_surface_terms(icc_surface* s, int64_t* rows, int64_t* cols, double* vals, double* force);
/* ImmersedProblem::add_surface_terms (rust/src/immersed.rs:541-548): K += the
 * triplets (as icc_op_add_triplets), force += the load */
int icc_immersed_mesh_add_surface(icc_immersed_mesh* h, icc_surface* s);
/* replaces the problem's force (ImmersedProblem::force, e.g. zeroed before
 * surface loads are added, immersed_cad_bc.rs:129) */
int icc_immersed_mesh_set_force(icc_immersed_mesh* h, const double* force);
/* the background grid of the problem (owned by h) */
int icc_immersed_mesh_grid(icc_immersed_mesh* h, icc_grid** out);

/* ---- immersed SIMP topology optimisation in a triangle B-rep ------------------ */
/* Mirrors ImmersedTopOpt::{new, run} (rust/src/topopt_immersed.rs:21-247):
 * p = 2 background grid on padded_bounds(brep, 0.05) with grid_res cells,
 * ray-parity classification with 4^3 sub-cells, design domain w >= 0.05,
 * CAD volume sum w over it; stiffness scale w (emin + (1 - emin) x^penal) on
 * the design domain and emin elsewhere, plus the stabilisation; bottom face
 * clamped, load -1/count on z at the top-face control points with
 * |i - round(n/2)| <= 2 (1-based); filter radius rmin * mean(h) over the
 * design domain; per iteration a warm-started Jacobi PCG (diagonal clamped at
 * 1e-6, relative tolerance min(pcg_tol |F| / |F - K u|, 0.5)), compliance F.u,
 * w-weighted sensitivities, the filter and the OC bisection (move 0.2,
 * bounds [0.001, 1], target volfrac * CAD volume).  Runs exactly max_iter
 * iterations like the reference. */
enum { ICC_STAB_NONE = 0, ICC_STAB_LEGACY = 1, ICC_STAB_GHOST = 2 };
typedef struct icc_imm_topopt icc_imm_topopt;
typedef struct icc_imm_topopt_config {
  int64_t grid_res[3];
  double volfrac, penal, rmin; /* rmin in units of the mean cell size */
  int64_t max_iter;
  double young, poisson, emin;
  int32_t stabilization; /* ICC_STAB_NONE, ICC_STAB_LEGACY or ICC_STAB_GHOST */
  int32_t solver;        /* 0: FP64 Jacobi PCG; 1: FP32 PCG with FP64 reliable updates; 2: the same with
                            the split-FP16 inner contraction (ICC_F16X) */
  double gamma;          /* stabilisation parameter */
  double pcg_tol;
  int64_t pcg_maxit;
} icc_imm_topopt_config;
/* grid 28x20x16, volfrac 0.35, penal 3, rmin 1.8, 30 iterations, E = 1,
 * nu = 0.3, emin 1e-4, ghost penalty gamma 1e-2, pcg_tol 1e-8,
 * pcg_maxit 20000, solver 0 (ImmersedTopOptConfig::default) */
void icc_imm_topopt_config_default(icc_imm_topopt_config* cfg);
int icc_imm_topopt_create(icc_ctx* ctx, const icc_imm_topopt_config* cfg, const double* vertices, int64_t nv,
                          const int64_t* tris, int64_t ntri, icc_imm_topopt** out);
int icc_imm_topopt_destroy(icc_imm_topopt* t);
/* out = nel, ndof, design cells, stabilised faces, free DOFs; CAD volume; grid bounds (any may be NULL) */
int icc_imm_topopt_sizes(const icc_imm_topopt* t, int64_t out[5], double* cad_volume, double grid_bounds[6]);
/* status, weights, active [nel]; force, free mask [ndof] (any may be NULL) */
int icc_imm_topopt_problem(const icc_imm_topopt* t, int8_t* status, double* weights, uint8_t* active, double* force,
                           uint8_t* free_mask);
typedef void (*icc_imm_report_fn)(int64_t iter, double compliance, double volume, double change, int64_t pcg_iters,
                                  void* user);
/* Runs max_iter iterations from x = volfrac on the design domain, u = 0.
 * History arrays [max_iter] and density [nel] may be NULL; report (may be
 * NULL) is called after every iteration like the reference's callback. */
int icc_imm_topopt_run(icc_imm_topopt* t, icc_imm_report_fn report, void* user, double* compliance, double* volume,
                       double* change, int64_t* pcg_iterations, double* density);
/* displacement of the last iteration [ndof] */
int icc_imm_topopt_state(icc_imm_topopt* t, double* u);

/* ---- batched instances (uneven sizes) ------------------------------------------ */
/* n structured p = 2 instances, instance k on [0,size_k0]x[0,size_k1]x[0,size_k2]
 * (size NULL: 2 x 1 x 1) with nel_k = nel[3k..3k+2]; default free mask clamps
 * x = 0, default force is zero.  The PCG iterations of every unfinished
 * instance (16 per instance) are captured into one HIP graph, replayed until
 * all have stopped; each instance stops on its own.  No counterpart in the
 * reference (it solves one instance at a time). */
typedef struct icc_batch icc_batch;
int icc_batch_create(icc_ctx* ctx, int64_t n, const int64_t* nel, const double* size, double young, double poisson,
                     icc_batch** out);
int icc_batch_destroy(icc_batch* b);
/* ndof [n], nel [n] (either may be NULL) */
int icc_batch_sizes(const icc_batch* b, int64_t* ndof, int64_t* nel);
/* instance k: element scale [nel_k] (NULL: keep), free mask [ndof_k] (NULL: clamp
 * x = 0), force [ndof_k] (NULL: keep; masked on entry) */
int icc_batch_set(icc_batch* b, int64_t k, const double* scale, const uint8_t* free_mask, const double* force);
/* instance k: force = consistent traction t on the face x = x_max, masked */
int icc_batch_set_traction(icc_batch* b, int64_t k, const double t[3]);
/* icc_grid_autotune on every instance (tile shapes per instance size) */
int icc_batch_autotune(icc_batch* b, int prec, int reps);
/* Solves every instance from u = 0.  prec ICC_F64: Jacobi PCG to |r|/|b| < tol;
 * ICC_F32: FP32 PCG with FP64 reliable updates every `delta` drop, to a true
 * relative residual <= tol.  u (host, may be NULL): instances concatenated.
 * stats (may be NULL) [n][4]: iterations, reliable updates, true relative
 * residual, status code.  Returns the first non-OK instance status, if any. */
int icc_batch_solve(icc_batch* b, int prec, double tol, double delta, int64_t maxit, double* u, double* stats);

/* ---- diagnostics ------------------------------------------------------------- */
/* runs one v_mfma_f64_16x16x4 and one v_mfma_f32_16x16x4 with
 * A[i][k] = i + 16k, B[k][j] = (k == j%4) and writes the 64x4 accumulator
 * registers of each (f64 first, then f32): pins the lane layouts in tests */
int icc_debug_mfma_layout(icc_ctx* ctx, double* out512);
/* Lane-layout check of v_mfma_f32_16x16x32_bf16 as the BF16 apply uses it:
 * out[4 l + r] = accumulator r of lane l for A[i][k] = (7 i + 3 k) % 17 - 8,
 * B[k][n] = (5 k + 11 n) % 13 - 6 loaded as lane l holding A[l%16][8 (l/16) + j]
 * and B[8 (l/16) + j][l%16]; equals D[4 (l/16) + r][l%16] = (A B) if the
 * assumed layout holds. */
int icc_debug_mfma_layout_bf16(icc_ctx* ctx, double* out256);
/* Lane-layout check of the two FP16 MFMAs of the split-FP16 apply, same
 * operands: out512[4 l + r] = register r of lane l after
 * v_mfma_f32_16x16x32_f16 (A[l%16][8 (l/16) + j], B[8 (l/16) + j][l%16]) and
 * out512[256 + 4 l + r] after v_mfma_f32_16x16x16_f16 (A[l%16][4 (l/16) + j],
 * B[4 (l/16) + j][l%16]); both equal D[4 (l/16) + r][l%16] if the assumed
 * layouts hold. */
int icc_debug_mfma_layout_f16(icc_ctx* ctx, double* out512);
/* Host only (no device work): the 81x81 element matrix k rounded as the FP32
 * contraction uses it (bits = 24): each row i on its grid 2^(e_i - bits),
 * |k_i.| < 2^e_i, with every 27x27 component block's row sums exactly zero
 * (so the rigid translations stay in its null space).
 * No counterpart in the reference (its contraction is FP64).  Returns the
 * largest entry deviation in its row's quanta, or -1 for bad arguments. */
double icc_debug_quantize_kmat(const double* k, int bits, double* out);
/* Host only (no device work): the 1D p = 2 WQ rule the formation tables are
 * built from, on the open uniform knot vector of nel elements over [0, 1]
 * (wq_rules_1d_layout, wq.rs:134-139).  which 0: the points; 1..4: the
 * weights W00, W10, W01, W11 of test function i on its points.  Returns the
 * length (out may be NULL), -ICC_EINVAL, or -ICC_ENOTSUP (Calabro, nel = 1). */
int64_t icc_debug_wq_rule(int64_t nel, int layout, int which, int64_t i, double* out);

#ifdef __cplusplus
}
#endif
#endif /* ICC_H */
