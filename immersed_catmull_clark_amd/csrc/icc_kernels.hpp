// Kernel launchers shared between the .hip translation units and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "icc_internal.hpp"

namespace icc {

// Device-resident PCG / refinement scalars.
struct PcgState {
  double rz[2];   // r.D^-1 r by iteration parity
  double norm_b;  // max(|b|, 1e-16); squared |f_free|^2 after residual_kernel
  double res;     // |r| / norm_b after the last update
  double rr;      // r.r (squared |r| after residual_kernel)
  double pap;     // p.Ap of the current iteration
  double tol;     // host-set relative tolerance
  int32_t iters;   // completed iterations (written by pcg_update2)
  int32_t done;    // 0 running, 1 converged, 2 breakdown, 3 iteration limit
  uint32_t ticket;
  int32_t maxit;   // host-set iteration limit
  int32_t it_cur;  // iteration in progress (written by pcg_update1)
  int32_t pad;
};

// Vector kernels use at most VEC_MAX_BLOCKS workgroups: each block of a
// reduction's consumer re-sums all partials, so few partials keep that cheap.
constexpr int VEC_MAX_BLOCKS = 512;

// k_apply.hip
// pst != null: the launch does nothing when pst->done is set
// workgroups of the persistent apply launch (one p.Ap partial per workgroup)
int apply_grid(const icc_grid* g);
hipError_t launch_apply_f64(const icc_grid* g, const icc_op* op, const double* x, double* y, double* pap,
                            const PcgState* pst, hipStream_t st);
hipError_t launch_apply_f32(const icc_grid* g, const icc_op* op, const float* x, float* y, double* pap,
                            const PcgState* pst, hipStream_t st);
// K and x rounded to BF16, FP32 accumulation (v_mfma_f32_16x16x32_bf16); float x, y, scale
hipError_t launch_apply_bf16(const icc_grid* g, const icc_op* op, const float* x, float* y, double* pap,
                             const PcgState* pst, hipStream_t st);
// k_split.hip: K and x split into FP16 hi + lo parts, three MFMA passes,
// FP32 accumulation; float x, y, scale
hipError_t launch_apply_split(const icc_grid* g, const icc_op* op, const float* x, float* y, double* pap,
                              const PcgState* pst, hipStream_t st);
hipError_t launch_layout_probe_f16(double* out, hipStream_t st);
// Batched apply over the concatenated tiles of several instances (one launch):
// y += M A x per instance, p.Ap into pap[TileInst::slot]
struct BatchApply {
  const Tile* tiles = nullptr;
  const int32_t* wg_first = nullptr;
  const TileInst* tinst = nullptr;
  const int32_t* tmask = nullptr;
  const double* kfrag64 = nullptr;  // concatenated A fragments (the tiles' types index them)
  const float* kfrag32 = nullptr;
  int G = 0;
};
hipError_t launch_apply_batch_f64(const BatchApply& b, const double* x, double* y, const double* scale,
                                  const uint8_t* freem, double* pap, hipStream_t st);
hipError_t launch_apply_batch_f32(const BatchApply& b, const float* x, float* y, const float* scale,
                                  const uint8_t* freem, double* pap, hipStream_t st);
// Deterministic apply (opt-in): y += M A M x by per-element outputs E
// [nel * 81] and a fixed-order gather; pap gets det_grid(g) partials
int det_grid(const icc_grid* g);
template <typename T>
hipError_t launch_apply_det(const icc_grid* g, const icc_op* op, const T* x, T* y, double* pap, const PcgState* pst,
                            T* E, hipStream_t st);
hipError_t launch_energies(const icc_grid* g, const double* u, double* ee, const double* x, double penal,
                           double emin, const double* w, double* dc, hipStream_t st);
hipError_t launch_diag_f64(const icc_grid* g, const double* scale, double* d, hipStream_t st);
hipError_t launch_diag_f32(const icc_grid* g, const float* scale, float* d, hipStream_t st);
hipError_t launch_layout_probe(double* out, hipStream_t st);
hipError_t launch_layout_probe_bf16(double* out, hipStream_t st);

// k_surface.hip: basis at surface quadrature points (key = first scalar
// index, 27 values per point) and the per-element 81 x 81 surface blocks and
// 81 loads of runs of points with one key
hipError_t launch_surface_basis(int64_t npts, const double* px, const double* const knots[3], const int nk[3],
                                const double lo[3], const double len[3], int64_t ncp0, int64_t ncp1, int64_t* key,
                                double* nval, hipStream_t st);
hipError_t launch_surface_blocks(int64_t nseg, const int64_t* seg, const int64_t* perm, const double* nval,
                                 const double* pw, const double* ks, const int8_t* fmode, const double* fs,
                                 const double* fv, double* blocks, double* fe, hipStream_t st);

// k_vec.hip
int vec_blocks(int64_t n);
template <typename T>
hipError_t launch_fill(T* y, T v, int64_t n, hipStream_t st);
template <typename T>
hipError_t launch_inv_diag(const double* d, T* dinv, int64_t n, double clamp_min, int do_clamp, hipStream_t st);
template <typename T, typename TB>
hipError_t launch_pcg_init(int64_t n, const TB* b, double bscale, const T* dinv, T* x, T* r, T* p, T* ap,
                           PcgState* st, double* parts, hipStream_t s);
template <typename T>
hipError_t launch_pcg_update(int64_t n, PcgState* st, const double* pap_parts, int npap, T* p, T* ap, T* x, T* r,
                             const T* dinv, double* parts, hipStream_t s);
// Batched PCG updates: one launch of nblk blocks over the segments of several
// instances (concatenated vectors; each instance's blocks, state, p.Ap slots
// and partials as in its single-instance launch, see VecSeg)
struct VecSeg {
  int64_t off, n;    // the instance's vector offset (a multiple of 64) and length
  int32_t blk0, nblk;  // its blocks (vec_blocks(n) of them) and partials at 2 blk0
  int32_t pap0, npap;  // its p.Ap slots
};
template <typename T>
hipError_t launch_pcg_update_batch(int nblk, const VecSeg* segs, const int32_t* seg_of, PcgState* st, const double* pap,
                                   T* p, T* ap, T* x, T* r, const T* dinv, double* parts, hipStream_t s);
hipError_t launch_pdir_update(int64_t n, const PcgState* st, float* p, const float* r, const float* dinv, float* ap,
                              hipStream_t s);
hipError_t launch_reliable_replace(int64_t n, const double* r64, const float* dinv, float* r, float* x, PcgState* st,
                                   double* parts, hipStream_t s);
hipError_t launch_residual(int64_t n, const double* f, const double* ap, const uint8_t* freem, double* r,
                           PcgState* st, double* parts, hipStream_t s);
template <typename T>
hipError_t launch_ir_update(int64_t n, double* u, const T* d, double scale, hipStream_t s);
// y = S x (overwrites y), masked; rows[nrows]: the rows with entries; rb
// blocks for those rows (one partial each in parts, if not null), the rest of
// the launch zeroes the other rows; y_zeroed: y is already zero (the PCG
// iteration's Ap, cleared by the previous update), so the zeroing is skipped
template <typename T, typename TD>
hipError_t launch_csr_spmv(int64_t n, const int32_t* indptr, const int32_t* indices, const TD* data,
                           const int32_t* rows, int64_t nrows, int rb, const T* x, T* y, const uint8_t* freem,
                           double* parts, bool y_zeroed, hipStream_t s);

hipError_t launch_clamp_min(int64_t n, const double* a, double lo, double* b, hipStream_t st);
hipError_t launch_simp_scale(int64_t n, const double* x, double penal, double emin, const double* w,
                             const uint8_t* active, double* s, hipStream_t st);
hipError_t launch_masked_add(int64_t n, const double* u0, const double* du, const uint8_t* freem, double* u,
                             hipStream_t st);

// k_misc.hip
hipError_t launch_filter(const icc_filter* f, int mode, const double* x, const double* v, double* out,
                         hipStream_t st);
hipError_t launch_to_f32(const double* a, float* b, int64_t n, hipStream_t st);
hipError_t launch_wq_diag(const icc_grid* g, const double* tab, const double* x, double penal, double emin,
                          double* d, hipStream_t st);

// k_noalloc.hip: fixed workspace of icc_pcg_dev_noalloc, in doubles:
// [PcgState (NA_STATE) | 2 * VEC_MAX_BLOCKS vector partials | NA_PAP p.Ap
// partials | r | p | Ap | D^-1 (NA_MAX_NDOF each)], every part 256-byte aligned
constexpr int64_t NA_MAX_NDOF = 1 << 20;
constexpr int64_t NA_STATE = 32;
constexpr int64_t NA_PAP = 4096;
constexpr int64_t NA_SCRATCH_DOUBLES = NA_STATE + 2 * VEC_MAX_BLOCKS + NA_PAP + 4 * NA_MAX_NDOF;
static_assert(sizeof(PcgState) <= NA_STATE * 8 && (2 * VEC_MAX_BLOCKS) % 32 == 0 && NA_PAP % 32 == 0, "workspace alignment");
// looks up this device's workspace (device current; once per device)
hipError_t noalloc_init(int device);
// the address found by noalloc_init, or null before it ran or if it failed
double* noalloc_scratch(int device);

// k_faces.hip: matrix-free face stabilisation, at most FACE_MAX_BLOCKS p.Ap partials
constexpr int FACE_MAX_BLOCKS = 512;
constexpr int FACE_LEGACY = 1, FACE_GHOST = 2;
constexpr int FACE_BLK = 4;  // faces are grouped by FACE_BLK^3 element blocks
int face_blocks(const icc_op* op);
template <typename T>
hipError_t launch_faces(const icc_op* op, const T* x, T* y, double* pap, const PcgState* pst, hipStream_t st);

// k_topopt.hip: immersed SIMP design update (filter on the design domain, OC bisection)
struct OcState {
  double l1, l2;
  int32_t done;   // 1 once the bisection interval test fails
  int32_t steps;  // bisection steps evaluated
};
int topopt_blocks(int64_t nel);  // blocks (= partials) of the OC and stats kernels
hipError_t launch_imm_filter(int mode, const int64_t nel[3], const int32_t* off, const double* ow, int noff,
                             const uint8_t* act, const double* x, const double* v, const double* hs, double* out,
                             hipStream_t st);
hipError_t launch_oc_init(int64_t nel, const uint8_t* act, const double* dc, const double* xold, double* xnew,
                          double* parts_max, hipStream_t st);
hipError_t launch_oc_step(int k, int64_t nel, const uint8_t* act, const double* w, const double* dvs,
                          const double* dc, const double* xold, double* xnew, double target, double move,
                          const double* parts_max, double* parts, OcState* st, hipStream_t s);
hipError_t launch_topopt_stats(int64_t n, const double* f, const double* u, int64_t nel, const uint8_t* act,
                               const double* w, const double* x, const double* xold, double* parts, hipStream_t st);
hipError_t launch_fill_active(int64_t nel, const uint8_t* act, double v, double* x, hipStream_t st);

// k_wq.hip: WQ stiffness formation (stencil storage [9 * 125][ndof_sc]) and its product;
// np[d] WQ points per direction, pstart / pbin per-direction tables (concatenated)
hipError_t launch_wq_simp(const icc_grid* g, int kind, const int np[3], const double* x, double penal, double emin,
                          const int* pbin, const int* pfirst, const double* pval, double* simp, hipStream_t st);
hipError_t launch_wq_form(const icc_grid* g, const int np[3], const int* pstart, const double* tabs,
                          const double* simp, double* S, hipStream_t st);
hipError_t launch_wq_spmv(const icc_grid* g, const double* S, const double* x, double* y, hipStream_t st);

// k_immersed.hip
hipError_t launch_classify_cylinders(const double box_lo[3], const double box_hi[3], const double* cyl_dev,
                                     int64_t ncyl, const double gb[6], const int64_t res[3], const int64_t sub[3],
                                     uint8_t* inside_dev, int8_t* status_dev, double* w_dev, hipStream_t st);
// triangle records for the ray-parity predicate: a, b, c (xyz) and the y-z box
constexpr int MESH_TRI_STRIDE = 13;
hipError_t launch_classify_mesh(const double* tri_dev, int64_t ntri, const double gb[6], const int64_t res[3],
                                const int64_t sub[3], uint8_t* inside_dev, int8_t* status_dev, double* w_dev,
                                hipStream_t st);

// wq.cpp
void wq_tables_1d(int64_t nel, std::vector<double>& A0, std::vector<double>& A1);

}  // namespace icc
