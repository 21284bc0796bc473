// Internal types of libicc (host side).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/icc.h"

namespace icc {

constexpr int NSH = 81;      // 3 components x 27 control points per p = 2 element
constexpr int NSC = 27;      // scalar functions per element
constexpr int MT = 6;        // 16-row M tiles covering 81 test rows (96 padded)
constexpr int KS = 21;       // 4-deep K steps covering 81 trial columns (84 padded)
constexpr int TILE_MAX_EL = 128;   // elements per tile (8 waves x 16)
constexpr int TILE_MAX_PTS = 640;  // control points staged per tile
constexpr int APPLY_MAX_WG_PER_CU = 4;  // apply workgroups per CU, at most
// Split-FP16 apply (k_split.hip): one workgroup of 4 MFMA + 4 memory waves
// per CU; tiles of <= 16 element groups (x-lines or 16-element groups).
constexpr int SP_THREADS = 512;
constexpr int SP_MAX_GROUPS = 16;   // element groups per tile
constexpr int SP_MAX_EL = 256;      // elements per tile
constexpr int SP_MAX_PTS = 720;     // staged control points per tile
constexpr int SP_MAX_HPTS = 4096;   // staged FP16 points per part, rows padded to 8 points
constexpr int SP_MAX_ROWS = 720;    // staged (component, y, z) rows per tile (3 (ny + 2) (nz + 2))
constexpr int SP_TASK_ROUNDS = 3;   // staging tasks per memory thread
constexpr int SP_SCALE_BITS = 15;   // K_t and x scaled to a maximum in [2^14, 2^15)

// Tile of a constant-type element sub-box: elements [i0, i0+nx) x [j0, j0+ny) x [k0, k0+nz).
// lines = 1: the tile is processed as x-lines of nx <= 16 elements (one MFMA
// N group per line); lines = 0: as groups of 16 consecutive tile elements.
struct Tile {
  int32_t i0, j0, k0, nx, ny, nz, type, lines;
};
// Instance of a tile of a batched (multi-instance) apply: the instance's
// offsets in the concatenated DOF and element vectors, its grid extents, and
// the p.Ap slot of (workgroup run, instance).
struct TileInst {
  int32_t dbase, ebase, ncp0, ncp1, ndof_sc, nel0, nel1, slot;
};

#if defined(__HIP__) || defined(__HIPCC__)
#define ICC_HD __host__ __device__
#else
#define ICC_HD
#endif

// Padded row length (points) of a split-apply tile's FP16 rows: at least
// nx + 3, the last element's 4-point run (x .. x + 3), rounded up to 8.
ICC_HD inline int sp_pxp(int nx) { return (nx + 3 + 7) & ~7; }

// Row of K held by accumulator register r (0..3) of row group kq (0..3) in
// M tile w (0..5), or -1 for padding.  Rows are permuted so that registers
// 0..2 hold the ax = 0, 1, 2 rows of one (component, ay, az) triplet
// t = 4 w + kq (t < 24); register 3 holds ax = w of triplet 24 + kq for
// w, kq < 3.  Within one accumulator register (one ds_add instruction) the
// four row groups then hold four different triplets.
ICC_HD inline int perm_row(int w, int kq, int r) {
  int t, ax;
  if (r < 3) {
    t = 4 * w + kq;
    ax = r;
  } else {
    if (w >= 3 || kq >= 3) return -1;
    t = 24 + kq;
    ax = w;
  }
  const int c = t / 9, rem = t - 9 * c, ay = rem % 3, az = rem / 3;
  return c * 27 + ax + 3 * ay + 9 * az;
}

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// the grid an icc_surface belongs to (api_surface.cpp)
const icc_grid* surface_grid(const icc_surface* s);

#define ICC_HIP(call)                                 \
  do {                                                \
    hipError_t e__ = (call);                          \
    if (e__ != hipSuccess) return icc::hip_fail(e__, #call); \
  } while (0)

// Host-side product setup: p = 2 B-spline space on a box, 1D element matrices,
// element types and the 81x81 matrices.  Independent of the test oracle.
struct GridSetup {
  int64_t nel[3];
  int64_t ncp[3];
  double lo[3], len[3];
  int64_t nel_tot, ndof_sc, ndof;
  double lambda, mu;
  // per direction: 1D type of each element and the 1D matrices of each type
  std::vector<int32_t> tid1[3];
  std::vector<std::array<double, 36>> mats1[3];  // [k][a][b] with k = 00,10,01,11
  // element types (first-appearance order) as (t0, t1, t2)
  std::vector<std::array<int32_t, 3>> type_list;
  std::vector<int32_t> type_of;  // index t0 + T0*(t1 + T1*t2) -> element type
  // the reference's numbering (rounded keys, elasticity.rs:73-98), reported by
  // icc_grid_type_id: 1D types, combination -> type, type -> element type above
  std::vector<int32_t> rtid1[3];
  std::vector<int32_t> rtype_of;  // index r0 + R0*(r1 + R1*r2)
  std::vector<int32_t> rtype_internal;
  std::vector<std::vector<double>> kmat;    // [ntypes][81*81]
  std::vector<std::vector<double>> kmat_q32;  // kmat on 24-bit row grids with exact zero block row sums (FP32 apply)
  std::vector<Tile> tiles;
  std::vector<Tile> stiles;  // split-FP16 apply tiles
  int split_max_groups = 16;  // element groups per split tile (<= SP_MAX_GROUPS)
};

int build_setup(const int64_t nel[3], const double bounds[6], double young, double poisson, GridSetup& s);
// K (81x81) with each row rounded onto its grid 2^(e_i - bits), |K_i.| <
// 2^e_i, and every 27x27 component block's row sums exactly zero; returns the
// largest entry deviation in its row's quanta.
double quantize_kmat(const double* K, int bits, double* out);
// (Re)builds s.tiles with tiling variant v (0 .. n_tile_variants() - 1; the
// shape order is v % 6, and v < 6 asks for the balanced partition).
void build_tiles(GridSetup& s, int variant);
// Tiles of the split-FP16 apply (s.stiles), within the SP_* limits.
void build_split_tiles(GridSetup& s);
// Work of an apply tile (element-group pairs plus staging and flush).
double apply_tile_cost(const Tile& t);
// Work of a split-apply tile (its element groups plus a per-tile overhead).
double split_tile_cost(const Tile& t);
// Re-partitions s.tiles for G persistent workgroups: walks the tiles in
// order, giving workgroup b a contiguous run that ends as close as possible
// to the cumulative target (b + 1) W / G, where W is the total work; the tile
// that crosses a target may be split along z (else y).  Falls back to equal
// whole-tile counts when that partition is no worse.  first[b] .. first[b + 1]
// are workgroup b's tiles (G + 1 entries).
void balance_tiles(GridSetup& s, int G, std::vector<int32_t>& first);
// Split-apply partition: contiguous runs of minimal maximum cost
// (split_tile_cost), tiles split where that lowers it.
void partition_split_tiles(std::vector<Tile>& tiles, int G, std::vector<int32_t>& first);
// The same partition of an arbitrary tile list under a cost model.
void balance_tile_list(std::vector<Tile>& tiles, int G, double (*cost)(const Tile&), std::vector<int32_t>& first);
int n_tile_variants();
int32_t element_type(const GridSetup& s, int64_t i, int64_t j, int64_t k);
// the element's type in the reference's numbering
int32_t reference_type(const GridSetup& s, int64_t i, int64_t j, int64_t k);

void ghost_tables_1d(int64_t nel, double length, std::vector<double>& jump, std::vector<double>& mass);
// Design-domain filter offsets (topopt_immersed.rs:117-131): radius
// r = rmin * mean(h); every (di, dj, dk), dk outermost, with |offset| <= r and
// weight r - |offset|.  Returns r.
double design_filter_offsets(const GridSetup& s, double rmin, std::vector<int32_t>& offs, std::vector<double>& ows);

// WQ rules for p = 2 (wq.cpp)
struct WqRulesP2 {
  int nel = 0, ndof = 0, npt = 0;
  std::vector<double> pts;                       // WQ points, ascending
  std::vector<int> bin;                          // element of each point (a knot goes to its right element)
  std::vector<std::vector<int>> ind;             // points of each test function (contiguous)
  // weights on ind of the four contractions: W00 (trial B, test B), W10 (trial B,
  // test B'), W01 (trial B', test B), W11 (trial B', test B')
  std::vector<std::vector<double>> w0, w1, w2, w3;
  std::vector<double> b, db;                     // [npt][ndof] basis values / parametric derivatives
};
enum { WQ_LAYOUT_INTERIOR = 0, WQ_LAYOUT_CALABRO = 1 };
void wq_rules_p2(int64_t nel, WqRulesP2& r);
void wq_rules_calabro_p2(int64_t nel, WqRulesP2& r);
// Formation tables M[((i * 4 + t) * 5 + jr) * 9 + qc] of one direction and the
// first point pstart[i] of each row (rows have at most 9 points).
void wq_form_tables(const WqRulesP2& r, std::vector<double>& M, std::vector<int>& pstart);
void wq_form_tables_1d(int64_t nel, std::vector<double>& M);

}  // namespace icc

struct icc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
};

struct icc_grid {
  icc_ctx* ctx = nullptr;
  icc::GridSetup s;
  int ntypes = 0;
  // device data
  double* d_kfrag64 = nullptr;  // [ntypes][MT][KS][64]
  float* d_kfrag32 = nullptr;   // [ntypes][MT][KS][64]
  uint16_t* d_kfragb = nullptr; // [ntypes][MT][KSB = 3][64][8] BF16 bits
  double* d_kdiag = nullptr;    // [ntypes][81]
  int32_t* d_tid1 = nullptr;    // [3][maxnel] 1D types
  int32_t* d_typeof = nullptr;  // [T0*T1*T2]
  icc::Tile* d_tiles = nullptr;
  int ntiles = 0;
  double* d_mats1 = nullptr;  // [3][T][36]
  double* d_wq = nullptr;     // WQ tables [A0x, A1x, A0y, A1y, A0z, A1z], built on first use
  int maxnel = 0;
  int T[3] = {0, 0, 0};
  int tile_variant = 0;  // build_tiles variant of s.tiles
  int wg_per_cu = 2;     // apply workgroups per CU (1 .. APPLY_MAX_WG_PER_CU)
  int apply_G = 1;       // persistent apply workgroups
  int32_t* d_wgfirst = nullptr;  // [apply_G + 1] tile range of each workgroup (balance_tiles)
  std::vector<struct icc_op*> ops;  // operators on this grid (their tile masks follow the tiling)
  // split-FP16 apply
  uint4* d_afrag16 = nullptr;     // [ntypes][2 halves][3 M tiles][4 K steps][64 lanes][hi, lo]
  int32_t* d_kexp16 = nullptr;    // [ntypes] K_t was scaled by 2^-kexp
  icc::Tile* d_stiles = nullptr;
  int nstiles = 0;
  int split_G = 1;                // workgroups (one per CU)
  int32_t* d_swgfirst = nullptr;  // [split_G + 1]
};

struct icc_op {
  icc_grid* g = nullptr;
  double* d_scale = nullptr;    // [nel]
  float* d_scale32 = nullptr;   // [nel]
  uint8_t* d_free = nullptr;    // [ndof], null = all free
  int32_t* d_tmask = nullptr;   // [ntiles]: 1 if the tile stages a fixed DOF
  int32_t* d_stmask = nullptr;  // [nstiles]: the same for the split-apply tiles
  std::vector<uint8_t> h_free;  // host copy of the free mask
  // optional sparse add-on (CSR, int32 indices)
  int64_t s_nnz = 0;
  int32_t* d_s_indptr = nullptr;
  int32_t* d_s_indices = nullptr;
  double* d_s_data = nullptr;
  float* d_s_data32 = nullptr;
  double* d_s_diag = nullptr;
  int32_t* d_s_rows = nullptr;  // {row, first entry, end, 0} per row with entries
  int64_t s_nrows = 0;
  std::vector<int64_t> h_s_indptr, h_s_indices;  // host copy of the add-on (plus_triplets)
  std::vector<double> h_s_data;
  // optional matrix-free cut-cell face stabilisation (k_faces.hip)
  int f_kind = 0;               // FACE_LEGACY or FACE_GHOST
  int64_t f_n = 0;              // faces
  int32_t* d_faces = nullptr;   // [f_n]: e1 * 4 + dir, neighbour e2 = e1 + unit step along dir, by block
  int64_t f_nblk = 0;           // element blocks holding faces
  int32_t* d_fblk = nullptr;    // [f_nblk] block index (x fastest)
  int32_t* d_fptr = nullptr;    // [f_nblk + 1] face range of each block
  double f_alpha = 0.0;         // legacy: alpha = gamma * mean(h); ghost: gamma E mean(h)^3
  double* d_f_diag = nullptr;   // [ndof] diagonal of the face term
  double* d_ftab = nullptr;     // ghost: jump tables then mass tables
  int ftab_joff[3] = {0, 0, 0}, ftab_moff[3] = {0, 0, 0}, ftab_mass = 0;
  std::vector<int32_t> h_faces;  // host copies (Frobenius norm of the face term)
  std::vector<double> h_ftab;
  // deterministic mode: element outputs [nel * 81] of the two-pass apply
  int det = 0;
  void* d_E = nullptr;
  // scratch
  double* d_partials = nullptr;  // reduction partials
  int partials_cap = 0;
  void* work = nullptr;  // solver workspace (api.cpp), created on first use
};

struct icc_filter {
  icc_grid* g = nullptr;
  double r = 0;
  int rad[3] = {0, 0, 0};
  // design-domain variant (topopt_immersed.rs:117-151): active cells and the
  // offset table (di, dj, dk) with weights r - dist in (dk, dj, di) order
  uint8_t* d_act = nullptr;  // [nel], null for the structured filter
  int32_t* d_off = nullptr;  // [noff][3]
  double* d_w = nullptr;     // [noff]
  int noff = 0;
  double* d_hs = nullptr;  // [nel] row sums
  double* d_tmp = nullptr;
};
