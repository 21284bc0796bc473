// Jacobi-PCG vector kernels, FP64 and FP32, with device-resident scalars.
//
// One PCG iteration = apply (p -> ap, per-tile partials of p.Ap) + pcg_update1
// (alpha, x += alpha p, r -= alpha ap, partials of r.r and r.D^-1 r) +
// pcg_update2 (convergence test, beta, p = D^-1 r + beta p, ap = 0).  The
// kernels take the iteration index from the device state and return at once
// when its `done` flag is set, so a batch of iterations is captured once as a
// HIP graph and replayed until the solver stops.
//
// Reductions are deterministic and fence-free inside the iteration: each
// block writes one partial and every block of the consuming kernel re-sums
// all partials in the same fixed order (kernel boundaries publish them).  A
// device-scope fence per block costs an L2 write-back on the multi-XCD part,
// so the last-block pattern is kept for the once-per-solve kernels only.
// Partial sums are accumulated in FP64; the FP64 residual norms of the
// refinement loops use compensated Dot2 accumulation (TwoProd by fma, TwoSum;
// Ogita, Rump and Oishi 2005).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

constexpr int VB = 256;  // threads per block of the vector kernels

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;  // valid on thread 0
}

// Last-block reduction: returns true on every thread of the block that
// finished last; that block then sums `parts` (ordered) with sum_parts.
__device__ __forceinline__ bool last_block(unsigned* ticket) {
  __shared__ bool is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned prev = atomicAdd(ticket, 1u);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last;
}

__device__ __forceinline__ double sum_parts(const double* parts, int n, double* sh) {
  double v = 0.0;
  // fixed assignment of partials to threads, then a fixed tree
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += __builtin_nontemporal_load(parts + i);
  return block_sum(v, sh);
}

template <typename T>
__global__ __launch_bounds__(VB) void fill_kernel(T* __restrict__ y, T v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = v;
}

// inv_diag = |d| > 1e-12 ? 1/d : 1 (solver.rs:24-27); optionally clamps d first
template <typename T>
__global__ __launch_bounds__(VB) void inv_diag_kernel(const double* __restrict__ d, T* __restrict__ dinv, int64_t n,
                                                      double clamp_min, int do_clamp) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double v = d[i];
    if (do_clamp) v = fmax(v, clamp_min);
    dinv[i] = (T)(fabs(v) > 1e-12 ? 1.0 / v : 1.0);
  }
}

// x = 0, r = scale * b (converted to T), z = dinv r, p = z, ap = 0; partials of r.z and b.b
template <typename T, typename TB>
__global__ __launch_bounds__(VB) void pcg_init1(int64_t n, const TB* __restrict__ b, double bscale,
                                                const T* __restrict__ dinv, T* __restrict__ x, T* __restrict__ r,
                                                T* __restrict__ p, T* __restrict__ ap, PcgState* st,
                                                double* __restrict__ parts) {
  __shared__ double sh[16];
  double rz = 0.0, bb = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T ri = (T)((double)b[i] * bscale);
    T zi = dinv[i] * ri;
    x[i] = T(0);
    r[i] = ri;
    p[i] = zi;
    ap[i] = T(0);
    rz += (double)ri * (double)zi;
    bb += (double)ri * (double)ri;
  }
  double s1 = block_sum(rz, sh);
  double s2 = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    parts[blockIdx.x] = s1;
    parts[gridDim.x + blockIdx.x] = s2;
  }
  if (last_block(&st->ticket)) {
    double rz_t = sum_parts(parts, gridDim.x, sh);
    double bb_t = sum_parts(parts + gridDim.x, gridDim.x, sh);
    if (threadIdx.x == 0) {
      st->rz[0] = rz_t;
      st->rz[1] = rz_t;
      st->norm_b = fmax(sqrt(bb_t), 1e-16);
      st->iters = 0;
      st->done = 0;
      st->res = 1.0;
      st->ticket = 0;
    }
  }
}

// Sum of n partials in a fixed order (each thread a fixed strided subset, then
// a fixed tree), so every block of a kernel obtains the identical value.
__device__ __forceinline__ double sum_parts_all(const double* __restrict__ parts, int n, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += parts[i];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  return s;  // valid on every thread
}

// 16-byte runs of V = 16 / sizeof(T) consecutive entries: the update kernels
// move whole runs (one dwordx4 per array), the last run of a vector entry
// by entry.  Each thread takes RU runs a grid stride apart per step, all
// their loads issued before any arithmetic.
constexpr int RU = 2;
template <typename T>
struct Run {
  T v[16 / sizeof(T)];
};
template <typename T>
__device__ __forceinline__ Run<T> ld_run(const T* __restrict__ a, int64_t e, int64_t n) {
  constexpr int V = 16 / (int)sizeof(T);
  Run<T> q;
  if (e + V <= n) {
    const uint4 w = *reinterpret_cast<const uint4*>(a + e);
    __builtin_memcpy(&q, &w, 16);
  } else {
#pragma unroll
    for (int j = 0; j < V; ++j) q.v[j] = e + j < n ? a[e + j] : T(0);
  }
  return q;
}
template <typename T>
__device__ __forceinline__ void st_run(T* __restrict__ a, int64_t e, int64_t n, const Run<T>& q) {
  constexpr int V = 16 / (int)sizeof(T);
  if (e + V <= n) {
    uint4 w;
    __builtin_memcpy(&w, &q, 16);
    *reinterpret_cast<uint4*>(a + e) = w;
  } else {
#pragma unroll
    for (int j = 0; j < V; ++j)
      if (e + j < n) a[e + j] = q.v[j];
  }
}

// alpha = rz / p.Ap (p.Ap summed from the apply's partials), x += alpha p,
// r -= alpha Ap, per-block partials of r.r and r.D^-1 r.  Only block 0 writes
// the state (it_cur, breakdown / limit flags).
// The bodies take the block's index lb among the nb blocks of its vector
// (blockIdx.x / gridDim.x for one instance; the block's place in its
// instance's segment for a batch).
template <typename T>
__device__ __forceinline__ void pcg_update1_body(int lb, int nb, int64_t n, PcgState* st,
                                                 const double* __restrict__ pap_parts, int npap,
                                                 const T* __restrict__ p, const T* __restrict__ ap,
                                                 T* __restrict__ x, T* __restrict__ r, const T* __restrict__ dinv,
                                                 double* __restrict__ parts, double* sh) {
  constexpr int V = 16 / (int)sizeof(T);
  const int64_t gs = (int64_t)nb * blockDim.x;
  Run<T> pv[RU], av[RU], xv[RU], rv[RU], dv[RU];
  auto load_step = [&](int64_t v) {
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int64_t e = (v + u * gs) * V;
      if (e < n) {
        pv[u] = ld_run(p, e, n);
        av[u] = ld_run(ap, e, n);
        xv[u] = ld_run(x, e, n);
        rv[u] = ld_run(r, e, n);
        dv[u] = ld_run(dinv, e, n);
      }
    }
  };
  const int32_t done = st->done;
  const int it = st->iters;
  const int maxit = st->maxit;
  int64_t v = lb * (int64_t)blockDim.x + threadIdx.x;
  // the first step's loads overlap the state checks and the p.Ap sum
  load_step(v);
  if (done) return;
  if (it >= maxit) {  // iteration limit
    if (lb == 0 && threadIdx.x == 0) st->done = 3;
    return;
  }
  const double pap = sum_parts_all(pap_parts, npap, sh);
  if (!(pap > 0.0)) {  // breakdown (solver.rs:47-49): stop before updating x
    if (lb == 0 && threadIdx.x == 0) {
      st->pap = pap;
      st->it_cur = it;
      st->iters = it + 1;  // the reference counts the breaking iteration (final_iter = iter + 1, solver.rs:41)
      st->done = 2;
    }
    return;
  }
  const T alpha = (T)(st->rz[it & 1] / pap);
  double rr = 0.0, rz = 0.0;
  for (; v * V < n; v += RU * gs) {
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int64_t e = (v + u * gs) * V;
      if (e < n) {
#pragma unroll
        for (int j = 0; j < V; ++j) {
          const T ri = rv[u].v[j] - alpha * av[u].v[j];
          xv[u].v[j] = xv[u].v[j] + alpha * pv[u].v[j];
          rv[u].v[j] = ri;
          rr += (double)ri * (double)ri;
          rz += (double)ri * (double)(dv[u].v[j] * ri);
        }
        st_run(x, e, n, xv[u]);
        st_run(r, e, n, rv[u]);
      }
    }
    if ((v + RU * gs) * V < n) load_step(v + RU * gs);
  }
  double s1 = block_sum(rr, sh);
  double s2 = block_sum(rz, sh);
  if (threadIdx.x == 0) {
    parts[lb] = s1;
    parts[nb + lb] = s2;
    if (lb == 0) {
      st->pap = pap;
      st->it_cur = it;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(VB) void pcg_update1(int64_t n, PcgState* st, const double* __restrict__ pap_parts,
                                                  int npap, const T* __restrict__ p, const T* __restrict__ ap,
                                                  T* __restrict__ x, T* __restrict__ r, const T* __restrict__ dinv,
                                                  double* __restrict__ parts) {
  __shared__ double sh[16];
  pcg_update1_body<T>(blockIdx.x, gridDim.x, n, st, pap_parts, npap, p, ap, x, r, dinv, parts, sh);
}

// Convergence test on |r| / norm_b (solver.rs:57-61), then
// beta = rz_new / rz_old, p = D^-1 r + beta p, ap = 0.  Block 0 writes the
// state (iters, rz_new, res, done).
template <typename T>
__device__ __forceinline__ void pcg_update2_body(int lb, int nb, int64_t n, PcgState* st,
                                                 const double* __restrict__ parts, T* __restrict__ p,
                                                 const T* __restrict__ r, const T* __restrict__ dinv,
                                                 T* __restrict__ ap, double* sh) {
  constexpr int V = 16 / (int)sizeof(T);
  const int64_t gs = (int64_t)nb * blockDim.x;
  Run<T> pv[RU], rv[RU], dv[RU];
  auto load_step = [&](int64_t v) {
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int64_t e = (v + u * gs) * V;
      if (e < n) {
        pv[u] = ld_run(p, e, n);
        rv[u] = ld_run(r, e, n);
        dv[u] = ld_run(dinv, e, n);
      }
    }
  };
  const int32_t done = st->done;
  const int it = st->it_cur;
  int64_t v = lb * (int64_t)blockDim.x + threadIdx.x;
  // the first step's loads overlap the state check and the sums
  load_step(v);
  if (done) return;
  const double rr = sum_parts_all(parts, nb, sh);
  const double rz = sum_parts_all(parts + nb, nb, sh);
  const double res = sqrt(rr) / st->norm_b;
  const bool conv = res < st->tol;
  if (lb == 0 && threadIdx.x == 0) {
    st->rr = rr;
    st->res = res;
    st->rz[(it + 1) & 1] = rz;
    st->iters = it + 1;
    if (conv) st->done = 1;
  }
  if (conv) return;
  const T beta = (T)(rz / st->rz[it & 1]);
  Run<T> zero;
#pragma unroll
  for (int j = 0; j < V; ++j) zero.v[j] = T(0);
  for (; v * V < n; v += RU * gs) {
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int64_t e = (v + u * gs) * V;
      if (e < n) {
#pragma unroll
        for (int j = 0; j < V; ++j) pv[u].v[j] = dv[u].v[j] * rv[u].v[j] + beta * pv[u].v[j];
        st_run(p, e, n, pv[u]);
        st_run(ap, e, n, zero);
      }
    }
    if ((v + RU * gs) * V < n) load_step(v + RU * gs);
  }
}

template <typename T>
__global__ __launch_bounds__(VB) void pcg_update2(int64_t n, PcgState* st, const double* __restrict__ parts,
                                                  T* __restrict__ p, const T* __restrict__ r,
                                                  const T* __restrict__ dinv, T* __restrict__ ap) {
  __shared__ double sh[16];
  pcg_update2_body<T>(blockIdx.x, gridDim.x, n, st, parts, p, r, dinv, ap, sh);
}

// Batched updates: block b works on instance seg_of[b]'s segment (its
// vectors at seg.off, its blocks seg.blk0 .. seg.blk0 + seg.nblk, its state,
// its p.Ap slots seg.pap0 .. seg.pap0 + seg.npap and partials at 2 seg.blk0).
template <typename T>
__global__ __launch_bounds__(VB) void pcg_update1_batch(const VecSeg* __restrict__ segs,
                                                        const int32_t* __restrict__ seg_of, PcgState* st,
                                                        const double* __restrict__ pap, const T* __restrict__ p,
                                                        const T* __restrict__ ap, T* __restrict__ x, T* __restrict__ r,
                                                        const T* __restrict__ dinv, double* __restrict__ parts) {
  __shared__ double sh[16];
  const int k = seg_of[blockIdx.x];
  const VecSeg sg = segs[k];
  pcg_update1_body<T>(blockIdx.x - sg.blk0, sg.nblk, sg.n, st + k, pap + sg.pap0, sg.npap, p + sg.off, ap + sg.off,
                      x + sg.off, r + sg.off, dinv + sg.off, parts + 2 * sg.blk0, sh);
}
template <typename T>
__global__ __launch_bounds__(VB) void pcg_update2_batch(const VecSeg* __restrict__ segs,
                                                        const int32_t* __restrict__ seg_of, PcgState* st,
                                                        const double* __restrict__ parts, T* __restrict__ p,
                                                        const T* __restrict__ r, const T* __restrict__ dinv,
                                                        T* __restrict__ ap) {
  __shared__ double sh[16];
  const int k = seg_of[blockIdx.x];
  const VecSeg sg = segs[k];
  pcg_update2_body<T>(blockIdx.x - sg.blk0, sg.nblk, sg.n, st + k, parts + 2 * sg.blk0, p + sg.off, r + sg.off,
                      dinv + sg.off, ap + sg.off, sh);
}

// p = D^-1 r + beta p, beta = rz[iters & 1] / rz[(iters - 1) & 1], ap = 0:
// the direction update the mixed solver completes after a reliable update.
__global__ __launch_bounds__(VB) void pdir_update_kernel(int64_t n, const PcgState* st, float* __restrict__ p,
                                                         const float* __restrict__ r, const float* __restrict__ dinv,
                                                         float* __restrict__ ap) {
  const int it = st->iters;
  const float beta = (float)(st->rz[it & 1] / st->rz[(it - 1) & 1]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    p[i] = dinv[i] * r[i] + beta * p[i];
    ap[i] = 0.0f;
  }
}

// Reliable update of the mixed solver: r32 = (float) r64, x32 = 0, and the
// recurrence scalars restart from the replaced residual:
// rz[iters & 1] = r.D^-1 r, norm_b = |r|, done = 0.
__global__ __launch_bounds__(VB) void reliable_replace_kernel(int64_t n, const double* __restrict__ r64,
                                                              const float* __restrict__ dinv, float* __restrict__ r,
                                                              float* __restrict__ x, PcgState* st,
                                                              double* __restrict__ parts) {
  __shared__ double sh[16];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float ri = (float)r64[i];
    r[i] = ri;
    x[i] = 0.0f;
    rz += (double)ri * (double)(dinv[i] * ri);
    rr += (double)ri * (double)ri;
  }
  double s1 = block_sum(rz, sh);
  double s2 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    parts[blockIdx.x] = s1;
    parts[gridDim.x + blockIdx.x] = s2;
  }
  if (last_block(&st->ticket)) {
    double rz_t = sum_parts(parts, gridDim.x, sh);
    double rr_t = sum_parts(parts + gridDim.x, gridDim.x, sh);
    if (threadIdx.x == 0) {
      st->rz[st->iters & 1] = rz_t;
      st->norm_b = fmax(sqrt(rr_t), 1e-300);
      st->rr = rr_t;
      st->res = 1.0;
      st->done = 0;
      st->ticket = 0;
    }
  }
}

// ---- refinement helpers ------------------------------------------------------
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Compensated block sum of (hi, lo) pairs: thread 0 gets hi + lo.
__device__ double block_sum_comp(double hi, double lo, double* sh) {
  __shared__ double shl[16];
  for (int off = 32; off > 0; off >>= 1) {
    double oh = __shfl_down(hi, off, 64), ol = __shfl_down(lo, off, 64);
    double s, e;
    two_sum(hi, oh, s, e);
    hi = s;
    lo = lo + ol + e;
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    sh[w] = hi;
    shl[w] = lo;
  }
  __syncthreads();
  double out = 0.0;
  if (threadIdx.x == 0) {
    double h = 0.0, lo2 = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      double s, e;
      two_sum(h, sh[i], s, e);
      h = s;
      lo2 += shl[i] + e;
    }
    out = h + lo2;
  }
  __syncthreads();
  return out;
}

// r = free ? f - ap : 0 ; |r|^2 and |f_free|^2 by Dot2; results in st->rr, st->norm_b (squared norms)
__global__ __launch_bounds__(VB) void residual_kernel(int64_t n, const double* __restrict__ f,
                                                      const double* __restrict__ ap, const uint8_t* __restrict__ freem,
                                                      double* __restrict__ r, PcgState* st, double* __restrict__ parts) {
  __shared__ double sh[16];
  double h1 = 0.0, l1 = 0.0, h2 = 0.0, l2 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    bool fr = freem == nullptr || freem[i] != 0;
    double fi = fr ? f[i] : 0.0;
    double ri = fr ? fi - ap[i] : 0.0;
    r[i] = ri;
    double p = ri * ri, pe = fma(ri, ri, -p), s, e;
    two_sum(h1, p, s, e);
    h1 = s;
    l1 += e + pe;
    p = fi * fi;
    pe = fma(fi, fi, -p);
    two_sum(h2, p, s, e);
    h2 = s;
    l2 += e + pe;
  }
  double s1 = block_sum_comp(h1, l1, sh);
  double s2 = block_sum_comp(h2, l2, sh);
  if (threadIdx.x == 0) {
    parts[blockIdx.x] = s1;
    parts[gridDim.x + blockIdx.x] = s2;
  }
  if (last_block(&st->ticket)) {
    // compensated sum of the block partials in index order
    double h = 0.0, lo = 0.0, hf = 0.0, lof = 0.0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
      double s, e;
      two_sum(h, __builtin_nontemporal_load(parts + i), s, e);
      h = s;
      lo += e;
      two_sum(hf, __builtin_nontemporal_load(parts + gridDim.x + i), s, e);
      hf = s;
      lof += e;
    }
    double t1 = block_sum_comp(h, lo, sh);
    double t2 = block_sum_comp(hf, lof, sh);
    if (threadIdx.x == 0) {
      st->rr = t1;
      st->norm_b = t2;
      st->ticket = 0;
    }
  }
}

// u += s * d  (d in T)
template <typename T>
__global__ __launch_bounds__(VB) void ir_update_kernel(int64_t n, double* __restrict__ u, const T* __restrict__ d,
                                                       double s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] += s * (double)d[i];
}

// ---- CSR add-on: y = S x (overwrites y), per-block partials of x.(S x) ----------
// The add-on's rows with entries (`rows`, nrows of them) are few (penalty
// and surface terms) and the rest of y must be zeroed, so the launch has two
// parts:
//   * blocks [0, rb): each wave takes four rows with entries per step (a
//     {row, first entry, end, 0} record each), the lanes over a row's
//     entries, the first 64 entries of all four loaded at once; each row sum
//     is reduced across the wave in a fixed order, and each block writes one
//     partial of x.(S x) (its waves' rows in order);
//   * blocks [rb, gridDim.x): y = 0 on the rows without entries, four rows per
//     thread.
// Masked: rows and columns with freem = 0 contribute nothing (y = 0 there).
template <typename T, typename TD>
__global__ __launch_bounds__(VB) void csr_spmv_kernel(int64_t n, const int32_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ indices, const TD* __restrict__ data,
                                                      const int32_t* __restrict__ rows, int64_t nrows, int rb,
                                                      const T* __restrict__ x, T* __restrict__ y,
                                                      const uint8_t* __restrict__ freem, double* __restrict__ parts) {
  __shared__ double sh[16];
  if ((int)blockIdx.x < rb) {
    // each wave takes CG listed rows per step ({row, first entry, end, 0}
    // records) and issues the first 64 entries' loads of all of them before
    // any arithmetic, so their dependent loads (entry -> column -> x) overlap
    constexpr int WPB = VB / 64, CG = 4;
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)rb * WPB;
    const int4* __restrict__ info = reinterpret_cast<const int4*>(rows);
    double acc = 0.0;
    for (int64_t q0 = ((int64_t)blockIdx.x * WPB + (threadIdx.x >> 6)) * CG; q0 < nrows; q0 += nw * CG) {
      int ri[CG], k0[CG], k1[CG], c[CG];
      TD dv[CG];
      T s[CG];
#pragma unroll
      for (int j = 0; j < CG; ++j) {
        if (q0 + j < nrows) {
          const int4 v = info[q0 + j];
          ri[j] = v.x;
          k0[j] = v.y;
          k1[j] = v.z;
        } else {
          ri[j] = -1;
          k0[j] = k1[j] = 0;
        }
      }
#pragma unroll
      for (int j = 0; j < CG; ++j) {
        const int k = k0[j] + lane;
        c[j] = k < k1[j] ? indices[k] : -1;
        dv[j] = k < k1[j] ? data[k] : TD(0);
      }
#pragma unroll
      for (int j = 0; j < CG; ++j) {
        const bool on = c[j] >= 0 && (freem == nullptr || freem[c[j] >= 0 ? c[j] : 0] != 0);
        const T xv = c[j] >= 0 ? x[c[j]] : T(0);
        s[j] = on ? (T)dv[j] * xv : T(0);
      }
      // rows of more than 64 entries: the rest in order, 64 at a time
#pragma unroll
      for (int j = 0; j < CG; ++j)
        for (int k = k0[j] + 64 + lane; k < k1[j]; k += 64) {
          const int cc = indices[k];
          const T xv = (freem == nullptr || freem[cc] != 0) ? x[cc] : T(0);
          s[j] += (T)data[k] * xv;
        }
#pragma unroll
      for (int j = 0; j < CG; ++j) {
        for (int off = 32; off > 0; off >>= 1) s[j] += __shfl_xor(s[j], off, 64);
        if (ri[j] >= 0) {
          const bool fr = freem == nullptr || freem[ri[j]] != 0;
          if (!fr) s[j] = T(0);
          if (lane == 0) {
            y[ri[j]] = s[j];
            acc += (double)x[ri[j]] * (double)s[j];
          }
        }
      }
    }
    if (parts) {
      const double b = block_sum(acc, sh);
      if (threadIdx.x == 0) parts[blockIdx.x] = b;
    }
  } else {
    const int64_t stride = (int64_t)(gridDim.x - rb) * VB * 4;
    for (int64_t i0 = ((int64_t)(blockIdx.x - rb) * VB + threadIdx.x) * 4; i0 < n; i0 += stride) {
      int32_t ip[5];
#pragma unroll
      for (int j = 0; j < 5; ++j) ip[j] = i0 + j <= n ? indptr[i0 + j] : 0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (i0 + j < n && ip[j] == ip[j + 1]) y[i0 + j] = T(0);
    }
  }
}

// s_e = emin + (1 - emin) x_e^penal  [* w_e, or emin where w_e is inactive]
// (topopt_iga.rs:132; topopt_immersed.rs:166 when w != null)
__global__ __launch_bounds__(VB) void simp_scale_kernel(int64_t n, const double* __restrict__ x, double penal,
                                                        double emin, const double* __restrict__ w,
                                                        const uint8_t* __restrict__ active, double* __restrict__ s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double v = emin + (1.0 - emin) * pow(x[i], penal);
    if (w != nullptr) v = (active == nullptr || active[i]) ? w[i] * v : emin;
    s[i] = v;
  }
}

// u = free ? u0 + du : 0   (topopt_iga.rs:160)
__global__ __launch_bounds__(VB) void masked_add_kernel(int64_t n, const double* __restrict__ u0,
                                                        const double* __restrict__ du,
                                                        const uint8_t* __restrict__ freem, double* __restrict__ u) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] = (freem == nullptr || freem[i] != 0) ? u0[i] + du[i] : 0.0;
}

__global__ __launch_bounds__(VB) void clamp_min_kernel(int64_t n, const double* __restrict__ a, double lo,
                                                       double* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = fmax(a[i], lo);
}

// ---- launchers ------------------------------------------------------------------
hipError_t launch_clamp_min(int64_t n, const double* a, double lo, double* b, hipStream_t st) {
  hipLaunchKernelGGL(clamp_min_kernel, dim3(vec_blocks(n)), dim3(VB), 0, st, n, a, lo, b);
  return hipGetLastError();
}

hipError_t launch_simp_scale(int64_t n, const double* x, double penal, double emin, const double* w,
                             const uint8_t* active, double* s, hipStream_t st) {
  hipLaunchKernelGGL(simp_scale_kernel, dim3(vec_blocks(n)), dim3(VB), 0, st, n, x, penal, emin, w, active, s);
  return hipGetLastError();
}

hipError_t launch_masked_add(int64_t n, const double* u0, const double* du, const uint8_t* freem, double* u,
                             hipStream_t st) {
  hipLaunchKernelGGL(masked_add_kernel, dim3(vec_blocks(n)), dim3(VB), 0, st, n, u0, du, freem, u);
  return hipGetLastError();
}

// Blocks of the vector kernels: at least four elements per thread, at most
// VEC_MAX_BLOCKS (the partial arrays hold 2 * VEC_MAX_BLOCKS values).
int vec_blocks(int64_t n) {
  int64_t nb = (n + VB * 4 - 1) / (VB * 4);
  if (nb < 1) nb = 1;
  if (nb > VEC_MAX_BLOCKS) nb = VEC_MAX_BLOCKS;
  return (int)nb;
}

template <typename T>
hipError_t launch_fill(T* y, T v, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL((fill_kernel<T>), dim3(vec_blocks(n)), dim3(VB), 0, st, y, v, n);
  return hipGetLastError();
}
template hipError_t launch_fill<double>(double*, double, int64_t, hipStream_t);
template hipError_t launch_fill<float>(float*, float, int64_t, hipStream_t);

template <typename T>
hipError_t launch_inv_diag(const double* d, T* dinv, int64_t n, double clamp_min, int do_clamp, hipStream_t st) {
  hipLaunchKernelGGL((inv_diag_kernel<T>), dim3(vec_blocks(n)), dim3(VB), 0, st, d, dinv, n, clamp_min, do_clamp);
  return hipGetLastError();
}
template hipError_t launch_inv_diag<double>(const double*, double*, int64_t, double, int, hipStream_t);
template hipError_t launch_inv_diag<float>(const double*, float*, int64_t, double, int, hipStream_t);

template <typename T, typename TB>
hipError_t launch_pcg_init(int64_t n, const TB* b, double bscale, const T* dinv, T* x, T* r, T* p, T* ap,
                           PcgState* st, double* parts, hipStream_t s) {
  hipLaunchKernelGGL((pcg_init1<T, TB>), dim3(vec_blocks(n)), dim3(VB), 0, s, n, b, bscale, dinv, x, r, p, ap, st,
                     parts);
  return hipGetLastError();
}
template hipError_t launch_pcg_init<double, double>(int64_t, const double*, double, const double*, double*, double*,
                                                    double*, double*, PcgState*, double*, hipStream_t);
template hipError_t launch_pcg_init<float, double>(int64_t, const double*, double, const float*, float*, float*,
                                                   float*, float*, PcgState*, double*, hipStream_t);

template <typename T>
hipError_t launch_pcg_update(int64_t n, PcgState* st, const double* pap_parts, int npap, T* p, T* ap, T* x, T* r,
                             const T* dinv, double* parts, hipStream_t s) {
  int nb = vec_blocks(n);
  hipLaunchKernelGGL((pcg_update1<T>), dim3(nb), dim3(VB), 0, s, n, st, pap_parts, npap, p, ap, x, r, dinv, parts);
  hipLaunchKernelGGL((pcg_update2<T>), dim3(nb), dim3(VB), 0, s, n, st, parts, p, r, dinv, ap);
  return hipGetLastError();
}
template hipError_t launch_pcg_update<double>(int64_t, PcgState*, const double*, int, double*, double*, double*,
                                              double*, const double*, double*, hipStream_t);

template <typename T>
hipError_t launch_pcg_update_batch(int nblk, const VecSeg* segs, const int32_t* seg_of, PcgState* st, const double* pap,
                                   T* p, T* ap, T* x, T* r, const T* dinv, double* parts, hipStream_t s) {
  hipLaunchKernelGGL((pcg_update1_batch<T>), dim3(nblk), dim3(VB), 0, s, segs, seg_of, st, pap, p, ap, x, r, dinv, parts);
  hipLaunchKernelGGL((pcg_update2_batch<T>), dim3(nblk), dim3(VB), 0, s, segs, seg_of, st, parts, p, r, dinv, ap);
  return hipGetLastError();
}
template hipError_t launch_pcg_update_batch<double>(int, const VecSeg*, const int32_t*, PcgState*, const double*,
                                                    double*, double*, double*, double*, const double*, double*,
                                                    hipStream_t);
template hipError_t launch_pcg_update_batch<float>(int, const VecSeg*, const int32_t*, PcgState*, const double*,
                                                   float*, float*, float*, float*, const float*, double*, hipStream_t);
template hipError_t launch_pcg_update<float>(int64_t, PcgState*, const double*, int, float*, float*, float*, float*,
                                             const float*, double*, hipStream_t);

hipError_t launch_pdir_update(int64_t n, const PcgState* st, float* p, const float* r, const float* dinv, float* ap,
                              hipStream_t s) {
  hipLaunchKernelGGL(pdir_update_kernel, dim3(vec_blocks(n)), dim3(VB), 0, s, n, st, p, r, dinv, ap);
  return hipGetLastError();
}

hipError_t launch_reliable_replace(int64_t n, const double* r64, const float* dinv, float* r, float* x, PcgState* st,
                                   double* parts, hipStream_t s) {
  hipLaunchKernelGGL(reliable_replace_kernel, dim3(vec_blocks(n)), dim3(VB), 0, s, n, r64, dinv, r, x, st, parts);
  return hipGetLastError();
}

hipError_t launch_residual(int64_t n, const double* f, const double* ap, const uint8_t* freem, double* r,
                           PcgState* st, double* parts, hipStream_t s) {
  hipLaunchKernelGGL(residual_kernel, dim3(vec_blocks(n)), dim3(VB), 0, s, n, f, ap, freem, r, st, parts);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ir_update(int64_t n, double* u, const T* d, double scale, hipStream_t s) {
  hipLaunchKernelGGL((ir_update_kernel<T>), dim3(vec_blocks(n)), dim3(VB), 0, s, n, u, d, scale);
  return hipGetLastError();
}
template hipError_t launch_ir_update<double>(int64_t, double*, const double*, double, hipStream_t);
template hipError_t launch_ir_update<float>(int64_t, double*, const float*, double, hipStream_t);

template <typename T, typename TD>
hipError_t launch_csr_spmv(int64_t n, const int32_t* indptr, const int32_t* indices, const TD* data,
                           const int32_t* rows, int64_t nrows, int rb, const T* x, T* y, const uint8_t* freem,
                           double* parts, bool y_zeroed, hipStream_t s) {
  const int64_t fb = y_zeroed ? 0 : std::min<int64_t>((n + 4 * VB - 1) / (4 * VB), 2048);
  hipLaunchKernelGGL((csr_spmv_kernel<T, TD>), dim3((unsigned)(rb + fb)), dim3(VB), 0, s, n, indptr, indices, data,
                     rows, nrows, rb, x, y, freem, parts);
  return hipGetLastError();
}
template hipError_t launch_csr_spmv<double, double>(int64_t, const int32_t*, const int32_t*, const double*,
                                                    const int32_t*, int64_t, int, const double*, double*,
                                                    const uint8_t*, double*, bool, hipStream_t);
template hipError_t launch_csr_spmv<float, float>(int64_t, const int32_t*, const int32_t*, const float*,
                                                  const int32_t*, int64_t, int, const float*, float*, const uint8_t*,
                                                  double*, bool, hipStream_t);

}  // namespace icc
