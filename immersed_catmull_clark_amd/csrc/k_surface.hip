// Surface boundary terms on triangle surfaces (penalty Dirichlet, traction,
// pressure, Robin): the basis at the surface quadrature points and the
// per-element surface blocks, formed on the device.
//
// surf_basis_kernel: one thread per quadrature point; the 27 nonzero scalar
// B-spline values at the point (scalar_basis_at, rust/src/immersed_bc.rs:25-50,
// Cox-de Boor of rust/src/iga/basis.rs:20-50 on the grid's knot vectors) and
// the global index of the first one, the key of the point's element block.
//
// surf_block_kernel: one workgroup per element block (a run of points with
// the same key, in the caller's point order); thread t owns the entries
// t, t + 256, ... of the 81 x 81 block m[(ci, a), (cj, b)] and adds
// (ks_q[ci][cj] w_q N_a) N_b point by point, skipping zero ks entries, as
// SurfaceAccumulator::add does (immersed_bc.rs:61-80); threads 0..80 add the
// point loads of entry (c, a).  Built with contraction off, so every block
// entry equals the host restatement's bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"
#include "surface_basis.hpp"

namespace icc {

struct SurfGeom {
  const double* knots[3];
  int nk[3];
  double lo[3], len[3];
  int64_t ncp0, ncp1;
};

__global__ void surf_basis_kernel(int64_t npts, const double* __restrict__ px, SurfGeom gm, int64_t* __restrict__ key,
                                  double* __restrict__ nval) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  int first[3];
  double v[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double xi = (px[3 * q + d] - gm.lo[d]) / gm.len[d];
    xi = xi < 0.0 ? 0.0 : (xi > 1.0 ? 1.0 : xi);  // f64::clamp(0, 1)
    const int s = surf_span(gm.knots[d], gm.nk[d], xi);
    surf_basis_p2(gm.knots[d], s, xi, v[d]);
    first[d] = s - 2;
  }
  key[q] = (int64_t)first[0] + gm.ncp0 * ((int64_t)first[1] + gm.ncp1 * (int64_t)first[2]);
  // a = ax + 3 ay + 9 az (unravel order, x fastest); w = ((1 vx) vy) vz
  for (int a = 0; a < 27; ++a) {
    const int ax = a % 3, ay = (a / 3) % 3, az = a / 9;
    double w = 1.0;
    w *= v[0][ax];
    w *= v[1][ay];
    w *= v[2][az];
    nval[27 * q + a] = w;
  }
}

// fmode 0: f[c] += ((fs w) N_a) fv[c];  fmode 1: f[c] -= ((w N_a) fs) fv[c];  fmode 2: no load
__global__ __launch_bounds__(256) void surf_block_kernel(const int64_t* __restrict__ seg, const int64_t* __restrict__ perm,
                                                         const double* __restrict__ nval, const double* __restrict__ pw,
                                                         const double* __restrict__ ks, const int8_t* __restrict__ fmode,
                                                         const double* __restrict__ fs, const double* __restrict__ fv,
                                                         double* __restrict__ blocks, double* __restrict__ fe) {
  constexpr int PER = (81 * 81 + 255) / 256;
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  double m[PER];
  int ra[PER], rb[PER], kk[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    m[j] = 0.0;
    const int e = tid + 256 * j;
    const int R = e / 81, Cc = e - 81 * (e / 81);
    ra[j] = R % 27;
    rb[j] = Cc % 27;
    kk[j] = (R / 27) * 3 + Cc / 27;
  }
  double f = 0.0;
  const int fc = tid / 27, fa = tid - 27 * (tid / 27);
  for (int64_t i = seg[b]; i < seg[b + 1]; ++i) {
    const int64_t q = perm[i];
    const double w = pw[q];
    const double* nq = nval + 27 * q;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if (tid + 256 * j < 81 * 81) {
        const double k = ks[9 * q + kk[j]] * w;
        if (k != 0.0) {
          const double ka = k * nq[ra[j]];
          m[j] += ka * nq[rb[j]];
        }
      }
    }
    if (tid < 81) {
      const int md = fmode[q];
      if (md == 0)
        f += fs[q] * w * nq[fa] * fv[3 * q + fc];
      else if (md == 1)
        f -= w * nq[fa] * fs[q] * fv[3 * q + fc];
    }
  }
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (tid + 256 * j < 81 * 81) blocks[b * 6561 + tid + 256 * j] = m[j];
  if (tid < 81) fe[b * 81 + tid] = f;
}

hipError_t launch_surface_basis(int64_t npts, const double* px, const double* const knots[3], const int nk[3],
                                const double lo[3], const double len[3], int64_t ncp0, int64_t ncp1, int64_t* key,
                                double* nval, hipStream_t st) {
  if (npts == 0) return hipSuccess;
  SurfGeom gm;
  for (int d = 0; d < 3; ++d) {
    gm.knots[d] = knots[d];
    gm.nk[d] = nk[d];
    gm.lo[d] = lo[d];
    gm.len[d] = len[d];
  }
  gm.ncp0 = ncp0;
  gm.ncp1 = ncp1;
  const int64_t blocks = (npts + 255) / 256;
  hipLaunchKernelGGL(surf_basis_kernel, dim3((unsigned)blocks), dim3(256), 0, st, npts, px, gm, key, nval);
  return hipGetLastError();
}

hipError_t launch_surface_blocks(int64_t nseg, const int64_t* seg, const int64_t* perm, const double* nval,
                                 const double* pw, const double* ks, const int8_t* fmode, const double* fs,
                                 const double* fv, double* blocks, double* fe, hipStream_t st) {
  if (nseg == 0) return hipSuccess;
  hipLaunchKernelGGL(surf_block_kernel, dim3((unsigned)nseg), dim3(256), 0, st, seg, perm, nval, pw, ks, fmode, fs, fv,
                     blocks, fe);
  return hipGetLastError();
}

}  // namespace icc
