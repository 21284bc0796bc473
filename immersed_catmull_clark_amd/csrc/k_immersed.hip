// Immersed background-cell classification and volume fractions on the GPU,
// for two inside predicates:
//   * CylGeom: "inside a box and outside every cylinder parallel to z"
//     (inside iff the signed distance is < 0, strict);
//   * MeshGeom: ray parity along +x against a closed triangle mesh, the
//     ParityRayCaster of rust/src/immersed.rs:29-76 (per-triangle y-z box
//     cull, barycentric test in the y-z plane with the 1e-15 guard, hit iff
//     x_hit > q_x; inside iff the hit count is odd).
//
// Semantics of rust/src/immersed.rs:79-155: corner coordinates from the
// MATLAB-compatible linspace(lo, hi, res + 1) (last point = hi exactly); a
// cell with 8 inside corners is INSIDE (1), with none OUTSIDE (0), else CUT
// (-1); a CUT cell's weight is the fraction of its sub^3 midpoints
// lo + (a + 0.5) dx inside, with lo = lo_d + idx h and dx = (hi - lo) / sub.
// Built with -ffp-contract=off so every coordinate, distance and barycentric
// coordinate rounds like the scalar CPU evaluation (f64 division and sqrt are
// correctly rounded on the device), which makes status and weights
// bit-identical to it.  The triangle loop index is wave-uniform, so the
// triangle records are read through scalar loads.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "icc_internal.hpp"
#include "icc_kernels.hpp"

namespace icc {

struct CylGeom {
  double lo[3], hi[3];
  int ncyl;
  const double* cyl;  // [ncyl][3] = (cx, cy, r)
  __device__ __forceinline__ bool inside(double q0, double q1, double q2) const {
    double s = -1e300;
    const double q[3] = {q0, q1, q2};
    for (int d = 0; d < 3; ++d) s = fmax(s, fmax(lo[d] - q[d], q[d] - hi[d]));
    for (int c = 0; c < ncyl; ++c) {
      const double dx = q0 - cyl[3 * c], dy = q1 - cyl[3 * c + 1];
      s = fmax(s, cyl[3 * c + 2] - sqrt(dx * dx + dy * dy));
    }
    return s < 0.0;
  }
};

struct MeshGeom {
  int ntri;
  const double* tri;  // [ntri][MESH_TRI_STRIDE]: a, b, c (xyz each), y-z box (ymin, ymax, zmin, zmax)
  __device__ __forceinline__ bool inside(double qx, double qy, double qz) const {
    int count = 0;
    for (int t = 0; t < ntri; ++t) {
      const double* T = tri + (int64_t)t * MESH_TRI_STRIDE;
      if (!(qy >= T[9] && qy <= T[10] && qz >= T[11] && qz <= T[12])) continue;
      const double v00 = T[7] - T[1], v01 = T[8] - T[2];  // c - a (y, z)
      const double v10 = T[4] - T[1], v11 = T[5] - T[2];  // b - a
      const double v20 = qy - T[1], v21 = qz - T[2];
      const double dot00 = v00 * v00 + v01 * v01;
      const double dot01 = v00 * v10 + v01 * v11;
      const double dot02 = v00 * v20 + v01 * v21;
      const double dot11 = v10 * v10 + v11 * v11;
      const double dot12 = v10 * v20 + v11 * v21;
      const double inv = 1.0 / (dot00 * dot11 - dot01 * dot01 + 1e-15);
      const double u = (dot11 * dot02 - dot01 * dot12) * inv;
      const double v = (dot00 * dot12 - dot01 * dot02) * inv;
      if (u >= 0.0 && v >= 0.0 && u + v <= 1.0) {
        const double x_hit = T[0] + v * (T[3] - T[0]) + u * (T[6] - T[0]);
        if (x_hit > qx) count += 1;
      }
    }
    return (count & 1) == 1;
  }
};

template <typename G>
__global__ __launch_bounds__(256) void corner_kernel(G g, double b0, double b1, double b2, double e0, double e1,
                                                     double e2, int nv0, int nv1, int nv2,
                                                     uint8_t* __restrict__ inside) {
  const int64_t nv = (int64_t)nv0 * nv1 * nv2;
  const double st0 = (e0 - b0) / (double)(nv0 - 1), st1 = (e1 - b1) / (double)(nv1 - 1),
               st2 = (e2 - b2) / (double)(nv2 - 1);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(i % nv0), b = (int)((i / nv0) % nv1), c = (int)(i / ((int64_t)nv0 * nv1));
    const double x = a == nv0 - 1 ? e0 : b0 + (double)a * st0;
    const double y = b == nv1 - 1 ? e1 : b1 + (double)b * st1;
    const double z = c == nv2 - 1 ? e2 : b2 + (double)c * st2;
    inside[i] = g.inside(x, y, z) ? 1 : 0;
  }
}

template <typename G>
__global__ __launch_bounds__(256) void cell_kernel(G g, double b0, double b1, double b2, double e0, double e1,
                                                   double e2, int r0, int r1, int r2, int s0, int s1, int s2,
                                                   const uint8_t* __restrict__ inside, int8_t* __restrict__ status,
                                                   double* __restrict__ w) {
  const int64_t nc = (int64_t)r0 * r1 * r2;
  const int nv0 = r0 + 1, nv1 = r1 + 1;
  const double h0 = (e0 - b0) / (double)r0, h1 = (e1 - b1) / (double)r1, h2 = (e2 - b2) / (double)r2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nc; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % r0), j = (int)((e / r0) % r1), k = (int)(e / ((int64_t)r0 * r1));
    int cnt = 0;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) cnt += inside[(i + di) + nv0 * ((j + dj) + (int64_t)nv1 * (k + dk))];
    if (cnt == 8) {
      status[e] = 1;
      w[e] = 1.0;
    } else if (cnt == 0) {
      status[e] = 0;
      w[e] = 0.0;
    } else {
      status[e] = -1;
      const double lo0 = b0 + (double)i * h0, lo1 = b1 + (double)j * h1, lo2 = b2 + (double)k * h2;
      const double hi0 = b0 + (double)(i + 1) * h0, hi1 = b1 + (double)(j + 1) * h1, hi2 = b2 + (double)(k + 1) * h2;
      const double dx0 = (hi0 - lo0) / (double)s0, dx1 = (hi1 - lo1) / (double)s1, dx2 = (hi2 - lo2) / (double)s2;
      int n_in = 0;
      for (int c = 0; c < s2; ++c)
        for (int b = 0; b < s1; ++b)
          for (int a = 0; a < s0; ++a)
            n_in += g.inside(lo0 + ((double)a + 0.5) * dx0, lo1 + ((double)b + 0.5) * dx1,
                             lo2 + ((double)c + 0.5) * dx2)
                        ? 1
                        : 0;
      w[e] = (double)n_in / (double)(s0 * s1 * s2);
    }
  }
}

template <typename G>
static hipError_t launch_classify(const G& g, const double gb[6], const int64_t res[3], const int64_t sub[3],
                                  uint8_t* inside_dev, int8_t* status_dev, double* w_dev, hipStream_t st) {
  const int nv0 = (int)res[0] + 1, nv1 = (int)res[1] + 1, nv2 = (int)res[2] + 1;
  const int64_t nv = (int64_t)nv0 * nv1 * nv2;
  int nb = (int)((nv + 255) / 256);
  if (nb > 16384) nb = 16384;
  hipLaunchKernelGGL(corner_kernel<G>, dim3(nb), dim3(256), 0, st, g, gb[0], gb[2], gb[4], gb[1], gb[3], gb[5], nv0,
                     nv1, nv2, inside_dev);
  const int64_t nc = res[0] * res[1] * res[2];
  nb = (int)((nc + 255) / 256);
  if (nb > 16384) nb = 16384;
  hipLaunchKernelGGL(cell_kernel<G>, dim3(nb), dim3(256), 0, st, g, gb[0], gb[2], gb[4], gb[1], gb[3], gb[5],
                     (int)res[0], (int)res[1], (int)res[2], (int)sub[0], (int)sub[1], (int)sub[2], inside_dev,
                     status_dev, w_dev);
  return hipGetLastError();
}

hipError_t launch_classify_cylinders(const double box_lo[3], const double box_hi[3], const double* cyl_dev,
                                     int64_t ncyl, const double gb[6], const int64_t res[3], const int64_t sub[3],
                                     uint8_t* inside_dev, int8_t* status_dev, double* w_dev, hipStream_t st) {
  CylGeom g;
  for (int d = 0; d < 3; ++d) {
    g.lo[d] = box_lo[d];
    g.hi[d] = box_hi[d];
  }
  g.ncyl = (int)ncyl;
  g.cyl = cyl_dev;
  return launch_classify(g, gb, res, sub, inside_dev, status_dev, w_dev, st);
}

hipError_t launch_classify_mesh(const double* tri_dev, int64_t ntri, const double gb[6], const int64_t res[3],
                                const int64_t sub[3], uint8_t* inside_dev, int8_t* status_dev, double* w_dev,
                                hipStream_t st) {
  MeshGeom g;
  g.ntri = (int)ntri;
  g.tri = tri_dev;
  return launch_classify(g, gb, res, sub, inside_dev, status_dev, w_dev, st);
}

}  // namespace icc
