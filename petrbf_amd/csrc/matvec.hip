// Truncated Gaussian matvec / evaluation (kernel families (b) and (c)).
//
//   w_i = sum_{j in trunc(c(i))} g_j exp(-|x_i - x_j|^2 / (2 sigma^2)) / (2 pi sigma^2)
//
// as 2d/matmult.cxx:28-43 (MatShell MULT) and 2d/get_vorticity.h:11-26
// (evaluation at a distinct target set).  One wavefront per cluster: the
// cluster's truncation list (the same for all of its targets,
// 2d/get_trunc.h:63) is staged in LDS as SoA FP64 tiles; each lane owns one
// target and a residue class of the source tile, accumulates in FP64 with
// FMA, and the per-target partial sums are combined in a fixed order, so the
// result is deterministic.  FP64 VALU-bound: one exp per pair.
#include "common.h"
#include "kernels.h"

namespace prbf {

constexpr int kTile = 256;  // sources staged per LDS tile

template <int DIM, int NRHS>
__global__ __launch_bounds__(64) void k_matvec_wave(
    long long ncell, const int* __restrict__ tsta, const int* __restrict__ tend,
    const long long* __restrict__ tptr, const int* __restrict__ tidx,
    const double* __restrict__ xi, const double* __restrict__ yi, const double* __restrict__ zi,
    const double* __restrict__ xj, const double* __restrict__ yj, const double* __restrict__ zj,
    const double* __restrict__ g, long long ldg, double* __restrict__ w, long long ldw,
    double inv_two_s2, double norm) {
  __shared__ double sx[kTile], sy[kTile], sz[DIM == 3 ? kTile : 1];
  __shared__ double sg[NRHS][kTile];
  __shared__ double part[NRHS][64];
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  const int t0 = tsta[c], t1 = tend[c];
  if (t0 > t1) return;
  const long long s0 = tptr[c], s1 = tptr[c + 1];
  const int T = t1 - t0 + 1;
  for (int tb = 0; tb < T; tb += 64) {
    const int Tc = min(64, T - tb);
    const int Q = 64 / Tc;             // source phases per target
    const int t = lane % Tc, q = lane / Tc;
    const bool active = q < Q;
    const int it = t0 + tb + t;
    double px = 0, py = 0, pz = 0;
    if (active) {
      px = xi[it];
      py = yi[it];
      if (DIM == 3) pz = zi[it];
    }
    double acc[NRHS];
#pragma unroll
    for (int r = 0; r < NRHS; ++r) acc[r] = 0.0;
    for (long long sb = s0; sb < s1; sb += kTile) {
      const int n = (int)min<long long>(kTile, s1 - sb);
      __syncthreads();
      for (int k = lane; k < n; k += 64) {
        int j = tidx[sb + k];
        sx[k] = xj[j];
        sy[k] = yj[j];
        if (DIM == 3) sz[k] = zj[j];
#pragma unroll
        for (int r = 0; r < NRHS; ++r) sg[r][k] = g[r * ldg + j];
      }
      __syncthreads();
      if (active) {
        for (int k = q; k < n; k += Q) {
          double dx = px - sx[k];
          double dy = py - sy[k];
          double r2 = dx * dx + dy * dy;
          if (DIM == 3) {
            double dz = pz - sz[k];
            r2 += dz * dz;
          }
          double e = exp(-r2 * inv_two_s2);
#pragma unroll
          for (int r = 0; r < NRHS; ++r) acc[r] = fma(sg[r][k], e, acc[r]);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NRHS; ++r) part[r][lane] = active ? acc[r] : 0.0;
    __syncthreads();
    if (lane < Tc) {
#pragma unroll
      for (int r = 0; r < NRHS; ++r) {
        double s = 0.0;
        for (int qq = 0; qq < Q; ++qq) s += part[r][qq * Tc + lane];
        w[r * ldw + t0 + tb + lane] = s * norm;
      }
    }
  }
}

void matvec(hipStream_t st, int dim, int nrhs, long long ncell, const int* tsta, const int* tend,
            const long long* tptr, const int* tidx, const double* xi, const double* yi, const double* zi,
            const double* xj, const double* yj, const double* zj, const double* g, long long ldg, double* w,
            long long ldw, double sigma) {
  Kern k = make_kern(sigma);
  dim3 grid((unsigned)ncell), block(64);
#define PRBF_MV(D, R)                                                                              \
  hipLaunchKernelGGL((k_matvec_wave<D, R>), grid, block, 0, st, ncell, tsta, tend, tptr, tidx, xi, yi, \
                     zi, xj, yj, zj, g, ldg, w, ldw, k.inv_two_s2, k.norm)
  if (dim == 2) {
    if (nrhs == 1) PRBF_MV(2, 1);
    else if (nrhs == 2) PRBF_MV(2, 2);
    else PRBF_MV(2, 3);
  } else {
    if (nrhs == 1) PRBF_MV(3, 1);
    else if (nrhs == 2) PRBF_MV(3, 2);
    else PRBF_MV(3, 3);
  }
#undef PRBF_MV
}

}  // namespace prbf
