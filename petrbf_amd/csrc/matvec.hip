// Truncated Gaussian matvec / evaluation (kernel families (b) and (c)).
//
//   w_i = sum_{j in trunc(c(i))} g_j exp(-|x_i - x_j|^2 / (2 sigma^2)) / (2 pi sigma^2)
//
// as 2d/matmult.cxx:28-43 (MatShell MULT) and 2d/get_vorticity.h:11-26
// (evaluation at a distinct target set).  One wavefront per cluster: the
// cluster's truncation list (the same for all of its targets,
// 2d/get_trunc.h:63) is staged in LDS as SoA FP64 tiles; each lane owns two
// targets and a residue class of the source tile, accumulates in FP64 with
// FMA, and the per-target partial sums are combined in a fixed order, so the
// result is deterministic.  FP64 VALU-bound: one exp (exp_nonpos_tab, with
// its 2^(j/64) table in LDS) per pair.
#include "common.h"
#include "kernels.h"

namespace prbf {

constexpr int kTile = 256;  // sources staged per LDS tile

// Lanes own two targets each (t and t + TP) and a residue class q of the
// staged sources, so every source read from LDS feeds two pair evaluations;
// the source loop is unrolled by two for independent exps.
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
  __shared__ double part[NRHS][128];
  __shared__ double etab[64];
  const long long c = blockIdx.x;
  const int lane = threadIdx.x;
  const int t0 = tsta[c], t1 = tend[c];
  if (t0 > t1) return;
  const long long s0 = tptr[c], s1 = tptr[c + 1];
  const int T = t1 - t0 + 1;
  const double nc = -inv_two_s2;
  etab[lane] = exp2(lane * (1.0 / 64));  // (visible after the first tile's barrier)
  for (int tb = 0; tb < T; tb += 128) {
    const int Tc = min(128, T - tb);
    const int TP = (Tc + 1) / 2;       // target pairs
    const int Q = 64 / TP;             // source phases
    const int pr = lane % TP, q = lane / TP;
    const bool active = q < Q;
    const int ta = tb + pr, tb2 = tb + pr + TP;
    const bool hasb = tb2 < tb + Tc;
    double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
    if (active) {
      ax = xi[t0 + ta];
      ay = yi[t0 + ta];
      if (DIM == 3) az = zi[t0 + ta];
      const int tbb = hasb ? tb2 : ta;
      bx = xi[t0 + tbb];
      by = yi[t0 + tbb];
      if (DIM == 3) bz = zi[t0 + tbb];
    }
    double accA[NRHS], accB[NRHS];
#pragma unroll
    for (int r = 0; r < NRHS; ++r) accA[r] = accB[r] = 0.0;
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
        int k = q;
        for (; k + Q < n; k += 2 * Q) {
          const int k2 = k + Q;
          const double x1 = sx[k], y1 = sy[k], x2 = sx[k2], y2 = sy[k2];
          double dxa1 = ax - x1, dya1 = ay - y1, dxb1 = bx - x1, dyb1 = by - y1;
          double dxa2 = ax - x2, dya2 = ay - y2, dxb2 = bx - x2, dyb2 = by - y2;
          double ra1 = fma(dxa1, dxa1, dya1 * dya1), rb1 = fma(dxb1, dxb1, dyb1 * dyb1);
          double ra2 = fma(dxa2, dxa2, dya2 * dya2), rb2 = fma(dxb2, dxb2, dyb2 * dyb2);
          if (DIM == 3) {
            const double z1 = sz[k], z2 = sz[k2];
            ra1 = fma(az - z1, az - z1, ra1);
            rb1 = fma(bz - z1, bz - z1, rb1);
            ra2 = fma(az - z2, az - z2, ra2);
            rb2 = fma(bz - z2, bz - z2, rb2);
          }
          const double ea1 = exp_nonpos_tab(ra1 * nc, etab), eb1 = exp_nonpos_tab(rb1 * nc, etab);
          const double ea2 = exp_nonpos_tab(ra2 * nc, etab), eb2 = exp_nonpos_tab(rb2 * nc, etab);
#pragma unroll
          for (int r = 0; r < NRHS; ++r) {
            const double g1 = sg[r][k], g2 = sg[r][k2];
            accA[r] = fma(g1, ea1, accA[r]);
            accB[r] = fma(g1, eb1, accB[r]);
            accA[r] = fma(g2, ea2, accA[r]);
            accB[r] = fma(g2, eb2, accB[r]);
          }
        }
        if (k < n) {
          const double x1 = sx[k], y1 = sy[k];
          double ra1 = fma(ax - x1, ax - x1, (ay - y1) * (ay - y1));
          double rb1 = fma(bx - x1, bx - x1, (by - y1) * (by - y1));
          if (DIM == 3) {
            const double z1 = sz[k];
            ra1 = fma(az - z1, az - z1, ra1);
            rb1 = fma(bz - z1, bz - z1, rb1);
          }
          const double ea1 = exp_nonpos_tab(ra1 * nc, etab), eb1 = exp_nonpos_tab(rb1 * nc, etab);
#pragma unroll
          for (int r = 0; r < NRHS; ++r) {
            accA[r] = fma(sg[r][k], ea1, accA[r]);
            accB[r] = fma(sg[r][k], eb1, accB[r]);
          }
        }
      }
    }
    __syncthreads();
    // partials: part[r][qq*2TP + p] for target p (< 2TP); combined in phase order
#pragma unroll
    for (int r = 0; r < NRHS; ++r) {
      if (active) {
        part[r][q * 2 * TP + pr] = accA[r];
        part[r][q * 2 * TP + TP + pr] = accB[r];
      }
    }
    __syncthreads();
    // target tt sits in slot tt (pair tt for tt < TP, else the B half of pair tt - TP)
    for (int tt = lane; tt < Tc; tt += 64) {
#pragma unroll
      for (int r = 0; r < NRHS; ++r) {
        double s = 0.0;
        for (int qq = 0; qq < Q; ++qq) s += part[r][qq * 2 * TP + tt];
        w[r * ldw + t0 + tb + tt] = s * norm;
      }
    }
    __syncthreads();
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
