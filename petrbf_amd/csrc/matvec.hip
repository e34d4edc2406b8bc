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
// result is deterministic.  FP64 VALU-bound: one exponential per pair.  The
// coordinates enter as x' = (x - o) S, o the cluster's first target and
// S = sqrt(64 / (2 sigma^2 ln2)), so that -|x_i' - x_j'|^2 is the exponent in
// units of ln2/64 and exp2_64_nonpos (its 2^(j/64) table in LDS) finishes it
// without a reduction constant: 15 FP64 operations per pair (2D).  Local
// coordinates keep the differences accurate: |x - o| is at most the
// truncation box's reach, so each x' carries a relative rounding of 2^-53 of
// a number of the differences' own size.
#include "common.h"
#include "kernels.h"

namespace prbf {

constexpr int kTile = 128;  // sources staged per LDS tile (smaller tiles: more wavefronts per SIMD)

// 2^(j/64), j = 0..63, correctly rounded (from 60-digit decimal arithmetic):
// read once per workgroup instead of 64 FP64 exp2 calls
__constant__ double kExp2Tab64[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};


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
    double scale, double norm) {
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
  // local origin: the cluster's first target
  const double ox = xi[t0], oy = yi[t0], oz = DIM == 3 ? zi[t0] : 0.0;
  etab[lane] = kExp2Tab64[lane];  // (visible after the first tile's barrier)
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
      ax = (xi[t0 + ta] - ox) * scale;
      ay = (yi[t0 + ta] - oy) * scale;
      if (DIM == 3) az = (zi[t0 + ta] - oz) * scale;
      const int tbb = hasb ? tb2 : ta;
      bx = (xi[t0 + tbb] - ox) * scale;
      by = (yi[t0 + tbb] - oy) * scale;
      if (DIM == 3) bz = (zi[t0 + tbb] - oz) * scale;
    }
    double accA[NRHS], accB[NRHS];
#pragma unroll
    for (int r = 0; r < NRHS; ++r) accA[r] = accB[r] = 0.0;
    for (long long sb = s0; sb < s1; sb += kTile) {
      const int n = (int)min<long long>(kTile, s1 - sb);
      __syncthreads();
      for (int k = lane; k < n; k += 64) {
        int j = tidx[sb + k];
        sx[k] = (xj[j] - ox) * scale;
        sy[k] = (yj[j] - oy) * scale;
        if (DIM == 3) sz[k] = (zj[j] - oz) * scale;
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
          const double ea1 = exp2_64_nonpos(-ra1, etab), eb1 = exp2_64_nonpos(-rb1, etab);
          const double ea2 = exp2_64_nonpos(-ra2, etab), eb2 = exp2_64_nonpos(-rb2, etab);
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
          const double ea1 = exp2_64_nonpos(-ra1, etab), eb1 = exp2_64_nonpos(-rb1, etab);
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
  // coordinate scale: |dx'|^2 = |dx|^2 (64/ln2) / (2 sigma^2)
  const double scale = std::sqrt(k.inv_two_s2 * 92.33248261689366);
  dim3 grid((unsigned)ncell), block(64);
#define PRBF_MV(D, R)                                                                              \
  hipLaunchKernelGGL((k_matvec_wave<D, R>), grid, block, 0, st, ncell, tsta, tend, tptr, tidx, xi, yi, \
                     zi, xj, yj, zj, g, ldg, w, ldw, scale, k.norm)
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
