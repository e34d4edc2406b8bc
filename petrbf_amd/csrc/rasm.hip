// Restricted additive Schwarz (RASM) blocks: batched setup and apply, FP64.
//
// Reference: the dense untruncated block over each cluster's buffer particles
// (2d/matmult.cxx:83-93), factorised by PETSc's sub-PC (-sub_pc_type lu,
// 2d/Makefile:7) and applied by PCASM in its restrict form with subdomains
// is = buffer, is_local = core prefix (2d/rbf_interpolation.cxx:213-229):
//     z = sum_c P_core(c) A_c^{-1} R_buffer(c) r.
// The restrict form only ever reads the core rows of A_c^{-1}.  The setup
// therefore computes, once, M_c = (A_c^{-1})[core, buffer] (T_c x n_c) and the
// apply is a batched gather + small GEMV that streams M_c from HBM: T_c n_c
// doubles per cluster instead of the n_c^2 of a stored factor.
//
// The setup kernel lives in blockchol.hip; this file holds the apply.
#include "common.h"
#include "kernels.h"

namespace prbf {

// Apply: one wavefront per cluster.  Lanes stride over buffer slots b,
// gather r[b] once, and accumulate T partial sums against the rows of M_c
// (row-major T x n, so each load is a contiguous wave access); partials are
// combined through LDS in a fixed order.  MT = double (plain) or float (mixed;
// products and sums still in FP64).  In mixed mode a cluster whose FP32 block
// broke down carries FP64 rows in M64 at offset fb[c] (fb[c] < 0 otherwise).
constexpr int kApplyWaves = 4;  // clusters per workgroup, one per wavefront
constexpr int kKB = 8;          // buffer slots per lane per pass (n <= 512 in one pass)

// acc[t] += sum over this lane's slots of M[t0+t][b] * r[b] for one pass
template <int TMAX, typename MT>
__device__ inline void rows_times(const MT* __restrict__ Mc, int n, int t0, int tc, int bb, int lane,
                                  const double (&rb)[kKB], double (&acc)[TMAX]) {
#pragma unroll
  for (int t = 0; t < TMAX; ++t) {
    if (t < tc) {
      const long long row = (long long)(t0 + t) * n;
#pragma unroll
      for (int k = 0; k < kKB; ++k) {
        const int b = bb + lane + 64 * k;
        if (b < n) acc[t] = fma((double)Mc[row + b], rb[k], acc[t]);
      }
    }
  }
}

template <int TMAX, typename MT>
__global__ __launch_bounds__(64 * kApplyWaves) void k_rasm_apply(
    long long ncell, const long long* __restrict__ bptr, const int* __restrict__ bidx,
    const int* __restrict__ ista, const int* __restrict__ iend, const long long* __restrict__ mptr,
    const MT* __restrict__ M, const long long* __restrict__ fb, const double* __restrict__ M64,
    const double* __restrict__ r, double* __restrict__ z) {
  __shared__ double red[kApplyWaves][TMAX][65];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const long long c = (long long)blockIdx.x * kApplyWaves + wave;
  if (c >= ncell) return;
  const long long b0 = bptr[c];
  const int n = (int)(bptr[c + 1] - b0);
  if (n == 0) return;
  const int T = iend[c] - ista[c] + 1;
  const long long f = fb ? fb[c] : -1;
  const MT* Mc = M + mptr[c];
  const double* Md = f >= 0 ? M64 + f : nullptr;
  double (*rw)[65] = red[wave];
  for (int t0 = 0; t0 < T; t0 += TMAX) {
    const int tc = min(TMAX, T - t0);
    double acc[TMAX];
#pragma unroll
    for (int t = 0; t < TMAX; ++t) acc[t] = 0.0;
    for (int bb = 0; bb < n; bb += 64 * kKB) {
      // all index loads, then all gathers, then the streamed rows
      int idx[kKB];
      double rb[kKB];
#pragma unroll
      for (int k = 0; k < kKB; ++k) {
        const int b = bb + lane + 64 * k;
        idx[k] = b < n ? bidx[b0 + b] : -1;
      }
#pragma unroll
      for (int k = 0; k < kKB; ++k) rb[k] = idx[k] >= 0 ? r[idx[k]] : 0.0;
      if (Md)
        rows_times<TMAX, double>(Md, n, t0, tc, bb, lane, rb, acc);
      else
        rows_times<TMAX, MT>(Mc, n, t0, tc, bb, lane, rb, acc);
    }
#pragma unroll
    for (int t = 0; t < TMAX; ++t) rw[t][lane] = acc[t];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < tc) {
      double s = 0.0;
      for (int l = 0; l < 64; ++l) s += rw[lane][l];
      z[ista[c] + t0 + lane] = s;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

static inline dim3 apply_grid(long long ncell) { return dim3((unsigned)((ncell + kApplyWaves - 1) / kApplyWaves)); }

void rasm_apply_fp64(hipStream_t st, long long ncell, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const long long* mptr, const double* M,
                     const double* r, double* z, int tmax) {
  (void)tmax;
  hipLaunchKernelGGL((k_rasm_apply<16, double>), apply_grid(ncell), dim3(64 * kApplyWaves), 0, st, ncell, bptr,
                     bidx, ista, iend, mptr, M, nullptr, nullptr, r, z);
}

void rasm_apply_fp32(hipStream_t st, long long ncell, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const long long* mptr, const float* M,
                     const long long* fb, const double* M64, const double* r, double* z) {
  hipLaunchKernelGGL((k_rasm_apply<16, float>), apply_grid(ncell), dim3(64 * kApplyWaves), 0, st, ncell, bptr,
                     bidx, ista, iend, mptr, M, fb, M64, r, z);
}

}  // namespace prbf
