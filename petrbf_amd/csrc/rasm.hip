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
// Setup per block (one workgroup of 256 threads, persistent over its block):
//   1. order the buffer with the core LAST and build the lower triangle of A'
//      (the same Gaussian, symmetric permutation) in a global workspace;
//   2. right-looking blocked Cholesky A' = L L^T, panel width 16, diagonal
//      tile and panel in LDS;
//   3. with the core last, L^{-1} E_core = [0; L_cc^{-1}], so
//      A'^{-1} E_core = L^{-T} [0; L_cc^{-1}]: a T x T inverse in LDS, then
//      a blocked back substitution over 16-row slabs;
//   4. write M_c in buffer order, layout [b][t] (t fastest).
// Cholesky instead of the reference's pivoted LU: the blocks are SPD and the
// solve is exact in FP64 either way; the apply only needs core rows.
#include "common.h"
#include "kernels.h"

namespace prbf {

constexpr int kNB = 16;       // panel width
constexpr int kThreads = 256;

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_rasm_factor_fp64(
    long long c0, const long long* __restrict__ bptr, const int* __restrict__ bidx,
    const int* __restrict__ ista, const int* __restrict__ iend, const double* __restrict__ xs,
    const double* __restrict__ ys, const double* __restrict__ zs, const long long* __restrict__ wptr,
    double* __restrict__ Wbase, const long long* __restrict__ mptr, double* __restrict__ M,
    double inv_two_s2, double norm, int* __restrict__ err) {
  extern __shared__ double dyn[];
  __shared__ double D[kNB][kNB + 1];
  __shared__ double piv_ok;
  const long long c = c0 + blockIdx.x;
  const long long b0 = bptr[c];
  const int n = (int)(bptr[c + 1] - b0);
  if (n == 0) return;
  const int T = iend[c] - ista[c] + 1;
  const int ncore0 = n - T;  // first position of the core in the permuted order
  const int tid = threadIdx.x;
  double* W = Wbase + wptr[c - c0];
  double* cx = dyn;
  double* cy = dyn + n;
  double* cz = dyn + 2 * n;
  double* work = dyn + 3 * n;  // max(kNB, T) * n doubles: panel, then X

  // permuted position p -> buffer slot: core (slots 0..T-1) moved to the end
  for (int p = tid; p < n; p += kThreads) {
    int b = p < ncore0 ? p + T : p - ncore0;
    int j = bidx[b0 + b];
    cx[p] = xs[j];
    cy[p] = ys[j];
    if (DIM == 3) cz[p] = zs[j];
  }
  if (tid == 0) piv_ok = 1.0;
  __syncthreads();

  // 1. lower triangle of A' (column-major, ld = n)
  for (int q = 0; q < n; ++q) {
    double qx = cx[q], qy = cy[q], qz = DIM == 3 ? cz[q] : 0.0;
    for (int p = q + tid; p < n; p += kThreads) {
      double dx = cx[p] - qx, dy = cy[p] - qy;
      double r2 = dx * dx + dy * dy;
      if (DIM == 3) {
        double dz = cz[p] - qz;
        r2 += dz * dz;
      }
      W[(long long)q * n + p] = exp(-r2 * inv_two_s2) * norm;
    }
  }
  __syncthreads();

  // 2. blocked right-looking Cholesky
  double* PT = work;  // PT[p * n + i]: panel transposed (row i of the panel below the diagonal tile)
  for (int k0 = 0; k0 < n; k0 += kNB) {
    const int kb = min(kNB, n - k0);
    if (tid < kb * kb) {
      int p = tid % kb, q = tid / kb;
      if (p >= q) D[p][q] = W[(long long)(k0 + q) * n + k0 + p];
    }
    __syncthreads();
    for (int p = 0; p < kb; ++p) {
      if (tid == 0) {
        double d = D[p][p];
        if (!(d > 0.0)) { piv_ok = 0.0; d = 1.0; }
        D[p][p] = sqrt(d);
      }
      __syncthreads();
      if (tid > p && tid < kb) D[tid][p] /= D[p][p];
      __syncthreads();
      if (tid > p && tid < kb)
        for (int q = p + 1; q <= tid; ++q) D[tid][q] -= D[tid][p] * D[q][p];
      __syncthreads();
    }
    if (tid < kb * kb) {
      int p = tid % kb, q = tid / kb;
      if (p >= q) W[(long long)(k0 + q) * n + k0 + p] = D[p][q];
    }
    const int m0 = k0 + kb, H = n - m0;
    // panel: x L_kk^T = a  for each row below the tile
    for (int ii = tid; ii < H; ii += kThreads) {
      double a[kNB];
#pragma unroll
      for (int p = 0; p < kNB; ++p)
        if (p < kb) a[p] = W[(long long)(k0 + p) * n + m0 + ii];
#pragma unroll
      for (int p = 0; p < kNB; ++p) {
        if (p < kb) {
          double x = a[p];
          for (int q = 0; q < p; ++q) x -= a[q] * D[p][q];
          a[p] = x / D[p][p];
          W[(long long)(k0 + p) * n + m0 + ii] = a[p];
          PT[p * n + ii] = a[p];
        }
      }
    }
    __syncthreads();
    // trailing lower update: A[i][j] -= sum_p P[i][p] P[j][p], i >= j
    const int wave = tid / kWave, lane = tid % kWave;
    for (int jj = wave; jj < H; jj += kThreads / kWave) {
      double pj[kNB];
#pragma unroll
      for (int p = 0; p < kNB; ++p) pj[p] = p < kb ? PT[p * n + jj] : 0.0;
      double* col = W + (long long)(m0 + jj) * n + m0;
      for (int ii = jj + lane; ii < H; ii += kWave) {
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < kNB; ++p)
          if (p < kb) s = fma(PT[p * n + ii], pj[p], s);
        col[ii] -= s;
      }
    }
    __syncthreads();
  }
  if (piv_ok == 0.0) {
    if (tid == 0) atomicOr(err, 2);
    return;
  }

  // 3a. X_c = L_cc^{-T} L_cc^{-1} (T x T) for the trailing core block.
  double* X = work;  // X[i * T + t], i = permuted position
  if (tid < T) {
    const int t = tid;
    // column t of Z = L_cc^{-1}, stored into X rows ncore0.. (zero above t)
    for (int i = 0; i < T; ++i) {
      double z;
      if (i < t) {
        z = 0.0;
      } else {
        z = (i == t) ? 1.0 : 0.0;
        for (int q = t; q < i; ++q) z -= W[(long long)(ncore0 + q) * n + ncore0 + i] * X[(ncore0 + q) * T + t];
        z /= W[(long long)(ncore0 + i) * n + ncore0 + i];
      }
      X[(ncore0 + i) * T + t] = z;
    }
    // back: L_cc^T x = z
    for (int i = T - 1; i >= 0; --i) {
      double x = X[(ncore0 + i) * T + t];
      for (int k = i + 1; k < T; ++k) x -= W[(long long)(ncore0 + i) * n + ncore0 + k] * X[(ncore0 + k) * T + t];
      X[(ncore0 + i) * T + t] = x / W[(long long)(ncore0 + i) * n + ncore0 + i];
    }
  }
  __syncthreads();
  // 3b. rows above the core: L^T X = [0; Z], slabs of kNB rows, bottom-up.
  for (int i1 = ncore0; i1 > 0; i1 -= kNB) {
    const int i0 = max(0, i1 - kNB), nb = i1 - i0;
    // X[i][t] = -sum_{k >= i1} L[k][i] X[k][t] for i in [i0, i1)
    for (int e = tid; e < nb * T; e += kThreads) {
      int r = e / T, t = e % T, i = i0 + r;
      const double* Lcol = W + (long long)i * n;
      double s = 0.0;
      for (int k = i1; k < n; ++k) s = fma(Lcol[k], X[k * T + t], s);
      X[i * T + t] = -s;
    }
    __syncthreads();
    // triangular part inside the slab: L_slab^T x = rhs, bottom-up
    for (int i = i1 - 1; i >= i0; --i) {
      if (tid < T) {
        const double* Lcol = W + (long long)i * n;
        double x = X[i * T + tid];
        for (int k = i + 1; k < i1; ++k) x -= Lcol[k] * X[k * T + tid];
        X[i * T + tid] = x / Lcol[i];
      }
      __syncthreads();
    }
  }
  // 4. M_c[b][t] = X[p(b)][t]
  double* Mc = M + mptr[c];
  for (int e = tid; e < n * T; e += kThreads) {
    int b = e / T, t = e % T;
    int p = b < T ? ncore0 + b : b - T;
    Mc[e] = X[p * T + t];
  }
}

void rasm_factor_fp64(hipStream_t st, int dim, long long c0, long long ncl, int nmax, int tmax,
                      const long long* bptr, const int* bidx, const int* ista, const int* iend,
                      const double* xs, const double* ys, const double* zs, const long long* wptr,
                      double* W, const long long* mptr, double* M, double sigma, int* err) {
  Kern k = make_kern(sigma);
  size_t lds = sizeof(double) * ((size_t)3 * nmax + (size_t)std::max(kNB, tmax) * nmax);
  static bool attr_set = false;
  if (!attr_set) {
    // allow up to the CU's LDS for the dynamic part (default cap is 64 KiB)
    PRBF_HIP(hipFuncSetAttribute((const void*)k_rasm_factor_fp64<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
    PRBF_HIP(hipFuncSetAttribute((const void*)k_rasm_factor_fp64<3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
    attr_set = true;
  }
  if (dim == 2)
    hipLaunchKernelGGL(k_rasm_factor_fp64<2>, dim3((unsigned)ncl), dim3(kThreads), lds, st, c0, bptr, bidx,
                       ista, iend, xs, ys, zs, wptr, W, mptr, M, k.inv_two_s2, k.norm, err);
  else
    hipLaunchKernelGGL(k_rasm_factor_fp64<3>, dim3((unsigned)ncl), dim3(kThreads), lds, st, c0, bptr, bidx,
                       ista, iend, xs, ys, zs, wptr, W, mptr, M, k.inv_two_s2, k.norm, err);
}

// Apply: one wavefront per cluster.  Lanes stride over buffer slots, gather
// r, and accumulate T partial sums against the rows of M_c; partials are then
// combined through LDS in a fixed order.
template <int TMAX>
__global__ __launch_bounds__(64) void k_rasm_apply_fp64(const long long* __restrict__ bptr,
                                                        const int* __restrict__ bidx,
                                                        const int* __restrict__ ista,
                                                        const int* __restrict__ iend,
                                                        const long long* __restrict__ mptr,
                                                        const double* __restrict__ M,
                                                        const double* __restrict__ r, double* __restrict__ z) {
  __shared__ double red[TMAX][65];
  const long long c = blockIdx.x;
  const long long b0 = bptr[c];
  const int n = (int)(bptr[c + 1] - b0);
  if (n == 0) return;
  const int T = iend[c] - ista[c] + 1;
  const int lane = threadIdx.x;
  const double* Mc = M + mptr[c];
  for (int t0 = 0; t0 < T; t0 += TMAX) {
    const int tc = min(TMAX, T - t0);
    double acc[TMAX];
#pragma unroll
    for (int t = 0; t < TMAX; ++t) acc[t] = 0.0;
    for (int b = lane; b < n; b += 64) {
      double rb = r[bidx[b0 + b]];
      const double* row = Mc + (long long)b * T + t0;
#pragma unroll
      for (int t = 0; t < TMAX; ++t)
        if (t < tc) acc[t] = fma(row[t], rb, acc[t]);
    }
#pragma unroll
    for (int t = 0; t < TMAX; ++t) red[t][lane] = acc[t];
    __syncthreads();
    if (lane < tc) {
      double s = 0.0;
      for (int l = 0; l < 64; ++l) s += red[lane][l];
      z[ista[c] + t0 + lane] = s;
    }
    __syncthreads();
  }
}

void rasm_apply_fp64(hipStream_t st, long long ncell, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const long long* mptr, const double* M,
                     const double* r, double* z, int tmax) {
  (void)tmax;
  hipLaunchKernelGGL(k_rasm_apply_fp64<16>, dim3((unsigned)ncell), dim3(64), 0, st, bptr, bidx, ista, iend,
                     mptr, M, r, z);
}

}  // namespace prbf
