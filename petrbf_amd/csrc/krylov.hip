// Vector kernels of the Krylov solvers (GMRES / FGMRES), FP64.
//
// Reductions use a fixed partition of the index range into kRedBlocks slices
// and combine the slice partials in slice order, so every dot product and norm
// is identical from run to run.  The multi-dot reads w once for all k basis
// vectors (classical Gram-Schmidt, the PETSc default orthogonalisation), and
// the multi-AXPY fuses the norm of the updated vector into the same pass.
#include "common.h"
#include "kernels.h"

namespace prbf {

constexpr int kRT = 256;     // threads per reduction block
constexpr int kMaxK = 64;    // max basis vectors per fused pass

// Sums of v[0..NV) over the workgroup (kRT threads): a butterfly over each
// wavefront, then the wavefronts' totals through LDS in wavefront order; the
// results are valid in thread 0.  sh holds NV * kRT / 64 values.
template <int NV>
__device__ inline void block_sums(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] += __shfl_xor(v[j], o);
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) sh[j * (kRT / 64) + wave] = v[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double s = sh[j * (kRT / 64)];
      for (int w = 1; w < kRT / 64; ++w) s += sh[j * (kRT / 64) + w];
      v[j] = s;
    }
  }
  __syncthreads();
}

__device__ inline void slice(long long n, long long* a, long long* b) {
  long long len = (n + kRedBlocks - 1) / kRedBlocks;
  *a = min(n, (long long)blockIdx.x * len);
  *b = min(n, *a + len);
}

// part[blk * kMaxK + j] = sum over the slice of w * V_j, for j < k; 16
// vectors per pass over the slice (w read once for up to 16), so that a
// thread has 17 loads in flight per element
constexpr int kDotChunk = 16;
__global__ __launch_bounds__(kRT) void k_mdot(long long n, const double* __restrict__ w,
                                              const double* __restrict__ V, long long ldv, int k,
                                              double* __restrict__ part) {
  __shared__ double sh[kDotChunk * (kRT / 64)];
  long long a, b;
  slice(n, &a, &b);
  for (int j0 = 0; j0 < k; j0 += kDotChunk) {
    double acc[kDotChunk];
#pragma unroll
    for (int j = 0; j < kDotChunk; ++j) acc[j] = 0.0;
    const int kk = min(kDotChunk, k - j0);
    for (long long i = a + threadIdx.x; i < b; i += kRT) {
      const double wi = w[i];
#pragma unroll
      for (int j = 0; j < kDotChunk; ++j)
        if (j < kk) acc[j] = fma(wi, V[(long long)(j0 + j) * ldv + i], acc[j]);
    }
    block_sums<kDotChunk>(acc, sh);
    if (threadIdx.x == 0)
      for (int j = 0; j < kk; ++j) part[(long long)blockIdx.x * kMaxK + j0 + j] = acc[j];
  }
}

// out[j] = sum of the kRedBlocks slice partials of column j: one wavefront
// per column, lane l summing slices l, l + 64, ... in order, the lane sums
// then combined by a butterfly (a fixed order: identical run to run)
__global__ __launch_bounds__(64) void k_reduce_parts(const double* __restrict__ part, double* __restrict__ out) {
  const int j = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int b = lane; b < kRedBlocks; b += 64) s += part[(long long)b * kMaxK + j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[j] = s;
}

void mdot(hipStream_t st, long long n, const double* w, const double* V, long long ldv, int k, double* part,
          double* out) {
  hipLaunchKernelGGL(k_mdot, dim3(kRedBlocks), dim3(kRT), 0, st, n, w, V, ldv, k, part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(k), dim3(64), 0, st, part, out);
}

// w += sum_j alpha*coef[j] V_j, then part = partial sums of w^2
__global__ __launch_bounds__(kRT) void k_maxpy_norm2(long long n, double* __restrict__ w,
                                                     const double* __restrict__ V, long long ldv, int k,
                                                     const double* __restrict__ coef, double alpha,
                                                     double* __restrict__ part) {
  __shared__ double sh[kRT / 64];
  __shared__ double cf[kMaxK];
  if (threadIdx.x < k) cf[threadIdx.x] = alpha * coef[threadIdx.x];
  __syncthreads();
  long long a, b;
  slice(n, &a, &b);
  double acc = 0.0;
  for (long long i = a + threadIdx.x; i < b; i += kRT) {
    double s = w[i];
    for (int j = 0; j < k; ++j) s = fma(cf[j], V[(long long)j * ldv + i], s);
    w[i] = s;
    acc = fma(s, s, acc);
  }
  double v[1] = {acc};
  block_sums<1>(v, sh);
  if (threadIdx.x == 0) part[(long long)blockIdx.x * kMaxK] = v[0];
}

void maxpy_norm2(hipStream_t st, long long n, double* w, const double* V, long long ldv, int k,
                 const double* coef, double alpha, double* part, double* out_norm2) {
  hipLaunchKernelGGL(k_maxpy_norm2, dim3(kRedBlocks), dim3(kRT), 0, st, n, w, V, ldv, k, coef, alpha, part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(64), 0, st, part, out_norm2);
}

__global__ __launch_bounds__(kRT) void k_norm2sq(long long n, const double* __restrict__ v,
                                                 double* __restrict__ part) {
  __shared__ double sh[kRT / 64];
  long long a, b;
  slice(n, &a, &b);
  double acc = 0.0;
  for (long long i = a + threadIdx.x; i < b; i += kRT) acc = fma(v[i], v[i], acc);
  double r[1] = {acc};
  block_sums<1>(r, sh);
  if (threadIdx.x == 0) part[(long long)blockIdx.x * kMaxK] = r[0];
}

void norm2sq(hipStream_t st, long long n, const double* v, double* part, double* out) {
  hipLaunchKernelGGL(k_norm2sq, dim3(kRedBlocks), dim3(kRT), 0, st, n, v, part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(64), 0, st, part, out);
}

__global__ void k_scale(long long n, double* __restrict__ v, double s) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= s;
}
__global__ void k_scale_copy(long long n, const double* __restrict__ a, double* __restrict__ b, double s) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i] * s;
}
__global__ void k_sub(long long n, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ r) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = a[i] - b[i];
}
__global__ void k_update_x(long long n, double* __restrict__ x, const double* __restrict__ V, long long ldv,
                           int k, const double* __restrict__ y) {
  __shared__ double cf[kMaxK];
  if (threadIdx.x < k) cf[threadIdx.x] = y[threadIdx.x];
  __syncthreads();
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s = fma(cf[j], V[(long long)j * ldv + i], s);
  x[i] += s;
}
__global__ void k_gather(long long n, const int* __restrict__ perm, const double* __restrict__ src,
                         double* __restrict__ dst) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
__global__ void k_scatter(long long n, const int* __restrict__ perm, const double* __restrict__ src,
                          double* __restrict__ dst) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}
__global__ void k_fill(long long n, double* __restrict__ v, double a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = a;
}

static inline dim3 g1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

void scale(hipStream_t st, long long n, double* v, double s) {
  hipLaunchKernelGGL(k_scale, g1(n), dim3(256), 0, st, n, v, s);
}
void scale_copy(hipStream_t st, long long n, const double* src, double* dst, double s) {
  hipLaunchKernelGGL(k_scale_copy, g1(n), dim3(256), 0, st, n, src, dst, s);
}
void xpay_sub(hipStream_t st, long long n, const double* b, const double* ax, double* r) {
  hipLaunchKernelGGL(k_sub, g1(n), dim3(256), 0, st, n, b, ax, r);
}
void update_x(hipStream_t st, long long n, double* x, const double* V, long long ldv, int k, const double* y) {
  hipLaunchKernelGGL(k_update_x, g1(n), dim3(256), 0, st, n, x, V, ldv, k, y);
}
void gather(hipStream_t st, long long n, const int* perm, const double* src, double* dst) {
  hipLaunchKernelGGL(k_gather, g1(n), dim3(256), 0, st, n, perm, src, dst);
}
void scatter(hipStream_t st, long long n, const int* perm, const double* src, double* dst) {
  hipLaunchKernelGGL(k_scatter, g1(n), dim3(256), 0, st, n, perm, src, dst);
}
void fill(hipStream_t st, long long n, double* v, double a) {
  hipLaunchKernelGGL(k_fill, g1(n), dim3(256), 0, st, n, v, a);
}


// ---- device-side Arnoldi bookkeeping (for iterations captured as graphs) --
// State of one restart cycle, column-major H with leading dimension m + 1:
//   ks[0 .. (m+1)^2)            H
//   ks[kc .. kc+m+1)            cosines, kc = (m+1)^2
//   ks[kc+m+1 .. kc+2m+2)       sines
//   ks[kc+2m+2 .. kc+3m+4)      rs (rotated right-hand side)
//   ks[kc+3m+4]                 1 / h_{j+1,j} (1 if zero) for the next basis vector
//   ks[kc+3m+5 .. kc+3m+8)      out: residual estimate, happy flag, breakdown flag
// The arithmetic is the host loop's, operation for operation and without
// contraction into fma, so both paths give the same bits.
long long arnoldi_state_len(int m) { return (long long)(m + 1) * (m + 1) + 3 * m + 8; }

__global__ void k_arnoldi_step(int j, int m, const double* __restrict__ dsc, double* __restrict__ ks) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int ld = m + 1, kc = ld * ld;
  double* H = ks;
  double* cs = ks + kc;
  double* sn = cs + m + 1;
  double* rs = sn + m + 1;
  double* inv = rs + m + 2;
  double* out = inv + 1;
  for (int i = 0; i <= j; ++i) H[j * ld + i] = dsc[i];
  const double tt = sqrt(dsc[64]);
  *inv = tt != 0.0 ? 1.0 / tt : 1.0;
  H[j * ld + j + 1] = tt;
  const double q = fabs(tt / rs[j]);
  const double hapbnd = 1e-30 < q ? 1e-30 : q;  // std::min(q, 1e-30) of the host loop
  const bool hapend = tt < hapbnd;
  for (int i = 1; i <= j; ++i) {
    const double a0 = H[j * ld + i - 1];
    H[j * ld + i - 1] = cs[i - 1] * a0 + sn[i - 1] * H[j * ld + i];
    H[j * ld + i] = cs[i - 1] * H[j * ld + i] - sn[i - 1] * a0;
  }
  double res = 0.0, breakdown = 0.0;
  if (!hapend) {
    const double a0 = H[j * ld + j], a1 = H[j * ld + j + 1];
    const double t = sqrt(a0 * a0 + a1 * a1);
    if (t == 0.0) {
      breakdown = 1.0;
    } else {
      cs[j] = a0 / t;
      sn[j] = a1 / t;
      rs[j + 1] = -(sn[j] * rs[j]);
      rs[j] = cs[j] * rs[j];
      H[j * ld + j] = cs[j] * a0 + sn[j] * a1;
      res = fabs(rs[j + 1]);
    }
  }
  out[0] = res;
  out[1] = hapend ? 1.0 : 0.0;
  out[2] = breakdown;
}

// y = H(0:it, 0:it)^{-1} rs(0:it) by back substitution (upper triangular)
__global__ void k_arnoldi_backsub(int it, int m, const double* __restrict__ ks, double* __restrict__ y) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int ld = m + 1, kc = ld * ld;
  const double* H = ks;
  const double* rs = ks + kc + 2 * (m + 1);
  y[it] = rs[it] / H[it * ld + it];
  for (int ii = 1; ii <= it; ++ii) {
    const int k = it - ii;
    double t = rs[k];
    for (int j = k + 1; j <= it; ++j) t = t - H[j * ld + k] * y[j];
    y[k] = t / H[k * ld + k];
  }
}

__global__ void k_scale_dev(long long n, double* __restrict__ v, const double* __restrict__ s) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= *s;
}

__global__ void k_set1(double* __restrict__ p, double a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = a;
}

void arnoldi_step(hipStream_t st, int j, int m, const double* dsc, double* ks) {
  hipLaunchKernelGGL(k_arnoldi_step, dim3(1), dim3(64), 0, st, j, m, dsc, ks);
}
void arnoldi_backsub(hipStream_t st, int it, int m, const double* ks, double* y) {
  hipLaunchKernelGGL(k_arnoldi_backsub, dim3(1), dim3(64), 0, st, it, m, ks, y);
}
void scale_dev(hipStream_t st, long long n, double* v, const double* s) {
  hipLaunchKernelGGL(k_scale_dev, g1(n), dim3(256), 0, st, n, v, s);
}
void set_scalar(hipStream_t st, double* p, double a) { hipLaunchKernelGGL(k_set1, dim3(1), dim3(64), 0, st, p, a); }

}  // namespace prbf
