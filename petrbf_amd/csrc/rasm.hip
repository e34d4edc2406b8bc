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
// values per cluster (T_c rounded up to 16) instead of the n_c^2 of a stored
// factor.
//
// The setup kernel lives in blockchol.hip; this file holds the apply.
#include "common.h"
#include "kernels.h"

namespace prbf {

// Apply: one wavefront per cluster.  Lane l owns the buffer slots l, l+64,
// ... (one per slot group), gathers r at those slots once, and accumulates 16
// row partials against the core rows, which it reads as 16-byte vectors from
// contiguous slabs of up to 1 KB (the storage order of apply_offset; only the
// rows stored, T rounded up to 4, are read).  The 16 x 64
// partials are combined by a butterfly over the wavefront (reduce-scatter by
// halves, then two full exchanges): a fixed order, so the result is identical
// run to run.  MT = double (plain) or float (mixed; products and sums still
// in FP64).  Each launch covers a list of clusters, so that the mixed
// preconditioner can serve its small blocks in FP64 and the learned one can
// leave its network clusters out.  Lanes past the end of
// the last group read a valid slot of the same group with r taken as 0, so
// the loads need no branches.
constexpr int kApplyWaves = 2;  // clusters per workgroup, one per wavefront (FP32: 2% faster than 4)

template <typename ET>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

// Residuals at this lane's buffer slots (one per slot group, up to kMaxGroups
// groups, i.e. n <= 64 kMaxGroups) into the wavefront's LDS slice
// rs[(v kMaxGroups + g) 64 + lane] for each of the NR vectors v (stride ldr
// in r): gathered once per cluster, before the core rows stream in, so the
// index -> residual dependency is paid once.
constexpr int kMaxGroups = 8;
template <int NR>
__device__ __forceinline__ void gather_slots(int n, int lane, const int* __restrict__ bi, const double* __restrict__ r,
                                             long long ldr, double* __restrict__ rs) {
  const int ngrp = (n + 63) / 64;
  int idx[kMaxGroups];
#pragma unroll
  for (int g = 0; g < kMaxGroups; ++g) {
    const int gg = g < ngrp ? g : ngrp - 1;
    const int ng = min(64, n - 64 * gg);
    idx[g] = bi[64 * gg + min(lane, ng - 1)];
  }
#pragma unroll
  for (int v = 0; v < NR; ++v) {
#pragma unroll
    for (int g = 0; g < kMaxGroups; ++g) {
      const int gg = g < ngrp ? g : ngrp - 1;
      const int ng = min(64, n - 64 * gg);
      const double x = r[v * ldr + idx[g]];
      rs[(v * kMaxGroups + g) * 64 + lane] = (g < ngrp && lane < ng) ? x : 0.0;
    }
  }
}

// acc[v][t] += sum over this lane's slots of M[t][b] r_v[b], the R stored rows
// of one chunk (rows R..15 of acc stay zero); KG slot groups per pass (all
// loads of a pass are issued before the FMAs), residuals from the LDS slice.
// The core rows are read once per apply, for all NR vectors: non-temporal
// loads keep them from displacing the residuals that neighbouring clusters
// gather.
template <typename ET, int KG, int NR>
__device__ __forceinline__ void chunk_times(const ET* __restrict__ Mk, int n, int R, int lane,
                                            const double* __restrict__ rs, double (&acc)[NR][kRowChunk]) {
  constexpr int V = 16 / (int)sizeof(ET), QN = kRowChunk / V;
  using VT = typename Vec16<ET>::type;
  const int ngrp = (n + 63) / 64, qn = R / V;
  for (int g0 = 0; g0 < ngrp; g0 += KG) {
    VT m[KG][QN];
#pragma unroll
    for (int k = 0; k < KG; ++k) {
      const int g = g0 + k < ngrp ? g0 + k : ngrp - 1;
      const int ng = min(64, n - 64 * g);
      const VT* p = reinterpret_cast<const VT*>(Mk + (long long)R * 64 * g) + min(lane, ng - 1);
#pragma unroll
      for (int q = 0; q < QN; ++q) m[k][q] = q < qn ? __builtin_nontemporal_load(p + q * ng) : VT{};
    }
#pragma unroll
    for (int k = 0; k < KG; ++k) {
#pragma unroll
      for (int w = 0; w < NR; ++w) {
        const double rb = g0 + k < ngrp ? rs[(w * kMaxGroups + g0 + k) * 64 + lane] : 0.0;
#pragma unroll
        for (int q = 0; q < QN; ++q) {
#pragma unroll
          for (int v = 0; v < V; ++v) acc[w][q * V + v] = fma((double)m[k][q][v], rb, acc[w][q * V + v]);
        }
      }
    }
  }
}

// The same for blocks of more than 64 kMaxGroups slots: indices and residuals
// gathered pass by pass, KG slot groups per pass
template <typename ET, int KG, int NR>
__device__ __forceinline__ void chunk_times_gather(const ET* __restrict__ Mk, int n, int R, int lane,
                                                   const int* __restrict__ bi, const double* __restrict__ r,
                                                   long long ldr, double (&acc)[NR][kRowChunk]) {
  constexpr int V = 16 / (int)sizeof(ET), QN = kRowChunk / V;
  using VT = typename Vec16<ET>::type;
  const int ngrp = (n + 63) / 64, qn = R / V;
  for (int g0 = 0; g0 < ngrp; g0 += KG) {
    int idx[KG];
    bool ok[KG];
    VT m[KG][QN];
#pragma unroll
    for (int k = 0; k < KG; ++k) {
      const int g = g0 + k < ngrp ? g0 + k : ngrp - 1;
      const int ng = min(64, n - 64 * g);
      const int l = min(lane, ng - 1);
      ok[k] = (g0 + k < ngrp) && lane < ng;
      idx[k] = bi[64 * g + l];
      const VT* p = reinterpret_cast<const VT*>(Mk + (long long)R * 64 * g) + l;
#pragma unroll
      for (int q = 0; q < QN; ++q) m[k][q] = q < qn ? p[q * ng] : VT{};
    }
    double rb[KG][NR];
#pragma unroll
    for (int k = 0; k < KG; ++k) {
#pragma unroll
      for (int w = 0; w < NR; ++w) {
        const double x = r[w * ldr + idx[k]];
        rb[k][w] = ok[k] ? x : 0.0;
      }
    }
#pragma unroll
    for (int k = 0; k < KG; ++k) {
#pragma unroll
      for (int w = 0; w < NR; ++w) {
#pragma unroll
        for (int q = 0; q < QN; ++q) {
          const ET* e = reinterpret_cast<const ET*>(&m[k][q]);
#pragma unroll
          for (int v = 0; v < V; ++v) acc[w][q * V + v] = fma((double)e[v], rb[k][w], acc[w][q * V + v]);
        }
      }
    }
  }
}

// Sum of acc[t] over the wavefront; on return lanes with (lane & 3) == 0 hold
// the total of row t = 8 b5 + 4 b4 + 2 b3 + b2 (b_i = bit i of the lane).
__device__ inline double wave_rows_sum(const double (&acc)[kRowChunk], int lane) {
  double v8[8], v4[4], v2[2];
  {
    const bool hi = lane & 32;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double keep = hi ? acc[8 + i] : acc[i], send = hi ? acc[i] : acc[8 + i];
      v8[i] = keep + __shfl_xor(send, 32);
    }
  }
  {
    const bool hi = lane & 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double keep = hi ? v8[4 + i] : v8[i], send = hi ? v8[i] : v8[4 + i];
      v4[i] = keep + __shfl_xor(send, 16);
    }
  }
  {
    const bool hi = lane & 8;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double keep = hi ? v4[2 + i] : v4[i], send = hi ? v4[i] : v4[2 + i];
      v2[i] = keep + __shfl_xor(send, 8);
    }
  }
  const bool hi = lane & 4;
  double v = (hi ? v2[1] : v2[0]) + __shfl_xor(hi ? v2[0] : v2[1], 4);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 1);
  return v;
}

template <typename ET, int KG, int NR>
__device__ __forceinline__ void apply_cluster(const ET* __restrict__ Mc, int n, int T, int lane, const int* __restrict__ bi,
                                              const double* __restrict__ r, long long ld, double* __restrict__ rs,
                                              double* __restrict__ zc) {
  const int trow = 8 * ((lane >> 5) & 1) + 4 * ((lane >> 4) & 1) + 2 * ((lane >> 3) & 1) + ((lane >> 2) & 1);
  const bool small = n <= 64 * kMaxGroups;
  if (small) gather_slots<NR>(n, lane, bi, r, ld, rs);
  for (int t0 = 0; t0 < T; t0 += kRowChunk) {
    double acc[NR][kRowChunk];
#pragma unroll
    for (int w = 0; w < NR; ++w)
#pragma unroll
      for (int t = 0; t < kRowChunk; ++t) acc[w][t] = 0.0;
    const int kc = t0 / kRowChunk, R = apply_chunk_rows(T, kc);
    const ET* Mk = Mc + apply_chunk_base(n, kc);
    if (small)
      chunk_times<ET, KG, NR>(Mk, n, R, lane, rs, acc);
    else
      chunk_times_gather<ET, KG, NR>(Mk, n, R, lane, bi, r, ld, acc);
#pragma unroll
    for (int w = 0; w < NR; ++w) {
      const double s = wave_rows_sum(acc[w], lane);
      if ((lane & 3) == 0 && t0 + trow < T) zc[w * ld + t0 + trow] = s;
    }
  }
}

// One launch over the clusters of `list` (list[0..nlist); all ncell clusters
// in order when list is null), for NR vectors at once (r and z hold them at
// stride ld).  One slot group in flight per pass keeps the single-vector
// kernels at 96 (FP32, five wavefronts per SIMD) or 128 VGPRs (FP64, four):
// the apply streams each core row once, and more resident clusters hide the
// gathers' latency better than more loads in flight per cluster.  Each
// vector's sums run in the same order whatever NR is, so a vector's result
// does not depend on the vectors applied beside it.
template <typename MT, int NR>
__global__ __launch_bounds__(64 * kApplyWaves, NR > 1 ? 2 : (sizeof(MT) == 4 ? 5 : 4)) void k_rasm_apply(
    long long nlist, const int* __restrict__ list, const long long* __restrict__ bptr, const int* __restrict__ bidx,
    const int* __restrict__ ista, const int* __restrict__ iend, const long long* __restrict__ mptr,
    const MT* __restrict__ M, const double* __restrict__ r, long long ld, double* __restrict__ z) {
  __shared__ double rs[kApplyWaves][NR * kMaxGroups * 64];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const long long i = (long long)xcd_block(blockIdx.x, gridDim.x) * kApplyWaves + wave;
  if (i >= nlist) return;
  const long long c = list ? (long long)list[i] : i;
  const long long b0 = bptr[c];
  const int n = (int)(bptr[c + 1] - b0);
  if (n == 0) return;
  const int s = ista[c];
  const int T = iend[c] - s + 1;
  apply_cluster<MT, 1, NR>(M + mptr[c], n, T, lane, bidx + b0, r, ld, rs[wave], z + s);
}

static inline dim3 apply_grid(long long n) { return dim3((unsigned)((n + kApplyWaves - 1) / kApplyWaves)); }

template <typename MT>
static void launch_apply(hipStream_t st, int nrhs, long long nlist, const int* list, const long long* bptr,
                         const int* bidx, const int* ista, const int* iend, const long long* mptr, const MT* M,
                         const double* r, long long ld, double* z) {
  if (nlist <= 0) return;
#define PRBF_AP(NR)                                                                                              \
  hipLaunchKernelGGL((k_rasm_apply<MT, NR>), apply_grid(nlist), dim3(64 * kApplyWaves), 0, st, nlist, list, bptr, \
                     bidx, ista, iend, mptr, M, r, ld, z)
  if (nrhs == 1) PRBF_AP(1);
  else if (nrhs == 2) PRBF_AP(2);
  else PRBF_AP(3);
#undef PRBF_AP
}

void rasm_apply_fp64(hipStream_t st, int nrhs, long long nlist, const int* list, const long long* bptr,
                     const int* bidx, const int* ista, const int* iend, const long long* mptr, const double* M,
                     const double* r, long long ld, double* z) {
  launch_apply<double>(st, nrhs, nlist, list, bptr, bidx, ista, iend, mptr, M, r, ld, z);
}

void rasm_apply_fp32(hipStream_t st, int nrhs, long long nlist, const int* list, const long long* bptr,
                     const int* bidx, const int* ista, const int* iend, const long long* mptr, const float* M,
                     const double* r, long long ld, double* z) {
  launch_apply<float>(st, nrhs, nlist, list, bptr, bidx, ista, iend, mptr, M, r, ld, z);
}

}  // namespace prbf
