// Batched RASM block setup on the matrix cores: one persistent launch whose
// workgroups each factorise one cluster block at a time, 16x16 tiles, MFMA
// for every O(n^3) and O(n^2 T) term.
//
// For cluster c with n buffer particles and T core particles the block is the
// dense untruncated Gaussian over the buffer (2d/matmult.cxx:83-93).  PCASM's
// restrict form (2d/rbf_interpolation.cxx:213-229) only ever reads the core
// rows of A_c^{-1}, so the setup produces M_c = (A_c^{-1})[core, buffer]
// (T x n, stored as apply_offset lays it out) and the apply is a batched GEMV
// over M_c.
//
//   phase 0  order the buffer with the core LAST and padding (the diagonal
//            1/(2 pi sigma^2) alone) up to NP = 16*ceil(n/16); write the lower
//            tiles of A' to a workspace in
//            the MFMA accumulator layout (one coalesced access per register);
//   phase 1  right-looking tiled Cholesky A' = L L^T over panels of NPC tile
//            columns: per column, the 16x16 diagonal tile by one wavefront
//            (in LDS) with L_kk^{-1} beside it and the column as
//            A_ik L_kk^{-T} on MFMA; then one trailing update
//            A_ij -= sum L_ic L_jc^T per panel on MFMA with the panel in LDS
//            (NPC = 2, 3 or 4 divides the passes over the trailing tiles at
//            the cost of LDS; a panel's later columns stay in LDS from their
//            in-panel update to their TRSM);
//   phase 2  with the core last, L^{-1} E_core is zero above the core, so
//            A'^{-1} E_core = L^{-T} (L^{-1} E_core) is a short forward and a
//            full backward tile substitution, both MFMA with the
//            L_II^{-1} tiles kept from phase 1, 16 right-hand sides at a
//            time held in LDS;
//   phase 3  write each 16-column slice of X to M_c in buffer order.
// The left-looking variant (factor_block_ll) replaces phases 0 and 1: each
// panel is brought up to date from the finished L rows left of it (those rows
// of the panel in LDS, the rest streamed once), so L tiles are written once
// instead of being re-read and re-written by every trailing update, and the
// entries of A' are made where they enter instead of being stored.
// T = double is the plain path (exact FP64 block solves); T = float runs the
// same kernel in FP32 on the FP32 MFMA (mixed precision).  Numerically
// singular blocks (clustered points) still give a preconditioner: a pivot at
// or below tau times the Gaussian's diagonal 1/(2 pi sigma^2) is replaced by
// that floor and the block is factorised again with a growing diagonal shift
// (diag_shift) until no pivot reaches the floor; its status records the
// attempt.  tau = 2^-14 in
// FP32 and 2^-40 in FP64 lies above the rounding noise of a block's Schur
// complements (~n eps of the diagonal for n <= ~800), so every multiplier of
// a replaced pivot stays O(1) (|s_ip|^2 <= s_ii s_pp + noise), and below the
// smallest pivots of well-separated points (lattice blocks at h/sigma = 0.8,
// 2D and 3D, and the reference's cylinder input: >= 2.4e-3 of the diagonal),
// which therefore never reach it.
#include <mutex>
#include <vector>

#include "../../include/petrbf.h"
#include "common.h"
#include "kernels.h"

namespace prbf {

template <typename T>
struct MF;
// Both 16x16x4 MFMAs take A as lane l -> A[l%16][l/16] and B as
// lane l -> B[l/16][l%16].  Their accumulators differ: register r of lane l
// holds row 4*(l/16) + r for f32 but row 4*r + l/16 for f64 (column l%16 in
// both).  row() gives that row; off(i, j) is where element (i, j) of a tile
// sits when tiles are stored register-major (register r of lane l at r*64 + l,
// one coalesced access per register).
template <>
struct MF<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int r, int lq) { return 4 * r + lq; }
  __device__ static inline int off(int i, int j) { return (i >> 2) * 64 + (i & 3) * 16 + j; }
};
template <>
struct MF<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static inline v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int r, int lq) { return 4 * lq + r; }
  __device__ static inline int off(int i, int j) { return (i & 3) * 64 + (i >> 2) * 16 + j; }
};

constexpr int kTE = 256;   // elements per 16x16 tile
constexpr int kLS = 17;    // LDS row stride (bank-conflict padding)
constexpr int kLT = 16 * kLS;

// Sets of per-wavefront partial-sum tiles the backward substitution keeps
// (FP32 two, one barrier per step; FP64 one, two barriers per step), and the
// tiles per wavefront in a set (FP32 steps over two tile rows at once).
template <typename T>
__host__ __device__ constexpr int sbk_sets() { return sizeof(T) == 4 ? 2 : 1; }
template <typename T>
__host__ __device__ constexpr int sbk_tiles(int nw, bool pair) { return sizeof(T) == 4 ? 2 * (pair ? 2 : 1) * nw : nw; }
// The FP32 left-looking kernels with 2-column panels run the backward
// substitution two tile rows per step (4x2: 1.3% faster); 4-column panels
// have no registers to spare for it, and at 1 column it would cost the
// fourth workgroup per CU.
template <typename T>
__host__ __device__ constexpr bool sbk_pair(bool ll, int npc) { return sizeof(T) == 4 && ll && npc == 2; }

// Lane-major tiles (the left-looking factorisation): lane l keeps its four
// values X[l%16][row(q, l/16)], q = 0..3, contiguous at 4 l.  One vector
// access per lane is then the A operand of X, the B operand of X^T (both with
// K index row(q, l/16)) and the accumulator of X^T.  lm_off(a, b) is where
// element (a, b) sits.
template <typename Real>
__device__ inline int lm_off(int a, int b) {
  if constexpr (sizeof(Real) == 4)
    return (16 * (b >> 2) + a) * 4 + (b & 3);
  else
    return (16 * (b & 3) + a) * 4 + (b >> 2);
}
template <typename Real, bool LM>
__device__ inline int tile_off(int a, int b) {
  if constexpr (LM)
    return lm_off<Real>(a, b);
  else
    return MF<Real>::off(a, b);
}


// Panel tiles (the current tile columns of L, read by every trailing
// update) and the product -P_i P_j^T of two of them into an accumulator.
// FP32: row stride 16 in 16-byte chunks, chunk c of row r stored at chunk
// c ^ f(r), f(r) = ((r >> 1) + (r >> 3)) & 3.  Lane (lr, lq) reads chunk lq
// of row lr of both tiles as one 16-byte load each and runs the four K steps
// with K index 4 lq + q on both operands; the chunk swizzle keeps these loads
// free of bank conflicts with 32 or 64 banks.  Dropping the padding column
// of the stride-17 layout is what lets three FP32 workgroups share a CU's LDS.
// FP64: stride 17, one 8-byte load per K step.
template <typename Real>
struct Pan;
template <>
struct Pan<float> {
  static constexpr int kT = 256;
  __device__ static inline int chunk(int r) { return ((r >> 1) + (r >> 3)) & 3; }
  __device__ static inline int idx(int r, int c) { return r * 16 + (((c >> 2) ^ chunk(r)) << 2) + (c & 3); }
  __device__ static inline MF<float>::v4 minus_prod(const float* pi, const float* pj, int lr, int lq,
                                                    MF<float>::v4 acc) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int o = lr * 16 + ((lq ^ chunk(lr)) << 2);
    const f4 a = *(const f4*)(pi + o);
    const f4 b = *(const f4*)(pj + o);
#pragma unroll
    for (int q = 0; q < 4; ++q) acc = MF<float>::mma(-a[q], b[q], acc);
    return acc;
  }
};
template <>
struct Pan<double> {
  static constexpr int kT = kLT;
  __device__ static inline int idx(int r, int c) { return r * kLS + c; }
  __device__ static inline MF<double>::v4 minus_prod(const double* pi, const double* pj, int lr, int lq,
                                                     MF<double>::v4 acc) {
#pragma unroll
    for (int q = 0; q < 4; ++q) acc = MF<double>::mma(-pi[lr * kLS + 4 * q + lq], pj[lr * kLS + 4 * q + lq], acc);
    return acc;
  }
};

__device__ inline int ltile(int I, int J) { return I * (I + 1) / 2 + J; }

__device__ inline void tri_decode(int t, int* I, int* J) {
  int i = (int)((sqrtf(8.0f * t + 1.0f) - 1.0f) * 0.5f);
  while (i * (i + 1) / 2 > t) --i;
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  *I = i;
  *J = t - i * (i + 1) / 2;
}

__device__ inline void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// Copies `bytes` (a multiple of 1 KB) from global memory to LDS, 1 KB per
// wavefront instruction (global_load_lds_dwordx4: 16 bytes per lane, lane l
// to dst + 16 l), without passing through registers; `dst` must be the same
// in every lane.  The data is in LDS once this wavefront's vector-memory
// counter is back to zero (lds_dma_wait).
typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;
__device__ __forceinline__ void lds_dma_1k(const void* src, void* dst, int lane) {
  __builtin_amdgcn_global_load_lds((glb_void_t*)((const char*)src + 16 * lane), (lds_void_t*)dst, 16, 0, 0);
}
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima)
__device__ __forceinline__ void lds_dma_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

template <typename T>
__device__ inline T gauss_entry(double dx, double dy, double dz, double inv_two_s2, double norm);
template <>
__device__ inline double gauss_entry<double>(double dx, double dy, double dz, double inv, double norm) {
  return exp_nonpos(-(dx * dx + dy * dy + dz * dz) * inv) * norm;
}
template <>
__device__ inline float gauss_entry<float>(double dx, double dy, double dz, double inv, double norm) {
  // differences in FP64 (coordinates ~1, distances ~sigma), the rest in FP32:
  // the exponent's relative error stays within a few FP32 ulp
  const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
  const float a = (fx * fx + fy * fy + fz * fz) * (float)inv;
  return __expf(-a) * (float)norm;
}

// 1/sqrt: the hardware instruction in FP32 (the FP32 blocks are a
// preconditioner), the library's full-precision one in FP64.
__device__ __forceinline__ float pivot_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double pivot_rsqrt(double x) { return rsqrt(x); }

// Smallest pivot the modified Cholesky keeps: tau times the Gaussian's
// diagonal norm = 1/(2 pi sigma^2), tau = 2^-14 (FP32) or 2^-40 (FP64).
template <typename Real>
__device__ __forceinline__ Real pivot_floor(double norm) {
  return sizeof(Real) == 4 ? (Real)(norm * 0x1p-14) : (Real)(norm * 0x1p-40);
}

// Diagonal shift of factorisation attempt a (0: none): a block whose pivots
// reach the floor is factorised again as A' + shift I (Cholesky with an added
// multiple of the identity, Nocedal & Wright, Numerical Optimization,
// Alg. 3.3), shift = norm 2^-12, 2^-8, 2^-4 in FP32 and norm 2^-36, 2^-28,
// 2^-20 in FP64; after the last attempt the floor stays in place.
constexpr int kShiftLevels = 4;
template <typename Real>
__device__ __forceinline__ Real diag_shift(double norm, int attempt) {
  if (attempt == 0) return Real(0);
  return sizeof(Real) == 4 ? (Real)ldexp(norm, -16 + 4 * attempt) : (Real)ldexp(norm, -44 + 8 * attempt);
}

// acc + (acc in lane P of this 16-lane row) * nm as one v_fmac_f32_dpp: the
// row_newbcast broadcast is the FMAC's DPP operand, with no separate move
// (the compiler emits a v_mov_b32_dpp, the FMA and a nop for each update).
// The DPP read-after-write hazard wants any VALU write of a register this
// instruction reads at least two wait states ahead; NOP puts an s_nop 1 in
// front where the elimination's order does not already keep such writes
// that far away (the step's multiplier, written just before its first
// update, and the right half's newest column, written by compiler code).
template <int P, bool NOP>
__device__ __forceinline__ float fmac_row_bcast(float acc, float nm) {
  if constexpr (NOP)
    asm volatile("s_nop 1\n\tv_fmac_f32_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "+v"(acc)
                 : "v"(nm), "n"(P));
  else
    asm volatile("v_fmac_f32_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "+v"(acc)
                 : "v"(nm), "n"(P));
  return acc;
}
// v in lane P of this 16-lane row, after two wait states (an asm statement,
// so that it stays behind the previous step's updates of v)
template <int P>
__device__ __forceinline__ float row_bcast_nop(float v) {
  float r;
  asm volatile("s_nop 1\n\tv_mov_b32_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
               : "=&v"(r)
               : "v"(v), "n"(P));
  return r;
}

// Step P of the symmetric elimination on [D | I] in FP32: every 16-lane row
// of the wavefront holds the whole 16x16 problem, lane i of the row holding
// row i.  Row P is read through row_newbcast; rows below P subtract
// (a_iP / a_PP) x row P from both halves (rows at or above P multiply by 0),
// and lane P keeps 1/sqrt(a_PP), the scale of its row of L^{-1}.  Column P of
// the right half is (i == P) until this step, so its update is a select.
// (FP32 4x4 setup kernel 1.3% faster than the move + FMA form; the same bits.)
template <typename Real, int P>
struct DiagStep {
  __device__ static __forceinline__ void run(Real (&a)[16], Real (&e)[16], Real& rs, bool& bad, int i, Real pfloor) {
    const Real piv = row_bcast_nop<P>(a[P]);
    const bool b = !(piv > pfloor);  // (NaN included)
    bad |= b;
    const Real rinv = pivot_rsqrt(b ? pfloor : piv);  // (branch-free: one block per step)
    rs = i == P ? rinv : rs;
    const Real nm = i > P ? -(a[P] * (rinv * rinv)) : Real(0);
    if constexpr (P + 1 < 16) a[P + 1] = fmac_row_bcast<P, true>(a[P + 1], nm);
#pragma unroll
    for (int q = P + 2; q < 16; ++q) a[q] = fmac_row_bcast<P, false>(a[q], nm);
#pragma unroll
    for (int q = 0; q + 1 < P; ++q) {
      if (q == 0 && P + 1 == 16) e[q] = fmac_row_bcast<P, true>(e[q], nm);  // (no a update at the last step)
      else e[q] = fmac_row_bcast<P, false>(e[q], nm);
    }
    if constexpr (P > 0) e[P - 1] = fmac_row_bcast<P, true>(e[P - 1], nm);  // (set by the select of step P-1)
    e[P] = (i == P) ? Real(1) : nm;
    DiagStep<Real, P + 1>::run(a, e, rs, bad, i, pfloor);
  }
};
template <typename Real>
struct DiagStep<Real, 16> {
  __device__ static __forceinline__ void run(Real (&)[16], Real (&)[16], Real&, bool&, int, Real) {}
};

// Value of v in lane l (l uniform) via readlane.
__device__ inline double bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// L^{-1} of the 16x16 diagonal tile D (row-major, row stride DS) into Li (row
// stride OS, zero above the diagonal), by one wavefront, with every pivot at
// or below pfloor replaced by pfloor.  Returns true (in every lane) when a
// pivot was replaced.  Li may be D: every lane holds its rows in registers
// before any lane writes.
//
// FP32: the elimination of DiagStep on every 16-lane row at once, without
// scaling rows as it goes (row i of L^{-1} is the right half's row i times
// 1/sqrt of its pivot, applied at the end).
//
// FP64: the same elimination on lanes 0-15 with readlane broadcasts of row p
// and row p scaled at its step: the DPP form needs ~30 more VGPRs in FP64,
// which the FP64 factorisation (at its 256-VGPR cap) would spill.
template <typename Real, int DS, int OS>
__device__ __forceinline__ bool diag_inverse(const Real* D, Real* Li, int lane, Real pfloor) {
  if constexpr (sizeof(Real) == 4) {
    const int i = lane & 15;
    Real a[16], e[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) a[q] = D[i * DS + q];
    Real rs = Real(1);
    bool bad = false;
    DiagStep<Real, 0>::run(a, e, rs, bad, i, pfloor);
    if (lane < 16) {
#pragma unroll
      for (int q = 0; q < 16; ++q) Li[i * OS + q] = q <= i ? e[q] * rs : Real(0);
    }
    return bad;
  } else {
    Real ra[16], re[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      ra[q] = lane < 16 ? D[lane * DS + q] : Real(0);
      re[q] = (q == lane) ? Real(1) : Real(0);
    }
    bool bad = false;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
      const Real piv = bcast(ra[p], p);
      const bool b = !(piv > pfloor);
      bad |= b;
      const Real rinv = pivot_rsqrt(b ? pfloor : piv);
      const Real m = ra[p] * rinv;
      if (lane == p) {
#pragma unroll
        for (int q = p; q < 16; ++q) ra[q] *= rinv;
#pragma unroll
        for (int q = 0; q <= p; ++q) re[q] *= rinv;
      }
#pragma unroll
      for (int q = p + 1; q < 16; ++q) {
        const Real v = bcast(ra[q], p);
        if (lane > p) ra[q] -= m * v;
      }
#pragma unroll
      for (int q = 0; q <= p; ++q) {
        const Real v = bcast(re[q], p);
        if (lane > p) re[q] -= m * v;
      }
    }
    if (lane < 16) {
#pragma unroll
      for (int q = 0; q < 16; ++q) Li[lane * OS + q] = q <= lane ? re[q] : Real(0);
    }
    return bad;
  }
}

// Workspace of one block, in elements of T:
//   lower tiles (nt = KT(KT+1)/2, 256 each) | L_II^{-1} (KT x 256, row-major)
//   | coordinates, core last (3 x 16 KT doubles, left-looking kernel)
__host__ __device__ inline long long blk_ws(int n, int T) {
  int KT = (n + 15) / 16;
  (void)T;
  return (long long)(KT * (KT + 1) / 2 + KT) * kTE + 6LL * 16 * KT;
}

// Panel tiles held in LDS for m columns of a block of KT tile rows: column pc
// of the panel starting at tile column k0 keeps the rows below its diagonal
// tile, k0+pc+1 .. KT-1 (the diagonal tile itself goes to D), and the first
// panel (k0 = 0) is the largest
__host__ __device__ inline int pan_tiles(int KT, int m) { return m * (KT - 1) - m * (m - 1) / 2; }

template <typename T>
__host__ __device__ inline size_t blk_lds_bytes(int nmax, int tmax, int nw, int npc) {
  int KT = (nmax + 15) / 16, NP = 16 * KT;
  (void)tmax;
  const int m = npc < KT ? npc : KT;
  size_t pan = (size_t)pan_tiles(KT, m) * Pan<T>::kT * sizeof(T);
  size_t crd = (size_t)3 * NP * sizeof(double);
  size_t red = ((size_t)NP * kLS + (size_t)sbk_tiles<T>(nw, false) * kTE) * sizeof(T);  // phase-2 right-hand sides, partial sums
  size_t region = pan > crd ? pan : crd;
  region = region > red ? region : red;
  region = (region + 15) & ~(size_t)15;
  return region + (size_t)kLT * sizeof(T) + 16;
}

// Phases 2 and 3 of one block whose factor L (tiles in Wb, L_II^{-1} in Wd)
// is complete: the core rows of A'^{-1} into M.  LM: tiles stored lane-major
// (lm_off) instead of register-major (MF::off).
template <typename Real, int NW, bool LM, bool PAIR>
__device__ __forceinline__ void core_rows_body(long long c, const Real* __restrict__ Wb, const Real* __restrict__ Wd,
                          unsigned char* __restrict__ lds_raw, int n, int T, const long long* __restrict__ mptr,
                          Real* __restrict__ M) {
  typedef typename MF<Real>::v4 v4;
  const int KT = (n + 15) / 16, TC = (T + 15) / 16;
  const int ncore0 = n - T;
  const int tid = threadIdx.x, lane = tid % 64;
  // FP32 takes the wavefront index as a scalar (fewer VGPRs: no spills at the
  // 168 cap, 3% faster); in FP64 the same change costs 10%, so it stays a vector
  const int wave = sizeof(Real) == 4 ? __builtin_amdgcn_readfirstlane(tid / 64) : tid / 64;
  const int lr = lane & 15, lq = lane >> 4;
  // per-wave partial sums of the backward substitution, past X: FP32 keeps
  // two sets used in alternate steps (one tile per wavefront, two in the
  // two-row form), so that a step needs one barrier (a wavefront writing the
  // next step's partials cannot overwrite what is still read of this one);
  // FP64 keeps one set (LDS for two workgroups per CU) and a second barrier
  constexpr int kSB = sbk_sets<Real>();
  static_assert(!PAIR || kSB == 2, "the two-row backward substitution is FP32's");
  Real* S = (Real*)lds_raw + (size_t)(16 * KT) * kLS;
  // ---- phase 2: X = L^{-T} L^{-1} E_core, 16 right-hand sides at a time,
  // X in LDS (row stride 17) in the region the panel used
  Real* X = (Real*)lds_raw;
  const int I0 = ncore0 / 16;
  Real* Mc = M + mptr[c];
  for (int tc = 0; tc < TC; ++tc) {
    // forward: Y_I = L_II^{-1} (E_I - sum_{I0<=J<I} L_IJ Y_J), tile rows I >= I0
    if (wave == 0) {
      __builtin_amdgcn_s_setprio(3);  // serial chain: raised issue priority (as in the factorisation)
      for (int I = I0; I < KT; ++I) {
        v4 acc = {0, 0, 0, 0};
        for (int J = I0; J < I; ++J) {
          const Real* tij = Wb + (long long)ltile(I, J) * kTE;
          if constexpr (LM) {
            const v4 t = *(const v4*)(tij + lane * 4);
#pragma unroll
            for (int q = 0; q < 4; ++q)
              acc = MF<Real>::mma(t[q], X[(16 * J + MF<Real>::row(q, lq)) * kLS + lr], acc);
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
              acc = MF<Real>::mma(tij[MF<Real>::off(lr, 4 * q + lq)], X[(16 * J + 4 * q + lq) * kLS + lr], acc);
          }
        }
        // R = E_I - sum in accumulator layout; register q is the B operand of
        // product q with K index row(q, lq), so no re-layout through LDS
        Real rv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int row = 16 * I + MF<Real>::row(r, lq), col = 16 * tc + lr;
          Real e = (col < T && row == ncore0 + col) ? Real(1) : Real(0);
          rv[r] = e - acc[r];
        }
        const Real* li = Wd + (long long)I * kTE;
        v4 y = {0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q) y = MF<Real>::mma(li[lr * 16 + MF<Real>::row(q, lq)], rv[q], y);
#pragma unroll
        for (int r = 0; r < 4; ++r) X[(16 * I + MF<Real>::row(r, lq)) * kLS + lr] = y[r];
        wave_fence();
      }
      __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();
    if constexpr (!PAIR) {
      // backward: X_I = L_II^{-T} (Y_I - sum_{J>I} L_JI^T X_J); the sum is split
      // over the wavefronts and combined by wavefront 0 through the scratch
      // tiles.  In FP64 wavefront 0, which also combines the sums and applies
      // the diagonal inverse, takes only J = I + 1 (whose X_J it wrote itself at
      // the previous step) and the others share J >= I + 2 (FP64 4x2 setup 3.8%
      // faster); FP32 deals J round-robin over all wavefronts (the split form
      // spills there).  Each wavefront's first kPF tiles of column I-1 (and
      // wavefront 0's L_{I-1,I-1}^{-1}) are loaded while step I runs, so the
      // global loads leave the step's critical path.
      constexpr bool kSplit = sizeof(Real) == 8;
      constexpr int kPF = kSplit ? 6 : 4;
      constexpr int kStride = kSplit ? NW - 1 : NW;
      auto jof = [&](int I, int g) {
        if constexpr (kSplit) return wave == 0 ? (g == 0 ? I + 1 : KT) : I + 1 + wave + g * (NW - 1);
        else return I + 1 + wave + g * NW;
      };
      Real nx[kPF][4] = {}, nli[4] = {};
      auto load_col = [&](int I) {
#pragma unroll
        for (int g = 0; g < kPF; ++g) {
          const int J = jof(I, g);
          if (J < KT) {
            const Real* tji = Wb + (long long)ltile(J, I) * kTE;
#pragma unroll
            for (int q = 0; q < 4; ++q) nx[g][q] = tji[tile_off<Real, LM>(4 * q + lq, lr)];
          }
        }
        if (wave == 0) {  // (L_II^{-1})^T as the A operand, K index row(q, lq)
          const Real* li = Wd + (long long)I * kTE;
#pragma unroll
          for (int q = 0; q < 4; ++q) nli[q] = li[MF<Real>::row(q, lq) * 16 + lr];
        }
      };
      load_col(KT - 1);
      for (int I = KT - 1; I >= 0; --I) {
        Real av[kPF][4], al[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          al[q] = nli[q];
#pragma unroll
          for (int g = 0; g < kPF; ++g) av[g][q] = nx[g][q];
        }
        if (I > 0) load_col(I - 1);
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int g = 0; g < kPF; ++g) {
          const int J = jof(I, g);
          if (J < KT) {
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = MF<Real>::mma(av[g][q], X[(16 * J + 4 * q + lq) * kLS + lr], acc);
          }
        }
        // (blocks beyond the prefetched tiles)
        for (int J = (kSplit && wave == 0) ? KT : I + 1 + wave + kPF * kStride; J < KT; J += kStride) {
          const Real* tji = Wb + (long long)ltile(J, I) * kTE;
          Real a[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) a[q] = tji[tile_off<Real, LM>(4 * q + lq, lr)];
#pragma unroll
          for (int q = 0; q < 4; ++q) acc = MF<Real>::mma(a[q], X[(16 * J + 4 * q + lq) * kLS + lr], acc);
        }
        Real* Ss = S + (kSB == 2 ? (I & 1) * NW : 0) * kTE;  // this step's set
#pragma unroll
        for (int r = 0; r < 4; ++r) Ss[wave * kTE + r * 64 + lane] = acc[r];
        __syncthreads();
        if (wave == 0) {
          Real rv[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            Real sum = Ss[r * 64 + lane];
            for (int w = 1; w < NW; ++w) sum += Ss[w * kTE + r * 64 + lane];
            Real yv = (I >= I0) ? X[(16 * I + MF<Real>::row(r, lq)) * kLS + lr] : Real(0);
            rv[r] = yv - sum;
          }
          v4 x = {0, 0, 0, 0};
#pragma unroll
          for (int q = 0; q < 4; ++q) x = MF<Real>::mma(al[q], rv[q], x);
#pragma unroll
          for (int r = 0; r < 4; ++r) X[(16 * I + MF<Real>::row(r, lq)) * kLS + lr] = x[r];
        }
        // X_I is read next by wavefront 0 itself (column I of step I-1) and by
        // the others only after step I-1's barrier
        if constexpr (kSB == 1) __syncthreads();
      }
      if constexpr (kSB == 2) __syncthreads();
    } else {
      // FP32: two tile rows per step from the top, (H, H-1), and row 0 alone
      // last when KT is odd; one barrier per step, half the dependent steps
      // of one row at a time.  Both columns' sums over J > H are split over
      // the wavefronts (J = H + 1 + wave + g NW; tiles (J, H-1) and (J, H)
      // are adjacent in the workspace and share the X_J operand) and written
      // as partial tiles; after the barrier wavefront 1 forms
      // X_H = L_HH^{-T} (Y_H - S_H), and wavefront 0 forms X_H the same way
      // and then X_{H-1} = L_{H-1,H-1}^{-T} (Y_{H-1} - S_{H-1} - L_{H,H-1}^T X_H).
      // At the next step wavefront 0 reads X_{H-1} first and wavefront 1
      // X_H, each its own; the others read X tiles written before this
      // step's barrier.  Each wavefront's first kPF rows of the next step
      // are loaded during the current step; the diagonal inverses and the
      // coupling tile L_{H,H-1} at the start of their own step.
      constexpr int kPF = 4;
      Real nx[kPF][2][4] = {};
      auto load_step = [&](int H) {
#pragma unroll
        for (int g = 0; g < kPF; ++g) {
          const int J = H + 1 + wave + g * NW;
          if (J < KT) {
            const Real* th = Wb + (long long)ltile(J, H) * kTE;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int o = tile_off<Real, LM>(4 * q + lq, lr);
              nx[g][0][q] = th[o];
              nx[g][1][q] = H > 0 ? th[o - kTE] : Real(0);  // tile (J, H-1)
            }
          }
        }
      };
      load_step(KT - 1);
      int u = 0;
      for (int H = KT - 1; H >= 0; H -= 2, ++u) {
        const bool two = H > 0;
        Real av[kPF][2][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
          for (int g = 0; g < kPF; ++g) {
            av[g][0][q] = nx[g][0][q];
            av[g][1][q] = nx[g][1][q];
          }
        }
        if (H >= 2) load_step(H - 2);
        // this step's diagonal inverses and coupling tile, needed after the
        // products and the barrier
        Real ah[4], al[4], ac[4];
        if (wave <= 1) {
          const Real* li = Wd + (long long)H * kTE;
#pragma unroll
          for (int q = 0; q < 4; ++q) ah[q] = li[MF<Real>::row(q, lq) * 16 + lr];
        }
        if (wave == 0 && two) {
          const Real* li = Wd + (long long)(H - 1) * kTE;
          const Real* tc = Wb + (long long)ltile(H, H - 1) * kTE;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            al[q] = li[MF<Real>::row(q, lq) * 16 + lr];
            ac[q] = tc[tile_off<Real, LM>(MF<Real>::row(q, lq), lr)];
          }
        }
        v4 sh = {0, 0, 0, 0}, sl = {0, 0, 0, 0};
#pragma unroll
        for (int g = 0; g < kPF; ++g) {
          const int J = H + 1 + wave + g * NW;
          if (J < KT) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const Real xb = X[(16 * J + 4 * q + lq) * kLS + lr];
              sh = MF<Real>::mma(av[g][0][q], xb, sh);
              if (two) sl = MF<Real>::mma(av[g][1][q], xb, sl);
            }
          }
        }
        // (rows beyond the prefetched ones)
        for (int J = H + 1 + wave + kPF * NW; J < KT; J += NW) {
          const Real* th = Wb + (long long)ltile(J, H) * kTE;
          Real a[4], b[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int o = tile_off<Real, LM>(4 * q + lq, lr);
            a[q] = th[o];
            b[q] = two ? th[o - kTE] : Real(0);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const Real xb = X[(16 * J + 4 * q + lq) * kLS + lr];
            sh = MF<Real>::mma(a[q], xb, sh);
            if (two) sl = MF<Real>::mma(b[q], xb, sl);
          }
        }
        Real* Ss = S + (size_t)(u & 1) * (2 * NW) * kTE;  // this step's set: (H, H-1) tiles per wavefront
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          Ss[(2 * wave) * kTE + r * 64 + lane] = sh[r];
          if (two) Ss[(2 * wave + 1) * kTE + r * 64 + lane] = sl[r];
        }
        __syncthreads();
        Real yh[4], yl[4];
        if (wave <= 1) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            yh[r] = H >= I0 ? X[(16 * H + MF<Real>::row(r, lq)) * kLS + lr] : Real(0);
            yl[r] = (two && H - 1 >= I0) ? X[(16 * (H - 1) + MF<Real>::row(r, lq)) * kLS + lr] : Real(0);
          }
        }
        // (in the steps over the core's tile rows wavefront 1 overwrites Y_H
        // with X_H only after wavefront 0 has read Y_H)
        if (H >= I0) __syncthreads();
        if (wave <= 1) {
          Real rv[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            Real sum = Ss[r * 64 + lane];
            for (int w = 1; w < NW; ++w) sum += Ss[(2 * w) * kTE + r * 64 + lane];
            rv[r] = yh[r] - sum;
          }
          v4 xh = {0, 0, 0, 0};
#pragma unroll
          for (int q = 0; q < 4; ++q) xh = MF<Real>::mma(ah[q], rv[q], xh);
          if (wave == (two ? 1 : 0)) {
#pragma unroll
            for (int r = 0; r < 4; ++r) X[(16 * H + MF<Real>::row(r, lq)) * kLS + lr] = xh[r];
          }
          if (wave == 0 && two) {
            v4 t = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; ++q) t = MF<Real>::mma(ac[q], xh[q], t);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              Real sum = Ss[kTE + r * 64 + lane];
              for (int w = 1; w < NW; ++w) sum += Ss[(2 * w + 1) * kTE + r * 64 + lane];
              rv[r] = yl[r] - sum - t[r];
            }
            v4 xl = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; ++q) xl = MF<Real>::mma(al[q], rv[q], xl);
#pragma unroll
            for (int r = 0; r < 4; ++r) X[(16 * (H - 1) + MF<Real>::row(r, lq)) * kLS + lr] = xl[r];
          }
        }
      }
      __syncthreads();
    }
    // ---- phase 3: M_c[16 tc + tt][b] = X[pos(b)][tt], rows past T zero, in
    // the apply's storage order (apply_offset), written contiguously
    const int tcn = min(16, T - 16 * tc), R = apply_chunk_rows(T, tc);
    constexpr int V = 16 / (int)sizeof(Real);
    Real* Mk = Mc + apply_chunk_base(n, tc);
    for (int e = tid; e < R * n; e += NW * 64) {
      const int g = e / (R * 64), rr = e % (R * 64);
      const int ng = min(64, n - 64 * g);
      const int q = rr / (ng * V), rem = rr % (ng * V);
      const int b = 64 * g + rem / V, tt = q * V + rem % V;
      const int p = b < T ? ncore0 + b : b - T;
      Mk[e] = tt < tcn ? X[p * kLS + tt] : Real(0);
    }
    __syncthreads();
  }
}

// Phases 2 and 3 as the factorisations call them: inlined in FP32 (a call
// spills there), a call of its own in FP64, where inlining it raises the
// factorisation's spills at the 256-VGPR cap from 17 to 33 (4x2 setup
// 250.8 against 253.0 ms)
template <typename Real, int NW, bool LM, bool PAIR>
__device__ __attribute__((noinline)) void core_rows_call(long long c, const Real* __restrict__ Wb,
                                                         const Real* __restrict__ Wd,
                                                         unsigned char* __restrict__ lds_raw, int n, int T,
                                                         const long long* __restrict__ mptr, Real* __restrict__ M) {
  core_rows_body<Real, NW, LM, PAIR>(c, Wb, Wd, lds_raw, n, T, mptr, M);
}
template <typename Real, int NW, bool LM, bool PAIR>
__device__ __forceinline__ void core_rows(long long c, const Real* __restrict__ Wb, const Real* __restrict__ Wd,
                                          unsigned char* __restrict__ lds_raw, int n, int T,
                                          const long long* __restrict__ mptr, Real* __restrict__ M) {
  if constexpr (sizeof(Real) == 4)
    core_rows_body<Real, NW, LM, PAIR>(c, Wb, Wd, lds_raw, n, T, mptr, M);
  else
    core_rows_call<Real, NW, LM, PAIR>(c, Wb, Wd, lds_raw, n, T, mptr, M);
}

// One block: cluster c, workspace Wb (blk_ws elements), LDS as laid out below.
template <typename Real, int DIM, int NW, int NPC>
__device__ void factor_block(
    long long c, Real* __restrict__ Wb, unsigned char* __restrict__ lds_raw, const long long* __restrict__ bptr,
    const int* __restrict__ bidx, const int* __restrict__ ista, const int* __restrict__ iend,
    const double* __restrict__ xs, const double* __restrict__ ys, const double* __restrict__ zs,
    const long long* __restrict__ mptr, Real* __restrict__ M, double inv_two_s2, double norm,
    int* __restrict__ blk_status) {
  typedef typename MF<Real>::v4 v4;
  const long long b0 = bptr[c];
  const int n = (int)(bptr[c + 1] - b0);
  if (n == 0) return;
  const int T = iend[c] - ista[c] + 1;
  const int KT = (n + 15) / 16, NP = 16 * KT;
  const int nt = KT * (KT + 1) / 2;
  const int ncore0 = n - T;
  const int tid = threadIdx.x, lane = tid % 64;
  // FP32 takes the wavefront index as a scalar (fewer VGPRs: no spills at the
  // 168 cap, 3% faster); in FP64 the same change costs 10%, so it stays a vector
  const int wave = sizeof(Real) == 4 ? __builtin_amdgcn_readfirstlane(tid / 64) : tid / 64;
  const int lr = lane & 15, lq = lane >> 4;

  Real* Wd = Wb + (long long)nt * kTE;       // L_II^{-1}, row-major

  // LDS carve-up: one region holding, by phase, the coordinates (0), the
  // panel (1), and the right-hand sides X with the per-wave scratch tiles
  // (2, 3); then the diagonal tile, which its inverse overwrites in place,
  // and the flag of regularised pivots
  typedef Pan<Real> PT;
  size_t pan = (size_t)pan_tiles(KT, min(NPC, KT)) * PT::kT * sizeof(Real), crd = (size_t)3 * NP * sizeof(double),
         red = ((size_t)NP * kLS + (size_t)sbk_tiles<Real>(NW, false) * kTE) * sizeof(Real);
  size_t region = pan > crd ? pan : crd;
  region = region > red ? region : red;
  region = (region + 15) & ~(size_t)15;
  Real* P = (Real*)lds_raw;                     // panel: pan_tiles tiles (PT layout)
  double* cx = (double*)lds_raw;          // phase 0 coordinates (same region)
  double* cy = cx + NP;
  double* cz = cy + NP;
  Real* D = (Real*)(lds_raw + region);       // diagonal tile L_kk, then L_kk^{-1}
  Real* Li = D;
  int* flag = (int*)(D + kLT);

  // phases 0 and 1 once more, on A' + shift I, while a pivot reaches the floor
  // (the panel overwrites the coordinates, so phase 0 is redone as well)
  for (int attempt = 0;; ++attempt) {
  const Real dshift = diag_shift<Real>(norm, attempt);
  // ---- phase 0: coordinates (core last) and the lower tiles of A'
  for (int p = tid; p < n; p += NW * 64) {
    int b = p < ncore0 ? p + T : p - ncore0;
    int j = bidx[b0 + b];
    cx[p] = xs[j];
    cy[p] = ys[j];
    cz[p] = DIM == 3 ? zs[j] : 0.0;
  }
  if (tid == 0) *flag = 0;
  __syncthreads();
  for (int t = wave; t < nt; t += NW) {
    int I, J;
    tri_decode(t, &I, &J);
    Real* tile = Wb + (long long)t * kTE;
    const int j = 16 * J + lr;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = 16 * I + MF<Real>::row(r, lq);
      Real v;
      if (i < n && j < n) {
        v = gauss_entry<Real>(cx[i] - cx[j], cy[i] - cy[j], DIM == 3 ? cz[i] - cz[j] : 0.0, inv_two_s2, norm);
        if (i == j) v += dshift;
      } else {
        v = (i == j) ? (Real)norm : Real(0);  // padding: decoupled, at the Gaussian's diagonal
      }
      tile[r * 64 + lane] = v;
    }
  }
  __syncthreads();

  // ---- phase 1: tiled right-looking Cholesky, NPC tile columns per panel
  for (int k0 = 0; k0 < KT; k0 += NPC) {
    const int npc = min(NPC, KT - k0);
    // panel slot of tile (i, k0 + pc), i > k0 + pc: the columns' row ranges
    // packed one after another
    auto pan = [&](int pc, int i) { return P + (pc * (KT - k0 - 2) - (pc * (pc - 1) >> 1) + i - k0 - 1) * PT::kT; };
    for (int cc = 0; cc < npc; ++cc) {
      const int kk = k0 + cc;
      if (cc > 0) {
        // bring column kk up to date with the earlier columns of this panel;
        // the result waits for the TRSM in its panel slot, not in the workspace
        for (int i = kk + wave; i < KT; i += NW) {
          const Real* tik = Wb + (long long)ltile(i, kk) * kTE;
          v4 acc;
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[r] = tik[r * 64 + lane];
          for (int pc = 0; pc < cc; ++pc) acc = PT::minus_prod(pan(pc, i), pan(pc, kk), lr, lq, acc);
          if (i == kk) {
#pragma unroll
            for (int r = 0; r < 4; ++r) D[MF<Real>::row(r, lq) * kLS + lr] = acc[r];
          } else {
            Real* pi = pan(cc, i);
#pragma unroll
            for (int r = 0; r < 4; ++r) pi[PT::idx(MF<Real>::row(r, lq), lr)] = acc[r];
          }
        }
        __syncthreads();
      }
      if (wave == 0) {
        if (kk == 0) {  // later diagonal tiles are left in D by the update that finished them
          const Real* tkk = Wb;
#pragma unroll
          for (int r = 0; r < 4; ++r) D[MF<Real>::row(r, lq) * kLS + lr] = tkk[r * 64 + lane];
        }
        wave_fence();
        if (diag_inverse<Real, kLS, kLS>(D, Li, lane, pivot_floor<Real>(norm)) && lane == 0) *flag = 1;
        wave_fence();
        Real* wd = Wd + (long long)kk * kTE;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int e = r * 64 + lane;
          wd[e] = Li[(e >> 4) * kLS + (e & 15)];
        }
      }
      __syncthreads();
      // panel column: L_ik = A_ik L_kk^{-T}; the A operand is read straight
      // from the stored tile (the first column of a panel) or from the panel
      // slot the in-panel update left it in, all of a wavefront's tiles
      // issued before use; L_ik replaces A_ik in the slot
      for (int i0 = kk + 1 + wave; i0 < KT; i0 += NW * 4) {
        Real av[4][4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int i = i0 + g * NW;
          if (i < KT) {
            if (cc == 0) {
              const Real* tik = Wb + (long long)ltile(i, kk) * kTE;
#pragma unroll
              for (int q = 0; q < 4; ++q) av[g][q] = tik[MF<Real>::off(lr, 4 * q + lq)];
            } else {
              const Real* pi = pan(cc, i);
#pragma unroll
              for (int q = 0; q < 4; ++q) av[g][q] = pi[PT::idx(lr, 4 * q + lq)];
            }
          }
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int i = i0 + g * NW;
          if (i < KT) {
            v4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = MF<Real>::mma(av[g][q], Li[lr * kLS + 4 * q + lq], acc);
            Real* tik = Wb + (long long)ltile(i, kk) * kTE;
            Real* pi = pan(cc, i);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              tik[r * 64 + lane] = acc[r];
              pi[PT::idx(MF<Real>::row(r, lq), lr)] = acc[r];
            }
          }
        }
      }
      __syncthreads();
    }
    // trailing update: A_ij -= sum_c L_ic L_jc^T for k0+npc <= j <= i < KT,
    // kG tiles in flight per wavefront (FP32: 4, within its 168-VGPR cap)
    const int j0 = k0 + npc;
    const int m = KT - j0;
    const int ntr = m * (m + 1) / 2;
    constexpr int kG = sizeof(Real) == 4 ? 4 : 8;
    for (int u0 = wave; u0 < ntr; u0 += NW * kG) {
      v4 acc[kG];
      Real* tp[kG];
      int ti[kG], tj[kG];
#pragma unroll
      for (int g = 0; g < kG; ++g) {
        const int u = u0 + g * NW;
        tp[g] = nullptr;
        if (u < ntr) {
          int ii, jj;
          tri_decode(u, &ii, &jj);
          ti[g] = j0 + ii;
          tj[g] = j0 + jj;
          tp[g] = Wb + (long long)ltile(ti[g], tj[g]) * kTE;
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[g][r] = tp[g][r * 64 + lane];
        }
      }
#pragma unroll
      for (int g = 0; g < kG; ++g) {
        if (tp[g]) {
          for (int pc = 0; pc < npc; ++pc) acc[g] = PT::minus_prod(pan(pc, ti[g]), pan(pc, tj[g]), lr, lq, acc[g]);
#pragma unroll
          for (int r = 0; r < 4; ++r) tp[g][r * 64 + lane] = acc[g][r];
          if (ti[g] == j0 && tj[g] == j0) {  // next diagonal tile: hand it over in D
#pragma unroll
            for (int r = 0; r < 4; ++r) D[MF<Real>::row(r, lq) * kLS + lr] = acc[g][r];
          }
        }
      }
    }
    __syncthreads();
  }
  // (the flag is final after the last panel's barrier; `again` is uniform)
  const bool flagged = *flag != 0, again = flagged && attempt + 1 < kShiftLevels;
  __syncthreads();
  if (!again) {
    if (tid == 0 && (attempt > 0 || flagged)) blk_status[c] = attempt + (flagged ? 4 : 0);
    break;
  }
  }
  core_rows<Real, NW, false, false>(c, Wb, Wd, lds_raw, n, T, mptr, M);
}

// ---- left-looking variant ------------------------------------------------
// Tiles of L rows the left-looking factorisation holds in LDS: rows k0 ..
// k0+npc-1 of L left of the panel at tile column k0, the most over its panels.
__host__ __device__ inline int ll_row_tiles(int KT, int npc) {
  int m = 0;
  for (int k0 = 0; k0 < KT; k0 += npc) {
    const int c = KT - k0 < npc ? KT - k0 : npc;
    if (c * k0 > m) m = c * k0;
  }
  return m;
}

// Panels of more than two columns (FP64: more than one) read their L rows
// straight from the workspace (L2-resident: just written, and read by every
// wavefront of the workgroup), one column ahead of their use, so that LDS
// stays within three FP32 or two FP64 workgroups per CU (FP64 4x2 timed 24%
// faster this way than with its panel rows in LDS at one workgroup per CU).
template <typename T>
__host__ __device__ constexpr bool ll_rows_global(int npc) { return npc >= (sizeof(T) == 8 ? 2 : 3); }

template <typename T>
__host__ __device__ inline size_t ll_region(int KT, int nw, int npc) {
  size_t rows = ll_rows_global<T>(npc) ? 0 : (size_t)ll_row_tiles(KT, npc) * kTE * sizeof(T);
  size_t red = ((size_t)16 * KT * kLS + (size_t)sbk_tiles<T>(nw, sbk_pair<T>(true, npc)) * kTE) * sizeof(T);
  size_t region = rows > red ? rows : red;
  return (region + 31) & ~(size_t)31;
}

template <typename T>
__host__ __device__ inline size_t ll_lds_bytes(int nmax, int nw, int npc) {
  const int KT = (nmax + 15) / 16;
  return ll_region<T>(KT, nw, npc) + (size_t)npc * kLT * sizeof(T) + (size_t)(npc * (npc - 1) / 2) * kTE * sizeof(T) +
         (size_t)3 * 16 * npc * sizeof(double) + 16;
}

// One block, left-looking: for each panel of NPC tile columns k0.., the L
// rows k0..k0+npc-1 left of the panel go to LDS; the panel's diagonal block
// A_{k0+r, k0+c} - sum_j L_{k0+r,j} L_{k0+c,j}^T is formed from LDS alone and
// factorised by one wavefront (inverses of the diagonal tiles, L of the
// block below them); then every tile row i below the diagonal block is
// brought up to date, A_{i,k} - sum_{j<k0} L_ij L_kj^T, with L_ij streamed
// once from the workspace and the panel rows from LDS, and finished in
// registers (the panel's earlier columns, then L_kk^{-1}).  Each L tile is
// written once and read once per later panel, against a read and a write of
// every trailing tile per panel in the right-looking order.  A' is never
// stored: its entries are made where they enter, from the block's
// coordinates (staged in the workspace; the panel columns' copy in LDS),
// while a pass's first tile loads are in flight.  All products run
// transposed, on X^T in accumulators, so every workspace and panel tile
// access is one vector access per lane (lane-major tiles, lm_off).
template <typename Real, int DIM, int NW, int NPC>
__device__ void factor_block_ll(
    long long c, Real* __restrict__ Wb, unsigned char* __restrict__ lds_raw, const long long* __restrict__ bptr,
    const int* __restrict__ bidx, const int* __restrict__ ista, const int* __restrict__ iend,
    const double* __restrict__ xs, const double* __restrict__ ys, const double* __restrict__ zs,
    const long long* __restrict__ mptr, Real* __restrict__ M, double inv_two_s2, double norm,
    int* __restrict__ blk_status) {
  typedef typename MF<Real>::v4 v4;
  constexpr bool PG = ll_rows_global<Real>(NPC);  // panel rows from the workspace, not LDS
  constexpr int G = sizeof(Real) == 4 ? 2 : (NPC > 2 ? 2 : 4);  // tile rows per wavefront pass
  const long long b0 = bptr[c];
  const int n = (int)(bptr[c + 1] - b0);
  if (n == 0) return;
  const int T = iend[c] - ista[c] + 1;
  const int KT = (n + 15) / 16, NP = 16 * KT;
  const int nt = KT * (KT + 1) / 2;
  const int ncore0 = n - T;
  const int tid = threadIdx.x, lane = tid % 64;
  // FP32 takes the wavefront index as a scalar (fewer VGPRs: no spills at the
  // 168 cap, 3% faster); in FP64 the same change costs 10%, so it stays a vector
  const int wave = sizeof(Real) == 4 ? __builtin_amdgcn_readfirstlane(tid / 64) : tid / 64;
  const int lr = lane & 15, lq = lane >> 4;
  Real* Wd = Wb + (long long)nt * kTE;  // L_II^{-1}, row-major
  double* cx = (double*)(Wd + (long long)KT * kTE);  // coordinates, core last
  double* cy = cx + NP;
  double* cz = cy + NP;

  // LDS: one region holding, by phase, the L rows of the current panel (1)
  // and the right-hand sides with the per-wave scratch tiles (2, 3); then the
  // panel's diagonal tiles (row-major, each replaced by its inverse), the L
  // tiles of its diagonal block below the diagonal (lane-major), the panel
  // columns' coordinates and the flag of regularised pivots
  const size_t region = ll_region<Real>(KT, NW, NPC);
  Real* Rw = (Real*)lds_raw;
  Real* Dg = (Real*)(lds_raw + region);
  Real* Pd = Dg + NPC * kLT;
  double* Cc = (double*)(Pd + (NPC * (NPC - 1) / 2) * kTE);  // x, y, z of the panel's 16 NPC columns
  int* flag = (int*)(Cc + 3 * 16 * NPC);
  auto pd = [&](int r, int q) { return Pd + ((r * (r - 1)) / 2 + q) * kTE; };  // tile (r, q), r > q
  auto wtile = [&](int I, int J) { return Wb + (long long)ltile(I, J) * kTE; };

  // ---- phase 0: the block's coordinates, core last, into the workspace
  for (int p = tid; p < n; p += NW * 64) {
    int b = p < ncore0 ? p + T : p - ncore0;
    int j = bidx[b0 + b];
    cx[p] = xs[j];
    cy[p] = ys[j];
    cz[p] = DIM == 3 ? zs[j] : 0.0;
  }
  __syncthreads();

  // phase 1 once more, on A' + shift I, while a pivot reaches the floor
  for (int attempt = 0;; ++attempt) {
  const Real dshift = diag_shift<Real>(norm, attempt);
  if (tid == 0) *flag = 0;
  __syncthreads();
  // ---- phase 1: left-looking Cholesky over panels of NPC tile columns
  for (int k0 = 0; k0 < KT; k0 += NPC) {
    const int npc = min(NPC, KT - k0);
    // the panel's L rows left of it, row r at Rw + r k0 tiles (rows are
    // contiguous in the workspace), copied by the LDS DMA, 1 KB per
    // instruction, every wavefront issuing its share before waiting once;
    // and the panel columns' coordinates
    if constexpr (!PG) {
      const int per_row = k0 * (int)(kTE * sizeof(Real) / 1024), total = npc * per_row;  // 1 KB pieces
      for (int b = wave; b < total; b += NW) {
        const int r = b / per_row, o = b - r * per_row;
        lds_dma_1k((const char*)wtile(k0 + r, 0) + 1024 * o, (char*)Rw + (size_t)1024 * b, lane);
      }
      lds_dma_wait();
    }
    for (int p = tid; p < 16 * npc; p += NW * 64) {
      const int gp = 16 * k0 + p;
      const bool in = gp < n;
      Cc[p] = in ? cx[gp] : 0.0;
      Cc[16 * NPC + p] = in ? cy[gp] : 0.0;
      Cc[32 * NPC + p] = (in && DIM == 3) ? cz[gp] : 0.0;
    }
    __syncthreads();
    // this lane's entries of tile (I, k0 + q2) of A' as the accumulator of
    // its transpose, A'(16 I + lr, 16 (k0 + q2) + row(q, lq)), from the row's
    // coordinates (xi, yi, zi) and the panel columns'; the padding decoupled
    auto aprime = [&](int I, int q2, double xi, double yi, double zi) {
      const int i = 16 * I + lr;
      v4 v;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int jj = 16 * q2 + MF<Real>::row(q, lq), j = 16 * k0 + jj;
        if (i < n && j < n) {
          v[q] = gauss_entry<Real>(xi - Cc[jj], yi - Cc[16 * NPC + jj], DIM == 3 ? zi - Cc[32 * NPC + jj] : 0.0,
                                   inv_two_s2, norm);
          if (i == j) v[q] += dshift;
        } else {
          v[q] = (i == j) ? (Real)norm : Real(0);  // padding: decoupled, at the Gaussian's diagonal
        }
      }
      return v;
    };
    // diagonal block: acc = (A_{k0+r,k0+q} - sum_j L_{k0+r,j} L_{k0+q,j}^T)^T.
    // Panels of up to 2 columns (at most 3 tiles): wavefront 0 sums them all
    // itself before factorising, while the others already stream their tile
    // rows (no barrier between); wider panels: one tile per wavefront.
    constexpr bool kW0D = NPC <= 2;
    auto diag_w0 = [&]() {
      const bool two = npc > 1;
      v4 e00 = aprime(k0, 0, Cc[lr], Cc[16 * NPC + lr], Cc[32 * NPC + lr]), e10 = {0, 0, 0, 0}, e11 = {0, 0, 0, 0};
      if (two) {
        e10 = aprime(k0 + 1, 0, Cc[16 + lr], Cc[16 * NPC + 16 + lr], Cc[32 * NPC + 16 + lr]);
        e11 = aprime(k0 + 1, 1, Cc[16 + lr], Cc[16 * NPC + 16 + lr], Cc[32 * NPC + 16 + lr]);
      }
      v4 o00 = {0, 0, 0, 0}, o10 = {0, 0, 0, 0}, o11 = {0, 0, 0, 0};
      const Real* r0 = (PG ? wtile(k0, 0) : Rw) + lane * 4;
      const Real* r1 = (PG ? wtile(k0 + 1, 0) : Rw + (size_t)k0 * kTE) + lane * 4;
      // two columns per step into even and odd accumulators, as the
      // one-tile-per-wavefront form sums them (the same bits)
      int j = 0;
      for (; j + 1 < k0; j += 2) {
        const v4 x0 = *(const v4*)(r0 + j * kTE), y0 = *(const v4*)(r0 + (j + 1) * kTE);
        v4 x1 = {0, 0, 0, 0}, y1 = {0, 0, 0, 0};
        if (two) {
          x1 = *(const v4*)(r1 + j * kTE);
          y1 = *(const v4*)(r1 + (j + 1) * kTE);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          e00 = MF<Real>::mma(-x0[q], x0[q], e00);
          o00 = MF<Real>::mma(-y0[q], y0[q], o00);
        }
        if (two) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            e10 = MF<Real>::mma(-x0[q], x1[q], e10);
            o10 = MF<Real>::mma(-y0[q], y1[q], o10);
            e11 = MF<Real>::mma(-x1[q], x1[q], e11);
            o11 = MF<Real>::mma(-y1[q], y1[q], o11);
          }
        }
      }
      if (j < k0) {
        const v4 x0 = *(const v4*)(r0 + j * kTE);
#pragma unroll
        for (int q = 0; q < 4; ++q) e00 = MF<Real>::mma(-x0[q], x0[q], e00);
        if (two) {
          const v4 x1 = *(const v4*)(r1 + j * kTE);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            e10 = MF<Real>::mma(-x0[q], x1[q], e10);
            e11 = MF<Real>::mma(-x1[q], x1[q], e11);
          }
        }
      }
      e00 += o00;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) Dg[MF<Real>::row(rr, lq) * kLS + lr] = e00[rr];
      if (two) {
        e10 += o10;
        e11 += o11;
        *(v4*)(pd(1, 0) + lane * 4) = e10;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) Dg[kLT + MF<Real>::row(rr, lq) * kLS + lr] = e11[rr];
      }
    };
    for (int u = wave; !kW0D && u < npc * (npc + 1) / 2; u += NW) {
      int r, q2;
      tri_decode(u, &r, &q2);
      v4 acc = aprime(k0 + r, q2, Cc[16 * r + lr], Cc[16 * NPC + 16 * r + lr], Cc[32 * NPC + 16 * r + lr]);
      const Real* ra = (PG ? wtile(k0 + q2, 0) : Rw + (size_t)q2 * k0 * kTE) + lane * 4;
      const Real* rb = (PG ? wtile(k0 + r, 0) : Rw + (size_t)r * k0 * kTE) + lane * 4;
      // two columns per step into two accumulators: both columns' loads are
      // in flight together and the two product chains are independent
      v4 acc2 = {0, 0, 0, 0};
      int j = 0;
      for (; j + 1 < k0; j += 2) {
        const v4 a = *(const v4*)(ra + j * kTE), a2 = *(const v4*)(ra + (j + 1) * kTE);
        const v4 b = *(const v4*)(rb + j * kTE), b2 = *(const v4*)(rb + (j + 1) * kTE);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          acc = MF<Real>::mma(-a[q], b[q], acc);
          acc2 = MF<Real>::mma(-a2[q], b2[q], acc2);
        }
      }
      if (j < k0) {
        const v4 a = *(const v4*)(ra + j * kTE), b = *(const v4*)(rb + j * kTE);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = MF<Real>::mma(-a[q], b[q], acc);
      }
      acc += acc2;
      if (r == q2) {
        Real* D = Dg + r * kLT;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) D[MF<Real>::row(rr, lq) * kLS + lr] = acc[rr];
      } else {
        *(v4*)(pd(r, q2) + lane * 4) = acc;
      }
    }
    // tile rows below the diagonal block, G per wavefront pass (rows ib,
    // ib + S, ..., ib + (G-1) S): the update from the L rows left of the
    // panel (L_ij streamed from the workspace, the next column's tiles loaded
    // during the current column's products), then per panel column the update
    // from the panel's earlier columns and the TRSM, all in registers.
    v4 acc[G][NPC];
    auto stream = [&](int ib, int S) {
      v4 b[G];
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (k0 > 0 && ib + g * S < KT) b[g] = *(const v4*)(wtile(ib + g * S, 0) + lane * 4);
      // A' of this pass's tiles, made while the first tiles load
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int i = ib + g * S, ii = 16 * i + lr;
        if (i < KT) {
          const bool in = ii < n;
          const double xi = in ? cx[ii] : 0.0, yi = in ? cy[ii] : 0.0, zi = (in && DIM == 3) ? cz[ii] : 0.0;
#pragma unroll
          for (int q2 = 0; q2 < NPC; ++q2)
            if (q2 < npc) acc[g][q2] = aprime(i, q2, xi, yi, zi);
        }
      }
      // panel tiles from the workspace are loaded one column ahead too
      v4 pa[PG ? NPC : 1];
      if constexpr (PG) {
#pragma unroll
        for (int q2 = 0; q2 < NPC; ++q2)
          if (k0 > 0 && q2 < npc) pa[q2] = *(const v4*)(wtile(k0 + q2, 0) + lane * 4);
      }
      for (int j = 0; j < k0; ++j) {
        v4 bn[G];
#pragma unroll
        for (int g = 0; g < G; ++g)
          if (j + 1 < k0 && ib + g * S < KT) bn[g] = *(const v4*)(wtile(ib + g * S, j + 1) + lane * 4);
        v4 pn[PG ? NPC : 1];
        if constexpr (PG) {
#pragma unroll
          for (int q2 = 0; q2 < NPC; ++q2)
            if (j + 1 < k0 && q2 < npc) pn[q2] = *(const v4*)(wtile(k0 + q2, j + 1) + lane * 4);
        }
#pragma unroll
        for (int q2 = 0; q2 < NPC; ++q2) {
          if (q2 < npc) {
            v4 a;
            if constexpr (PG) a = pa[q2];
            else a = *(const v4*)(Rw + ((size_t)q2 * k0 + j) * kTE + lane * 4);
#pragma unroll
            for (int g = 0; g < G; ++g)
              if (ib + g * S < KT) {
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[g][q2] = MF<Real>::mma(-a[q], b[g][q], acc[g][q2]);
              }
          }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) b[g] = bn[g];
        if constexpr (PG) {
#pragma unroll
          for (int q2 = 0; q2 < NPC; ++q2) pa[q2] = pn[q2];
        }
      }
    };
    // the panel column's update from the panel's earlier columns and the
    // TRSM with the diagonal inverses, for the pass streamed at ib
    auto finish = [&](int ib, int S) {
#pragma unroll
      for (int q2 = 0; q2 < NPC; ++q2) {
        if (q2 < npc) {
#pragma unroll
          for (int p2 = 0; p2 < q2; ++p2) {
            const v4 a = *(const v4*)(pd(q2, p2) + lane * 4);
#pragma unroll
            for (int g = 0; g < G; ++g)
              if (ib + g * S < KT) {
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[g][q2] = MF<Real>::mma(-a[q], acc[g][p2][q], acc[g][q2]);
              }
          }
          const Real* D = Dg + q2 * kLT;
          Real li[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) li[q] = D[lr * kLS + MF<Real>::row(q, lq)];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            const int i = ib + g * S;
            if (i < KT) {
              v4 v = {0, 0, 0, 0};
#pragma unroll
              for (int q = 0; q < 4; ++q) v = MF<Real>::mma(li[q], acc[g][q2][q], v);
              acc[g][q2] = v;
              *(v4*)(wtile(i, k0 + q2) + lane * 4) = v;
            }
          }
        }
      }
    };
    // the diagonal block's sums are in LDS; wavefront 0 factorises it while
    // the others stream the first (NW - 1) G tile rows below it (the streamed
    // updates do not need the diagonal inverses), then every wavefront
    // finishes its rows and takes the remaining ones NW G at a time
    if constexpr (!kW0D) __syncthreads();
    constexpr int SO = NW - 1;
    const int ibo = k0 + npc + wave - 1;
    if (wave > 0 && ibo < KT) stream(ibo, SO);
    // factorise the diagonal block (one wavefront): per column q2 the inverse
    // of its diagonal tile, L below it (L^T = L_kk^{-1} A^T), and the update
    // of the block's later columns
    if (wave == 0) {
      // the block's serial chain: raised issue priority against the other
      // workgroups' wavefronts on this SIMD while it runs
      __builtin_amdgcn_s_setprio(3);
      if constexpr (kW0D) {
        diag_w0();
        wave_fence();
      }
      for (int q2 = 0; q2 < npc; ++q2) {
        Real* D = Dg + q2 * kLT;
        wave_fence();
        if (diag_inverse<Real, kLS, kLS>(D, D, lane, pivot_floor<Real>(norm)) && lane == 0) *flag = 1;
        wave_fence();
        Real* wd = Wd + (long long)(k0 + q2) * kTE;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int e = r * 64 + lane;
          wd[e] = D[(e >> 4) * kLS + (e & 15)];
        }
        Real li[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) li[q] = D[lr * kLS + MF<Real>::row(q, lq)];
        for (int r = q2 + 1; r < npc; ++r) {
          const v4 u = *(const v4*)(pd(r, q2) + lane * 4);
          v4 v = {0, 0, 0, 0};
#pragma unroll
          for (int q = 0; q < 4; ++q) v = MF<Real>::mma(li[q], u[q], v);
          *(v4*)(pd(r, q2) + lane * 4) = v;
          *(v4*)(wtile(k0 + r, k0 + q2) + lane * 4) = v;
        }
        wave_fence();
        for (int r = q2 + 1; r < npc; ++r) {
          const v4 b = *(const v4*)(pd(r, q2) + lane * 4);
          for (int r2 = q2 + 1; r2 <= r; ++r2) {
            const v4 a = *(const v4*)(pd(r2, q2) + lane * 4);
            if (r2 == r) {
              Real* Dr = Dg + r * kLT;
              v4 acc;
#pragma unroll
              for (int rr = 0; rr < 4; ++rr) acc[rr] = Dr[MF<Real>::row(rr, lq) * kLS + lr];
#pragma unroll
              for (int q = 0; q < 4; ++q) acc = MF<Real>::mma(-a[q], b[q], acc);
#pragma unroll
              for (int rr = 0; rr < 4; ++rr) Dr[MF<Real>::row(rr, lq) * kLS + lr] = acc[rr];
            } else {
              v4 acc = *(const v4*)(pd(r, r2) + lane * 4);
#pragma unroll
              for (int q = 0; q < 4; ++q) acc = MF<Real>::mma(-a[q], b[q], acc);
              *(v4*)(pd(r, r2) + lane * 4) = acc;
            }
          }
        }
        wave_fence();
      }
      __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();
    if (wave > 0 && ibo < KT) finish(ibo, SO);
    for (int ib = k0 + npc + SO * G + wave; ib < KT; ib += NW * G) {
      stream(ib, NW);
      finish(ib, NW);
    }
    __syncthreads();
  }
  // (the flag is final after the last panel's barrier; `again` is uniform)
  const bool flagged = *flag != 0, again = flagged && attempt + 1 < kShiftLevels;
  __syncthreads();
  if (!again) {
    if (tid == 0 && (attempt > 0 || flagged)) blk_status[c] = attempt + (flagged ? 4 : 0);
    break;
  }
  }
  core_rows<Real, NW, true, sbk_pair<Real>(true, NPC)>(c, Wb, Wd, lds_raw, n, T, mptr, M);
}

// Persistent batched launch: each workgroup owns one workspace slot and takes
// the next block from a global counter until the work list is exhausted (the
// list is ordered largest block first, so uneven blocks balance over the
// workgroups).  Every wavefront leaves through the same counter test.
// Register caps the compiler meets without spilling: FP32 at most 168 VGPRs
// (three wavefronts per SIMD), FP64 at most 256 (two), so that LDS, not
// registers, bounds the workgroups per CU.
template <typename Real, int DIM, int NW, int NPC, bool LL>
__global__ __launch_bounds__(NW * 64, sizeof(Real) == 4 ? 3 : 2) void k_block_chol(
    long long nwork, const int* __restrict__ order, int* __restrict__ next, long long ws_slot,
    const long long* __restrict__ bptr, const int* __restrict__ bidx, const int* __restrict__ ista,
    const int* __restrict__ iend, const double* __restrict__ xs, const double* __restrict__ ys,
    const double* __restrict__ zs, Real* __restrict__ Wbase, const long long* __restrict__ mptr, Real* __restrict__ M,
    double inv_two_s2, double norm, size_t item_slot, int* __restrict__ blk_status) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int* s_item = (int*)(lds_raw + item_slot);
  Real* Wb = Wbase + (long long)blockIdx.x * ws_slot;
  for (;;) {
    if (threadIdx.x == 0) *s_item = atomicAdd(next, 1);
    __syncthreads();
    const long long item = *s_item;
    __syncthreads();
    if (item >= nwork) break;
    if constexpr (LL)
      factor_block_ll<Real, DIM, NW, NPC>(order[item], Wb, lds_raw, bptr, bidx, ista, iend, xs, ys, zs, mptr, M,
                                          inv_two_s2, norm, blk_status);
    else
      factor_block<Real, DIM, NW, NPC>(order[item], Wb, lds_raw, bptr, bidx, ista, iend, xs, ys, zs, mptr, M,
                                       inv_two_s2, norm, blk_status);
  }
}

// Instantiated launch configurations (wavefronts per workgroup, tile columns
// per panel, left-looking): FP32 right-looking also has 6-wavefront
// workgroups (two per CU fill its three wavefronts per SIMD) and 3-column
// panels (two per CU in LDS); the left-looking kernels have 4 or 8
// wavefronts with 1 or 2 columns.
template <typename T, int D>
static const void* chol_fn_dim(int nw, int npc, bool ll) {
#define PRBF_FN(NWV, NPCV, LLV) \
  if (nw == NWV && npc == NPCV && ll == LLV) return (const void*)k_block_chol<T, D, NWV, NPCV, LLV>;
  PRBF_FN(4, 1, false) PRBF_FN(4, 2, false) PRBF_FN(4, 4, false) PRBF_FN(8, 1, false) PRBF_FN(8, 2, false)
  PRBF_FN(8, 4, false)
  PRBF_FN(4, 1, true) PRBF_FN(4, 2, true) PRBF_FN(4, 4, true) PRBF_FN(8, 1, true) PRBF_FN(8, 2, true)
  PRBF_FN(8, 4, true)
  if constexpr (sizeof(T) == 4) {
    PRBF_FN(4, 3, false) PRBF_FN(6, 1, false) PRBF_FN(6, 2, false) PRBF_FN(6, 3, false) PRBF_FN(6, 4, false)
    PRBF_FN(8, 3, false)
  }
#undef PRBF_FN
  return nullptr;
}

template <typename T>
static const void* chol_fn(int dim, int nw, int npc, bool ll) {
  return dim == 3 ? chol_fn_dim<T, 3>(nw, npc, ll) : chol_fn_dim<T, 2>(nw, npc, ll);
}

template <typename T>
static size_t chol_lds_bytes(int nmax, int tmax, int nw, int npc, bool ll) {
  return ll ? ll_lds_bytes<T>(nmax, nw, npc) : blk_lds_bytes<T>(nmax, tmax, nw, npc);
}

// Raises the dynamic-LDS limit of a kernel to the CU's 160 KB, once per
// (device, kernel); the record is process-wide, hence the lock.
static void set_lds_attr(const void* fn) {
  static std::mutex m;
  static std::vector<std::pair<int, const void*>> done;
  int dev = 0;
  PRBF_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(m);
  for (const auto& d : done)
    if (d.first == dev && d.second == fn) return;
  PRBF_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  done.push_back({dev, fn});
}

template <typename T>
static int resident_workgroups(int dim, int nw, int npc, bool ll, int nmax, int tmax) {
  const void* fn = chol_fn<T>(dim, nw, npc, ll);
  if (!fn) return 0;
  set_lds_attr(fn);
  size_t lds = chol_lds_bytes<T>(nmax, tmax, nw, npc, ll);
  int per_cu = 0, dev = 0, ncu = 0;
  PRBF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nw * 64, lds));
  PRBF_HIP(hipGetDevice(&dev));
  PRBF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  return (per_cu < 1 ? 1 : per_cu) * ncu;
}

template <typename T>
static void launch_block_chol(hipStream_t st, int dim, int nw, int npc, bool ll, long long nwork, const int* order,
                              int* next,
                              int nwg, long long ws_slot, int nmax, int tmax, const long long* bptr, const int* bidx,
                              const int* ista, const int* iend, const double* xs, const double* ys, const double* zs,
                              T* W, const long long* mptr, T* M, double sigma, int* blk_status) {
  const void* fn = chol_fn<T>(dim, nw, npc, ll);
  if (!fn) throw ApiError(PRBF_ERR_ARG, "no block factorisation kernel for this launch configuration");
  Kern k = make_kern(sigma);
  set_lds_attr(fn);
  size_t lds = chol_lds_bytes<T>(nmax, tmax, nw, npc, ll);
  size_t item_slot = lds - 8;  // past the flag word of the largest block
  PRBF_HIP(hipMemsetAsync(next, 0, sizeof(int), st));
  double inv_two_s2 = k.inv_two_s2, norm = k.norm;
  void* args[] = {&nwork, &order, &next, &ws_slot, &bptr, &bidx,       &ista, &iend, &xs,  &ys,
                  &zs,    &W,     &mptr, &M,       &inv_two_s2, &norm, &item_slot, &blk_status};
  PRBF_HIP(hipLaunchKernel(fn, dim3((unsigned)nwg), dim3(nw * 64), args, lds, st));
}

// list[0..*count) = the clusters of order[0..nwork) whose block reported a
// breakdown (in no particular order; *count must be zero on entry)
__global__ void k_failed_blocks(long long nwork, const int* __restrict__ order, const int* __restrict__ status,
                                int* __restrict__ list, int* __restrict__ count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwork) return;
  const int c = order[i];
  if (status[c]) list[atomicAdd(count, 1)] = c;
}

void failed_blocks(hipStream_t st, long long nwork, const int* order, const int* status, int* list, int* count) {
  if (nwork <= 0) return;
  hipLaunchKernelGGL(k_failed_blocks, dim3((unsigned)((nwork + 255) / 256)), dim3(256), 0, st, nwork, order, status,
                     list, count);
}

// Dense untruncated block of one cluster, A[i][j] = G(x_bi - x_bj) over its
// buffer in list order (2d/matmult.cxx:83-93), with the entry formula the FP64
// factorisation uses; row-major, one entry per thread.
template <int DIM>
__global__ void k_dense_block(long long b0, int n, const int* __restrict__ bidx, const double* __restrict__ xs,
                              const double* __restrict__ ys, const double* __restrict__ zs, double inv_two_s2,
                              double norm, double* __restrict__ A) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  const int pi = bidx[b0 + i], pj = bidx[b0 + j];
  A[e] = gauss_entry<double>(xs[pi] - xs[pj], ys[pi] - ys[pj], DIM == 3 ? zs[pi] - zs[pj] : 0.0, inv_two_s2, norm);
}

void dense_block(hipStream_t st, int dim, long long b0, int n, const int* bidx, const double* xs, const double* ys,
                 const double* zs, double sigma, double* A) {
  if (n <= 0) return;
  const Kern k = make_kern(sigma);
  const dim3 grid((unsigned)(((long long)n * n + 255) / 256));
  if (dim == 3)
    hipLaunchKernelGGL(k_dense_block<3>, grid, dim3(256), 0, st, b0, n, bidx, xs, ys, zs, k.inv_two_s2, k.norm, A);
  else
    hipLaunchKernelGGL(k_dense_block<2>, grid, dim3(256), 0, st, b0, n, bidx, xs, ys, zs, k.inv_two_s2, k.norm, A);
}

long long block_workspace(int n, int T) { return blk_ws(n, T); }

bool block_chol_has_config(bool fp32, int nw, int npc, bool ll) {
  return (fp32 ? chol_fn<float>(2, nw, npc, ll) : chol_fn<double>(2, nw, npc, ll)) != nullptr;
}

size_t block_lds_bytes(bool fp32, int nmax, int tmax, int nw, int npc, bool ll) {
  return fp32 ? chol_lds_bytes<float>(nmax, tmax, nw, npc, ll) : chol_lds_bytes<double>(nmax, tmax, nw, npc, ll);
}

int block_chol_workgroups(bool fp32, int dim, int nw, int npc, bool ll, int nmax, int tmax) {
  return fp32 ? resident_workgroups<float>(dim, nw, npc, ll, nmax, tmax)
              : resident_workgroups<double>(dim, nw, npc, ll, nmax, tmax);
}

void block_chol_fp64(hipStream_t st, int dim, int nw, int npc, bool ll, long long nwork, const int* order, int* next,
                     int nwg, long long ws_slot, int nmax, int tmax, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const double* xs, const double* ys, const double* zs, double* W,
                     const long long* mptr, double* M, double sigma, int* blk_status) {
  launch_block_chol<double>(st, dim, nw, npc, ll, nwork, order, next, nwg, ws_slot, nmax, tmax, bptr, bidx, ista, iend,
                            xs, ys, zs, W, mptr, M, sigma, blk_status);
}

void block_chol_fp32(hipStream_t st, int dim, int nw, int npc, bool ll, long long nwork, const int* order, int* next,
                     int nwg, long long ws_slot, int nmax, int tmax, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const double* xs, const double* ys, const double* zs, float* W,
                     const long long* mptr, float* M, double sigma, int* blk_status) {
  launch_block_chol<float>(st, dim, nw, npc, ll, nwork, order, next, nwg, ws_slot, nmax, tmax, bptr, bidx, ista, iend,
                           xs, ys, zs, W, mptr, M, sigma, blk_status);
}

}  // namespace prbf
