// Learned preconditioner: a small MLP maps a cluster's residual patch to a
// correction on the cluster's core particles.
//
//   patch   2D: the 8x8 lattice slots around a cluster of 4x4 lattice
//           spacings (its own 16 particles plus a ring of two spacings);
//           3D: the 5x5x5 slots around a cluster of 3x3x3 spacings (its 27
//           particles plus a ring of one spacing).  Slot coordinate
//           floor((x - xm)/hl + ps/2) per axis with hl = box/cs (the lattice
//           spacing) and xm the mean of the cluster's particles; a cluster is
//           eligible when every slot holds exactly one particle and its core
//           particles fill the inner slots;
//   input   the np residuals divided by their RMS s, then the np offsets of
//           the particles from their slot centres in units of sigma (x, y,
//           then z), zero-padded to a multiple of 16 (2D 192, 3D 512);
//   network nin -> 256 -> 256 -> 256 -> nout, GELU (tanh form) after each
//           hidden layer; BF16 weights and activations on the BF16 MFMA
//           (v_mfma_f32_16x16x16_bf16) with FP32 accumulation, FP32 master
//           weights trained with Adam;
//   output  the nc core corrections (2D 16, 3D 27; 3D pads to 32 outputs
//           trained toward zero), multiplied back by s and a fixed output
//           scale.
//
// The GEMM kernel computes C = A B for A(m, k) = A[m*sam + k*sak] and
// B(k, n) = B[k*sbk + n*sbn] in BF16 with FP32 accumulation: 64x64 output
// tiles per workgroup, 32x32 per wavefront (2x2 MFMA tiles), operand
// fragments read straight from memory (the operands are small and stay in
// L2).  Epilogues fuse the bias, GELU, GELU' and BF16 conversions.
#include "common.h"
#include "kernels.h"

namespace prbf {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float x) {  // round to nearest even
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// GELU, tanh form, and its derivative
__device__ inline float gelu(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  const float t = tanhf(k0 * (x + k1 * x * x * x));
  return 0.5f * x * (1.0f + t);
}
__device__ inline float gelu_d(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  const float t = tanhf(k0 * (x + k1 * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k0 * (1.0f + 3.0f * k1 * x * x);
}

enum { EPI_F32 = 0, EPI_GELU = 1, EPI_DGELU = 2 };

// EPI_F32:   Cf = acc (+ bias)
// EPI_GELU:  Cf = acc + bias (pre-activation), Cb = bf16(gelu(Cf))
// EPI_DGELU: v = acc * gelu'(aux), Cf = v, Cb = bf16(v)
template <int EPI, bool ACONT, bool BCONT>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const unsigned short* __restrict__ A,
                                                   long long sam, long long sak, const unsigned short* __restrict__ B,
                                                   long long sbk, long long sbn, const float* __restrict__ bias,
                                                   const float* __restrict__ aux, float* __restrict__ Cf,
                                                   unsigned short* __restrict__ Cb, long long ldc) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int lr = lane & 15, lq = lane >> 4;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32, n0 = blockIdx.x * 64 + (wave & 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
  // fragment rows / columns of this lane (clamped into range; out-of-range
  // rows and columns are computed on clamped data and never stored)
  int mr[2], nc[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    mr[t] = min(m0 + 16 * t + lr, M - 1);
    nc[t] = min(n0 + 16 * t + lr, N - 1);
  }
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int kb = k0 + 4 * lq;
    bf4 af[2], bfr[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const unsigned short* pa = A + (long long)mr[t] * sam + (long long)kb * sak;
      if (ACONT) {
        const uint2 v = *reinterpret_cast<const uint2*>(pa);
        af[t] = bf4{(short)(v.x & 0xffff), (short)(v.x >> 16), (short)(v.y & 0xffff), (short)(v.y >> 16)};
      } else {
        af[t] = bf4{(short)pa[0], (short)pa[sak], (short)pa[2 * sak], (short)pa[3 * sak]};
      }
      const unsigned short* pb = B + (long long)kb * sbk + (long long)nc[t] * sbn;
      if (BCONT) {
        const uint2 v = *reinterpret_cast<const uint2*>(pb);
        bfr[t] = bf4{(short)(v.x & 0xffff), (short)(v.x >> 16), (short)(v.y & 0xffff), (short)(v.y >> 16)};
      } else {
        bfr[t] = bf4{(short)pb[0], (short)pb[sbk], (short)pb[2 * sbk], (short)pb[3 * sbk]};
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = n0 + 16 * b + lr;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * a + 4 * lq + r;
        if (m >= M) continue;
        const long long o = (long long)m * ldc + n;
        float v = acc[a][b][r];
        if (EPI == EPI_F32) {
          if (bias) v += bias[n];
          Cf[o] = v;
        } else if (EPI == EPI_GELU) {
          v += bias[n];
          Cf[o] = v;
          Cb[o] = f2bf(gelu(v));
        } else {
          v *= gelu_d(aux[o]);
          Cf[o] = v;
          Cb[o] = f2bf(v);
        }
      }
    }
}

template <int EPI>
static void gemm_launch(hipStream_t st, int M, int N, int K, const unsigned short* A, long long sam, long long sak,
                        const unsigned short* B, long long sbk, long long sbn, const float* bias, const float* aux,
                        float* Cf, unsigned short* Cb, long long ldc) {
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  const bool ac = (sak == 1), bc = (sbk == 1);
#define PRBF_G(AC, BC) \
  hipLaunchKernelGGL((k_gemm_bf16<EPI, AC, BC>), grid, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, bias, aux, Cf, Cb, ldc)
  if (ac && bc) PRBF_G(true, true);
  else if (ac) PRBF_G(true, false);
  else if (bc) PRBF_G(false, true);
  else PRBF_G(false, false);
#undef PRBF_G
}

// ---------------------------------------------------------------- element-wise
__global__ void k_to_bf16(long long n, const float* __restrict__ a, unsigned short* __restrict__ b) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

// W (K x N, row-major FP32) -> Wb (K x N) and Wt (N x K), both BF16
__global__ void k_weights_bf16(int K, int N, const float* __restrict__ W, unsigned short* __restrict__ Wb,
                               unsigned short* __restrict__ Wt) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)K * N) return;
  const int k = (int)(i / N), n = (int)(i % N);
  const unsigned short b = f2bf(W[i]);
  Wb[i] = b;
  Wt[(long long)n * K + k] = b;
}

// dY = 2 (Y - T) / (M * N) (FP32 and BF16); part[blk] = block's sum of (Y - T)^2
__global__ __launch_bounds__(256) void k_mse_grad(long long n, float scale, const float* __restrict__ Y,
                                                  const float* __restrict__ T, float* __restrict__ dYf,
                                                  unsigned short* __restrict__ dYb, float* __restrict__ part) {
  __shared__ float sh[256];
  float acc = 0.0f;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float d = Y[i] - T[i];
    acc += d * d;
    const float g = 2.0f * d * scale;
    dYf[i] = g;
    dYb[i] = f2bf(g);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_parts(int n, const float* __restrict__ part, float* __restrict__ out, float scale) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.0f;
  for (int i = 0; i < n; ++i) s += part[i];
  *out = s * scale;
}

// db[n] = sum_m G[m*N + n], fixed order per column
__global__ void k_colsum(int M, int N, const float* __restrict__ G, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += G[(long long)m * N + n];
  db[n] = s;
}

__global__ void k_adam(long long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// rows idx[0..M) of X (ld cols) into Xb (BF16) and of T into Tb
__global__ void k_gather_rows(int M, int cx, int ct, const long long* __restrict__ idx, const float* __restrict__ X,
                              const float* __restrict__ T, unsigned short* __restrict__ Xb, float* __restrict__ Tb) {
  const int m = blockIdx.x;
  if (m >= M) return;
  const long long r = idx[m];
  for (int j = threadIdx.x; j < cx; j += blockDim.x) Xb[(long long)m * cx + j] = f2bf(X[r * cx + j]);
  if (T)
    for (int j = threadIdx.x; j < ct; j += blockDim.x) Tb[(long long)m * ct + j] = T[r * ct + j];
}

static inline dim3 g256(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// ---------------------------------------------------------------- the network
// Layer l maps K_l -> N_l: (nin, 256), (256, 256), (256, 256), (256, nout);
// 2D (192, 16), 3D (512, 32).  Parameters (FP32): W_0, b_0, W_1, b_1, ...
// with W_l row-major K_l x N_l.
void mlp_dims(const LGeo& q, int l, int* K, int* N) {
  *K = l == 0 ? q.nin : kMlpHidden;
  *N = l == kMlpLayers - 1 ? q.nout : kMlpHidden;
}
long long mlp_w_off(const LGeo& q, int l) {
  long long o = 0;
  for (int i = 0; i < l; ++i) {
    int K, N;
    mlp_dims(q, i, &K, &N);
    o += (long long)K * N + N;
  }
  return o;
}
long long mlp_b_off(const LGeo& q, int l) {
  int K, N;
  mlp_dims(q, l, &K, &N);
  return mlp_w_off(q, l) + (long long)K * N;
}
long long mlp_nparams(const LGeo& q) { return mlp_w_off(q, kMlpLayers); }

void mlp_weights_bf16(hipStream_t st, const LGeo& q, const float* P, unsigned short* Wb, unsigned short* Wt) {
  for (int l = 0; l < kMlpLayers; ++l) {
    int K, N;
    mlp_dims(q, l, &K, &N);
    const long long off = mlp_w_off(q, l);
    hipLaunchKernelGGL(k_weights_bf16, g256((long long)K * N), dim3(256), 0, st, K, N, P + off, Wb + off, Wt + off);
  }
}

// Forward over M rows of Xb (BF16, M x nin).  act: per hidden layer M x 256
// BF16; pre: per hidden layer M x 256 FP32; Y: M x nout FP32.
void mlp_forward(hipStream_t st, const LGeo& q, int M, const float* P, const unsigned short* Wt,
                 const unsigned short* Xb, float* const pre[3], unsigned short* const act[3], float* Y) {
  const unsigned short* in = Xb;
  for (int l = 0; l < kMlpLayers; ++l) {
    int K, N;
    mlp_dims(q, l, &K, &N);
    const unsigned short* Bt = Wt + mlp_w_off(q, l);  // N x K: B(k, n) = Bt[n*K + k]
    const float* bias = P + mlp_b_off(q, l);
    if (l < kMlpLayers - 1) {
      gemm_launch<EPI_GELU>(st, M, N, K, in, K, 1, Bt, 1, K, bias, nullptr, pre[l], act[l], N);
      in = act[l];
    } else {
      gemm_launch<EPI_F32>(st, M, N, K, in, K, 1, Bt, 1, K, bias, nullptr, Y, nullptr, N);
    }
  }
}

// Backward of the mean-square loss for targets T: gradient into G (same
// packing as P).  dY buffers: M x nout; dact/dpre: M x 256 (FP32 and BF16).
void mlp_backward(hipStream_t st, const LGeo& q, int M, const float* P, const unsigned short* Wb,
                  const unsigned short* Xb, float* const pre[3], unsigned short* const act[3], const float* Y,
                  const float* T, float* G, float* dYf, unsigned short* dYb, float* dHf, unsigned short* dHb,
                  float* part, float* loss) {
  (void)P;
  const long long ny = (long long)M * q.nout;
  const int nblk = 256;
  hipLaunchKernelGGL(k_mse_grad, dim3(nblk), dim3(256), 0, st, ny, 1.0f / (float)ny, Y, T, dYf, dYb, part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, st, nblk, part, loss, 1.0f / (float)ny);
  // gradient of the current layer's output pre-activation: (gf, gb), width N
  const float* gf = dYf;
  const unsigned short* gb = dYb;
  for (int l = kMlpLayers - 1; l >= 0; --l) {
    int K, N;
    mlp_dims(q, l, &K, &N);
    const unsigned short* in = l == 0 ? Xb : act[l - 1];
    // dW_l = in^T (K x M) * g (M x N): A(k, m) = in[m*K + k], B(m, n) = g[m*N + n]
    gemm_launch<EPI_F32>(st, K, N, M, in, 1, K, gb, N, 1, nullptr, nullptr, G + mlp_w_off(q, l), nullptr, N);
    hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256), dim3(256), 0, st, M, N, gf, G + mlp_b_off(q, l));
    if (l == 0) break;
    // d pre_{l-1} = (g W_l^T) * gelu'(pre_{l-1}): A(m, n) = g[m*N + n], B(n, k) = W_l[k*N + n]
    float* hf = dHf + (long long)(l & 1) * M * kMlpHidden;
    unsigned short* hb = dHb + (long long)(l & 1) * M * kMlpHidden;
    gemm_launch<EPI_DGELU>(st, M, K, N, gb, N, 1, Wb + mlp_w_off(q, l), 1, N, nullptr, pre[l - 1], hf, hb, K);
    gf = hf;
    gb = hb;
  }
}

void mlp_gather_rows(hipStream_t st, const LGeo& q, int M, const long long* idx, const float* X, const float* T,
                     unsigned short* Xb, float* Tb) {
  hipLaunchKernelGGL(k_gather_rows, dim3(M), dim3(256), 0, st, M, q.nin, q.nout, idx, X, T, Xb, Tb);
}

void mlp_to_bf16(hipStream_t st, long long n, const float* a, unsigned short* b) {
  hipLaunchKernelGGL(k_to_bf16, g256(n), dim3(256), 0, st, n, a, b);
}

void mlp_adam(hipStream_t st, long long n, float* p, const float* g, float* m1, float* m2, float lr, int t) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, g256(n), dim3(256), 0, st, n, p, g, m1, m2, lr, b1, b2, eps, c1, c2);
}

// ---------------------------------------------------------------- patches
// One thread per cluster.  The slots are anchored at the mean of the
// cluster's nc particles, the centre of its block of cs^dim lattice sites:
// slot coordinate floor((x - xm)/hl + ps/2) per axis, so lattice sites sit
// mid-slot whatever the offset between the cell grid and the lattice (the
// 1e-6 added to each cell side accumulates across the grid).  Slot s =
// sx ps + sy (2D), (sx ps + sy) ps + sz (3D); core slots are those with every
// coordinate in [ring, ring + cs), ring = (ps - cs)/2.  Writes the np patch
// and nc core slots and the anchor, or pok[c] = 0 when the cluster is not
// eligible (a slot empty or doubly filled, or the core slots not exactly the
// cluster's own particles).
template <int DIM>
__global__ void k_patch_build(Grid g, LGeo q, double hl, const int* __restrict__ sta, const int* __restrict__ end,
                              const double* __restrict__ xs, const double* __restrict__ ys,
                              const double* __restrict__ zs, int* __restrict__ pidx, int* __restrict__ cidx,
                              double* __restrict__ pctr, unsigned char* __restrict__ pok) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncell) return;
  const int ps = q.ps, cs = q.cs, ring = (q.ps - q.cs) / 2;
  int* P = pidx + c * q.np;
  int* C = cidx + c * q.nc;
  for (int s = 0; s < q.np; ++s) P[s] = -1;
  for (int s = 0; s < q.nc; ++s) C[s] = -1;
  bool ok = end[c] - sta[c] + 1 == q.nc;
  double xm = 0.0, ym = 0.0, zm = 0.0;
  if (ok) {
    for (int j = sta[c]; j <= end[c]; ++j) {
      xm += xs[j];
      ym += ys[j];
      if (DIM == 3) zm += zs[j];
    }
    xm /= q.nc;
    ym /= q.nc;
    zm /= q.nc;
  }
  pctr[3 * c] = xm;
  pctr[3 * c + 1] = ym;
  pctr[3 * c + 2] = zm;
  const double half = 0.5 * ps;
  int ix, iy, iz;
  cell_coords(g, c, &ix, &iy, &iz);
  for (int jz = DIM == 3 ? iz - 1 : 0; jz <= (DIM == 3 ? iz + 1 : 0) && ok; ++jz)
    for (int jx = ix - 1; jx <= ix + 1 && ok; ++jx)
      for (int jy = iy - 1; jy <= iy + 1 && ok; ++jy) {
        if (jx < 0 || jy < 0 || jx >= g.nx || jy >= g.ny) continue;
        if (DIM == 3 && (jz < 0 || jz >= g.nz)) continue;
        const long long jc = cell_index(g, jx, jy, jz);
        for (int j = sta[jc]; j <= end[jc]; ++j) {
          const double fx = floor((xs[j] - xm) / hl + half), fy = floor((ys[j] - ym) / hl + half);
          const double fz = DIM == 3 ? floor((zs[j] - zm) / hl + half) : 0.0;
          if (!(fx >= 0.0 && fx < ps && fy >= 0.0 && fy < ps && fz >= 0.0 && fz < (DIM == 3 ? ps : 1))) continue;
          const int sx = (int)fx, sy = (int)fy, sz = (int)fz;
          const int s = DIM == 3 ? (sx * ps + sy) * ps + sz : sx * ps + sy;
          if (P[s] >= 0) ok = false;
          P[s] = j;
          const bool incore = sx >= ring && sx < ring + cs && sy >= ring && sy < ring + cs &&
                              (DIM == 2 || (sz >= ring && sz < ring + cs));
          const bool mine = j >= sta[c] && j <= end[c];
          if (incore != mine) ok = false;
          if (incore) {
            const int o = DIM == 3 ? ((sx - ring) * cs + (sy - ring)) * cs + (sz - ring) : (sx - ring) * cs + (sy - ring);
            C[o] = j;
          }
        }
      }
  for (int s = 0; s < q.np && ok; ++s) ok = P[s] >= 0;
  pok[c] = ok ? 1 : 0;
}

void patch_build(hipStream_t st, const Grid& g, const LGeo& q, double hl, const int* sta, const int* end,
                 const double* xs, const double* ys, const double* zs, int* pidx, int* cidx, double* pctr,
                 unsigned char* pok) {
  if (q.dim == 3)
    hipLaunchKernelGGL(k_patch_build<3>, g256(g.ncell), dim3(256), 0, st, g, q, hl, sta, end, xs, ys, zs, pidx, cidx,
                       pctr, pok);
  else
    hipLaunchKernelGGL(k_patch_build<2>, g256(g.ncell), dim3(256), 0, st, g, q, hl, sta, end, xs, ys, zs, pidx, cidx,
                       pctr, pok);
}

// Features of the listed clusters for the vector r: one wavefront per
// cluster, lane l handling slots l, l + 64, ...  X row: np scaled residuals,
// np x-offsets, np y-offsets (3D: np z-offsets) in units of sigma, zeros up
// to nin; s_out[k] = RMS of the patch residual.  With z != nullptr also the
// targets T[k][o] = z[core o] / (s * out_scale), zeros for o >= nc.
template <int DIM>
__global__ __launch_bounds__(64) void k_features(LGeo q, long long nk, const int* __restrict__ list, double hl,
                                                 double inv_sigma, const int* __restrict__ pidx,
                                                 const int* __restrict__ cidx, const double* __restrict__ pctr,
                                                 const double* __restrict__ xs, const double* __restrict__ ys,
                                                 const double* __restrict__ zs, const double* __restrict__ r,
                                                 const double* __restrict__ z, double out_scale, float* __restrict__ X,
                                                 float* __restrict__ T, double* __restrict__ s_out) {
  const long long k = blockIdx.x;
  if (k >= nk) return;
  const int lane = threadIdx.x;
  const long long c = list[k];
  const double xc = pctr[3 * c], yc = pctr[3 * c + 1], zc = pctr[3 * c + 2];
  const int* P = pidx + c * q.np;
  double ss = 0.0;
  for (int s = lane; s < q.np; s += 64) {
    const double rv = r[P[s]];
    ss = fma(rv, rv, ss);
  }
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
  const double sc = sqrt(ss / q.np);
  const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
  const double half = 0.5 * q.ps;
  float* Xk = X + k * q.nin;
  for (int s = lane; s < q.np; s += 64) {
    const int j = P[s];
    const int sx = DIM == 3 ? s / (q.ps * q.ps) : s / q.ps;
    const int sy = DIM == 3 ? (s / q.ps) % q.ps : s % q.ps;
    const int sz = DIM == 3 ? s % q.ps : 0;
    Xk[s] = (float)(r[j] * inv);
    Xk[q.np + s] = (float)((xs[j] - (xc + (sx - half + 0.5) * hl)) * inv_sigma);
    Xk[2 * q.np + s] = (float)((ys[j] - (yc + (sy - half + 0.5) * hl)) * inv_sigma);
    if (DIM == 3) Xk[3 * q.np + s] = (float)((zs[j] - (zc + (sz - half + 0.5) * hl)) * inv_sigma);
  }
  for (int i = (DIM + 1) * q.np + lane; i < q.nin; i += 64) Xk[i] = 0.0f;
  if (lane == 0) s_out[k] = sc;
  if (z)
    for (int o = lane; o < q.nout; o += 64)
      T[k * q.nout + o] = o < q.nc ? (float)(z[cidx[c * q.nc + o]] * inv / out_scale) : 0.0f;
}

void patch_features(hipStream_t st, const LGeo& q, long long nk, const int* list, double hl, double sigma,
                    const int* pidx, const int* cidx, const double* pctr, const double* xs, const double* ys,
                    const double* zs, const double* r, const double* z, double out_scale, float* X, float* T,
                    double* s_out) {
  if (nk <= 0) return;
  if (q.dim == 3)
    hipLaunchKernelGGL(k_features<3>, dim3((unsigned)nk), dim3(64), 0, st, q, nk, list, hl, 1.0 / sigma, pidx, cidx,
                       pctr, xs, ys, zs, r, z, out_scale, X, T, s_out);
  else
    hipLaunchKernelGGL(k_features<2>, dim3((unsigned)nk), dim3(64), 0, st, q, nk, list, hl, 1.0 / sigma, pidx, cidx,
                       pctr, xs, ys, zs, r, z, out_scale, X, T, s_out);
}

// z[core o of cluster list[k]] = Y[k][o] * s[k] * out_scale, o < nc
__global__ void k_core_scatter(LGeo q, long long nk, const int* __restrict__ list, const int* __restrict__ cidx,
                               const float* __restrict__ Y, const double* __restrict__ s, double out_scale,
                               double* __restrict__ z) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk * q.nc) return;
  const long long k = i / q.nc;
  const int o = (int)(i % q.nc);
  z[cidx[(long long)list[k] * q.nc + o]] = (double)Y[k * q.nout + o] * s[k] * out_scale;
}

void core_scatter(hipStream_t st, const LGeo& q, long long nk, const int* list, const int* cidx, const float* Y,
                  const double* s, double out_scale, double* z) {
  if (nk <= 0) return;
  hipLaunchKernelGGL(k_core_scatter, g256(nk * q.nc), dim3(256), 0, st, q, nk, list, cidx, Y, s, out_scale, z);
}

// Local residual reduction of a core correction on its patch:
// rho[k] = |r_patch - A(patch, core) z_core| / |r_patch| (A the Gaussian
// block, 1/(2 pi sigma^2) normalisation), one wavefront per cluster.
template <int DIM>
__global__ __launch_bounds__(64) void k_patch_rho(LGeo q, long long nk, const int* __restrict__ list,
                                                  const int* __restrict__ pidx, const int* __restrict__ cidx,
                                                  const double* __restrict__ xs, const double* __restrict__ ys,
                                                  const double* __restrict__ zs, const double* __restrict__ r,
                                                  const double* __restrict__ z, double inv_two_s2, double norm,
                                                  double* __restrict__ rho) {
  const long long k = blockIdx.x;
  if (k >= nk) return;
  const int lane = threadIdx.x;
  const long long c = list[k];
  const int* C = cidx + c * q.nc;
  double num = 0.0, den = 0.0;
  for (int s = lane; s < q.np; s += 64) {
    const int j = pidx[c * q.np + s];
    const double xj = xs[j], yj = ys[j], zj = DIM == 3 ? zs[j] : 0.0;
    double az = 0.0;
    for (int o = 0; o < q.nc; ++o) {
      const int jc = C[o];
      const double dx = xj - xs[jc], dy = yj - ys[jc], dz = DIM == 3 ? zj - zs[jc] : 0.0;
      az = fma(exp_nonpos(-(dx * dx + dy * dy + dz * dz) * inv_two_s2) * norm, z[jc], az);
    }
    const double rv = r[j];
    num = fma(rv - az, rv - az, num);
    den = fma(rv, rv, den);
  }
  for (int off = 32; off > 0; off >>= 1) {
    num += __shfl_xor(num, off);
    den += __shfl_xor(den, off);
  }
  if (lane == 0) rho[k] = den > 0.0 ? sqrt(num / den) : 0.0;
}

void patch_rho(hipStream_t st, const LGeo& q, long long nk, const int* list, const int* pidx, const int* cidx,
               const double* xs, const double* ys, const double* zs, const double* r, const double* z, double sigma,
               double* rho) {
  if (nk <= 0) return;
  Kern kk = make_kern(sigma);
  if (q.dim == 3)
    hipLaunchKernelGGL(k_patch_rho<3>, dim3((unsigned)nk), dim3(64), 0, st, q, nk, list, pidx, cidx, xs, ys, zs, r, z,
                       kk.inv_two_s2, kk.norm, rho);
  else
    hipLaunchKernelGGL(k_patch_rho<2>, dim3((unsigned)nk), dim3(64), 0, st, q, nk, list, pidx, cidx, xs, ys, zs, r, z,
                       kk.inv_two_s2, kk.norm, rho);
}

}  // namespace prbf

namespace prbf {

// idx[i] = a seeded pseudo-random row in [0, npairs) (splitmix64 of seed, step, i)
__global__ void k_random_rows(int m, long long npairs, unsigned long long seed, long long step,
                              long long* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  unsigned long long z = seed + 0x9E3779B97F4A7C15ULL * (unsigned long long)(step * (long long)m + i + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  idx[i] = (long long)(z % (unsigned long long)npairs);
}

void mlp_random_rows(hipStream_t st, int m, long long npairs, unsigned long long seed, long long step,
                     long long* idx) {
  hipLaunchKernelGGL(k_random_rows, g256(m), dim3(256), 0, st, m, npairs, seed, step, idx);
}

__global__ void k_copy_f32(const float* __restrict__ a, float* __restrict__ b) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *b = *a;
}
void copy_f32(hipStream_t st, const float* a, float* b) { hipLaunchKernelGGL(k_copy_f32, dim3(1), dim3(64), 0, st, a, b); }

}  // namespace prbf

namespace prbf {

// out[k] = sum over the nc core particles of cluster list[k] of r^2
__global__ void k_core_norms(LGeo q, long long nk, const int* __restrict__ list, const int* __restrict__ cidx,
                             const double* __restrict__ r, double* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  const int* cs = cidx + (long long)list[k] * q.nc;
  double s = 0.0;
  for (int o = 0; o < q.nc; ++o) s = fma(r[cs[o]], r[cs[o]], s);
  out[k] = s;
}

void core_norms(hipStream_t st, const LGeo& q, long long nk, const int* list, const int* cidx, const double* r,
                double* out) {
  if (nk <= 0) return;
  hipLaunchKernelGGL(k_core_norms, g256(nk), dim3(256), 0, st, q, nk, list, cidx, r, out);
}

}  // namespace prbf
