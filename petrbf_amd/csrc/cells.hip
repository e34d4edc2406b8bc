// Cell list, stable sort and cluster lists (kernel family (a)).
//
// Device restatement of 2d/get_cluster.h:20-166 (single rank), 2d/get_buffer.h:11-81
// and 2d/get_trunc.h:14-78, with the 3D twins.  Results are bit-identical to the
// reference's: cell ids use the same IEEE divide and floor, box tests the same
// fabs(x - xc) < len in FP64, and list order is core first, then neighbour
// cells jx-outer / jy-inner (3D: jx, jy, jz) with particles in sorted order.
#include "common.h"
#include "kernels.h"

namespace prbf {

// ---------------------------------------------------------------- bounding box
__global__ void k_bbox_partial(long long n, const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, int dim, double* __restrict__ part) {
  __shared__ double smin[3][256], smax[3][256];
  int t = threadIdx.x;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (long long p = (long long)blockIdx.x * blockDim.x + t; p < n; p += (long long)gridDim.x * blockDim.x) {
    double v[3] = {x[p], y[p], dim == 3 ? z[p] : 0.0};
    for (int d = 0; d < 3; ++d) { lo[d] = fmin(lo[d], v[d]); hi[d] = fmax(hi[d], v[d]); }
  }
  for (int d = 0; d < 3; ++d) { smin[d][t] = lo[d]; smax[d][t] = hi[d]; }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s)
      for (int d = 0; d < 3; ++d) {
        smin[d][t] = fmin(smin[d][t], smin[d][t + s]);
        smax[d][t] = fmax(smax[d][t], smax[d][t + s]);
      }
    __syncthreads();
  }
  if (t == 0)
    for (int d = 0; d < 3; ++d) { part[blockIdx.x * 6 + d] = smin[d][0]; part[blockIdx.x * 6 + 3 + d] = smax[d][0]; }
}

__global__ void k_bbox_final(int nparts, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = 0; b < nparts; ++b)
    for (int d = 0; d < 3; ++d) { lo[d] = fmin(lo[d], part[b * 6 + d]); hi[d] = fmax(hi[d], part[b * 6 + 3 + d]); }
  for (int d = 0; d < 3; ++d) { out[d] = lo[d]; out[3 + d] = hi[d]; }
}

void bbox(hipStream_t s, long long n, const double* x, const double* y, const double* z, int dim,
          double* d_part /*>= 6*256*/, double* d_out /*6*/) {
  int nb = (int)std::min<long long>(256, (n + 255) / 256);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, s, n, x, y, z, dim, d_part);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, s, nb, d_part, d_out);
}

// ---------------------------------------------------------------- cell of each particle
__global__ void k_cell_assign(long long n, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, Grid g, int* __restrict__ cell,
                              int* __restrict__ count, int* __restrict__ err) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  // (int)floor((x - xmin)/box), 2d/get_cluster.h:78-79 -- IEEE divide, no contraction.
  double fx = floor((x[p] - g.xmin) / g.box);
  double fy = floor((y[p] - g.ymin) / g.box);
  double fz = g.dim == 3 ? floor((z[p] - g.zmin) / g.box) : 0.0;
  // range test in FP64 first, so NaN and out-of-grid points are rejected
  // before any conversion to int
  if (!(fx >= 0.0 && fx < g.nx && fy >= 0.0 && fy < g.ny && fz >= 0.0 && fz < g.nz)) {
    atomicOr(err, 1);
    cell[p] = -1;
    return;
  }
  int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  int c = (int)cell_index(g, ix, iy, iz);
  cell[p] = c;
  atomicAdd(&count[c], 1);
}

// Exclusive scan of int counts into int64 offsets, out[n] = total; integer
// sums, so the result does not depend on the order.  Three passes over
// 1024-element blocks, with no scratch beyond out: (1) each block scans
// itself by wavefront shuffles and leaves its total in the slot of its first
// element (whose block-local exclusive value is 0); (2) one workgroup turns
// those totals into block offsets in place and writes out[n]; (3) each block
// adds its offset to its other elements.
constexpr int kScanB = 1024;

__device__ inline long long wave_incl_scan(long long v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// inclusive scan of v over the workgroup's 1024 threads (wsum: 16 slots)
__device__ inline long long block_incl_scan(long long v, long long* wsum) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  v = wave_incl_scan(v, lane);
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  if (w == 0) {
    long long s = lane < kScanB / 64 ? wsum[lane] : 0;
    s = wave_incl_scan(s, lane);
    if (lane < kScanB / 64) wsum[lane] = s;
  }
  __syncthreads();
  if (w > 0) v += wsum[w - 1];
  return v;
}

__global__ __launch_bounds__(kScanB) void k_scan_blocks(long long n, const int* __restrict__ in,
                                                        long long* __restrict__ out) {
  __shared__ long long wsum[kScanB / 64];
  const long long i = (long long)blockIdx.x * kScanB + threadIdx.x;
  const long long v = i < n ? in[i] : 0;
  const long long incl = block_incl_scan(v, wsum);
  if (threadIdx.x == kScanB - 1) out[(long long)blockIdx.x * kScanB] = incl;  // block total
  else if (i + 1 < n) out[i + 1] = incl;                                       // exclusive of i + 1
}

__global__ __launch_bounds__(kScanB) void k_scan_offsets(long long n, long long nb, long long* __restrict__ out) {
  __shared__ long long wsum[kScanB / 64];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (long long b0 = 0; b0 < nb; b0 += kScanB) {
    const long long b = b0 + threadIdx.x;
    const long long v = b < nb ? out[b * kScanB] : 0;
    const long long incl = block_incl_scan(v, wsum);
    const long long base = carry;
    if (b < nb) out[b * kScanB] = base + incl - v;
    __syncthreads();
    if (threadIdx.x == kScanB - 1) carry = base + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(kScanB) void k_scan_add(long long n, long long* __restrict__ out) {
  const long long b0 = (long long)blockIdx.x * kScanB, i = b0 + threadIdx.x;
  if (threadIdx.x > 0 && i < n) out[i] += out[b0];
}

void scan_excl(hipStream_t s, long long n, const int* in, long long* out) {
  if (n <= 0) {
    PRBF_HIP(hipMemsetAsync(out, 0, sizeof(long long), s));
    return;
  }
  const long long nb = (n + kScanB - 1) / kScanB;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kScanB), 0, s, n, in, out);
  hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(kScanB), 0, s, n, nb, out);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanB), 0, s, n, out);
}

// Unordered placement into cell slots, then per-cell ordering by original
// index: the result equals the reference's stable counting sort
// (2d/get_cluster.h:123-130: slots filled in increasing particle index).
__global__ void k_cell_scatter(long long n, const int* __restrict__ cell, int* __restrict__ cursor,
                               int* __restrict__ perm) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int c = cell[p];
  if (c < 0) return;  // rejected point; the caller reports the error
  int s = atomicAdd(&cursor[c], 1);
  perm[s] = (int)p;
}

// One wavefront per cell: each particle's place in its cell is the number of
// the cell's particles with a smaller original index (indices are distinct),
// read from the unordered slots `slot` and written to `perm`.  k^2/64 steps
// per cell of k particles on one wavefront, so a crowded cell costs what its
// size says and no more.
__global__ __launch_bounds__(256) void k_cell_sortseg(long long ncell, const long long* __restrict__ start,
                                                      const int* __restrict__ slot, int* __restrict__ perm) {
  const int lane = threadIdx.x & 63;
  const long long c = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncell) return;
  const long long a = start[c], b = start[c + 1];
  for (long long i0 = a; i0 < b; i0 += 64) {
    const long long i = i0 + lane;
    const int v = i < b ? slot[i] : 0;
    long long rank = 0;
    for (long long j = a; j < b; ++j) rank += slot[j] < v;
    if (i < b) perm[a + rank] = v;
  }
}

__global__ void k_cell_ranges(long long ncell, const long long* __restrict__ start, int* __restrict__ sta,
                              int* __restrict__ end) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  sta[c] = (int)start[c];
  end[c] = (int)start[c + 1] - 1;  // iend = ista + count - 1 (2d/get_cluster.h:100-107)
}

__global__ void k_gather_sorted(long long n, const int* __restrict__ perm, const double* __restrict__ x,
                                const double* __restrict__ y, const double* __restrict__ z,
                                double* __restrict__ xs, double* __restrict__ ys, double* __restrict__ zs) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int p = perm[s];
  xs[s] = x[p];
  ys[s] = y[p];
  if (zs) zs[s] = z[p];
}

// Stable counting sort of n particles by key[p] in [0, nkey) with the
// per-key counts already in d_count: key offsets (d_start, nkey + 1), the
// permutation (sorted position -> original index), key ranges sta/end and the
// sorted coordinates.  d_slot is n ints of scratch.
void sort_by_key(hipStream_t st, long long nkey, long long n, const int* key, const int* d_count, long long* d_start,
                 int* d_cursor, int* d_slot, int* d_perm, int* d_sta, int* d_end, const double* x, const double* y,
                 const double* z, double* xs, double* ys, double* zs) {
  const int nbp = (int)((n + 255) / 256);
  const int nbc = (int)((nkey + 255) / 256);
  scan_excl(st, nkey, d_count, d_start);
  hipLaunchKernelGGL(k_copy_ll_to_int, dim3(nbc), dim3(256), 0, st, nkey, d_start, d_cursor);
  // slots of rejected points stay 0 (a valid index) until the caller reports the error
  PRBF_HIP(hipMemsetAsync(d_slot, 0, sizeof(int) * n, st));
  PRBF_HIP(hipMemsetAsync(d_perm, 0, sizeof(int) * n, st));
  hipLaunchKernelGGL(k_cell_scatter, dim3(nbp), dim3(256), 0, st, n, key, d_cursor, d_slot);
  hipLaunchKernelGGL(k_cell_sortseg, dim3((unsigned)((nkey + 3) / 4)), dim3(256), 0, st, nkey, d_start, d_slot,
                     d_perm);
  hipLaunchKernelGGL(k_cell_ranges, dim3(nbc), dim3(256), 0, st, nkey, d_start, d_sta, d_end);
  hipLaunchKernelGGL(k_gather_sorted, dim3(nbp), dim3(256), 0, st, n, d_perm, x, y, z, xs, ys, zs);
}

void cell_assign(hipStream_t st, const Grid& g, long long n, const double* x, const double* y, const double* z,
                 int* d_cell, int* d_count, int* d_err) {
  const int nbp = (int)((n + 255) / 256);
  PRBF_HIP(hipMemsetAsync(d_count, 0, sizeof(int) * g.ncell, st));
  hipLaunchKernelGGL(k_cell_assign, dim3(nbp), dim3(256), 0, st, n, x, y, z, g, d_cell, d_count, d_err);
}

void sort_points(hipStream_t st, const Grid& g, long long n, const double* x, const double* y, const double* z,
                 int* d_cell, int* d_count, long long* d_start, int* d_cursor, int* d_slot, int* d_perm,
                 int* d_sta, int* d_end, double* xs, double* ys, double* zs, int* d_err) {
  cell_assign(st, g, n, x, y, z, d_cell, d_count, d_err);
  sort_by_key(st, g.ncell, n, d_cell, d_count, d_start, d_cursor, d_slot, d_perm, d_sta, d_end, x, y, z, xs, ys, zs);
}

__global__ void k_copy_ll_to_int(long long n, const long long* __restrict__ a, int* __restrict__ b) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (int)a[i];
}

// ---------------------------------------------------------------- cluster lists
// Cluster lists, one wavefront per cluster: TWO, the buffer list (reach rb,
// half-width lb) and the truncation list (rt, lt) of a cluster structure
// whose targets are its sources (the solve), in one scan; otherwise one list
// (rb, lb) of source particles for the clusters of a target set (evaluation
// at distinct targets; tsta/tend are the targets, an empty cluster gets an
// empty list, 2d/get_buffer.h:79-81).  The candidates are the source
// particles of the cluster's own cell, taken unconditionally, then those of
// the neighbour cells in the reference's loop order (jx outer, jy, jz inner)
// that a list reaches, tested against the box around the cell centre with
// the reference's FP64 test fabs(x - xc) < len per axis
// (2d/get_buffer.h:64, 2d/get_trunc.h:63).  Lanes take consecutive
// candidates and ballots place the hits, so each list comes out in the
// reference's order with coalesced writes.  Count pass (FILL false):
// per-cluster sizes; fill pass: the indices at pb / pt.
template <int DIM, bool FILL, bool TWO>
__global__ __launch_bounds__(256) void k_lists(Grid g, int rb, double lb, int rt, double lt,
                                                const int* __restrict__ tsta, const int* __restrict__ tend,
                                                const int* __restrict__ sta, const int* __restrict__ end,
                                                const double* __restrict__ xs, const double* __restrict__ ys,
                                                const double* __restrict__ zs, int* __restrict__ cb,
                                                int* __restrict__ ct, const long long* __restrict__ pb,
                                                int* __restrict__ ib, const long long* __restrict__ pt,
                                                int* __restrict__ it) {
  __shared__ int s_off[4][65];           // per wavefront: first candidate of each cell of a chunk
  __shared__ int s_j0[4][64];            // the cell's first particle
  __shared__ unsigned char s_fl[4][64];  // 1: buffer list reaches it, 2: truncation list, 4: own cell
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long c = __builtin_amdgcn_readfirstlane((int)((long long)blockIdx.x * 4 + w));
  if (c >= g.ncell) return;  // (wavefronts never wait for each other)
  const unsigned long long below = (1ull << lane) - 1;
  long long kb = 0, kt = 0;
  if (tsta[c] <= tend[c]) {
    int ix, iy, iz;
    cell_coords(g, c, &ix, &iy, &iz);
    const double xc = cell_centre(g.xmin, ix, g.box);
    const double yc = cell_centre(g.ymin, iy, g.box);
    const double zc = DIM == 3 ? cell_centre(g.zmin, iz, g.box) : 0.0;
    const int R = TWO ? max(rb, rt) : rb, W = 2 * R + 1;
    const int M = DIM == 3 ? W * W * W : W * W;             // cells visited, own first
    const int own = DIM == 3 ? (R * W + R) * W + R : R * W + R;  // own cell's place in loop order
    const long long bb = FILL ? pb[c] : 0, bt = (FILL && TWO) ? pt[c] : 0;
    for (int q0 = 0; q0 < M; q0 += 64) {
      const int q = q0 + lane;
      int cnt = 0, j0 = 0, fl = 0;
      if (q < M) {
        const int id = q == 0 ? own : (q - 1 < own ? q - 1 : q);
        const int ox = (DIM == 3 ? id / (W * W) : id / W) - R;
        const int oy = (DIM == 3 ? (id / W) % W : id % W) - R;
        const int oz = DIM == 3 ? id % W - R : 0;
        const int jx = ix + ox, jy = iy + oy, jz = iz + oz;
        const bool inside = jx >= 0 && jx < g.nx && jy >= 0 && jy < g.ny && (DIM == 2 || (jz >= 0 && jz < g.nz));
        const int ax = abs(ox), ay = abs(oy), az = abs(oz);
        const bool fb = ax <= rb && ay <= rb && az <= rb;
        const bool ft = TWO && ax <= rt && ay <= rt && az <= rt;
        if (inside && (q == 0 || fb || ft)) {
          const long long jc = cell_index(g, jx, jy, jz);
          j0 = sta[jc];
          cnt = max(0, end[jc] - j0 + 1);
          fl = q == 0 ? 4 : (fb ? 1 : 0) | (ft ? 2 : 0);
        }
      }
      int incl = cnt;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
      }
      const int tot = __shfl(incl, 63);
      s_off[w][lane] = incl - cnt;
      s_j0[w][lane] = j0;
      s_fl[w][lane] = (unsigned char)fl;
      if (lane == 63) s_off[w][64] = tot;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (int k0 = 0; k0 < tot; k0 += 64) {
        const int k = k0 + lane;
        bool hb = false, ht = false;
        int j = 0;
        if (k < tot) {
          int lo = 0, hi = 63;  // the last cell whose first candidate is <= k
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[w][mid] <= k) lo = mid;
            else hi = mid - 1;
          }
          j = s_j0[w][lo] + (k - s_off[w][lo]);
          const int f = s_fl[w][lo];
          if (f & 4) {
            hb = ht = true;
          } else {
            const double dx = fabs(xs[j] - xc), dy = fabs(ys[j] - yc), dz = DIM == 3 ? fabs(zs[j] - zc) : 0.0;
            hb = (f & 1) && dx < lb && dy < lb;
            ht = (f & 2) && dx < lt && dy < lt;
            if (DIM == 3) {
              hb = hb && dz < lb;
              ht = ht && dz < lt;
            }
          }
        }
        const unsigned long long mb = __ballot(hb), mt = TWO ? __ballot(ht) : 0ull;
        if (FILL) {
          if (hb) ib[bb + kb + __popcll(mb & below)] = j;
          if (TWO && ht) it[bt + kt + __popcll(mt & below)] = j;
        }
        kb += __popcll(mb);
        kt += __popcll(mt);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (!FILL && lane == 0) {
    cb[c] = (int)kb;
    if (TWO) ct[c] = (int)kt;
  }
}

void lists2_count(hipStream_t st, const Grid& g, const int* sta, const int* end, const double* xs, const double* ys,
                  const double* zs, int* count_b, int* count_t) {
  const unsigned nb = (unsigned)((g.ncell + 3) / 4);
  if (g.dim == 2)
    hipLaunchKernelGGL((k_lists<2, false, true>), dim3(nb), dim3(256), 0, st, g, g.nb_buffer, g.buffer_len,
                       g.nb_trunc, g.trunc_len, sta, end, sta, end, xs, ys, zs, count_b, count_t, nullptr, nullptr,
                       nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_lists<3, false, true>), dim3(nb), dim3(256), 0, st, g, g.nb_buffer, g.buffer_len,
                       g.nb_trunc, g.trunc_len, sta, end, sta, end, xs, ys, zs, count_b, count_t, nullptr, nullptr,
                       nullptr, nullptr);
}

void lists2_fill(hipStream_t st, const Grid& g, const int* sta, const int* end, const double* xs, const double* ys,
                 const double* zs, const long long* ptr_b, int* idx_b, const long long* ptr_t, int* idx_t) {
  const unsigned nb = (unsigned)((g.ncell + 3) / 4);
  if (g.dim == 2)
    hipLaunchKernelGGL((k_lists<2, true, true>), dim3(nb), dim3(256), 0, st, g, g.nb_buffer, g.buffer_len,
                       g.nb_trunc, g.trunc_len, sta, end, sta, end, xs, ys, zs, nullptr, nullptr, ptr_b, idx_b,
                       ptr_t, idx_t);
  else
    hipLaunchKernelGGL((k_lists<3, true, true>), dim3(nb), dim3(256), 0, st, g, g.nb_buffer, g.buffer_len,
                       g.nb_trunc, g.trunc_len, sta, end, sta, end, xs, ys, zs, nullptr, nullptr, ptr_b, idx_b,
                       ptr_t, idx_t);
}

void list_count(hipStream_t st, const Grid& g, int reach, double len, const int* tsta, const int* tend,
                const int* sta, const int* end, const double* xs, const double* ys, const double* zs,
                int* count) {
  const unsigned nb = (unsigned)((g.ncell + 3) / 4);
  if (g.dim == 2)
    hipLaunchKernelGGL((k_lists<2, false, false>), dim3(nb), dim3(256), 0, st, g, reach, len, 0, 0.0, tsta, tend,
                       sta, end, xs, ys, zs, count, nullptr, nullptr, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_lists<3, false, false>), dim3(nb), dim3(256), 0, st, g, reach, len, 0, 0.0, tsta, tend,
                       sta, end, xs, ys, zs, count, nullptr, nullptr, nullptr, nullptr, nullptr);
}

void list_fill(hipStream_t st, const Grid& g, int reach, double len, const int* tsta, const int* tend,
               const int* sta, const int* end, const double* xs, const double* ys, const double* zs,
               const long long* ptr, int* idx) {
  const unsigned nb = (unsigned)((g.ncell + 3) / 4);
  if (g.dim == 2)
    hipLaunchKernelGGL((k_lists<2, true, false>), dim3(nb), dim3(256), 0, st, g, reach, len, 0, 0.0, tsta, tend,
                       sta, end, xs, ys, zs, nullptr, nullptr, ptr, idx, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_lists<3, true, false>), dim3(nb), dim3(256), 0, st, g, reach, len, 0, 0.0, tsta, tend,
                       sta, end, xs, ys, zs, nullptr, nullptr, ptr, idx, nullptr, nullptr);
}

}  // namespace prbf
