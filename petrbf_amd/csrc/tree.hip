// Adaptive cluster structure: the reference's uniform grid (2d/get_cluster.h:
// 20-166) as level 0, cells of more than `cap` particles split into 2^dim
// children of half the side, recursively; the leaves are the clusters.
//
//   * child k of a node with centre c at depth d: k = 2 qx + qy (3D 4 qz +
//     2 qx + qy, the grid's own index order) with q = (coordinate >= c) per
//     axis; the child's centre is c +- box 2^-(d+2) per axis;
//   * leaves in depth-first order, children in index order, empty ones kept;
//     particles sorted by leaf, then by original index;
//   * a leaf at depth d has buffer half-width buffer_len 2^-d and truncation
//     half-width (trunc_len - box/2) + box 2^-(d+1) (trunc_len at depth 0);
//   * a leaf's list is its own particles, then those of the level-0 cells
//     within the reference's reach of its level-0 cell, in the reference's
//     scan order (jx outer, jy, jz inner) and sorted order, except its own,
//     that pass the box test fabs(x - xc) < len per axis around the leaf
//     centre (2d/get_buffer.h:64, 2d/get_trunc.h:63).
// With no cell split this is the reference's structure, lists included.
//
// The refinement runs level by level: the particles of the current frontier
// are counted per node on the device, the host (which holds the few
// thousand-to-million node records) decides the splits and the children's
// centres, and the particles of split nodes descend one level.  The lists
// are built like the uniform ones: one wavefront per leaf, candidates in scan
// order, ballots place the hits.
#include "common.h"
#include "kernels.h"

namespace prbf {

// cnt[v - lo] += 1 for the particles whose node v is in the frontier [lo, hi)
__global__ void k_node_count(long long n, const int* __restrict__ node, int lo, int hi, int* __restrict__ cnt) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int v = node[p];
  if (v >= lo && v < hi) atomicAdd(&cnt[v - lo], 1);
}

// Particles of split frontier nodes (cbase[v - lo] >= 0) move to their child.
template <int DIM>
__global__ void k_descend(long long n, const double* __restrict__ x, const double* __restrict__ y,
                          const double* __restrict__ z, int* __restrict__ node, int lo, int hi,
                          const int* __restrict__ cbase, const double* __restrict__ cx, const double* __restrict__ cy,
                          const double* __restrict__ cz) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int v = node[p];
  if (v < lo || v >= hi) return;
  const int f = v - lo, b = cbase[f];
  if (b < 0) return;
  const int qx = x[p] >= cx[f], qy = y[p] >= cy[f];
  const int k = DIM == 3 ? 4 * (z[p] >= cz[f]) + 2 * qx + qy : 2 * qx + qy;
  node[p] = b + k;
}

// key[p] = leaf of the particle's node; counts per leaf
__global__ void k_leaf_key(long long n, const int* __restrict__ node, const int* __restrict__ leaf_of,
                           int* __restrict__ key, int* __restrict__ cnt) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int l = leaf_of[node[p]];
  key[p] = l;
  atomicAdd(&cnt[l], 1);
}

void tree_count(hipStream_t st, long long n, const int* node, int lo, int hi, int* cnt) {
  PRBF_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * (hi - lo), st));
  hipLaunchKernelGGL(k_node_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, node, lo, hi, cnt);
}

void tree_descend(hipStream_t st, int dim, long long n, const double* x, const double* y, const double* z, int* node,
                  int lo, int hi, const int* cbase, const double* cx, const double* cy, const double* cz) {
  const dim3 grid((unsigned)((n + 255) / 256));
  if (dim == 3)
    hipLaunchKernelGGL(k_descend<3>, grid, dim3(256), 0, st, n, x, y, z, node, lo, hi, cbase, cx, cy, cz);
  else
    hipLaunchKernelGGL(k_descend<2>, grid, dim3(256), 0, st, n, x, y, z, node, lo, hi, cbase, cx, cy, cz);
}

void tree_leaf_keys(hipStream_t st, long long n, long long nleaf, const int* node, const int* leaf_of, int* key,
                    int* cnt) {
  PRBF_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * nleaf, st));
  hipLaunchKernelGGL(k_leaf_key, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, node, leaf_of, key, cnt);
}

// Both lists of every leaf in one scan, one wavefront per leaf (the uniform
// k_lists with the leaf's own centre and half-widths).  Candidate segments in
// order: the leaf's own particles (taken unconditionally), then the level-0
// cells of the scan window, the leaf's own level-0 cell split around the
// leaf's range into the particles before it and after it.
template <int DIM, bool FILL>
__global__ __launch_bounds__(256) void k_lists_tree(
    Grid g, long long nleaf, const int* __restrict__ sta, const int* __restrict__ end,
    const int* __restrict__ c0sta, const int* __restrict__ c0end, const int* __restrict__ lcell,
    const double* __restrict__ lcx, const double* __restrict__ lcy, const double* __restrict__ lcz,
    const double* __restrict__ lbl, const double* __restrict__ ltl, const double* __restrict__ xs,
    const double* __restrict__ ys, const double* __restrict__ zs, int* __restrict__ cb, int* __restrict__ ct,
    const long long* __restrict__ pb, int* __restrict__ ib, const long long* __restrict__ pt, int* __restrict__ it) {
  __shared__ int s_off[4][65];
  __shared__ int s_j0[4][64];
  __shared__ unsigned char s_fl[4][64];  // 1: buffer reaches it, 2: truncation, 4: own leaf
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long c = __builtin_amdgcn_readfirstlane((int)((long long)blockIdx.x * 4 + w));
  if (c >= nleaf) return;
  const unsigned long long below = (1ull << lane) - 1;
  long long kb = 0, kt = 0;
  const int ls = sta[c], le = end[c];
  if (ls <= le) {
    int ix, iy, iz;
    const long long c0 = lcell[c];
    cell_coords(g, c0, &ix, &iy, &iz);
    const double xc = lcx[c], yc = lcy[c], zc = DIM == 3 ? lcz[c] : 0.0;
    const double lb = lbl[c], lt = ltl[c];
    const int rb = g.nb_buffer, rt = g.nb_trunc;
    const int R = max(rb, rt), W = 2 * R + 1;
    const int M = DIM == 3 ? W * W * W : W * W;
    const int own = DIM == 3 ? (R * W + R) * W + R : R * W + R;
    const int S = M + 2;  // own leaf, the window's cells, the own cell in two parts
    const long long bb = FILL ? pb[c] : 0, bt = FILL ? pt[c] : 0;
    for (int q0 = 0; q0 < S; q0 += 64) {
      const int q = q0 + lane;
      int cnt = 0, j0 = 0, fl = 0;
      if (q == 0) {
        j0 = ls;
        cnt = le - ls + 1;
        fl = 4;
      } else if (q < S) {
        const int pos = q - 1;
        const int id = pos <= own ? pos : pos - 1;  // positions own and own + 1 are the own cell
        const int ox = (DIM == 3 ? id / (W * W) : id / W) - R;
        const int oy = (DIM == 3 ? (id / W) % W : id % W) - R;
        const int oz = DIM == 3 ? id % W - R : 0;
        const int jx = ix + ox, jy = iy + oy, jz = iz + oz;
        const bool inside = jx >= 0 && jx < g.nx && jy >= 0 && jy < g.ny && (DIM == 2 || (jz >= 0 && jz < g.nz));
        const int ax = abs(ox), ay = abs(oy), az = abs(oz);
        const bool fb = ax <= rb && ay <= rb && az <= rb;
        const bool ft = ax <= rt && ay <= rt && az <= rt;
        if (inside && (fb || ft)) {
          const long long jc = cell_index(g, jx, jy, jz);
          int a = c0sta[jc], b = c0end[jc];
          if (pos == own) b = ls - 1;          // own cell before the leaf
          else if (pos == own + 1) a = le + 1;  // own cell after the leaf
          j0 = a;
          cnt = max(0, b - a + 1);
          fl = (fb ? 1 : 0) | (ft ? 2 : 0);
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
          int lo = 0, hi = 63;  // the last segment whose first candidate is <= k
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
        const unsigned long long mb = __ballot(hb), mt = __ballot(ht);
        if (FILL) {
          if (hb) ib[bb + kb + __popcll(mb & below)] = j;
          if (ht) it[bt + kt + __popcll(mt & below)] = j;
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
    ct[c] = (int)kt;
  }
}

void tree_lists(hipStream_t st, const Grid& g, bool fill, long long nleaf, const int* sta, const int* end,
                const int* c0sta, const int* c0end, const int* lcell, const double* lcx, const double* lcy,
                const double* lcz, const double* lbl, const double* ltl, const double* xs, const double* ys,
                const double* zs, int* cb, int* ct, const long long* pb, int* ib, const long long* pt, int* it) {
  const dim3 grid((unsigned)((nleaf + 3) / 4));
#define PRBF_TL(D, F)                                                                                            \
  hipLaunchKernelGGL((k_lists_tree<D, F>), grid, dim3(256), 0, st, g, nleaf, sta, end, c0sta, c0end, lcell, lcx, \
                     lcy, lcz, lbl, ltl, xs, ys, zs, cb, ct, pb, ib, pt, it)
  if (g.dim == 3) {
    if (fill) PRBF_TL(3, true);
    else PRBF_TL(3, false);
  } else {
    if (fill) PRBF_TL(2, true);
    else PRBF_TL(2, false);
  }
#undef PRBF_TL
}

}  // namespace prbf
