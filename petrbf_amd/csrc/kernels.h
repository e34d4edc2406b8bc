// Host-side launchers of the device kernels (all stream-ordered, no syncs).
#pragma once
#include "common.h"

namespace prbf {

// cells.hip -------------------------------------------------------------------
__global__ void k_copy_ll_to_int(long long n, const long long* a, int* b);
void bbox(hipStream_t s, long long n, const double* x, const double* y, const double* z, int dim,
          double* d_part, double* d_out);
void scan_excl(hipStream_t s, long long n, const int* in, long long* out);
void sort_points(hipStream_t st, const Grid& g, long long n, const double* x, const double* y, const double* z,
                 int* d_cell, int* d_count, long long* d_start, int* d_cursor, int* d_slot, int* d_perm,
                 int* d_sta, int* d_end, double* xs, double* ys, double* zs, int* d_err);
// cell of each particle (-1 and *d_err set when outside the grid) and the counts per cell
void cell_assign(hipStream_t st, const Grid& g, long long n, const double* x, const double* y, const double* z,
                 int* d_cell, int* d_count, int* d_err);
// stable counting sort by a precomputed key with its counts (see cells.hip)
void sort_by_key(hipStream_t st, long long nkey, long long n, const int* key, const int* d_count, long long* d_start,
                 int* d_cursor, int* d_slot, int* d_perm, int* d_sta, int* d_end, const double* x, const double* y,
                 const double* z, double* xs, double* ys, double* zs);
void list_count(hipStream_t st, const Grid& g, int reach, double len, const int* tsta, const int* tend,
                const int* sta, const int* end, const double* xs, const double* ys, const double* zs,
                int* count);
void list_fill(hipStream_t st, const Grid& g, int reach, double len, const int* tsta, const int* tend,
               const int* sta, const int* end, const double* xs, const double* ys, const double* zs,
               const long long* ptr, int* idx);

// both lists in one scan when the targets are the sources
void lists2_count(hipStream_t st, const Grid& g, const int* sta, const int* end, const double* xs, const double* ys,
                  const double* zs, int* count_b, int* count_t);
void lists2_fill(hipStream_t st, const Grid& g, const int* sta, const int* end, const double* xs, const double* ys,
                 const double* zs, const long long* ptr_b, int* idx_b, const long long* ptr_t, int* idx_t);

// tree.hip (adaptive cluster structure) -------------------------------------
// counts of the particles whose node lies in the frontier [lo, hi)
void tree_count(hipStream_t st, long long n, const int* node, int lo, int hi, int* cnt);
// particles of split frontier nodes (cbase >= 0) move to their child
void tree_descend(hipStream_t st, int dim, long long n, const double* x, const double* y, const double* z, int* node,
                  int lo, int hi, const int* cbase, const double* cx, const double* cy, const double* cz);
// key[p] = leaf_of[node[p]], with the counts per leaf
void tree_leaf_keys(hipStream_t st, long long n, long long nleaf, const int* node, const int* leaf_of, int* key,
                    int* cnt);
// count (fill false: cb, ct) or fill (pb/ib, pt/it) both lists of every leaf
void tree_lists(hipStream_t st, const Grid& g, bool fill, long long nleaf, const int* sta, const int* end,
                const int* c0sta, const int* c0end, const int* lcell, const double* lcx, const double* lcy,
                const double* lcz, const double* lbl, const double* ltl, const double* xs, const double* ys,
                const double* zs, int* cb, int* ct, const long long* pb, int* ib, const long long* pt, int* it);

// matvec.hip ------------------------------------------------------------------
void matvec(hipStream_t st, int dim, int nrhs, long long ncell, const int* tsta, const int* tend,
            const long long* tptr, const int* tidx, const double* xi, const double* yi, const double* zi,
            const double* xj, const double* yj, const double* zj, const double* g, long long ldg, double* w,
            long long ldw, double sigma);

// blockchol.hip / rasm.hip ----------------------------------------------------
// Storage of the core rows M_c (T x n) for the apply.  The rows come in
// chunks of 16, the last one cut to R rows, T rounded up to a multiple of 4
// (its rows past T are zero); inside a chunk of R rows the buffer slots form
// groups of 64 (the last group holds ng = n - 64 g slots), and a group stores
// its R x ng values as R/V vector slabs [q][slot][V], V values (16 bytes) per
// slot and slab.  Lane l of a wavefront thus reads 16 bytes per slab at l*16
// bytes and every slab load is one contiguous access of up to 1 KB.
constexpr int kRowChunk = 16;
constexpr int kRowQuant = 4;
// rows stored for chunk kc of T core rows
__host__ __device__ inline int apply_chunk_rows(int T, int kc) {
  const int r = T - kRowChunk * kc;
  return r >= kRowChunk ? kRowChunk : (r + kRowQuant - 1) / kRowQuant * kRowQuant;
}
__host__ __device__ inline long long apply_chunk_base(int n, int kc) { return (long long)kRowChunk * kc * n; }
__host__ __device__ inline long long apply_elems(int n, int T) {
  if (n == 0 || T <= 0) return 0;
  const int kc = (T - 1) / kRowChunk;
  return apply_chunk_base(n, kc) + (long long)apply_chunk_rows(T, kc) * n;
}
// offset of the value of row t (all chunks) and slot b; V = 16 / sizeof(element)
template <int V>
__host__ __device__ inline long long apply_offset(int n, int T, int t, int b) {
  const int kc = t / kRowChunk, tt = t % kRowChunk, R = apply_chunk_rows(T, kc);
  const int g = b / 64, l = b % 64, ng = n - 64 * g < 64 ? n - 64 * g : 64;
  return apply_chunk_base(n, kc) + (long long)R * 64 * g + (long long)((tt / V) * ng + l) * V + tt % V;
}

// Dense untruncated block of the cluster whose buffer list starts at b0 (n
// entries), row-major into A (n x n, device)
void dense_block(hipStream_t st, int dim, long long b0, int n, const int* bidx, const double* xs, const double* ys,
                 const double* zs, double sigma, double* A);
// Workspace (elements) and LDS bytes of one block factorisation (ll: the
// left-looking kernels).
long long block_workspace(int n, int T);
size_t block_lds_bytes(bool fp32, int nmax, int tmax, int nw, int npc, bool ll);
// Clusters of order[0..nwork) with status[c] != 0 (blocks whose factorisation
// regularised a pivot) into list (unordered), their number in *count.
void failed_blocks(hipStream_t st, long long nwork, const int* order, const int* status, int* list, int* count);
// Whether the block factorisation has a kernel for (precision, wavefronts,
// panel columns, order).
bool block_chol_has_config(bool fp32, int nw, int npc, bool ll);
// Workgroups of the persistent block factorisation that fit on the device at once.
int block_chol_workgroups(bool fp32, int dim, int nw, int npc, bool ll, int nmax, int tmax);
// Factorises the blocks of clusters order[0..nwork) and writes their core rows:
// nwg persistent workgroups, each with a workspace slot of ws_slot elements
// of W, take blocks from the counter *next (reset by the launcher).
void block_chol_fp64(hipStream_t st, int dim, int nw, int npc, bool ll, long long nwork, const int* order, int* next,
                     int nwg, long long ws_slot, int nmax, int tmax, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const double* xs, const double* ys, const double* zs, double* W,
                     const long long* mptr, double* M, double sigma, int* blk_status);
void block_chol_fp32(hipStream_t st, int dim, int nw, int npc, bool ll, long long nwork, const int* order, int* next,
                     int nwg, long long ws_slot, int nmax, int tmax, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const double* xs, const double* ys, const double* zs, float* W,
                     const long long* mptr, float* M, double sigma, int* blk_status);
// z[core(c)] = M_c r[buffer(c)] for the clusters list[0..nlist) (all clusters
// 0..nlist-1 when list is null), for nrhs (1..3) vectors stored at stride ld
void rasm_apply_fp64(hipStream_t st, int nrhs, long long nlist, const int* list, const long long* bptr,
                     const int* bidx, const int* ista, const int* iend, const long long* mptr, const double* M,
                     const double* r, long long ld, double* z);
void rasm_apply_fp32(hipStream_t st, int nrhs, long long nlist, const int* list, const long long* bptr,
                     const int* bidx, const int* ista, const int* iend, const long long* mptr, const float* M,
                     const double* r, long long ld, double* z);

// krylov.hip ------------------------------------------------------------------
constexpr int kRedBlocks = 1024;  // fixed partition => run-to-run identical sums
void mdot(hipStream_t st, long long n, const double* w, const double* V, long long ldv, int k,
          double* part, double* out);
void maxpy_norm2(hipStream_t st, long long n, double* w, const double* V, long long ldv, int k,
                 const double* coef, double alpha, double* part, double* out_norm2);
void norm2sq(hipStream_t st, long long n, const double* v, double* part, double* out);
void scale(hipStream_t st, long long n, double* v, double s);
void scale_copy(hipStream_t st, long long n, const double* src, double* dst, double s);
void xpay_sub(hipStream_t st, long long n, const double* b, const double* ax, double* r);  // r = b - ax
void update_x(hipStream_t st, long long n, double* x, const double* V, long long ldv, int k, const double* y);
void gather(hipStream_t st, long long n, const int* perm, const double* src, double* dst);   // dst[s] = src[perm[s]]
void scatter(hipStream_t st, long long n, const int* perm, const double* src, double* dst);  // dst[perm[s]] = src[s]
void fill(hipStream_t st, long long n, double* v, double a);
// Device-side Arnoldi bookkeeping of one GMRES cycle (state layout in krylov.hip).
long long arnoldi_state_len(int m);
// H(:, j) from the dots dsc[0..j] and |w|^2 = dsc[64], Givens rotations, residual estimate
void arnoldi_step(hipStream_t st, int j, int m, const double* dsc, double* ks);
void arnoldi_backsub(hipStream_t st, int it, int m, const double* ks, double* y);
void scale_dev(hipStream_t st, long long n, double* v, const double* s);  // v *= *s
void set_scalar(hipStream_t st, double* p, double a);

// learned.hip -----------------------------------------------------------------
// Patch geometry of the learned preconditioner: a cluster of cs^dim lattice
// sites (its core) inside a patch of ps^dim lattice slots.  2D: 4x4 core in an
// 8x8 patch; 3D: 3x3x3 core in a 5x5x5 patch.  The network maps nin inputs
// (np scaled residuals, then dim x np slot offsets, zero-padded to a multiple
// of 16) through 3 hidden layers of 256 to nout outputs: the nc core
// corrections, zero-padded to a multiple of 16 (the pad outputs' targets are
// zero and they are never used).
struct LGeo {
  int dim, ps, cs, np, nc, nin, nout;
};
inline LGeo lgeo(int dim) {
  return dim == 3 ? LGeo{3, 5, 3, 125, 27, 512, 32} : LGeo{2, 8, 4, 64, 16, 192, 16};
}
constexpr int kMlpHidden = 256;
constexpr int kMlpLayers = 4;
constexpr int kMaxPatch = 125;  // the largest np
constexpr int kMaxCore = 27;    // the largest nc
void mlp_dims(const LGeo& q, int l, int* K, int* N);
long long mlp_w_off(const LGeo& q, int l);
long long mlp_b_off(const LGeo& q, int l);
long long mlp_nparams(const LGeo& q);
void mlp_weights_bf16(hipStream_t st, const LGeo& q, const float* P, unsigned short* Wb, unsigned short* Wt);
void mlp_forward(hipStream_t st, const LGeo& q, int M, const float* P, const unsigned short* Wt,
                 const unsigned short* Xb, float* const pre[3], unsigned short* const act[3], float* Y);
void mlp_backward(hipStream_t st, const LGeo& q, int M, const float* P, const unsigned short* Wb,
                  const unsigned short* Xb, float* const pre[3], unsigned short* const act[3], const float* Y,
                  const float* T, float* G, float* dYf, unsigned short* dYb, float* dHf, unsigned short* dHb,
                  float* part, float* loss);
void mlp_gather_rows(hipStream_t st, const LGeo& q, int M, const long long* idx, const float* X, const float* T,
                     unsigned short* Xb, float* Tb);
void mlp_to_bf16(hipStream_t st, long long n, const float* a, unsigned short* b);
void mlp_adam(hipStream_t st, long long n, float* p, const float* g, float* m1, float* m2, float lr, int t);
void mlp_random_rows(hipStream_t st, int m, long long npairs, unsigned long long seed, long long step,
                     long long* idx);
void copy_f32(hipStream_t st, const float* a, float* b);
// pidx: ncell x np slot particles; cidx: ncell x nc core particles; pctr:
// ncell x 3 anchors; pok[c] = 1 when cluster c is eligible
void patch_build(hipStream_t st, const Grid& g, const LGeo& q, double hl, const int* sta, const int* end,
                 const double* xs, const double* ys, const double* zs, int* pidx, int* cidx, double* pctr,
                 unsigned char* pok);
void patch_features(hipStream_t st, const LGeo& q, long long nk, const int* list, double hl, double sigma,
                    const int* pidx, const int* cidx, const double* pctr, const double* xs, const double* ys,
                    const double* zs, const double* r, const double* z, double out_scale, float* X, float* T,
                    double* s_out);
void core_scatter(hipStream_t st, const LGeo& q, long long nk, const int* list, const int* cidx, const float* Y,
                  const double* s, double out_scale, double* z);
void core_norms(hipStream_t st, const LGeo& q, long long nk, const int* list, const int* cidx, const double* r,
                double* out);
void patch_rho(hipStream_t st, const LGeo& q, long long nk, const int* list, const int* pidx, const int* cidx,
               const double* xs, const double* ys, const double* zs, const double* r, const double* z, double sigma,
               double* rho);

}  // namespace prbf
