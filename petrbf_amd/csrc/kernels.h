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
                 int* d_cell, int* d_count, long long* d_start, int* d_cursor, int* d_perm,
                 int* d_sta, int* d_end, double* xs, double* ys, double* zs, int* d_err);
void list_count(hipStream_t st, const Grid& g, int reach, double len, const int* tsta, const int* tend,
                const int* sta, const int* end, const double* xs, const double* ys, const double* zs,
                int* count);
void list_fill(hipStream_t st, const Grid& g, int reach, double len, const int* tsta, const int* tend,
               const int* sta, const int* end, const double* xs, const double* ys, const double* zs,
               const long long* ptr, int* idx);

// matvec.hip ------------------------------------------------------------------
void matvec(hipStream_t st, int dim, int nrhs, long long ncell, const int* tsta, const int* tend,
            const long long* tptr, const int* tidx, const double* xi, const double* yi, const double* zi,
            const double* xj, const double* yj, const double* zj, const double* g, long long ldg, double* w,
            long long ldw, double sigma);

// blockchol.hip / rasm.hip ----------------------------------------------------
// Workspace (elements) and LDS bytes of one block factorisation.
long long block_workspace(int n, int T);
size_t block_lds_bytes(bool fp32, int nmax, int tmax, int nw, int npc);
// Factorises the blocks of clusters [c0, c0+ncl) and writes their core rows.
// clist != nullptr: factorise clusters clist[0..ncl) instead of [c0, c0+ncl).
void block_chol_fp64(hipStream_t st, int dim, int nw, int npc, long long c0, const int* clist, long long ncl,
                     int nmax, int tmax, const long long* bptr, const int* bidx, const int* ista, const int* iend,
                     const double* xs, const double* ys, const double* zs, const long long* wptr, double* W,
                     const long long* mptr, double* M, double sigma, int* err);
void block_chol_fp32(hipStream_t st, int dim, int nw, int npc, long long c0, long long ncl, int nmax, int tmax,
                     const long long* bptr, const int* bidx, const int* ista, const int* iend, const double* xs,
                     const double* ys, const double* zs, const long long* wptr, float* W, const long long* mptr,
                     float* M, double sigma, int* blk_status);
void rasm_apply_fp64(hipStream_t st, long long ncell, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const long long* mptr, const double* M,
                     const double* r, double* z, int tmax);
void rasm_apply_fp32(hipStream_t st, long long ncell, const long long* bptr, const int* bidx,
                     const int* ista, const int* iend, const long long* mptr, const float* M,
                     const long long* fb, const double* M64, const double* r, double* z);

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

}  // namespace prbf
