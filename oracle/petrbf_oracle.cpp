// petrbf_oracle.cpp -- CPU restatement of barbagroup/petrbf's Gaussian-RBF
// interpolation path, used ONLY as test infrastructure: the checker for the
// HIP library (tests/, __graft_entry__.smoke()) and the timed host-core
// baseline in bench.py.  Nothing in the product path links or calls this file.
//
// What it restates (paths relative to the reference tree):
//   cell list + stable sort ........ 2d/get_cluster.h:20-114  (3d/get_cluster.h:16-130)
//   buffer (RASM subdomain) list ... 2d/get_buffer.h:11-81    (3d/get_buffer.h)
//   truncation (source) list ....... 2d/get_trunc.h:14-78     (3d/get_trunc.h)
//   truncated Gaussian matvec ...... 2d/matmult.cxx:28-43     (3d/matmult.cxx:28-44)
//   evaluation at targets .......... 2d/get_vorticity.h:11-26, 2d/vorticity_evaluation.cxx:59-70
//   dense RASM blocks .............. 2d/matmult.cxx:83-93
//   block LU (PETSc -sub_pc_type lu -sub_mat_type dense -> LAPACK dgetf2/dgetrs)
//   PCASM restrict apply ........... subdomains of 2d/rbf_interpolation.cxx:213-229
//   GMRES(30), left PC, classical Gram-Schmidt, zero guess, PETSc default
//   convergence test ............... KSPSolve call at 2d/rbf_interpolation.cxx:232
//   FGMRES(m), right PC (Saad 2003, Alg. 9.6) for the flexible solver.
// Single rank: the MPI scans reduce to identities, no ghost cells, and the
// local cluster range is [0, n_cells) (2d/mpi_range.cxx:4-17 with nprocs=1).
//
// Generalisation: the reference's nsigma_box / sigma_buffer / sigma_trunc are
// ints in units of sigma (2d/par.h:72-74); here they are doubles, with every
// formula evaluated in the same operation order.
//
// Build: oracle/build_oracle.sh (g++ -O2 -fopenmp -ffp-contract=off).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

const double kEpsf = 1e-6;  // 2d/par.h:153

struct Oracle {
  int dim = 2;
  double sigma = 0, nsigma_box = 0, sigma_buffer = 0, sigma_trunc = 0;
  int nb_buffer = 0, nb_trunc = 0;
  // grid
  double cmin[3] = {0, 0, 0}, cmax[3] = {0, 0, 0}, box = 0;
  int nc[3] = {1, 1, 1};
  long long ncell = 0;
  std::vector<int> cix[3];
  std::vector<double> ctr[3];
  // targets (i-set) and sources (j-set), in cluster-sorted order
  long long ni = 0, nj = 0;
  std::vector<double> xi[3], xj[3];
  std::vector<int> perm_i, perm_j;  // sorted position -> original index
  std::vector<int> ista, iend, jsta, jend;
  bool same_sets = false;
  // lists (CSR over cells)
  std::vector<long long> bptr, tptr;
  std::vector<int> bidx, tidx;
  // RASM factors
  std::vector<long long> fptr;   // offset of block c in fac (n_c^2 doubles)
  std::vector<double> fac;
  std::vector<int> piv;          // offset bptr[c]
  bool rasm_ready = false;
  long long rasm_stride = 1;
  // adaptive structure: clusters are the leaves of an occupancy-capped
  // subdivision of the grid (ncell then counts leaves); per leaf its centre,
  // depth, grid cell, and per grid cell its particle range in sorted order
  bool tree = false;
  long long ncell0 = 0;
  std::vector<double> lctr[3];
  std::vector<int> ldepth;
  std::vector<long long> lcell0;
  std::vector<int> c0sta, c0end;
};

inline long long cell_id(const Oracle& o, int ix, int iy, int iz) {
  // 2d/get_cluster.h:78-80 (ic = ix*ny+iy); 3d/get_cluster.h:101 (iz*nx*ny+ix*ny+iy)
  if (o.dim == 2) return (long long)ix * o.nc[1] + iy;
  return (long long)iz * o.nc[0] * o.nc[1] + (long long)ix * o.nc[1] + iy;
}

// Cell of one point; returns false when the point falls outside the grid.
inline bool point_cell(const Oracle& o, const double* p, long long* ic) {
  int k[3] = {0, 0, 0};
  for (int d = 0; d < o.dim; ++d) {
    k[d] = (int)floor((p[d] - o.cmin[d]) / o.box);  // 2d/get_cluster.h:78-79
    if (k[d] < 0 || k[d] >= o.nc[d]) return false;
  }
  *ic = cell_id(o, k[0], k[1], k[2]);
  return true;
}

// Stable counting sort of one point set by cell (2d/get_cluster.h:77-138 with
// one rank: the per-cell offsets are exclusive prefix sums of the counts and
// particles are placed in increasing original index).
bool sort_set(const Oracle& o, long long n, const double* const* src,
              std::vector<double>* dst, std::vector<int>& perm,
              std::vector<int>& sta, std::vector<int>& end) {
  std::vector<long long> cell(n);
  std::vector<int> cnt(o.ncell, 0);
  for (long long p = 0; p < n; ++p) {
    double q[3] = {src[0][p], src[1][p], o.dim == 3 ? src[2][p] : 0.0};
    if (!point_cell(o, q, &cell[p])) return false;
    cnt[cell[p]]++;
  }
  sta.assign(o.ncell, 0);
  end.assign(o.ncell, -1);
  std::vector<long long> off(o.ncell);
  long long id = 0;
  for (long long c = 0; c < o.ncell; ++c) {
    off[c] = id;
    sta[c] = (int)id;
    id += cnt[c];
    end[c] = (int)(id - 1);
  }
  perm.assign(n, 0);
  for (int d = 0; d < o.dim; ++d) dst[d].assign(n, 0.0);
  for (long long p = 0; p < n; ++p) {
    long long s = off[cell[p]]++;
    perm[s] = (int)p;
    for (int d = 0; d < o.dim; ++d) dst[d][s] = src[d][p];
  }
  return true;
}

// Grid over the bounding box of both sets (2d/get_cluster.h:20-72;
// 2d/vorticity_evaluation.cxx:59-70 takes the union of target and source boxes).
void make_grid(Oracle& o, const double* lo, const double* hi) {
  double nb = o.nsigma_box;
  o.nb_buffer = (int)ceil((o.sigma_buffer - nb + kEpsf) / 2 / nb);  // get_cluster.h:20
  o.nb_trunc = (int)ceil((o.sigma_trunc - nb + kEpsf) / 2 / nb);    // get_cluster.h:21
  o.box = o.nsigma_box * o.sigma + kEpsf;                           // get_cluster.h:29
  for (int d = 0; d < 3; ++d) { o.nc[d] = 1; o.cmin[d] = 0; o.cmax[d] = 0; }
  for (int d = 0; d < o.dim; ++d) {
    o.cmin[d] = lo[d] - kEpsf;
    o.cmax[d] = hi[d] + kEpsf;
    o.nc[d] = (int)ceil((o.cmax[d] - o.cmin[d]) / o.box);           // get_cluster.h:32-33
  }
  o.ncell = (long long)o.nc[0] * o.nc[1] * o.nc[2];
  for (int d = 0; d < 3; ++d) { o.cix[d].assign(o.ncell, 0); o.ctr[d].assign(o.ncell, 0.0); }
  // centres: 2D loops ix, iy; 3D loops iz, ix, iy (get_cluster.h:55-72)
  for (int iz = 0; iz < o.nc[2]; ++iz)
    for (int ix = 0; ix < o.nc[0]; ++ix)
      for (int iy = 0; iy < o.nc[1]; ++iy) {
        long long c = cell_id(o, ix, iy, iz);
        o.cix[0][c] = ix; o.cix[1][c] = iy; o.cix[2][c] = iz;
        o.ctr[0][c] = o.cmin[0] + (ix + 0.5) * o.box;
        o.ctr[1][c] = o.cmin[1] + (iy + 0.5) * o.box;
        if (o.dim == 3) o.ctr[2][c] = o.cmin[2] + (iz + 0.5) * o.box;
      }
}

// Neighbour scan of cluster c in the reference order: jx outer, jy inner (3D:
// jx, jy, jz), skipping c itself (get_buffer.h:31-76, get_trunc.h:45-76).
template <class F>
void for_neighbours(const Oracle& o, long long c, int reach, F&& f) {
  int ix = o.cix[0][c], iy = o.cix[1][c], iz = o.cix[2][c];
  int lo[3], hi[3];
  int k[3] = {ix, iy, iz};
  for (int d = 0; d < 3; ++d) {
    if (d < o.dim) {
      lo[d] = std::max(0, k[d] - reach);
      hi[d] = std::min(o.nc[d] - 1, k[d] + reach);
    } else {
      lo[d] = hi[d] = 0;
    }
  }
  for (int jx = lo[0]; jx <= hi[0]; ++jx)
    for (int jy = lo[1]; jy <= hi[1]; ++jy)
      for (int jz = lo[2]; jz <= hi[2]; ++jz) {
        if (jx == ix && jy == iy && jz == iz) continue;
        f(cell_id(o, jx, jy, jz));
      }
}

inline bool in_box(const Oracle& o, long long c, const std::vector<double>* X, long long p, double len) {
  for (int d = 0; d < o.dim; ++d)
    if (!(fabs(X[d][p] - o.ctr[d][c]) < len)) return false;
  return true;
}

// Buffer list of cluster c over the i-set: core first, then box-filtered
// neighbours (get_buffer.h:11-81).  Empty cluster -> empty list.
template <class Emit>
void buffer_of(const Oracle& o, long long c, Emit&& emit) {
  double len = o.sigma_buffer * o.sigma / 2 + kEpsf;  // get_buffer.h:11
  if (o.ista[c] > o.iend[c]) return;
  for (int j = o.ista[c]; j <= o.iend[c]; ++j) emit(j);
  for_neighbours(o, c, o.nb_buffer, [&](long long jc) {
    for (int j = o.ista[jc]; j <= o.iend[jc]; ++j)
      if (in_box(o, c, o.xi, j, len)) emit(j);
  });
}

// Truncation list of cluster c over the j-set (get_trunc.h:14-78); empty when
// the cluster holds no target.
template <class Emit>
void trunc_of(const Oracle& o, long long c, Emit&& emit) {
  double len = o.sigma_trunc * o.sigma / 2 + kEpsf;  // get_trunc.h:14
  if (o.ista[c] > o.iend[c]) return;
  for (int j = o.jsta[c]; j <= o.jend[c]; ++j) emit(j);
  for_neighbours(o, c, o.nb_trunc, [&](long long jc) {
    for (int j = o.jsta[jc]; j <= o.jend[jc]; ++j)
      if (in_box(o, c, o.xj, j, len)) emit(j);
  });
}

// ---- adaptive cluster structure ---------------------------------------------
// The reference clusters particles on a uniform grid (2d/get_cluster.h:20-166)
// and builds each cluster's lists from the neighbour cells
// (2d/get_buffer.h:11-81, 2d/get_trunc.h:14-78).  For strongly non-uniform
// particle sets the adaptive structure keeps that grid as level 0 and splits
// any cell holding more than `cap` particles into 2^dim children of half its
// side, recursively, down to depth `dmax`; the leaves are the clusters.
//   * child k of a node with centre c at depth d: k = 2 qx + qy (3D: 4 qz +
//     2 qx + qy, the grid's own index order), q = (coordinate >= c) per axis,
//     centre c +- box 2^-(d+2) per axis; level-0 centres are the reference's;
//   * leaves in depth-first order, children in index order, empty ones kept;
//     particles sorted by leaf, then by original index;
//   * a leaf at depth d has buffer half-width buffer_len 2^-d (the block keeps
//     its shape relative to the cluster) and truncation half-width
//     (trunc_len - box/2) + box 2^-(d+1) (the same cutoff beyond the cluster;
//     trunc_len itself at depth 0);
//   * a leaf's list is its own particles, then those of the level-0 cells
//     within the reference's reach of its level-0 cell, in the reference's
//     scan order and sorted order, except its own, that pass the per-axis box
//     test around the leaf centre.
// With no cell split the structure and its lists are the reference's.
struct TreeLeaf {
  double c[3];
  int depth;
  long long cell0;
};

void tree_split(const Oracle& o, const std::vector<int>& pts, const double* const* src, const double c[3],
                int depth, long long cell0, int cap, int dmax, std::vector<TreeLeaf>& leaves,
                std::vector<std::vector<int>>& members) {
  if ((int)pts.size() <= cap || depth >= dmax) {
    leaves.push_back({{c[0], c[1], c[2]}, depth, cell0});
    members.push_back(pts);
    return;
  }
  const int nch = 1 << o.dim;
  std::vector<std::vector<int>> ch(nch);
  for (int p : pts) {
    const int qx = src[0][p] >= c[0], qy = src[1][p] >= c[1], qz = o.dim == 3 && src[2][p] >= c[2];
    ch[o.dim == 3 ? 4 * qz + 2 * qx + qy : 2 * qx + qy].push_back(p);
  }
  const double q = ldexp(o.box, -(depth + 2));
  for (int k = 0; k < nch; ++k) {
    const int qx = (k >> 1) & 1, qy = k & 1, qz = (k >> 2) & 1;
    const double cc[3] = {c[0] + (qx ? q : -q), c[1] + (qy ? q : -q), o.dim == 3 ? c[2] + (qz ? q : -q) : 0.0};
    tree_split(o, ch[k], src, cc, depth + 1, cell0, cap, dmax, leaves, members);
  }
}

// Leaves, sorted order and level-0 ranges of one point set (the i-set = the
// j-set).  Returns false when a point falls outside the grid.
bool tree_build(Oracle& o, long long n, const double* const* src, int cap, int dmax) {
  std::vector<std::vector<int>> in_cell(o.ncell);
  for (long long p = 0; p < n; ++p) {
    double q[3] = {src[0][p], src[1][p], o.dim == 3 ? src[2][p] : 0.0};
    long long ic;
    if (!point_cell(o, q, &ic)) return false;
    in_cell[ic].push_back((int)p);
  }
  std::vector<TreeLeaf> leaves;
  std::vector<std::vector<int>> members;
  std::vector<long long> first_leaf(o.ncell + 1, 0);
  for (long long c = 0; c < o.ncell; ++c) {
    first_leaf[c] = (long long)leaves.size();
    const double c0[3] = {o.ctr[0][c], o.ctr[1][c], o.dim == 3 ? o.ctr[2][c] : 0.0};
    tree_split(o, in_cell[c], src, c0, 0, c, cap, dmax, leaves, members);
  }
  first_leaf[o.ncell] = (long long)leaves.size();
  const long long nl = (long long)leaves.size();
  o.ncell0 = o.ncell;
  o.ista.assign(nl, 0);
  o.iend.assign(nl, -1);
  o.perm_i.assign(n, 0);
  for (int d = 0; d < o.dim; ++d) o.xi[d].assign(n, 0.0);
  for (int d = 0; d < 3; ++d) o.lctr[d].assign(nl, 0.0);
  o.ldepth.assign(nl, 0);
  o.lcell0.assign(nl, 0);
  long long s = 0;
  for (long long l = 0; l < nl; ++l) {
    o.ista[l] = (int)s;
    for (int p : members[l]) {
      o.perm_i[s] = p;
      for (int d = 0; d < o.dim; ++d) o.xi[d][s] = src[d][p];
      ++s;
    }
    o.iend[l] = (int)s - 1;
    for (int d = 0; d < 3; ++d) o.lctr[d][l] = leaves[l].c[d];
    o.ldepth[l] = leaves[l].depth;
    o.lcell0[l] = leaves[l].cell0;
  }
  o.c0sta.assign(o.ncell0, 0);
  o.c0end.assign(o.ncell0, -1);
  for (long long c = 0; c < o.ncell0; ++c) {
    o.c0sta[c] = o.ista[first_leaf[c]];
    o.c0end[c] = first_leaf[c + 1] < nl ? o.ista[first_leaf[c + 1]] - 1 : (int)n - 1;
  }
  o.ncell = nl;
  return true;
}

// Half-widths of a leaf's buffer and truncation boxes at depth d.
inline double tree_blen(const Oracle& o, int d) { return ldexp(o.sigma_buffer * o.sigma / 2 + kEpsf, -d); }
inline double tree_tlen(const Oracle& o, int d) {
  const double t = o.sigma_trunc * o.sigma / 2 + kEpsf;
  return d == 0 ? t : (t - o.box * 0.5) + ldexp(o.box, -(d + 1));
}

template <class Emit>
void tree_list_of(const Oracle& o, long long l, int reach, double len, Emit&& emit) {
  if (o.ista[l] > o.iend[l]) return;
  for (int j = o.ista[l]; j <= o.iend[l]; ++j) emit(j);
  const long long c0 = o.lcell0[l];
  const int k[3] = {o.cix[0][c0], o.cix[1][c0], o.cix[2][c0]};
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = d < o.dim ? std::max(0, k[d] - reach) : 0;
    hi[d] = d < o.dim ? std::min(o.nc[d] - 1, k[d] + reach) : 0;
  }
  for (int jx = lo[0]; jx <= hi[0]; ++jx)
    for (int jy = lo[1]; jy <= hi[1]; ++jy)
      for (int jz = lo[2]; jz <= hi[2]; ++jz) {
        const long long jc = cell_id(o, jx, jy, jz);
        for (int j = o.c0sta[jc]; j <= o.c0end[jc]; ++j) {
          if (j >= o.ista[l] && j <= o.iend[l]) continue;
          bool in = true;
          for (int d = 0; d < o.dim; ++d) in = in && fabs(o.xi[d][j] - o.lctr[d][l]) < len;
          if (in) emit(j);
        }
      }
}

template <class ListFn>
void build_csr(const Oracle& o, ListFn fn, std::vector<long long>& ptr, std::vector<int>& idx) {
  ptr.assign(o.ncell + 1, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (long long c = 0; c < o.ncell; ++c) {
    long long k = 0;
    fn(o, c, [&](int) { ++k; });
    ptr[c + 1] = k;
  }
  for (long long c = 0; c < o.ncell; ++c) ptr[c + 1] += ptr[c];
  idx.assign(ptr[o.ncell], 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (long long c = 0; c < o.ncell; ++c) {
    long long k = ptr[c];
    fn(o, c, [&](int j) { idx[k++] = j; });
  }
}

// Gaussian with the reference's operation order (2d/matmult.cxx:39-40; 3D
// keeps the 2D normalisation 1/(2 pi sigma^2), 3d/matmult.cxx:40-41).
inline double gauss(const Oracle& o, double dx, double dy, double dz) {
  double r2 = (o.dim == 3) ? (dx * dx + dy * dy + dz * dz) : (dx * dx + dy * dy);
  return exp(-r2 / (2 * o.sigma * o.sigma)) / (2 * M_PI * o.sigma * o.sigma);
}

// w_i = sum_{j in trunc(c(i))} g_j G(x_i - x_j), accumulated in list order
// (2d/matmult.cxx:33-43).  nrhs right-hand sides share the list.
void matvec(const Oracle& o, int nrhs, const double* g, long long ldg, double* w, long long ldw) {
#pragma omp parallel for schedule(dynamic, 64)
  for (long long c = 0; c < o.ncell; ++c) {
    long long t0 = o.tptr[c], t1 = o.tptr[c + 1];
    for (int i = o.ista[c]; i <= o.iend[c]; ++i) {
      for (int r = 0; r < nrhs; ++r) {
        double acc = 0;
        const double* gr = g + r * ldg;
        for (long long t = t0; t < t1; ++t) {
          int j = o.tidx[t];
          double dz = (o.dim == 3) ? o.xi[2][i] - o.xj[2][j] : 0.0;
          double dx = o.xi[0][i] - o.xj[0][j];
          double dy = o.xi[1][i] - o.xj[1][j];
          double r2 = (o.dim == 3) ? (dx * dx + dy * dy + dz * dz) : (dx * dx + dy * dy);
          acc += gr[j] * exp(-r2 / (2 * o.sigma * o.sigma)) / (2 * M_PI * o.sigma * o.sigma);
        }
        w[r * ldw + i] = acc;
      }
    }
  }
}

// LAPACK dgetf2 semantics on a row-major n x n matrix (the blocks are
// symmetric, so row/column storage only transposes the factor layout):
// pivot = first index of max |a(i,k)|, i >= k; rows swapped; multipliers
// scaled by the reciprocal pivot; rank-1 update of the trailing block.
// Stored as column-major to mirror LAPACK: a[col*n + row].
bool lu_factor(int n, double* a, int* ipiv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = fabs(a[(long long)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      double v = fabs(a[(long long)k * n + i]);
      if (v > amax) { amax = v; p = i; }
    }
    ipiv[k] = p;
    if (a[(long long)k * n + p] == 0.0) return false;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[(long long)j * n + k], a[(long long)j * n + p]);
    double rp = 1.0 / a[(long long)k * n + k];
    for (int i = k + 1; i < n; ++i) a[(long long)k * n + i] *= rp;
    for (int j = k + 1; j < n; ++j) {
      double u = a[(long long)j * n + k];
      if (u == 0.0) continue;
      double* col = a + (long long)j * n;
      const double* l = a + (long long)k * n;
      for (int i = k + 1; i < n; ++i) col[i] -= l[i] * u;
    }
  }
  return true;
}

// dgetrs (no transpose): apply the row interchanges, then L (unit) and U solves.
void lu_solve(int n, const double* a, const int* ipiv, double* b) {
  for (int k = 0; k < n; ++k)
    if (ipiv[k] != k) std::swap(b[k], b[ipiv[k]]);
  for (int j = 0; j < n; ++j) {
    double bj = b[j];
    if (bj != 0.0)
      for (int i = j + 1; i < n; ++i) b[i] -= bj * a[(long long)j * n + i];
  }
  for (int j = n - 1; j >= 0; --j) {
    if (b[j] != 0.0) {
      b[j] /= a[(long long)j * n + j];
      double bj = b[j];
      for (int i = 0; i < j; ++i) b[i] -= bj * a[(long long)j * n + i];
    }
  }
}

// Dense untruncated block over the buffer particles (2d/matmult.cxx:83-93).
void block_of(const Oracle& o, long long c, double* A) {
  long long b0 = o.bptr[c];
  int n = (int)(o.bptr[c + 1] - b0);
  for (int i = 0; i < n; ++i) {
    int pi = o.bidx[b0 + i];
    for (int j = 0; j < n; ++j) {
      int pj = o.bidx[b0 + j];
      double dx = o.xi[0][pi] - o.xi[0][pj];
      double dy = o.xi[1][pi] - o.xi[1][pj];
      double dz = (o.dim == 3) ? o.xi[2][pi] - o.xi[2][pj] : 0.0;
      A[(long long)i * n + j] = gauss(o, dx, dy, dz);
    }
  }
}

// stride > 1 factorises only clusters c with c % stride == 0 (a timing sample).
int rasm_setup(Oracle& o, long long stride = 1) {
  o.fptr.assign(o.ncell + 1, 0);
  for (long long c = 0; c < o.ncell; ++c) {
    long long n = (c % stride == 0) ? o.bptr[c + 1] - o.bptr[c] : 0;
    o.fptr[c + 1] = o.fptr[c] + n * n;
  }
  o.fac.assign(o.fptr[o.ncell], 0.0);
  o.piv.assign(o.bptr[o.ncell], 0);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : bad)
  for (long long c = 0; c < o.ncell; ++c) {
    int n = (int)(o.bptr[c + 1] - o.bptr[c]);
    if (n == 0 || c % stride != 0) continue;
    double* A = o.fac.data() + o.fptr[c];
    block_of(o, c, A);  // symmetric: row-major == column-major
    if (!lu_factor(n, A, o.piv.data() + o.bptr[c])) bad++;
  }
  o.rasm_ready = (bad == 0) && stride == 1;
  o.rasm_stride = stride;
  return bad == 0 ? 0 : 5;
}

// PCASM, restrict variant: z = sum_c P_core(c) A_c^{-1} R_buffer(c) r.  Cores
// partition the unknowns, so every z entry is written exactly once.
void rasm_apply(const Oracle& o, const double* r, double* z) {
  const long long stride = o.rasm_stride;
#pragma omp parallel
  {
    std::vector<double> buf;
#pragma omp for schedule(dynamic, 64)
    for (long long c = 0; c < o.ncell; ++c) {
      long long b0 = o.bptr[c];
      int n = (int)(o.bptr[c + 1] - b0);
      if (n == 0 || c % stride != 0) continue;
      buf.resize(n);
      for (int i = 0; i < n; ++i) buf[i] = r[o.bidx[b0 + i]];
      lu_solve(n, o.fac.data() + o.fptr[c], o.piv.data() + b0, buf.data());
      int ncore = o.iend[c] - o.ista[c] + 1;
      for (int i = 0; i < ncore; ++i) z[o.bidx[b0 + i]] = buf[i];
    }
  }
}

// ---- deterministic vector kernels (fixed chunking, ordered combination) ----
const long long kChunk = 1 << 15;

double dot(long long n, const double* a, const double* b) {
  long long nch = (n + kChunk - 1) / kChunk;
  std::vector<double> part(nch, 0.0);
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < nch; ++k) {
    long long i0 = k * kChunk, i1 = std::min(n, i0 + kChunk);
    double s = 0;
    for (long long i = i0; i < i1; ++i) s += a[i] * b[i];
    part[k] = s;
  }
  double s = 0;
  for (long long k = 0; k < nch; ++k) s += part[k];
  return s;
}

double norm2(long long n, const double* a) { return sqrt(dot(n, a, a)); }

struct GmresOut {
  int its = 0;
  int reason = 0;  // >0 converged, <0 diverged (PETSc KSPConvergedReason numbering)
  double rnorm = 0;
  std::vector<double> hist;
};

// PETSc KSPConvergedDefault for the preconditioned norm with a zero guess.
struct ConvTest {
  double rtol, abstol, dtol, rnorm0 = 0, ttol = 0;
  int operator()(int n, double rnorm) {
    if (n == 0) { rnorm0 = rnorm; ttol = std::max(rtol * rnorm0, abstol); }
    if (rnorm != rnorm) return -9;                 // KSP_DIVERGED_NANORINF
    if (rnorm <= ttol) return rnorm < abstol ? 3 : 2;  // CONVERGED_ATOL / RTOL
    if (rnorm >= dtol * rnorm0) return -4;         // KSP_DIVERGED_DTOL
    return 0;
  }
};

// Restarted GMRES, left preconditioning, classical Gram-Schmidt without
// refinement, Givens rotations; mirrors PETSc's KSPSolve_GMRES /
// KSPGMRESCycle / KSPGMRESUpdateHessenberg / KSPGMRESBuildSoln with default
// options (restart 30, haptol 1e-30, abstol 1e-50, dtol 1e5).
template <class ApplyA, class ApplyB>
GmresOut gmres_left(long long n, ApplyA A, ApplyB B, const double* b, double* x,
                    double rtol, int restart, int max_it) {
  GmresOut out;
  const int m = restart;
  const double haptol = 1e-30;
  ConvTest conv{rtol, 1e-50, 1e5};
  std::vector<std::vector<double>> V(m + 1, std::vector<double>(n));
  std::vector<double> tmp(n), tmp2(n);
  std::vector<double> hh((m + 1) * (m + 1), 0.0), cc(m + 1), ss(m + 1), rs(m + 2), lhh(m + 1), nrs(m + 1);
  auto HH = [&](int i, int j) -> double& { return hh[(long long)j * (m + 1) + i]; };
  std::fill(x, x + n, 0.0);
  bool guess_zero = true;
  int its = 0, itcount = 0, reason = 0;
  while (!reason) {
    // KSPInitialResidual: VV(0) = B (b - A x)
    if (guess_zero) {
      std::memcpy(tmp.data(), b, sizeof(double) * n);
    } else {
      A(x, tmp2.data());
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < n; ++i) tmp[i] = b[i] - tmp2[i];
    }
    B(tmp.data(), V[0].data());
    // KSPGMRESCycle
    double res = norm2(n, V[0].data());
    {
      double s = 1.0 / res;
      if (res != 0.0) {
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < n; ++i) V[0][i] *= s;
      }
    }
    rs[0] = res;
    out.hist.push_back(res);
    if (res == 0.0) { reason = 3; break; }
    reason = conv(its, res);
    int loc = 0;
    bool hapend = false;
    while (!reason && loc < m && its < max_it) {
      A(V[loc].data(), tmp.data());
      B(tmp.data(), V[loc + 1].data());
      // classical Gram-Schmidt: lhh = V^T w; w -= V lhh
      double* w = V[loc + 1].data();
      for (int j = 0; j <= loc; ++j) lhh[j] = dot(n, w, V[j].data());
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < n; ++i) {
        double s = w[i];
        for (int j = 0; j <= loc; ++j) s += (-lhh[j]) * V[j][i];
        w[i] = s;
      }
      for (int j = 0; j <= loc; ++j) HH(j, loc) = lhh[j];
      double tt = norm2(n, w);
      if (tt != 0.0) {
        double s = 1.0 / tt;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < n; ++i) w[i] *= s;
      }
      HH(loc + 1, loc) = tt;
      double hapbnd = fabs(tt / rs[loc]);
      if (hapbnd > haptol) hapbnd = haptol;
      if (tt < hapbnd) hapend = true;
      // KSPGMRESUpdateHessenberg
      for (int j = 1; j <= loc; ++j) {
        double t0 = HH(j - 1, loc);
        HH(j - 1, loc) = cc[j - 1] * t0 + ss[j - 1] * HH(j, loc);
        HH(j, loc) = cc[j - 1] * HH(j, loc) - ss[j - 1] * t0;
      }
      if (!hapend) {
        double a0 = HH(loc, loc), a1 = HH(loc + 1, loc);
        double t = sqrt(a0 * a0 + a1 * a1);
        if (t == 0.0) { reason = -2; break; }  // KSP_DIVERGED_NULL
        cc[loc] = a0 / t;
        ss[loc] = a1 / t;
        rs[loc + 1] = -(ss[loc] * rs[loc]);
        rs[loc] = cc[loc] * rs[loc];
        HH(loc, loc) = cc[loc] * a0 + ss[loc] * a1;
        res = fabs(rs[loc + 1]);
      } else {
        res = 0.0;
      }
      loc++;
      its++;
      out.hist.push_back(res);
      reason = conv(its, res);
      if (hapend && !reason) reason = 8;  // KSP_CONVERGED_HAPPY_BREAKDOWN
    }
    // KSPGMRESBuildSoln with it = loc-1
    int it = loc - 1;
    if (it >= 0) {
      nrs[it] = rs[it] / HH(it, it);
      for (int ii = 1; ii <= it; ++ii) {
        int k = it - ii;
        double t = rs[k];
        for (int j = k + 1; j <= it; ++j) t = t - HH(k, j) * nrs[j];
        nrs[k] = t / HH(k, k);
      }
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j <= it; ++j) s += nrs[j] * V[j][i];
        x[i] += s;
      }
    }
    itcount += loc;
    if (itcount >= max_it && !reason) { reason = -3; break; }  // KSP_DIVERGED_ITS
    guess_zero = false;
  }
  out.its = its;
  out.reason = reason;
  out.rnorm = out.hist.empty() ? 0.0 : out.hist.back();
  return out;
}

// Flexible GMRES(m) with right preconditioning (Saad 2003, Alg. 9.6): the
// basis Z of preconditioned directions is kept, so B may change per step.
// Residual estimate = true unpreconditioned residual (in exact arithmetic).
template <class ApplyA, class ApplyB>
GmresOut fgmres_right(long long n, ApplyA A, ApplyB B, const double* b, double* x,
                      double rtol, int restart, int max_it) {
  GmresOut out;
  const int m = restart;
  ConvTest conv{rtol, 1e-50, 1e5};
  std::vector<std::vector<double>> V(m + 1, std::vector<double>(n)), Z(m, std::vector<double>(n));
  std::vector<double> tmp(n);
  std::vector<double> hh((m + 1) * (m + 1), 0.0), cc(m + 1), ss(m + 1), rs(m + 2), lhh(m + 1), nrs(m + 1);
  auto HH = [&](int i, int j) -> double& { return hh[(long long)j * (m + 1) + i]; };
  std::fill(x, x + n, 0.0);
  bool guess_zero = true;
  int its = 0, itcount = 0, reason = 0;
  while (!reason) {
    if (guess_zero) {
      std::memcpy(V[0].data(), b, sizeof(double) * n);
    } else {
      A(x, tmp.data());
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < n; ++i) V[0][i] = b[i] - tmp[i];
    }
    double res = norm2(n, V[0].data());
    if (res != 0.0) {
      double s = 1.0 / res;
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < n; ++i) V[0][i] *= s;
    }
    rs[0] = res;
    out.hist.push_back(res);
    if (res == 0.0) { reason = 3; break; }
    reason = conv(its, res);
    int loc = 0;
    bool hapend = false;
    while (!reason && loc < m && its < max_it) {
      B(V[loc].data(), Z[loc].data());
      A(Z[loc].data(), V[loc + 1].data());
      double* w = V[loc + 1].data();
      for (int j = 0; j <= loc; ++j) lhh[j] = dot(n, w, V[j].data());
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < n; ++i) {
        double s = w[i];
        for (int j = 0; j <= loc; ++j) s += (-lhh[j]) * V[j][i];
        w[i] = s;
      }
      for (int j = 0; j <= loc; ++j) HH(j, loc) = lhh[j];
      double tt = norm2(n, w);
      if (tt != 0.0) {
        double s = 1.0 / tt;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < n; ++i) w[i] *= s;
      }
      HH(loc + 1, loc) = tt;
      double hapbnd = std::min(fabs(tt / rs[loc]), 1e-30);
      if (tt < hapbnd) hapend = true;
      for (int j = 1; j <= loc; ++j) {
        double t0 = HH(j - 1, loc);
        HH(j - 1, loc) = cc[j - 1] * t0 + ss[j - 1] * HH(j, loc);
        HH(j, loc) = cc[j - 1] * HH(j, loc) - ss[j - 1] * t0;
      }
      if (!hapend) {
        double a0 = HH(loc, loc), a1 = HH(loc + 1, loc);
        double t = sqrt(a0 * a0 + a1 * a1);
        if (t == 0.0) { reason = -2; break; }
        cc[loc] = a0 / t;
        ss[loc] = a1 / t;
        rs[loc + 1] = -(ss[loc] * rs[loc]);
        rs[loc] = cc[loc] * rs[loc];
        HH(loc, loc) = cc[loc] * a0 + ss[loc] * a1;
        res = fabs(rs[loc + 1]);
      } else {
        res = 0.0;
      }
      loc++;
      its++;
      out.hist.push_back(res);
      reason = conv(its, res);
      if (hapend && !reason) reason = 8;
    }
    int it = loc - 1;
    if (it >= 0) {
      nrs[it] = rs[it] / HH(it, it);
      for (int ii = 1; ii <= it; ++ii) {
        int k = it - ii;
        double t = rs[k];
        for (int j = k + 1; j <= it; ++j) t = t - HH(k, j) * nrs[j];
        nrs[k] = t / HH(k, k);
      }
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j <= it; ++j) s += nrs[j] * Z[j][i];
        x[i] += s;
      }
    }
    itcount += loc;
    if (itcount >= max_it && !reason) { reason = -3; break; }  // KSP_DIVERGED_ITS
    guess_zero = false;
  }
  out.its = its;
  out.reason = reason;
  out.rnorm = out.hist.empty() ? 0.0 : out.hist.back();
  return out;
}

}  // namespace

// ============================== C ABI =====================================
extern "C" {

void* orc_create(int dim, double sigma, double nsigma_box, double sigma_buffer, double sigma_trunc) {
  if (dim != 2 && dim != 3) return nullptr;
  Oracle* o = new Oracle;
  o->dim = dim;
  o->sigma = sigma;
  o->nsigma_box = nsigma_box;
  o->sigma_buffer = sigma_buffer;
  o->sigma_trunc = sigma_trunc;
  return o;
}

void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

int orc_num_threads(void) { return omp_get_max_threads(); }

// Build the grid over the union bounding box of both sets, sort both, build
// the buffer (i-set) and truncation (j-set) lists.  xj == nullptr means the
// source set is the target set (the solve).
int orc_set_points(void* h, long long ni, const double* xi, const double* yi, const double* zi,
                   long long nj, const double* xj, const double* yj, const double* zj) {
  Oracle& o = *static_cast<Oracle*>(h);
  o.same_sets = (xj == nullptr);
  o.tree = false;
  if (o.same_sets) { nj = ni; xj = xi; yj = yi; zj = zi; }
  if (ni <= 0 || nj <= 0) return 1;
  const double* si[3] = {xi, yi, zi};
  const double* sj[3] = {xj, yj, zj};
  double lo[3], hi[3];
  for (int d = 0; d < o.dim; ++d) {
    lo[d] = si[d][0]; hi[d] = si[d][0];
    for (long long p = 0; p < ni; ++p) { lo[d] = std::min(lo[d], si[d][p]); hi[d] = std::max(hi[d], si[d][p]); }
    for (long long p = 0; p < nj; ++p) { lo[d] = std::min(lo[d], sj[d][p]); hi[d] = std::max(hi[d], sj[d][p]); }
  }
  make_grid(o, lo, hi);
  o.ni = ni;
  o.nj = nj;
  if (!sort_set(o, ni, si, o.xi, o.perm_i, o.ista, o.iend)) return 1;
  if (!sort_set(o, nj, sj, o.xj, o.perm_j, o.jsta, o.jend)) return 1;
  build_csr(o, [](const Oracle& oo, long long c, auto&& e) { buffer_of(oo, c, e); }, o.bptr, o.bidx);
  build_csr(o, [](const Oracle& oo, long long c, auto&& e) { trunc_of(oo, c, e); }, o.tptr, o.tidx);
  o.rasm_ready = false;
  return 0;
}

// Adaptive structure over one point set (the solve): the grid of
// orc_set_points, cells of more than `cap` particles split recursively down to
// depth `dmax`, the leaves as clusters (see tree_build).
int orc_set_points_tree(void* h, long long n, const double* x, const double* y, const double* z, int cap,
                        int dmax) {
  Oracle& o = *static_cast<Oracle*>(h);
  if (n <= 0 || cap < 1 || dmax < 0 || dmax > 20) return 1;
  const double* si[3] = {x, y, z};
  double lo[3], hi[3];
  for (int d = 0; d < o.dim; ++d) {
    lo[d] = si[d][0]; hi[d] = si[d][0];
    for (long long p = 0; p < n; ++p) { lo[d] = std::min(lo[d], si[d][p]); hi[d] = std::max(hi[d], si[d][p]); }
  }
  make_grid(o, lo, hi);
  o.same_sets = true;
  o.tree = true;
  o.ni = o.nj = n;
  if (!tree_build(o, n, si, cap, dmax)) return 1;
  o.perm_j = o.perm_i;
  o.jsta = o.ista;
  o.jend = o.iend;
  for (int d = 0; d < 3; ++d) o.xj[d] = o.xi[d];
  build_csr(o, [](const Oracle& oo, long long l, auto&& e) {
    tree_list_of(oo, l, oo.nb_buffer, tree_blen(oo, oo.ldepth[l]), e); }, o.bptr, o.bidx);
  build_csr(o, [](const Oracle& oo, long long l, auto&& e) {
    tree_list_of(oo, l, oo.nb_trunc, tree_tlen(oo, oo.ldepth[l]), e); }, o.tptr, o.tidx);
  o.rasm_ready = false;
  return 0;
}

// Leaf geometry of the adaptive structure: centres, depths, level-0 cells,
// buffer and truncation half-widths.
void orc_get_leaves(void* h, double* xc, double* yc, double* zc, int* depth, long long* cell0, double* blen,
                    double* tlen) {
  Oracle& o = *static_cast<Oracle*>(h);
  for (long long l = 0; l < o.ncell; ++l) {
    xc[l] = o.lctr[0][l];
    yc[l] = o.lctr[1][l];
    if (zc) zc[l] = o.lctr[2][l];
    depth[l] = o.ldepth[l];
    cell0[l] = o.lcell0[l];
    blen[l] = tree_blen(o, o.ldepth[l]);
    tlen[l] = tree_tlen(o, o.ldepth[l]);
  }
}

// ints[0..7] = nx, ny, nz, nb_buffer, nb_trunc, (ncell lo 31 bits) ...; use orc_ncell for counts.
long long orc_ncell(void* h) { return static_cast<Oracle*>(h)->ncell; }
long long orc_buffer_nnz(void* h) { auto& o = *static_cast<Oracle*>(h); return o.bptr.empty() ? 0 : o.bptr.back(); }
long long orc_trunc_nnz(void* h) { auto& o = *static_cast<Oracle*>(h); return o.tptr.empty() ? 0 : o.tptr.back(); }

void orc_get_grid(void* h, int* ints, double* reals) {
  Oracle& o = *static_cast<Oracle*>(h);
  ints[0] = o.nc[0]; ints[1] = o.nc[1]; ints[2] = o.nc[2];
  ints[3] = o.nb_buffer; ints[4] = o.nb_trunc;
  for (int d = 0; d < 3; ++d) reals[d] = o.cmin[d];
  reals[3] = o.box;
}

void orc_get_centres(void* h, double* xc, double* yc, double* zc) {
  Oracle& o = *static_cast<Oracle*>(h);
  const std::vector<double>* c = o.tree ? o.lctr : o.ctr;  // cluster centres: leaves or cells
  std::memcpy(xc, c[0].data(), sizeof(double) * o.ncell);
  std::memcpy(yc, c[1].data(), sizeof(double) * o.ncell);
  if (zc) std::memcpy(zc, c[2].data(), sizeof(double) * o.ncell);
}

void orc_get_perm(void* h, int* perm_i, int* perm_j) {
  Oracle& o = *static_cast<Oracle*>(h);
  std::memcpy(perm_i, o.perm_i.data(), sizeof(int) * o.ni);
  if (perm_j) std::memcpy(perm_j, o.perm_j.data(), sizeof(int) * o.nj);
}

void orc_get_cells(void* h, int* ista, int* iend, int* jsta, int* jend) {
  Oracle& o = *static_cast<Oracle*>(h);
  std::memcpy(ista, o.ista.data(), sizeof(int) * o.ncell);
  std::memcpy(iend, o.iend.data(), sizeof(int) * o.ncell);
  if (jsta) std::memcpy(jsta, o.jsta.data(), sizeof(int) * o.ncell);
  if (jend) std::memcpy(jend, o.jend.data(), sizeof(int) * o.ncell);
}

void orc_get_buffer(void* h, long long* ptr, int* idx) {
  Oracle& o = *static_cast<Oracle*>(h);
  std::memcpy(ptr, o.bptr.data(), sizeof(long long) * (o.ncell + 1));
  std::memcpy(idx, o.bidx.data(), sizeof(int) * o.bptr.back());
}

void orc_get_trunc(void* h, long long* ptr, int* idx) {
  Oracle& o = *static_cast<Oracle*>(h);
  std::memcpy(ptr, o.tptr.data(), sizeof(long long) * (o.ncell + 1));
  std::memcpy(idx, o.tidx.data(), sizeof(int) * o.tptr.back());
}

// Sorted coordinates (i-set), for callers that work in cluster order.
void orc_get_sorted(void* h, double* x, double* y, double* z) {
  Oracle& o = *static_cast<Oracle*>(h);
  std::memcpy(x, o.xi[0].data(), sizeof(double) * o.ni);
  std::memcpy(y, o.xi[1].data(), sizeof(double) * o.ni);
  if (z && o.dim == 3) std::memcpy(z, o.xi[2].data(), sizeof(double) * o.ni);
}

// MatShell MULT equivalent in sorted order: w[i-set] = A~ g[j-set]; nrhs
// vectors of lengths nj / ni stored back to back.
int orc_matvec(void* h, int nrhs, const double* g_sorted, double* w_sorted) {
  Oracle& o = *static_cast<Oracle*>(h);
  matvec(o, nrhs, g_sorted, o.nj, w_sorted, o.ni);
  return 0;
}

int orc_rasm_setup(void* h) {
  Oracle& o = *static_cast<Oracle*>(h);
  if (!o.same_sets) return 4;
  return rasm_setup(o);
}

// Timing sample: factorise (and later apply) only every stride-th cluster.
int orc_rasm_setup_sample(void* h, long long stride) {
  Oracle& o = *static_cast<Oracle*>(h);
  if (!o.same_sets || stride < 1) return 4;
  return rasm_setup(o, stride);
}

int orc_rasm_apply_sample(void* h, const double* r, double* z) {
  Oracle& o = *static_cast<Oracle*>(h);
  if (o.fptr.empty()) return 4;
  rasm_apply(o, r, z);
  return 0;
}

int orc_rasm_apply(void* h, const double* r, double* z) {
  Oracle& o = *static_cast<Oracle*>(h);
  if (!o.rasm_ready) return 4;
  rasm_apply(o, r, z);
  return 0;
}

// Dense block of cluster c (untruncated Gaussian over its buffer), row-major.
int orc_get_block(void* h, long long c, double* A) {
  Oracle& o = *static_cast<Oracle*>(h);
  block_of(o, c, A);
  return 0;
}

// LU factor (column-major, LAPACK dgetrf pivots 0-based) of block c.
int orc_get_lu(void* h, long long c, double* LU, int* ipiv) {
  Oracle& o = *static_cast<Oracle*>(h);
  if (!o.rasm_ready) return 4;
  long long n = o.bptr[c + 1] - o.bptr[c];
  std::memcpy(LU, o.fac.data() + o.fptr[c], sizeof(double) * n * n);
  std::memcpy(ipiv, o.piv.data() + o.bptr[c], sizeof(int) * n);
  return 0;
}

// Dense helpers exposed for known-answer tests of the factorisation.
int orc_dense_lu(int n, double* a_colmajor, int* ipiv) { return lu_factor(n, a_colmajor, ipiv) ? 0 : 5; }
void orc_dense_lu_solve(int n, const double* a, const int* ipiv, double* b) { lu_solve(n, a, ipiv, b); }

// GMRES on a small dense operator (column-major A, optional dense left PC B):
// known-answer tests of the Krylov solver alone.
int orc_dense_gmres(int n, const double* A, const double* B, const double* b, double* x,
                    double rtol, int restart, int max_it, int flexible,
                    int* its, double* hist, int hist_cap, int* nhist) {
  auto mv = [&](const double* M, const double* v, double* y) {
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int j = 0; j < n; ++j) s += M[(long long)j * n + i] * v[j];
      y[i] = s;
    }
  };
  auto opA = [&](const double* v, double* y) { mv(A, v, y); };
  auto opB = [&](const double* v, double* y) {
    if (B) mv(B, v, y); else std::memcpy(y, v, sizeof(double) * n);
  };
  GmresOut r = flexible ? fgmres_right(n, opA, opB, b, x, rtol, restart, max_it)
                        : gmres_left(n, opA, opB, b, x, rtol, restart, max_it);
  *its = r.its;
  int k = std::min<int>(hist_cap, (int)r.hist.size());
  for (int i = 0; i < k; ++i) hist[i] = r.hist[i];
  *nhist = (int)r.hist.size();
  return r.reason;
}

// Full plain solve in sorted order: GMRES(restart), left RASM PC, zero guess.
// res[0] = final residual estimate, res[1] = ||B b||.  Returns the PETSc-style
// converged reason (>0 converged).
int orc_solve_sorted(void* h, const double* w_sorted, double* g_sorted, double rtol, int restart,
                     int max_it, int flexible, int* its, double* hist, int hist_cap, int* nhist) {
  Oracle& o = *static_cast<Oracle*>(h);
  if (!o.rasm_ready) {
    int st = orc_rasm_setup(h);
    if (st) return -100 - st;
  }
  long long n = o.ni;
  auto opA = [&](const double* v, double* y) { matvec(o, 1, v, o.nj, y, o.ni); };
  auto opB = [&](const double* v, double* y) { rasm_apply(o, v, y); };
  GmresOut r = flexible ? fgmres_right(n, opA, opB, w_sorted, g_sorted, rtol, restart, max_it)
                        : gmres_left(n, opA, opB, w_sorted, g_sorted, rtol, restart, max_it);
  *its = r.its;
  int k = std::min<int>(hist_cap, (int)r.hist.size());
  for (int i = 0; i < k; ++i) hist[i] = r.hist[i];
  *nhist = (int)r.hist.size();
  return r.reason;
}

}  // extern "C"
