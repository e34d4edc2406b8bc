// C ABI (include/petrbf.h): context, setup, matvec, RASM, Krylov drivers.
//
// Host side of the drop-in boundary.  The context plays the role of the
// reference's BOTH{PARTICLE, CLUSTER} (2d/par.h:148-151) handed to the PETSc
// MatShell callbacks; here it owns the device copies of the sorted particles,
// cell ranges, buffer / truncation lists and RASM core rows.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/petrbf.h"
#include "common.h"
#include "kernels.h"

using namespace prbf;

namespace {

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n && p) return;
    release();
    size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    hipError_t e = hipMalloc((void**)&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw ApiError(PRBF_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    n = std::max<size_t>(1, count);
  }
  operator T*() const { return p; }
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// The learned preconditioner's network: FP32 master parameters with their
// gradient and Adam moments, BF16 weight copies in both orientations, and the
// activations of the largest batch run so far.
struct prbf_mlp {
  LGeo q = lgeo(2);  // patch geometry and layer sizes (2D or 3D network)
  hipStream_t st = nullptr;
  std::string err;
  DBuf<float> P, G, m1, m2;
  DBuf<unsigned short> Wb, Wt;
  double out_scale = 1.0;
  int t = 0;  // Adam steps taken
  long long cap = 0;
  DBuf<unsigned short> Xb, act, dYb, dHb;
  DBuf<float> pre, Y, Tb, dYf, dHf, part, loss, lhist;
  DBuf<long long> idx;
  void ensure(long long m) {
    if (m <= cap) return;
    Xb.ensure((size_t)m * q.nin);
    act.ensure((size_t)3 * m * kMlpHidden);
    pre.ensure((size_t)3 * m * kMlpHidden);
    Y.ensure((size_t)m * q.nout);
    Tb.ensure((size_t)m * q.nout);
    dYf.ensure((size_t)m * q.nout);
    dYb.ensure((size_t)m * q.nout);
    dHf.ensure((size_t)2 * m * kMlpHidden);
    dHb.ensure((size_t)2 * m * kMlpHidden);
    part.ensure(256);
    loss.ensure(1);
    idx.ensure((size_t)m);
    cap = m;
  }
  void layers(float* pr[3], unsigned short* ac[3]) {
    for (int l = 0; l < 3; ++l) {
      pr[l] = pre.p + (long long)l * cap * kMlpHidden;
      ac[l] = act.p + (long long)l * cap * kMlpHidden;
    }
  }
};

struct prbf_ctx {
  prbf_params prm{};
  Grid g{};
  hipStream_t st = nullptr;
  bool own_stream = false;
  std::string err;
  bool has_points = false, same = true;
  long long ni = 0, nj = 0;
  // this context's device's copy of the static scratch array (allocation-free
  // evaluation), resolved at prbf_create
  double* scratch = nullptr;
  // clusters: the grid's cells, or the leaves of the adaptive structure
  // (tree), with the leaves' geometry and the grid cells' particle ranges
  long long nclus = 0;
  bool tree = false;
  int tree_cap = 0, tree_dmax = 0;
  DBuf<int> lcell, c0sta, c0end;
  DBuf<double> lcx, lcy, lcz, lbl, ltl;
  std::vector<int> h_ldepth;
  // sorted particles and cell structure
  DBuf<double> xi, yi, zi, xj, yj, zj;
  DBuf<int> perm_i, perm_j, ista, iend, jsta, jend;
  DBuf<long long> bptr, tptr;
  DBuf<int> bidx, tidx;
  long long buffer_nnz = 0, trunc_nnz = 0;
  int max_buffer = 0, max_core = 0, max_trunc = 0;
  std::vector<long long> h_bptr;
  std::vector<int> h_ista, h_iend;
  // RASM: core rows of the FP32 blocks (M32, offsets mptr) and of the FP64
  // blocks (M, offsets mptr64); per cluster the precision of its block
  // (h_prec: 0 none, 1 FP64, 2 FP32) and, per precision, the list of the
  // clusters the apply serves (null list: every cluster)
  bool rasm_ready = false;
  DBuf<long long> mptr, mptr64;
  DBuf<double> M;
  std::vector<long long> h_mptr, h_mptr64;
  std::vector<unsigned char> h_prec;
  DBuf<int> l32, l64;
  long long n32 = 0, n64 = 0;
  bool l32_all = false, l64_all = false;
  // scratch
  DBuf<int> i_scr1, i_scr2, i_scr3, i_scr4, d_err;
  DBuf<long long> l_scr;
  DBuf<double> part, scal, v1, v2, v3;
  double* pin = nullptr;  // pinned host scalars
  // Krylov
  DBuf<double> V, Z;
  int chol_waves = 8;  // wavefronts per block-factorisation workgroup (when forced)
  int chol_panel = 2;  // tile columns per panel (when forced)
  int chol_order = PRBF_CHOL_RIGHT;  // right- or left-looking (when forced)
  bool tuning_forced = false;
  int used_nw = 0, used_npc = 0, used_order = 0;  // configuration of the last factorisation
  double t_chol_s = 0, chol_flops = 0;  // its kernel time and algorithmic FLOPs
  int rasm_pc = 0;
  int n_regularised = 0;
  DBuf<float> M32;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // block factorisation: work list, counter and workspace slots
  DBuf<int> fb_order, fb_next, fb_status, fb_failed;
  DBuf<unsigned char> fb_ws;
  // Arnoldi steps captured as graphs: device state of the cycle, one
  // executable graph per step index, and the launch arguments they captured
  int last_restart = 0;  // basis size of the last Krylov solve
  // learned preconditioner: patches of the eligible clusters, the attached
  // network, the clusters it serves (h_skip[c] = 1) and inference buffers
  bool patches_ready = false, learned_ready = false;
  DBuf<int> pidx, cidx, elig, mlp_list;
  DBuf<double> pctr;
  DBuf<unsigned char> pok;
  long long n_elig = 0, n_mlp = 0;
  int n_net_fallback = 0;
  std::vector<unsigned char> h_skip;
  std::vector<int> h_mlp_list;
  DBuf<double> lnorm;
  prbf_mlp* net = nullptr;
  DBuf<float> lX;
  DBuf<double> ls;
  DBuf<double> kst;
  std::vector<hipGraphExec_t> kgraphs;
  std::vector<unsigned long long> kgraph_key;
  void drop_graphs() {
    for (hipGraphExec_t e : kgraphs) (void)hipGraphExecDestroy(e);
    kgraphs.clear();
    kgraph_key.clear();
  }

  const double* xj_() const { return same ? xi.p : xj.p; }
  const double* yj_() const { return same ? yi.p : yj.p; }
  const double* zj_() const { return same ? zi.p : zj.p; }
  const int* jsta_() const { return same ? ista.p : jsta.p; }
  const int* jend_() const { return same ? iend.p : jend.p; }
  const int* perm_j_() const { return same ? perm_i.p : perm_j.p; }
};

namespace {

prbf_status fail(prbf_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return (prbf_status)code;
}

template <class F>
prbf_status guarded(prbf_ctx* c, F&& f) {
  if (!c) return PRBF_ERR_ARG;
  try {
    f();
    c->err.clear();
    return PRBF_OK;
  } catch (const ApiError& e) {
    return fail(c, e.code, e.what);
  } catch (const HipError& e) {
    return fail(c, PRBF_ERR_HIP, e.what);
  } catch (const std::bad_alloc&) {
    return fail(c, PRBF_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(c, PRBF_ERR_ARG, e.what());
  } catch (...) {
    return fail(c, PRBF_ERR_HIP, "unexpected error");
  }
}

void sync(prbf_ctx* c) { PRBF_HIP(hipStreamSynchronize(c->st)); }

void need(bool ok, int code, const char* msg) {
  if (!ok) throw ApiError(code, msg);
}

// Grid parameters from the bounding box (2d/get_cluster.h:20-34), evaluated
// on the host in the reference's operation order.
Grid make_grid(const prbf_params& p, const double* lo, const double* hi) {
#pragma clang fp contract(off)
  Grid g{};
  g.dim = p.dim;
  double nb = p.nsigma_box;
  g.nb_buffer = (int)ceil((p.sigma_buffer - nb + kEpsf) / 2 / nb);
  g.nb_trunc = (int)ceil((p.sigma_trunc - nb + kEpsf) / 2 / nb);
  g.box = p.nsigma_box * p.sigma + kEpsf;
  double mn[3] = {0, 0, 0};
  int nc[3] = {1, 1, 1};
  for (int d = 0; d < p.dim; ++d) {
    mn[d] = lo[d] - kEpsf;
    double mx = hi[d] + kEpsf;
    nc[d] = (int)ceil((mx - mn[d]) / g.box);
  }
  g.xmin = mn[0];
  g.ymin = mn[1];
  g.zmin = mn[2];
  g.nx = nc[0];
  g.ny = nc[1];
  g.nz = nc[2];
  g.ncell = (long long)nc[0] * nc[1] * nc[2];
  g.buffer_len = p.sigma_buffer * p.sigma / 2 + kEpsf;
  g.trunc_len = p.sigma_trunc * p.sigma / 2 + kEpsf;
  return g;
}

void bbox_of(prbf_ctx* c, long long n, const double* x, const double* y, const double* z, double* lo, double* hi) {
  c->part.ensure(6 * 256);
  c->scal.ensure(128);
  bbox(c->st, n, x, y, z, c->prm.dim, c->part, c->scal);
  PRBF_HIP(hipMemcpyAsync(c->pin, c->scal, 6 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  sync(c);
  for (int d = 0; d < 3; ++d) {
    lo[d] = c->pin[d];
    hi[d] = c->pin[3 + d];
  }
}

void build_lists(prbf_ctx* c, bool buffer) {
  const Grid& g = c->g;
  int reach = buffer ? g.nb_buffer : g.nb_trunc;
  double len = buffer ? g.buffer_len : g.trunc_len;
  const int* sta = buffer ? c->ista.p : c->jsta_();
  const int* end = buffer ? c->iend.p : c->jend_();
  const double* xs = buffer ? c->xi.p : c->xj_();
  const double* ys = buffer ? c->yi.p : c->yj_();
  const double* zs = buffer ? c->zi.p : c->zj_();
  DBuf<long long>& ptr = buffer ? c->bptr : c->tptr;
  DBuf<int>& idx = buffer ? c->bidx : c->tidx;
  c->i_scr1.ensure(g.ncell);
  ptr.ensure(g.ncell + 1);
  list_count(c->st, g, reach, len, c->ista, c->iend, sta, end, xs, ys, zs, c->i_scr1);
  scan_excl(c->st, g.ncell, c->i_scr1, ptr);
  long long total = 0;
  PRBF_HIP(hipMemcpyAsync(&total, ptr.p + g.ncell, sizeof(long long), hipMemcpyDeviceToHost, c->st));
  sync(c);
  idx.ensure(total);
  list_fill(c->st, g, reach, len, c->ista, c->iend, sta, end, xs, ys, zs, ptr, idx);
  (buffer ? c->buffer_nnz : c->trunc_nnz) = total;
}

// Buffer and truncation lists of one particle set in a single scan.
void build_lists_same(prbf_ctx* c) {
  const Grid& g = c->g;
  c->i_scr1.ensure(g.ncell);
  c->i_scr2.ensure(g.ncell);
  c->bptr.ensure(g.ncell + 1);
  c->tptr.ensure(g.ncell + 1);
  lists2_count(c->st, g, c->ista, c->iend, c->xi, c->yi, c->zi, c->i_scr1, c->i_scr2);
  scan_excl(c->st, g.ncell, c->i_scr1, c->bptr);
  scan_excl(c->st, g.ncell, c->i_scr2, c->tptr);
  long long tot[2] = {0, 0};
  PRBF_HIP(hipMemcpyAsync(&tot[0], c->bptr.p + g.ncell, sizeof(long long), hipMemcpyDeviceToHost, c->st));
  PRBF_HIP(hipMemcpyAsync(&tot[1], c->tptr.p + g.ncell, sizeof(long long), hipMemcpyDeviceToHost, c->st));
  sync(c);
  c->bidx.ensure(tot[0]);
  c->tidx.ensure(tot[1]);
  lists2_fill(c->st, g, c->ista, c->iend, c->xi, c->yi, c->zi, c->bptr, c->bidx, c->tptr, c->tidx);
  c->buffer_nnz = tot[0];
  c->trunc_nnz = tot[1];
}

// Host copies of the cluster structure (list pointers, ranges) for chunking
// and size limits.
void host_cluster_copies(prbf_ctx* c) {
  const long long nc = c->nclus;
  c->h_bptr.resize(nc + 1);
  c->h_ista.resize(nc);
  c->h_iend.resize(nc);
  std::vector<long long> h_tptr(nc + 1);
  PRBF_HIP(hipMemcpyAsync(c->h_bptr.data(), c->bptr, sizeof(long long) * (nc + 1), hipMemcpyDeviceToHost, c->st));
  PRBF_HIP(hipMemcpyAsync(h_tptr.data(), c->tptr, sizeof(long long) * (nc + 1), hipMemcpyDeviceToHost, c->st));
  PRBF_HIP(hipMemcpyAsync(c->h_ista.data(), c->ista, sizeof(int) * nc, hipMemcpyDeviceToHost, c->st));
  PRBF_HIP(hipMemcpyAsync(c->h_iend.data(), c->iend, sizeof(int) * nc, hipMemcpyDeviceToHost, c->st));
  sync(c);
  c->max_buffer = c->max_core = c->max_trunc = 0;
  for (long long k = 0; k < nc; ++k) {
    c->max_buffer = std::max<int>(c->max_buffer, (int)(c->h_bptr[k + 1] - c->h_bptr[k]));
    c->max_trunc = std::max<int>(c->max_trunc, (int)(h_tptr[k + 1] - h_tptr[k]));
    c->max_core = std::max(c->max_core, c->h_iend[k] - c->h_ista[k] + 1);
  }
}

void set_points_impl(prbf_ctx* c, long long ni, const double* xi, const double* yi, const double* zi,
                     long long nj, const double* xj, const double* yj, const double* zj) {
  const int dim = c->prm.dim;
  need(ni > 0, PRBF_ERR_ARG, "ni must be positive");
  need(xi && yi && (dim == 2 || zi), PRBF_ERR_ARG, "missing target coordinates");
  need(ni < (1LL << 31), PRBF_ERR_ARG, "ni exceeds int32 indexing");
  bool same = (xj == nullptr);
  if (!same) {
    need(nj > 0 && yj && (dim == 2 || zj), PRBF_ERR_ARG, "missing source coordinates");
    need(nj < (1LL << 31), PRBF_ERR_ARG, "nj exceeds int32 indexing");
  } else {
    nj = ni;
  }
  c->has_points = false;
  c->rasm_ready = false;
  c->patches_ready = false;
  c->learned_ready = false;
  c->tree = false;
  double lo[3], hi[3], lo2[3], hi2[3];
  bbox_of(c, ni, xi, yi, zi, lo, hi);
  if (!same) {
    bbox_of(c, nj, xj, yj, zj, lo2, hi2);
    for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], lo2[d]); hi[d] = std::max(hi[d], hi2[d]); }
  }
  for (int d = 0; d < dim; ++d) need(std::isfinite(lo[d]) && std::isfinite(hi[d]), PRBF_ERR_ARG, "non-finite coordinate");
  c->g = make_grid(c->prm, lo, hi);
  const Grid& g = c->g;
  need(g.ncell > 0 && g.ncell < (1LL << 31), PRBF_ERR_LIMIT, "cell grid too large");
  c->same = same;
  c->ni = ni;
  c->nj = nj;
  c->d_err.ensure(1);
  PRBF_HIP(hipMemsetAsync(c->d_err, 0, sizeof(int), c->st));
  c->l_scr.ensure(g.ncell + 1);
  c->i_scr1.ensure(std::max<long long>(ni, nj));
  c->i_scr4.ensure(std::max<long long>(ni, nj));
  c->i_scr2.ensure(g.ncell);
  c->i_scr3.ensure(g.ncell);
  auto sort_set = [&](long long n, const double* x, const double* y, const double* z, DBuf<double>& xs,
                      DBuf<double>& ys, DBuf<double>& zs, DBuf<int>& perm, DBuf<int>& sta, DBuf<int>& end) {
    xs.ensure(n);
    ys.ensure(n);
    if (dim == 3) zs.ensure(n);
    perm.ensure(n);
    sta.ensure(g.ncell);
    end.ensure(g.ncell);
    sort_points(c->st, g, n, x, y, dim == 3 ? z : nullptr, c->i_scr1, c->i_scr2, c->l_scr, c->i_scr3, c->i_scr4,
                perm, sta, end, xs, ys, dim == 3 ? zs.p : nullptr, c->d_err);
  };
  sort_set(ni, xi, yi, zi, c->xi, c->yi, c->zi, c->perm_i, c->ista, c->iend);
  if (!same) sort_set(nj, xj, yj, zj, c->xj, c->yj, c->zj, c->perm_j, c->jsta, c->jend);
  int herr = 0;
  PRBF_HIP(hipMemcpyAsync(&herr, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->st));
  sync(c);
  need(herr == 0, PRBF_ERR_OUTSIDE, "a point fell outside the cell grid");
  if (same) {
    build_lists_same(c);
  } else {
    build_lists(c, true);
    build_lists(c, false);
  }
  c->nclus = g.ncell;
  host_cluster_copies(c);
  c->has_points = true;
}

// Half-widths of an adaptive leaf's buffer and truncation boxes at depth d:
// the buffer keeps its shape relative to the cluster (buffer_len 2^-d), the
// truncation keeps the cutoff beyond the cluster ((trunc_len - box/2) +
// box 2^-(d+1); trunc_len itself at depth 0).
double tree_blen(const Grid& g, int d) { return ldexp(g.buffer_len, -d); }
double tree_tlen(const Grid& g, int d) {
#pragma clang fp contract(off)
  return d == 0 ? g.trunc_len : (g.trunc_len - g.box * 0.5) + ldexp(g.box, -(d + 1));
}

// The adaptive structure over one point set (the solve): the reference's grid
// as level 0, cells of more than cap particles split into 2^dim children down
// to depth dmax, the leaves as clusters (rules in tree.hip).  The frontier's
// node counts come to the host level by level; the host keeps the node
// records, decides the splits, numbers the children and gives their centres.
void set_points_tree_impl(prbf_ctx* c, long long n, const double* x, const double* y, const double* z, int cap,
                          int dmax) {
  const int dim = c->prm.dim;
  need(n > 0, PRBF_ERR_ARG, "n must be positive");
  need(x && y && (dim == 2 || z), PRBF_ERR_ARG, "missing coordinates");
  need(n < (1LL << 31), PRBF_ERR_ARG, "n exceeds int32 indexing");
  need(cap >= 1 && dmax >= 0 && dmax <= 20, PRBF_ERR_ARG, "cap must be positive and max_depth in 0..20");
  c->has_points = false;
  c->rasm_ready = false;
  c->patches_ready = false;
  c->learned_ready = false;
  c->tree = false;
  double lo[3], hi[3];
  bbox_of(c, n, x, y, z, lo, hi);
  for (int d = 0; d < dim; ++d) need(std::isfinite(lo[d]) && std::isfinite(hi[d]), PRBF_ERR_ARG, "non-finite coordinate");
  c->g = make_grid(c->prm, lo, hi);
  const Grid& g = c->g;
  need(g.ncell > 0 && g.ncell < (1LL << 31), PRBF_ERR_LIMIT, "cell grid too large");
  c->same = true;
  c->ni = c->nj = n;
  const long long nc0 = g.ncell;
  // level 0: the grid cells (node v < nc0 is cell v)
  c->d_err.ensure(1);
  PRBF_HIP(hipMemsetAsync(c->d_err, 0, sizeof(int), c->st));
  DBuf<int> node, cnt;
  node.ensure(n);
  cnt.ensure(nc0);
  cell_assign(c->st, g, n, x, y, z, node, cnt, c->d_err);
  struct Node {
    double c[3];
    int depth;
    int child;  // first child's node id, -1 for a leaf
    long long cell0;
  };
  std::vector<Node> nodes(nc0);
  for (long long k = 0; k < nc0; ++k) {
    int ix, iy, iz;
    cell_coords(g, k, &ix, &iy, &iz);
    nodes[k] = {{cell_centre(g.xmin, ix, g.box), cell_centre(g.ymin, iy, g.box),
                 dim == 3 ? cell_centre(g.zmin, iz, g.box) : 0.0},
                0, -1, k};
  }
  std::vector<int> hcnt(nc0);
  PRBF_HIP(hipMemcpyAsync(hcnt.data(), cnt, sizeof(int) * nc0, hipMemcpyDeviceToHost, c->st));
  int herr = 0;
  PRBF_HIP(hipMemcpyAsync(&herr, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->st));
  sync(c);
  need(herr == 0, PRBF_ERR_OUTSIDE, "a point fell outside the cell grid");
  const int nch = 1 << dim;
  long long lo_f = 0, hi_f = nc0;  // the frontier: nodes [lo_f, hi_f)
  DBuf<int> cbase;
  DBuf<double> fcx, fcy, fcz;
  for (int d = 0; d < dmax; ++d) {
    const long long nf = hi_f - lo_f;
    std::vector<int> hb(nf, -1);
    std::vector<double> hx(nf), hy(nf), hz(nf);
    bool any = false;
    for (long long f = 0; f < nf; ++f) {
      const Node v = nodes[lo_f + f];  // (a copy: the children's push_back may move the records)
      hx[f] = v.c[0];
      hy[f] = v.c[1];
      hz[f] = v.c[2];
      if (hcnt[f] <= cap) continue;
      need((long long)nodes.size() + nch < (1LL << 31), PRBF_ERR_LIMIT, "adaptive structure too large");
      nodes[lo_f + f].child = (int)nodes.size();
      hb[f] = (int)nodes.size();
      any = true;
      const double q = ldexp(g.box, -(d + 2));
      for (int k = 0; k < nch; ++k) {
        const int qx = (k >> 1) & 1, qy = k & 1, qz = (k >> 2) & 1;
        nodes.push_back({{v.c[0] + (qx ? q : -q), v.c[1] + (qy ? q : -q), dim == 3 ? v.c[2] + (qz ? q : -q) : 0.0},
                         d + 1, -1, v.cell0});
      }
    }
    if (!any) break;
    cbase.ensure(nf);
    fcx.ensure(nf);
    fcy.ensure(nf);
    fcz.ensure(nf);
    PRBF_HIP(hipMemcpyAsync(cbase, hb.data(), sizeof(int) * nf, hipMemcpyHostToDevice, c->st));
    PRBF_HIP(hipMemcpyAsync(fcx, hx.data(), sizeof(double) * nf, hipMemcpyHostToDevice, c->st));
    PRBF_HIP(hipMemcpyAsync(fcy, hy.data(), sizeof(double) * nf, hipMemcpyHostToDevice, c->st));
    PRBF_HIP(hipMemcpyAsync(fcz, hz.data(), sizeof(double) * nf, hipMemcpyHostToDevice, c->st));
    tree_descend(c->st, dim, n, x, y, z, node, (int)lo_f, (int)hi_f, cbase, fcx, fcy, fcz);
    lo_f = hi_f;
    hi_f = (long long)nodes.size();
    cnt.ensure(hi_f - lo_f);
    tree_count(c->st, n, node, (int)lo_f, (int)hi_f, cnt);
    PRBF_HIP(hipGetLastError());
    hcnt.assign(hi_f - lo_f, 0);
    PRBF_HIP(hipMemcpyAsync(hcnt.data(), cnt, sizeof(int) * (hi_f - lo_f), hipMemcpyDeviceToHost, c->st));
    sync(c);  // the host vectors of this level are released on return
  }
  // leaves in depth-first order, children in index order
  std::vector<int> leaf_of(nodes.size(), -1), first_leaf(nc0 + 1, 0);
  std::vector<long long> order;
  order.reserve(nodes.size());
  {
    std::vector<long long> stack;
    for (long long k = 0; k < nc0; ++k) {
      first_leaf[k] = (int)order.size();
      stack.push_back(k);
      while (!stack.empty()) {
        const long long v = stack.back();
        stack.pop_back();
        if (nodes[v].child < 0) {
          leaf_of[v] = (int)order.size();
          order.push_back(v);
        } else {
          for (int j = nch - 1; j >= 0; --j) stack.push_back(nodes[v].child + j);
        }
      }
    }
    first_leaf[nc0] = (int)order.size();
  }
  const long long nl = (long long)order.size();
  c->nclus = nl;
  // particles sorted by leaf (stable in the original index)
  DBuf<int> lof;
  lof.ensure(nodes.size());
  PRBF_HIP(hipMemcpyAsync(lof, leaf_of.data(), sizeof(int) * nodes.size(), hipMemcpyHostToDevice, c->st));
  c->i_scr1.ensure(n);
  c->i_scr2.ensure(nl);
  c->i_scr3.ensure(nl);
  c->i_scr4.ensure(n);
  c->l_scr.ensure(nl + 1);
  tree_leaf_keys(c->st, n, nl, node, lof, c->i_scr1, c->i_scr2);
  c->xi.ensure(n);
  c->yi.ensure(n);
  if (dim == 3) c->zi.ensure(n);
  c->perm_i.ensure(n);
  c->ista.ensure(nl);
  c->iend.ensure(nl);
  sort_by_key(c->st, nl, n, c->i_scr1, c->i_scr2, c->l_scr, c->i_scr3, c->i_scr4, c->perm_i, c->ista, c->iend, x, y,
              dim == 3 ? z : nullptr, c->xi, c->yi, dim == 3 ? c->zi.p : nullptr);
  // leaf geometry and the grid cells' particle ranges
  std::vector<long long> start(nl + 1);
  PRBF_HIP(hipMemcpyAsync(start.data(), c->l_scr, sizeof(long long) * (nl + 1), hipMemcpyDeviceToHost, c->st));
  sync(c);
  std::vector<int> hl0(nl), s0(nc0), e0(nc0);
  std::vector<double> hx(nl), hy(nl), hz(nl), hbl(nl), htl(nl);
  c->h_ldepth.resize(nl);
  for (long long l = 0; l < nl; ++l) {
    const Node& v = nodes[order[l]];
    hl0[l] = (int)v.cell0;
    hx[l] = v.c[0];
    hy[l] = v.c[1];
    hz[l] = v.c[2];
    hbl[l] = tree_blen(g, v.depth);
    htl[l] = tree_tlen(g, v.depth);
    c->h_ldepth[l] = v.depth;
  }
  for (long long k = 0; k < nc0; ++k) {
    s0[k] = (int)start[first_leaf[k]];
    e0[k] = (int)start[first_leaf[k + 1]] - 1;
  }
  c->lcell.ensure(nl);
  c->lcx.ensure(nl);
  c->lcy.ensure(nl);
  c->lcz.ensure(nl);
  c->lbl.ensure(nl);
  c->ltl.ensure(nl);
  c->c0sta.ensure(nc0);
  c->c0end.ensure(nc0);
  PRBF_HIP(hipMemcpyAsync(c->lcell, hl0.data(), sizeof(int) * nl, hipMemcpyHostToDevice, c->st));
  PRBF_HIP(hipMemcpyAsync(c->lcx, hx.data(), sizeof(double) * nl, hipMemcpyHostToDevice, c->st));
  PRBF_HIP(hipMemcpyAsync(c->lcy, hy.data(), sizeof(double) * nl, hipMemcpyHostToDevice, c->st));
  PRBF_HIP(hipMemcpyAsync(c->lcz, hz.data(), sizeof(double) * nl, hipMemcpyHostToDevice, c->st));
  PRBF_HIP(hipMemcpyAsync(c->lbl, hbl.data(), sizeof(double) * nl, hipMemcpyHostToDevice, c->st));
  PRBF_HIP(hipMemcpyAsync(c->ltl, htl.data(), sizeof(double) * nl, hipMemcpyHostToDevice, c->st));
  PRBF_HIP(hipMemcpyAsync(c->c0sta, s0.data(), sizeof(int) * nc0, hipMemcpyHostToDevice, c->st));
  PRBF_HIP(hipMemcpyAsync(c->c0end, e0.data(), sizeof(int) * nc0, hipMemcpyHostToDevice, c->st));
  // buffer and truncation lists of every leaf
  c->bptr.ensure(nl + 1);
  c->tptr.ensure(nl + 1);
  DBuf<int> cb, ct;
  cb.ensure(nl);
  ct.ensure(nl);
  tree_lists(c->st, g, false, nl, c->ista, c->iend, c->c0sta, c->c0end, c->lcell, c->lcx, c->lcy, c->lcz, c->lbl,
             c->ltl, c->xi, c->yi, c->zi, cb, ct, nullptr, nullptr, nullptr, nullptr);
  scan_excl(c->st, nl, cb, c->bptr);
  scan_excl(c->st, nl, ct, c->tptr);
  long long tot[2] = {0, 0};
  PRBF_HIP(hipMemcpyAsync(&tot[0], c->bptr.p + nl, sizeof(long long), hipMemcpyDeviceToHost, c->st));
  PRBF_HIP(hipMemcpyAsync(&tot[1], c->tptr.p + nl, sizeof(long long), hipMemcpyDeviceToHost, c->st));
  sync(c);
  c->bidx.ensure(tot[0]);
  c->tidx.ensure(tot[1]);
  tree_lists(c->st, g, true, nl, c->ista, c->iend, c->c0sta, c->c0end, c->lcell, c->lcx, c->lcy, c->lcz, c->lbl,
             c->ltl, c->xi, c->yi, c->zi, nullptr, nullptr, c->bptr, c->bidx, c->tptr, c->tidx);
  PRBF_HIP(hipGetLastError());
  c->buffer_nnz = tot[0];
  c->trunc_nnz = tot[1];
  host_cluster_copies(c);  // (synchronises: the host arrays above are released on return)
  c->tree = true;
  c->tree_cap = cap;
  c->tree_dmax = dmax;
  c->has_points = true;
}

void matvec_impl(prbf_ctx* c, int nrhs, const double* g, double* w) {
  need(c->has_points, PRBF_ERR_STATE, "set_points first");
  need(nrhs >= 1 && nrhs <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
  need(g && w, PRBF_ERR_ARG, "null vector");
  matvec(c->st, c->prm.dim, nrhs, c->nclus, c->ista, c->iend, c->tptr, c->tidx, c->xi, c->yi, c->zi,
         c->xj_(), c->yj_(), c->zj_(), g, c->nj, w, c->ni, c->prm.sigma);
  PRBF_HIP(hipGetLastError());
}

constexpr size_t kMaxLdsBytes = 160 * 1024;

struct CholCfg {
  int nw, npc;
  bool ll;  // left-looking kernel
};

// Launch configurations the tuner chooses among (wavefronts per workgroup,
// tile columns per panel, left-looking).
// (FP32 also 6 wavefronts and 3-column panels; a candidate without a kernel
// for the precision is skipped.)
const CholCfg kCholCfgs[] = {{4, 1, false}, {8, 1, false}, {4, 2, false}, {8, 2, false}, {4, 4, false},
                             {8, 4, false}, {4, 3, false}, {6, 1, false}, {6, 2, false}, {6, 3, false},
                             {6, 4, false}, {8, 3, false}, {4, 1, true},  {4, 2, true},  {8, 1, true},
                             {8, 2, true},  {4, 4, true},  {8, 4, true}};

// Process-wide cache of tuned configurations, keyed by device, precision and
// the largest block's tile count; shared by every context, so guarded (two
// threads tuning the same key at once both tune, and the first result stays).
std::mutex& tuned_mutex() {
  static std::mutex m;
  return m;
}
std::vector<std::pair<long long, CholCfg>>& tuned_cache() {
  static std::vector<std::pair<long long, CholCfg>> v;
  return v;
}

// Algorithmic FLOPs of the block setups of the clusters in `order`: per block
// of n unknowns with T core members, the Cholesky factorisation (n^3/3), the
// backward substitution of the T core right-hand sides through all n rows
// (n^2 T) and their forward substitution, which starts at the core because
// the core is ordered last (T^3: T right-hand sides over T rows).
double block_flops(const prbf_ctx* c, const std::vector<int>& order) {
  double fl = 0;
  for (int k : order) {
    const double nb = (double)(c->h_bptr[k + 1] - c->h_bptr[k]), T = c->h_iend[k] - c->h_ista[k] + 1;
    fl += nb * nb * nb / 3.0 + nb * nb * T + T * T * T;
  }
  return fl;
}

template <typename Real>
void launch_blocks(prbf_ctx* c, bool fp32, const CholCfg& cfg, long long nwork, const int* order, int* next, int nwg,
                   long long ws_slot, int nmax, int tmax, Real* W, const long long* mptr, Real* Mout, int* status) {
  const Grid& g = c->g;
  if (fp32)
    block_chol_fp32(c->st, g.dim, cfg.nw, cfg.npc, cfg.ll, nwork, order, next, nwg, ws_slot, nmax, tmax, c->bptr, c->bidx,
                    c->ista, c->iend, c->xi, c->yi, c->zi, (float*)W, mptr, (float*)Mout, c->prm.sigma, status);
  else
    block_chol_fp64(c->st, g.dim, cfg.nw, cfg.npc, cfg.ll, nwork, order, next, nwg, ws_slot, nmax, tmax, c->bptr, c->bidx,
                    c->ista, c->iend, c->xi, c->yi, c->zi, (double*)W, mptr, (double*)Mout, c->prm.sigma, status);
  PRBF_HIP(hipGetLastError());
}

// Factorises the blocks of the clusters in `order` (largest first) in one
// persistent launch, each resident workgroup reusing one workspace slot.
// fp32: blocks in FP32 on the FP32 MFMA, per-block status in `status`.  The
// launch configuration is forced, cached, or, on the first call for a
// (precision, block size), timed for every configuration that fits in LDS on
// the first min(12288, nwork) blocks of the list, keeping the fastest (the
// project's autotuner).
template <typename Real>
void factor_blocks(prbf_ctx* c, bool fp32, const std::vector<int>& order, int nmax, int tmax, const long long* mptr,
                   Real* Mout, int* status, bool tune_allowed) {
  const Grid& g = c->g;
  const long long nwork = (long long)order.size();
  if (nwork == 0) return;
  DBuf<int>& d_order = c->fb_order;
  DBuf<int>& next = c->fb_next;
  d_order.ensure(order.size());
  next.ensure(1);
  PRBF_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, c->st));
  const long long ws_slot = block_workspace(nmax, tmax);
  size_t free_b = 0, total_b = 0;
  PRBF_HIP(hipMemGetInfo(&free_b, &total_b));
  const long long cap = (long long)(free_b / 3 / sizeof(Real)) / ws_slot;
  need(cap >= 1, PRBF_ERR_NOMEM, "not enough device memory for one RASM block");
  auto workgroups = [&](const CholCfg& cf) {
    long long w = block_chol_workgroups(fp32, g.dim, cf.nw, cf.npc, cf.ll, nmax, tmax);
    return (int)std::max<long long>(1, std::min(std::min(w, nwork), cap));
  };
  const int ktmax = (nmax + 15) / 16;
  int dev = 0;
  PRBF_HIP(hipGetDevice(&dev));
  const long long key = dev * 10000000LL + (fp32 ? 1000000LL : 0LL) + ktmax * 1000LL + g.dim;
  CholCfg cfg{c->chol_waves, c->chol_panel, c->chol_order == PRBF_CHOL_LEFT};
  bool tune = tune_allowed && !c->tuning_forced;
  if (tune) {
    std::lock_guard<std::mutex> lock(tuned_mutex());
    for (auto& kv : tuned_cache())
      if (kv.first == key) {
        cfg = kv.second;
        tune = false;
      }
  }
  // the workspace slots, kept by the context from one setup to the next
  struct {
    DBuf<unsigned char>& raw;
    Real* p = nullptr;
    void ensure(size_t count) {
      raw.ensure(count * sizeof(Real));
      p = (Real*)raw.p;
    }
  } W{c->fb_ws};
  if (tune) {
    // ~16 blocks per resident workgroup: fewer let the launch's tail decide
    // (at 4096 the 4x2 and 8x4 FP32 configurations timed within 7% of each
    // other, against 12% over the full launch)
    const long long ns = std::min<long long>(12288, nwork);
    float best = 1e30f;
    for (const CholCfg& cand : kCholCfgs) {
      if (block_lds_bytes(fp32, nmax, tmax, cand.nw, cand.npc, cand.ll) > kMaxLdsBytes) continue;
      if (!block_chol_has_config(fp32, cand.nw, cand.npc, cand.ll)) continue;
      const int nwg = workgroups(cand);
      W.ensure((size_t)nwg * ws_slot);
      PRBF_HIP(hipEventRecord(c->ev0, c->st));
      launch_blocks<Real>(c, fp32, cand, ns, d_order, next, nwg, ws_slot, nmax, tmax, W.p, mptr, Mout, status);
      PRBF_HIP(hipEventRecord(c->ev1, c->st));
      PRBF_HIP(hipEventSynchronize(c->ev1));
      float ms = 0;
      PRBF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      if (ms < best) { best = ms; cfg = cand; }
    }
    std::lock_guard<std::mutex> lock(tuned_mutex());
    bool have = false;
    for (auto& kv : tuned_cache())
      if (kv.first == key) {
        cfg = kv.second;
        have = true;
      }
    if (!have) tuned_cache().push_back({key, cfg});
  }
  if (block_lds_bytes(fp32, nmax, tmax, cfg.nw, cfg.npc, cfg.ll) > kMaxLdsBytes ||
      !block_chol_has_config(fp32, cfg.nw, cfg.npc, cfg.ll))
    cfg = CholCfg{4, 1, false};
  need(block_lds_bytes(fp32, nmax, tmax, cfg.nw, cfg.npc, cfg.ll) <= kMaxLdsBytes, PRBF_ERR_LIMIT,
       fp32 ? "RASM block too large for the LDS panel of the block factorisation"
            : "RASM block too large for the LDS panel of the FP64 block factorisation");
  const int nwg = workgroups(cfg);
  W.ensure((size_t)nwg * ws_slot);
  PRBF_HIP(hipEventRecord(c->ev0, c->st));
  launch_blocks<Real>(c, fp32, cfg, nwork, d_order, next, nwg, ws_slot, nmax, tmax, W.p, mptr, Mout, status);
  PRBF_HIP(hipEventRecord(c->ev1, c->st));
  c->used_nw = cfg.nw;
  c->used_npc = cfg.npc;
  c->used_order = cfg.ll ? PRBF_CHOL_LEFT : PRBF_CHOL_RIGHT;
  sync(c);
  float ms = 0;
  PRBF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->t_chol_s = ms * 1e-3;
  c->chol_flops = block_flops(c, order);
}

// Blocks of fewer unknowns than this are factorised and applied in FP64 by
// the mixed-precision preconditioners; larger ones in FP32 on the matrix cores.
constexpr int kFp32MinUnknowns = 32;

// Largest block first (stable counting sort by block size): the persistent
// workgroups of the factorisation then finish together.
void largest_first(const prbf_ctx* c, std::vector<int>& order) {
  const int nb = c->max_buffer + 1;
  auto sz = [&](int k) { return (int)(c->h_bptr[k + 1] - c->h_bptr[k]); };
  std::vector<int> cnt(nb + 1, 0), sorted(order.size());
  for (int k : order) ++cnt[nb - 1 - sz(k) + 1];
  for (int i = 0; i < nb; ++i) cnt[i + 1] += cnt[i];
  for (int k : order) sorted[cnt[nb - 1 - sz(k)]++] = k;
  order.swap(sorted);
}

// The apply's cluster lists from h_prec and the network membership h_skip:
// the clusters with a block of each precision that the network does not
// serve, in cluster order; a list that holds every nonempty cluster is
// replaced by a launch over all clusters (empty ones return at once).
void build_apply_lists(prbf_ctx* c, bool learned) {
  const long long nc = c->nclus;
  std::vector<int> a32, a64;
  long long nonempty = 0;
  for (long long k = 0; k < nc; ++k) {
    if (!c->h_prec[k]) continue;
    ++nonempty;
    if (learned && c->h_skip[k]) continue;
    (c->h_prec[k] == 2 ? a32 : a64).push_back((int)k);
  }
  c->l32_all = (long long)a32.size() == nonempty;
  c->l64_all = (long long)a64.size() == nonempty;
  c->n32 = c->l32_all ? nc : (long long)a32.size();
  c->n64 = c->l64_all ? nc : (long long)a64.size();
  c->l32.ensure(std::max<size_t>(1, a32.size()));
  c->l64.ensure(std::max<size_t>(1, a64.size()));
  if (!c->l32_all && !a32.empty())
    PRBF_HIP(hipMemcpyAsync(c->l32, a32.data(), sizeof(int) * a32.size(), hipMemcpyHostToDevice, c->st));
  if (!c->l64_all && !a64.empty())
    PRBF_HIP(hipMemcpyAsync(c->l64, a64.data(), sizeof(int) * a64.size(), hipMemcpyHostToDevice, c->st));
  sync(c);  // the host vectors are released on return
}

// Factorises the blocks of `order` (any precisions, per h_prec) and writes
// their core rows; the blocks whose modified Cholesky replaced a pivot by its
// floor are counted in c->n_regularised.
void factor_clusters(prbf_ctx* c, const std::vector<int>& order) {
  std::vector<int> o32, o64;
  int nmax32 = 0, tmax32 = 0, nmax64 = 0, tmax64 = 0;
  for (int k : order) {
    const int n = (int)(c->h_bptr[k + 1] - c->h_bptr[k]), T = c->h_iend[k] - c->h_ista[k] + 1;
    if (c->h_prec[k] == 2) {
      o32.push_back(k);
      nmax32 = std::max(nmax32, n);
      tmax32 = std::max(tmax32, T);
    } else if (c->h_prec[k] == 1) {
      o64.push_back(k);
      nmax64 = std::max(nmax64, n);
      tmax64 = std::max(tmax64, T);
    }
  }
  largest_first(c, o32);
  largest_first(c, o64);
  const long long nc = c->nclus;
  DBuf<int>& status = c->fb_status;
  status.ensure(nc);
  // blocks of order whose pivots were regularised, counted on the device
  auto regularised = [&](const std::vector<int>& o) {
    DBuf<int>& cnt = c->fb_next;  // the counter of the finished launch
    c->fb_failed.ensure(o.size());
    PRBF_HIP(hipMemsetAsync(cnt, 0, sizeof(int), c->st));
    failed_blocks(c->st, (long long)o.size(), c->fb_order, status, c->fb_failed, cnt);
    int nf = 0;
    PRBF_HIP(hipMemcpyAsync(&nf, cnt, sizeof(int), hipMemcpyDeviceToHost, c->st));
    sync(c);
    return nf;
  };
  double t_k = 0, fl = 0;
  int used[3] = {0, 0, 0};
  if (!o64.empty()) {
    PRBF_HIP(hipMemsetAsync(status, 0, sizeof(int) * nc, c->st));
    factor_blocks<double>(c, false, o64, nmax64, tmax64, c->mptr64, c->M.p, status, true);
    c->n_regularised += regularised(o64);
    t_k += c->t_chol_s;
    fl += c->chol_flops;
    used[0] = c->used_nw, used[1] = c->used_npc, used[2] = c->used_order;
  }
  if (!o32.empty()) {
    PRBF_HIP(hipMemsetAsync(status, 0, sizeof(int) * nc, c->st));
    factor_blocks<float>(c, true, o32, nmax32, tmax32, c->mptr, c->M32.p, status, true);
    c->n_regularised += regularised(o32);
    t_k += c->t_chol_s;
    fl += c->chol_flops;
  } else if (used[0]) {
    c->used_nw = used[0], c->used_npc = used[1], c->used_order = used[2];
  }
  c->t_chol_s = t_k;
  c->chol_flops = fl;
}

void rasm_setup_impl(prbf_ctx* c, int pc) {
  need(c->has_points, PRBF_ERR_STATE, "set_points first");
  need(c->same, PRBF_ERR_STATE, "RASM needs the source set equal to the target set");
  need(pc == PRBF_PC_RASM_FP64 || pc == PRBF_PC_RASM_FP32 || pc == PRBF_PC_LEARNED, PRBF_ERR_ARG,
       "unknown preconditioner");
  const bool mixed = (pc != PRBF_PC_RASM_FP64);
  // with the learned preconditioner only the clusters off the network get blocks
  const bool learned = (pc == PRBF_PC_LEARNED);
  need(!learned || c->h_skip.size() == (size_t)c->nclus, PRBF_ERR_STATE, "prbf_learned_attach first");
  const long long nc = c->nclus;
  c->rasm_ready = false;
  // block precision per cluster, and storage for every block of each
  // precision (network clusters too, so a later fallback can add theirs)
  c->h_prec.assign(nc, 0);
  c->h_mptr.assign(nc + 1, 0);
  c->h_mptr64.assign(nc + 1, 0);
  std::vector<int> order;
  order.reserve(nc);
  for (long long k = 0; k < nc; ++k) {
    const long long n = c->h_bptr[k + 1] - c->h_bptr[k];
    const long long T = n ? (c->h_iend[k] - c->h_ista[k] + 1) : 0;
    const long long e = apply_elems((int)n, (int)T);
    c->h_prec[k] = n == 0 ? 0 : (mixed && n >= kFp32MinUnknowns) ? 2 : 1;
    c->h_mptr[k + 1] = c->h_mptr[k] + (c->h_prec[k] == 2 ? e : 0);
    c->h_mptr64[k + 1] = c->h_mptr64[k] + (c->h_prec[k] == 1 ? e : 0);
    if (n && !(learned && c->h_skip[k])) order.push_back((int)k);
  }
  need(block_lds_bytes(mixed, c->max_buffer, c->max_core, 4, 1, false) <= kMaxLdsBytes, PRBF_ERR_LIMIT,
       "RASM block too large for the LDS panel of the block factorisation");
  c->mptr.ensure(nc + 1);
  c->mptr64.ensure(nc + 1);
  PRBF_HIP(hipMemcpyAsync(c->mptr, c->h_mptr.data(), sizeof(long long) * (nc + 1), hipMemcpyHostToDevice, c->st));
  PRBF_HIP(hipMemcpyAsync(c->mptr64, c->h_mptr64.data(), sizeof(long long) * (nc + 1), hipMemcpyHostToDevice,
                          c->st));
  c->M.ensure(c->h_mptr64[nc]);
  c->M32.ensure(c->h_mptr[nc]);
  c->n_regularised = 0;
  build_apply_lists(c, learned);
  factor_clusters(c, order);
  c->rasm_ready = true;
  c->rasm_pc = pc;
}

// The network's corrections on the clusters in `list` for the sorted vector r.
void network_apply(prbf_ctx* c, const int* list, long long n, const double* r, double* z) {
  if (n <= 0) return;
  prbf_mlp* net = c->net;
  const LGeo& q = net->q;
  patch_features(c->st, q, n, list, c->g.box / q.cs, c->prm.sigma, c->pidx, c->cidx, c->pctr, c->xi, c->yi, c->zi,
                 r, nullptr, net->out_scale, c->lX, nullptr, c->ls);
  mlp_to_bf16(c->st, n * q.nin, c->lX, net->Xb);
  float* pr[3];
  unsigned short* ac[3];
  net->layers(pr, ac);
  mlp_forward(c->st, q, (int)n, net->P, net->Wt, net->Xb, pr, ac, net->Y);
  core_scatter(c->st, q, n, list, c->cidx, net->Y, c->ls, net->out_scale, z);
}

// z = B r for nrhs vectors stored back to back (length ni each); every block's
// core rows are streamed once for all of them.
void rasm_apply_impl(prbf_ctx* c, const double* r, double* z, int nrhs = 1) {
  need(c->rasm_ready, PRBF_ERR_STATE, "rasm_setup first");
  need(r && z, PRBF_ERR_ARG, "null vector");
  need(nrhs >= 1 && nrhs <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
  if (c->rasm_pc == PRBF_PC_LEARNED) need(c->learned_ready && c->net, PRBF_ERR_STATE, "prbf_learned_attach first");
  rasm_apply_fp32(c->st, nrhs, c->n32, c->l32_all ? nullptr : c->l32.p, c->bptr, c->bidx, c->ista, c->iend, c->mptr,
                  c->M32, r, c->ni, z);
  rasm_apply_fp64(c->st, nrhs, c->n64, c->l64_all ? nullptr : c->l64.p, c->bptr, c->bidx, c->ista, c->iend,
                  c->mptr64, c->M, r, c->ni, z);
  if (c->rasm_pc == PRBF_PC_LEARNED)
    for (int k = 0; k < nrhs; ++k)
      network_apply(c, c->mlp_list, c->n_mlp, r + (long long)k * c->ni, z + (long long)k * c->ni);
  PRBF_HIP(hipGetLastError());
}
// ---------------------------------------------------------------- Krylov
struct Conv {  // PETSc KSPConvergedDefault, preconditioned norm, zero guess
  double rtol, abstol, dtol, rnorm0 = 0, ttol = 0;
  int operator()(int n, double rnorm) {
    if (n == 0) {
      rnorm0 = rnorm;
      ttol = std::max(rtol * rnorm0, abstol);
    }
    if (rnorm != rnorm) return -9;
    if (rnorm <= ttol) return rnorm < abstol ? 3 : 2;
    if (rnorm >= dtol * rnorm0) return -4;
    return 0;
  }
};

struct Timed {
  prbf_ctx* c;
  double* acc;
  int* cnt;
  Timed(prbf_ctx* c_, double* a, int* n) : c(c_), acc(a), cnt(n) { PRBF_HIP(hipEventRecord(c->ev0, c->st)); }
  ~Timed() noexcept(false) {
    PRBF_HIP(hipEventRecord(c->ev1, c->st));
    PRBF_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    PRBF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *acc += ms * 1e-3;
    ++*cnt;
  }
};

// Moves clusters from the network to exact blocks (the learned
// preconditioner's per-cluster fallback) and factorises their blocks.  The
// context is marked not set up until the blocks are in place, so a failed
// factorisation leaves a state that the next solve sets up again.
void learned_fallback(prbf_ctx* c, const std::vector<int>& newc) {
  if (newc.empty()) return;
  c->rasm_ready = false;
  for (int k : newc) c->h_skip[k] = 0;
  std::vector<int> list;
  for (long long k = 0; k < c->nclus; ++k)
    if (c->h_skip[k]) list.push_back((int)k);
  c->n_mlp = (long long)list.size();
  if (!list.empty())
    PRBF_HIP(hipMemcpyAsync(c->mlp_list, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, c->st));
  c->h_mlp_list = list;
  c->n_net_fallback += (int)newc.size();
  build_apply_lists(c, true);
  factor_clusters(c, newc);
  c->rasm_ready = true;
}

// Preconditioner application without timing (usable under graph capture).
void pc_apply_raw(prbf_ctx* c, int pc, long long n, const double* v, double* y) {
  if (pc == PRBF_PC_NONE)
    PRBF_HIP(hipMemcpyAsync(y, v, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
  else
    rasm_apply_impl(c, v, y);
}

// Offsets in the Arnoldi state (layout in krylov.hip).
inline long long ks_rs(int m) { return (long long)(m + 1) * (m + 1) + 2 * (m + 1); }
inline long long ks_inv(int m) { return ks_rs(m) + m + 2; }
inline long long ks_out(int m) { return ks_inv(m) + 1; }
constexpr int kPinStep = 256;  // pinned slots receiving a step's residual estimate and flags
// least core-residual reduction per restart cycle for a cluster to stay on the
// network: 1000x (exact FP32 blocks reduce the residual by far more per cycle,
// so a cluster that does not is moved at the first restart that exposes it)
constexpr double kNetCycleDrop = 1e-3;

// Enqueues Arnoldi step j of a cycle: the new basis vector, its Gram-Schmidt
// coefficients and norm, the Givens update of H and the residual estimate,
// the normalisation, and the copy of the estimate to pinned memory.
void enqueue_arnoldi_step(prbf_ctx* c, int pc, bool flex, long long n, int m, int j) {
  double* V = c->V;
  double* w = V + (long long)(j + 1) * n;
  double* dsc = c->scal;
  if (flex) {
    double* zj = c->Z.p + (long long)j * n;
    pc_apply_raw(c, pc, n, V + (long long)j * n, zj);
    matvec_impl(c, 1, zj, w);
  } else {
    matvec_impl(c, 1, V + (long long)j * n, c->v1);
    pc_apply_raw(c, pc, n, c->v1, w);
  }
  mdot(c->st, n, w, V, n, j + 1, c->part, dsc);
  maxpy_norm2(c->st, n, w, V, n, j + 1, dsc, -1.0, c->part, dsc + 64);
  arnoldi_step(c->st, j, m, dsc, c->kst);
  scale_dev(c->st, n, w, c->kst.p + ks_inv(m));
  PRBF_HIP(hipMemcpyAsync(c->pin + kPinStep, c->kst.p + ks_out(m), 3 * sizeof(double), hipMemcpyDeviceToHost,
                          c->st));
}

// (Re)captures the m Arnoldi-step graphs when anything they captured changed:
// sizes, the preconditioner, or the address of any buffer their kernels read.
void ensure_graphs(prbf_ctx* c, int pc, bool flex, long long n, int m) {
  auto u = [](const void* p) { return (unsigned long long)(uintptr_t)p; };
  unsigned long long sig;
  std::memcpy(&sig, &c->prm.sigma, sizeof(sig));
  std::vector<unsigned long long> key = {
      (unsigned long long)n, (unsigned long long)m, flex ? 1ULL : 0ULL, (unsigned long long)pc,
      (unsigned long long)c->rasm_pc, (unsigned long long)c->nclus, (unsigned long long)c->ni,
      (unsigned long long)c->nj, (unsigned long long)c->prm.dim, sig, u(c->V.p), u(c->Z.p), u(c->v1.p),
      u(c->part.p), u(c->scal.p), u(c->kst.p), u(c->pin), u(c->ista.p), u(c->iend.p), u(c->tptr.p), u(c->tidx.p),
      u(c->xi.p), u(c->yi.p), u(c->zi.p), u(c->xj_()), u(c->yj_()), u(c->zj_()), u(c->bptr.p), u(c->bidx.p),
      u(c->mptr.p), u(c->M.p), u(c->M32.p), u(c->mptr64.p), u(c->l32.p), u(c->l64.p),
      (unsigned long long)c->n32, (unsigned long long)c->n64, (unsigned long long)c->l32_all,
      (unsigned long long)c->l64_all, u(c->mlp_list.p),
      (unsigned long long)c->n_mlp, u(c->pidx.p), u(c->cidx.p), u(c->pctr.p), u(c->lX.p), u(c->ls.p), u(c->net)};
  if (c->net) {
    unsigned long long os;
    std::memcpy(&os, &c->net->out_scale, sizeof(os));
    for (unsigned long long v : {os, u(c->net->P.p), u(c->net->Wt.p), u(c->net->Xb.p), u(c->net->act.p),
                                 u(c->net->pre.p), u(c->net->Y.p), (unsigned long long)c->net->cap})
      key.push_back(v);
  }
  if (key == c->kgraph_key && (int)c->kgraphs.size() == m) return;
  c->drop_graphs();
  for (int j = 0; j < m; ++j) {
    PRBF_HIP(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    hipGraph_t gr = nullptr;
    try {
      enqueue_arnoldi_step(c, pc, flex, n, m, j);
    } catch (...) {
      (void)hipStreamEndCapture(c->st, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      (void)hipGetLastError();
      throw;
    }
    PRBF_HIP(hipStreamEndCapture(c->st, &gr));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    PRBF_HIP(e);
    c->kgraphs.push_back(ex);
  }
  c->kgraph_key = key;
}

// Restarted GMRES (left PC) or FGMRES (right PC), PETSc's KSPGMRESCycle
// structure: classical Gram-Schmidt without refinement, Givens rotations on
// the host, solution update at the end of each cycle.
void krylov_impl(prbf_ctx* c, const prbf_solve_opts& o, const double* b, double* x, prbf_solve_info* info,
                 double* hist, int hist_cap) {
  const long long n = c->ni;
  const int m = o.restart;
  need(m >= 1 && m < 63, PRBF_ERR_ARG, "restart must be in 1..62");
  need(o.pc == PRBF_PC_NONE || c->rasm_ready, PRBF_ERR_STATE, "rasm_setup first");
  const bool flex = (o.ksp == PRBF_KSP_FGMRES);
  need(flex || o.pc != PRBF_PC_LEARNED, PRBF_ERR_ARG, "the learned preconditioner needs FGMRES");
  const bool graph = o.graph != 0 && c->st != nullptr;  // the legacy default stream cannot be captured
  c->last_restart = m;
  c->V.ensure((size_t)(m + 1) * n);
  if (flex) c->Z.ensure((size_t)m * n);
  c->v1.ensure(n);
  c->v2.ensure(n);
  c->part.ensure(kRedBlocks * 64);
  c->scal.ensure(256);
  if (graph) {
    c->kst.ensure(arnoldi_state_len(m));
    ensure_graphs(c, o.pc, flex, n, m);
  }
  double* V = c->V;
  auto Vj = [&](int j) { return V + (long long)j * n; };
  double* tmp = c->v1;
  double* tmp2 = c->v2;
  double* dsc = c->scal;  // [0..63] dots, [64] norm^2, [128..191] y
  double* pin = c->pin;
  prbf_solve_info inf{};
  auto A = [&](const double* v, double* y) {
    Timed t(c, &inf.t_matvec_s, &inf.n_matvec);
    matvec_impl(c, 1, v, y);
  };
  auto B = [&](const double* v, double* y) {
    if (o.pc == PRBF_PC_NONE) {
      PRBF_HIP(hipMemcpyAsync(y, v, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
      return;
    }
    Timed t(c, &inf.t_pc_s, &inf.n_pc);
    rasm_apply_impl(c, v, y);
  };
  auto read_norm = [&]() {
    PRBF_HIP(hipMemcpyAsync(pin, dsc + 64, sizeof(double), hipMemcpyDeviceToHost, c->st));
    sync(c);
    return std::sqrt(pin[0]);
  };
  std::vector<double> hh((size_t)(m + 1) * (m + 1), 0.0), cc(m + 1), ss(m + 1), rs(m + 2), nrs(m + 1);
  auto HH = [&](int i, int j) -> double& { return hh[(size_t)j * (m + 1) + i]; };
  Conv conv{o.rtol, o.abstol, o.dtol};
  int nh = 0;
  double last_res = 0.0;
  auto log_res = [&](double r) {
    if (hist && nh < hist_cap) hist[nh] = r;
    ++nh;
    last_res = r;
  };
  const double t0 = now_s();
  fill(c->st, n, x, 0.0);
  std::vector<double> prev_core;  // network clusters' core residuals at the last cycle start
  bool guess_zero = true;
  int its = 0, itcount = 0, reason = 0;
  while (!reason) {
    // initial residual of the cycle
    if (guess_zero) {
      PRBF_HIP(hipMemcpyAsync(tmp, b, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
    } else {
      A(x, tmp2);
      xpay_sub(c->st, n, b, tmp2, tmp);
    }
    if (o.pc == PRBF_PC_LEARNED && c->n_mlp > 0) {
      // FGMRES's residual judges the network per cluster: a cluster whose core
      // residual fell by less than kNetCycleDrop over the last cycle moves to
      // its exact block
      const long long nm = c->n_mlp;
      std::vector<double> cur(nm);
      core_norms(c->st, c->net->q, nm, c->mlp_list, c->cidx, tmp, c->lnorm);
      PRBF_HIP(hipMemcpyAsync(cur.data(), c->lnorm, sizeof(double) * nm, hipMemcpyDeviceToHost, c->st));
      sync(c);
      std::vector<int> newc;
      if (!prev_core.empty())
        for (long long k = 0; k < nm; ++k) {
          const int cl = c->h_mlp_list[k];
          if (!(cur[k] <= kNetCycleDrop * kNetCycleDrop * prev_core[cl])) newc.push_back(cl);
        }
      prev_core.assign(c->nclus, 0.0);
      for (long long k = 0; k < nm; ++k) prev_core[c->h_mlp_list[k]] = cur[k];
      if (!newc.empty()) {
        learned_fallback(c, newc);
        if (graph) ensure_graphs(c, o.pc, flex, n, m);
      }
    }
    if (flex) {
      PRBF_HIP(hipMemcpyAsync(Vj(0), tmp, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
    } else {
      B(tmp, Vj(0));
    }
    norm2sq(c->st, n, Vj(0), c->part, dsc + 64);
    double res = read_norm();
    if (res != 0.0) scale(c->st, n, Vj(0), 1.0 / res);
    rs[0] = res;
    if (graph) set_scalar(c->st, c->kst.p + ks_rs(m), res);
    log_res(res);
    if (res == 0.0) {
      reason = 3;
      break;
    }
    reason = conv(its, res);
    int loc = 0;
    bool hapend = false;
    while (!reason && loc < m && its < o.max_it) {
      if (graph) {
        PRBF_HIP(hipGraphLaunch(c->kgraphs[loc], c->st));
        sync(c);
        const double* out = pin + kPinStep;
        ++inf.n_matvec;
        if (o.pc != PRBF_PC_NONE) ++inf.n_pc;
        if (out[2] != 0.0) {
          reason = -2;
          break;
        }
        hapend = out[1] != 0.0;
        res = out[0];
        ++loc;
        ++its;
        log_res(res);
        reason = conv(its, res);
        if (hapend && !reason) reason = 8;
        continue;
      }
      if (flex) {
        B(Vj(loc), c->Z.p + (long long)loc * n);
        A(c->Z.p + (long long)loc * n, Vj(loc + 1));
      } else {
        A(Vj(loc), tmp);
        B(tmp, Vj(loc + 1));
      }
      double* w = Vj(loc + 1);
      mdot(c->st, n, w, V, n, loc + 1, c->part, dsc);
      maxpy_norm2(c->st, n, w, V, n, loc + 1, dsc, -1.0, c->part, dsc + 64);
      PRBF_HIP(hipMemcpyAsync(pin, dsc, sizeof(double) * 65, hipMemcpyDeviceToHost, c->st));
      sync(c);
      for (int j = 0; j <= loc; ++j) HH(j, loc) = pin[j];
      double tt = std::sqrt(pin[64]);
      if (tt != 0.0) scale(c->st, n, w, 1.0 / tt);
      HH(loc + 1, loc) = tt;
      double hapbnd = std::min(std::fabs(tt / rs[loc]), 1e-30);
      if (tt < hapbnd) hapend = true;
      for (int j = 1; j <= loc; ++j) {
        double a0 = HH(j - 1, loc);
        HH(j - 1, loc) = cc[j - 1] * a0 + ss[j - 1] * HH(j, loc);
        HH(j, loc) = cc[j - 1] * HH(j, loc) - ss[j - 1] * a0;
      }
      if (!hapend) {
        double a0 = HH(loc, loc), a1 = HH(loc + 1, loc);
        double t = std::sqrt(a0 * a0 + a1 * a1);
        if (t == 0.0) {
          reason = -2;
          break;
        }
        cc[loc] = a0 / t;
        ss[loc] = a1 / t;
        rs[loc + 1] = -(ss[loc] * rs[loc]);
        rs[loc] = cc[loc] * rs[loc];
        HH(loc, loc) = cc[loc] * a0 + ss[loc] * a1;
        res = std::fabs(rs[loc + 1]);
      } else {
        res = 0.0;
      }
      ++loc;
      ++its;
      log_res(res);
      reason = conv(its, res);
      if (hapend && !reason) reason = 8;
    }
    int it = loc - 1;
    if (it >= 0 && graph) {
      arnoldi_backsub(c->st, it, m, c->kst, dsc + 128);
      update_x(c->st, n, x, flex ? c->Z.p : V, n, it + 1, dsc + 128);
    } else if (it >= 0) {
      nrs[it] = rs[it] / HH(it, it);
      for (int ii = 1; ii <= it; ++ii) {
        int k = it - ii;
        double t = rs[k];
        for (int j = k + 1; j <= it; ++j) t = t - HH(k, j) * nrs[j];
        nrs[k] = t / HH(k, k);
      }
      for (int j = 0; j <= it; ++j) pin[128 + j] = nrs[j];
      PRBF_HIP(hipMemcpyAsync(dsc + 128, pin + 128, sizeof(double) * (it + 1), hipMemcpyHostToDevice, c->st));
      update_x(c->st, n, x, flex ? c->Z.p : V, n, it + 1, dsc + 128);
      sync(c);  // pinned y consumed before the next cycle reuses it
    }
    itcount += loc;
    if (itcount >= o.max_it && !reason) {
      reason = -3;
      break;
    }
    guess_zero = false;
  }
  sync(c);
  inf.its = its;
  inf.reason = reason;
  inf.rnorm0 = conv.rnorm0;
  inf.rnorm = last_res;
  inf.t_solve_s = now_s() - t0;
  if (info) *info = inf;
}

// Restarted GMRES (left PC) or FGMRES (right PC) on k = 1..3 right-hand
// sides at once, sorted order, b and x holding them back to back.  Each
// right-hand side keeps its own Krylov basis, Hessenberg matrix, Givens
// rotations and convergence test, exactly as krylov_impl's kernel-by-kernel
// path runs them (same kernels, same reduction order, host Givens), while
// every step's matvec is one launch over the k vectors (one exponential per
// pair for all of them) and every preconditioner apply streams each block's
// core rows once.  The bases interleave the right-hand sides, V[j][r][n], so
// the k vectors of one basis index are contiguous.  A right-hand side that
// has converged stops taking part in the Gram-Schmidt steps; the cycles of all
// of them restart together, as each would on its own.
void krylov_batch_impl(prbf_ctx* c, const prbf_solve_opts& o, int k, const double* b, double* x,
                       prbf_solve_info* info, double* const* hist, int hist_cap) {
  const long long n = c->ni, kn = (long long)k * n;
  const int m = o.restart;
  need(k >= 1 && k <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
  need(m >= 1 && m < 63, PRBF_ERR_ARG, "restart must be in 1..62");
  need(o.pc == PRBF_PC_NONE || c->rasm_ready, PRBF_ERR_STATE, "rasm_setup first");
  const bool flex = (o.ksp == PRBF_KSP_FGMRES);
  need(flex || o.pc != PRBF_PC_LEARNED, PRBF_ERR_ARG, "the learned preconditioner needs FGMRES");
  c->last_restart = 0;  // the interleaved bases are not single-vector bases
  c->V.ensure((size_t)(m + 1) * kn);
  if (flex) c->Z.ensure((size_t)m * kn);
  c->v1.ensure(kn);
  c->v2.ensure(kn);
  c->part.ensure(kRedBlocks * 64);
  c->scal.ensure(256 * 3);
  double* V = c->V;
  auto Vr = [&](int j, int r) { return V + (long long)j * kn + (long long)r * n; };
  double* tmp = c->v1;
  double* tmp2 = c->v2;
  double* pin = c->pin;  // [r 65 + j] dots and norms of right-hand side r; [256 + 64 r + j] its y
  auto dsc = [&](int r) { return c->scal.p + 256 * r; };
  prbf_solve_info inf[3] = {};
  // (n_matvec and n_pc count, per right-hand side, the operations it used)
  auto A = [&](const double* v, double* y) { matvec_impl(c, k, v, y); };
  auto B = [&](const double* v, double* y) {
    if (o.pc == PRBF_PC_NONE)
      PRBF_HIP(hipMemcpyAsync(y, v, sizeof(double) * kn, hipMemcpyDeviceToDevice, c->st));
    else
      rasm_apply_impl(c, v, y, k);
  };
  struct State {
    std::vector<double> hh, cc, ss, rs, nrs;
    Conv conv{0, 0, 0};
    int its = 0, itcount = 0, reason = 0, nh = 0, loc = 0;
    bool hapend = false, guess_zero = true;
    double last_res = 0.0;
  };
  std::vector<State> S(k);
  for (State& q : S) {
    q.hh.assign((size_t)(m + 1) * (m + 1), 0.0);
    q.cc.assign(m + 1, 0.0);
    q.ss.assign(m + 1, 0.0);
    q.rs.assign(m + 2, 0.0);
    q.nrs.assign(m + 1, 0.0);
    q.conv = Conv{o.rtol, o.abstol, o.dtol};
  }
  auto HH = [&](State& q, int i, int j) -> double& { return q.hh[(size_t)j * (m + 1) + i]; };
  auto log_res = [&](int r, double v) {
    State& q = S[r];
    if (hist && hist[r] && q.nh < hist_cap) hist[r][q.nh] = v;
    ++q.nh;
    q.last_res = v;
  };
  const double t0 = now_s();
  fill(c->st, kn, x, 0.0);
  std::vector<double> prev_core;  // network clusters' core residuals at the last cycle start
  auto any_active = [&]() {
    for (const State& q : S)
      if (!q.reason) return true;
    return false;
  };
  while (any_active()) {
    // initial residuals of the cycle (the matvec covers all k; only the
    // unfinished right-hand sides use it)
    bool need_ax = false;
    for (const State& q : S) need_ax |= !q.reason && !q.guess_zero;
    if (need_ax) A(x, tmp2);
    for (int r = 0; r < k; ++r) {
      const State& q = S[r];
      if (q.reason) continue;
      if (!q.guess_zero) ++inf[r].n_matvec;
      if (!flex && o.pc != PRBF_PC_NONE) ++inf[r].n_pc;
      if (q.guess_zero)
        PRBF_HIP(hipMemcpyAsync(tmp + r * n, b + r * n, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
      else
        xpay_sub(c->st, n, b + r * n, tmp2 + r * n, tmp + r * n);
    }
    if (o.pc == PRBF_PC_LEARNED && c->n_mlp > 0) {
      // as krylov_impl, on the network clusters' core residuals summed over
      // the right-hand sides still running
      const long long nm = c->n_mlp;
      std::vector<double> cur(nm, 0.0), one(nm);
      for (int r = 0; r < k; ++r) {
        if (S[r].reason) continue;
        core_norms(c->st, c->net->q, nm, c->mlp_list, c->cidx, tmp + r * n, c->lnorm);
        PRBF_HIP(hipMemcpyAsync(one.data(), c->lnorm, sizeof(double) * nm, hipMemcpyDeviceToHost, c->st));
        sync(c);
        for (long long i = 0; i < nm; ++i) cur[i] += one[i];
      }
      std::vector<int> newc;
      if (!prev_core.empty())
        for (long long i = 0; i < nm; ++i) {
          const int cl = c->h_mlp_list[i];
          if (!(cur[i] <= kNetCycleDrop * kNetCycleDrop * prev_core[cl])) newc.push_back(cl);
        }
      prev_core.assign(c->nclus, 0.0);
      for (long long i = 0; i < nm; ++i) prev_core[c->h_mlp_list[i]] = cur[i];
      learned_fallback(c, newc);
    }
    if (flex)
      PRBF_HIP(hipMemcpyAsync(Vr(0, 0), tmp, sizeof(double) * kn, hipMemcpyDeviceToDevice, c->st));
    else
      B(tmp, Vr(0, 0));
    for (int r = 0; r < k; ++r)
      if (!S[r].reason) norm2sq(c->st, n, Vr(0, r), c->part, dsc(r) + 64);
    for (int r = 0; r < k; ++r)
      if (!S[r].reason) PRBF_HIP(hipMemcpyAsync(pin + 65 * r + 64, dsc(r) + 64, sizeof(double), hipMemcpyDeviceToHost,
                                                c->st));
    sync(c);
    std::vector<bool> run(k, false);
    for (int r = 0; r < k; ++r) {
      State& q = S[r];
      if (q.reason) continue;
      const double res = std::sqrt(pin[65 * r + 64]);
      if (res != 0.0) scale(c->st, n, Vr(0, r), 1.0 / res);
      q.rs[0] = res;
      log_res(r, res);
      if (res == 0.0) {
        q.reason = 3;
        continue;
      }
      q.reason = q.conv(q.its, res);
      q.loc = 0;
      q.hapend = false;
      run[r] = !q.reason && q.its < o.max_it;
    }
    // the cycle: steps while any right-hand side still runs it
    for (int loc = 0; loc < m; ++loc) {
      bool any = false;
      for (int r = 0; r < k; ++r) any |= run[r];
      if (!any) break;
      if (flex) {
        B(Vr(loc, 0), c->Z.p + (long long)loc * kn);
        A(c->Z.p + (long long)loc * kn, Vr(loc + 1, 0));
      } else {
        A(Vr(loc, 0), tmp);
        B(tmp, Vr(loc + 1, 0));
      }
      for (int r = 0; r < k; ++r) {
        if (!run[r]) continue;
        ++inf[r].n_matvec;
        if (o.pc != PRBF_PC_NONE) ++inf[r].n_pc;
        double* w = Vr(loc + 1, r);
        mdot(c->st, n, w, Vr(0, r), kn, loc + 1, c->part, dsc(r));
        maxpy_norm2(c->st, n, w, Vr(0, r), kn, loc + 1, dsc(r), -1.0, c->part, dsc(r) + 64);
        PRBF_HIP(hipMemcpyAsync(pin + 65 * r, dsc(r), sizeof(double) * 65, hipMemcpyDeviceToHost, c->st));
      }
      sync(c);
      for (int r = 0; r < k; ++r) {
        if (!run[r]) continue;
        State& q = S[r];
        double* w = Vr(loc + 1, r);
        const double* pr = pin + 65 * r;
        for (int j = 0; j <= loc; ++j) HH(q, j, loc) = pr[j];
        const double tt = std::sqrt(pr[64]);
        if (tt != 0.0) scale(c->st, n, w, 1.0 / tt);
        HH(q, loc + 1, loc) = tt;
        const double hapbnd = std::min(std::fabs(tt / q.rs[loc]), 1e-30);
        if (tt < hapbnd) q.hapend = true;
        for (int j = 1; j <= loc; ++j) {
          const double a0 = HH(q, j - 1, loc);
          HH(q, j - 1, loc) = q.cc[j - 1] * a0 + q.ss[j - 1] * HH(q, j, loc);
          HH(q, j, loc) = q.cc[j - 1] * HH(q, j, loc) - q.ss[j - 1] * a0;
        }
        double res;
        if (!q.hapend) {
          const double a0 = HH(q, loc, loc), a1 = HH(q, loc + 1, loc);
          const double t = std::sqrt(a0 * a0 + a1 * a1);
          if (t == 0.0) {
            q.reason = -2;
            run[r] = false;
            continue;
          }
          q.cc[loc] = a0 / t;
          q.ss[loc] = a1 / t;
          q.rs[loc + 1] = -(q.ss[loc] * q.rs[loc]);
          q.rs[loc] = q.cc[loc] * q.rs[loc];
          HH(q, loc, loc) = q.cc[loc] * a0 + q.ss[loc] * a1;
          res = std::fabs(q.rs[loc + 1]);
        } else {
          res = 0.0;
        }
        q.loc = loc + 1;
        ++q.its;
        log_res(r, res);
        q.reason = q.conv(q.its, res);
        if (q.hapend && !q.reason) q.reason = 8;
        run[r] = !q.reason && q.its < o.max_it;
      }
    }
    // solution updates, each with its own step count
    for (int r = 0; r < k; ++r) {
      State& q = S[r];
      const int it = q.loc - 1;  // steps this cycle (finished right-hand sides: none)
      if (it >= 0) {
        q.nrs[it] = q.rs[it] / HH(q, it, it);
        for (int ii = 1; ii <= it; ++ii) {
          const int kk = it - ii;
          double t = q.rs[kk];
          for (int j = kk + 1; j <= it; ++j) t = t - HH(q, kk, j) * q.nrs[j];
          q.nrs[kk] = t / HH(q, kk, kk);
        }
        for (int j = 0; j <= it; ++j) pin[256 + 64 * r + j] = q.nrs[j];
        PRBF_HIP(hipMemcpyAsync(dsc(r) + 128, pin + 256 + 64 * r, sizeof(double) * (it + 1), hipMemcpyHostToDevice,
                                c->st));
        update_x(c->st, n, x + r * n, flex ? c->Z.p + r * n : Vr(0, r), kn, it + 1, dsc(r) + 128);
      }
      q.itcount += q.loc;
      q.loc = 0;
      if (!q.reason && q.itcount >= o.max_it) q.reason = -3;
      q.guess_zero = false;
    }
    sync(c);  // pinned y consumed before the next cycle reuses it
  }
  sync(c);
  for (int r = 0; r < k; ++r) {
    inf[r].its = S[r].its;
    inf[r].reason = S[r].reason;
    inf[r].rnorm0 = S[r].conv.rnorm0;
    inf[r].rnorm = S[r].last_res;
    inf[r].t_solve_s = now_s() - t0;
    if (o.pc == PRBF_PC_LEARNED) {
      inf[r].n_network = (int)c->n_mlp;
      inf[r].n_net_fallback = c->n_net_fallback;
    }
    if (info) info[r] = inf[r];
  }
}

// The solver options (defaults when opts is null), validated; the
// preconditioner set up when it is not (its wall time in *t_setup).
prbf_solve_opts checked_opts(prbf_ctx* c, const prbf_solve_opts* opts, int hist_cap, double* t_setup) {
  need(c->has_points && c->same, PRBF_ERR_STATE, "set_points (one set) first");
  prbf_solve_opts o;
  prbf_solve_opts_default(&o);
  if (opts) o = *opts;
  *t_setup = 0.0;
  need(o.ksp == PRBF_KSP_GMRES || o.ksp == PRBF_KSP_FGMRES, PRBF_ERR_ARG, "unknown Krylov method");
  need(o.pc == PRBF_PC_NONE || o.pc == PRBF_PC_RASM_FP64 || o.pc == PRBF_PC_RASM_FP32 || o.pc == PRBF_PC_LEARNED,
       PRBF_ERR_ARG, "unknown preconditioner");
  need(o.max_it >= 0 && o.max_it <= (1 << 24), PRBF_ERR_ARG, "max_it must be in 0..2^24");
  need(o.rtol >= 0 && o.abstol >= 0 && o.dtol > 0, PRBF_ERR_ARG, "tolerances must be non-negative (dtol positive)");
  need(hist_cap >= 0, PRBF_ERR_ARG, "hist_cap must be non-negative");
  need(o.pc != PRBF_PC_LEARNED || c->learned_ready, PRBF_ERR_STATE, "prbf_learned_attach first");
  if (o.pc != PRBF_PC_NONE && (!c->rasm_ready || c->rasm_pc != o.pc)) {
    double t0 = now_s();
    rasm_setup_impl(c, o.pc);
    *t_setup = now_s() - t0;
  }
  return o;
}

void solve_multi_dev_impl(prbf_ctx* c, const prbf_solve_opts* opts, int k, const double* w, double* gamma,
                          prbf_solve_info* info, double* hist, int hist_cap) {
  need(w && gamma, PRBF_ERR_ARG, "null vector");
  need(k >= 1 && k <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
  double t_setup = 0.0;
  const prbf_solve_opts o = checked_opts(c, opts, hist_cap, &t_setup);
  const long long n = c->ni;
  c->v3.ensure((size_t)2 * k * n);
  double* bs = c->v3;
  double* xs = c->v3.p + (long long)k * n;
  for (int r = 0; r < k; ++r) gather(c->st, n, c->perm_i, w + r * n, bs + r * n);
  double* hp[3] = {nullptr, nullptr, nullptr};
  for (int r = 0; r < k && hist; ++r) hp[r] = hist + (long long)r * hist_cap;
  prbf_solve_info local[3] = {};
  krylov_batch_impl(c, o, k, bs, xs, local, hp, hist_cap);
  for (int r = 0; r < k; ++r) scatter(c->st, n, c->perm_i, xs + r * n, gamma + r * n);
  for (int r = 0; r < k; ++r) {
    local[r].t_setup_s = t_setup;
    local[r].n_regularised = c->n_regularised;
    if (info) info[r] = local[r];
  }
}

void solve_dev_impl(prbf_ctx* c, const prbf_solve_opts* opts, const double* w, double* gamma,
                    prbf_solve_info* info, double* hist, int hist_cap) {
  need(w && gamma, PRBF_ERR_ARG, "null vector");
  double t_setup = 0.0;
  const prbf_solve_opts o = checked_opts(c, opts, hist_cap, &t_setup);
  const long long n = c->ni;
  c->v3.ensure(2 * n);
  double* bs = c->v3;
  double* xs = c->v3.p + n;
  gather(c->st, n, c->perm_i, w, bs);
  std::vector<double> hbuf;
  double* h = hist;
  int cap = hist_cap;
  if (!h) {
    hbuf.resize(o.max_it + 64);
    h = hbuf.data();
    cap = (int)hbuf.size();
  }
  prbf_solve_info local{};
  krylov_impl(c, o, bs, xs, &local, h, cap);
  scatter(c->st, n, c->perm_i, xs, gamma);
  local.t_setup_s = t_setup;
  local.n_regularised = c->n_regularised;
  if (o.pc == PRBF_PC_LEARNED) {
    local.n_network = (int)c->n_mlp;
    local.n_net_fallback = c->n_net_fallback;
  }
  if (info) *info = local;
}

// ------------------------------------------------ allocation-free evaluation
// The staging of prbf_evaluate_dev_noalloc: the strengths in sorted order
// and the field before it is scattered back, nrhs (nj + ni) doubles at most
// kScratchDoubles (16 MiB).  One copy per GPU, held until the process ends;
// the host reaches it only through the address hipGetSymbolAddress gives on
// that device (the array's name is a host-side stand-in).
constexpr long long kScratchDoubles = 1LL << 21;
static __device__ __align__(256) double scratch[kScratchDoubles];

// The scratch array's address on the current device, resolved once per device
// (with that device current) and kept; null on failure.
static double* device_scratch() {
  static std::mutex m;
  static std::vector<std::pair<int, double*>> known;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(m);
  for (const auto& k : known)
    if (k.first == dev) return k.second;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return nullptr;
  }
  known.push_back({dev, (double*)ptr});
  return (double*)ptr;
}

}  // namespace

// ============================================================== C ABI
extern "C" {

void prbf_solve_opts_default(prbf_solve_opts* o) {
  if (!o) return;
  o->ksp = PRBF_KSP_GMRES;
  o->pc = PRBF_PC_RASM_FP64;
  o->restart = 30;
  o->max_it = 1000;
  o->rtol = 1e-10;
  o->abstol = 1e-50;
  o->dtol = 1e5;
  o->graph = 1;
}

prbf_status prbf_create(prbf_ctx** out, const prbf_params* p) {
  if (!out || !p) return PRBF_ERR_ARG;
  *out = nullptr;
  if ((p->dim != 2 && p->dim != 3) || !(p->sigma > 0) || !(p->nsigma_box > 0) || !(p->sigma_buffer > 0) ||
      !(p->sigma_trunc > 0))
    return PRBF_ERR_ARG;
  prbf_ctx* c = new (std::nothrow) prbf_ctx;
  if (!c) return PRBF_ERR_NOMEM;
  c->prm = *p;
  const bool ok = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) == hipSuccess &&
                  hipHostMalloc((void**)&c->pin, 512 * sizeof(double), hipHostMallocDefault) == hipSuccess &&
                  hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess;
  c->own_stream = c->st != nullptr;
  c->scratch = ok ? device_scratch() : nullptr;
  if (!ok || !c->scratch) {
    (void)hipGetLastError();
    prbf_destroy(c);
    return PRBF_ERR_HIP;
  }
  *out = c;
  return PRBF_OK;
}

prbf_status prbf_destroy(prbf_ctx* c) {
  if (!c) return PRBF_ERR_ARG;
  if (c->st) (void)hipStreamSynchronize(c->st);
  c->drop_graphs();
  if (c->pin) (void)hipHostFree(c->pin);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own_stream && c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return PRBF_OK;
}

static thread_local std::string g_free_err;

const char* prbf_last_error(const prbf_ctx* c) { return c ? c->err.c_str() : g_free_err.c_str(); }

prbf_status prbf_set_stream(prbf_ctx* c, void* stream) {
  return guarded(c, [&] {
    sync(c);
    if (c->own_stream) PRBF_HIP(hipStreamDestroy(c->st));
    c->st = (hipStream_t)stream;
    c->own_stream = false;
  });
}

prbf_status prbf_synchronize(prbf_ctx* c) {
  return guarded(c, [&] { sync(c); });
}

prbf_status prbf_set_points_dev(prbf_ctx* c, int64_t ni, const double* xi, const double* yi, const double* zi,
                                int64_t nj, const double* xj, const double* yj, const double* zj) {
  return guarded(c, [&] { set_points_impl(c, ni, xi, yi, zi, nj, xj, yj, zj); });
}

prbf_status prbf_set_points(prbf_ctx* c, int64_t ni, const double* xi, const double* yi, const double* zi,
                            int64_t nj, const double* xj, const double* yj, const double* zj) {
  return guarded(c, [&] {
    need(ni > 0, PRBF_ERR_ARG, "ni must be positive");
    const int dim = c->prm.dim;
    need(xi && yi && (dim == 2 || zi), PRBF_ERR_ARG, "missing target coordinates");
    bool same = (xj == nullptr);
    long long nsrc = same ? 0 : nj;
    need(same || nj > 0, PRBF_ERR_ARG, "nj must be positive");
    need(same || (yj && (dim == 2 || zj)), PRBF_ERR_ARG, "missing source coordinates");
    DBuf<double> buf;
    buf.ensure((size_t)3 * (ni + nsrc));
    double* p = buf;
    auto up = [&](double* dst, const double* src, long long n) {
      if (src) PRBF_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    };
    up(p, xi, ni);
    up(p + ni, yi, ni);
    if (dim == 3) up(p + 2 * ni, zi, ni);
    double* q = p + 3 * ni;
    if (!same) {
      up(q, xj, nsrc);
      up(q + nsrc, yj, nsrc);
      if (dim == 3) up(q + 2 * nsrc, zj, nsrc);
    }
    set_points_impl(c, ni, p, p + ni, dim == 3 ? p + 2 * ni : nullptr, nj, same ? nullptr : q,
                    same ? nullptr : q + nsrc, (!same && dim == 3) ? q + 2 * nsrc : nullptr);
    sync(c);
  });
}

static void d2h(prbf_ctx* c, void* dst, const void* src, size_t bytes) {
  if (dst && bytes) PRBF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->st));
}

prbf_status prbf_set_points_adaptive_dev(prbf_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                                         int cap, int max_depth) {
  return guarded(c, [&] { set_points_tree_impl(c, n, x, y, z, cap, max_depth); });
}

prbf_status prbf_set_points_adaptive(prbf_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                                     int cap, int max_depth) {
  return guarded(c, [&] {
    need(n > 0, PRBF_ERR_ARG, "n must be positive");
    const int dim = c->prm.dim;
    need(x && y && (dim == 2 || z), PRBF_ERR_ARG, "missing coordinates");
    DBuf<double> buf;
    buf.ensure((size_t)3 * n);
    double* p = buf;
    PRBF_HIP(hipMemcpyAsync(p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    PRBF_HIP(hipMemcpyAsync(p + n, y, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    if (dim == 3) PRBF_HIP(hipMemcpyAsync(p + 2 * n, z, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    set_points_tree_impl(c, n, p, p + n, dim == 3 ? p + 2 * n : nullptr, cap, max_depth);
    sync(c);
  });
}

prbf_status prbf_get_leaves(prbf_ctx* c, double* xc, double* yc, double* zc, int32_t* depth, int32_t* cell0,
                            double* blen, double* tlen) {
  return guarded(c, [&] {
    need(c->has_points && c->tree, PRBF_ERR_STATE, "prbf_set_points_adaptive first");
    const size_t nl = (size_t)c->nclus;
    d2h(c, xc, c->lcx, sizeof(double) * nl);
    d2h(c, yc, c->lcy, sizeof(double) * nl);
    if (c->prm.dim == 3) d2h(c, zc, c->lcz, sizeof(double) * nl);
    d2h(c, cell0, c->lcell, sizeof(int) * nl);
    d2h(c, blen, c->lbl, sizeof(double) * nl);
    d2h(c, tlen, c->ltl, sizeof(double) * nl);
    sync(c);
    if (depth) std::memcpy(depth, c->h_ldepth.data(), sizeof(int) * nl);
  });
}

prbf_status prbf_get_info(prbf_ctx* c, int64_t* info, double* geom) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    const Grid& g = c->g;
    if (info) {
      int64_t v[16] = {c->nclus, g.nx, g.ny, g.nz, g.nb_buffer, g.nb_trunc, c->buffer_nnz, c->trunc_nnz,
                       c->max_buffer, c->max_core, c->max_trunc, c->ni, c->nj, c->used_nw, c->used_npc,
                       c->used_order};
      std::memcpy(info, v, sizeof(v));
    }
    if (geom) {
      double v[8] = {g.xmin, g.ymin, g.zmin, g.box, g.buffer_len, g.trunc_len, c->t_chol_s, c->chol_flops};
      std::memcpy(geom, v, sizeof(v));
    }
  });
}


prbf_status prbf_get_permutation(prbf_ctx* c, int32_t* pi, int32_t* pj) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    d2h(c, pi, c->perm_i, sizeof(int) * c->ni);
    d2h(c, pj, c->perm_j_(), sizeof(int) * c->nj);
    sync(c);
  });
}

prbf_status prbf_get_cells(prbf_ctx* c, int32_t* ista, int32_t* iend, int32_t* jsta, int32_t* jend) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    size_t b = sizeof(int) * c->nclus;
    d2h(c, ista, c->ista, b);
    d2h(c, iend, c->iend, b);
    d2h(c, jsta, c->jsta_(), b);
    d2h(c, jend, c->jend_(), b);
    sync(c);
  });
}

prbf_status prbf_get_buffer_csr(prbf_ctx* c, int64_t* ptr, int32_t* idx) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    d2h(c, ptr, c->bptr, sizeof(long long) * (c->nclus + 1));
    d2h(c, idx, c->bidx, sizeof(int) * c->buffer_nnz);
    sync(c);
  });
}

prbf_status prbf_get_trunc_csr(prbf_ctx* c, int64_t* ptr, int32_t* idx) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    d2h(c, ptr, c->tptr, sizeof(long long) * (c->nclus + 1));
    d2h(c, idx, c->tidx, sizeof(int) * c->trunc_nnz);
    sync(c);
  });
}

prbf_status prbf_get_sorted_points(prbf_ctx* c, double* x, double* y, double* z) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    d2h(c, x, c->xi, sizeof(double) * c->ni);
    d2h(c, y, c->yi, sizeof(double) * c->ni);
    if (c->prm.dim == 3) d2h(c, z, c->zi, sizeof(double) * c->ni);
    sync(c);
  });
}

prbf_status prbf_matvec_dev(prbf_ctx* c, int nrhs, const double* g, double* w) {
  return guarded(c, [&] { matvec_impl(c, nrhs, g, w); });
}

prbf_status prbf_matvec(prbf_ctx* c, int nrhs, const double* g, double* w) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    need(nrhs >= 1 && nrhs <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
    need(g && w, PRBF_ERR_ARG, "null vector");
    DBuf<double> dg, dw;
    dg.ensure((size_t)nrhs * c->nj);
    dw.ensure((size_t)nrhs * c->ni);
    PRBF_HIP(hipMemcpyAsync(dg, g, sizeof(double) * nrhs * c->nj, hipMemcpyHostToDevice, c->st));
    matvec_impl(c, nrhs, dg, dw);
    PRBF_HIP(hipMemcpyAsync(w, dw, sizeof(double) * nrhs * c->ni, hipMemcpyDeviceToHost, c->st));
    sync(c);
  });
}

static void evaluate_dev_impl(prbf_ctx* c, int nrhs, const double* gj, double* wi) {
  need(c->has_points, PRBF_ERR_STATE, "set_points first");
  need(nrhs >= 1 && nrhs <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
  need(gj && wi, PRBF_ERR_ARG, "null vector");
  DBuf<double> gs, ws;
  gs.ensure((size_t)nrhs * c->nj);
  ws.ensure((size_t)nrhs * c->ni);
  for (int r = 0; r < nrhs; ++r) gather(c->st, c->nj, c->perm_j_(), gj + (long long)r * c->nj, gs.p + (long long)r * c->nj);
  matvec_impl(c, nrhs, gs, ws);
  for (int r = 0; r < nrhs; ++r) scatter(c->st, c->ni, c->perm_i, ws.p + (long long)r * c->ni, wi + (long long)r * c->ni);
  sync(c);  // scratch buffers are released on return
}

prbf_status prbf_evaluate_dev(prbf_ctx* c, int nrhs, const double* gj, double* wi) {
  return guarded(c, [&] { evaluate_dev_impl(c, nrhs, gj, wi); });
}

prbf_status prbf_evaluate(prbf_ctx* c, int nrhs, const double* gj, double* wi) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    need(nrhs >= 1 && nrhs <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
    need(gj && wi, PRBF_ERR_ARG, "null vector");
    DBuf<double> dg, dw;
    dg.ensure((size_t)nrhs * c->nj);
    dw.ensure((size_t)nrhs * c->ni);
    PRBF_HIP(hipMemcpyAsync(dg, gj, sizeof(double) * nrhs * c->nj, hipMemcpyHostToDevice, c->st));
    evaluate_dev_impl(c, nrhs, dg, dw);
    PRBF_HIP(hipMemcpyAsync(wi, dw, sizeof(double) * nrhs * c->ni, hipMemcpyDeviceToHost, c->st));
    sync(c);
  });
}

prbf_status prbf_evaluate_dev_noalloc(prbf_ctx* c, int nrhs, const double* gj, double* wi) {
  return guarded(c, [&] {
    need(c->has_points, PRBF_ERR_STATE, "set_points first");
    need(nrhs >= 1 && nrhs <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
    need(gj && wi, PRBF_ERR_ARG, "null vector");
    need((long long)nrhs * (c->nj + c->ni) <= kScratchDoubles, PRBF_ERR_LIMIT,
         "nrhs (nj + ni) exceeds the 2^21 doubles of the allocation-free scratch; use prbf_evaluate_dev");
    double* gs = c->scratch;  // the device address resolved at prbf_create, never the array's name
    double* ws = gs + (long long)nrhs * c->nj;
    for (int r = 0; r < nrhs; ++r)
      gather(c->st, c->nj, c->perm_j_(), gj + (long long)r * c->nj, gs + (long long)r * c->nj);
    matvec_impl(c, nrhs, gs, ws);
    for (int r = 0; r < nrhs; ++r)
      scatter(c->st, c->ni, c->perm_i, ws + (long long)r * c->ni, wi + (long long)r * c->ni);
    PRBF_HIP(hipGetLastError());
  });
}

prbf_status prbf_rasm_setup(prbf_ctx* c) {
  return guarded(c, [&] { rasm_setup_impl(c, PRBF_PC_RASM_FP64); });
}

prbf_status prbf_rasm_setup_pc(prbf_ctx* c, int pc) {
  return guarded(c, [&] { rasm_setup_impl(c, pc); });
}

prbf_status prbf_set_tuning(prbf_ctx* c, int key, int value) {
  return guarded(c, [&] {
    if (key == PRBF_TUNE_CHOL_WAVES) {
      need(value == 4 || value == 6 || value == 8, PRBF_ERR_ARG, "waves per block factorisation must be 4, 6 or 8");
      c->chol_waves = value;
    } else if (key == PRBF_TUNE_CHOL_PANEL) {
      need(value >= 1 && value <= 4, PRBF_ERR_ARG, "panel width must be 1 to 4 tile columns");
      c->chol_panel = value;
    } else if (key == PRBF_TUNE_CHOL_ORDER) {
      need(value == PRBF_CHOL_RIGHT || value == PRBF_CHOL_LEFT, PRBF_ERR_ARG,
           "factorisation order must be PRBF_CHOL_RIGHT or PRBF_CHOL_LEFT");
      c->chol_order = value;
    } else {
      need(false, PRBF_ERR_ARG, "unknown tuning key");
    }
    c->tuning_forced = true;
  });
}

prbf_status prbf_rasm_apply_dev(prbf_ctx* c, const double* r, double* z) {
  return guarded(c, [&] { rasm_apply_impl(c, r, z); });
}

prbf_status prbf_rasm_apply(prbf_ctx* c, const double* r, double* z) {
  return guarded(c, [&] {
    need(c->rasm_ready, PRBF_ERR_STATE, "rasm_setup first");
    need(r && z, PRBF_ERR_ARG, "null vector");
    DBuf<double> dr, dz;
    dr.ensure(c->ni);
    dz.ensure(c->ni);
    PRBF_HIP(hipMemcpyAsync(dr, r, sizeof(double) * c->ni, hipMemcpyHostToDevice, c->st));
    rasm_apply_impl(c, dr, dz);
    PRBF_HIP(hipMemcpyAsync(z, dz, sizeof(double) * c->ni, hipMemcpyDeviceToHost, c->st));
    sync(c);
  });
}

prbf_status prbf_rasm_get_core_rows(prbf_ctx* c, int64_t k, double* M, int32_t* n, int32_t* T) {
  return guarded(c, [&] {
    need(c->rasm_ready && c->rasm_pc == PRBF_PC_RASM_FP64, PRBF_ERR_STATE, "FP64 rasm_setup first");
    need(k >= 0 && k < c->nclus, PRBF_ERR_ARG, "cluster index out of range");
    int nn = (int)(c->h_bptr[k + 1] - c->h_bptr[k]);
    int tt = nn ? c->h_iend[k] - c->h_ista[k] + 1 : 0;
    if (n) *n = nn;
    if (T) *T = tt;
    if (M && nn) {
      std::vector<double> st((size_t)apply_elems(nn, tt));
      d2h(c, st.data(), c->M.p + c->h_mptr64[k], sizeof(double) * st.size());
      sync(c);
      for (int t = 0; t < tt; ++t)
        for (int b = 0; b < nn; ++b) M[(size_t)t * nn + b] = st[apply_offset<2>(nn, tt, t, b)];
    }
  });
}

prbf_status prbf_rasm_get_block(prbf_ctx* c, int64_t k, double* A, int32_t* idx, int32_t* n) {
  return guarded(c, [&] {
    need(c->has_points && c->same, PRBF_ERR_STATE, "set_points (one set) first");
    need(k >= 0 && k < c->nclus, PRBF_ERR_ARG, "cluster index out of range");
    const long long b0 = c->h_bptr[k];
    const int nn = (int)(c->h_bptr[k + 1] - b0);
    if (n) *n = nn;
    if (idx && nn) d2h(c, idx, c->bidx.p + b0, sizeof(int) * nn);
    if (A && nn) {
      DBuf<double> dA;
      dA.ensure((size_t)nn * nn);
      dense_block(c->st, c->prm.dim, b0, nn, c->bidx, c->xi, c->yi, c->zi, c->prm.sigma, dA);
      PRBF_HIP(hipGetLastError());
      d2h(c, A, dA.p, sizeof(double) * nn * nn);
      sync(c);  // before the device copy is released
    }
    sync(c);
  });
}

prbf_status prbf_solve_dev(prbf_ctx* c, const prbf_solve_opts* opts, const double* w, double* gamma,
                           prbf_solve_info* info, double* hist, int hist_cap) {
  return guarded(c, [&] { solve_dev_impl(c, opts, w, gamma, info, hist, hist_cap); });
}

prbf_status prbf_solve_multi_dev(prbf_ctx* c, const prbf_solve_opts* opts, int nrhs, const double* w, double* gamma,
                                 prbf_solve_info* info, double* hist, int hist_cap) {
  return guarded(c, [&] { solve_multi_dev_impl(c, opts, nrhs, w, gamma, info, hist, hist_cap); });
}

prbf_status prbf_solve_multi(prbf_ctx* c, const prbf_solve_opts* opts, int nrhs, const double* w, double* gamma,
                             prbf_solve_info* info, double* hist, int hist_cap) {
  return guarded(c, [&] {
    need(c->has_points && c->same, PRBF_ERR_STATE, "set_points (one set) first");
    need(w && gamma, PRBF_ERR_ARG, "null vector");
    need(nrhs >= 1 && nrhs <= 3, PRBF_ERR_ARG, "nrhs must be 1..3");
    DBuf<double> dw, dg;
    dw.ensure((size_t)nrhs * c->ni);
    dg.ensure((size_t)nrhs * c->ni);
    PRBF_HIP(hipMemcpyAsync(dw, w, sizeof(double) * nrhs * c->ni, hipMemcpyHostToDevice, c->st));
    solve_multi_dev_impl(c, opts, nrhs, dw, dg, info, hist, hist_cap);
    PRBF_HIP(hipMemcpyAsync(gamma, dg, sizeof(double) * nrhs * c->ni, hipMemcpyDeviceToHost, c->st));
    sync(c);
  });
}

prbf_status prbf_solve(prbf_ctx* c, const prbf_solve_opts* opts, const double* w, double* gamma,
                       prbf_solve_info* info, double* hist, int hist_cap) {
  return guarded(c, [&] {
    need(c->has_points && c->same, PRBF_ERR_STATE, "set_points (one set) first");
    need(w && gamma, PRBF_ERR_ARG, "null vector");
    DBuf<double> dw, dg;
    dw.ensure(c->ni);
    dg.ensure(c->ni);
    PRBF_HIP(hipMemcpyAsync(dw, w, sizeof(double) * c->ni, hipMemcpyHostToDevice, c->st));
    solve_dev_impl(c, opts, dw, dg, info, hist, hist_cap);
    PRBF_HIP(hipMemcpyAsync(gamma, dg, sizeof(double) * c->ni, hipMemcpyDeviceToHost, c->st));
    sync(c);
  });
}

static prbf_status rbf_interpolation_f32_impl(const prbf_params* p, int64_t ng, const float* xg, const float* yg,
                                              const float* zg, float* wg, int64_t np_, const float* xp,
                                              const float* yp, const float* zp, const float* gp,
                                              const prbf_solve_opts* opts, int* its) {
  const int dim = p->dim;
  auto up = [](const float* a, int64_t n) {
    std::vector<double> v(a ? n : 0);
    for (int64_t i = 0; a && i < n; ++i) v[i] = a[i];
    return v;
  };
  std::vector<double> gx = up(xg, ng), gy = up(yg, ng), gz = up(dim == 3 ? zg : nullptr, ng), w = up(wg, ng);
  std::vector<double> g(ng);
  prbf_ctx* c = nullptr;
  prbf_status st = prbf_create(&c, p);
  if (st) {
    g_free_err = "prbf_create failed";
    return st;
  }
  struct Owner {  // destroys the context on every way out, exceptions included
    prbf_ctx* c;
    ~Owner() { prbf_destroy(c); }
  } owner{c};
  const bool remesh = np_ > 0 && xp && std::fabs((double)xg[0] - (double)xp[0]) > kEpsf;
  if (remesh) {
    std::vector<double> px = up(xp, np_), py = up(yp, np_), pz = up(dim == 3 ? zp : nullptr, np_), pg = up(gp, np_);
    st = prbf_set_points(c, ng, gx.data(), gy.data(), dim == 3 ? gz.data() : nullptr, np_, px.data(), py.data(),
                         dim == 3 ? pz.data() : nullptr);
    if (!st) st = prbf_evaluate(c, 1, pg.data(), w.data());
  }
  prbf_solve_info info{};
  if (!st) st = prbf_set_points(c, ng, gx.data(), gy.data(), dim == 3 ? gz.data() : nullptr, 0, nullptr, nullptr, nullptr);
  if (!st) st = prbf_solve(c, opts, w.data(), g.data(), &info, nullptr, 0);
  if (st) {
    g_free_err = c->err;
  } else {
    for (int64_t i = 0; i < ng; ++i) wg[i] = (float)g[i];
    if (its) *its = info.its;
  }
  return st;
}

// RBFInterpolation (3d/wrapper.cxx:20-105): FP32 grid values in, grid
// strengths out.  When the grid is not the particle set (first coordinates
// differ by more than 1e-6, 3d/wrapper.cxx:93) the field of the particles
// (strengths gp) is first evaluated at the grid; then the grid strengths
// solving A~ g = w are written back over wg.  Host allocation failures return
// PRBF_ERR_NOMEM; no exception leaves the call.
prbf_status prbf_rbf_interpolation_f32(const prbf_params* p, int64_t ng, const float* xg, const float* yg,
                                       const float* zg, float* wg, int64_t np_, const float* xp, const float* yp,
                                       const float* zp, const float* gp, const prbf_solve_opts* opts, int* its) {
  if (!p || ng <= 0 || !xg || !yg || !wg || (p->dim == 3 && !zg) || (np_ > 0 && xp && (!yp || !gp)) ||
      (np_ > 0 && xp && p->dim == 3 && !zp)) {
    g_free_err = "invalid arguments";
    return PRBF_ERR_ARG;
  }
  try {
    return rbf_interpolation_f32_impl(p, ng, xg, yg, zg, wg, np_, xp, yp, zp, gp, opts, its);
  } catch (const std::bad_alloc&) {
    g_free_err = "host allocation failed";
    return PRBF_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_free_err = e.what();
    return PRBF_ERR_ARG;
  } catch (...) {
    g_free_err = "unexpected error";
    return PRBF_ERR_HIP;
  }
}

// ------------------------------------------------------------ learned preconditioner

}  // extern "C"

namespace {

template <class F>
prbf_status guarded_mlp(prbf_mlp* net, F&& f) {
  if (!net) return PRBF_ERR_ARG;
  try {
    f();
    net->err.clear();
    return PRBF_OK;
  } catch (const ApiError& e) {
    net->err = e.what;
    return (prbf_status)e.code;
  } catch (const HipError& e) {
    net->err = e.what;
    return PRBF_ERR_HIP;
  } catch (const std::bad_alloc&) {
    net->err = "host allocation failed";
    return PRBF_ERR_NOMEM;
  } catch (const std::exception& e) {
    net->err = e.what();
    return PRBF_ERR_ARG;
  }
}

void mlp_sync(prbf_mlp* net) { PRBF_HIP(hipStreamSynchronize(net->st)); }

void mlp_refresh_bf16(prbf_mlp* net) { mlp_weights_bf16(net->st, net->q, net->P, net->Wb, net->Wt); }

// Patch slots of every cluster and the list of eligible clusters.
void ensure_patches(prbf_ctx* c) {
  need(c->has_points && c->same, PRBF_ERR_STATE, "set_points (one set) first");
  need(!c->tree, PRBF_ERR_ARG, "the learned preconditioner works on the uniform grid's clusters");
  if (c->patches_ready) return;
  const LGeo q = lgeo(c->prm.dim);
  const long long nc = c->g.ncell;
  c->pidx.ensure((size_t)nc * q.np);
  c->cidx.ensure((size_t)nc * q.nc);
  c->pok.ensure((size_t)nc);
  c->pctr.ensure((size_t)3 * nc);
  patch_build(c->st, c->g, q, c->g.box / q.cs, c->ista, c->iend, c->xi, c->yi, c->zi, c->pidx, c->cidx, c->pctr,
              c->pok);
  PRBF_HIP(hipGetLastError());
  std::vector<unsigned char> h(nc);
  PRBF_HIP(hipMemcpyAsync(h.data(), c->pok, nc, hipMemcpyDeviceToHost, c->st));
  sync(c);
  std::vector<int> e;
  for (long long k = 0; k < nc; ++k)
    if (h[k]) e.push_back((int)k);
  c->n_elig = (long long)e.size();
  c->elig.ensure(std::max<size_t>(1, e.size()));
  if (!e.empty())
    PRBF_HIP(hipMemcpyAsync(c->elig, e.data(), sizeof(int) * e.size(), hipMemcpyHostToDevice, c->st));
  sync(c);
  c->patches_ready = true;
}

}  // namespace

extern "C" {

int64_t prbf_mlp_num_params(void) { return mlp_nparams(lgeo(2)); }

int64_t prbf_mlp_num_params_dim(int dim) { return dim == 2 || dim == 3 ? mlp_nparams(lgeo(dim)) : -1; }

prbf_status prbf_mlp_create(prbf_mlp** out, uint64_t seed) { return prbf_mlp_create_dim(out, seed, 2); }

prbf_status prbf_mlp_create_dim(prbf_mlp** out, uint64_t seed, int dim) {
  if (!out) return PRBF_ERR_ARG;
  *out = nullptr;
  if (dim != 2 && dim != 3) return PRBF_ERR_ARG;
  prbf_mlp* net = new (std::nothrow) prbf_mlp;
  if (!net) return PRBF_ERR_NOMEM;
  net->q = lgeo(dim);
  if (hipStreamCreateWithFlags(&net->st, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete net;
    return PRBF_ERR_HIP;
  }
  prbf_status st = guarded_mlp(net, [&] {
    const LGeo& q = net->q;
    const long long np = mlp_nparams(q);
    // Glorot-uniform weights, zero biases, from a splitmix64 stream (3D: the
    // weights into the pad outputs are zero, so those outputs start at zero)
    std::vector<float> h(np, 0.0f);
    unsigned long long z = seed;
    auto next = [&]() {
      z += 0x9E3779B97F4A7C15ULL;
      unsigned long long v = z;
      v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
      v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
      v ^= v >> 31;
      return (double)(v >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < kMlpLayers; ++l) {
      int K, N;
      mlp_dims(q, l, &K, &N);
      const double a = std::sqrt(6.0 / (K + (l == kMlpLayers - 1 ? q.nc : N)));
      for (long long i = 0; i < (long long)K * N; ++i) {
        const double v = (2.0 * next() - 1.0) * a;
        h[mlp_w_off(q, l) + i] = (l == kMlpLayers - 1 && i % N >= q.nc) ? 0.0f : (float)v;
      }
    }
    net->P.ensure(np);
    net->G.ensure(np);
    net->m1.ensure(np);
    net->m2.ensure(np);
    net->Wb.ensure(np);
    net->Wt.ensure(np);
    PRBF_HIP(hipMemcpyAsync(net->P, h.data(), sizeof(float) * np, hipMemcpyHostToDevice, net->st));
    PRBF_HIP(hipMemsetAsync(net->m1, 0, sizeof(float) * np, net->st));
    PRBF_HIP(hipMemsetAsync(net->m2, 0, sizeof(float) * np, net->st));
    mlp_refresh_bf16(net);
    mlp_sync(net);
  });
  if (st) {
    prbf_mlp_destroy(net);
    return st;
  }
  *out = net;
  return PRBF_OK;
}

prbf_status prbf_mlp_destroy(prbf_mlp* net) {
  if (!net) return PRBF_ERR_ARG;
  if (net->st) {
    (void)hipStreamSynchronize(net->st);
    (void)hipStreamDestroy(net->st);
  }
  delete net;
  return PRBF_OK;
}

const char* prbf_mlp_last_error(const prbf_mlp* net) { return net ? net->err.c_str() : ""; }

prbf_status prbf_mlp_set_params(prbf_mlp* net, const float* params, double out_scale) {
  return guarded_mlp(net, [&] {
    need(params != nullptr && out_scale > 0, PRBF_ERR_ARG, "params and a positive out_scale required");
    const long long np = mlp_nparams(net->q);
    PRBF_HIP(hipMemcpyAsync(net->P, params, sizeof(float) * np, hipMemcpyHostToDevice, net->st));
    PRBF_HIP(hipMemsetAsync(net->m1, 0, sizeof(float) * np, net->st));
    PRBF_HIP(hipMemsetAsync(net->m2, 0, sizeof(float) * np, net->st));
    net->t = 0;
    net->out_scale = out_scale;
    mlp_refresh_bf16(net);
    mlp_sync(net);
  });
}

prbf_status prbf_mlp_get_params(prbf_mlp* net, float* params, double* out_scale) {
  return guarded_mlp(net, [&] {
    if (params)
      PRBF_HIP(hipMemcpyAsync(params, net->P, sizeof(float) * mlp_nparams(net->q), hipMemcpyDeviceToHost, net->st));
    if (out_scale) *out_scale = net->out_scale;
    mlp_sync(net);
  });
}

prbf_status prbf_mlp_forward_dev(prbf_mlp* net, int64_t m, const float* X, float* Y) {
  return guarded_mlp(net, [&] {
    need(m > 0 && X && Y, PRBF_ERR_ARG, "invalid arguments");
    net->ensure(m);
    float* pr[3];
    unsigned short* ac[3];
    net->layers(pr, ac);
    mlp_to_bf16(net->st, m * net->q.nin, X, net->Xb);
    mlp_forward(net->st, net->q, (int)m, net->P, net->Wt, net->Xb, pr, ac, net->Y);
    PRBF_HIP(hipMemcpyAsync(Y, net->Y, sizeof(float) * m * net->q.nout, hipMemcpyDeviceToDevice, net->st));
    PRBF_HIP(hipGetLastError());
    mlp_sync(net);
  });
}

prbf_status prbf_mlp_loss_grad_dev(prbf_mlp* net, int64_t m, const float* X, const float* T, float* loss,
                                   float* grad) {
  return guarded_mlp(net, [&] {
    need(m > 0 && m % 16 == 0 && X && T, PRBF_ERR_ARG, "m must be a positive multiple of 16");
    net->ensure(m);
    float* pr[3];
    unsigned short* ac[3];
    net->layers(pr, ac);
    mlp_to_bf16(net->st, m * net->q.nin, X, net->Xb);
    mlp_forward(net->st, net->q, (int)m, net->P, net->Wt, net->Xb, pr, ac, net->Y);
    mlp_backward(net->st, net->q, (int)m, net->P, net->Wb, net->Xb, pr, ac, net->Y, T, net->G, net->dYf, net->dYb,
                 net->dHf, net->dHb, net->part, net->loss);
    if (grad)
      PRBF_HIP(hipMemcpyAsync(grad, net->G, sizeof(float) * mlp_nparams(net->q), hipMemcpyDeviceToDevice, net->st));
    float h = 0.0f;
    PRBF_HIP(hipMemcpyAsync(&h, net->loss, sizeof(float), hipMemcpyDeviceToHost, net->st));
    PRBF_HIP(hipGetLastError());
    mlp_sync(net);
    if (loss) *loss = h;
  });
}

prbf_status prbf_mlp_train_dev(prbf_mlp* net, int64_t npairs, const float* X, const float* T, int batch, int steps,
                               float lr, uint64_t seed, float* loss_hist) {
  return guarded_mlp(net, [&] {
    need(npairs > 0 && X && T && batch > 0 && batch % 16 == 0 && steps >= 0, PRBF_ERR_ARG,
         "batch must be a positive multiple of 16");
    net->ensure(batch);
    net->lhist.ensure(std::max(1, steps));
    float* pr[3];
    unsigned short* ac[3];
    net->layers(pr, ac);
    const long long np = mlp_nparams(net->q);
    for (int s = 0; s < steps; ++s) {
      mlp_random_rows(net->st, batch, npairs, seed, s, net->idx);
      mlp_gather_rows(net->st, net->q, batch, net->idx, X, T, net->Xb, net->Tb);
      mlp_forward(net->st, net->q, batch, net->P, net->Wt, net->Xb, pr, ac, net->Y);
      mlp_backward(net->st, net->q, batch, net->P, net->Wb, net->Xb, pr, ac, net->Y, net->Tb, net->G, net->dYf, net->dYb,
                   net->dHf, net->dHb, net->part, net->loss);
      copy_f32(net->st, net->loss, net->lhist.p + s);
      mlp_adam(net->st, np, net->P, net->G, net->m1, net->m2, lr, ++net->t);
      mlp_refresh_bf16(net);
    }
    PRBF_HIP(hipGetLastError());
    if (loss_hist && steps > 0)
      PRBF_HIP(hipMemcpyAsync(loss_hist, net->lhist, sizeof(float) * steps, hipMemcpyDeviceToHost, net->st));
    mlp_sync(net);
  });
}

prbf_status prbf_learned_eligible(prbf_ctx* c, int64_t* n) {
  return guarded(c, [&] {
    ensure_patches(c);
    if (n) *n = c->n_elig;
  });
}

prbf_status prbf_learned_pairs_dev(prbf_ctx* c, const double* r, const double* z, double out_scale, float* X,
                                   float* T, double* s, int64_t cap, int64_t* n) {
  return guarded(c, [&] {
    need(r && X && s && cap >= 0 && out_scale > 0 && (!z || T), PRBF_ERR_ARG, "invalid arguments");
    ensure_patches(c);
    const long long m = std::min<long long>(cap, c->n_elig);
    const LGeo q = lgeo(c->prm.dim);
    patch_features(c->st, q, m, c->elig, c->g.box / q.cs, c->prm.sigma, c->pidx, c->cidx, c->pctr, c->xi, c->yi,
                   c->zi, r, z, out_scale, X, T, s);
    PRBF_HIP(hipGetLastError());
    if (n) *n = m;
  });
}

prbf_status prbf_learned_rho_dev(prbf_ctx* c, const double* r, const double* z, double* rho, int64_t cap,
                                 int64_t* n) {
  return guarded(c, [&] {
    need(r && z && rho && cap >= 0, PRBF_ERR_ARG, "invalid arguments");
    ensure_patches(c);
    const long long m = std::min<long long>(cap, c->n_elig);
    patch_rho(c->st, lgeo(c->prm.dim), m, c->elig, c->pidx, c->cidx, c->xi, c->yi, c->zi, r, z, c->prm.sigma, rho);
    PRBF_HIP(hipGetLastError());
    if (n) *n = m;
  });
}

prbf_status prbf_learned_attach(prbf_ctx* c, prbf_mlp* net, double tau, int64_t* n_network) {
  return guarded(c, [&] {
    need(net != nullptr && tau >= 0, PRBF_ERR_ARG, "a network and tau >= 0 required");
    need(net->q.dim == c->prm.dim, PRBF_ERR_ARG, "the network's dimension differs from the context's");
    ensure_patches(c);
    c->learned_ready = false;
    c->net = net;
    const long long ne = c->n_elig, n = c->ni, nc = c->g.ncell;
    std::vector<unsigned char> accept(ne, 1);
    if (ne > 0) {
      mlp_sync(net);  // the network's own work is done before the context's stream uses it
      net->ensure(ne);
      c->lX.ensure((size_t)ne * net->q.nin);
      c->ls.ensure((size_t)ne);
      DBuf<double> g, r, z, rho;
      g.ensure(n);
      r.ensure(n);
      z.ensure(n);
      rho.ensure(ne);
      std::vector<double> hg(n), hr(ne);
      for (int probe = 0; probe < 2; ++probe) {
        // probe residual r = A g for a seeded g uniform in (-1, 1)
        unsigned long long st = 0x5EED5EEDULL + probe;
        for (long long i = 0; i < n; ++i) {
          st += 0x9E3779B97F4A7C15ULL;
          unsigned long long v = st;
          v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
          v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
          v ^= v >> 31;
          hg[i] = 2.0 * ((double)(v >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
        }
        PRBF_HIP(hipMemcpyAsync(g, hg.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->st));
        matvec_impl(c, 1, g, r);
        fill(c->st, n, z, 0.0);
        network_apply(c, c->elig, ne, r, z);
        patch_rho(c->st, net->q, ne, c->elig, c->pidx, c->cidx, c->xi, c->yi, c->zi, r, z, c->prm.sigma, rho);
        PRBF_HIP(hipMemcpyAsync(hr.data(), rho, sizeof(double) * ne, hipMemcpyDeviceToHost, c->st));
        sync(c);
        for (long long k = 0; k < ne; ++k)
          if (!(hr[k] <= tau)) accept[k] = 0;
      }
    }
    std::vector<int> he(std::max<long long>(ne, 1));
    if (ne > 0) PRBF_HIP(hipMemcpy(he.data(), c->elig, sizeof(int) * ne, hipMemcpyDeviceToHost));
    c->h_skip.assign(nc, 0);
    std::vector<int> list;
    for (long long k = 0; k < ne; ++k)
      if (accept[k]) {
        list.push_back(he[k]);
        c->h_skip[he[k]] = 1;
      }
    c->n_mlp = (long long)list.size();
    c->h_mlp_list = list;
    c->n_net_fallback = 0;
    c->lnorm.ensure(std::max<size_t>(1, list.size()));
    c->mlp_list.ensure(std::max<size_t>(1, list.size()));
    if (!list.empty())
      PRBF_HIP(hipMemcpyAsync(c->mlp_list, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, c->st));
    sync(c);
    rasm_setup_impl(c, PRBF_PC_LEARNED);
    c->learned_ready = true;
    if (n_network) *n_network = c->n_mlp;
  });
}

prbf_status prbf_krylov_vectors_dev(prbf_ctx* c, int j, double* v, double* z) {
  return guarded(c, [&] {
    need(c->last_restart > 0 && j >= 0 && j < c->last_restart, PRBF_ERR_STATE, "no such basis vector");
    need(c->Z.p != nullptr && c->Z.n >= (size_t)c->last_restart * c->ni, PRBF_ERR_STATE, "run an FGMRES solve first");
    const size_t b = sizeof(double) * c->ni;
    if (v) PRBF_HIP(hipMemcpyAsync(v, c->V.p + (long long)j * c->ni, b, hipMemcpyDeviceToDevice, c->st));
    if (z) PRBF_HIP(hipMemcpyAsync(z, c->Z.p + (long long)j * c->ni, b, hipMemcpyDeviceToDevice, c->st));
  });
}

}  // extern "C"
