/* A plain C host program on the C ABI: the reference's 2D driver sequence
 * (2d/main.cxx: particles and RHS in, rbf_interpolation, gamma and the
 * iteration count out) with host buffers.
 *
 *   c1_caller <in.bin> <out.bin>
 * in.bin:  int64 n; double sigma, nsigma_box, sigma_buffer, sigma_trunc;
 *          double x[n], y[n], w[n]
 * out.bin: int32 status, its, reason; double gamma[n]; double Ag[n]
 *          (Ag = the operator applied to gamma through prbf_evaluate)
 */
#include <stdio.h>
#include <stdlib.h>

#include "petrbf.h"

static int rd(FILE* f, void* p, size_t sz, size_t n) { return fread(p, sz, n, f) == n; }

int main(int argc, char** argv) {
  if (argc != 3) return 64;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 66;
  int64_t n;
  double prm[4];
  if (!rd(f, &n, sizeof n, 1) || n <= 0 || !rd(f, prm, sizeof(double), 4)) return 65;
  double* x = malloc(sizeof(double) * n);
  double* y = malloc(sizeof(double) * n);
  double* w = malloc(sizeof(double) * n);
  double* g = malloc(sizeof(double) * n);
  double* ag = malloc(sizeof(double) * n);
  if (!x || !y || !w || !g || !ag) return 71;
  if (!rd(f, x, sizeof(double), n) || !rd(f, y, sizeof(double), n) || !rd(f, w, sizeof(double), n)) return 65;
  fclose(f);

  prbf_params p = {2, prm[0], prm[1], prm[2], prm[3]};
  prbf_ctx* ctx = NULL;
  prbf_solve_opts o;
  prbf_solve_info info = {0};
  prbf_solve_opts_default(&o); /* GMRES(30), FP64 RASM, rtol 1e-10 */
  int st = prbf_create(&ctx, &p);
  if (st == PRBF_OK) st = prbf_set_points(ctx, n, x, y, NULL, 0, NULL, NULL, NULL);
  if (st == PRBF_OK) st = prbf_solve(ctx, &o, w, g, &info, NULL, 0);
  if (st == PRBF_OK) st = prbf_evaluate(ctx, 1, g, ag);
  if (st != PRBF_OK) fprintf(stderr, "c1_caller: status %d: %s\n", st, prbf_last_error(ctx));
  if (ctx) prbf_destroy(ctx);

  FILE* out = fopen(argv[2], "wb");
  if (!out) return 73;
  int32_t hdr[3] = {st, info.its, info.reason};
  fwrite(hdr, sizeof hdr[0], 3, out);
  fwrite(g, sizeof(double), n, out);
  fwrite(ag, sizeof(double), n, out);
  fclose(out);
  free(x); free(y); free(w); free(g); free(ag);
  return st == PRBF_OK ? 0 : 1;
}
