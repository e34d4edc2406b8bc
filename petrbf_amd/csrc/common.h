// Shared definitions for the petrbf HIP library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

namespace prbf {

constexpr double kEpsf = 1e-6;  // absolute slack on box widths (2d/par.h:153)
constexpr int kWave = 64;       // CDNA wavefront width

// Geometry of the uniform cell grid, by value into kernels.
struct Grid {
  int dim;
  int nx, ny, nz;
  long long ncell;
  double xmin, ymin, zmin;  // grid origin = bbox min - eps (2d/get_cluster.h:25-28)
  double box;               // cell side = nsigma_box*sigma + eps (2d/get_cluster.h:29)
  int nb_buffer, nb_trunc;  // neighbour reach in cells (2d/get_cluster.h:20-21)
  double buffer_len;        // buffer half-width (2d/get_buffer.h:11)
  double trunc_len;         // truncation half-width (2d/get_trunc.h:14)
};

__host__ __device__ inline long long cell_index(const Grid& g, int ix, int iy, int iz) {
  // 2d/get_cluster.h:78-80; 3d/get_cluster.h:101
  return g.dim == 2 ? (long long)ix * g.ny + iy
                    : (long long)iz * g.nx * g.ny + (long long)ix * g.ny + iy;
}

__host__ __device__ inline void cell_coords(const Grid& g, long long c, int* ix, int* iy, int* iz) {
  if (g.dim == 2) {
    *iz = 0;
    *ix = (int)(c / g.ny);
    *iy = (int)(c % g.ny);
  } else {
    long long nxy = (long long)g.nx * g.ny;
    *iz = (int)(c / nxy);
    long long r = c % nxy;
    *ix = (int)(r / g.ny);
    *iy = (int)(r % g.ny);
  }
}

// Cell centre, evaluated as xmin + (ix+0.5)*box with two roundings
// (2d/get_cluster.h:55-72); contraction into an fma would change the bits.
__host__ __device__ inline double cell_centre(double xmin, int k, double box) {
#pragma clang fp contract(off)
  double off = (k + 0.5) * box;
  return xmin + off;
}

// The dispatcher deals workgroups round-robin over the 8 XCDs (workgroup b
// runs on XCD b % 8), each with its own L2.  This maps workgroup b to a
// logical index so that every XCD receives one contiguous range of logical
// indices (neighbouring clusters share the buffer and truncation particles
// they gather); a bijection on [0, n) for any n.
constexpr unsigned kXcds = 8;
__device__ inline unsigned xcd_block(unsigned b, unsigned n) {
  const unsigned q = n / kXcds, r = n % kXcds;
  const unsigned x = b % kXcds, k = b / kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Gaussian normalisation 1/(2 pi sigma^2) as the reference writes it,
// exp(-r2/(2 sigma sigma)) / (2 pi sigma sigma) (2d/matmult.cxx:39-40); the
// 3D tree uses the same 2D constant (3d/matmult.cxx:40-41).
struct Kern {
  double two_s2;    // 2*sigma*sigma
  double inv_two_s2;
  double norm_div;  // 2*pi*sigma*sigma
  double norm;      // 1/(2 pi sigma^2)
};

inline Kern make_kern(double sigma) {
  Kern k;
  k.two_s2 = 2 * sigma * sigma;
  k.inv_two_s2 = 1.0 / k.two_s2;
  k.norm_div = 2 * M_PI * sigma * sigma;
  k.norm = 1.0 / k.norm_div;
  return k;
}

// exp(x) for -745 < x <= 0 in FP64: Cody-Waite reduction x = k ln2 + r with a
// two-part ln2 (the high part has 33 significant bits, so k*ln2_hi is exact
// for |k| < 2^20), |r| <= ln2/2, then the degree-13 Taylor polynomial of e^r
// (truncation error < 1e-17 relative) and a scaling by 2^k.  No range checks:
// every caller passes -r^2/(2 sigma^2) of points inside a cluster's boxes.
__device__ inline double exp_nonpos(double x) {
  const double kf = rint(x * 1.4426950408889634);
  double r = fma(-kf, 0.6931471804855391, x);
  r = fma(-kf, 7.440617110012397e-11, r);
  double p = 1.6059043836821613e-10;
  p = fma(p, r, 2.08767569878681e-09);
  p = fma(p, r, 2.505210838544172e-08);
  p = fma(p, r, 2.755731922398589e-07);
  p = fma(p, r, 2.7557319223985893e-06);
  p = fma(p, r, 2.48015873015873e-05);
  p = fma(p, r, 0.0001984126984126984);
  p = fma(p, r, 0.001388888888888889);
  p = fma(p, r, 0.008333333333333333);
  p = fma(p, r, 0.041666666666666664);
  p = fma(p, r, 0.16666666666666666);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)kf);
}

// 2^(y/256) for -2^20 < y <= 0 in FP64, for an argument already in units of
// ln2/256 (the matvec scales its coordinates so that y = -|dx|^2 (256/ln2) /
// (2 sigma^2) comes straight out of the squared distance): k = rint(y) by the
// 1.5 * 2^52 shift (adding it leaves k in the sum's low word, |k| < 2^31, in
// place of a round and a conversion), r = y - k EXACTLY (|r| <= 1/2,
// Sterbenz), 2^(r/256) by its degree-4 Taylor polynomial in r (coefficients
// (ln2/256)^n / n!, correctly rounded from 60-digit decimal arithmetic;
// truncation < 4e-17 relative), times tab[k & 255] = 2^((k & 255)/256) and
// 2^(k >> 8).  With no reduction constant there is no reduction error: the
// result's error is the polynomial's and the two products' roundings (9 FP64
// operations; a 64-entry table needs a degree-5 polynomial for the same
// truncation, a table-driven exp of x = -r^2/(2 sigma^2) 11, two of them for
// a two-part ln2/64 reduction).
__device__ inline double exp2_256_nonpos(double y, const double* tab) {
  constexpr double kShift = 6755399441055744.0;  // 1.5 * 2^52
  const double kd = y + kShift;
  const int k = __double2loint(kd);
  const double r = y - (kd - kShift);
  double p = fma(r, 0x1.3b2ab6fba4e77p-39, 0x1.c6b08d704a0c0p-29);
  p = fma(p, r, 0x1.ebfbdff82c58fp-19);
  p = fma(p, r, 0x1.62e42fefa39efp-9);
  p = fma(p, r, 1.0);
  return ldexp(p * tab[k & 255], k >> 8);
}

}  // namespace prbf

#define PRBF_HIP(call)                                                         \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess) {                                                    \
      throw prbf::HipError(_e, #call, __FILE__, __LINE__);                     \
    }                                                                          \
  } while (0)

namespace prbf {
struct HipError {
  hipError_t err;
  std::string what;
  HipError(hipError_t e, const char* call, const char* file, int line) : err(e) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", call, hipGetErrorString(e), file, line);
    what = buf;
  }
};
struct ApiError {
  int code;
  std::string what;
  ApiError(int c, std::string w) : code(c), what(std::move(w)) {}
};
}  // namespace prbf
