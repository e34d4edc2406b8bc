// Shared definitions for the petrbf HIP library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
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
