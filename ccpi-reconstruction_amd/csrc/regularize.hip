// Regularisation terms of the regularised CGLS: Tikhonov and TV.
//
// Values follow Core/src/tikhonov.cpp:4-32 and Core/src/total_v.cpp:21-195
// operation for operation:
//   * a neighbour past a face is mirrored to the other side (i+1 at the
//     last plane reads i-1, i-1 at the first reads i+1), so every extent
//     must be at least 2;
//   * Tikhonov: float sum of the four x/y neighbours, scaled by 0.25 and
//     differenced in double, stored in float; z is not coupled;
//   * TV: forward differences, min-mod terms formed in double and stored in
//     float, T = sqrt(own^2 + minmod^2 + minmod^2 + FLT_EPSILON) in float,
//     D = diff / T, then the backward divergence of D.  Square roots and
//     quotients are taken in double and rounded once to float, which yields
//     the correctly rounded float result whatever the compiler's float
//     division / sqrt mode.
// Layout: [nx][ny][nz], z contiguous.  One thread per (j, k) of a plane,
// planes strided over blockIdx.y, so x / y neighbour reads hit the lines
// that the neighbouring planes and rows fetch.  TV runs as two passes
// through a 3-volume scratch array (D1, D2, D3); the terms are a few ms at
// 512^3 against ~180 ms of projections per CGLS iteration.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.hpp"
#include "solvers.hpp"

namespace ccpi {

namespace {

constexpr int kRegThreads = 256;

__device__ __forceinline__ int mirror_up(int i, int n) { return i + 1 < n ? i + 1 : i - 1; }
__device__ __forceinline__ int mirror_dn(int i) { return i > 0 ? i - 1 : i + 1; }

__global__ __launch_bounds__(kRegThreads) void tikhonov_kernel(const float *__restrict__ a,
                                                               float *__restrict__ b, int nx,
                                                               int ny, int nz) {
  const int plane = ny * nz;
  const int r = blockIdx.x * kRegThreads + threadIdx.x;
  if (r >= plane) return;
  const int j = r / nz, k = r - j * nz;
  const int rj1 = mirror_up(j, ny) * nz + k, rj2 = mirror_dn(j) * nz + k;
  for (int i = blockIdx.y; i < nx; i += gridDim.y) {
    const float *ai = a + (size_t)i * plane;
    const float s = a[(size_t)mirror_up(i, nx) * plane + r] +
                    a[(size_t)mirror_dn(i) * plane + r] + ai[rj1] + ai[rj2];
    b[(size_t)i * plane + r] = (float)(0.25 * (double)s - (double)ai[r]);
  }
}

// (0.5 (sign p + sign m) min(|p|, |m|))^2, the product in double, stored
// in float before squaring (total_v.cpp:117-139)
__device__ __forceinline__ float minmod_sq(float p, float m) {
  const int sp = (p > 0.f) - (p < 0.f), sm = (m > 0.f) - (m < 0.f);
  const float ap = fabsf(p), am = fabsf(m);
  const float v = (float)(0.5 * (double)(sp + sm) * (double)(am < ap ? am : ap));
  return v * v;
}

__device__ __forceinline__ float tv_ratio(float num, float t) {
  const float T = (float)sqrt((double)t);
  return (float)((double)num / (double)T);
}

// D1 = dx+/T1, D2 = dy+/T2, D3 = dz+/T3 (total_v.cpp:86-153)
__global__ __launch_bounds__(kRegThreads) void tv_d_kernel(const float *__restrict__ a,
                                                           float *__restrict__ D, size_t nvox,
                                                           int nx, int ny, int nz) {
  const int plane = ny * nz;
  const int r = blockIdx.x * kRegThreads + threadIdx.x;
  if (r >= plane) return;
  const int j = r / nz, k = r - j * nz;
  const int rj1 = mirror_up(j, ny) * nz + k, rj2 = mirror_dn(j) * nz + k;
  const int rk1 = j * nz + mirror_up(k, nz), rk2 = j * nz + mirror_dn(k);
  for (int i = blockIdx.y; i < nx; i += gridDim.y) {
    const size_t c = (size_t)i * plane + r;
    const float *ai = a + (size_t)i * plane;
    const float A = ai[r];
    const float x1 = a[(size_t)mirror_up(i, nx) * plane + r] - A;
    const float x0 = A - a[(size_t)mirror_dn(i) * plane + r];
    const float y1 = ai[rj1] - A, y0 = A - ai[rj2];
    const float z1 = ai[rk1] - A, z0 = A - ai[rk2];
    const float mx = minmod_sq(x1, x0), my = minmod_sq(y1, y0), mz = minmod_sq(z1, z0);
    D[c] = tv_ratio(x1, x1 * x1 + my + mz + FLT_EPSILON);
    D[nvox + c] = tv_ratio(y1, y1 * y1 + mx + mz + FLT_EPSILON);
    D[2 * nvox + c] = tv_ratio(z1, z1 * z1 + mx + my + FLT_EPSILON);
  }
}

// b = (D1 - D1(i-1)) + (D2 - D2(j-1)) + (D3 - D3(k-1)) (total_v.cpp:157-195)
__global__ __launch_bounds__(kRegThreads) void tv_div_kernel(const float *__restrict__ D,
                                                             float *__restrict__ b, size_t nvox,
                                                             int nx, int ny, int nz) {
  const int plane = ny * nz;
  const int r = blockIdx.x * kRegThreads + threadIdx.x;
  if (r >= plane) return;
  const int j = r / nz, k = r - j * nz;
  const int rj2 = mirror_dn(j) * nz + k, rk2 = j * nz + mirror_dn(k);
  const float *D1 = D, *D2 = D + nvox, *D3 = D + 2 * nvox;
  for (int i = blockIdx.y; i < nx; i += gridDim.y) {
    const size_t base = (size_t)i * plane, c = base + r;
    const float dv1 = D1[c] - D1[(size_t)mirror_dn(i) * plane + r];
    const float dv2 = D2[c] - D2[base + rj2];
    const float dv3 = D3[c] - D3[base + rk2];
    b[c] = dv1 + dv2 + dv3;
  }
}

dim3 reg_grid(int nx, int ny, int nz) {
  const int plane = ny * nz;
  return dim3((unsigned)((plane + kRegThreads - 1) / kRegThreads),
              (unsigned)std::min(nx, 4096));
}

}  // namespace

void check_reg_dims(int nx, int ny, int nz) {
  CCPI_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, CCPI_E_ARG,
               "regularisation terms mirror neighbours: every volume extent must be >= 2");
  CCPI_REQUIRE((long long)ny * nz < (1LL << 31), CCPI_E_LIMIT,
               "regularisation terms: ny * nz must be below 2^31");
}

size_t reg_scratch_floats(int kind, size_t nvox) { return kind == kRegTV ? 3 * nvox : 0; }

void reg_term(int kind, const float *a, float *out, float *scratch, int nx, int ny, int nz,
              hipStream_t st) {
  check_reg_dims(nx, ny, nz);
  const dim3 grid = reg_grid(nx, ny, nz);
  if (kind == kRegTikhonov) {
    hipLaunchKernelGGL(tikhonov_kernel, grid, dim3(kRegThreads), 0, st, a, out, nx, ny, nz);
  } else {
    CCPI_REQUIRE(kind == kRegTV, CCPI_E_ARG, "unknown regularisation kind");
    CCPI_REQUIRE(scratch != nullptr, CCPI_E_ARG, "TV needs a 3-volume scratch array");
    const size_t nvox = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(tv_d_kernel, grid, dim3(kRegThreads), 0, st, a, scratch, nvox, nx, ny,
                       nz);
    hipLaunchKernelGGL(tv_div_kernel, grid, dim3(kRegThreads), 0, st, scratch, out, nvox, nx,
                       ny, nz);
  }
  CCPI_HIP(hipGetLastError());
}

}  // namespace ccpi
