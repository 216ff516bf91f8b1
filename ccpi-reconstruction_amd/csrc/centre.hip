// XTek centre-of-rotation search on the device.
//
// Nikon_XTek::find_centre (Core/src/Readers/xtek.cpp:671-758; T. Liu, Opt.
// Eng. 48 (2009) 046501): five refinement levels of 21 candidate centres;
// for each candidate, the mean over (angle, h) of the squared difference
// between a ray and its conjugate ray interpolated from the sinogram row.
// The reference spends na * nh interpolations per candidate on one host
// thread (seconds at 720 x 768); here they are one launch per level.
//
// Arithmetic follows the reference value for value:
//   * per-candidate tables (atan / tan of the candidate and pixel angles)
//     are formed on the host in double, as the reference;
//   * the interpolation (linear / bilinear in h and angle, angles wrapping
//     at 2 pi, xtek.cpp:566-669) runs in double and is stored in float, the
//     difference and its square in float;
//   * the sum over (angle, h) is the reference's sequential float sum: one
//     wavefront per candidate loads 64 terms at a time and adds them in
//     order through readlane, so M and the chosen centre are bit-identical.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "plan.hpp"
#include "solvers.hpp"

namespace ccpi {

namespace {

constexpr int kCorSamples = 21;
constexpr int kCorLevels = 5;
constexpr int kCorThreads = 256;
constexpr double kTwoPi = 2.0 * M_PI;

__device__ __forceinline__ float cor_lin(double x1, double x2, float f1, float f2, double x) {
  return (float)(((double)f1 * (x2 - x) + (double)f2 * (x - x1)) / (x2 - x1));
}

__device__ __forceinline__ float cor_bil(double x1, double x2, double y1, double y2, float f11,
                                         float f12, float f21, float f22, double x, double y) {
  return (float)(((double)f11 * (x2 - x) * (y2 - y) + (double)f21 * (x - x1) * (y2 - y) +
                  (double)f12 * (x2 - x) * (y - y1) + (double)f22 * (x - x1) * (y - y1)) /
                 ((x2 - x1) * (y2 - y1)));
}

// First index i in [0, n-1) with v[i+1] > x, else n-1: the reference's
// linear scan.  Binary search when v is non-decreasing (the predicate is
// then monotone), the scan itself otherwise.
__device__ __forceinline__ int cor_find(const double *v, int n, double x, bool sorted) {
  if (sorted) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (v[m + 1] > x)
        hi = m;
      else
        lo = m + 1;
    }
    return lo;
  }
  int i = 0;
  while (i < n - 1 && !(v[i + 1] > x)) i++;
  return i;
}

// interpolate2D (xtek.cpp:651-669) with angles_linear / angles_bilinear
// (xtek.cpp:582-649) on d [na][nh]
__device__ float cor_interp(const double *__restrict__ h, const double *__restrict__ ang,
                            const float *__restrict__ d, int nh, int na, double nh_, double na_,
                            bool ang_sorted) {
  const int p1 = cor_find(h, nh, nh_, true);
  auto D = [&](int a, int k) { return d[(size_t)a * nh + k]; };
  if (h[p1] == nh_) {
    if (na_ < ang[0]) return cor_lin(ang[na - 1] - kTwoPi, ang[0], D(na - 1, p1), D(0, p1), na_);
    const int a1 = cor_find(ang, na, na_, ang_sorted);
    if (ang[a1] == na_) return D(a1, p1);
    if (a1 == na - 1) return cor_lin(ang[a1], ang[0] + kTwoPi, D(a1, p1), D(0, p1), na_);
    return cor_lin(ang[a1], ang[a1 + 1], D(a1, p1), D(a1 + 1, p1), na_);
  }
  if (na_ < ang[0])
    return cor_bil(h[p1], h[p1 + 1], ang[na - 1] - kTwoPi, ang[0], D(na - 1, p1), D(0, p1),
                   D(na - 1, p1 + 1), D(0, p1 + 1), nh_, na_);
  const int a1 = cor_find(ang, na, na_, ang_sorted);
  if (ang[a1] == na_) return cor_lin(h[p1], h[p1 + 1], D(a1, p1), D(a1, p1 + 1), nh_);
  if (a1 == na - 1)
    return cor_bil(h[p1], h[p1 + 1], ang[a1], ang[0] + kTwoPi, D(a1, p1), D(0, p1),
                   D(a1, p1 + 1), D(0, p1 + 1), nh_, na_);
  return cor_bil(h[p1], h[p1 + 1], ang[a1], ang[a1 + 1], D(a1, p1), D(a1 + 1, p1),
                 D(a1, p1 + 1), D(a1 + 1, p1 + 1), nh_, na_);
}

// row v_slice of [na][nh][nv] -> d [na][nh]
__global__ void cor_slice_kernel(const float *__restrict__ pix, float *__restrict__ d, int nrays,
                                 int nv, int v) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrays) d[r] = pix[(size_t)r * nv + v];
}

// tt[j][a nh + k] = (d[a][k] - conjugate)^2 for candidate j, 0 where the
// conjugate ray falls off the detector (xtek.cpp:720-735)
__global__ __launch_bounds__(kCorThreads) void cor_terms_kernel(
    const float *__restrict__ d, const double *__restrict__ h, const double *__restrict__ ang,
    const double *__restrict__ beta, const double *__restrict__ s2, float *__restrict__ tt,
    int na, int nh, int ang_sorted) {
  const int j = blockIdx.y;
  const int r = blockIdx.x * kCorThreads + threadIdx.x;
  if (r >= na * nh) return;
  const int a = r / nh, k = r - a * nh;
  const double s = s2[j * nh + k];
  float out = 0.f;
  if (s >= h[0] && s <= h[nh - 1]) {
    double ab = ang[a] + beta[j * nh + k];
    while (ab < ang[0]) ab += kTwoPi;
    while (ab > ang[na - 1]) ab -= kTwoPi;
    const float p = cor_interp(h, ang, d, nh, na, s, ab, ang_sorted != 0);
    const float t = d[r] - p;
    out = t * t;
  }
  tt[(size_t)j * na * nh + r] = out;
}

// sums[j] = the float sum of tt[j][0 .. n) in index order: one wavefront
// per candidate; lane 0's order of additions is the reference's loop order
__global__ __launch_bounds__(64) void cor_sum_kernel(const float *__restrict__ tt, int n,
                                                     float *__restrict__ sums) {
  const int j = blockIdx.x;
  const float *t = tt + (size_t)j * n;
  float acc = 0.f;
  for (int base = 0; base < n; base += 64) {
    const int q = base + (int)threadIdx.x;
    const float v = q < n ? t[q] : 0.f;
    const int bits = __float_as_int(v);
#pragma unroll
    for (int i = 0; i < 64; i++) acc = acc + __int_as_float(__builtin_amdgcn_readlane(bits, i));
  }
  if (threadIdx.x == 0) sums[j] = acc;
}

}  // namespace

double xtek_find_centre(const Geometry &G, const float *pixels, int v_slice, hipStream_t st) {
  const int na = G.na, nh = G.nh, nv = G.nv;
  CCPI_REQUIRE(G.kind == CCPI_CONE, CCPI_E_ARG, "find_centre needs a cone-beam geometry");
  CCPI_REQUIRE(nh >= 2 && na >= 1, CCPI_E_ARG, "find_centre needs >= 2 h pixels");
  CCPI_REQUIRE(v_slice >= 0 && v_slice < nv, CCPI_E_ARG, "v_slice outside the detector");
  CCPI_REQUIRE((long long)na * nh * kCorSamples < (1LL << 31), CCPI_E_LIMIT,
               "find_centre: na * nh too large");
  const int n = na * nh;
  const double dist = G.det_x - G.source[0];
  bool ang_sorted = true;
  for (int a = 0; a + 1 < na; a++) ang_sorted = ang_sorted && !(G.angles[a + 1] < G.angles[a]);

  DeviceBuffer d, dh, dang, dbeta, ds2, tt, sums;
  d.alloc((size_t)n * sizeof(float));
  dh.upload(G.h_pix, st);
  dang.upload(G.angles, st);
  dbeta.alloc((size_t)kCorSamples * nh * sizeof(double));
  ds2.alloc((size_t)kCorSamples * nh * sizeof(double));
  tt.alloc((size_t)kCorSamples * n * sizeof(float));
  sums.alloc(kCorSamples * sizeof(float));
  hipLaunchKernelGGL(cor_slice_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pixels,
                     d.as<float>(), n, nv, v_slice);
  CCPI_HIP(hipGetLastError());

  std::vector<double> gi(nh), beta((size_t)kCorSamples * nh), s2((size_t)kCorSamples * nh);
  for (int k = 0; k < nh; k++) gi[k] = std::atan(G.h_pix[k] / dist);
  const double prec[kCorLevels] = {1, 0.1, 0.01, 0.001, 0.0001};
  double mid = 0.0;
  float hsum[kCorSamples];
  for (int i = 0; i < kCorLevels; i++) {
    const double scor = mid - (double)(kCorSamples / 2) * prec[i];
    int count[kCorSamples];
    for (int j = 0; j < kCorSamples; j++) {
      const double gc = std::atan((mid + (j - kCorSamples / 2) * prec[i]) / dist);
      int c = 0;
      for (int k = 0; k < nh; k++) {
        beta[(size_t)j * nh + k] = M_PI + 2.0 * (gi[k] - gc);
        const double s = dist * std::tan(2.0 * gc - gi[k]);
        s2[(size_t)j * nh + k] = s;
        if (s >= G.h_pix[0] && s <= G.h_pix[nh - 1]) c++;
      }
      count[j] = c * na;
    }
    CCPI_HIP(hipMemcpyAsync(dbeta.p, beta.data(), beta.size() * sizeof(double),
                            hipMemcpyHostToDevice, st));
    CCPI_HIP(hipMemcpyAsync(ds2.p, s2.data(), s2.size() * sizeof(double), hipMemcpyHostToDevice,
                            st));
    hipLaunchKernelGGL(cor_terms_kernel, dim3((n + kCorThreads - 1) / kCorThreads, kCorSamples),
                       dim3(kCorThreads), 0, st, d.as<float>(), dh.as<double>(),
                       dang.as<double>(), dbeta.as<double>(), ds2.as<double>(), tt.as<float>(),
                       na, nh, ang_sorted ? 1 : 0);
    hipLaunchKernelGGL(cor_sum_kernel, dim3(kCorSamples), dim3(64), 0, st, tt.as<float>(), n,
                       sums.as<float>());
    CCPI_HIP(hipGetLastError());
    CCPI_HIP(hipMemcpyAsync(hsum, sums.p, sizeof hsum, hipMemcpyDeviceToHost, st));
    CCPI_HIP(hipStreamSynchronize(st));
    // minimum of M (xtek.cpp:737-748); 0 / 0 = NaN never wins
    int ind = -1;
    double mmin = 1e20;
    for (int j = 0; j < kCorSamples; j++) {
      const float M = hsum[j] / (float)count[j];
      if (mmin > M) {
        ind = j;
        mmin = M;
      }
    }
    CCPI_REQUIRE(ind >= 0, CCPI_E_GEOMETRY, "Failure in XTek find centre");
    mid = scor + ind * prec[i];
  }
  return mid * G.source[0] / dist;
}

}  // namespace ccpi
