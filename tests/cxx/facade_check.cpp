// C++ caller of the drop-in facade (include/ccpi_hip.hpp): a small cone-beam
// FP, BP, CGLS and Tikhonov-regularised CGLS through the reference-shaped
// classes.  Prints one line "facade ok <fp_sum> <bp_sum> <norm0> <tik_norm0>"
// or exits non-zero.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ccpi_hip.hpp"

// HIP runtime calls of this check; a failure exits with status 3
#define CK(call)                                                        \
  do {                                                                  \
    if ((call) != hipSuccess) {                                         \
      std::printf("facade hip error at line %d\n", __LINE__);           \
      std::exit(3);                                                     \
    }                                                                   \
  } while (0)

int main() {
  const int n = 10, na = 12, nh = 16, nv = 14;
  ccpi::hip::cone_beam dev;
  std::vector<double> phi(na), h(nh), v(nv);
  for (int a = 0; a < na; a++) phi[a] = 2.0 * M_PI * a / na;
  for (int i = 0; i < nh; i++) h[i] = (i - (nh - 1) / 2.0) * 1.5;
  for (int i = 0; i < nv; i++) v[i] = (i - (nv - 1) / 2.0) * 1.5;
  dev.set_phi(phi);
  dev.set_h_pixels(h);
  dev.set_v_pixels(v);
  dev.set_source(-30.0, 0.0, 0.0);
  dev.set_detector(30.0);
  const double origin[3] = {-5, -5, -5}, width[3] = {1, 1, 1};
  const size_t nvox = (size_t)n * n * n, npix = (size_t)na * nh * nv;
  std::vector<float> x(nvox, 1.0f);
  float *dx = nullptr, *dy = nullptr;
  if (hipMalloc(&dx, nvox * 4) != hipSuccess || hipMalloc(&dy, npix * 4) != hipSuccess) return 2;
  CK(hipMemcpy(dx, x.data(), nvox * 4, hipMemcpyHostToDevice));
  CK(hipMemset(dy, 0, npix * 4));
  try {
    dev.forward_project(dy, dx, origin, width, n, n, n);
    std::vector<float> y(npix);
    CK(hipMemcpy(y.data(), dy, npix * 4, hipMemcpyDeviceToHost));
    double fps = 0;
    for (float q : y) fps += q;
    CK(hipMemset(dx, 0, nvox * 4));
    dev.backward_project(dy, dx, origin, width, n, n, n);
    CK(hipMemcpy(x.data(), dx, nvox * 4, hipMemcpyDeviceToHost));
    double bps = 0;
    for (float q : x) bps += q;
    CK(hipMemset(dx, 0, nvox * 4));
    ccpi::hip::cgls_3d alg(3);
    alg.reconstruct(dev, dy, dx, origin, width, n, n, n);
    CK(hipMemcpy(dy, y.data(), npix * 4, hipMemcpyHostToDevice));
    CK(hipMemset(dx, 0, nvox * 4));
    ccpi::hip::cgls_tikhonov tik(3, 0.2);
    tik.reconstruct(dev, dy, dx, origin, width, n, n, n);
    std::printf("facade ok %.9g %.9g %.9g %.9g\n", fps, bps, alg.convergence_data()[0],
                tik.convergence_data()[0]);
  } catch (const ccpi::hip::error &e) {
    std::printf("facade error %d %s\n", e.status, e.what());
    return 1;
  }
  CK(hipFree(dx));
  CK(hipFree(dy));
  return 0;
}
