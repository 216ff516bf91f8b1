// C++ caller of the drop-in facade (include/ccpi_hip.hpp): a small cone-beam
// FP, BP, CGLS and Tikhonov-regularised CGLS through the reference-shaped
// classes.  Writes the whole result arrays as raw float32 / float64 files
// into the directory given as argv[1] (fp.f32, bp.f32, cgls_x.f32,
// cgls_norms.f64, tik_x.f32, tik_norms.f64) and prints "facade ok", or
// exits non-zero.
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

static void dump(const char *dir, const char *name, const void *p, size_t bytes) {
  char path[4096];
  std::snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = std::fopen(path, "wb");
  if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
    std::printf("facade cannot write %s\n", path);
    std::exit(4);
  }
  std::fclose(f);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::printf("usage: facade_check <output dir>\n");
    return 2;
  }
  const char *dir = argv[1];
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
  // deterministic non-uniform volume: x[q] = (q * 37 mod 101) / 100
  std::vector<float> x(nvox);
  for (size_t q = 0; q < nvox; q++) x[q] = (float)((q * 37) % 101) / 100.0f;
  float *dx = nullptr, *dy = nullptr;
  if (hipMalloc(&dx, nvox * 4) != hipSuccess || hipMalloc(&dy, npix * 4) != hipSuccess) return 2;
  CK(hipMemcpy(dx, x.data(), nvox * 4, hipMemcpyHostToDevice));
  CK(hipMemset(dy, 0, npix * 4));
  try {
    dev.forward_project(dy, dx, origin, width, n, n, n);
    std::vector<float> y(npix);
    CK(hipMemcpy(y.data(), dy, npix * 4, hipMemcpyDeviceToHost));
    dump(dir, "fp.f32", y.data(), npix * 4);
    CK(hipMemset(dx, 0, nvox * 4));
    dev.backward_project(dy, dx, origin, width, n, n, n);
    CK(hipMemcpy(x.data(), dx, nvox * 4, hipMemcpyDeviceToHost));
    dump(dir, "bp.f32", x.data(), nvox * 4);
    CK(hipMemset(dx, 0, nvox * 4));
    ccpi::hip::cgls_3d alg(3);
    alg.reconstruct(dev, dy, dx, origin, width, n, n, n);
    CK(hipMemcpy(x.data(), dx, nvox * 4, hipMemcpyDeviceToHost));
    dump(dir, "cgls_x.f32", x.data(), nvox * 4);
    dump(dir, "cgls_norms.f64", alg.convergence_data().data(), 3 * sizeof(double));
    CK(hipMemcpy(dy, y.data(), npix * 4, hipMemcpyHostToDevice));
    CK(hipMemset(dx, 0, nvox * 4));
    ccpi::hip::cgls_tikhonov tik(3, 0.2);
    tik.reconstruct(dev, dy, dx, origin, width, n, n, n);
    CK(hipMemcpy(x.data(), dx, nvox * 4, hipMemcpyDeviceToHost));
    dump(dir, "tik_x.f32", x.data(), nvox * 4);
    dump(dir, "tik_norms.f64", tik.convergence_data().data(), 3 * sizeof(double));
    std::printf("facade ok\n");
  } catch (const ccpi::hip::error &e) {
    std::printf("facade error %d %s\n", e.status, e.what());
    return 1;
  }
  CK(hipFree(dx));
  CK(hipFree(dy));
  return 0;
}
