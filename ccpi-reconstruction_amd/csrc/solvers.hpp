// Device-resident iteration bodies (see solvers.hip).
#pragma once

#include <vector>

#include "plan.hpp"

namespace ccpi {

// Scalars of one CGLS run, kept in device memory.
struct CglsScalars {
  double gamma;      // ||A^T r||^2 (FP64 mode)
  float gamma_f;     // same, float mode
  float alpha_f;     // step applied to x and r
  float beta_f;      // direction update factor
  float pad;
};

// Work vectors reused across calls on one geometry.
struct Workspace {
  DeviceBuffer v1, v2, v3;   // volume-sized
  DeviceBuffer p1, p2;       // sinogram-sized
  DeviceBuffer part;         // reduction partials (and projector norm partials)
  DeviceBuffer fold;         // norm partials folded to kRedBlocks
  DeviceBuffer scal;         // CglsScalars
  DeviceBuffer v4, part2;    // regularised CGLS: Ad_vox, its partials
  DeviceBuffer tvd;          // TV term: D1, D2, D3
  DeviceBuffer dscale;       // cone BP input scaled by d_conv (padded rows)
  ~Workspace();
  void ensure(const Projector &P);
  // the workspace's BP scratch (nullptr when the projector needs none)
  float *bp_scratch() const { return dscale.p ? dscale.as<float>() : nullptr; }
};

// regularize.hip: Tikhonov / TV terms (tikhonov.cpp, total_v.cpp)
constexpr int kRegTikhonov = 0;
constexpr int kRegTV = 1;
void check_reg_dims(int nx, int ny, int nz);
size_t reg_scratch_floats(int kind, size_t nvox);
// out = L(a); scratch holds reg_scratch_floats(kind, nvox) floats
void reg_term(int kind, const float *a, float *out, float *scratch, int nx, int ny, int nz,
              hipStream_t st);

// HIP events bracketing projector launches (resolved on demand).
struct EventLog {
  std::vector<hipEvent_t> fp, bp;   // pairs: start, stop
  ~EventLog();
  void mark(std::vector<hipEvent_t> &v, hipStream_t st);
  static double total_ms(const std::vector<hipEvent_t> &v);
};

// One CGLS run split in launch-only steps (no host sync inside
// start / iterate), so a caller may capture iterate() in a graph.
struct CglsRun {
  const Projector &P;
  Workspace &W;
  float *b;      // sinogram in, residual out
  float *x;      // volume in/out
  int iters;
  bool fp32;
  DeviceBuffer norms_dev;
  EventLog *log = nullptr;
  void start(hipStream_t st);
  void iterate(int it, hipStream_t st);
  void finish(double *norms, hipStream_t st);
};

void dev_fill(float *p, size_t n, float v, hipStream_t st);
void run_cgls(const Projector &P, float *b, float *x, int iters, double *norms,
              bool fp32, Workspace &W, hipStream_t st);
// centre.hip: XTek centre-of-rotation search (xtek.cpp:671-758) on row
// v_slice of the device sinogram; returns the y shift of source and pixels
double xtek_find_centre(const Geometry &G, const float *pixels, int v_slice, hipStream_t st);
// regularised CGLS (cgls.cpp:343-435), kind kRegTikhonov or kRegTV
void run_cgls_reg(const Projector &P, float *b, float *x, int iters, int kind, double lambda,
                  double *norms, bool fp32, Workspace &W, hipStream_t st);
void run_sirt(const Projector &P, const float *b, float *x, int iters,
              Workspace &W, hipStream_t st);
void run_mlem(const Projector &P, const float *b, float *x, int iters,
              Workspace &W, hipStream_t st, bool f64 = false);
unsigned long long projector_nnz(const Projector &P, hipStream_t st);
// batched.hip: parallel-beam problems of uneven sizes, one persistent launch
// per step, iteration replayed from a hipGraph (FP64 reductions)
void run_pb_cgls_batched(const std::vector<Geometry> &geoms, float *const *b,
                         float *const *x, int iters, double *norms_out, hipStream_t st);

}  // namespace ccpi
