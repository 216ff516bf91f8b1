// CGLS / SIRT / MLEM iteration bodies on device-resident vectors.
//
// Bodies follow Core/src/Algorithms/cgls.cpp:26-157 (cgls_3d),
// sirt.cpp:15-80 and mlem.cpp:15-80 with the vector operations of
// Core/src/blas.hpp:11-466.  Differences by design:
//   * reductions (||Ad||^2, ||A^T r||^2) are deterministic two-pass
//     reductions: per-workgroup partials, then one fixed-order pass.  By
//     default partials and the CGLS scalars are FP64 (the steps whose
//     stability needs it); CCPI_REDUCE_FP32 keeps float partials and a float
//     alpha like the reference (cgls.cpp:118-119);
//   * alpha / beta never leave the device, so an iteration issues no host
//     synchronisation and is graph-capturable;
//   * element-wise steps are fused (b - r then * R; x + t * C then clamp;
//     ratio then clamp; x * N * t then clamp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "plan.hpp"
#include "solvers.hpp"

namespace ccpi {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

__global__ void fill_kernel(float *__restrict__ p, size_t n, float val) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    p[q] = val;
}

template <class Acc>
__device__ Acc block_sum(Acc v) {
  __shared__ Acc sh[kRedThreads / 64];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  Acc t = 0;
  if (threadIdx.x == 0)
    for (int q = 0; q < kRedThreads / 64; q++) t += sh[q];
  return t;
}

// Per-workgroup partial sums of x^2 (fixed grid: deterministic).
template <class Acc>
__global__ __launch_bounds__(kRedThreads) void sumsq_kernel(const float *__restrict__ x,
                                                            size_t n, Acc *__restrict__ part) {
  Acc s = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const Acc v = (Acc)x[q];
    s += v * v;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Folds n FP64 partials into kRedBlocks: workgroup b sums the contiguous
// range b of them (its threads strided over it, then block_sum), so the
// result is a fixed function of the partials.
__global__ __launch_bounds__(kRedThreads) void fold_partials_kernel(const double *__restrict__ part,
                                                                    size_t n,
                                                                    double *__restrict__ out) {
  const size_t chunk = (n + kRedBlocks - 1) / kRedBlocks;
  const size_t q0 = (size_t)blockIdx.x * chunk, q1 = q0 + chunk < n ? q0 + chunk : n;
  double s = 0.0;
  for (size_t q = q0 + threadIdx.x; q < q1; q += kRedThreads) s += part[q];
  s = block_sum(s);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

template <class Acc>
__device__ Acc sum_partials(const Acc *part, int n) {
  // n partials in one block: each thread a fixed strided set in order, then
  // a fixed-order pairwise tree
  __shared__ Acc sh[kRedThreads];
  Acc s = 0;
  for (int q = threadIdx.x; q < n; q += kRedThreads) s += part[q];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  return sh[0];
}

// gamma = ||d||^2 at start (cgls.cpp:48-49)
template <class Acc>
__global__ __launch_bounds__(kRedThreads) void cgls_init_kernel(const Acc *part, int n,
                                                                CglsScalars *sc) {
  const Acc s = sum_partials(part, n);
  if (threadIdx.x == 0) {
    sc->gamma = (double)s;
    sc->gamma_f = (float)s;
  }
}

// alpha = gamma / ||Ad||^2 (cgls.cpp:118-119)
template <class Acc>
__global__ __launch_bounds__(kRedThreads) void cgls_alpha_kernel(const Acc *part, int n,
                                                                 CglsScalars *sc, int fp32) {
  const Acc s = sum_partials(part, n);
  if (threadIdx.x == 0) {
    if (fp32) {
      const float a = sc->gamma_f / (float)s;
      sc->alpha_f = a;
    } else {
      // a zero direction (b in the null space of A^T) stops instead of
      // producing 0/0 as the float reference does
      sc->alpha_f = (double)s > 0.0 ? (float)(sc->gamma / (double)s) : 0.f;
    }
  }
}

// beta = ||s||^2 / gamma, gamma <- ||s||^2, record ||A^T r|| (cgls.cpp:153-156)
template <class Acc>
__global__ __launch_bounds__(kRedThreads) void cgls_beta_kernel(const Acc *part, int n,
                                                                CglsScalars *sc,
                                                                double *norms, int it,
                                                                int fp32) {
  const Acc s = sum_partials(part, n);
  if (threadIdx.x == 0) {
    if (fp32) {
      const float nn = (float)s;
      const double beta = (double)nn / (double)sc->gamma_f;
      sc->beta_f = (float)beta;
      sc->gamma_f = nn;
      norms[it] = (double)sqrtf(nn);
    } else {
      const double nn = (double)s;
      sc->beta_f = sc->gamma > 0.0 ? (float)(nn / sc->gamma) : 0.f;
      sc->gamma = nn;
      norms[it] = sqrt(nn);
    }
  }
}

// alpha = gamma / (||Ad||^2 + ||Ad_vox||^2) (cgls.cpp:374-377); in float
// mode the two float norms are added in float as the reference does
template <class Acc>
__global__ __launch_bounds__(kRedThreads) void cgls_alpha_reg_kernel(const Acc *part_pix,
                                                                     int n_pix,
                                                                     const Acc *part_vox,
                                                                     CglsScalars *sc,
                                                                     int fp32) {
  const Acc sp = sum_partials(part_pix, n_pix);
  __syncthreads();
  const Acc sv = sum_partials(part_vox, kRedBlocks);
  if (threadIdx.x == 0) {
    if (fp32) {
      const float den = (float)sp + (float)sv;
      sc->alpha_f = sc->gamma_f / den;
    } else {
      const double den = (double)sp + (double)sv;
      sc->alpha_f = den > 0.0 ? (float)(sc->gamma / den) : 0.f;
    }
  }
}

// y += a x, with per-workgroup partial sums of the new y^2 (sum_axpy then
// norm_voxels, cgls.cpp:373-376)
template <class Acc>
__global__ __launch_bounds__(kRedThreads) void axpy_sumsq_kernel(float *__restrict__ y,
                                                                 const float *__restrict__ x,
                                                                 size_t n, float a,
                                                                 Acc *__restrict__ part) {
  Acc s = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const float v = y[q] + a * x[q];
    y[q] = v;
    const Acc w = (Acc)v;
    s += w * w;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// y += a x (sum_axpy, blas.hpp:307-322)
__global__ void axpy_kernel(float *__restrict__ y, const float *__restrict__ x, size_t n,
                            float a) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    y[q] = y[q] + a * x[q];
}

// x += alpha d (voxels) and r -= alpha Ad (pixels) (cgls.cpp:120-121)
__global__ void cgls_update_xr_kernel(float *__restrict__ x, const float *__restrict__ d,
                                      size_t nvox, float *__restrict__ r,
                                      const float *__restrict__ Ad, size_t npix,
                                      const CglsScalars *__restrict__ sc) {
  const float a = sc->alpha_f;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t q = t0; q < nvox; q += stride) x[q] = x[q] + a * d[q];
  for (size_t q = t0; q < npix; q += stride) r[q] = r[q] + (-a) * Ad[q];
}

// The same update for the cone beam, fused with the BP's input: x += alpha
// d, r -= alpha Ad, and dpix = r * d_conv in the BP's padded row layout
// (the reference's scaled copy, c2D.cpp:1105-1114; cone_scale_kernel's
// float product), so the BP of the new residual needs no scaling pass.
// Grid-stride over the voxels, then over whole rays (a workgroup's threads
// over a ray's rows: no per-element division for the row and ray).
__global__ void cgls_update_xr_scale_kernel(float *__restrict__ x, const float *__restrict__ d,
                                            size_t nvox, float *__restrict__ r,
                                            const float *__restrict__ Ad, size_t nrays, int nh,
                                            int nv, int nvp, const float *__restrict__ dconv,
                                            float *__restrict__ dpix,
                                            const CglsScalars *__restrict__ sc) {
  const float a = sc->alpha_f;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t q = t0; q < nvox; q += stride) x[q] = x[q] + a * d[q];
  for (size_t ray = blockIdx.x; ray < nrays; ray += gridDim.x) {
    const float *__restrict__ dc = dconv + (ray % (size_t)nh) * (size_t)nv;
    float *__restrict__ rr = r + ray * (size_t)nv;
    const float *__restrict__ ar = Ad + ray * (size_t)nv;
    float *__restrict__ dp = dpix + ray * (size_t)nvp;
    for (int v = threadIdx.x; v < nvp; v += blockDim.x) {
      float out = 0.0f;
      if (v < nv) {
        const float rn = rr[v] + (-a) * ar[v];
        rr[v] = rn;
        out = rn * dc[v];
      }
      dp[v] = out;
    }
  }
}

// d = s + beta d (scal_xby, blas.hpp:362-376)
__global__ void cgls_update_d_kernel(float *__restrict__ d, const float *__restrict__ s,
                                     size_t n, const CglsScalars *__restrict__ sc) {
  const float b = sc->beta_f;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    d[q] = b * d[q] + s[q];
}

// x = 1/x where x > m, else m (invert_min_x, blas.hpp:100-116)
__global__ void invert_min_kernel(float *__restrict__ x, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    x[q] = x[q] > 0.f ? (float)(1.0 / (double)x[q]) : 0.f;
}

// x = 1/x where x != 0 (invert_x, blas.hpp:86-98)
__global__ void invert_nz_kernel(float *__restrict__ x, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    if (x[q] != 0.f) x[q] = (float)(1.0 / (double)x[q]);
}

// r = (b - r) * R (sirt.cpp:60-62)
__global__ void sirt_residual_kernel(float *__restrict__ r, const float *__restrict__ b,
                                     const float *__restrict__ R, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const float t = -1.0f * r[q] + b[q];
    r[q] = t * R[q];
  }
}

// x = max(x + t * C, 0) (sirt.cpp:70-74)
__global__ void sirt_update_kernel(float *__restrict__ x, const float *__restrict__ t,
                                   const float *__restrict__ C, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const float v = x[q] + 1.0f * (t[q] * C[q]);
    x[q] = v < 0.f ? 0.f : v;
  }
}

// ratio = max(b / r or 0, 0) (mlem.cpp:54-65)
__global__ void mlem_ratio_kernel(float *__restrict__ r, const float *__restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const float e = r[q];
    const float v = e != 0.f ? b[q] / e : 0.f;
    r[q] = v < 0.f ? 0.f : v;
  }
}

// x = max(x * (N * t), 0) (mlem.cpp:72-74)
__global__ void mlem_update_kernel(float *__restrict__ x, const float *__restrict__ N,
                                   const float *__restrict__ t, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const float v = x[q] * (N[q] * t[q]);
    x[q] = v < 0.f ? 0.f : v;
  }
}

// ---------------------------------------------------------------------------

static int grid_for(size_t n) {
  const size_t b = (n + 255) / 256;
  return (int)std::min<size_t>(std::max<size_t>(b, 1), 8192);
}

void dev_fill(float *p, size_t n, float v, hipStream_t st) {
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, n, v);
  CCPI_HIP(hipGetLastError());
}

Workspace::~Workspace() = default;

void Workspace::ensure(const Projector &P) {
  const size_t nvox = P.geometry().n_vox(), npix = P.geometry().n_pix();
  if (const size_t nd = P.bp_scratch_floats()) dscale.alloc(nd * 4);
  v1.alloc(nvox * 4);
  v2.alloc(nvox * 4);
  v3.alloc(nvox * 4);
  p1.alloc(npix * 4);
  p2.alloc(npix * 4);
  part.alloc(std::max<size_t>({(size_t)kRedBlocks, P.fp_norm_parts(), P.bp_norm_parts()}) *
             sizeof(double));
  fold.alloc(kRedBlocks * sizeof(double));
  scal.alloc(sizeof(CglsScalars));
}

// The projector epilogues leave one norm partial per workgroup (or column).
// The single-workgroup reductions sum up to kFoldAbove of them quickly;
// more are first folded into kRedBlocks (into `folded`).  Returns the
// partials to reduce and sets n to their count.
constexpr size_t kFoldAbove = 16 * kRedThreads;
static const double *fold_parts(const double *part, size_t &n, double *folded, hipStream_t st) {
  if (n <= kFoldAbove) return part;
  hipLaunchKernelGGL(fold_partials_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, st, part, n,
                     folded);
  CCPI_HIP(hipGetLastError());
  n = kRedBlocks;
  return folded;
}

template <class Acc>
static void sumsq(const float *x, size_t n, void *part, hipStream_t st) {
  hipLaunchKernelGGL(sumsq_kernel<Acc>, dim3(kRedBlocks), dim3(kRedThreads), 0, st, x, n,
                     (Acc *)part);
  CCPI_HIP(hipGetLastError());
}

void CglsRun::start(hipStream_t st) {
  const size_t nvox = P.geometry().n_vox();
  W.ensure(P);
  norms_dev.alloc((size_t)std::max(iters, 1) * sizeof(double));
  float *d = W.v1.as<float>();
  if (fp32) {
    // the reference's float partial sums: a separate pass over d
    dev_fill(d, nvox, 0.f, st);
    P.bp(b, d, st, W.bp_scratch());
    sumsq<float>(d, nvox, W.part.p, st);
    hipLaunchKernelGGL(cgls_init_kernel<float>, dim3(1), dim3(kRedThreads), 0, st,
                       W.part.as<float>(), kRedBlocks, W.scal.as<CglsScalars>());
  } else {
    // d = A^T b written over d, its FP64 norm partials from the BP's epilogue
    P.bp(b, d, st, W.bp_scratch(), false, false, W.part.as<double>());
    size_t n = P.bp_norm_parts();
    const double *pp = fold_parts(W.part.as<double>(), n, W.fold.as<double>(), st);
    hipLaunchKernelGGL(cgls_init_kernel<double>, dim3(1), dim3(kRedThreads), 0, st, pp, (int)n,
                       W.scal.as<CglsScalars>());
  }
  CCPI_HIP(hipGetLastError());
}

EventLog::~EventLog() {
  for (hipEvent_t e : fp) (void)hipEventDestroy(e);
  for (hipEvent_t e : bp) (void)hipEventDestroy(e);
}

void EventLog::mark(std::vector<hipEvent_t> &v, hipStream_t st) {
  hipEvent_t e;
  CCPI_HIP(hipEventCreate(&e));
  v.push_back(e);
  CCPI_HIP(hipEventRecord(e, st));
}

double EventLog::total_ms(const std::vector<hipEvent_t> &v) {
  double t = 0.0;
  for (size_t q = 0; q + 1 < v.size(); q += 2) {
    float ms = 0.f;
    CCPI_HIP(hipEventSynchronize(v[q + 1]));
    CCPI_HIP(hipEventElapsedTime(&ms, v[q], v[q + 1]));
    t += ms;
  }
  return t;
}

void CglsRun::iterate(int it, hipStream_t st) {
  const size_t nvox = P.geometry().n_vox(), npix = P.geometry().n_pix();
  float *d = W.v1.as<float>(), *s = W.v2.as<float>(), *Ad = W.p1.as<float>();
  CglsScalars *sc = W.scal.as<CglsScalars>();
  // FP64 mode: the projectors write their outputs over Ad and s and leave
  // the FP64 norm partials of ||Ad||^2 and ||s||^2 in W.part (epilogue
  // norms); FP32 mode keeps the reference's float partial sums, a pass of
  // their own
  double *norm = fp32 ? nullptr : W.part.as<double>();
  if (fp32) dev_fill(Ad, npix, 0.f, st);
  if (log) log->mark(log->fp, st);
  P.fp(d, Ad, st, false, fp32, norm);
  if (log) log->mark(log->fp, st);
  if (fp32) {
    sumsq<float>(Ad, npix, W.part.p, st);
    hipLaunchKernelGGL(cgls_alpha_kernel<float>, dim3(1), dim3(kRedThreads), 0, st,
                       W.part.as<float>(), kRedBlocks, sc, 1);
  } else {
    size_t n = P.fp_norm_parts();
    const double *pp = fold_parts(W.part.as<double>(), n, W.fold.as<double>(), st);
    hipLaunchKernelGGL(cgls_alpha_kernel<double>, dim3(1), dim3(kRedThreads), 0, st, pp, (int)n,
                       sc, 0);
  }
  if (P.cone) {
    // the residual update writes the cone BP's scaled input as well
    const ConeArgs &A = P.cone->args;
    const size_t nrays = (size_t)A.na * A.nh;
    float *dpix = W.bp_scratch();
    hipLaunchKernelGGL(cgls_update_xr_scale_kernel,
                       dim3(grid_for(std::max(nvox, nrays * (size_t)A.nvp))), dim3(256), 0, st,
                       x, d, nvox, b, Ad, nrays, A.nh, A.nv, A.nvp, A.dconv, dpix, sc);
    if (fp32) dev_fill(s, nvox, 0.f, st);
    if (log) log->mark(log->bp, st);
    P.bp_scaled(dpix, s, st, fp32, norm);
    if (log) log->mark(log->bp, st);
  } else {
    hipLaunchKernelGGL(cgls_update_xr_kernel, dim3(grid_for(std::max(nvox, npix))), dim3(256),
                       0, st, x, d, nvox, b, Ad, npix, sc);
    if (fp32) dev_fill(s, nvox, 0.f, st);
    if (log) log->mark(log->bp, st);
    P.bp(b, s, st, W.bp_scratch(), false, fp32, norm);
    if (log) log->mark(log->bp, st);
  }
  if (fp32) {
    sumsq<float>(s, nvox, W.part.p, st);
    hipLaunchKernelGGL(cgls_beta_kernel<float>, dim3(1), dim3(kRedThreads), 0, st,
                       W.part.as<float>(), kRedBlocks, sc, norms_dev.as<double>(), it, 1);
  } else {
    size_t n = P.bp_norm_parts();
    const double *pp = fold_parts(W.part.as<double>(), n, W.fold.as<double>(), st);
    hipLaunchKernelGGL(cgls_beta_kernel<double>, dim3(1), dim3(kRedThreads), 0, st, pp, (int)n,
                       sc, norms_dev.as<double>(), it, 0);
  }
  hipLaunchKernelGGL(cgls_update_d_kernel, dim3(grid_for(nvox)), dim3(256), 0, st, d, s, nvox,
                     sc);
  CCPI_HIP(hipGetLastError());
}

void CglsRun::finish(double *norms, hipStream_t st) {
  if (norms && iters > 0) {
    CCPI_HIP(hipMemcpyAsync(norms, norms_dev.p, (size_t)iters * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    CCPI_HIP(hipStreamSynchronize(st));
  }
}

void run_cgls(const Projector &P, float *b, float *x, int iters, double *norms, bool fp32,
              Workspace &W, hipStream_t st) {
  CglsRun run{P, W, b, x, iters, fp32};
  run.start(st);
  for (int it = 0; it < iters; it++) run.iterate(it, st);
  run.finish(norms, st);
}

// cgls_regularize::reconstruct (cgls.cpp:343-423).  Per iteration:
// Ad = A d; Ad_vox += lambda L(d) (never reset); alpha = gamma / (||Ad||^2 +
// ||Ad_vox||^2); x += alpha d; r -= alpha Ad; x += lambda L(x); s = A^T r;
// beta, d = s + beta d.  The reference's update of an unused voxel residual
// (cgls.cpp:390, indexed with the sinogram's extents) has no effect on x or
// the norms and is not performed.
void run_cgls_reg(const Projector &P, float *b, float *x, int iters, int kind, double lambda,
                  double *norms, bool fp32, Workspace &W, hipStream_t st) {
  const Geometry &G = P.geometry();
  check_reg_dims(G.nx, G.ny, G.nz);
  const size_t nvox = G.n_vox(), npix = G.n_pix();
  W.ensure(P);
  W.v4.alloc(nvox * 4);
  W.part2.alloc(kRedBlocks * sizeof(double));
  const size_t nscr = reg_scratch_floats(kind, nvox);
  if (nscr) W.tvd.alloc(nscr * 4);
  DeviceBuffer norms_dev;
  norms_dev.alloc((size_t)std::max(iters, 1) * sizeof(double));
  float *d = W.v1.as<float>(), *s = W.v2.as<float>(), *L = W.v3.as<float>();
  float *adv = W.v4.as<float>(), *Ad = W.p1.as<float>();
  float *scr = nscr ? W.tvd.as<float>() : nullptr;
  CglsScalars *sc = W.scal.as<CglsScalars>();
  const float lam = (float)lambda;
  const int gv = grid_for(nvox);
  // FP64 mode: projector outputs written over d, Ad, s with their norm
  // partials from the projectors' epilogues, as in run_cgls (so lambda = 0
  // gives run_cgls bit for bit); FP32 mode: float partial sums in passes of
  // their own
  double *norm = fp32 ? nullptr : W.part.as<double>();
  dev_fill(adv, nvox, 0.f, st);
  if (fp32) dev_fill(d, nvox, 0.f, st);
  P.bp(b, d, st, W.bp_scratch(), false, fp32, norm);
  if (fp32) {
    sumsq<float>(d, nvox, W.part.p, st);
    hipLaunchKernelGGL(cgls_init_kernel<float>, dim3(1), dim3(kRedThreads), 0, st,
                       W.part.as<float>(), kRedBlocks, sc);
  } else {
    size_t n = P.bp_norm_parts();
    const double *pp = fold_parts(W.part.as<double>(), n, W.fold.as<double>(), st);
    hipLaunchKernelGGL(cgls_init_kernel<double>, dim3(1), dim3(kRedThreads), 0, st, pp, (int)n,
                       sc);
  }
  for (int it = 0; it < iters; it++) {
    if (fp32) dev_fill(Ad, npix, 0.f, st);
    P.fp(d, Ad, st, false, fp32, norm);
    reg_term(kind, d, L, scr, G.nx, G.ny, G.nz, st);
    if (fp32) {
      sumsq<float>(Ad, npix, W.part.p, st);
      hipLaunchKernelGGL(axpy_sumsq_kernel<float>, dim3(kRedBlocks), dim3(kRedThreads), 0, st,
                         adv, L, nvox, lam, W.part2.as<float>());
      hipLaunchKernelGGL(cgls_alpha_reg_kernel<float>, dim3(1), dim3(kRedThreads), 0, st,
                         W.part.as<float>(), kRedBlocks, W.part2.as<float>(), sc, 1);
    } else {
      hipLaunchKernelGGL(axpy_sumsq_kernel<double>, dim3(kRedBlocks), dim3(kRedThreads), 0,
                         st, adv, L, nvox, lam, W.part2.as<double>());
      size_t n = P.fp_norm_parts();
      const double *pp = fold_parts(W.part.as<double>(), n, W.fold.as<double>(), st);
      hipLaunchKernelGGL(cgls_alpha_reg_kernel<double>, dim3(1), dim3(kRedThreads), 0, st, pp,
                         (int)n, W.part2.as<double>(), sc, 0);
    }
    hipLaunchKernelGGL(cgls_update_xr_kernel, dim3(grid_for(std::max(nvox, npix))), dim3(256),
                       0, st, x, d, nvox, b, Ad, npix, sc);
    reg_term(kind, x, L, scr, G.nx, G.ny, G.nz, st);
    hipLaunchKernelGGL(axpy_kernel, dim3(gv), dim3(256), 0, st, x, L, nvox, lam);
    if (fp32) dev_fill(s, nvox, 0.f, st);
    P.bp(b, s, st, W.bp_scratch(), false, fp32, norm);
    if (fp32) {
      sumsq<float>(s, nvox, W.part.p, st);
      hipLaunchKernelGGL(cgls_beta_kernel<float>, dim3(1), dim3(kRedThreads), 0, st,
                         W.part.as<float>(), kRedBlocks, sc, norms_dev.as<double>(), it, 1);
    } else {
      size_t n = P.bp_norm_parts();
      const double *pp = fold_parts(W.part.as<double>(), n, W.fold.as<double>(), st);
      hipLaunchKernelGGL(cgls_beta_kernel<double>, dim3(1), dim3(kRedThreads), 0, st, pp,
                         (int)n, sc, norms_dev.as<double>(), it, 0);
    }
    hipLaunchKernelGGL(cgls_update_d_kernel, dim3(gv), dim3(256), 0, st, d, s, nvox, sc);
    CCPI_HIP(hipGetLastError());
  }
  if (norms && iters > 0) {
    CCPI_HIP(hipMemcpyAsync(norms, norms_dev.p, (size_t)iters * sizeof(double),
                            hipMemcpyDeviceToHost, st));
  }
  CCPI_HIP(hipStreamSynchronize(st));
}

void run_sirt(const Projector &P, const float *b, float *x, int iters, Workspace &W,
              hipStream_t st) {
  const size_t nvox = P.geometry().n_vox(), npix = P.geometry().n_pix();
  W.ensure(P);
  float *C = W.v1.as<float>(), *t = W.v2.as<float>(), *ones = W.v3.as<float>();
  float *R = W.p1.as<float>(), *r = W.p2.as<float>();
  // R = 1 / (A 1), C = 1 / (A^T 1) (sirt.cpp:31-48)
  // the projections write over their outputs (the reference adds to
  // zeroed arrays: the same sums)
  dev_fill(ones, nvox, 1.f, st);
  P.fp(ones, R, st, false, false);
  hipLaunchKernelGGL(invert_min_kernel, dim3(grid_for(npix)), dim3(256), 0, st, R, npix);
  dev_fill(r, npix, 1.f, st);
  P.bp(r, C, st, W.bp_scratch(), false, false);
  hipLaunchKernelGGL(invert_min_kernel, dim3(grid_for(nvox)), dim3(256), 0, st, C, nvox);
  for (int it = 0; it < iters; it++) {
    P.fp(x, r, st, false, false);
    hipLaunchKernelGGL(sirt_residual_kernel, dim3(grid_for(npix)), dim3(256), 0, st, r, b, R,
                       npix);
    P.bp(r, t, st, W.bp_scratch(), false, false);
    hipLaunchKernelGGL(sirt_update_kernel, dim3(grid_for(nvox)), dim3(256), 0, st, x, t, C,
                       nvox);
  }
  CCPI_HIP(hipGetLastError());
}

// f64: the projections' sums (A x and the ratio backprojection A^T (b / A x),
// and the sensitivity A^T 1) accumulated in double, each rounded to float
// once; the element-wise steps are unchanged.
void run_mlem(const Projector &P, const float *b, float *x, int iters, Workspace &W,
              hipStream_t st, bool f64) {
  const size_t nvox = P.geometry().n_vox(), npix = P.geometry().n_pix();
  W.ensure(P);
  float *N = W.v1.as<float>(), *t = W.v2.as<float>();
  float *r = W.p1.as<float>();
  // x = 1; N = 1 / (A^T 1) keeping zeros (mlem.cpp:30-41)
  // the projections write over their outputs (the reference adds to
  // zeroed arrays: the same sums)
  dev_fill(x, nvox, 1.f, st);
  dev_fill(r, npix, 1.f, st);
  P.bp(r, N, st, W.bp_scratch(), f64, false);
  hipLaunchKernelGGL(invert_nz_kernel, dim3(grid_for(nvox)), dim3(256), 0, st, N, nvox);
  for (int it = 0; it < iters; it++) {
    P.fp(x, r, st, f64, false);
    hipLaunchKernelGGL(mlem_ratio_kernel, dim3(grid_for(npix)), dim3(256), 0, st, r, b, npix);
    P.bp(r, t, st, W.bp_scratch(), f64, false);
    hipLaunchKernelGGL(mlem_update_kernel, dim3(grid_for(nvox)), dim3(256), 0, st, x, N, t,
                       nvox);
  }
  CCPI_HIP(hipGetLastError());
}

}  // namespace ccpi

namespace ccpi {

unsigned long long projector_nnz(const Projector &P, hipStream_t st) {
  DeviceBuffer cnt;
  cnt.alloc(sizeof(unsigned long long));
  CCPI_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), st));
  if (P.cone)
    cone_nnz(*P.cone, cnt.as<unsigned long long>(), st);
  else
    par_nnz(*P.par, cnt.as<unsigned long long>(), st);
  unsigned long long h = 0;
  CCPI_HIP(hipMemcpyAsync(&h, cnt.p, sizeof h, hipMemcpyDeviceToHost, st));
  CCPI_HIP(hipStreamSynchronize(st));
  return h;
}

}  // namespace ccpi
