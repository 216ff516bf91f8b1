// Parallel-beam forward projector and matched backprojector for gfx950.
//
// Discrete model: CCPi::parallel_beam::f2D / b2D (Core/src/p2D.cpp:22-1087).
//   * FP: 2-D Siddon walk of the detector line through the xy grid, the
//     xy length of each crossed cell weighting voxel (i, j, map(v)) for
//     every detector row v (p2D.cpp:63-466);
//   * BP: per voxel column and angle, the exact chord of each detector line
//     through the column square - a trapezoid in h (p2D.cpp:775-1049).
// Geometry arithmetic is the reference's (double, rounded to float weights,
// no contraction).  The FP walks the whole grid in one pass instead of the
// reference's 32 x 32 tiles, so only the float summation order differs;
// the BP visits angles and columns in the reference's order.
//
// Mapping to CDNA4.  FP: `lr` lanes share one ray (lr = nv / 4 rounded up
// to a power of two, at most 16), each lane walks the ray and keeps 4
// consecutive detector rows in registers, so the lanes of a ray read one
// contiguous run of each crossed column (16 bytes per lane) and a wave
// covers 64 / lr adjacent rays of one angle.  BP: one voxel column per
// lane, kRows slices in registers, pixel rows read once per (angle, h)
// hit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "plan.hpp"
#include "par_body.hpp"

namespace ccpi {

// slices per BP pass: 8 keeps the kernel at 91 VGPRs (5 waves per SIMD);
// 16 needs 121 (4 waves) and was 5 % slower on the 64-problem workload
constexpr int kRows = 8;

// Acc: float sums (the reference's) or double (FP64-accumulation option).
// acc_in: y += A x (the reference's f2D) or y = A x; norm: null, or one
// partial of ||y||^2 per workgroup (epilogue norms, common.hpp).
template <class Acc>
__global__ __launch_bounds__(256) void par_fp_kernel(ParArgs g, int lr,
                                                     const float *__restrict__ vol,
                                                     float *__restrict__ sino, int acc_in,
                                                     double *__restrict__ norm) {
  const double ss =
      par_fp_lanes<Acc>(g, blockIdx.x, blockIdx.y, threadIdx.x, lr, vol, sino, acc_in != 0);
  if (norm) block_part256(ss, norm, (size_t)blockIdx.y * gridDim.x + blockIdx.x);
}

template <class Acc>
__global__ __launch_bounds__(256) void par_bp_kernel(ParArgs g,
                                                     const float *__restrict__ sino,
                                                     float *__restrict__ vol, int acc_in,
                                                     double *__restrict__ norm) {
  const double ss = par_bp_thread<kRows, Acc>(g, blockIdx.x, blockIdx.y, threadIdx.x, sino,
                                              vol, acc_in != 0);
  if (norm) block_part256(ss, norm, (size_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// The tiled BP as one launch: blockIdx.x = tile, y = slice chunk, z = angle
// group; one group writes vol (adding to it when acc_in), several write
// their partial sums to part[group] for par_bp_sum_kernel.  norm (one
// group only): one partial of ||vol||^2 per workgroup.  The body waits on
// its staged rows and footprints, not on issue: 5 waves per SIMD (94
// VGPRs, a few spills in the set-up) instead of the 4 that 103 VGPRs allow
// took the 64-problem mix's BP from 21.9 to 18.8 ms per round (6 waves:
// 19.2), the 256 x 256 x 16 BP unchanged.  The FP64-accumulation instances
// keep their registers.
template <class Acc>
constexpr int par_bp_lds_min_waves() { return std::is_same<Acc, float>::value ? 5 : 1; }

template <int R, class Acc>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(par_bp_lds_min_waves<Acc>(), 8))) void par_bp_lds_kernel(
    ParArgs g, int hmax, int ab, int groups, const float *__restrict__ sino,
    float *__restrict__ vol, float *__restrict__ part, int acc_in, double *__restrict__ norm) {
  extern __shared__ float smem[];
  const int grp = blockIdx.z;
  const int a_begin = (int)((long long)g.na * grp / groups);
  const int a_end = (int)((long long)g.na * (grp + 1) / groups);
  float *dst = groups == 1 ? vol : part + (size_t)grp * g.nx * g.ny * g.nz;
  const double ss = par_bp_tile<R, Acc>(g, blockIdx.x, blockIdx.y, a_begin, a_end, hmax, ab,
                                        groups == 1 && acc_in != 0, sino, dst, smem);
  if (norm) block_part256(ss, norm, (size_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// Workgroups of par_bp_sum_kernel (a fixed grid: its norm partials are a
// fixed set).
constexpr int kParSumBlocks = 2048;

// vol (+)= sum of the angle groups' partials, in group order (in Acc);
// norm: one partial of ||vol||^2 per workgroup.
template <class Acc>
__global__ __launch_bounds__(256) void par_bp_sum_kernel(const float *__restrict__ part,
                                                         float *__restrict__ vol, size_t n,
                                                         int groups, int acc_in,
                                                         double *__restrict__ norm) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  double ss = 0.0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    Acc s = (Acc)part[q];
    for (int p = 1; p < groups; p++) s = s + (Acc)part[(size_t)p * n + q];
    const float o = acc_in ? (float)((Acc)vol[q] + s) : (float)s;
    vol[q] = o;
    ss += (double)o * (double)o;
  }
  if (norm) block_part256(ss, norm, blockIdx.x);
}

// Walk table of the plan, pass 1: the number of cells of every ray (one
// lane per ray, the FP's own walk).
__global__ __launch_bounds__(256) void par_walk_count_kernel(ParArgs g,
                                                             unsigned long long *__restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.na * g.nh) return;
  const int a = r / g.nh, h = r - a * g.nh;
  unsigned long long n = 0;
  par_walk_groups<kParFpGroup>(g, a, h, [&](const int *, const float *, int m) { n += m; });
  cnt[r] = n;
}

// Pass 2: the cells themselves, at the ray's offset.
__global__ __launch_bounds__(256) void par_walk_fill_kernel(ParArgs g,
                                                            const unsigned long long *__restrict__ off,
                                                            uint2 *__restrict__ ent) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.na * g.nh) return;
  const int a = r / g.nh, h = r - a * g.nh;
  unsigned long long e = off[r];
  par_walk_groups<kParFpGroup>(g, a, h, [&](const int *cid, const float *l, int m) {
    for (int q = 0; q < m; q++)
      ent[e + q] = make_uint2((unsigned)cid[q], __builtin_bit_cast(unsigned, l[q]));
    e += m;
  });
}

// Nonzero system-matrix entries: per ray, crossed cells with nonzero xy
// length times the detector rows that map into the volume.
__global__ __launch_bounds__(256) void par_nnz_kernel(ParArgs g,
                                                      unsigned long long *__restrict__ count) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n = 0;
  if (r < g.na * g.nh) {
    const int a = r / g.nh, h = r - a * g.nh;
    int rows = 0;
    for (int v = 0; v < g.nv; v++) rows += (g.map[v] >= 0 && g.map[v] < g.nz) ? 1 : 0;
    const double c = g.atab[5 * a], s = g.atab[5 * a + 1];
    const double p2x = c * g.D - s * g.h_pix[h];
    const double p2y = s * g.D + c * g.h_pix[h];
    int cells = 0;
    if (fabs(c) < kEps) {
      const int i = (int)floor((p2x - g.ox) / g.dx);
      if (fabs(s) >= kEps && i >= 0 && i < g.nx) cells = g.ny;
    } else if (fabs(s) < kEps) {
      const int j = (int)floor((p2y - g.oy) / g.dy);
      if (j >= 0 && j < g.ny) cells = g.nx;
    } else {
      const double ix = 1.0 / c, iy = 1.0 / s;
      const double ax0 = (g.ox - p2x) * ix, ay0 = (g.oy - p2y) * iy;
      const double axn = (g.ox + (double)g.nx * g.dx - p2x) * ix;
      const double ayn = (g.oy + (double)g.ny * g.dy - p2y) * iy;
      const double amin = fmax(fmax(fmin(ax0, axn), fmin(ay0, ayn)), -g.dconv);
      const double amax = fmin(fmin(fmax(ax0, axn), fmax(ay0, ayn)), 0.0);
      if (amin < amax - kEps) {
        // |cells crossed| = steps until the walk leaves the grid
        const int sx = c > 0.0 ? 1 : -1, sy = s > 0.0 ? 1 : -1;
        int x = (int)floor((p2x + amin * c - g.ox) / g.dx);
        int y = (int)floor((p2y + amin * s - g.oy) / g.dy);
        x = min(max(x, 0), g.nx - 1);
        y = min(max(y, 0), g.ny - 1);
        const int bxs = sx > 0, bys = sy > 0;
        double ap = amin;
        while (x >= 0 && x < g.nx && y >= 0 && y < g.ny) {
          const double nxa = (g.ox + (double)(x + bxs) * g.dx - p2x) * ix;
          const double nya = (g.oy + (double)(y + bys) * g.dy - p2y) * iy;
          double nap;
          if (nxa < nya - kEps) { nap = nxa; x += sx; }
          else if (nya < nxa - kEps) { nap = nya; y += sy; }
          else { nap = fmax(nxa, nya); x += sx; y += sy; }
          cells += ((float)(nap - ap) != 0.f) ? 1 : 0;
          ap = nap;
        }
      }
    }
    n = (unsigned long long)cells * (unsigned long long)rows;
  }
  __shared__ unsigned long long s_tot;
  if (threadIdx.x == 0) s_tot = 0;
  __syncthreads();
  atomicAdd(&s_tot, n);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(count, s_tot);
}

void par_nnz(const ParPlan &P, unsigned long long *count, hipStream_t st) {
  const ParArgs &A = P.args;
  const int rays = A.na * A.nh;
  hipLaunchKernelGGL(par_nnz_kernel, dim3((rays + 255) / 256), dim3(256), 0, st, A, count);
  CCPI_HIP(hipGetLastError());
}

// Host tables of one parallel geometry and the scalars of its argument
// block (table pointers left for the caller to set).
struct ParTables {
  std::vector<double> atab;
  std::vector<int> map, rlo, rhi;
};

static ParTables par_tables(const Geometry &G, ParArgs &A) {
  A = ParArgs{};
  A.na = G.na; A.nh = G.nh; A.nv = G.nv; A.nx = G.nx; A.ny = G.ny; A.nz = G.nz;
  A.ox = G.origin[0]; A.oy = G.origin[1];
  A.dx = G.vsize[0]; A.dy = G.vsize[1];
  A.D = 2.0 * std::max(std::fabs(G.origin[0]),
                       std::max(std::fabs(G.origin[1]), std::fabs(G.origin[2])));
  A.dconv = (double)(float)(3.0 * A.D);
  A.ihp = 1.0 / (G.h_pix[1] - G.h_pix[0]);
  A.hp0 = G.h_pix[0] / (G.h_pix[1] - G.h_pix[0]);
  ParTables T;
  // row -> slice map and its pattern (p2D.cpp:22-61)
  std::vector<int> &map = T.map, &rlo = T.rlo, &rhi = T.rhi;
  map.resize(G.nv);
  rlo.assign(G.nz, 0);
  rhi.assign(G.nz, 0);
  for (int v = 0; v < G.nv; v++)
    map[v] = (int)std::floor((G.v_pix[v] - G.origin[2]) / G.vsize[2]);
  A.map_type = 0;
  for (int ppv = 1; ppv <= 4; ppv *= 2) {
    bool ok = true;
    for (int v = 0; v < G.nv && ok; v++) ok = map[v] == v / ppv;
    if (ok) { A.map_type = ppv; break; }
  }
  for (int k = 0; k < G.nz; k++) {
    int lo = G.nv, hi = G.nv;
    for (int v = 0; v < G.nv; v++)
      if (map[v] >= k) { lo = v; break; }
    for (int v = lo; v < G.nv; v++)
      if (map[v] > k) { hi = v; break; }
    rlo[k] = lo;
    rhi[k] = std::max(lo, hi);
  }
  // per-angle trapezoid parameters (p2D.cpp:996-1049)
  std::vector<double> &atab = T.atab;
  atab.resize((size_t)G.na * 5);
  for (int a = 0; a < G.na; a++) {
    const double c = std::cos(G.angles[a]), s = std::sin(G.angles[a]);
    const double *o = G.origin, *d = G.vsize;
    const double y00 = o[1] * c - o[0] * s;
    const double y01 = (o[1] + d[1]) * c - o[0] * s;
    const double y10 = o[1] * c - (o[0] + d[0]) * s;
    const double y11 = (o[1] + d[1]) * c - (o[0] + d[0]) * s;
    double ymin, ybot;
    if (c > 0.0) {
      if (s > 0.0) { ymin = y10; ybot = y11 < y00 ? y11 : y00; }
      else         { ymin = y00; ybot = y01 < y10 ? y01 : y10; }
    } else {
      if (s > 0.0) { ymin = y11; ybot = y01 < y10 ? y01 : y10; }
      else         { ymin = y01; ybot = y11 < y00 ? y11 : y00; }
    }
    double *t = &atab[(size_t)a * 5];
    t[0] = c;
    t[1] = s;
    t[2] = ybot - ymin;
    t[3] = 1.0 / (ybot - ymin);
    t[4] = std::fabs(c) > std::fabs(s) ? d[0] / std::fabs(c) : d[1] / std::fabs(s);
  }
  return T;
}

// LDS of one tiled-BP workgroup: the staged rows and the per-angle spans.
static size_t par_bp_lds_bytes(const ParBpShape &S) {
  return sizeof(float) * (size_t)S.ab * S.hmax * S.slices + 2 * sizeof(int) * (size_t)S.ab;
}

// Shape of the tiled BP for a geometry, or tiled = 0.  The staged span of a
// 16 x 16 tile is its corners' detector span plus 2 columns each side; in
// column-index units (the first pixel step, as the footprints use) it is at
// most the tile's diagonal times 1 / step, so hmax bounds it with a margin.
// Angle groups are added while the tiles alone give fewer than 4096 waves
// (four per SIMD on 256 CUs).
static ParBpShape par_bp_shape(const Geometry &G, const ParArgs &A) {
  ParBpShape S;
  if (A.map_type != 1) return S;
  S.slices = G.nz <= 8 ? 8 : 16;
  const double span = std::hypot(16.0 * G.vsize[0], 16.0 * G.vsize[1]) * std::fabs(A.ihp);
  if (!(span < 4096.0)) return S;
  S.hmax = (int)std::ceil(span) + 6;
  const size_t per_angle = sizeof(float) * (size_t)S.hmax * S.slices + 2 * sizeof(int);
  if (per_angle > 32 * 1024) return S;
  S.ab = (int)std::max<size_t>(1, std::min<size_t>(8, (32 * 1024) / per_angle));
  const long long tiles = (long long)((G.nx + 15) / 16) * ((G.ny + 15) / 16);
  const long long waves = tiles * ((G.nz + S.slices - 1) / S.slices) * 4;
  S.groups = (int)std::min<long long>(std::min<long long>(8, G.na),
                                      std::max<long long>(1, (4096 + waves - 1) / waves));
  S.tiled = 1;
  return S;
}

void par_build_plan(ParPlan &P, hipStream_t st) {
  const Geometry &G = P.g;
  ParArgs A;
  const ParTables T = par_tables(G, A);
  P.atab.upload(T.atab, st);
  P.hpix.upload(G.h_pix, st);
  P.map.upload(T.map, st);
  P.rlo.upload(T.rlo, st);
  P.rhi.upload(T.rhi, st);
  A.atab = P.atab.as<double>();
  A.h_pix = P.hpix.as<double>();
  A.map = P.map.as<int>();
  A.rlo = P.rlo.as<int>();
  A.rhi = P.rhi.as<int>();
  // walk table: every ray's cells, so the FP reads them instead of walking
  // the grid in double each time.  Only for FPs of fewer than
  // kParWalkTableWaves wavefronts: those leave SIMDs idle and are bound by
  // the walk's FP64 work (256 x 256 x 16, 180 angles: 0.182 -> 0.122 ms);
  // larger ones are bound by their gathers' latency, which the table's
  // dependent load lengthens (the 64-problem mix: 25.5 -> 27.0 ms).
  const size_t rays = (size_t)G.na * G.nh;
  if (par_walk_table_wanted(G)) {
    P.woff.alloc((rays + 1) * sizeof(unsigned long long));
    unsigned long long *off = P.woff.as<unsigned long long>();
    const int blocks = (int)((rays + 255) / 256);
    hipLaunchKernelGGL(par_walk_count_kernel, dim3(blocks), dim3(256), 0, st, A, off);
    CCPI_HIP(hipGetLastError());
    std::vector<unsigned long long> h(rays + 1);
    CCPI_HIP(hipMemcpyAsync(h.data(), off, rays * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, st));
    CCPI_HIP(hipStreamSynchronize(st));
    unsigned long long total = 0;
    for (size_t q = 0; q < rays; q++) {
      const unsigned long long n = h[q];
      h[q] = total;
      total += n;
    }
    h[rays] = total;
    P.woff.upload(h, st);
    P.went.alloc(std::max<unsigned long long>(total, 1) * sizeof(uint2));
    hipLaunchKernelGGL(par_walk_fill_kernel, dim3(blocks), dim3(256), 0, st, A,
                       P.woff.as<unsigned long long>(), P.went.as<uint2>());
    CCPI_HIP(hipGetLastError());
    A.woff = P.woff.as<unsigned long long>();
    A.went = P.went.as<uint2>();
  }
  CCPI_HIP(hipStreamSynchronize(st));  // h is the upload's source
  P.args = A;
  P.bp = par_bp_shape(G, A);
}

static dim3 par_fp_grid(const ParArgs &A, int &lr) {
  const int rays = A.na * A.nh;
  lr = par_fp_lanes_per_ray(A.nv);
  const int rpb = 256 / lr;
  return dim3((unsigned)((rays + rpb - 1) / rpb), (unsigned)((A.nv + 4 * lr - 1) / (4 * lr)));
}

size_t par_fp_norm_parts(const ParPlan &P) {
  int lr;
  const dim3 g = par_fp_grid(P.args, lr);
  return (size_t)g.x * g.y;
}

void par_fp(const ParPlan &P, const float *vol, float *sino, hipStream_t st, bool f64,
            bool acc_in, double *norm) {
  const ParArgs &A = P.args;
  int lr;
  const dim3 grid = par_fp_grid(A, lr);
  const int ai = acc_in ? 1 : 0;
  if (f64)
    hipLaunchKernelGGL(par_fp_kernel<double>, grid, dim3(256), 0, st, A, lr, vol, sino, ai,
                       norm);
  else
    hipLaunchKernelGGL(par_fp_kernel<float>, grid, dim3(256), 0, st, A, lr, vol, sino, ai,
                       norm);
  CCPI_HIP(hipGetLastError());
}

size_t par_bp_scratch_floats(const ParPlan &P) {
  return P.bp.tiled && P.bp.groups > 1 ? (size_t)P.bp.groups * P.g.n_vox() : 0;
}

// Grids of the BP's launches: the column-per-lane kernel, or the tiled
// kernel and (with angle groups) the group sum, which writes the volume.
static dim3 par_bp_grid(const ParPlan &P) {
  const ParArgs &A = P.args;
  const ParBpShape &S = P.bp;
  if (!S.tiled)
    return dim3((unsigned)((A.nx * A.ny + 255) / 256), (unsigned)((A.nz + kRows - 1) / kRows));
  const unsigned tiles = (unsigned)(((A.nx + 15) / 16) * ((A.ny + 15) / 16));
  return dim3(tiles, (unsigned)((A.nz + S.slices - 1) / S.slices), (unsigned)S.groups);
}

static int par_bp_sum_blocks(const ParPlan &P) {
  return (int)std::min<size_t>((P.g.n_vox() + 255) / 256, kParSumBlocks);
}

size_t par_bp_norm_parts(const ParPlan &P) {
  if (P.bp.tiled && P.bp.groups > 1) return (size_t)par_bp_sum_blocks(P);
  const dim3 g = par_bp_grid(P);
  return (size_t)g.x * g.y;
}

template <class Acc>
static void par_bp_launch(const ParPlan &P, const float *sino, float *vol, float *scratch,
                          hipStream_t st, bool acc_in, double *norm) {
  const ParArgs &A = P.args;
  const ParBpShape &S = P.bp;
  const dim3 grid = par_bp_grid(P);
  const int ai = acc_in ? 1 : 0;
  if (!S.tiled) {
    hipLaunchKernelGGL(par_bp_kernel<Acc>, grid, dim3(256), 0, st, A, sino, vol, ai, norm);
    CCPI_HIP(hipGetLastError());
    return;
  }
  CCPI_REQUIRE(S.groups == 1 || scratch, CCPI_E_ARG, "parallel BP: angle groups need scratch");
  const size_t lds = par_bp_lds_bytes(S);
  // with angle groups the sum kernel writes the volume (and its norm)
  double *tile_norm = S.groups == 1 ? norm : nullptr;
  if (S.slices == 8)
    hipLaunchKernelGGL((par_bp_lds_kernel<8, Acc>), grid, dim3(256), lds, st, A, S.hmax, S.ab,
                       S.groups, sino, vol, scratch, ai, tile_norm);
  else
    hipLaunchKernelGGL((par_bp_lds_kernel<16, Acc>), grid, dim3(256), lds, st, A, S.hmax,
                       S.ab, S.groups, sino, vol, scratch, ai, tile_norm);
  CCPI_HIP(hipGetLastError());
  if (S.groups > 1) {
    hipLaunchKernelGGL(par_bp_sum_kernel<Acc>, dim3(par_bp_sum_blocks(P)), dim3(256), 0, st,
                       scratch, vol, P.g.n_vox(), S.groups, ai, norm);
    CCPI_HIP(hipGetLastError());
  }
}

void par_bp(const ParPlan &P, const float *sino, float *vol, float *scratch, hipStream_t st,
            bool f64, bool acc_in, double *norm) {
  if (f64)
    par_bp_launch<double>(P, sino, vol, scratch, st, acc_in, norm);
  else
    par_bp_launch<float>(P, sino, vol, scratch, st, acc_in, norm);
}

// ---------------------------------------------------------------------------
// allocation-free forward projection
// ---------------------------------------------------------------------------

// The per-call tables of par_fp_noalloc live in this file-scope device
// array instead of a cached plan: [atab | h_pix | map | rlo | rhi], each
// part 256-byte aligned.  Every device the library runs on holds one copy
// until the process ends; calls on one device share it, so calls that
// overlap on one device (different streams, host threads) race.  In host
// code `scratch` names a host-side stand-in: kernels get its device address
// from hipGetSymbolAddress (scratch_address below), never the name.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch_ptr;  // device -> address of `scratch`

// Device address of the current device's `scratch`, looked up once per
// device (with that device current).
unsigned char *scratch_address() {
  int dev = 0;
  CCPI_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void *p = nullptr;
  CCPI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  CCPI_REQUIRE(p != nullptr, CCPI_E_HIP, "no device address for the projector scratch");
  g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  return static_cast<unsigned char *>(p);
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
}  // namespace

size_t par_noalloc_bytes(const Geometry &G) {
  return align256((size_t)G.na * 5 * sizeof(double)) + align256((size_t)G.nh * sizeof(double)) +
         align256((size_t)G.nv * sizeof(int)) + 2 * align256((size_t)G.nz * sizeof(int));
}

void par_fp_noalloc(const Geometry &G, const float *vol, float *sino, hipStream_t st) {
  const size_t need = par_noalloc_bytes(G);
  CCPI_REQUIRE(need <= kNoAllocScratchBytes, CCPI_E_LIMIT,
               "allocation-free parallel FP: the geometry's tables need " +
                   std::to_string(need) + " bytes, more than its " +
                   std::to_string(kNoAllocScratchBytes) + "-byte scratch");
  ParArgs A;
  const ParTables T = par_tables(G, A);
  // one host image of the tables, copied into the scratch in one transfer
  std::vector<unsigned char> img(need, 0);
  size_t off = 0;
  auto place = [&](const void *src, size_t bytes) {
    const size_t at = off;
    if (bytes) std::memcpy(img.data() + at, src, bytes);
    off += align256(bytes);
    return at;
  };
  const size_t o_atab = place(T.atab.data(), T.atab.size() * sizeof(double));
  const size_t o_hpix = place(G.h_pix.data(), G.h_pix.size() * sizeof(double));
  const size_t o_map = place(T.map.data(), T.map.size() * sizeof(int));
  const size_t o_rlo = place(T.rlo.data(), T.rlo.size() * sizeof(int));
  const size_t o_rhi = place(T.rhi.data(), T.rhi.size() * sizeof(int));
  unsigned char *base = scratch_address();
  // stream-ordered: an earlier call's kernel on this stream has read the
  // previous tables before these land
  CCPI_HIP(hipMemcpyAsync(base, img.data(), need, hipMemcpyHostToDevice, st));
  A.atab = reinterpret_cast<const double *>(base + o_atab);
  A.h_pix = reinterpret_cast<const double *>(base + o_hpix);
  A.map = reinterpret_cast<const int *>(base + o_map);
  A.rlo = reinterpret_cast<const int *>(base + o_rlo);
  A.rhi = reinterpret_cast<const int *>(base + o_rhi);
  ParPlan P;
  P.args = A;
  par_fp(P, vol, sino, st);
}

}  // namespace ccpi
