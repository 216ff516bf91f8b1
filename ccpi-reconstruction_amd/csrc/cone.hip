// Cone-beam forward projector and matched backprojector for gfx950.
//
// Discrete model: CCPi::cone_beam::f2D / b2D (Core/src/c2D.cpp:17-1119).
//   * xy: Siddon walk of the source->pixel line in cell units, shared by
//     every detector row of one (angle, h) column (c2D.cpp:193-467);
//   * z: per row and xy segment, the entry slice k and at most one slice
//     crossing split the segment between k and k-/+1 (c2D.cpp:17-191);
//   * output scaled by the source->pixel distance d_conv (c2D.cpp:616-623).
// Weights are computed with the reference's arithmetic (double geometry,
// float alpha, no contraction: the library is built -ffp-contract=off), so
// weights agree bit for bit with the model and only summation order
// differs.
//
// Mapping to CDNA4:
//   plan     one lane per (angle, h) ray walks the xy grid in double and
//            writes the ray's (cell, alpha) list to HBM once per geometry.
//   FP       one workgroup per ray column and 256-row block; lanes are
//            detector rows, so the xy walk is uniform (scalar loads of the
//            entry list) and each segment's voxel column reads are
//            near-contiguous in z across the wave.
//   BP       one wavefront per voxel column: the plan lists the column's
//            (angle, h) rays (double footprints); the list is swept once
//            per block of 64*S detector rows, S consecutive rows per lane,
//            and each row's two slice parts go into an LDS copy of the
//            column by plain read-add-writes of distinct cells (one LDS
//            array per row slot, or, past 4 rows per slice, one array
//            written slot by slot in lane phases).  One wave owns a column,
//            so the BP needs no global atomics and is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "plan.hpp"

namespace ccpi {

// ---------------------------------------------------------------------------
// plan: xy walk of every ray
// ---------------------------------------------------------------------------

// Walk state for one ray; FILL writes entries and tracks ncom.
template <bool FILL>
struct WalkSink {
  ConeEnt *out;
  int n = 0;
  // ncom bookkeeping (c2D.cpp:49-78)
  float prev = 0.f;
  int min_xy = -1, max_xy = -1;
  float pzbz, inv_dz, dz0, dzn;
  int nzm1;

  __device__ void emit(uint32_t col, float a) {
    if (FILL) {
      out[n] = ConeEnt{col, a};
      if (n >= 1) {
        const float ai = prev * inv_dz;
        if (min_xy < 0) {
          const int k = (int)floorf(pzbz + ai * dz0);
          if (k <= 0) min_xy = (k == 0) ? n : 0;
        }
        if (max_xy < 0) {
          const int k = (int)floorf(pzbz + ai * dzn);
          if (k >= nzm1) max_xy = (k == nzm1) ? n : 0;
        }
      }
      prev = a;
    }
    ++n;
  }
};

// Mirrors the cone xy walk (c2D.cpp:227-457) with one unified step rule
// for the four direction cases.
template <class Sink>
__device__ void cone_walk(double q1x, double q1y, double q2x, double q2y,
                          int nx, int ny, Sink &sk) {
  const double ddx = q2x - q1x, ddy = q2y - q1y;
  const double ix = 1.0 / ddx, iy = 1.0 / ddy;
  if (fabs(ddx) < kEps) {
    if (fabs(ddy) >= kEps) {
      const int i = (int)floor(q1x);
      if (i >= 0 && i < nx) {
        if (ddy < 0.0) {
          sk.emit((uint32_t)(i * ny + ny - 1), (float)(((double)ny - q1y) * iy));
          for (int j = ny - 1; j >= 0; j--)
            sk.emit((uint32_t)(i * ny + j), (float)(((double)j - q1y) * iy));
        } else {
          sk.emit((uint32_t)(i * ny), (float)((0.0 - q1y) * iy));
          for (int j = 0; j < ny; j++)
            sk.emit((uint32_t)(i * ny + j), (float)(((double)(j + 1) - q1y) * iy));
        }
      }
    }
    return;
  }
  if (fabs(ddy) < kEps) {
    const int j = (int)floor(q1y);
    if (j >= 0 && j < ny) {
      if (ddx < 0.0) {
        sk.emit((uint32_t)((nx - 1) * ny + j), (float)(((double)nx - q1x) * ix));
        for (int i = nx - 1; i >= 0; i--)
          sk.emit((uint32_t)(i * ny + j), (float)(((double)i - q1x) * ix));
      } else {
        sk.emit((uint32_t)j, (float)((0.0 - q1x) * ix));
        for (int i = 0; i < nx; i++)
          sk.emit((uint32_t)(i * ny + j), (float)(((double)(i + 1) - q1x) * ix));
      }
    }
    return;
  }
  const double ax0 = (0.0 - q1x) * ix, ay0 = (0.0 - q1y) * iy;
  const double axn = ((double)nx - q1x) * ix, ayn = ((double)ny - q1y) * iy;
  const double amin = fmax(fmax(fmin(ax0, axn), fmin(ay0, ayn)), 0.0);
  const double amax = fmin(fmin(fmax(ax0, axn), fmax(ay0, ayn)), 1.0);
  if (!(amin < amax - kEps)) return;
  const int sx = ddx > 0.0 ? 1 : -1, sy = ddy > 0.0 ? 1 : -1;
  int x, y;
  {
    const double xface = sx > 0 ? ax0 : axn, yface = sy > 0 ? ay0 : ayn;
    const int xin = sx > 0 ? 0 : nx - 1, yin = sy > 0 ? 0 : ny - 1;
    bool on_x = amin == xface, on_y = amin == yface;
    if (sx > 0 && sy < 0) {
      if (on_y) on_x = false;
    } else if (!(sx < 0 && sy < 0)) {
      if (on_x) on_y = false;
    }
    if (on_x) {
      x = xin;
      y = on_y ? yin : (int)floor(q1y + amin * ddy);
    } else if (on_y) {
      y = yin;
      x = (int)floor(q1x + amin * ddx);
    } else {
      x = 0;
      y = 0;
    }
    x = min(max(x, 0), nx - 1);
    y = min(max(y, 0), ny - 1);
  }
  sk.emit((uint32_t)(x * ny + y), (float)amin);
  const int bxs = sx > 0, bys = sy > 0;
  while (x >= 0 && x < nx && y >= 0 && y < ny) {
    const double nxa = ((double)(x + bxs) - q1x) * ix;
    const double nya = ((double)(y + bys) - q1y) * iy;
    const uint32_t col = (uint32_t)(x * ny + y);
    if (nxa < nya - kEps) {
      sk.emit(col, (float)nxa);
      x += sx;
    } else if (nya < nxa - kEps) {
      sk.emit(col, (float)nya);
      y += sy;
    } else {
      sk.emit(col, (float)fmax(nxa, nya));
      x += sx;
      y += sy;
    }
  }
}

template <bool FILL>
__global__ __launch_bounds__(256) void cone_walk_kernel(
    ConeArgs g, const uint64_t *__restrict__ offs, uint32_t *__restrict__ counts,
    ConeRay *__restrict__ rays, ConeEnt *__restrict__ ents) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.na * g.nh) return;
  const int a = r / g.nh, h = r - a * g.nh;
  const int hlo = g.hrange[2 * a], hhi = g.hrange[2 * a + 1];
  WalkSink<FILL> sk;
  sk.out = FILL ? ents + offs[r] : nullptr;
  sk.pzbz = g.pzbz;
  sk.inv_dz = g.inv_dz;
  sk.dz0 = g.dz[0];
  sk.dzn = g.dz[g.nv - 1];
  sk.nzm1 = g.nz - 1;
  if (h >= hlo && h <= hhi) {
    const double *t = g.atab + 8 * a;
    const double c = t[0], s = t[1], p1x = t[2], p1y = t[3];
    const double p2x = c * g.det_x - s * g.h_pix[h];
    const double p2y = s * g.det_x + c * g.h_pix[h];
    cone_walk((p1x - g.vx) / g.dx, (p1y - g.vy) / g.dy, (p2x - g.vx) / g.dx,
              (p2y - g.vy) / g.dy, g.nx, g.ny, sk);
  }
  if (!FILL) {
    counts[r] = (uint32_t)sk.n;
    return;
  }
  ConeRay R;
  R.off = offs[r];
  R.n = sk.n;
  R.ncom = 1;
  R.min_v = g.mid;
  R.max_v = g.mid - 1;
  R.pad[0] = R.pad[1] = 0;
  if (sk.n >= 2) {
    const int n = sk.n;
    int mx = sk.min_xy < 0 ? n : sk.min_xy;
    int Mx = sk.max_xy < 0 ? n : sk.max_xy;
    R.ncom = max(min(mx, Mx), 1);
    // rows inside through the start of the last segment (c2D.cpp:80-95);
    // the slice index is non-decreasing in v, so bisect.
    const float ai = ents[R.off + n - 2].alpha * g.inv_dz;
    int lo = 0, hi = g.mid;  // first v < mid with floor(...) > 0
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if ((int)floorf(g.pzbz + ai * g.dz[m]) > 0) hi = m; else lo = m + 1;
    }
    R.min_v = lo;
    lo = g.mid;
    hi = g.nv;  // first v >= mid with floor(...) >= nz - 1
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if ((int)floorf(g.pzbz + ai * g.dz[m]) >= g.nz - 1) hi = m; else lo = m + 1;
    }
    R.max_v = lo - 1;
  }
  rays[r] = R;
}

// ---------------------------------------------------------------------------
// forward projection
// ---------------------------------------------------------------------------

constexpr int kFpRows = 256;

// Workgroups go to the 8 XCDs round-robin (workgroup id mod 8), each XCD
// with its own L2.  xcd_block maps the hardware id to a logical id so that
// every XCD runs one contiguous range of logical ids: neighbouring work
// (adjacent voxel columns) then shares an L2.
constexpr int kXcds = 8;
__device__ inline int xcd_block(int bid, int n) {
  const int xcd = bid % kXcds, local = bid / kXcds;
  const int per = n / kXcds, rem = n % kXcds;
  return xcd * per + min(xcd, rem) + local;
}
// dword 3 of a raw (untyped, stride 0) buffer resource on gfx9 CDNA
constexpr int kRsrcWord3 = 0x00020000;

// 32-bit raw buffer load reinterpreted as float (the builtin returns the
// bits as an unsigned int); an offset outside the resource reads 0.
__device__ inline float buf_load_f32(__amdgpu_buffer_rsrc_t r, int byte_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}

// v_med3_i32: x clamped to [lo, hi] in one instruction (lo <= hi).
__device__ inline int med3_i32(int x, int lo, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
  return r;
}

// v_med3_i32 with lo = 0 (inline constant) and a uniform hi: no copy of hi
// into a VGPR
__device__ inline int med3_0s(int x, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(x), "s"(hi));
  return r;
}

// v_med3_i32 with lo = -2 (inline constant) and a uniform hi
__device__ inline int med3_m2s(int x, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, -2, %2" : "=v"(r) : "v"(x), "s"(hi));
  return r;
}

// LDS byte address of a shared-memory pointer, and back
typedef __attribute__((address_space(3))) float lds_f32;
__device__ inline uint32_t lds_addr(const float *p) {
  return (uint32_t)(uintptr_t)(const lds_f32 *)p;
}
__device__ inline lds_f32 *lds_ptr(uint32_t a) { return (lds_f32 *)(uintptr_t)a; }
// (k << 2) + base in one v_lshl_add_u32
__device__ inline uint32_t k4_add(int k, uint32_t base) {
  uint32_t r;
  asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(r) : "v"(k), "v"(base));
  return r;
}

// min(a, b) for a uniform a as one v_min_f32 (the select form needs a
// copy of a into a VGPR); operands are finite alphas, so this is the
// reference's `a < b ? a : b`.
__device__ inline float min_sv(float a, float b) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "s"(a), "v"(b));
  return r;
}

// Per-lane constants of one detector row in the FP segment loop.
struct FpRow {
  float dzv, izv;      // row's z run per unit alpha and its inverse
  float kb;            // slice base: pzbz, or pzbz + 1 for a floor-form row
  int ks;              // 1 for a floor-form row (c2D.cpp:153-167), else 0
  int klo, khi;        // clamp of k for the boundary plane (shifted above)
  const float *zb;     // s_vz, or s_vz + 1 above the source
  int dir4;            // byte step to the second slice: -4 below, +4 above
};

// One xy segment [a0, a1] of column `col`, ai = a0 / voxel z size: the
// row's two slices (c2D.cpp:104-143, 153-187) weighted by their alpha
// lengths.  Slices outside [0, nz) read 0 through the buffer range check.
template <bool EDGE>
__device__ inline void fp_seg(const FpRow &R, float a0, float a1, float ai,
                              __amdgpu_buffer_rsrc_t col, float &c0, float &c1,
                              float &w0, float &w1) {
  int k = (int)(R.kb + ai * R.dzv);
  if (EDGE) k -= R.ks;
  const float az = R.zb[med3_i32(k, R.klo, R.khi)] * R.izv;
  const float mz = min_sv(a1, az);
  w0 = mz - a0;
  w1 = a1 - mz;
  const int off = k * 4;
  c0 = buf_load_f32(col, off);
  c1 = buf_load_f32(col, off + R.dir4);
}

__device__ inline __amdgpu_buffer_rsrc_t column_rsrc(const float *vol, uint32_t col, int nz) {
  const uint32_t c = __builtin_amdgcn_readfirstlane(col);
  return __builtin_amdgcn_make_buffer_rsrc((void *)(vol + (size_t)c * nz), (short)0, nz * 4,
                                           kRsrcWord3);
}

constexpr int kFpGroup = 8;

// Segments [mb, me) of one ray for the wave's rows.  Groups of kFpGroup
// segments issue all their gathers before the in-order accumulation (the
// sum order is the reference's).  The wave stops once the segment start
// passes every row's exit parameter (amax_bits: a non-negative float's
// bits, so the integer compare is the float compare).
// Acc is the row accumulator: float (the reference's sum, c2D.cpp:104-143)
// or double (the FP64-accumulation option: each segment's float
// contribution is added in double).
template <bool EDGE, class Acc>
__device__ inline Acc fp_run(const FpRow &R, const ConeEnt *__restrict__ e,
                             const float *__restrict__ ai, const float *vol, int nz,
                             int mb, int me, int amax_bits, Acc acc) {
  int m = mb;
  for (; m + kFpGroup <= me; m += kFpGroup) {
    if (__float_as_int(e[m - 1].alpha) >= amax_bits) return acc;
    float c0[kFpGroup], c1[kFpGroup], w0[kFpGroup], w1[kFpGroup];
#pragma unroll
    for (int j = 0; j < kFpGroup; ++j) {
      const ConeEnt en = e[m + j];
      fp_seg<EDGE>(R, e[m + j - 1].alpha, en.alpha, ai[m + j], column_rsrc(vol, en.col, nz),
                   c0[j], c1[j], w0[j], w1[j]);
    }
#pragma unroll
    for (int j = 0; j < kFpGroup; ++j) acc = acc + (Acc)(c0[j] * w0[j] + c1[j] * w1[j]);
  }
  for (; m < me; ++m) {
    if (__float_as_int(e[m - 1].alpha) >= amax_bits) return acc;
    const ConeEnt en = e[m];
    float c0, c1, w0, w1;
    fp_seg<EDGE>(R, e[m - 1].alpha, en.alpha, ai[m], column_rsrc(vol, en.col, nz), c0, c1,
                 w0, w1);
    acc = acc + (Acc)(c0 * w0 + c1 * w1);
  }
  return acc;
}

// One workgroup per ray and 256 detector rows, one row per lane.  Acc as
// fp_run; with double the result is rounded once, after the d_conv scaling.
template <class Acc>
__global__ __launch_bounds__(kFpRows) void cone_fp_kernel(
    ConeArgs g, const ConeRay *__restrict__ rays, const ConeEnt *__restrict__ ents,
    const float *__restrict__ eai, const float *__restrict__ vol, float *__restrict__ sino,
    int acc_in, double *__restrict__ norm) {
  extern __shared__ float s_vz[];
  for (int q = threadIdx.x; q <= g.nz; q += kFpRows) s_vz[q] = g.vz[q];
  __syncthreads();
  // (ray, row block); grouping rays per XCD measured slower here than the
  // hardware's round-robin spread
  const int r = blockIdx.x;
  const int v0 = blockIdx.y * kFpRows + threadIdx.x;
  // lanes past the last row mirror row nv - 1 and store nothing, so the
  // wave-wide reduction below sees every lane
  const bool live = v0 < g.nv;
  const int v = live ? v0 : g.nv - 1;
  const int h = r % g.nh;
  const ConeRay R = rays[r];
  float *out = sino + (size_t)r * g.nv + v;
  Acc acc = acc_in ? (Acc)*out : (Acc)0;
  // A row whose ray has left the volume's z range at the start of a
  // segment adds nothing from there on (c2D.cpp:153-187 break): the wave
  // stops at the largest such walk parameter of its rows.
  float amax = g.alim[v];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) amax = fmaxf(amax, __shfl_xor(amax, d, 64));
  const int amax_bits = __builtin_amdgcn_readfirstlane(__float_as_int(amax));
  if (R.n > 1) {
    const int nz = g.nz;
    const bool lower = v < g.mid;
    FpRow row;
    row.dzv = g.dz[v];
    row.izv = g.iz[v];
    row.klo = lower ? 0 : -1;
    row.khi = lower ? nz : nz - 1;
    row.zb = s_vz + (lower ? 0 : 1);
    row.dir4 = lower ? -4 : 4;
    row.kb = g.pzbz;
    row.ks = 0;
    const ConeEnt *__restrict__ e = ents + R.off;
    const float *__restrict__ ai = eai + R.off;
    // segments before ncom: every row uses the truncating form
    const int ncom = min(R.ncom, R.n);
    acc = fp_run<false, Acc>(row, e, ai, vol, nz, 1, ncom, amax_bits, acc);
    // from ncom on, rows below min_v use the floor form (c2D.cpp:153-167)
    const bool edge = v < R.min_v;
    row.kb = edge ? g.pzbz1 : g.pzbz;
    row.ks = edge ? 1 : 0;
    acc = fp_run<true, Acc>(row, e, ai, vol, nz, ncom, R.n, amax_bits, acc);
  }
  const float o = (float)(acc * (Acc)g.dconv[(size_t)h * g.nv + v]);
  if (live) *out = o;
  if (norm) block_part256(live ? (double)o * (double)o : 0.0, norm,
                          (size_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// eai[q] = ents[q - 1].alpha / voxel z size: the slice-position factor of
// the segment ending at entry q (the same float product the FP would form).
__global__ void cone_ai_kernel(const ConeEnt *__restrict__ ents, uint64_t n, float inv_dz,
                               float *__restrict__ eai) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    eai[q] = q ? ents[q - 1].alpha * inv_dz : 0.0f;
}

// Counts the nonzero-weight voxel operands of cone_fp_kernel (same row
// rules), i.e. nnz of the system matrix.
__global__ __launch_bounds__(kFpRows) void cone_nnz_kernel(
    ConeArgs g, const ConeRay *__restrict__ rays, const ConeEnt *__restrict__ ents,
    unsigned long long *__restrict__ count) {
  extern __shared__ float s_vz[];
  __shared__ unsigned long long s_tot;
  for (int q = threadIdx.x; q <= g.nz; q += kFpRows) s_vz[q] = g.vz[q];
  if (threadIdx.x == 0) s_tot = 0;
  __syncthreads();
  const int r = blockIdx.x;
  const int v = blockIdx.y * kFpRows + threadIdx.x;
  const ConeRay R = rays[r];
  unsigned long long n = 0;
  if (v < g.nv && R.n > 1) {
    const float dzv = g.dz[v], izv = g.iz[v];
    const bool lower = v < g.mid;
    const int dir = lower ? -1 : 1;
    const int nz = g.nz;
    const ConeEnt *__restrict__ e = ents + R.off;
    float a0 = e[0].alpha;
    for (int m = 1; m < R.n; ++m) {
      const float a1 = e[m].alpha;
      const float ai = a0 * g.inv_dz;
      const bool edge = (m >= R.ncom) && (v < R.min_v);
      const int k = (int)((edge ? g.pzbz1 : g.pzbz) + ai * dzv) - (edge ? 1 : 0);
      const int kz = min(max(lower ? k : k + 1, 0), nz);
      const float az = s_vz[kz] * izv;
      const float mz = a1 < az ? a1 : az;
      const int k2 = k + dir;
      n += (k >= 0 && k < nz && mz - a0 != 0.f) ? 1 : 0;
      n += (k2 >= 0 && k2 < nz && a1 - mz != 0.f) ? 1 : 0;
      a0 = a1;
    }
  }
  atomicAdd(&s_tot, n);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(count, s_tot);
}

// ---------------------------------------------------------------------------
// backprojection
// ---------------------------------------------------------------------------

constexpr int kBpWaves = 4;
// BP column tile: consecutive workgroups of one XCD take kBpTile x kBpTile
// adjacent columns, which share most of their rays' detector rows in L2
// (BP at 512^3, tiles of 32 / 16 / 8: 92.0 / 90.5 / 90.6 ms; wide cone
// 768^3: 684 / 674 / 663 ms; 64: 97.6 / 691 ms)
constexpr int kBpTile = 8;
// records per step of the BP sweep (512^3 cone, S = 2, at most 64 VGPRs):
// 5 is the most without register spills, 87.6 ms against 89.2 for 6
// (spills); on another GPU 4 / 6 / 8 took 91.1 / 90.7 / 91.0 ms
template <int S>
constexpr int bp_recs() { return S <= 2 ? 5 : (S <= 4 ? 4 : (S <= 8 ? 2 : 1)); }
// Trash cells on each side of a BP column array (cells -kBpPad .. -1 and
// nz .. nz + kBpPad - 1 take the contributions of slices outside the
// volume); the slice-plane table has the same margin, clamped to planes 0
// and nz.
constexpr int kBpPad = 4;

// Detector columns covered by the four corner lines of one voxel column
// (c2D.cpp:837-901).
__device__ inline void cone_col_hrange(const ConeArgs &g, const double *t,
                                       double x0, double xn, double y0,
                                       double yn, int &hmin, int &hmax) {
  const double c = t[0], s = t[1], p1x = t[2], p1y = t[3];
  const double dx0 = x0 - p1x, dy0 = y0 - p1y;
  const double dxn = xn - p1x, dyn = yn - p1y;
  const double cx0 = c * x0 - g.sx, cxn = c * xn - g.sx;
  const double sy0 = s * y0, syn = s * yn;
  double y00, y01, y10, y11;
  if (fabs(s) > fabs(c)) {
    const double il = t[7];
    y00 = (c - dx0 / (sy0 + cx0)) * il;
    y01 = (c - dx0 / (syn + cx0)) * il;
    y10 = (c - dxn / (sy0 + cxn)) * il;
    y11 = (c - dxn / (syn + cxn)) * il;
  } else {
    const double il = t[6];
    y00 = (-s + dy0 / (cx0 + sy0)) * il;
    y01 = (-s + dyn / (cx0 + syn)) * il;
    y10 = (-s + dy0 / (cxn + sy0)) * il;
    y11 = (-s + dyn / (cxn + syn)) * il;
  }
  const int h00 = (int)floor(y00 * g.ipix + g.hpix0);
  const int h01 = (int)floor(y01 * g.ipix + g.hpix0);
  const int h10 = (int)floor(y10 * g.ipix + g.hpix0);
  const int h11 = (int)floor(y11 * g.ipix + g.hpix0);
  hmin = max(min(min(h00, h01), min(h10, h11)), 0);
  hmax = min(max(max(h00, h01), max(h10, h11)), g.nh - 1);
}

// Alpha interval of ray (a, h) inside the voxel column (c2D.cpp:904-936).
__device__ inline bool cone_col_alpha(const ConeArgs &g, const double *t, int h,
                                      double x0, double xn, double y0, double yn,
                                      float &a0, float &a1) {
  const double c = t[0], s = t[1], p1x = t[2], p1y = t[3];
  const double p2x = t[4] - s * g.h_pix[h];
  const double p2y = t[5] + c * g.h_pix[h];
  const double ddx = p2x - p1x, ddy = p2y - p1y;
  if (fabs(ddx) < kEps || fabs(ddy) < kEps) return false;
  const double ix = 1.0 / ddx, iy = 1.0 / ddy;
  const double ax0 = (x0 - p1x) * ix, ay0 = (y0 - p1y) * iy;
  const double axn = (xn - p1x) * ix, ayn = (yn - p1y) * iy;
  const double amin = fmax(fmax(fmin(ax0, axn), fmin(ay0, ayn)), 0.0);
  const double amax = fmin(fmin(fmax(ax0, axn), fmax(ay0, ayn)), 1.0);
  if (!(amin < amax - kEps)) return false;
  a0 = (float)amin;
  a1 = (float)amax;
  return true;
}

__device__ inline int wave_excl_scan(int x, int lane, int &total) {
  // inclusive scan over 64 lanes with shuffles
  int incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  total = __shfl(incl, 63, 64);
  return incl - x;
}

// Rows whose ray is inside the volume's z range at walk parameter a0:
// alim is increasing below the source row and decreasing above it.
__device__ inline void live_rows(const ConeArgs &g, float a0, int &vlo, int &vhi) {
  int lo = 0, hi = g.mid;  // first lower row with alim > a0
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (g.alim[m] > a0) hi = m; else lo = m + 1;
  }
  vlo = lo;
  lo = g.mid;
  hi = g.nv;  // first upper row with alim <= a0
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (g.alim[m] <= a0) hi = m; else lo = m + 1;
  }
  vhi = lo;
}

__device__ inline int lds_pad4(int n) { return (n + 3) & ~3; }

// Column extent in world units.
struct ColBox {
  double x0, xn, y0, yn;
};
__device__ inline ColBox col_box(const ConeArgs &g, int colid) {
  const int i = colid / g.ny, j = colid - (colid / g.ny) * g.ny;
  return ColBox{g.vx + (double)i * g.dx, g.vx + (double)(i + 1) * g.dx,
                g.vy + (double)j * g.dy, g.vy + (double)(j + 1) * g.dy};
}

// Plan pass 1, one wavefront per column, lanes over angles: the number of
// (angle, h) rays crossing the column, the row-set bounds min_v / max_v
// (c2D.cpp:657-676) and the smallest entry walk parameter (the column's
// lane phases in the phased BP).
__global__ __launch_bounds__(256) void cone_bpcount_kernel(ConeArgs g, BpCol *__restrict__ cols) {
  const int lane = threadIdx.x & 63;
  const int colid = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (colid >= g.nx * g.ny) return;
  const ColBox B = col_box(g, colid);
  const float pzdeps = g.pzdv * (float)(1.0 + kEps);
  int lo_max = 0, hi_min = g.nv, cnt = 0;
  float amin = 3.0e38f;
  for (int a = lane; a < g.na; a += 64) {
    const double *t = g.atab + 8 * a;
    int hmin, hmax;
    cone_col_hrange(g, t, B.x0, B.xn, B.y0, B.yn, hmin, hmax);
    for (int h = hmin; h <= hmax; h++) {
      float a0, a1;
      if (cone_col_alpha(g, t, h, B.x0, B.xn, B.y0, B.yn, a0, a1)) {
        lo_max = max(lo_max, (int)ceilf(pzdeps + g.z1 / a0));
        hi_min = min(hi_min, (int)floorf(g.pzdv + g.znm / a0));
        amin = fminf(amin, a0);
        ++cnt;
      }
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    lo_max = max(lo_max, __shfl_xor(lo_max, d, 64));
    hi_min = min(hi_min, __shfl_xor(hi_min, d, 64));
    amin = fminf(amin, __shfl_xor(amin, d, 64));
    cnt += __shfl_xor(cnt, d, 64);
  }
  if (lane == 0) {
    BpCol c{};
    c.n = cnt;
    c.min_v = lo_max;
    c.max_v = hi_min;
    c.amin = amin;
    cols[colid] = c;
  }
}

// Plan pass 2: the records of each column in angle-then-h order.
__global__ __launch_bounds__(256) void cone_bpfill_kernel(ConeArgs g, const BpCol *__restrict__ cols,
                                                          BpRec *__restrict__ recs) {
  const int lane = threadIdx.x & 63;
  const int colid = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (colid >= g.nx * g.ny) return;
  const ColBox B = col_box(g, colid);
  const BpCol C = cols[colid];
  // Rows past their exit parameter add nothing when every row is in one
  // set and the source is inside the z range; otherwise the window is all
  // rows (with overlapping sets a row can take a lower-set form above the
  // source, c2D.cpp:687-749).
  const bool cut = g.bp_main_safe && C.min_v <= g.mid && C.max_v >= g.mid;
  BpRec *out = recs + C.off;
  int base_pos = 0;
  for (int base = 0; base < g.na; base += 64) {
    const int a = base + lane;
    int hmin = 0, hmax = -1;
    const double *t = g.atab + 8 * (a < g.na ? a : 0);
    if (a < g.na) cone_col_hrange(g, t, B.x0, B.xn, B.y0, B.yn, hmin, hmax);
    int cnt = 0;
    for (int h = hmin; h <= hmax; h++) {
      float a0, a1;
      cnt += cone_col_alpha(g, t, h, B.x0, B.xn, B.y0, B.yn, a0, a1) ? 1 : 0;
    }
    int total;
    int q = base_pos + wave_excl_scan(cnt, lane, total);
    for (int h = hmin; h <= hmax; h++) {
      float a0, a1;
      if (cone_col_alpha(g, t, h, B.x0, B.xn, B.y0, B.yn, a0, a1)) {
        int vlo = 0, vhi = g.nv;
        if (cut) live_rows(g, a0, vlo, vhi);
        out[q] = BpRec{(uint32_t)(a * g.nh + h), __builtin_bit_cast(uint32_t, a0),
                       __builtin_bit_cast(uint32_t, a1),
                       (uint32_t)vlo | ((uint32_t)vhi << 16)};
        ++q;
      }
    }
    base_pos += total;
  }
}

// Per-lane constants of one row slot of the sweep: lane L of slot (vb, t)
// takes row vb + S*L + t.
struct BpSlot {
  float dzv, izv;  // row's z run per unit alpha and its inverse
  float kb;        // slice base of the row's set; kBpOff for lanes not in it
  int ks;          // 1 for the floor-form set (c2D.cpp:687-703), else 0
  int dir;         // second slice: -1 for the lower sets, +1 for the upper
};

// A slice base that sends k (and k +- 1) far below 0: the contribution
// lands in the trash cell.
constexpr float kBpOff = -1.0e9f;

// Row sets of c2D.cpp:687-749: 0 lower edge (v < min_v, floor form),
// 1 lower main, 2 upper main, 3 upper edge.  With one_set every row is in
// exactly one set; otherwise pass p sweeps set p over its rows.
__device__ inline BpSlot bp_slot(const ConeArgs &g, int v, int pass, bool one_set,
                                 int min_v, int max_v) {
  BpSlot s;
  const bool valid = v < g.nv;
  const int vv = valid ? v : g.nv - 1;
  const int mid = g.mid;
  int set;
  bool member;
  if (one_set) {
    set = v < min_v ? 0 : (v < mid ? 1 : (v < max_v ? 2 : 3));
    member = valid;
  } else {
    set = pass;
    member = valid && ((set == 0)   ? (v < min_v)
                       : (set == 1) ? (v >= min_v && v < mid)
                       : (set == 2) ? (v >= mid && v < max_v)
                                    : (v >= max_v));
  }
  s.dzv = g.dz[vv];
  s.izv = g.iz[vv];
  s.kb = member ? (set == 0 ? g.pzbz1 : g.pzbz) : kBpOff;
  s.ks = set == 0 ? 1 : 0;
  s.dir = set <= 1 ? -1 : 1;
  return s;
}

// Backprojection, one wavefront per voxel column: the column's list is
// swept once per block of 64*S detector rows, the block's S row slots (lane
// L of slot t takes row vb + S*L + t) holding their row constants in
// registers (c2D.cpp:657-749 row sets, 1105-1114 weights).  A record skips
// the blocks outside its live row window.
//
// Accumulation: S bounds the rows per slice, so the 64 rows of one slot
// fall on distinct slices, and so do their second slices; slot t has its own
// LDS array, so the slots' first-slice updates are plain read-add-writes
// over distinct cells, issued together, followed by their second-slice
// updates.  The kBpPad cells on each side of an array are trash for slices
// outside the volume.  PHASED (rows per slice beyond 4, or arrays beyond
// the LDS budget): S = 4, 8, 12 or 16 slots and one array; the slots update
// one after another, each in read-add-write rounds, and within a slot the
// lanes go in phases (lane L in phase L mod phases), whose lanes are
// phases * S >= rows per slice rows apart, so on distinct slices.  The
// phases are the column's own (bp_col_phases), at most the layout's
// g.bp_phases.  (LDS float atomics, ds_add_f32, made the kernel 2 to 19
// times slower than these rounds, and rows interleaved over passes re-read
// each record's detector row once per pass.)
// Waves per SIMD the register allocation must allow: the extra waves hide
// the LDS read-add-write latency; for the phased strides the few spills
// they force sit in the per-row-block set-up.  Strides 1 and 2: 8 (S = 2
// at 512^3: 98 ms at 6 waves, 93 at 8).
// Phased, measured on wide cones (ms per BP, waves per SIMD): S = 8 at
// 384^3 32.2 (6) / 29.8 (7) / 29.1 (8); S = 12 at 768^3 442.6 (5) / 403.4 (6)
// / 422.7 (7) / 602.0 (8); S = 16 at 1000^3 576.2 (4) / 503.0 (5) / 868.0
// (6).  Strides 3 and 4 keep the registers their slots need.
constexpr int bp_min_waves(int S) {
  return S <= 2 ? 8 : (S == 8 ? 8 : (S == 12 ? 6 : (S == 16 ? 5 : 1)));
}

// Lane phases of the phased form for one column.  Its rays enter it at walk
// parameter >= C.amin, where adjacent rows are amin * pitch apart in z, so
// at most 1.05 * bp_rps1 / amin rows share a slice in it: the layout's
// bound (cone_bp_layout) taken over the column instead of the volume.
// Columns away from the source need fewer phases than the layout's.
template <int S>
__device__ inline int bp_col_phases(const ConeArgs &g, const BpCol &C) {
  // fminf drops a NaN (no ray enters the column): the cap then applies
  const float want = fminf(1.05f * g.bp_rps1 / C.amin, 1.0e6f);
  const int w = max(1, (int)ceilf(want));
  return min(g.bp_phases, (w + S - 1) / S);
}

// One block of the phased form: lane L holds rows vb + S*L + t, t < S;
// every record's S rows are computed and written slot by slot into the one
// array (cells acc1[-1] and acc1[nz] take the trash).  MAIN: every row of the
// block below nv is a main row (as in the slot form's main path); otherwise
// each row takes its row set's form, pass by pass when the sets overlap (as
// bp_slot).
template <int S, class AccT>
__device__ void bp_phased_block(const ConeArgs &g, const BpRec *__restrict__ R, int n,
                                const float *__restrict__ dpix, const float *s_vz1,
                                AccT *acc1, int vb, int lane, int nph, int ph, float inv_dz,
                                bool main, bool one_set, int min_v, int max_v) {
  constexpr int span = 64 * S;
  const int nz = g.nz, nv = g.nv, mid = g.mid;
  const int vlane = vb + S * lane;
  float dzv[S], izv[S];
#pragma unroll
  for (int t = 0; t < S; ++t) {
    const int vv = min(vlane + t, nv - 1);
    dzv[t] = g.dz[vv];
    izv[t] = g.iz[vv];
  }
  static_assert(S % 4 == 0, "the phased form loads 4 rows at a time");
  auto live = [&](const BpRec &rec) {
    return vb < (int)(rec.w >> 16) && vb + span > (int)(rec.w & 0xffffu);
  };
  auto gather = [&](const BpRec &rec, float *p) {
    // rows vlane .. vlane + S - 1 of the record's padded row, 16 bytes per
    // load (nvp is a multiple of 64 * S, so every row is in range; one
    // dword per load made the wide-cone BP 1.8x slower)
    const float4 *src = (const float4 *)(dpix + (size_t)rec.x * g.nvp + vlane);
#pragma unroll
    for (int j = 0; j < S / 4; ++j) {
      const float4 q = src[j];
      p[4 * j] = q.x;
      p[4 * j + 1] = q.y;
      p[4 * j + 2] = q.z;
      p[4 * j + 3] = q.w;
    }
  };
  auto update = [&](const BpRec &rec, const float *p, int pass, auto MAIN, auto ONE) {
    const float a0 = __builtin_bit_cast(float, rec.y);
    const float a1 = __builtin_bit_cast(float, rec.z);
    const float ai = a0 * inv_dz;
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const int v = vlane + t;
      int i0, i1, pl;
      if constexpr (decltype(MAIN)::value) {
        // padded rows (v >= nv, zero data) write the trash cell: with the
        // last row's constants they would share its cell in this round
        const int k = med3_0s((int)(g.pzbz + ai * dzv[t]), nz - 1);
        const bool lower = v < mid, valid = v < nv;
        i0 = valid ? k : -1;
        i1 = valid ? (lower ? k - 1 : k + 1) : -1;
        pl = lower ? k : k + 1;
      } else {
        const bool valid = v < nv;
        int set;
        bool member;
        if (one_set) {
          set = v < min_v ? 0 : (v < mid ? 1 : (v < max_v ? 2 : 3));
          member = valid;
        } else {
          set = pass;
          member = valid && ((set == 0)   ? (v < min_v)
                             : (set == 1) ? (v >= min_v && v < mid)
                             : (set == 2) ? (v >= mid && v < max_v)
                                          : (v >= max_v));
        }
        const float kb = member ? (set == 0 ? g.pzbz1 : g.pzbz) : kBpOff;
        const int k = (int)(kb + ai * dzv[t]) - (set == 0 ? 1 : 0);
        i0 = med3_i32(k, -1, nz);
        i1 = med3_i32(k + (set <= 1 ? -1 : 1), -1, nz);
        pl = max(i0, i1);
      }
      const float az = s_vz1[pl] * izv[t];
      const float mz = min_sv(a1, az);
      const float c0 = p[t] * (mz - a0);
      const float c1 = p[t] * (a1 - mz);
      // a segment inside one slice (c1 = 0) leaves the second alone
      if constexpr (decltype(ONE)::value) {
        acc1[i0] += c0;
        __builtin_amdgcn_wave_barrier();
        if (c1 != 0.0f) acc1[i1] += c1;
        __builtin_amdgcn_wave_barrier();
      } else {
        for (int q = 0; q < nph; ++q) {
          if (ph == q) acc1[i0] += c0;
          __builtin_amdgcn_wave_barrier();
          if (ph == q && c1 != 0.0f) acc1[i1] += c1;
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
  };
  auto sweep = [&](int pass, auto MAIN, auto ONE) {
    constexpr int kBpRecs = bp_recs<S>();
    int e = 0;
    for (; e + kBpRecs <= n; e += kBpRecs) {
      BpRec r[kBpRecs];
      float p[kBpRecs][S];
#pragma unroll
      for (int u = 0; u < kBpRecs; ++u) r[u] = R[e + u];
#pragma unroll
      for (int u = 0; u < kBpRecs; ++u)
        if (live(r[u])) gather(r[u], p[u]);
#pragma unroll
      for (int u = 0; u < kBpRecs; ++u)
        if (live(r[u])) update(r[u], p[u], pass, MAIN, ONE);
    }
    for (; e < n; ++e) {
      const BpRec r0 = R[e];
      float p0[S];
      if (live(r0)) {
        gather(r0, p0);
        update(r0, p0, pass, MAIN, ONE);
      }
    }
  };
  // ONE: the column runs one lane phase (most do, away from the source),
  // and its rounds go without the phase loop and its lane masks (wide
  // cones: 768^3 374 -> 305 ms, 384^3 27.0 -> 23.8, 1000^3 484 -> 363)
  auto run = [&](auto ONE) {
    if (main) {
      sweep(0, std::true_type{}, ONE);
    } else {
      const int npass = one_set ? 1 : 4;
      for (int pass = 0; pass < npass; ++pass) sweep(pass, std::false_type{}, ONE);
    }
  };
  if (nph == 1)
    run(std::true_type{});
  else
    run(std::false_type{});
}

// AccT is the type of the LDS column arrays: float (the reference's
// accumulation, calc_ah_z c2D.cpp:640-751), or double for the FP64-
// accumulation option, which also leaves out the main-row path (its LDS
// addressing is float-only) and takes the general row-set path instead.
template <int S, bool PHASED, class AccT>
__global__ __launch_bounds__(64 * kBpWaves)
__attribute__((amdgpu_waves_per_eu(bp_min_waves(S), 8))) void cone_bp_kernel(
    ConeArgs g, const BpCol *__restrict__ cols, const BpRec *__restrict__ recs,
    const float *__restrict__ dpix, float *__restrict__ vol, int acc_in,
    double *__restrict__ norm) {
  constexpr int narr = PHASED ? 1 : S;  // LDS arrays per wavefront
  constexpr bool kFloatAcc = std::is_same<AccT, float>::value;
  extern __shared__ float smem[];
  const int nz = g.nz, nv = g.nv;
  // slice planes, clamped at both ends: s_vz1[q] = plane min(max(q, 0), nz)
  // for q in [-kBpPad, nz + kBpPad)
  float *s_vzx = smem;
  for (int q = threadIdx.x; q < nz + 2 * kBpPad; q += 64 * kBpWaves)
    s_vzx[q] = g.vz[min(max(q - kBpPad, 0), nz)];
  __syncthreads();
  const float *s_vz1 = s_vzx + kBpPad;
  const int astride = lds_pad4(nz + 2 * kBpPad);
  // the wave index is uniform: keep the column's loads scalar
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  // logical block -> kBpTile x kBpTile column tile, then kBpWaves adjacent
  // columns of one tile row
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  constexpr int kTile = kBpTile, kTileBlocks = kTile * kTile / kBpWaves;
  const int tile = lb / kTileBlocks, within = lb - tile * kTileBlocks;
  const int ntj = (g.ny + kTile - 1) / kTile;
  const int ci = (tile / ntj) * kTile + within / (kTile / kBpWaves);
  const int cj = (tile % ntj) * kTile + (within % (kTile / kBpWaves)) * kBpWaves + wave;
  if (ci >= g.nx || cj >= g.ny) return;
  const int colid = ci * g.ny + cj;
  const BpCol C = cols[colid];
  float *__restrict__ vcol = vol + (size_t)colid * nz;
  if (C.n == 0) {
    // nothing crosses the column: it stays as it is (zero when not
    // accumulating); its norm partial is that of what it holds
    double ss = 0.0;
    for (int k = lane; k < nz; k += 64) {
      const float o = acc_in ? vcol[k] : 0.0f;
      if (!acc_in) vcol[k] = o;
      ss += (double)o * (double)o;
    }
    if (norm) {
      ss = wave_sum64(ss);
      if (lane == 0) norm[colid] = ss;
    }
    return;
  }
  AccT *acc = reinterpret_cast<AccT *>(smem + astride) + wave * narr * astride;
  for (int q = lane; q < narr * astride; q += 64) acc[q] = (AccT)0;
  __builtin_amdgcn_wave_barrier();
  if (acc_in)
    for (int k = lane; k < nz; k += 64) acc[k + kBpPad] = vcol[k];
  __builtin_amdgcn_wave_barrier();

  const int min_v = C.min_v, max_v = C.max_v;
  const bool one_set = (min_v <= g.mid) && (max_v >= g.mid);
  const int npass = one_set ? 1 : 4;
  // lane phases of the phased form (the column's; uniform in the wave)
  const int nph = PHASED ? bp_col_phases<S>(g, C) : 1;
  const int ph = PHASED ? lane % nph : 0;
  constexpr int span = 64 * S;
  const uint32_t lane_off = 4u * (uint32_t)(S * lane);
  // 1 / voxel z size in a VGPR: a0 * inv_dz is then one v_mul with the
  // record's a0 in an SGPR
  float inv_dz;
  asm("v_mov_b32 %0, %1" : "=v"(inv_dz) : "s"(g.inv_dz));
  const BpRec *__restrict__ R = recs + C.off;
  const int nvp4 = 4 * g.nvp;
  for (int vb = 0; vb < nv; vb += span) {
    if constexpr (PHASED) {
      bp_phased_block<S, AccT>(g, R, C.n, dpix, s_vz1, acc + kBpPad, vb, lane, nph, ph, inv_dz,
                         one_set && g.bp_main_safe && vb >= min_v && min(vb + span, nv) <= max_v,
                         one_set, min_v, max_v);
      continue;
    } else if (kFloatAcc && one_set && g.bp_main_safe) {
      // Every row in one row set (c2D.cpp:657-749): the floor form below
      // min_v (set 0), the truncating form from there on, the second slice
      // below the row's first under the source row g.mid and above it from
      // there on.  A lane's slot constants fold its row's set into
      // addresses: with k = (int)(kb + ai * dz) clamped to [-2, nz + 2],
      // the first slice k - ks is at b0 + 4k, the second at b1 + 4k and the
      // boundary plane (k - ks under the source row, k + 1 from it) at
      // bp + 4k.  Those cells stay within the kBpPad trash cells on each
      // side and the plane table's clamped ends, and the real cells they
      // reach are those of the clamped row-set form below, with the same
      // weights: a k past either clamp sends both updates to trash there
      // too.  Block rows past nv write trash (kb far below slice 0).
      float kb[S], dzv[S], izv[S];
      uint32_t b0[S], b1[S], bz[S];
#pragma unroll
      for (int t = 0; t < S; ++t) {
        const int v = vb + S * lane + t;
        const int vv = min(v, nv - 1);
        const bool lower = v < g.mid, floor_form = v < min_v;
        kb[t] = v < nv ? (floor_form ? g.pzbz1 : g.pzbz) : kBpOff;
        dzv[t] = g.dz[vv];
        izv[t] = g.iz[vv];
        const uint32_t ks4 = floor_form ? 4u : 0u;
        b0[t] = lds_addr(reinterpret_cast<const float *>(acc + t * astride + kBpPad)) - ks4;
        b1[t] = lower ? b0[t] - 4u : b0[t] + 4u;
        bz[t] = lds_addr(s_vz1) - ks4 + (lower ? 0u : 4u);
      }
      const int khi = nz + 2;
      auto live = [&](const BpRec &rec) {
        return vb < (int)(rec.w >> 16) && vb + span > (int)(rec.w & 0xffffu);
      };
      auto gather = [&](const BpRec &rec, float *p) {
        const __amdgpu_buffer_rsrc_t prow = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(dpix + (size_t)rec.x * g.nvp), (short)0, nvp4, kRsrcWord3);
#pragma unroll
        for (int t = 0; t < S; ++t)
          p[t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                              prow, (int)lane_off, 4 * (vb + t), 0));
      };
      // DIR -1 / +1: every row of the block under / from the source row,
      // the second slice one cell below / above the first (an immediate
      // offset); 0: both kinds of rows
      auto update = [&](const BpRec &rec, const float *p, auto DIR) {
        constexpr int dir = decltype(DIR)::value;
        const float a0 = __builtin_bit_cast(float, rec.y);
        const float a1 = __builtin_bit_cast(float, rec.z);
        const float ai = a0 * inv_dz;
        // the S slots write S separate arrays: their first-slice updates
        // go together, then (after the lanes' first slices are written)
        // their second-slice updates.  A row whose segment stays in one
        // slice (mz = a1) has c1 = 0 and leaves its second slice alone:
        // adding a zero changes no cell, and the lanes left out of an LDS
        // round take no bank cycles (most segments cross no plane)
        lds_f32 *c0p[S], *c1p[S];
        float c0[S], c1[S], o0[S];
#pragma unroll
        for (int t = 0; t < S; ++t) {
          const int k = med3_m2s((int)(kb[t] + ai * dzv[t]), khi);
          if constexpr (dir < 0) {
            c1p[t] = lds_ptr(k4_add(k, b1[t]));
            c0p[t] = c1p[t] + 1;
          } else if constexpr (dir > 0) {
            c0p[t] = lds_ptr(k4_add(k, b0[t]));
            c1p[t] = c0p[t] + 1;
          } else {
            c0p[t] = lds_ptr(k4_add(k, b0[t]));
            c1p[t] = lds_ptr(k4_add(k, b1[t]));
          }
          const float az = *lds_ptr(k4_add(k, bz[t])) * izv[t];
          const float mz = min_sv(a1, az);
          c0[t] = p[t] * (mz - a0);
          c1[t] = p[t] * (a1 - mz);
        }
#pragma unroll
        for (int t = 0; t < S; ++t) o0[t] = *c0p[t];
#pragma unroll
        for (int t = 0; t < S; ++t) *c0p[t] = o0[t] + c0[t];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < S; ++t)
          if (c1[t] != 0.0f) *c1p[t] = *c1p[t] + c1[t];
        __builtin_amdgcn_wave_barrier();
      };
      // kBpRecs records at a time: all their gathers are in flight before
      // the first update (the updates stay in record order); the live
      // tests go into one uniform bit mask
      auto sweep = [&](auto DIR) {
        int e = 0;
        constexpr int kBpRecs = bp_recs<S>();
        for (; e + kBpRecs <= C.n; e += kBpRecs) {
          BpRec r[kBpRecs];
          float p[kBpRecs][S];
          uint32_t lm = 0;
#pragma unroll
          for (int u = 0; u < kBpRecs; ++u) r[u] = R[e + u];
#pragma unroll
          for (int u = 0; u < kBpRecs; ++u) lm |= live(r[u]) ? 1u << u : 0u;
#pragma unroll
          for (int u = 0; u < kBpRecs; ++u)
            if (lm & (1u << u)) gather(r[u], p[u]);
#pragma unroll
          for (int u = 0; u < kBpRecs; ++u)
            if (lm & (1u << u)) update(r[u], p[u], DIR);
        }
        for (; e < C.n; ++e) {
          const BpRec r0 = R[e];
          float p0[S];
          if (live(r0)) {
            gather(r0, p0);
            update(r0, p0, DIR);
          }
        }
      };
      if (vb + span <= g.mid)
        sweep(std::integral_constant<int, -1>{});
      else if (vb >= g.mid)
        sweep(std::integral_constant<int, 1>{});
      else
        sweep(std::integral_constant<int, 0>{});
      continue;
    }
    for (int pass = 0; pass < npass; ++pass) {
      BpSlot sl[S];
#pragma unroll
      for (int t = 0; t < S; ++t)
        sl[t] = bp_slot(g, vb + S * lane + t, pass, one_set, min_v, max_v);
      auto live = [&](const BpRec &rec) {
        return vb < (int)(rec.w >> 16) && vb + span > (int)(rec.w & 0xffffu);
      };
      auto gather = [&](const BpRec &rec, float *p) {
        const __amdgpu_buffer_rsrc_t prow = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(dpix + (size_t)rec.x * g.nvp), (short)0, nvp4, kRsrcWord3);
#pragma unroll
        for (int t = 0; t < S; ++t)
          p[t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                              prow, (int)lane_off, 4 * (vb + t), 0));
      };
      auto update = [&](const BpRec &rec, const float *p) {
        const float a0 = __builtin_bit_cast(float, rec.y);
        const float a1 = __builtin_bit_cast(float, rec.z);
        const float ai = a0 * inv_dz;
        AccT *c0p[S], *c1p[S];
        float c0[S], c1[S];
#pragma unroll
        for (int t = 0; t < S; ++t) {
          const int k = (int)(sl[t].kb + ai * sl[t].dzv) - sl[t].ks;
          const int i0 = med3_i32(k, -1, nz), i1 = med3_i32(k + sl[t].dir, -1, nz);
          // boundary plane between the two slices (k below the source,
          // k + 1 above it); clamping the larger index equals clamping the
          // plane index to [0, nz]
          const float az = s_vz1[max(i0, i1)] * sl[t].izv;
          const float mz = min_sv(a1, az);
          c0[t] = p[t] * (mz - a0);
          c1[t] = p[t] * (a1 - mz);
          AccT *ac = acc + t * astride + kBpPad;
          c0p[t] = ac + i0;
          c1p[t] = ac + i1;
        }
        {
          // separate arrays per slot: first-slice updates together, then
          // the second-slice updates of the segments that cross a plane
          AccT o[S];
#pragma unroll
          for (int t = 0; t < S; ++t) o[t] = *c0p[t];
#pragma unroll
          for (int t = 0; t < S; ++t) *c0p[t] = o[t] + c0[t];
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int t = 0; t < S; ++t)
            if (c1[t] != 0.0f) *c1p[t] = *c1p[t] + c1[t];
          __builtin_amdgcn_wave_barrier();
        }
      };
      // kBpRecs records at a time: all their gathers are in flight before
      // the first update (the updates stay in record order)
      int e = 0;
      constexpr int kBpRecs = bp_recs<S>();
      for (; e + kBpRecs <= C.n; e += kBpRecs) {
        BpRec r[kBpRecs];
        float p[kBpRecs][S];
#pragma unroll
        for (int u = 0; u < kBpRecs; ++u) r[u] = R[e + u];
#pragma unroll
        for (int u = 0; u < kBpRecs; ++u)
          if (live(r[u])) gather(r[u], p[u]);
#pragma unroll
        for (int u = 0; u < kBpRecs; ++u)
          if (live(r[u])) update(r[u], p[u]);
      }
      for (; e < C.n; ++e) {
        const BpRec r0 = R[e];
        float p0[S];
        if (live(r0)) {
          gather(r0, p0);
          update(r0, p0);
        }
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  double ss = 0.0;
  for (int k = lane; k < nz; k += 64) {
    AccT s = acc[k + kBpPad];
#pragma unroll
    for (int t = 1; t < narr; ++t) s += acc[t * astride + k + kBpPad];
    const float o = (float)s;
    vcol[k] = o;
    ss += (double)o * (double)o;
  }
  if (norm) {
    ss = wave_sum64(ss);
    if (lane == 0) norm[colid] = ss;
  }
}

// dpix[r][v] = pix[r][v] * d_conv[h][v] for v < nv, 0 up to the padded row
// stride nvp (c2D.cpp:1105-1114).  A workgroup takes whole rays, its
// threads the rows, 16 bytes at a time where the rows allow it (dconv and
// dpix are the plan's and a workspace's: 256-byte aligned).
__global__ void cone_scale_kernel(const float *__restrict__ pix,
                                  const float *__restrict__ dconv,
                                  float *__restrict__ dpix, size_t nrays, int nh, int nv,
                                  int nvp) {
  const bool vec = ((nv & 3) | (reinterpret_cast<uintptr_t>(pix) & 15)) == 0;
  for (size_t r = blockIdx.x; r < nrays; r += gridDim.x) {
    const float *__restrict__ dc = dconv + (r % (size_t)nh) * (size_t)nv;
    const float *__restrict__ pr = pix + r * (size_t)nv;
    float *__restrict__ dp = dpix + r * (size_t)nvp;
    if (vec) {
      const int nv4 = nv >> 2, nvp4 = nvp >> 2;
      for (int v = threadIdx.x; v < nvp4; v += blockDim.x) {
        float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
        if (v < nv4) {
          const float4 pv = reinterpret_cast<const float4 *>(pr)[v];
          const float4 cv = reinterpret_cast<const float4 *>(dc)[v];
          out = make_float4(pv.x * cv.x, pv.y * cv.y, pv.z * cv.z, pv.w * cv.w);
        }
        reinterpret_cast<float4 *>(dp)[v] = out;
      }
    } else {
      for (int v = threadIdx.x; v < nvp; v += blockDim.x) dp[v] = v < nv ? pr[v] * dc[v] : 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// largest row stride of the slot form (one LDS array per slot, one kernel
// each); beyond it the phased form
constexpr int kBpMaxStride = 4;

// LDS of one BP workgroup: the slice planes and `narr` column arrays per
// wavefront of `acc_bytes`-byte cells.
static size_t bp_lds_bytes(int nz, int narr, size_t acc_bytes = sizeof(float)) {
  const size_t pad = (size_t)((nz + 2 * kBpPad + 3) & ~3);
  return sizeof(float) * pad + acc_bytes * (size_t)kBpWaves * (size_t)narr * pad;
}

// BP accumulation layout of a cone geometry (host only, no device work).
// Row stride: at least the largest number of detector rows per slice over
// all columns.  The walk parameter of any voxel is at least
// (|source| - R) / L, R the farthest voxel from the axis and L the longest
// source -> pixel xy distance.
static int bp_rows_bound(double rows_per_slice) {
  return std::max(1, (int)std::ceil(std::min(rows_per_slice * 1.05, 1.0e6)));
}

BpLayout cone_bp_layout(const Geometry &G) {
  const int nh = G.nh, nv = G.nv, nx = G.nx, ny = G.ny, nz = G.nz;
  const double sx = G.source[0], sy = G.source[1], sz = G.source[2];
  BpLayout L;
  double R = 0.0;
  for (int cx = 0; cx < 2; cx++)
    for (int cy = 0; cy < 2; cy++)
      R = std::max(R, std::hypot(G.origin[0] + cx * nx * G.vsize[0],
                                 G.origin[1] + cy * ny * G.vsize[1]));
  double Lmax = 0.0;
  for (int h : {0, nh - 1})
    Lmax = std::max(Lmax, std::hypot(G.det_x - sx, G.h_pix[h] - sy));
  // (positive: the geometry has the source outside the corner circle)
  const double amin = (std::hypot(sx, sy) - R) / Lmax;
  double pitch = 1e300;
  for (int v = 1; v < nv; v++) pitch = std::min(pitch, G.v_pix[v] - G.v_pix[v - 1]);
  // a source close to the corner circle gives a tiny amin and a bound
  // past 64 * 16 rows: the phased form then runs one lane per phase,
  // which keeps every round's cells distinct whatever the rows per slice
  L.rps1 = G.vsize[2] / pitch;
  const double rows_per_slice = L.rps1 / amin;
  const int want = bp_rows_bound(rows_per_slice);
  L.stride = std::min(16, want);
  // lanes S rows apart are >= 1.05 slices apart: distinct slices
  L.unique = want <= 16 ? 1 : 0;
  // a source inside the volume's z range keeps main rows' slices in range
  // and uniformly spaced rows make min_v / max_v exact row bounds
  // (they are computed from the first pixel step, c2D.cpp:995-1000)
  bool uniform = true;
  const double st0 = G.v_pix[1] - G.v_pix[0];
  for (int v = 1; v < nv; v++)
    uniform = uniform && std::fabs((G.v_pix[v] - G.v_pix[v - 1]) - st0) <= 1e-9 * std::fabs(st0);
  L.main_safe =
      (uniform && sz >= G.origin[2] && sz < G.origin[2] + (double)nz * G.vsize[2]) ? 1 : 0;
  // one LDS array per row slot while they fit two workgroups per CU
  // beside each other; otherwise one array, updated in lane phases
  L.phased =
      (L.unique && L.stride <= kBpMaxStride && bp_lds_bytes(nz, L.stride) <= 64 * 1024)
          ? 0
          : 1;
  L.phases = 1;
  if (L.phased) {
    // 4, 8, 12 or 16 slots; lanes of one phase are phases * S >= want rows
    // apart.  Past 4 rows per slice, the S of least (padded rows) x (1 +
    // 0.11 per extra phase), the smaller S on a tie: measured on wide
    // cones, 768 rows: S = 16 681 ms, S = 8 in 2 phases 756 ms; 384 rows:
    // 83.6 and 47.9 ms.  These are the layout's phases, for the column
    // nearest the source; each column runs its own (bp_col_phases).
    auto phases = [&](int S) { return std::min(64, (want + S - 1) / S); };
    auto cost = [&](int S) {
      const double rows = (double)((nv + 64 * S - 1) / (64 * S)) * 64.0 * S;
      return rows * (1.0 + 0.11 * (phases(S) - 1));
    };
    L.stride = 4;
    if (want > 4) {
      L.stride = 8;
      for (int S : {12, 16})
        if (cost(S) < cost(L.stride)) L.stride = S;
    }
    L.phases = phases(L.stride);
  }
  CCPI_REQUIRE(bp_lds_bytes(nz, L.phased ? 1 : L.stride) <= 160 * 1024, CCPI_E_LIMIT,
               "cone BP: nz too large for LDS");
  // rows of one BP block: 64 lanes bp_stride rows apart (both forms)
  const int span = 64 * L.stride;
  L.nvp = (nv + span - 1) / span * span;
  CCPI_REQUIRE(nv < 65536, CCPI_E_LIMIT, "cone beam: at most 65535 detector rows");
  L.want = want;
  return L;
}

void cone_build_plan(ConePlan &P, hipStream_t st) {
  const Geometry &G = P.g;
  const int na = G.na, nh = G.nh, nv = G.nv, nx = G.nx, ny = G.ny, nz = G.nz;
  const double sx = G.source[0], sy = G.source[1], sz = G.source[2];
  ConeArgs A{};
  A.na = na; A.nh = nh; A.nv = nv; A.nx = nx; A.ny = ny; A.nz = nz;
  // tables (c2D.cpp:480-516, 993-1010)
  A.mid = nv;
  for (int v = 0; v < nv; v++)
    if (G.v_pix[v] > sz) { A.mid = v; break; }
  A.inv_dz = (float)(1.0 / G.vsize[2]);
  A.pzbz = (float)((sz - G.origin[2]) / G.vsize[2]);
  A.pzbz1 = A.pzbz + 1.0f;
  const double vstep = G.v_pix[1] - G.v_pix[0];
  A.pzdv = (float)((sz - G.v_pix[0]) / vstep);
  A.z1 = (float)((G.origin[2] + G.vsize[2] - sz) / vstep);
  A.znm = (float)((G.origin[2] + (double)(nz - 1) * G.vsize[2] - sz) / vstep);
  A.vx = G.origin[0]; A.vy = G.origin[1];
  A.dx = G.vsize[0]; A.dy = G.vsize[1];
  A.sx = sx; A.sy = sy; A.det_x = G.det_x;
  A.L = G.det_x - sx;
  const double step = G.h_pix[1] - G.h_pix[0];
  A.ipix = 1.0 / step;
  A.hpix0 = (sy - G.h_pix[0]) / step;

  std::vector<float> dz(nv), iz(nv), vz(nz + 1), dconv((size_t)nh * nv);
  for (int v = 0; v < nv; v++) {
    dz[v] = (float)(G.v_pix[v] - sz);
    iz[v] = (float)(1.0 / (double)dz[v]);
  }
  for (int k = 0; k <= nz; k++) vz[k] = (float)(G.origin[2] + (double)k * G.vsize[2] - sz);
  {
    // the FP forms byte offsets 4 * (slice +- 1) in 32-bit integers
    double kmax = std::fabs((double)A.pzbz) + 2.0;
    for (int v = 0; v < nv; v++) kmax = std::max(kmax, std::fabs((double)A.pzbz) + 2.0 +
                                                           std::fabs((double)dz[v]) * A.inv_dz);
    CCPI_REQUIRE(kmax < (double)(1 << 28), CCPI_E_LIMIT,
                 "cone geometry: detector rows span too many voxel slices");
  }
  const double x2 = (G.det_x - sx) * (G.det_x - sx);
  for (int h = 0; h < nh; h++) {
    const double xy2 = x2 + (G.h_pix[h] - sy) * (G.h_pix[h] - sy);
    for (int v = 0; v < nv; v++)
      dconv[(size_t)h * nv + v] =
          (float)std::sqrt(xy2 + (G.v_pix[v] - sz) * (G.v_pix[v] - sz));
  }
  // per-angle tables and whole-volume footprint (c2D.cpp:549-600, 978-991)
  std::vector<double> atab((size_t)na * 8);
  std::vector<int> hr((size_t)na * 2);
  const double vx = G.origin[0], wx = G.origin[0] + (double)nx * G.vsize[0];
  const double vy = G.origin[1], wy = G.origin[1] + (double)ny * G.vsize[1];
  for (int a = 0; a < na; a++) {
    const double c = std::cos(G.angles[a]), s = std::sin(G.angles[a]);
    double *t = &atab[(size_t)a * 8];
    t[0] = c; t[1] = s;
    t[2] = c * sx - s * sy;
    t[3] = s * sx + c * sy;
    t[4] = c * G.det_x;
    t[5] = s * G.det_x;
    t[6] = A.L / c;
    t[7] = A.L / s;
    const double p1x = t[2], p1y = t[3];
    const double dx0 = vx - p1x, dy0 = vy - p1y, dxn = wx - p1x, dyn = wy - p1y;
    const double cx0 = c * vx - sx, cxn = c * wx - sx, sy0 = s * vy, syn = s * wy;
    double y00, y01, y10, y11;
    if (std::fabs(s) > std::fabs(c)) {
      const double il = A.L / s;
      y00 = (c - dx0 / (sy0 + cx0)) * il;
      y01 = (c - dx0 / (syn + cx0)) * il;
      y10 = (c - dxn / (sy0 + cxn)) * il;
      y11 = (c - dxn / (syn + cxn)) * il;
    } else {
      const double il = A.L / c;
      y00 = (-s + dy0 / (cx0 + sy0)) * il;
      y01 = (-s + dyn / (cx0 + syn)) * il;
      y10 = (-s + dy0 / (cxn + sy0)) * il;
      y11 = (-s + dyn / (cxn + syn)) * il;
    }
    const int h00 = (int)std::floor(y00 * A.ipix + A.hpix0);
    const int h01 = (int)std::floor(y01 * A.ipix + A.hpix0);
    const int h10 = (int)std::floor(y10 * A.ipix + A.hpix0);
    const int h11 = (int)std::floor(y11 * A.ipix + A.hpix0);
    hr[2 * a] = std::max(std::min(std::min(h00, h01), std::min(h10, h11)), 0);
    hr[2 * a + 1] = std::min(std::max(std::max(h00, h01), std::max(h10, h11)), nh - 1);
  }
  {
    const BpLayout L = cone_bp_layout(G);
    A.bp_stride = L.stride;
    A.bp_unique = L.unique;
    A.bp_main_safe = L.main_safe;
    A.bp_phased = L.phased;
    A.bp_phases = L.phases;
    A.bp_rps1 = (float)L.rps1;
    A.nvp = L.nvp;
  }
  // Walk parameter past which a row's ray is outside the volume's z range
  // for good: a row below the source is kLimSlices slices under the bottom
  // face (its floor-form slice index, and the one below, are then < 0), a
  // row above it kLimSlices slices over the top (both indices >= nz).  The
  // margin covers the float rounding of the kernels' slice position.
  // Without a source inside the z range the cut is off (2 > any alpha).
  A.zb = (float)(G.origin[2] - sz);
  A.zt = (float)(G.origin[2] + (double)nz * G.vsize[2] - sz);
  constexpr double kLimSlices = 0.05;
  std::vector<float> alim(nv, 2.0f);
  if (A.bp_main_safe)
    for (int v = 0; v < nv; v++) {
      const double d = (double)dz[v];
      if (d == 0.0) continue;
      const double lim = v < A.mid ? ((double)A.zb - kLimSlices * G.vsize[2]) / d
                                   : ((double)A.zt + kLimSlices * G.vsize[2]) / d;
      alim[v] = (float)std::min(2.0, lim);
    }
  P.alim.upload(alim, st);
  A.alim = P.alim.as<float>();
  P.atab.upload(atab, st);
  P.hpix.upload(G.h_pix, st);
  P.hrange.upload(hr, st);
  P.dz.upload(dz, st);
  P.iz.upload(iz, st);
  P.vz.upload(vz, st);
  P.dconv.upload(dconv, st);
  A.atab = P.atab.as<double>();
  A.h_pix = P.hpix.as<double>();
  A.hrange = P.hrange.as<int>();
  A.dz = P.dz.as<float>();
  A.iz = P.iz.as<float>();
  A.vz = P.vz.as<float>();
  A.dconv = P.dconv.as<float>();

  // walk: count, scan, fill
  const size_t nrays = (size_t)na * nh;
  DeviceBuffer counts;
  counts.alloc(nrays * sizeof(uint32_t));
  P.offs.alloc(nrays * sizeof(uint64_t));
  P.rays.alloc(nrays * sizeof(ConeRay));
  const int blocks = (int)((nrays + 255) / 256);
  hipLaunchKernelGGL(cone_walk_kernel<false>, dim3(blocks), dim3(256), 0, st, A,
                     nullptr, counts.as<uint32_t>(), nullptr, nullptr);
  CCPI_HIP(hipGetLastError());
  std::vector<uint32_t> hc(nrays);
  CCPI_HIP(hipMemcpyAsync(hc.data(), counts.p, nrays * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, st));
  CCPI_HIP(hipStreamSynchronize(st));
  std::vector<uint64_t> ho(nrays);
  uint64_t tot = 0;
  for (size_t q = 0; q < nrays; q++) {
    ho[q] = tot;
    tot += hc[q];
  }
  P.n_entries = tot;
  P.offs.upload(ho, st);
  P.ents.alloc(std::max<uint64_t>(tot, 1) * sizeof(ConeEnt));
  hipLaunchKernelGGL(cone_walk_kernel<true>, dim3(blocks), dim3(256), 0, st, A,
                     P.offs.as<uint64_t>(), nullptr, P.rays.as<ConeRay>(),
                     P.ents.as<ConeEnt>());
  CCPI_HIP(hipGetLastError());
  P.eai.alloc(std::max<uint64_t>(tot, 1) * sizeof(float));
  hipLaunchKernelGGL(cone_ai_kernel, dim3(2048), dim3(256), 0, st, P.ents.as<ConeEnt>(),
                     (uint64_t)tot, A.inv_dz, P.eai.as<float>());
  CCPI_HIP(hipGetLastError());

  // backprojection lists: per-column count and row-set bounds, scan, fill
  {
    const size_t ncols = (size_t)nx * ny;
    P.bpcol.alloc(ncols * sizeof(BpCol));
    const int cblocks = (int)((ncols * 64 + 255) / 256);
    hipLaunchKernelGGL(cone_bpcount_kernel, dim3(cblocks), dim3(256), 0, st, A,
                       P.bpcol.as<BpCol>());
    CCPI_HIP(hipGetLastError());
    std::vector<BpCol> hcol(ncols);
    CCPI_HIP(hipMemcpyAsync(hcol.data(), P.bpcol.p, ncols * sizeof(BpCol),
                            hipMemcpyDeviceToHost, st));
    CCPI_HIP(hipStreamSynchronize(st));
    uint64_t nrec = 0;
    for (auto &c : hcol) {
      c.off = nrec;
      nrec += (uint64_t)c.n;
    }
    P.bpcol.upload(hcol, st);
    P.bprec.alloc(std::max<uint64_t>(nrec, 1) * sizeof(BpRec));
    hipLaunchKernelGGL(cone_bpfill_kernel, dim3(cblocks), dim3(256), 0, st, A,
                       P.bpcol.as<BpCol>(), P.bprec.as<BpRec>());
    CCPI_HIP(hipGetLastError());
    CCPI_HIP(hipStreamSynchronize(st));  // hcol is the upload's source
  }
  CCPI_HIP(hipStreamSynchronize(st));
  P.args = A;
}

size_t cone_fp_norm_parts(const ConePlan &P) {
  const ConeArgs &A = P.args;
  return (size_t)A.na * A.nh * (size_t)((A.nv + kFpRows - 1) / kFpRows);
}

void cone_fp(const ConePlan &P, const float *vol, float *sino, hipStream_t st, bool f64,
             bool acc_in, double *norm) {
  const ConeArgs &A = P.args;
  dim3 grid((unsigned)(A.na * A.nh), (unsigned)((A.nv + kFpRows - 1) / kFpRows));
  const size_t lds = sizeof(float) * (size_t)(A.nz + 1);
  const int ai = acc_in ? 1 : 0;
  if (f64)
    hipLaunchKernelGGL(cone_fp_kernel<double>, grid, dim3(kFpRows), lds, st, A,
                       P.rays.as<ConeRay>(), P.ents.as<ConeEnt>(), P.eai.as<float>(), vol,
                       sino, ai, norm);
  else
    hipLaunchKernelGGL(cone_fp_kernel<float>, grid, dim3(kFpRows), lds, st, A,
                       P.rays.as<ConeRay>(), P.ents.as<ConeEnt>(), P.eai.as<float>(), vol,
                       sino, ai, norm);
  CCPI_HIP(hipGetLastError());
}

void cone_nnz(const ConePlan &P, unsigned long long *count, hipStream_t st) {
  const ConeArgs &A = P.args;
  dim3 grid((unsigned)(A.na * A.nh), (unsigned)((A.nv + kFpRows - 1) / kFpRows));
  const size_t lds = sizeof(float) * (size_t)(A.nz + 1);
  hipLaunchKernelGGL(cone_nnz_kernel, grid, dim3(kFpRows), lds, st, A,
                     P.rays.as<ConeRay>(), P.ents.as<ConeEnt>(), count);
  CCPI_HIP(hipGetLastError());
}

size_t cone_dpix_floats(const ConePlan &P) {
  return (size_t)P.args.na * P.args.nh * (size_t)P.args.nvp;
}

void cone_scale(const ConePlan &P, const float *pix, float *dpix, hipStream_t st) {
  const ConeArgs &A = P.args;
  const size_t nrays = (size_t)A.na * A.nh;
  const int blocks = (int)std::min<size_t>(nrays, 16384);
  hipLaunchKernelGGL(cone_scale_kernel, dim3(blocks), dim3(256), 0, st, pix, A.dconv, dpix,
                     nrays, A.nh, A.nv, A.nvp);
  CCPI_HIP(hipGetLastError());
}

size_t cone_bp_norm_parts(const ConePlan &P) { return (size_t)P.args.nx * P.args.ny; }

template <class AccT>
static void cone_bp_launch(const ConePlan &P, const float *dpix, float *vol, hipStream_t st,
                           bool acc_in, double *norm) {
  const int ai = acc_in ? 1 : 0;
  const ConeArgs &A = P.args;
  // kBpTile x kBpTile column tiles of kBpTile^2 / kBpWaves workgroups
  // (kernel mapping)
  const int blocks = ((A.nx + kBpTile - 1) / kBpTile) * ((A.ny + kBpTile - 1) / kBpTile) *
                     (kBpTile * kBpTile / kBpWaves);
  const int narr = A.bp_phased ? 1 : A.bp_stride;  // LDS arrays per wavefront
  const size_t lds = bp_lds_bytes(A.nz, narr, sizeof(AccT));
  CCPI_REQUIRE(lds <= 160 * 1024, CCPI_E_LIMIT,
               "cone BP: nz too large for LDS with FP64 accumulation");
  const BpCol *bc = P.bpcol.as<BpCol>();
  const BpRec *br = P.bprec.as<BpRec>();
  const dim3 grid(blocks), block(64 * kBpWaves);
  if (A.bp_phased && A.bp_stride == 4)
    hipLaunchKernelGGL((cone_bp_kernel<4, true, AccT>), grid, block, lds, st, A, bc, br, dpix,
                       vol, ai, norm);
  else if (A.bp_phased && A.bp_stride == 8)
    hipLaunchKernelGGL((cone_bp_kernel<8, true, AccT>), grid, block, lds, st, A, bc, br, dpix,
                       vol, ai, norm);
  else if (A.bp_phased && A.bp_stride == 12)
    hipLaunchKernelGGL((cone_bp_kernel<12, true, AccT>), grid, block, lds, st, A, bc, br, dpix,
                       vol, ai, norm);
  else if (A.bp_phased)
    hipLaunchKernelGGL((cone_bp_kernel<16, true, AccT>), grid, block, lds, st, A, bc, br, dpix,
                       vol, ai, norm);
  else if (narr == 1)
    hipLaunchKernelGGL((cone_bp_kernel<1, false, AccT>), grid, block, lds, st, A, bc, br, dpix,
                       vol, ai, norm);
  else if (narr == 2)
    hipLaunchKernelGGL((cone_bp_kernel<2, false, AccT>), grid, block, lds, st, A, bc, br, dpix,
                       vol, ai, norm);
  else if (narr == 3)
    hipLaunchKernelGGL((cone_bp_kernel<3, false, AccT>), grid, block, lds, st, A, bc, br, dpix,
                       vol, ai, norm);
  else
    hipLaunchKernelGGL((cone_bp_kernel<4, false, AccT>), grid, block, lds, st, A, bc, br, dpix,
                       vol, ai, norm);
  CCPI_HIP(hipGetLastError());
}

void cone_bp_scaled(const ConePlan &P, const float *dpix, float *vol, hipStream_t st,
                    bool f64, bool acc_in, double *norm) {
  if (f64)
    cone_bp_launch<double>(P, dpix, vol, st, acc_in, norm);
  else
    cone_bp_launch<float>(P, dpix, vol, st, acc_in, norm);
}

}  // namespace ccpi
