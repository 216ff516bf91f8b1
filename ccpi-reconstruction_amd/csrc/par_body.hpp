// Per-thread bodies of the parallel-beam projectors, shared by the
// one-problem kernels (parallel.hip) and the batched persistent kernels
// (batched.hip).  Model: Core/src/p2D.cpp (see parallel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace ccpi {

// Device memory reached through a pointer typed as global: loads and stores
// compile to global (not flat) accesses.  The batched kernels read their
// problems' pointers from descriptors in device memory, which leaves them
// generic to the compiler; generic accesses wait on the scalar and vector
// memory counters both and need 64-bit addresses in vector registers.  The
// bodies below reach all their arrays through gmem.
#define CCPI_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ inline CCPI_GLOBAL T *gmem(T *p) {
  return (CCPI_GLOBAL T *)p;
}

// 16 bytes at p (4-byte aligned) as one load.
__device__ inline float4 load_f4(const CCPI_GLOBAL float *p) {
  float4 v;
  __builtin_memcpy(&v, p, sizeof(v));
  return v;
}

// One 8-byte walk-table entry.
__device__ inline uint2 load_u2(const CCPI_GLOBAL uint2 *p) {
  uint2 v;
  __builtin_memcpy(&v, p, sizeof(v));
  return v;
}

// The detector line of ray (a, h) walked through the xy grid (p2D.cpp:
// 215-458): add(x, y, l) for every crossed cell, l its xy length.
template <class Add>
__device__ inline void par_walk(const ParArgs &g, int a, int h, Add &&add) {
  const int nx = g.nx, ny = g.ny;
  const double c = gmem(g.atab)[5 * a], s = gmem(g.atab)[5 * a + 1];
  const double hp = gmem(g.h_pix)[h];
  const double p2x = c * g.D - s * hp;
  const double p2y = s * g.D + c * hp;
  if (fabs(c) < kEps) {
    if (fabs(s) >= kEps) {  // line parallel to y (p2D.cpp:233-260)
      const int i = (int)floor((p2x - g.ox) / g.dx);
      if (i >= 0 && i < nx) {
        const float l = (float)g.dy;
        if (s < 0.0)
          for (int j = ny - 1; j >= 0; j--) add(i, j, l);
        else
          for (int j = 0; j < ny; j++) add(i, j, l);
      }
    }
  } else if (fabs(s) < kEps) {  // line parallel to x (p2D.cpp:261-283)
    const int j = (int)floor((p2y - g.oy) / g.dy);
    if (j >= 0 && j < ny) {
      const float l = (float)g.dx;
      if (c < 0.0)
        for (int i = nx - 1; i >= 0; i--) add(i, j, l);
      else
        for (int i = 0; i < nx; i++) add(i, j, l);
    }
  } else {  // general line (p2D.cpp:284-458)
    const double ix = 1.0 / c, iy = 1.0 / s;
    const double ax0 = (g.ox - p2x) * ix, ay0 = (g.oy - p2y) * iy;
    const double axn = (g.ox + (double)nx * g.dx - p2x) * ix;
    const double ayn = (g.oy + (double)ny * g.dy - p2y) * iy;
    const double amin = fmax(fmax(fmin(ax0, axn), fmin(ay0, ayn)), -g.dconv);
    const double amax = fmin(fmin(fmax(ax0, axn), fmax(ay0, ayn)), 0.0);
    if (amin < amax - kEps) {
      const int sx = c > 0.0 ? 1 : -1, sy = s > 0.0 ? 1 : -1;
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
          y = on_y ? yin : (int)floor((p2y + amin * s - g.oy) / g.dy);
        } else if (on_y) {
          y = yin;
          x = (int)floor((p2x + amin * c - g.ox) / g.dx);
        } else {
          x = 0;
          y = 0;
        }
        x = min(max(x, 0), nx - 1);
        y = min(max(y, 0), ny - 1);
      }
      const int bxs = sx > 0, bys = sy > 0;
      double ap = amin;
      while (x >= 0 && x < nx && y >= 0 && y < ny) {
        const double nxa = (g.ox + (double)(x + bxs) * g.dx - p2x) * ix;
        const double nya = (g.oy + (double)(y + bys) * g.dy - p2y) * iy;
        const int cx = x, cy = y;
        double nap;
        if (nxa < nya - kEps) {
          nap = nxa;
          x += sx;
        } else if (nya < nxa - kEps) {
          nap = nya;
          y += sy;
        } else {
          nap = fmax(nxa, nya);
          x += sx;
          y += sy;
        }
        add(cx, cy, (float)(nap - ap));
        ap = nap;
      }
    }
  }
}

// The walk of par_walk, G cells at a time: emit(cid, l, n) receives the
// next n <= G cells (column index x * ny + y and xy length; n < G only for
// the last group) in the walk's order, so a caller can issue the group's
// gathers before it accumulates them.
template <int G, class Emit>
__device__ inline void par_walk_groups(const ParArgs &g, int a, int h, Emit &&emit) {
  const int nx = g.nx, ny = g.ny;
  const double c = gmem(g.atab)[5 * a], s = gmem(g.atab)[5 * a + 1];
  const double hp = gmem(g.h_pix)[h];
  const double p2x = c * g.D - s * hp;
  const double p2y = s * g.D + c * hp;
  int cid[G];
  float l[G];
  if (fabs(c) < kEps) {
    if (fabs(s) >= kEps) {  // line parallel to y (p2D.cpp:233-260)
      const int i = (int)floor((p2x - g.ox) / g.dx);
      if (i >= 0 && i < nx) {
        for (int j0 = 0; j0 < ny; j0 += G) {
#pragma unroll
          for (int q = 0; q < G; q++) {
            cid[q] = i * ny + (s < 0.0 ? ny - 1 - (j0 + q) : j0 + q);
            l[q] = (float)g.dy;
          }
          emit(cid, l, min(G, ny - j0));
        }
      }
    }
  } else if (fabs(s) < kEps) {  // line parallel to x (p2D.cpp:261-283)
    const int j = (int)floor((p2y - g.oy) / g.dy);
    if (j >= 0 && j < ny) {
      for (int i0 = 0; i0 < nx; i0 += G) {
#pragma unroll
        for (int q = 0; q < G; q++) {
          cid[q] = (c < 0.0 ? nx - 1 - (i0 + q) : i0 + q) * ny + j;
          l[q] = (float)g.dx;
        }
        emit(cid, l, min(G, nx - i0));
      }
    }
  } else {  // general line (p2D.cpp:284-458), steps as par_walk
    const double ix = 1.0 / c, iy = 1.0 / s;
    const double ax0 = (g.ox - p2x) * ix, ay0 = (g.oy - p2y) * iy;
    const double axn = (g.ox + (double)nx * g.dx - p2x) * ix;
    const double ayn = (g.oy + (double)ny * g.dy - p2y) * iy;
    const double amin = fmax(fmax(fmin(ax0, axn), fmin(ay0, ayn)), -g.dconv);
    const double amax = fmin(fmin(fmax(ax0, axn), fmax(ay0, ayn)), 0.0);
    if (amin < amax - kEps) {
      const int sx = c > 0.0 ? 1 : -1, sy = s > 0.0 ? 1 : -1;
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
          y = on_y ? yin : (int)floor((p2y + amin * s - g.oy) / g.dy);
        } else if (on_y) {
          y = yin;
          x = (int)floor((p2x + amin * c - g.ox) / g.dx);
        } else {
          x = 0;
          y = 0;
        }
        x = min(max(x, 0), nx - 1);
        y = min(max(y, 0), ny - 1);
      }
      const int bxs = sx > 0, bys = sy > 0;
      double ap = amin;
      while (x >= 0 && x < nx && y >= 0 && y < ny) {
        int n = 0;
#pragma unroll
        for (int q = 0; q < G; q++) {
          const bool in = x >= 0 && x < nx && y >= 0 && y < ny;
          cid[q] = in ? x * ny + y : 0;
          l[q] = 0.f;
          if (in) {
            const double nxa = (g.ox + (double)(x + bxs) * g.dx - p2x) * ix;
            const double nya = (g.oy + (double)(y + bys) * g.dy - p2y) * iy;
            double nap;
            if (nxa < nya - kEps) {
              nap = nxa;
              x += sx;
            } else if (nya < nxa - kEps) {
              nap = nya;
              y += sy;
            } else {
              nap = fmax(nxa, nya);
              x += sx;
              y += sy;
            }
            l[q] = (float)(nap - ap);
            ap = nap;
            n = q + 1;
          }
        }
        emit(cid, l, n);
      }
    }
  }
}

// The cells of ray r (angle a, column h), G at a time as par_walk_groups
// emits them: read from the plan's walk table when it has one (the same
// cells and lengths, written at plan time by par_walk_groups), else walked.
template <int G, class Emit>
__device__ inline void par_cells(const ParArgs &g, int r, int a, int h, Emit &&emit) {
  if (g.woff) {
    const unsigned long long e0 = gmem(g.woff)[r], e1 = gmem(g.woff)[r + 1];
    int cid[G];
    float l[G];
    for (unsigned long long e = e0; e < e1; e += G) {
      const int n = (int)min((unsigned long long)G, e1 - e);
#pragma unroll
      for (int q = 0; q < G; q++) {
        uint2 t = make_uint2(0u, 0u);
        if (q < n) t = load_u2(gmem(g.went) + e + q);
        cid[q] = (int)t.x;
        l[q] = __builtin_bit_cast(float, t.y);
      }
      emit(cid, l, n);
    }
  } else {
    par_walk_groups<G>(g, a, h, emit);
  }
}

// Cells per group of the FP's walk (or walk-table reads): in the
// one-row-per-slice FP the group's column reads are all in flight before
// the in-order accumulation
constexpr int kParFpGroup = 8;

// FP of detector rows [v0, v0 + RC) of ray r (clipped to nv): `acc_in`
// selects y += A x (reference f2D) or y = A x.  Returns the sum of squares
// of the rows written.  Acc: float row sums (the reference's), or double
// (the FP64-accumulation option; each float product added in double).
template <int RC, class Acc = float>
__device__ inline double par_fp_rows(const ParArgs &g, int r, int v0,
                                     const float *__restrict__ vol,
                                     float *__restrict__ sino, bool acc_in) {
  static_assert(RC % 4 == 0, "rows are read in groups of four");
  const int a = r / g.nh, h = r - a * g.nh;
  const int nz = g.nz;
  CCPI_GLOBAL float *__restrict__ out = gmem(sino) + (size_t)r * g.nv + v0;
  const int nrow = min(RC, g.nv - v0);
  Acc acc[RC];
  int kmap[RC];
#pragma unroll
  for (int q = 0; q < RC; q++) {
    acc[q] = (acc_in && q < nrow) ? (Acc)out[q] : (Acc)0;
    const int k = q < nrow ? gmem(g.map)[v0 + q] : -1;
    kmap[q] = (k >= 0 && k < nz) ? k : -1;
  }
  // one row per slice (map_type 1): the rows are the contiguous slices
  // v0 .. v0 + nval - 1 of each column, read 16 bytes at a time
  const int nval = g.map_type == 1 ? max(0, min(nrow, nz - v0)) : 0;
  if (g.map_type == 1) {
    // kParFpGroup cells at a time: all their column reads are issued before
    // the accumulation, which adds the cells in the walk's order (the same
    // float operations as one cell at a time)
    constexpr int Q4 = RC / 4;
    par_cells<kParFpGroup>(g, r, a, h, [&](const int *cid, const float *l, int n) {
      float4 f[kParFpGroup][Q4];
#pragma unroll
      for (int q = 0; q < kParFpGroup; q++) {
        const CCPI_GLOBAL float *__restrict__ cv = gmem(vol) + (size_t)cid[q] * nz + v0;
#pragma unroll
        for (int q4 = 0; q4 < Q4; q4++) {
          float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
          if (q < n) {
            if (4 * q4 + 3 < nval) {
              v = load_f4(cv + 4 * q4);
            } else {
              if (4 * q4 < nval) v.x = cv[4 * q4];
              if (4 * q4 + 1 < nval) v.y = cv[4 * q4 + 1];
              if (4 * q4 + 2 < nval) v.z = cv[4 * q4 + 2];
            }
          }
          f[q][q4] = v;
        }
      }
#pragma unroll
      for (int q = 0; q < kParFpGroup; q++) {
        if (q < n) {
#pragma unroll
          for (int q4 = 0; q4 < Q4; q4++) {
            const float4 v = f[q][q4];
            if (4 * q4 < nval) acc[4 * q4] = acc[4 * q4] + (Acc)(v.x * l[q]);
            if (4 * q4 + 1 < nval) acc[4 * q4 + 1] = acc[4 * q4 + 1] + (Acc)(v.y * l[q]);
            if (4 * q4 + 2 < nval) acc[4 * q4 + 2] = acc[4 * q4 + 2] + (Acc)(v.z * l[q]);
            if (4 * q4 + 3 < nval) acc[4 * q4 + 3] = acc[4 * q4 + 3] + (Acc)(v.w * l[q]);
          }
        }
      }
    });
  } else {
    par_cells<kParFpGroup>(g, r, a, h, [&](const int *cid, const float *l, int n) {
      for (int q = 0; q < n; q++) {
        const CCPI_GLOBAL float *__restrict__ col = gmem(vol) + (size_t)cid[q] * nz;
#pragma unroll
        for (int k = 0; k < RC; k++)
          if (kmap[k] >= 0) acc[k] = acc[k] + (Acc)(col[kmap[k]] * l[q]);
      }
    });
  }
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < RC; q++)
    if (q < nrow) {
      const float o = (float)acc[q];
      out[q] = o;
      ss += (double)o * (double)o;
    }
  return ss;
}

// One thread of the FP, one ray per lane: ray bx * 256 + tid, rows
// [by * RC, by * RC + RC).
template <int RC>
__device__ inline double par_fp_thread(const ParArgs &g, int bx, int by, int tid,
                                       const float *__restrict__ vol,
                                       float *__restrict__ sino, bool acc_in) {
  const int r = bx * 256 + tid;
  if (r >= g.na * g.nh) return 0.0;
  return par_fp_rows<RC>(g, r, by * RC, vol, sino, acc_in);
}

// FP, ray-shared layout: `lr` lanes (a power of two, at most 16) share one
// ray, each taking 4 consecutive detector rows, so the lanes of a ray read
// one contiguous run of every crossed column.  Block bx holds rays
// [bx * 256 / lr, ...), row block by the rows [by * 4 lr, (by + 1) * 4 lr).
template <class Acc = float>
__device__ inline double par_fp_lanes(const ParArgs &g, int bx, int by, int tid, int lr,
                                      const float *__restrict__ vol,
                                      float *__restrict__ sino, bool acc_in) {
  const int r = bx * (256 / lr) + tid / lr;
  if (r >= g.na * g.nh) return 0.0;
  const int v0 = (by * lr + (tid & (lr - 1))) * 4;
  if (v0 >= g.nv) return 0.0;
  return par_fp_rows<4, Acc>(g, r, v0, vol, sino, acc_in);
}

// Lanes per ray of the FP for nv detector rows: enough for the rows of one
// ray (4 per lane), at most 16.
inline int par_fp_lanes_per_ray(int nv) {
  int lr = 1;
  while (lr < 16 && 4 * lr < nv) lr *= 2;
  return lr;
}

// Footprint of the voxel column [x0, xn] x [y0, yn] on the detector at the
// angle whose table row is t (p2D.cpp:808-915): detector columns [lo, hi],
// kind 0 (general: trapezoid over [ymin, ymax]), 1 or 2 (axis-aligned:
// columns whose +-h lies in [lim0, lim1)).
struct ParFoot {
  int lo, hi, kind;
  double ymin, ymax, lim0, lim1;
};

template <class TabPtr>
__device__ inline ParFoot par_footprint(const ParArgs &g, TabPtr t, double x0, double xn,
                                        double y0, double yn) {
  const double c = t[0], s = t[1];
  ParFoot f;
  f.ymin = f.ymax = f.lim0 = f.lim1 = 0.0;
  if (fabs(c) < kEps) {  // p2D.cpp:808-844
    f.kind = s < 0.0 ? 1 : 2;
    f.lim0 = x0;
    f.lim1 = xn;
    if (s < 0.0) {
      f.lo = (int)floor(x0 * g.ihp - g.hp0);
      f.hi = (int)floor((xn - kEps) * g.ihp - g.hp0);
    } else {
      f.lo = (int)ceil((-xn + kEps) * g.ihp - g.hp0);
      f.hi = (int)ceil(-x0 * g.ihp - g.hp0);
    }
    f.lo = max(f.lo, 0);
    f.hi = min(f.hi, g.nh - 1);
  } else if (fabs(s) < kEps) {  // p2D.cpp:845-880
    f.kind = c < 0.0 ? 2 : 1;
    f.lim0 = y0;
    f.lim1 = yn;
    if (c < 0.0) {
      f.lo = (int)ceil((-yn + kEps) * g.ihp - g.hp0);
      f.hi = (int)ceil(-y0 * g.ihp - g.hp0);
    } else {
      f.lo = (int)floor(y0 * g.ihp - g.hp0);
      f.hi = (int)floor((yn - kEps) * g.ihp - g.hp0);
    }
    f.lo = max(f.lo, 0);
    f.hi = min(f.hi, g.nh - 1);
  } else {  // p2D.cpp:881-915
    f.kind = 0;
    if (c > 0.0) {
      if (s > 0.0) { f.ymin = y0 * c - xn * s; f.ymax = yn * c - x0 * s; }
      else         { f.ymin = y0 * c - x0 * s; f.ymax = yn * c - xn * s; }
    } else {
      if (s > 0.0) { f.ymin = yn * c - xn * s; f.ymax = y0 * c - x0 * s; }
      else         { f.ymin = yn * c - x0 * s; f.ymax = y0 * c - xn * s; }
    }
    // clamp at 0: the floor can round to -1 when ymin ~ h_pix[0]
    f.lo = f.ymin < gmem(g.h_pix)[0] ? 0 : max((int)floor(f.ymin * g.ihp - g.hp0), 0);
    f.hi = min((int)floor(f.ymax * g.ihp - g.hp0), g.nh - 1);
  }
  return f;
}

// Weight of detector column position hp in footprint f (p2D.cpp:916-938):
// 1 with w set, 0 to skip the column, -1 past the footprint's far edge.
template <class TabPtr>
__device__ inline int par_weight(const ParArgs &g, TabPtr t, const ParFoot &f, double hp,
                                 double &w) {
  if (f.kind != 0) {
    const double tt = f.kind == 1 ? hp : -hp;
    if (!(tt >= f.lim0 && tt < f.lim1)) return 0;
    w = fabs(t[0]) < kEps ? g.dy : g.dx;
    return 1;
  }
  const double ybot = f.ymin + t[2], ytop = f.ymax - t[2], l = t[4];
  if (!(hp > f.ymin + kEps)) return 0;
  if (hp < ybot)
    w = l * (hp - f.ymin) * t[3];
  else if (hp <= ytop)
    w = l;
  else if (hp < f.ymax - kEps)
    w = l * (f.ymax - hp) * t[3];
  else
    return -1;
  return 1;
}

// One thread of the BP: column bx * 256 + tid, slices [by * RC, by * RC +
// RC).  `acc_in` selects x += A^T y (reference b2D) or x = A^T y.  Returns
// the sum of squares of the slices written.
template <int RC, class Acc = float>
__device__ inline double par_bp_thread(const ParArgs &g, int bx, int by, int tid,
                                       const float *__restrict__ sino,
                                       float *__restrict__ vol, bool acc_in) {
  static_assert(RC % 4 == 0, "rows are read in groups of four");
  const int colid = bx * 256 + tid;
  if (colid >= g.nx * g.ny) return 0.0;
  const int i = colid / g.ny, j = colid - (colid / g.ny) * g.ny;
  const int k0 = by * RC;
  const int nz = g.nz, nv = g.nv, nh = g.nh;
  const int nk = min(RC, nz - k0);
  CCPI_GLOBAL float *__restrict__ vcol = gmem(vol) + (size_t)colid * nz + k0;
  Acc acc[RC];
#pragma unroll
  for (int q = 0; q < RC; q++) acc[q] = (acc_in && q < nk) ? (Acc)vcol[q] : (Acc)0;
  const double x0 = g.ox + (double)i * g.dx, xn = g.ox + (double)(i + 1) * g.dx;
  const double y0 = g.oy + (double)j * g.dy, yn = g.oy + (double)(j + 1) * g.dy;
  const int T = g.map_type;
  const int nvalid = max(0, min(nk, nv - k0));  // rows of the slices (T == 1)

  for (int a = 0; a < g.na; a++) {
    const CCPI_GLOBAL double *t = gmem(g.atab) + 5 * a;
    const double c = t[0], s = t[1];
    int lo, hi, kind;
    double ymin = 0, ymax = 0, lim0 = 0, lim1 = 0;
    if (fabs(c) < kEps) {  // p2D.cpp:808-844
      kind = s < 0.0 ? 1 : 2;
      lim0 = x0;
      lim1 = xn;
      if (s < 0.0) {
        lo = (int)floor(x0 * g.ihp - g.hp0);
        hi = (int)floor((xn - kEps) * g.ihp - g.hp0);
      } else {
        lo = (int)ceil((-xn + kEps) * g.ihp - g.hp0);
        hi = (int)ceil(-x0 * g.ihp - g.hp0);
      }
      lo = max(lo, 0);
      hi = min(hi, nh - 1);
    } else if (fabs(s) < kEps) {  // p2D.cpp:845-880
      kind = c < 0.0 ? 2 : 1;
      lim0 = y0;
      lim1 = yn;
      if (c < 0.0) {
        lo = (int)ceil((-yn + kEps) * g.ihp - g.hp0);
        hi = (int)ceil(-y0 * g.ihp - g.hp0);
      } else {
        lo = (int)floor(y0 * g.ihp - g.hp0);
        hi = (int)floor((yn - kEps) * g.ihp - g.hp0);
      }
      lo = max(lo, 0);
      hi = min(hi, nh - 1);
    } else {  // p2D.cpp:881-915
      kind = 0;
      if (c > 0.0) {
        if (s > 0.0) { ymin = y0 * c - xn * s; ymax = yn * c - x0 * s; }
        else         { ymin = y0 * c - x0 * s; ymax = yn * c - xn * s; }
      } else {
        if (s > 0.0) { ymin = yn * c - xn * s; ymax = y0 * c - x0 * s; }
        else         { ymin = yn * c - x0 * s; ymax = y0 * c - xn * s; }
      }
      // clamp at 0: the floor can round to -1 when ymin ~ h_pix[0]
      lo = ymin < gmem(g.h_pix)[0] ? 0 : max((int)floor(ymin * g.ihp - g.hp0), 0);
      hi = min((int)floor(ymax * g.ihp - g.hp0), nh - 1);
    }
    for (int h = lo; h <= hi; h++) {
      const double hp = gmem(g.h_pix)[h];
      double w;
      if (kind != 0) {
        const double tt = kind == 1 ? hp : -hp;
        if (!(tt >= lim0 && tt < lim1)) continue;
        w = fabs(c) < kEps ? g.dy : g.dx;
      } else {  // trapezoid (p2D.cpp:916-938)
        const double ybot = ymin + t[2], ytop = ymax - t[2], l = t[4];
        if (!(hp > ymin + kEps)) continue;
        if (hp < ybot)
          w = l * (hp - ymin) * t[3];
        else if (hp <= ytop)
          w = l;
        else if (hp < ymax - kEps)
          w = l * (ymax - hp) * t[3];
        else
          break;
      }
      const float wf = (float)w;
      const CCPI_GLOBAL float *__restrict__ p = gmem(sino) + ((size_t)a * nh + h) * nv;
      if (T == 1) {
        // rows k0 .. k0 + nvalid - 1 of the ray, read 16 bytes at a time
        const CCPI_GLOBAL float *__restrict__ pr = p + k0;
#pragma unroll
        for (int q4 = 0; q4 < RC / 4; q4++) {
          if (4 * q4 + 3 < nvalid) {
            const float4 f = load_f4(pr + 4 * q4);
            acc[4 * q4] = acc[4 * q4] + (Acc)(f.x * wf);
            acc[4 * q4 + 1] = acc[4 * q4 + 1] + (Acc)(f.y * wf);
            acc[4 * q4 + 2] = acc[4 * q4 + 2] + (Acc)(f.z * wf);
            acc[4 * q4 + 3] = acc[4 * q4 + 3] + (Acc)(f.w * wf);
          } else {
#pragma unroll
            for (int q = 4 * q4; q < 4 * q4 + 4; q++)
              if (q < nvalid) acc[q] = acc[q] + (Acc)(pr[q] * wf);
          }
        }
      } else if (T == 2 || T == 4) {  // rows summed before scaling (p2D.cpp:694-761)
#pragma unroll
        for (int q = 0; q < RC; q++) {
          const int k = k0 + q;
          if (q < nk && (k + 1) * T <= nv) {
            float sum = p[k * T];
            for (int u = 1; u < T; u++) sum = sum + p[k * T + u];
            acc[q] = acc[q] + (Acc)(sum * wf);
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < RC; q++) {
          if (q < nk) {
            const int k = k0 + q;
            const int vlo = gmem(g.rlo)[k], vhi = gmem(g.rhi)[k];
            for (int v = vlo; v < vhi; v++) acc[q] = acc[q] + (Acc)(p[v] * wf);
          }
        }
      }
    }
  }
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < RC; q++)
    if (q < nk) {
      const float o = (float)acc[q];
      vcol[q] = o;
      ss += (double)o * (double)o;
    }
  return ss;
}

// Tiled BP (one row per slice): a workgroup owns a 16 x 16 tile of voxel
// columns, one column per lane and R consecutive slices per lane.  Angles go
// in blocks of `ab`: for each angle of the block, the detector columns that
// the tile's square can reach (its corners' span plus a margin of 2) are
// staged in LDS with the R rows of the lane's slices (hmax columns at most),
// one coalesced copy per tile instead of each lane re-reading its columns'
// rows; every lane then walks its own column's footprint (the reference's
// FP64 arithmetic, par_footprint / par_weight) and reads the rows from LDS.
// A footprint column outside the staged span (never expected; rounding
// guard) is read from global memory.  Angles are visited in order, so with
// one angle group the result is bit-identical to par_bp_kernel; with G
// groups (blockIdx.z) each group writes its partial sums of its angle range
// to part[group] and par_bp_sum_kernel adds them in group order.
//
// One tile item of the tiled BP, run by all 256 threads of a workgroup
// (it synchronises the workgroup): tile `tile` (16 x 16 columns, row-major
// over the column grid), slices [zc * R, zc * R + R), angles [a_begin,
// a_end).  The result goes to `dst` (a volume-layout array), added to its
// contents when acc_in.  Returns the thread's sum of squares of the values
// it wrote (0 for threads outside the volume).  smem: par_bp_lds_bytes.
// FOLD (the batched kernel's tiled items; acc_in must be false): the
// angles are split into `groups` ranges as par_bp_lds_kernel splits them,
// each summed from zero, and the running sum of the groups is kept in
// `dst` (each lane re-reads only its own writes), so one item gives bit for
// bit what the grouped launch and par_bp_sum_kernel give into a zeroed
// volume, without registers for a second set of sums.
template <int R, class Acc, bool FOLD = false>
__device__ double par_bp_tile(const ParArgs &g, int tile, int zc, int a_begin, int a_end,
                              int hmax, int ab, bool acc_in, const float *__restrict__ sino,
                              float *__restrict__ dst_vol, float *smem, int groups = 1) {
  static_assert(R % 4 == 0, "rows are staged 4 at a time");
  float *s_rows = smem;                                 // [ab][hmax][R]
  int *s_lo = reinterpret_cast<int *>(smem + (size_t)ab * hmax * R);
  int *s_cnt = s_lo + ab;
  const int tid = threadIdx.x;
  const int ntj = (g.ny + 15) >> 4;
  const int ti0 = (tile / ntj) * 16, tj0 = (tile % ntj) * 16;
  const int i = ti0 + (tid >> 4), j = tj0 + (tid & 15);
  const bool active = i < g.nx && j < g.ny;
  const int k0 = zc * R;
  const int nz = g.nz, nv = g.nv, nh = g.nh;
  const int nk = min(R, nz - k0);
  const size_t colid = (size_t)(active ? i : 0) * g.ny + (active ? j : 0);
  Acc acc[R];
  CCPI_GLOBAL float *__restrict__ vcol = gmem(dst_vol) + colid * nz + k0;
#pragma unroll
  for (int q = 0; q < R; q++)
    acc[q] = (!FOLD && acc_in && active && q < nk) ? (Acc)vcol[q] : (Acc)0;
  const double x0 = g.ox + (double)i * g.dx, xn = g.ox + (double)(i + 1) * g.dx;
  const double y0 = g.oy + (double)j * g.dy, yn = g.oy + (double)(j + 1) * g.dy;
  // tile square
  const double X0 = g.ox + (double)ti0 * g.dx, XN = g.ox + (double)min(ti0 + 16, g.nx) * g.dx;
  const double Y0 = g.oy + (double)tj0 * g.dy, YN = g.oy + (double)min(tj0 + 16, g.ny) * g.dy;
  // rows of the lane's slices that exist (one row per slice)
  const int nrow = max(0, min(R, min(nv, nz) - k0));
  // rows of 16-byte aligned start (nv a multiple of 4) are staged 16 bytes
  // at a time, others one float at a time
  const bool aligned = (nv & 3) == 0;
  constexpr int Q4 = R / 4;
  const int ngrp = FOLD ? groups : 1;
  for (int grp = 0; grp < ngrp; grp++) {
    const int ga = FOLD ? a_begin + (int)((long long)(a_end - a_begin) * grp / groups) : a_begin;
    const int gb =
        FOLD ? a_begin + (int)((long long)(a_end - a_begin) * (grp + 1) / groups) : a_end;
    if (FOLD) {
#pragma unroll
      for (int q = 0; q < R; q++) acc[q] = (Acc)0;
    }
    for (int a0 = ga; a0 < gb; a0 += ab) {
      const int nblk = min(ab, gb - a0);
      if (tid < nblk) {
        const CCPI_GLOBAL double *t = gmem(g.atab) + 5 * (a0 + tid);
        const double c = t[0], s = t[1];
        const double p0 = Y0 * c - X0 * s, p1 = YN * c - X0 * s;
        const double p2 = Y0 * c - XN * s, p3 = YN * c - XN * s;
        const double tmin = fmin(fmin(p0, p1), fmin(p2, p3));
        const double tmax = fmax(fmax(p0, p1), fmax(p2, p3));
        const int lo = max((int)floor(tmin * g.ihp - g.hp0) - 2, 0);
        const int hi = min((int)floor(tmax * g.ihp - g.hp0) + 2, nh - 1);
        s_lo[tid] = lo;
        s_cnt[tid] = max(0, min(hi - lo + 1, hmax));
      }
      __syncthreads();
      const int total = nblk * hmax * Q4;
      for (int e = tid; e < total; e += 256) {
        const int q4 = e % Q4, r = e / Q4;
        const int hh = r % hmax, t = r / hmax;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (hh < s_cnt[t]) {
          const CCPI_GLOBAL float *p =
              gmem(sino) + ((size_t)(a0 + t) * nh + s_lo[t] + hh) * nv + k0 + 4 * q4;
          if (aligned && 4 * q4 + 3 < nrow) {
            v = load_f4(p);
          } else {
            if (4 * q4 < nrow) v.x = p[0];
            if (4 * q4 + 1 < nrow) v.y = p[1];
            if (4 * q4 + 2 < nrow) v.z = p[2];
            if (4 * q4 + 3 < nrow) v.w = p[3];
          }
        }
        *reinterpret_cast<float4 *>(s_rows + ((size_t)t * hmax + hh) * R + 4 * q4) = v;
      }
      __syncthreads();
      if (active) {
        for (int t = 0; t < nblk; t++) {
          const int a = a0 + t;
          const CCPI_GLOBAL double *tab = gmem(g.atab) + 5 * a;
          const ParFoot f = par_footprint(g, tab, x0, xn, y0, yn);
          const int slo = s_lo[t], scnt = s_cnt[t];
          for (int h = f.lo; h <= f.hi; h++) {
            double w;
            const int use = par_weight(g, tab, f, gmem(g.h_pix)[h], w);
            if (use == 0) continue;
            if (use < 0) break;
            const float wf = (float)w;
            const int hh = h - slo;
            if (hh >= 0 && hh < scnt) {
              const float *pr = s_rows + ((size_t)t * hmax + hh) * R;
#pragma unroll
              for (int q4 = 0; q4 < Q4; q4++) {
                const float4 v = *reinterpret_cast<const float4 *>(pr + 4 * q4);
                acc[4 * q4] = acc[4 * q4] + (Acc)(v.x * wf);
                acc[4 * q4 + 1] = acc[4 * q4 + 1] + (Acc)(v.y * wf);
                acc[4 * q4 + 2] = acc[4 * q4 + 2] + (Acc)(v.z * wf);
                acc[4 * q4 + 3] = acc[4 * q4 + 3] + (Acc)(v.w * wf);
              }
            } else {
              const CCPI_GLOBAL float *pr = gmem(sino) + ((size_t)a * nh + h) * nv + k0;
#pragma unroll
              for (int q = 0; q < R; q++)
                if (q < nrow) acc[q] = acc[q] + (Acc)(pr[q] * wf);
            }
          }
        }
      }
      __syncthreads();
    }
    if (FOLD && active && grp + 1 < ngrp) {
      // running sum of the groups so far (the last group's is added below)
#pragma unroll
      for (int q = 0; q < R; q++)
        if (q < nk) vcol[q] = grp == 0 ? (float)acc[q] : (float)((Acc)vcol[q] + acc[q]);
    }
  }
  if (!active) return 0.0;
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < R; q++)
    if (q < nk) {
      const float o =
          (FOLD && groups > 1) ? (float)((Acc)vcol[q] + acc[q]) : (float)acc[q];
      vcol[q] = o;
      ss += (double)o * (double)o;
    }
  return ss;
}

}  // namespace ccpi
