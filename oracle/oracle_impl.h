/*
 * CPU oracle for the CCPi-Reconstruction projector + solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * It restates, in C, the algorithm of the reference's dispatched projectors
 * and iteration bodies:
 *   parallel beam  f2D / b2D     Core/src/p2D.cpp:22-1087
 *   cone beam      f2D / b2D     Core/src/c2D.cpp:17-1119
 *   vector ops                   Core/src/blas.hpp:11-466
 *   CGLS (cgls_3d)               Core/src/Algorithms/cgls.cpp:26-157
 *   SIRT                         Core/src/Algorithms/sirt.cpp:15-80
 *   MLEM                         Core/src/Algorithms/mlem.cpp:15-80
 *   Tikhonov / TV terms          Core/src/tikhonov.cpp:4-32,
 *                                Core/src/total_v.cpp:21-195
 *   regularised CGLS             Core/src/Algorithms/cgls.cpp:343-435
 * The float instantiation (RT = float) keeps the reference's precision
 * choices: geometry in double, weights stored as float, float accumulators,
 * float CGLS alpha.  The double instantiation (RT = double) is the FP64
 * oracle of the same discrete model.
 *
 * Defined where the reference is undefined (each is also a documented
 * behaviour of the HIP path):
 *   - cone f2D: a ray whose xy entry point already lies below/above the
 *     volume makes the reference index alpha_inv[-1] (c2D.cpp:123-126 with
 *     ncom == 0).  Here ncom is raised to 1, which clips such rays at the
 *     top/bottom faces like the reference's own edge loops do.
 *   - b2D footprint lists are unbounded (the reference's fixed-size lists
 *     p2D.cpp:795, c2D.cpp:811 can overflow).
 *   - walk entry cells are clamped into the grid (a ray entering exactly at a
 *     grid corner can floor to index -1 in p2D.cpp:320 / c2D.cpp:323).
 *   - parallel detector rows whose mapped slice lies outside [0, nz) are
 *     skipped (p2D.cpp:209 would read outside the column).
 *
 * Parity pins (tests/test_oracle_pins.py): brute-force exact line/voxel
 * intersection lengths, axis-aligned closed forms, chord sums, the d_conv
 * closed form, the transpose/adjoint identity, and textbook CGLS/SIRT/MLEM
 * on the dense matrix the projectors define.
 *
 * This header is included twice by ccpi_oracle.c with
 *   RT      element type (float / double)
 *   FN(x)   symbol name decoration
 */

#ifndef RT
#error "RT must be defined"
#endif

/* FLT_EPSILON is the model tolerance in both instantiations
   (p2D.cpp:20, c2D.cpp:15). */
#define OR_EPS ((double)FLT_EPSILON)

/* ------------------------------------------------------------------ */
/* parallel beam                                                       */
/* ------------------------------------------------------------------ */

/* Row -> slice map and its ppv pattern (p2D.cpp:22-61). */
static int FN(pb_mapping)(int *map, const double *v_pix, int nv,
                          double z0, double dz)
{
  int v, type;
  for (v = 0; v < nv; v++)
    map[v] = (int)floor((v_pix[v] - z0) / dz);
  for (type = 1; type <= 4; type *= 2) {
    int ok = 1;
    for (v = 0; v < nv; v++)
      if (map[v] != v / type) { ok = 0; break; }
    if (ok)
      return type;
  }
  return 0;
}

/* One ray of the parallel walk restricted to a tile (p2D.cpp:215-466).
   The ray passes through detector point p2 with unit direction (c, s);
   the tile starts at (bx, by) and holds tnx x tny cells.  Emits cell
   offsets (relative to the tile's first column) and xy lengths. */
static int FN(pb_walk)(double p2x, double p2y, double bx, double by,
                       double dx, double dy, int tnx, int tny, long nyz,
                       long nz, double dconv, double c, double s,
                       long *off, RT *len)
{
  int n = 0;
  if (fabs(c) < OR_EPS) {
    if (fabs(s) >= OR_EPS) {
      int i = (int)floor((p2x - bx) / dx);
      if (i >= 0 && i < tnx) {
        int j, step = s < 0.0 ? -1 : 1, j0 = s < 0.0 ? tny - 1 : 0;
        for (j = 0; j < tny; j++) {
          off[n] = (long)i * nyz + (long)(j0 + step * j) * nz;
          len[n++] = (RT)dy;
        }
      }
    }
    return n;
  }
  if (fabs(s) < OR_EPS) {
    int j = (int)floor((p2y - by) / dy);
    if (j >= 0 && j < tny) {
      int i, step = c < 0.0 ? -1 : 1, i0 = c < 0.0 ? tnx - 1 : 0;
      for (i = 0; i < tnx; i++) {
        off[n] = (long)(i0 + step * i) * nyz + (long)j * nz;
        len[n++] = (RT)dx;
      }
    }
    return n;
  }
  {
    const double ix = 1.0 / c, iy = 1.0 / s;
    const double ax0 = (bx - p2x) * ix, ay0 = (by - p2y) * iy;
    const double axn = (bx + (double)tnx * dx - p2x) * ix;
    const double ayn = (by + (double)tny * dy - p2y) * iy;
    const double amin = fmax(fmax(fmin(ax0, axn), fmin(ay0, ayn)), -dconv);
    const double amax = fmin(fmin(fmax(ax0, axn), fmax(ay0, ayn)), 0.0);
    const int sx = c > 0.0 ? 1 : -1, sy = s > 0.0 ? 1 : -1;
    int x, y;
    double ap;
    if (!(amin < amax - OR_EPS))
      return 0;
    /* entry cell: the face the ray enters through fixes one index
       (p2D.cpp:316-431, same face priority per direction) */
    {
      const double xface = sx > 0 ? ax0 : axn, yface = sy > 0 ? ay0 : ayn;
      const int xin = sx > 0 ? 0 : tnx - 1, yin = sy > 0 ? 0 : tny - 1;
      int on_x = amin == xface, on_y = amin == yface;
      if (sx > 0 && sy < 0) {          /* y face checked first */
        if (on_y) on_x = 0;
      } else if (!(sx < 0 && sy < 0)) {
        if (on_x) on_y = 0;
      }
      if (on_x) {
        x = xin;
        y = (on_y) ? yin : (int)floor((p2y + amin * s - by) / dy);
      } else if (on_y) {
        y = yin;
        x = (int)floor((p2x + amin * c - bx) / dx);
      } else {
        x = 0; y = 0;   /* reference reports an error and walks from 0,0 */
      }
      if (x < 0) x = 0;
      if (x >= tnx) x = tnx - 1;
      if (y < 0) y = 0;
      if (y >= tny) y = tny - 1;
    }
    ap = amin;
    while (x >= 0 && x < tnx && y >= 0 && y < tny) {
      const double nxa = (bx + (double)(x + (sx > 0)) * dx - p2x) * ix;
      const double nya = (by + (double)(y + (sy > 0)) * dy - p2y) * iy;
      off[n] = (long)x * nyz + (long)y * nz;
      if (nxa < nya - OR_EPS) {
        len[n] = (RT)(nxa - ap); ap = nxa; x += sx;
      } else if (nya < nxa - OR_EPS) {
        len[n] = (RT)(nya - ap); ap = nya; y += sy;
      } else {
        const double m = fmax(nxa, nya);
        len[n] = (RT)(m - ap); ap = m; x += sx; y += sy;
      }
      n++;
    }
  }
  return n;
}

/* y[a][h][v] += sum over the walk of l * x[cell][map(v)]
   (p2D.cpp:63-213 general mapping form, p2D.cpp:468-611 driver with
   32 x 32 xy tiles and the per-tile detector h range). */
void FN(or_pb_f2D)(const double *h_pix, int nh, const double *v_pix, int nv,
                   const double *ang, int na, const double origin[3],
                   const double vsize[3], int nx, int ny, int nz,
                   const RT *vox, RT *pix)
{
  const double D = 2.0 * fmax(fabs(origin[0]), fmax(fabs(origin[1]),
                                                    fabs(origin[2])));
  const RT dconv = (RT)(3.0 * D);
  const long nyz = (long)ny * nz;
  const double ihp = 1.0 / (h_pix[1] - h_pix[0]);
  const double hp0 = h_pix[0] / (h_pix[1] - h_pix[0]);
  int *map = (int *)malloc(sizeof(int) * (size_t)nv);
  int bxi, byi;
  FN(pb_mapping)(map, v_pix, nv, origin[2], vsize[2]);
  for (bxi = 0; bxi < nx; bxi += 32) {
    const int tnx = nx - bxi < 32 ? nx - bxi : 32;
    const double vx = origin[0] + (double)bxi * vsize[0];
    const double wx = origin[0] + (double)(bxi + tnx) * vsize[0];
    for (byi = 0; byi < ny; byi += 32) {
      const int tny = ny - byi < 32 ? ny - byi : 32;
      const double vy = origin[1] + (double)byi * vsize[1];
      const double wy = origin[1] + (double)(byi + tny) * vsize[1];
      const long base = (long)bxi * nyz + (long)byi * nz;
      int a;
#pragma omp parallel for schedule(dynamic)
      for (a = 0; a < na; a++) {
        const double c = cos(ang[a]), s = sin(ang[a]);
        int hmin, hmax, h;
        long *off = (long *)malloc(sizeof(long) * (size_t)(2 * (tnx + tny) + 2));
        RT *len = (RT *)malloc(sizeof(RT) * (size_t)(2 * (tnx + tny) + 2));
        /* detector columns that can meet the tile (p2D.cpp:527-596) */
        if (fabs(c) < OR_EPS) {
          if (s < 0.0) {
            hmin = (int)floor(vx * ihp - hp0);
            hmax = (int)floor((wx - OR_EPS) * ihp - hp0);
          } else {
            hmin = (int)ceil((-wx + OR_EPS) * ihp - hp0);
            hmax = (int)ceil(-vx * ihp - hp0);
          }
          if (hmin < 0) hmin = 0;
          if (hmax >= nh) hmax = nh - 1;
        } else if (fabs(s) < OR_EPS) {
          if (c < 0.0) {
            hmin = (int)ceil((-wy + OR_EPS) * ihp - hp0);
            hmax = (int)ceil(-vy * ihp - hp0);
          } else {
            hmin = (int)floor(vy * ihp - hp0);
            hmax = (int)floor((wy - OR_EPS) * ihp - hp0);
          }
          if (hmin < 0) hmin = 0;
          if (hmax >= nh) hmax = nh - 1;
        } else {
          double ymin, ymax;
          if (c > 0.0) {
            if (s > 0.0) { ymin = vy * c - wx * s; ymax = wy * c - vx * s; }
            else         { ymin = vy * c - vx * s; ymax = wy * c - wx * s; }
          } else {
            if (s > 0.0) { ymin = wy * c - wx * s; ymax = vy * c - vx * s; }
            else         { ymin = wy * c - vx * s; ymax = vy * c - wx * s; }
          }
          hmin = ymin < h_pix[0] ? 0 : (int)floor(ymin * ihp - hp0);
          if (hmin < 0) hmin = 0;   /* the floor can round to -1 */
          hmax = (int)floor(ymax * ihp - hp0);
          if (hmax > nh - 1) hmax = nh - 1;
        }
        for (h = hmin; h <= hmax; h++) {
          const double p2x = c * D - s * h_pix[h];
          const double p2y = s * D + c * h_pix[h];
          const int n = FN(pb_walk)(p2x, p2y, vx, vy, vsize[0], vsize[1],
                                    tnx, tny, nyz, nz, (double)dconv, c, s,
                                    off, len);
          RT *row = pix + ((long)a * nh + h) * nv;
          int m, v;
          for (m = 0; m < n; m++) {
            const RT *col = vox + base + off[m];
            const RT l = len[m];
            for (v = 0; v < nv; v++)
              if (map[v] >= 0 && map[v] < nz)
                row[v] += col[map[v]] * l;
          }
        }
        free(off);
        free(len);
      }
    }
  }
  free(map);
}

/* x[i][j][map(v)] += y[a][h][v] * w, w the exact chord of the detector line
   through the column square: a trapezoid in h (p2D.cpp:775-1087). */
void FN(or_pb_b2D)(const double *h_pix, int nh, const double *v_pix, int nv,
                   const double *ang, int na, const double origin[3],
                   const double vsize[3], int nx, int ny, int nz,
                   const RT *pix, RT *vox)
{
  double *cs = (double *)malloc(sizeof(double) * (size_t)na * 5);
  int *map = (int *)malloc(sizeof(int) * (size_t)nv);
  const int type = FN(pb_mapping)(map, v_pix, nv, origin[2], vsize[2]);
  const double ihp = 1.0 / (h_pix[1] - h_pix[0]);
  const double hp0 = h_pix[0] / (h_pix[1] - h_pix[0]);
  const long nhv = (long)nh * nv;
  int a, i;
  /* per-angle trapezoid offset and flat-top length (p2D.cpp:996-1049) */
  for (a = 0; a < na; a++) {
    const double c = cos(ang[a]), s = sin(ang[a]);
    const double y00 = origin[1] * c - origin[0] * s;
    const double y01 = (origin[1] + vsize[1]) * c - origin[0] * s;
    const double y10 = origin[1] * c - (origin[0] + vsize[0]) * s;
    const double y11 = (origin[1] + vsize[1]) * c - (origin[0] + vsize[0]) * s;
    double ymin, ybot;
    if (c > 0.0) {
      if (s > 0.0) { ymin = y10; ybot = y11 < y00 ? y11 : y00; }
      else         { ymin = y00; ybot = y01 < y10 ? y01 : y10; }
    } else {
      if (s > 0.0) { ymin = y11; ybot = y01 < y10 ? y01 : y10; }
      else         { ymin = y01; ybot = y11 < y00 ? y11 : y00; }
    }
    cs[5 * a + 0] = c;
    cs[5 * a + 1] = s;
    cs[5 * a + 2] = ybot - ymin;
    cs[5 * a + 3] = 1.0 / (ybot - ymin);
    cs[5 * a + 4] = fabs(c) > fabs(s) ? vsize[0] / fabs(c) : vsize[1] / fabs(s);
  }
#pragma omp parallel for schedule(dynamic)
  for (i = 0; i < nx; i++) {
    const double x0 = origin[0] + (double)i * vsize[0];
    const double xn = origin[0] + (double)(i + 1) * vsize[0];
    int j;
    for (j = 0; j < ny; j++) {
      const double y0 = origin[1] + (double)j * vsize[1];
      const double yn = origin[1] + (double)(j + 1) * vsize[1];
      RT *col = vox + ((long)i * ny + j) * nz;
      int a;
      for (a = 0; a < na; a++) {
        const double c = cs[5 * a], s = cs[5 * a + 1];
        const RT *arow = pix + (long)a * nhv;
        int h, lo, hi, kind;
        double ymin = 0.0, ymax = 0.0, lim0 = 0.0, lim1 = 0.0;
        /* footprint of the column on the detector line (p2D.cpp:808-915) */
        if (fabs(c) < OR_EPS) {
          kind = s < 0.0 ? 1 : 2;
          lim0 = x0; lim1 = xn;
          if (s < 0.0) {
            lo = (int)floor(x0 * ihp - hp0);
            hi = (int)floor((xn - OR_EPS) * ihp - hp0);
          } else {
            lo = (int)ceil((-xn + OR_EPS) * ihp - hp0);
            hi = (int)ceil(-x0 * ihp - hp0);
          }
          if (lo < 0) lo = 0;
          if (hi >= nh) hi = nh - 1;
        } else if (fabs(s) < OR_EPS) {
          kind = c < 0.0 ? 2 : 1;
          lim0 = y0; lim1 = yn;
          if (c < 0.0) {
            lo = (int)ceil((-yn + OR_EPS) * ihp - hp0);
            hi = (int)ceil(-y0 * ihp - hp0);
          } else {
            lo = (int)floor(y0 * ihp - hp0);
            hi = (int)floor((yn - OR_EPS) * ihp - hp0);
          }
          if (lo < 0) lo = 0;
          if (hi >= nh) hi = nh - 1;
        } else {
          kind = 0;
          if (c > 0.0) {
            if (s > 0.0) { ymin = y0 * c - xn * s; ymax = yn * c - x0 * s; }
            else         { ymin = y0 * c - x0 * s; ymax = yn * c - xn * s; }
          } else {
            if (s > 0.0) { ymin = yn * c - xn * s; ymax = y0 * c - x0 * s; }
            else         { ymin = yn * c - x0 * s; ymax = y0 * c - xn * s; }
          }
          lo = ymin < h_pix[0] ? 0 : (int)floor(ymin * ihp - hp0);
          if (lo < 0) lo = 0;   /* the floor can round to -1 */
          hi = (int)floor(ymax * ihp - hp0);
          if (hi > nh - 1) hi = nh - 1;
        }
        for (h = lo; h <= hi; h++) {
          double w = 0.0;
          if (kind != 0) {
            /* axis-aligned: full voxel width for the covered columns */
            const double t = kind == 1 ? h_pix[h] : -h_pix[h];
            if (!(t >= lim0 && t < lim1))
              continue;
            w = fabs(c) < OR_EPS ? vsize[1] : vsize[0];
          } else {
            /* rising edge, flat top, falling edge (p2D.cpp:916-938) */
            const double ybot = ymin + cs[5 * a + 2];
            const double ytop = ymax - cs[5 * a + 2];
            const double l = cs[5 * a + 4];
            if (!(h_pix[h] > ymin + OR_EPS))
              continue;
            if (h_pix[h] < ybot)
              w = l * (h_pix[h] - ymin) * cs[5 * a + 3];
            else if (h_pix[h] <= ytop)
              w = l;
            else if (h_pix[h] < ymax - OR_EPS)
              w = l * (ymax - h_pix[h]) * cs[5 * a + 3];
            else
              break;
          }
          {
            /* p2D.cpp:613-773: ppv 2 and 4 sum their rows before scaling */
            const RT wf = (RT)w;
            const RT *p = arow + (long)h * nv;
            int v, k;
            if (type == 2 || type == 4) {
              for (k = 0; k < nz && (k + 1) * type <= nv; k++) {
                RT acc = p[k * type];
                int q;
                for (q = 1; q < type; q++)
                  acc += p[k * type + q];
                col[k] += acc * wf;
              }
            } else {
              for (v = 0; v < nv; v++)
                if (map[v] >= 0 && map[v] < nz)
                  col[map[v]] += p[v] * wf;
            }
          }
        }
      }
    }
  }
  free(map);
  free(cs);
}

/* ------------------------------------------------------------------ */
/* cone beam                                                           */
/* ------------------------------------------------------------------ */

/* Per-row and per-slice tables shared by cone f2D and b2D
   (c2D.cpp:480-516, 993-1010). */
typedef struct FN(cb_tab) {
  int mid;          /* first detector row above the source */
  RT inv_dz, pzbz;
  RT *delta_z;      /* v_pix[v] - source_z */
  RT *inv_delz;     /* 1 / delta_z[v] */
  RT *vox_z;        /* slice planes relative to source_z, nz + 1 */
} FN(cb_tab);

static void FN(cb_tab_make)(FN(cb_tab) *t, double sz, const double *v_pix,
                            int nv, const double origin[3],
                            const double vsize[3], int nz)
{
  int v, k;
  t->mid = nv;   /* reference leaves -1 (undefined use); nv = all rows below */
  for (v = 0; v < nv; v++)
    if (v_pix[v] > sz) { t->mid = v; break; }
  t->inv_dz = (RT)(1.0 / vsize[2]);
  t->pzbz = (RT)((sz - origin[2]) / vsize[2]);
  t->delta_z = (RT *)malloc(sizeof(RT) * (size_t)nv);
  t->inv_delz = (RT *)malloc(sizeof(RT) * (size_t)nv);
  t->vox_z = (RT *)malloc(sizeof(RT) * (size_t)(nz + 1));
  for (v = 0; v < nv; v++) {
    t->delta_z[v] = (RT)(v_pix[v] - sz);
    t->inv_delz[v] = (RT)(1.0 / (double)t->delta_z[v]);
  }
  for (k = 0; k <= nz; k++)
    t->vox_z[k] = (RT)(origin[2] + (double)k * vsize[2] - sz);
}

static void FN(cb_tab_free)(FN(cb_tab) *t)
{
  free(t->delta_z);
  free(t->inv_delz);
  free(t->vox_z);
}

/* Source -> pixel distance per detector pixel (c2D.cpp:495-503). */
static RT *FN(cb_dconv)(double sx, double sy, double sz, double det_x,
                        const double *h_pix, int nh, const double *v_pix,
                        int nv)
{
  RT *d = (RT *)malloc(sizeof(RT) * (size_t)nh * nv);
  const double x2 = (det_x - sx) * (det_x - sx);
  int h, v;
  for (h = 0; h < nh; h++) {
    const double xy2 = x2 + (h_pix[h] - sy) * (h_pix[h] - sy);
    for (v = 0; v < nv; v++)
      d[(long)h * nv + v] = (RT)sqrt(xy2 + (v_pix[v] - sz) * (v_pix[v] - sz));
  }
  return d;
}

/* xy walk of one source->pixel line in cell units (c2D.cpp:193-467).
   alpha[0] is the entry parameter, alpha[m] (m >= 1) the exit parameter of
   cell off[m]; alpha runs 0 (source) .. 1 (pixel). */
static int FN(cb_walk)(double q1x, double q1y, double q2x, double q2y,
                       int nx, int ny, long nyz, long nz, long *off, RT *alpha)
{
  const double ddx = q2x - q1x, ddy = q2y - q1y;
  const double ix = 1.0 / ddx, iy = 1.0 / ddy;
  int n = 0;
  if (fabs(ddx) < OR_EPS) {
    if (fabs(ddy) >= OR_EPS) {
      const int i = (int)floor(q1x);
      if (i >= 0 && i < nx) {
        int j;
        if (ddy < 0.0) {
          alpha[n] = (RT)(((double)ny - q1y) * iy);
          off[n++] = (long)i * nyz + (long)(ny - 1) * nz;
          for (j = ny - 1; j >= 0; j--) {
            alpha[n] = (RT)(((double)j - q1y) * iy);
            off[n++] = (long)i * nyz + (long)j * nz;
          }
        } else {
          alpha[n] = (RT)((0.0 - q1y) * iy);
          off[n++] = (long)i * nyz;
          for (j = 0; j < ny; j++) {
            alpha[n] = (RT)(((double)(j + 1) - q1y) * iy);
            off[n++] = (long)i * nyz + (long)j * nz;
          }
        }
      }
    }
    return n;
  }
  if (fabs(ddy) < OR_EPS) {
    const int j = (int)floor(q1y);
    if (j >= 0 && j < ny) {
      int i;
      if (ddx < 0.0) {
        alpha[n] = (RT)(((double)nx - q1x) * ix);
        off[n++] = (long)(nx - 1) * nyz + (long)j * nz;
        for (i = nx - 1; i >= 0; i--) {
          alpha[n] = (RT)(((double)i - q1x) * ix);
          off[n++] = (long)i * nyz + (long)j * nz;
        }
      } else {
        alpha[n] = (RT)((0.0 - q1x) * ix);
        off[n++] = (long)j * nz;
        for (i = 0; i < nx; i++) {
          alpha[n] = (RT)(((double)(i + 1) - q1x) * ix);
          off[n++] = (long)i * nyz + (long)j * nz;
        }
      }
    }
    return n;
  }
  {
    const double ax0 = (0.0 - q1x) * ix, ay0 = (0.0 - q1y) * iy;
    const double axn = ((double)nx - q1x) * ix, ayn = ((double)ny - q1y) * iy;
    const double amin = fmax(fmax(fmin(ax0, axn), fmin(ay0, ayn)), 0.0);
    const double amax = fmin(fmin(fmax(ax0, axn), fmax(ay0, ayn)), 1.0);
    const int sx = ddx > 0.0 ? 1 : -1, sy = ddy > 0.0 ? 1 : -1;
    int x, y;
    if (!(amin < amax - OR_EPS))
      return 0;
    {
      const double xface = sx > 0 ? ax0 : axn, yface = sy > 0 ? ay0 : ayn;
      const int xin = sx > 0 ? 0 : nx - 1, yin = sy > 0 ? 0 : ny - 1;
      int on_x = amin == xface, on_y = amin == yface;
      if (sx > 0 && sy < 0) {
        if (on_y) on_x = 0;
      } else if (!(sx < 0 && sy < 0)) {
        if (on_x) on_y = 0;
      }
      if (on_x) {
        x = xin;
        y = on_y ? yin : (int)floor(q1y + amin * ddy);
      } else if (on_y) {
        y = yin;
        x = (int)floor(q1x + amin * ddx);
      } else {
        x = 0; y = 0;
      }
      if (x < 0) x = 0;
      if (x >= nx) x = nx - 1;
      if (y < 0) y = 0;
      if (y >= ny) y = ny - 1;
    }
    alpha[n] = (RT)amin;
    off[n++] = (long)x * nyz + (long)y * nz;
    while (x >= 0 && x < nx && y >= 0 && y < ny) {
      const double nxa = ((double)(x + (sx > 0)) - q1x) * ix;
      const double nya = ((double)(y + (sy > 0)) - q1y) * iy;
      off[n] = (long)x * nyz + (long)y * nz;
      if (nxa < nya - OR_EPS) {
        alpha[n] = (RT)nxa; x += sx;
      } else if (nya < nxa - OR_EPS) {
        alpha[n] = (RT)nya; y += sy;
      } else {
        alpha[n] = (RT)fmax(nxa, nya); x += sx; y += sy;
      }
      n++;
    }
  }
  return n;
}

/* Two-voxel z split of every walk segment for every detector row
   (c2D.cpp:17-191).  Rows below the source step down in z, rows above step
   up; a segment crosses at most one slice plane. */
static void FN(cb_zsplit)(RT *pix, const RT *vox, const long *off,
                          const RT *alpha, int n, const FN(cb_tab) *t,
                          int nv, int nz, RT *ainv, int *kv)
{
  const int nzm1 = nz - 1, midp = t->mid;
  const RT *dz = t->delta_z, *iz = t->inv_delz, *vz = t->vox_z;
  const RT pzbz = t->pzbz, pzbz1 = t->pzbz + (RT)1.0;
  int m, v, ncom, min_xy = n, max_xy = n, min_v = midp, max_v = midp - 1;
  RT a0;
  for (m = 0; m < n; m++)
    ainv[m] = alpha[m] * t->inv_dz;
  /* first segment where the lowest / highest row reaches the bottom / top
     slice (c2D.cpp:49-78) */
  for (m = 1; m < n; m++) {
    const int k = (int)floor(pzbz + ainv[m - 1] * dz[0]);
    if (k <= 0) { min_xy = k == 0 ? m : 0; break; }
  }
  for (m = 1; m < n; m++) {
    const int k = (int)floor(pzbz + ainv[m - 1] * dz[nv - 1]);
    if (k >= nzm1) { max_xy = k == nzm1 ? m : 0; break; }
  }
  /* rows that stay inside through the start of the last segment
     (c2D.cpp:80-95) */
  for (v = 0; v < midp; v++)
    if ((int)floor(pzbz + ainv[n - 2] * dz[v]) > 0) { min_v = v; break; }
  for (v = nv - 1; v >= midp; v--)
    if ((int)floor(pzbz + ainv[n - 2] * dz[v]) < nzm1) { max_v = v; break; }
  ncom = min_xy < max_xy ? min_xy : max_xy;
  if (ncom < 1)
    ncom = 1;   /* reference: alpha_inv[-1]; see header */
  /* segments before ncom: every row inside (c2D.cpp:100-122) */
  a0 = alpha[0];
  for (m = 1; m < ncom; m++) {
    const RT *col = vox + off[m];
    const RT a1 = alpha[m];
    for (v = 0; v < nv; v++)
      kv[v] = (int)(pzbz + ainv[m - 1] * dz[v]);
    for (v = 0; v < midp; v++) {
      const int k = kv[v];
      const RT az = vz[k] * iz[v];
      const RT mz = az < a1 ? az : a1;
      pix[v] += (col[k] * (mz - a0) + col[k - 1] * (a1 - mz));
    }
    for (v = midp; v < nv; v++) {
      const int k = kv[v];
      const RT az = vz[k + 1] * iz[v];
      const RT mz = az < a1 ? az : a1;
      pix[v] += (col[k] * (mz - a0) + col[k + 1] * (a1 - mz));
    }
    a0 = a1;
  }
  /* remaining segments, rows that stay inside (c2D.cpp:123-145) */
  for (m = ncom; m < n; m++) {
    const RT *col = vox + off[m];
    const RT a1 = alpha[m];
    for (v = min_v; v <= max_v; v++)
      kv[v] = (int)(pzbz + ainv[m - 1] * dz[v]);
    for (v = min_v; v < midp; v++) {
      const int k = kv[v];
      const RT az = vz[k] * iz[v];
      const RT mz = az < a1 ? az : a1;
      pix[v] += (col[k] * (mz - a0) + col[k - 1] * (a1 - mz));
    }
    for (v = midp; v <= max_v; v++) {
      const int k = kv[v];
      const RT az = vz[k + 1] * iz[v];
      const RT mz = az < a1 ? az : a1;
      pix[v] += (col[k] * (mz - a0) + col[k + 1] * (a1 - mz));
    }
    a0 = a1;
  }
  /* rows leaving through the bottom / top face, clipped there
     (c2D.cpp:146-190) */
  a0 = alpha[ncom - 1];
  for (m = ncom; m < n; m++) {
    const RT *col = vox + off[m];
    const RT ai = ainv[m - 1], a1 = alpha[m];
    for (v = min_v - 1; v >= 0; v--) {
      const int k = (int)(pzbz1 + ai * dz[v]) - 1;
      if (k > 0) {
        const RT az = vz[k] * iz[v];
        const RT mz = az < a1 ? az : a1;
        pix[v] += (col[k] * (mz - a0) + col[k - 1] * (a1 - mz));
      } else if (k == 0) {
        const RT az = vz[0] * iz[v];
        const RT mz = az < a1 ? az : a1;
        pix[v] += col[0] * (mz - a0);
      } else
        break;
    }
    a0 = a1;
  }
  a0 = alpha[ncom - 1];
  for (m = ncom; m < n; m++) {
    const RT *col = vox + off[m];
    const RT ai = ainv[m - 1], a1 = alpha[m];
    for (v = max_v + 1; v < nv; v++) {
      const int k = (int)(pzbz + ai * dz[v]);
      if (k < nzm1) {
        const RT az = vz[k + 1] * iz[v];
        const RT mz = az < a1 ? az : a1;
        pix[v] += (col[k] * (mz - a0) + col[k + 1] * (a1 - mz));
      } else if (k == nzm1) {
        const RT az = vz[nz] * iz[v];
        const RT mz = az < a1 ? az : a1;
        pix[v] += col[nzm1] * (mz - a0);
      } else
        break;
    }
    a0 = a1;
  }
}

/* y = (y + A x) * d_conv over the whole sinogram (c2D.cpp:469-625). */
void FN(or_cb_f2D)(double sx, double sy, double sz, double det_x,
                   const double *h_pix, int nh, const double *v_pix, int nv,
                   const double *ang, int na, const double origin[3],
                   const double vsize[3], int nx, int ny, int nz,
                   const RT *vox, RT *pix)
{
  FN(cb_tab) t;
  RT *dconv = FN(cb_dconv)(sx, sy, sz, det_x, h_pix, nh, v_pix, nv);
  const double step = h_pix[1] - h_pix[0];
  const double ipix = 1.0 / step, hpix0 = (sy - h_pix[0]) / step;
  const double L = det_x - sx;
  const long nyz = (long)ny * nz;
  const double vx = origin[0], wx = origin[0] + (double)nx * vsize[0];
  const double vy = origin[1], wy = origin[1] + (double)ny * vsize[1];
  const int cap = 2 * (nx > ny ? nx : ny) + 2;
  int a;
  FN(cb_tab_make)(&t, sz, v_pix, nv, origin, vsize, nz);
#pragma omp parallel for schedule(dynamic)
  for (a = 0; a < na; a++) {
    const double c = cos(ang[a]), s = sin(ang[a]);
    const double p1x = c * sx - s * sy, p1y = s * sx + c * sy;
    double y00, y01, y10, y11;
    int h, hmin, hmax, h00, h01, h10, h11;
    long *off = (long *)malloc(sizeof(long) * (size_t)cap);
    RT *alpha = (RT *)malloc(sizeof(RT) * (size_t)cap);
    RT *ainv = (RT *)malloc(sizeof(RT) * (size_t)cap);
    int *kv = (int *)malloc(sizeof(int) * (size_t)nv);
    /* detector columns hit by the volume's four corner lines
       (c2D.cpp:558-600) */
    {
      const double dx0 = vx - p1x, dy0 = vy - p1y;
      const double dxn = wx - p1x, dyn = wy - p1y;
      const double cx0 = c * vx - sx, cxn = c * wx - sx;
      const double sy0 = s * vy, syn = s * wy;
      if (fabs(s) > fabs(c)) {
        const double il = L / s;
        y00 = (c - dx0 / (sy0 + cx0)) * il;
        y01 = (c - dx0 / (syn + cx0)) * il;
        y10 = (c - dxn / (sy0 + cxn)) * il;
        y11 = (c - dxn / (syn + cxn)) * il;
      } else {
        const double il = L / c;
        y00 = (-s + dy0 / (cx0 + sy0)) * il;
        y01 = (-s + dyn / (cx0 + syn)) * il;
        y10 = (-s + dy0 / (cxn + sy0)) * il;
        y11 = (-s + dyn / (cxn + syn)) * il;
      }
    }
    h00 = (int)floor(y00 * ipix + hpix0);
    h01 = (int)floor(y01 * ipix + hpix0);
    h10 = (int)floor(y10 * ipix + hpix0);
    h11 = (int)floor(y11 * ipix + hpix0);
    hmin = h00 < h01 ? h00 : h01;
    hmax = h00 > h01 ? h00 : h01;
    if (h10 < hmin) hmin = h10;
    if (h10 > hmax) hmax = h10;
    if (h11 < hmin) hmin = h11;
    if (h11 > hmax) hmax = h11;
    if (hmin < 0) hmin = 0;
    if (hmax > nh - 1) hmax = nh - 1;
    for (h = hmin; h <= hmax; h++) {
      const double p2x = c * det_x - s * h_pix[h];
      const double p2y = s * det_x + c * h_pix[h];
      const int n = FN(cb_walk)((p1x - vx) / vsize[0], (p1y - vy) / vsize[1],
                                (p2x - vx) / vsize[0], (p2y - vy) / vsize[1],
                                nx, ny, nyz, nz, off, alpha);
      if (n > 0)
        FN(cb_zsplit)(pix + ((long)a * nh + h) * nv, vox, off, alpha, n, &t,
                      nv, nz, ainv, kv);
    }
    free(off);
    free(alpha);
    free(ainv);
    free(kv);
  }
#pragma omp parallel for schedule(dynamic)
  for (a = 0; a < na; a++) {
    RT *p = pix + (long)a * nh * nv;
    long q;
    for (q = 0; q < (long)nh * nv; q++)
      p[q] *= dconv[q];
  }
  FN(cb_tab_free)(&t);
  free(dconv);
}

/* Mirror of the z split for one voxel column: scatter every (ray, row)
   back into the column's slices (c2D.cpp:640-751). */
static void FN(cb_zgather)(RT *col, const RT *const *rows, const RT *a0s,
                           const RT *a1s, int n, const FN(cb_tab) *t, int nv,
                           int nz, RT pzdv, RT z1, RT znm)
{
  const int nzm1 = nz - 1, midp = t->mid;
  const RT *dz = t->delta_z, *iz = t->inv_delz, *vz = t->vox_z;
  const RT pzbz = t->pzbz, pzbz1 = t->pzbz + (RT)1.0;
  const RT pzdeps = pzdv * (RT)(1.0 + OR_EPS);
  int m, v, min_v = 0, max_v = nv;
  for (m = 0; m < n; m++) {
    const int lo = (int)ceil(pzdeps + z1 / a0s[m]);
    const int hi = (int)floor(pzdv + znm / a0s[m]);
    if (lo > min_v) min_v = lo;
    if (hi < max_v) max_v = hi;
  }
  for (m = 0; m < n; m++) {
    const RT *p = rows[m];
    const RT a0 = a0s[m], a1 = a1s[m], ai = a0 * t->inv_dz;
    for (v = min_v; v < midp; v++) {
      const int k = (int)(pzbz + ai * dz[v]);
      const RT az = vz[k] * iz[v];
      const RT mz = az < a1 ? az : a1;
      col[k] += p[v] * (mz - a0);
      col[k - 1] += p[v] * (a1 - mz);
    }
    for (v = midp; v < max_v; v++) {
      const int k = (int)(pzbz + ai * dz[v]);
      const RT az = vz[k + 1] * iz[v];
      const RT mz = az < a1 ? az : a1;
      col[k] += p[v] * (mz - a0);
      col[k + 1] += p[v] * (a1 - mz);
    }
  }
  for (m = 0; m < n; m++) {
    const RT *p = rows[m];
    const RT a0 = a0s[m], a1 = a1s[m], ai = a0 * t->inv_dz;
    for (v = min_v - 1; v >= 0; v--) {
      const int k = (int)(pzbz1 + ai * dz[v]) - 1;
      if (k > 0) {
        const RT az = vz[k] * iz[v];
        const RT mz = az < a1 ? az : a1;
        col[k] += p[v] * (mz - a0);
        col[k - 1] += p[v] * (a1 - mz);
      } else if (k == 0) {
        const RT az = vz[0] * iz[v];
        const RT mz = az < a1 ? az : a1;
        col[0] += p[v] * (mz - a0);
      } else
        break;
    }
    for (v = max_v; v < nv; v++) {
      const int k = (int)(pzbz + ai * dz[v]);
      if (k < nzm1) {
        const RT az = vz[k + 1] * iz[v];
        const RT mz = az < a1 ? az : a1;
        col[k] += p[v] * (mz - a0);
        col[k + 1] += p[v] * (a1 - mz);
      } else if (k == nzm1) {
        const RT az = vz[nz] * iz[v];
        const RT mz = az < a1 ? az : a1;
        col[nzm1] += p[v] * (mz - a0);
      } else
        break;
    }
  }
}

/* x += A^T (y * d_conv); y is left unchanged (c2D.cpp:788-1050 and the
   limited_memory = false branch c2D.cpp:1095-1117). */
void FN(or_cb_b2D)(double sx, double sy, double sz, double det_x,
                   const double *h_pix, int nh, const double *v_pix, int nv,
                   const double *ang, int na, const double origin[3],
                   const double vsize[3], int nx, int ny, int nz,
                   const RT *pix, RT *vox)
{
  FN(cb_tab) t;
  RT *dconv = FN(cb_dconv)(sx, sy, sz, det_x, h_pix, nh, v_pix, nv);
  RT *dpix = (RT *)malloc(sizeof(RT) * (size_t)na * nh * nv);
  double *tab = (double *)malloc(sizeof(double) * (size_t)na * 8);
  const double step = h_pix[1] - h_pix[0];
  const double ipix = 1.0 / step, hpix0 = (sy - h_pix[0]) / step;
  const double L = det_x - sx;
  const double vstep = v_pix[1] - v_pix[0];
  const RT pzdv = (RT)((sz - v_pix[0]) / vstep);
  const RT z1 = (RT)((origin[2] + vsize[2] - sz) / vstep);
  const RT znm = (RT)((origin[2] + (double)(nz - 1) * vsize[2] - sz) / vstep);
  long q;
  int a, i;
  FN(cb_tab_make)(&t, sz, v_pix, nv, origin, vsize, nz);
  for (q = 0; q < (long)na * nh * nv; q++)
    dpix[q] = pix[q] * dconv[q % ((long)nh * nv)];
  for (a = 0; a < na; a++) {
    const double c = cos(ang[a]), s = sin(ang[a]);
    tab[8 * a + 0] = c;
    tab[8 * a + 1] = s;
    tab[8 * a + 2] = c * sx - s * sy;   /* source x */
    tab[8 * a + 3] = s * sx + c * sy;   /* source y */
    tab[8 * a + 4] = c * det_x;
    tab[8 * a + 5] = s * det_x;
    tab[8 * a + 6] = L / c;
    tab[8 * a + 7] = L / s;
  }
#pragma omp parallel for schedule(dynamic)
  for (i = 0; i < nx; i++) {
    const double x0 = origin[0] + (double)i * vsize[0];
    const double xn = origin[0] + (double)(i + 1) * vsize[0];
    int cap = 4 * na + 16, j;
    const RT **rows = (const RT **)malloc(sizeof(RT *) * (size_t)cap);
    RT *a0s = (RT *)malloc(sizeof(RT) * (size_t)cap);
    RT *a1s = (RT *)malloc(sizeof(RT) * (size_t)cap);
    for (j = 0; j < ny; j++) {
      const double y0 = origin[1] + (double)j * vsize[1];
      const double yn = origin[1] + (double)(j + 1) * vsize[1];
      int n = 0, a;
      for (a = 0; a < na; a++) {
        const double c = tab[8 * a], s = tab[8 * a + 1];
        const double p1x = tab[8 * a + 2], p1y = tab[8 * a + 3];
        const double dx0 = x0 - p1x, dy0 = y0 - p1y;
        const double dxn = xn - p1x, dyn = yn - p1y;
        const double cx0 = c * x0 - sx, cxn = c * xn - sx;
        const double sy0 = s * y0, syn = s * yn;
        double y00, y01, y10, y11;
        int h, hmin, hmax, h00, h01, h10, h11;
        /* column footprint from its four corner lines (c2D.cpp:837-901) */
        if (fabs(s) > fabs(c)) {
          const double il = tab[8 * a + 7];
          y00 = (c - dx0 / (sy0 + cx0)) * il;
          y01 = (c - dx0 / (syn + cx0)) * il;
          y10 = (c - dxn / (sy0 + cxn)) * il;
          y11 = (c - dxn / (syn + cxn)) * il;
        } else {
          const double il = tab[8 * a + 6];
          y00 = (-s + dy0 / (cx0 + sy0)) * il;
          y01 = (-s + dyn / (cx0 + syn)) * il;
          y10 = (-s + dy0 / (cxn + sy0)) * il;
          y11 = (-s + dyn / (cxn + syn)) * il;
        }
        h00 = (int)floor(y00 * ipix + hpix0);
        h01 = (int)floor(y01 * ipix + hpix0);
        h10 = (int)floor(y10 * ipix + hpix0);
        h11 = (int)floor(y11 * ipix + hpix0);
        hmin = h00 < h01 ? h00 : h01;
        hmax = h00 > h01 ? h00 : h01;
        if (h10 < hmin) hmin = h10;
        if (h10 > hmax) hmax = h10;
        if (h11 < hmin) hmin = h11;
        if (h11 > hmax) hmax = h11;
        if (hmin < 0) hmin = 0;
        if (hmax > nh - 1) hmax = nh - 1;
        for (h = hmin; h <= hmax; h++) {
          /* alpha interval of this ray inside the column (c2D.cpp:904-936) */
          const double p2x = tab[8 * a + 4] - s * h_pix[h];
          const double p2y = tab[8 * a + 5] + c * h_pix[h];
          const double ddx = p2x - p1x, ddy = p2y - p1y;
          if (fabs(ddx) < OR_EPS || fabs(ddy) < OR_EPS)
            continue;
          {
            const double ix = 1.0 / ddx, iy = 1.0 / ddy;
            const double ax0 = dx0 * ix, ay0 = dy0 * iy;
            const double axn = dxn * ix, ayn = dyn * iy;
            const double amin = fmax(fmax(fmin(ax0, axn), fmin(ay0, ayn)), 0.0);
            const double amax = fmin(fmin(fmax(ax0, axn), fmax(ay0, ayn)), 1.0);
            if (amin < amax - OR_EPS) {
              if (n == cap) {
                cap *= 2;
                rows = (const RT **)realloc((void *)rows, sizeof(RT *) * (size_t)cap);
                a0s = (RT *)realloc(a0s, sizeof(RT) * (size_t)cap);
                a1s = (RT *)realloc(a1s, sizeof(RT) * (size_t)cap);
              }
              a0s[n] = (RT)amin;
              a1s[n] = (RT)amax;
              rows[n++] = dpix + ((long)a * nh + h) * nv;
            }
          }
        }
      }
      if (n > 0)
        FN(cb_zgather)(vox + ((long)i * ny + j) * nz, rows, a0s, a1s, n, &t,
                       nv, nz, pzdv, z1, znm);
    }
    free((void *)rows);
    free(a0s);
    free(a1s);
  }
  FN(cb_tab_free)(&t);
  free(tab);
  free(dpix);
  free(dconv);
}

/* ------------------------------------------------------------------ */
/* vector ops and iteration bodies                                     */
/* ------------------------------------------------------------------ */

static void FN(fp)(const or_geom *g, const RT *x, RT *y)
{
  if (g->cone)
    FN(or_cb_f2D)(g->sx, g->sy, g->sz, g->det_x, g->h_pix, g->nh, g->v_pix,
                  g->nv, g->angles, g->na, g->origin, g->vsize, g->nx, g->ny,
                  g->nz, x, y);
  else
    FN(or_pb_f2D)(g->h_pix, g->nh, g->v_pix, g->nv, g->angles, g->na,
                  g->origin, g->vsize, g->nx, g->ny, g->nz, x, y);
}

static void FN(bp)(const or_geom *g, const RT *y, RT *x)
{
  if (g->cone)
    FN(or_cb_b2D)(g->sx, g->sy, g->sz, g->det_x, g->h_pix, g->nh, g->v_pix,
                  g->nv, g->angles, g->na, g->origin, g->vsize, g->nx, g->ny,
                  g->nz, y, x);
  else
    FN(or_pb_b2D)(g->h_pix, g->nh, g->v_pix, g->nv, g->angles, g->na,
                  g->origin, g->vsize, g->nx, g->ny, g->nz, y, x);
}

/* sum of squares with per-outer-index partials (blas.hpp:204-284) */
static RT FN(sumsq)(const RT *p, long n0, long n1, long n2)
{
  RT tot = 0;
  long i;
#pragma omp parallel for schedule(dynamic)
  for (i = 0; i < n0; i++) {
    RT s1 = 0;
    long j, k;
    for (j = 0; j < n1; j++) {
      const RT *q = p + (i * n1 + j) * n2;
      RT s2 = 0;
      for (k = 0; k < n2; k++)
        s2 += q[k] * q[k];
      s1 += s2;
    }
#pragma omp atomic
    tot += s1;
  }
  return tot;
}

/* root mean square of x - truth over n values, in double (fixed order per
   thread block, a double atomic sum: only the last bits vary with the
   thread count) */
static double FN(rmse)(const RT *x, const float *truth, long n)
{
  double tot = 0.0;
  long q;
#pragma omp parallel for schedule(static) reduction(+ : tot)
  for (q = 0; q < n; q++) {
    const double e = (double)x[q] - (double)truth[q];
    tot += e * e;
  }
  return sqrt(tot / (double)n);
}

/* CGLS as cgls_base::reconstruct + cgls_3d (cgls.cpp:26-157).  b is the
   instrument's sinogram and is overwritten with the residual; x receives
   x + CGLS(b) (x is not used to form the first residual).  norms[it] gets
   sqrt(||A^T r||^2) like norm_r.  With truth and rmse given (may be NULL),
   rmse[it] = RMSE(x - truth) after iteration it; iter_s (may be NULL)
   receives each iteration's wall time (projections and vector updates,
   without the RMSE). */
int FN(or_cgls_track)(const or_geom *g, RT *b, RT *x, int niter, double *norms,
                      const float *truth, double *rmse, double *iter_s)
{
  const long nvox = (long)g->nx * g->ny * g->nz;
  const long npix = (long)g->na * g->nh * g->nv;
  RT *d = (RT *)calloc((size_t)nvox, sizeof(RT));
  RT *s = (RT *)malloc(sizeof(RT) * (size_t)nvox);
  RT *Ad = (RT *)malloc(sizeof(RT) * (size_t)npix);
  RT normr2;
  int it;
  long q;
  FN(bp)(g, b, d);
  normr2 = FN(sumsq)(d, g->nx, g->ny, g->nz);
  for (it = 0; it < niter; it++) {
    RT alpha;
    double nn, beta;
    const double t_start = omp_get_wtime();
    memset(Ad, 0, sizeof(RT) * (size_t)npix);
    FN(fp)(g, d, Ad);
    alpha = FN(sumsq)(Ad, g->na, g->nh, g->nv);
    alpha = normr2 / alpha;
#pragma omp parallel for schedule(static)
    for (q = 0; q < nvox; q++)
      x[q] += alpha * d[q];
#pragma omp parallel for schedule(static)
    for (q = 0; q < npix; q++)
      b[q] += (-alpha) * Ad[q];
    memset(s, 0, sizeof(RT) * (size_t)nvox);
    FN(bp)(g, b, s);
    nn = (double)FN(sumsq)(s, g->nx, g->ny, g->nz);
    beta = nn / (double)normr2;
    normr2 = (RT)nn;
    {
      const RT bf = (RT)beta;
#pragma omp parallel for schedule(static)
      for (q = 0; q < nvox; q++)
        d[q] = bf * d[q] + s[q];
    }
    if (norms)
      norms[it] = sqrt((double)normr2);
    if (iter_s)
      iter_s[it] = omp_get_wtime() - t_start;
    if (truth && rmse)
      rmse[it] = FN(rmse)(x, truth, nvox);
  }
  free(d);
  free(s);
  free(Ad);
  return 0;
}

int FN(or_cgls)(const or_geom *g, RT *b, RT *x, int niter, double *norms)
{
  return FN(or_cgls_track)(g, b, x, niter, norms, NULL, NULL, NULL);
}

/* SIRT (sirt.cpp:15-80): R = 1/(A 1), C = 1/(A^T 1) with entries <= 0 set
   to 0 (invert_min_x, blas.hpp:100-116); x starts from the passed x. */
int FN(or_sirt)(const or_geom *g, const RT *b, RT *x, int niter)
{
  const long nvox = (long)g->nx * g->ny * g->nz;
  const long npix = (long)g->na * g->nh * g->nv;
  RT *R = (RT *)calloc((size_t)npix, sizeof(RT));
  RT *C = (RT *)calloc((size_t)nvox, sizeof(RT));
  RT *r = (RT *)malloc(sizeof(RT) * (size_t)npix);
  RT *t = (RT *)malloc(sizeof(RT) * (size_t)(npix > nvox ? npix : nvox));
  int it;
  long q;
  for (q = 0; q < nvox; q++) t[q] = (RT)1.0;
  FN(fp)(g, t, R);
  for (q = 0; q < npix; q++) R[q] = R[q] > 0 ? (RT)(1.0 / (double)R[q]) : (RT)0;
  for (q = 0; q < npix; q++) t[q] = (RT)1.0;
  FN(bp)(g, t, C);
  for (q = 0; q < nvox; q++) C[q] = C[q] > 0 ? (RT)(1.0 / (double)C[q]) : (RT)0;
  for (it = 0; it < niter; it++) {
    memset(r, 0, sizeof(RT) * (size_t)npix);
    FN(fp)(g, x, r);
#pragma omp parallel for schedule(static)
    for (q = 0; q < npix; q++)
      r[q] = ((RT)-1.0 * r[q] + b[q]) * R[q];
    memset(t, 0, sizeof(RT) * (size_t)nvox);
    FN(bp)(g, r, t);
#pragma omp parallel for schedule(static)
    for (q = 0; q < nvox; q++) {
      RT v = x[q] + (RT)1.0 * (t[q] * C[q]);
      x[q] = v < 0 ? (RT)0 : v;
    }
  }
  free(R);
  free(C);
  free(r);
  free(t);
  return 0;
}

/* MLEM (mlem.cpp:15-80): x <- 1; N = 1/(A^T 1) keeping zeros
   (invert_x, blas.hpp:86-98); ratio = b/(A x) with 0 where A x == 0
   (div_xyz, blas.hpp:152-171), clamped >= 0; x *= N * A^T ratio, x >= 0. */
int FN(or_mlem)(const or_geom *g, const RT *b, RT *x, int niter)
{
  const long nvox = (long)g->nx * g->ny * g->nz;
  const long npix = (long)g->na * g->nh * g->nv;
  RT *N = (RT *)calloc((size_t)nvox, sizeof(RT));
  RT *r = (RT *)malloc(sizeof(RT) * (size_t)npix);
  RT *t = (RT *)malloc(sizeof(RT) * (size_t)nvox);
  int it;
  long q;
  for (q = 0; q < nvox; q++) x[q] = (RT)1.0;
  for (q = 0; q < npix; q++) r[q] = (RT)1.0;
  FN(bp)(g, r, N);
  for (q = 0; q < nvox; q++) if (N[q] != 0) N[q] = (RT)(1.0 / (double)N[q]);
  for (it = 0; it < niter; it++) {
    memset(r, 0, sizeof(RT) * (size_t)npix);
    FN(fp)(g, x, r);
#pragma omp parallel for schedule(static)
    for (q = 0; q < npix; q++) {
      RT v = r[q] != 0 ? b[q] / r[q] : (RT)0;
      r[q] = v < 0 ? (RT)0 : v;
    }
    memset(t, 0, sizeof(RT) * (size_t)nvox);
    FN(bp)(g, r, t);
#pragma omp parallel for schedule(static)
    for (q = 0; q < nvox; q++) {
      RT v = x[q] * (N[q] * t[q]);
      x[q] = v < 0 ? (RT)0 : v;
    }
  }
  free(N);
  free(r);
  free(t);
  return 0;
}

/* ------------------------------------------------------------------ */
/* regularisation terms and the regularised CGLS                       */
/* ------------------------------------------------------------------ */

#define OR_MIRROR(i, n, d) ((i) + (d) >= 0 && (i) + (d) < (n) ? (i) + (d) : (i) - (d))

/* Tikhonov term (tikhonov.cpp:4-32): b = 0.25 (a(i+1) + a(i-1) + a(j+1) +
   a(j-1)) - a, a neighbour past a face mirrored to the other side; z is not
   coupled.  The neighbour sum is in RT, the scaling in double (0.25 is a
   double literal).  Needs nx, ny >= 2. */
void FN(or_tikhonov)(const RT *a, RT *b, int nx, int ny, int nz)
{
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < nx; i++) {
    const long i1 = OR_MIRROR(i, nx, 1), i2 = OR_MIRROR(i, nx, -1);
    long j, k;
    for (j = 0; j < ny; j++) {
      const long j1 = OR_MIRROR(j, ny, 1), j2 = OR_MIRROR(j, ny, -1);
      const RT *p1 = a + (i1 * ny + j) * nz, *p2 = a + (i2 * ny + j) * nz;
      const RT *p3 = a + (i * ny + j1) * nz, *p4 = a + (i * ny + j2) * nz;
      const RT *pc = a + (i * ny + j) * nz;
      RT *q = b + (i * ny + j) * nz;
      for (k = 0; k < nz; k++)
        q[k] = (RT)(0.25 * (double)(p1[k] + p2[k] + p3[k] + p4[k]) - (double)pc[k]);
    }
  }
}

static int FN(or_sgn)(RT x) { return (x > 0) - (x < 0); }

/* 0.5 (sign(p) + sign(m)) min(|p|, |m|) in double, stored in RT
   (total_v.cpp:62-67). */
static RT FN(or_minmod)(RT p, RT m)
{
  const RT ap = p < 0 ? -p : p, am = m < 0 ? -m : m;
  return (RT)(0.5 * (double)(FN(or_sgn)(p) + FN(or_sgn)(m)) * (double)(ap < am ? ap : am));
}

/* TV term (total_v.cpp:21-195): the divergence of the normalised forward
   differences, D1 = dx+/T1, D2 = dy+/T2, D3 = dz+/T3, with the reference's
   pairing of min-mod terms in T1..T3 (T2 reuses the z term of T1, T3 the y
   term of T1), mirrored neighbours at the faces, and
   b = (D1 - D1(i-1)) + (D2 - D2(j-1)) + (D3 - D3(k-1)).  Needs nx, ny,
   nz >= 2. */
void FN(or_tv)(const RT *a, RT *b, int nx, int ny, int nz)
{
  const long n = (long)nx * ny * nz;
  RT *D = (RT *)malloc(sizeof(RT) * 3 * (size_t)n);
  long i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < nx; i++) {
    const long i1 = OR_MIRROR(i, nx, 1), i2 = OR_MIRROR(i, nx, -1);
    long j, k;
    for (j = 0; j < ny; j++) {
      const long j1 = OR_MIRROR(j, ny, 1), j2 = OR_MIRROR(j, ny, -1);
      for (k = 0; k < nz; k++) {
        const long k1 = OR_MIRROR(k, nz, 1), k2 = OR_MIRROR(k, nz, -1);
        const long c = (i * ny + j) * nz + k;
        const RT A = a[c];
        const RT x1 = a[(i1 * ny + j) * nz + k] - A, x0 = A - a[(i2 * ny + j) * nz + k];
        const RT y1 = a[(i * ny + j1) * nz + k] - A, y0 = A - a[(i * ny + j2) * nz + k];
        const RT z1 = a[(i * ny + j) * nz + k1] - A, z0 = A - a[(i * ny + j) * nz + k2];
        const RT d1x = x1 * x1, d1y = y1 * y1, d1z = z1 * z1;
        RT d2x = FN(or_minmod)(y1, y0), d2y = FN(or_minmod)(x1, x0);
        RT d3x = FN(or_minmod)(z1, z0);
        d2x = d2x * d2x;
        d2y = d2y * d2y;
        d3x = d3x * d3x;
        {
          const RT d2z = d2y, d3y = d3x, d3z = d2x;
          const RT T1 = (RT)sqrt(d1x + d2x + d3x + (RT)FLT_EPSILON);
          const RT T2 = (RT)sqrt(d1y + d2y + d3y + (RT)FLT_EPSILON);
          const RT T3 = (RT)sqrt(d1z + d2z + d3z + (RT)FLT_EPSILON);
          D[c] = x1 / T1;
          D[n + c] = y1 / T2;
          D[2 * n + c] = z1 / T3;
        }
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (i = 0; i < nx; i++) {
    const long i2 = OR_MIRROR(i, nx, -1);
    long j, k;
    for (j = 0; j < ny; j++) {
      const long j2 = OR_MIRROR(j, ny, -1);
      for (k = 0; k < nz; k++) {
        const long k2 = OR_MIRROR(k, nz, -1);
        const long c = (i * ny + j) * nz + k;
        const RT dv1 = D[c] - D[(i2 * ny + j) * nz + k];
        const RT dv2 = D[n + c] - D[n + (i * ny + j2) * nz + k];
        const RT dv3 = D[2 * n + c] - D[2 * n + (i * ny + j) * nz + k2];
        b[c] = dv1 + dv2 + dv3;
      }
    }
  }
  free(D);
}

/* cgls_regularize::reconstruct (cgls.cpp:343-423) with regularize = the
   Tikhonov (kind 0) or TV (kind 1) term (cgls.cpp:425-435).  Each
   iteration: Ad = A d; Ad_vox += lambda L(d) (Ad_vox is never reset);
   alpha = gamma / (||Ad||^2 + ||Ad_vox||^2); x += alpha d; r -= alpha Ad;
   x += lambda L(x); s = A^T r; beta, d = s + beta d.  The reference also
   updates an unused voxel residual with the sinogram's extents
   (cgls.cpp:390), which is undefined behaviour and has no effect on x or
   the norms; it is left out. */
int FN(or_cgls_reg)(const or_geom *g, RT *b, RT *x, int niter, int kind,
                    double lambda, double *norms)
{
  const long nvox = (long)g->nx * g->ny * g->nz;
  const long npix = (long)g->na * g->nh * g->nv;
  RT *d = (RT *)calloc((size_t)nvox, sizeof(RT));
  RT *adv = (RT *)calloc((size_t)nvox, sizeof(RT));
  RT *L = (RT *)malloc(sizeof(RT) * (size_t)nvox);
  RT *s = (RT *)malloc(sizeof(RT) * (size_t)nvox);
  RT *Ad = (RT *)malloc(sizeof(RT) * (size_t)npix);
  const RT lam = (RT)lambda;
  RT normr2;
  int it;
  long q;
  FN(bp)(g, b, d);
  normr2 = FN(sumsq)(d, g->nx, g->ny, g->nz);
  for (it = 0; it < niter; it++) {
    RT alpha;
    double nn, beta;
    memset(Ad, 0, sizeof(RT) * (size_t)npix);
    FN(fp)(g, d, Ad);
    if (kind == 0) FN(or_tikhonov)(d, L, g->nx, g->ny, g->nz);
    else FN(or_tv)(d, L, g->nx, g->ny, g->nz);
    for (q = 0; q < nvox; q++) adv[q] += lam * L[q];
    alpha = FN(sumsq)(Ad, g->na, g->nh, g->nv) + FN(sumsq)(adv, g->nx, g->ny, g->nz);
    alpha = normr2 / alpha;
    for (q = 0; q < nvox; q++) x[q] += alpha * d[q];
    for (q = 0; q < npix; q++) b[q] += (-alpha) * Ad[q];
    if (kind == 0) FN(or_tikhonov)(x, L, g->nx, g->ny, g->nz);
    else FN(or_tv)(x, L, g->nx, g->ny, g->nz);
    for (q = 0; q < nvox; q++) x[q] += lam * L[q];
    memset(s, 0, sizeof(RT) * (size_t)nvox);
    FN(bp)(g, b, s);
    nn = (double)FN(sumsq)(s, g->nx, g->ny, g->nz);
    beta = nn / (double)normr2;
    normr2 = (RT)nn;
    {
      const RT bf = (RT)beta;
      for (q = 0; q < nvox; q++) d[q] = bf * d[q] + s[q];
    }
    if (norms) norms[it] = sqrt((double)normr2);
  }
  free(d);
  free(adv);
  free(L);
  free(s);
  free(Ad);
  return 0;
}
#undef OR_MIRROR
