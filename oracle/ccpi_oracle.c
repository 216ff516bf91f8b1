/*
 * CPU oracle library (TEST INFRASTRUCTURE ONLY, never linked into the
 * product).  Instantiates oracle_impl.h for float (the reference's own
 * precision choices) and for double (the FP64 oracle of the same model).
 */
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* Geometry of one projector instance, host memory.  Angles in radians,
   h/v pixel positions and lengths in mm, volume [x][y][z] with z fastest,
   sinogram [angle][h][v] with v fastest (Core/src/base_types.hpp:60-63). */
typedef struct or_geom {
  int cone;
  int na, nh, nv;
  const double *angles, *h_pix, *v_pix;
  double sx, sy, sz, det_x;
  double origin[3], vsize[3];
  int nx, ny, nz;
} or_geom;

#define RT float
#define FN(x) x##_f32
#include "oracle_impl.h"
#undef RT
#undef FN
#undef OR_EPS

#define RT double
#define FN(x) x##_f64
#include "oracle_impl.h"
#undef RT
#undef FN

/* ------------------------------------------------------------------ */
/* XTek centre-of-rotation search (Core/src/Readers/xtek.cpp:566-758,  */
/* T. Liu, Opt. Eng. 48 (2009) 046501), float data as the reference.   */
/* ------------------------------------------------------------------ */

/* M_PI of <math.h> (not declared under -std=c99) */
#define OC_PI 3.14159265358979323846

/* (f1 (x2 - x) + f2 (x - x1)) / (x2 - x1) in double, stored in float
   (xtek.cpp:566-570) */
static float oc_lin(double x1, double x2, float f1, float f2, double x)
{
  return (float)((f1 * (x2 - x) + f2 * (x - x1)) / (x2 - x1));
}

/* bilinear in (x, y), f_ab = f(x_a, y_b) (xtek.cpp:572-580) */
static float oc_bil(double x1, double x2, double y1, double y2, float f11, float f12,
                    float f21, float f22, double x, double y)
{
  return (float)((f11 * (x2 - x) * (y2 - y) + f21 * (x - x1) * (y2 - y)
                  + f12 * (x2 - x) * (y - y1) + f22 * (x - x1) * (y - y1))
                 / ((x2 - x1) * (y2 - y1)));
}

/* interpolate2D (xtek.cpp:582-669): d[a * nh + k] at (new_h, new_a); the
   angle axis wraps at 2 pi.  The searches take the first index whose
   successor exceeds the value, as the reference's linear scans. */
static float oc_interp(const double *h, const double *ang, const float *d, int nh, int na,
                       double new_h, double new_a)
{
  const double tp = 2.0 * OC_PI;
  int p1 = 0, a1 = 0;
  while (p1 < nh - 1 && !(h[p1 + 1] > new_h)) p1++;
#define D(a, k) d[(size_t)(a) * nh + (k)]
  if (h[p1] == new_h) {
    if (new_a < ang[0]) return oc_lin(ang[na - 1] - tp, ang[0], D(na - 1, p1), D(0, p1), new_a);
    while (a1 < na - 1 && !(ang[a1 + 1] > new_a)) a1++;
    if (ang[a1] == new_a) return D(a1, p1);
    if (a1 == na - 1) return oc_lin(ang[a1], ang[0] + tp, D(a1, p1), D(0, p1), new_a);
    return oc_lin(ang[a1], ang[a1 + 1], D(a1, p1), D(a1 + 1, p1), new_a);
  }
  if (new_a < ang[0])
    return oc_bil(h[p1], h[p1 + 1], ang[na - 1] - tp, ang[0], D(na - 1, p1), D(0, p1),
                  D(na - 1, p1 + 1), D(0, p1 + 1), new_h, new_a);
  while (a1 < na - 1 && !(ang[a1 + 1] > new_a)) a1++;
  if (ang[a1] == new_a) return oc_lin(h[p1], h[p1 + 1], D(a1, p1), D(a1, p1 + 1), new_h);
  if (a1 == na - 1)
    return oc_bil(h[p1], h[p1 + 1], ang[a1], ang[0] + tp, D(a1, p1), D(0, p1), D(a1, p1 + 1),
                  D(0, p1 + 1), new_h, new_a);
  return oc_bil(h[p1], h[p1 + 1], ang[a1], ang[a1 + 1], D(a1, p1), D(a1 + 1, p1),
                D(a1, p1 + 1), D(a1 + 1, p1 + 1), new_h, new_a);
#undef D
}

/* find_centre (xtek.cpp:671-758) on row v_slice of pixels [na][nh][nv]:
   five refinement levels of 21 candidate centres; for each, the mean
   squared difference between every ray and its interpolated conjugate ray
   (float sum in (angle, h) order).  Writes the y shift that the reference
   adds to source_y and the h pixels; returns -1 when no candidate has a
   conjugate ray on the detector. */
int or_find_centre(const or_geom *g, const float *pixels, int v_slice, double *y_centre,
                   double *m_out)
{
  const int ns = 21, na = g->na, nh = g->nh, nv = g->nv;
  const double prec[5] = {1, 0.1, 0.01, 0.001, 0.0001};
  const double *h = g->h_pix, *ph = g->angles;
  const double dist = g->det_x - g->sx;
  double mid = 0.0;
  double *gi = (double *)malloc(sizeof(double) * nh), *be = (double *)malloc(sizeof(double) * nh);
  double *s2 = (double *)malloc(sizeof(double) * nh);
  float *d = (float *)malloc(sizeof(float) * (size_t)na * nh);
  int i, j, k, a;
  for (a = 0; a < na; a++)
    for (k = 0; k < nh; k++) d[(size_t)a * nh + k] = pixels[((size_t)a * nh + k) * nv + v_slice];
  for (k = 0; k < nh; k++) gi[k] = atan(h[k] / dist);
  for (i = 0; i < 5; i++) {
    const double scor = mid - (double)(ns / 2) * prec[i];
    float M[21];
    int ind = -1;
    double mmin = 1e20;
    for (j = 0; j < ns; j++) {
      const double gc = atan((mid + (j - ns / 2) * prec[i]) / dist);
      float sum = 0;
      int count = 0;
      for (k = 0; k < nh; k++) {
        be[k] = OC_PI + 2.0 * (gi[k] - gc);
        s2[k] = dist * tan(2.0 * gc - gi[k]);
      }
      for (a = 0; a < na; a++)
        for (k = 0; k < nh; k++) {
          if (s2[k] >= h[0] && s2[k] <= h[nh - 1]) {
            double ab = ph[a] + be[k];
            float p, t;
            while (ab < ph[0]) ab += 2.0 * OC_PI;
            while (ab > ph[na - 1]) ab -= 2.0 * OC_PI;
            p = oc_interp(h, ph, d, nh, na, s2[k], ab);
            t = d[(size_t)a * nh + k] - p;
            sum += t * t;
            count++;
          }
        }
      M[j] = sum / (float)count;
      if (m_out) m_out[i * ns + j] = M[j];
    }
    for (j = 0; j < ns; j++)
      if (mmin > M[j]) {
        ind = j;
        mmin = M[j];
      }
    if (ind < 0) {
      free(gi); free(be); free(s2); free(d);
      return -1;
    }
    mid = scor + ind * prec[i];
  }
  *y_centre = mid * g->sx / dist;
  free(gi); free(be); free(s2); free(d);
  return 0;
}
