/*
 * CPU oracle library (TEST INFRASTRUCTURE ONLY, never linked into the
 * product).  Instantiates oracle_impl.h for float (the reference's own
 * precision choices) and for double (the FP64 oracle of the same model).
 */
#include <float.h>
#include <math.h>
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
