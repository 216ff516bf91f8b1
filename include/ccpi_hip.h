/*
 * ccpi_hip.h - C ABI of the MI355X projector + iterative-solver library
 * (libccpi_hip.so).
 *
 * Conventions (same as the reference, Core/src/base_types.hpp:31-63):
 *   volume    float [nx][ny][nz], z fastest      ("voxels")
 *   sinogram  float [n_angles][n_h][n_v], v fastest ("pixels")
 *   geometry  double, host memory: angles in radians, detector pixel
 *             positions and volume origin / voxel size in mm.
 * Every float array argument is DEVICE memory on the current HIP device.
 * Calls are asynchronous on `stream` (a hipStream_t, NULL = default
 * stream); host outputs (norm arrays) are valid when the call returns.
 * Every entry point returns 0 on success or a negative CCPI_E_* code; no
 * C++ exception crosses this boundary.  No entry point frees or keeps
 * caller memory.
 */
#ifndef CCPI_HIP_H
#define CCPI_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CCPI_OK              0
#define CCPI_E_ARG          -1  /* null pointer / non-positive size */
#define CCPI_E_GEOMETRY     -2  /* geometry the model cannot represent */
#define CCPI_E_HIP          -3  /* HIP runtime error */
#define CCPI_E_NOMEM        -4  /* device allocation failed */
#define CCPI_E_LIMIT        -5  /* size beyond a kernel limit */

/* flags for the solver entry points */
#define CCPI_REDUCE_FP32     1  /* alpha/beta and norms in float, as the
                                   reference (cgls.cpp:118-121, blas.hpp:204-305);
                                   default is FP64 reductions */

#define CCPI_PARALLEL 0
#define CCPI_CONE     1

/* One projector instance (instrument state of CCPi::parallel_beam /
   CCPi::cone_beam, Core/src/instruments.hpp:130-143, 194-199). */
typedef struct ccpi_geometry {
  int kind;                 /* CCPI_PARALLEL or CCPI_CONE */
  int n_angles, n_h, n_v;
  const double *angles;     /* [n_angles] radians, host */
  const double *h_pixels;   /* [n_h] mm, host, strictly increasing */
  const double *v_pixels;   /* [n_v] mm, host, strictly increasing */
  double source[3];         /* cone: source at angle 0 (x, y, z) mm */
  double detector_x;        /* cone: detector plane x at angle 0, mm */
  double origin[3];         /* volume corner, mm */
  double voxel_size[3];     /* mm */
  int nx, ny, nz;
} ccpi_geometry;

/* ---- projectors: the static f2D / b2D contract ------------------------ */

/* y += A_par x.  Replaces CCPi::parallel_beam::f2D
   (Core/src/instruments.hpp:266-271, Core/src/p2D.cpp:468-611). */
int ccpi_pb_f2D(const double *h_pixels, const double *v_pixels,
                const double *angles, int n_angles, int n_h, int n_v,
                const double vox_origin[3], const double vox_size[3],
                int nx, int ny, int nz, float *pixels, const float *voxels,
                void *stream);

/* ccpi_pb_f2D that allocates no device memory: no counterpart in the
   reference.  For callers that cannot allocate (code under HIP graph
   capture, serving loops, harnesses that forbid hipMalloc).  Same
   arguments and result as ccpi_pb_f2D; instead of a cached plan the
   geometry's tables go into a fixed 1 MiB device array of the library
   (per GPU, held until the process ends), so the geometry is limited to
   40 n_angles + 8 n_h + 4 n_v + 8 nz bytes (each part rounded up to 256)
   <= 1 MiB; larger ones return CCPI_E_LIMIT before any device work.
   Calls on one GPU share that array: calls that overlap on one GPU (other
   streams, other host threads) race; calls on one stream are ordered. */
int ccpi_pb_f2D_noalloc(const double *h_pixels, const double *v_pixels,
                        const double *angles, int n_angles, int n_h, int n_v,
                        const double vox_origin[3], const double vox_size[3],
                        int nx, int ny, int nz, float *pixels,
                        const float *voxels, void *stream);

/* x += A_par^T y.  Replaces CCPi::parallel_beam::b2D
   (Core/src/instruments.hpp:272-277, Core/src/p2D.cpp:950-1087). */
int ccpi_pb_b2D(const double *h_pixels, const double *v_pixels,
                const double *angles, const float *pixels, float *voxels,
                int n_angles, int n_h, int n_v, const double vox_origin[3],
                const double vox_size[3], int nx, int ny, int nz,
                void *stream);

/* y = (y + A_cone x) * d_conv.  Replaces CCPi::cone_beam::f2D
   (Core/src/instruments.hpp:164-171, Core/src/c2D.cpp:469-625). */
int ccpi_cb_f2D(double source_x, double source_y, double source_z,
                double detector_x, const double *h_pixels,
                const double *v_pixels, const double *angles, float *pixels,
                const float *voxels, int n_angles, int n_h, int n_v,
                const double vox_origin[3], const double vox_size[3],
                int nx, int ny, int nz, void *stream);

/* x += A_cone^T (y * d_conv), y untouched (limited_memory = false).
   Replaces CCPi::cone_beam::b2D (Core/src/instruments.hpp:172-179,
   Core/src/c2D.cpp:948-1119). */
int ccpi_cb_b2D(double source_x, double source_y, double source_z,
                double detector_x, const double *h_pixels,
                const double *v_pixels, const double *angles,
                const float *pixels, float *voxels, int n_angles, int n_h,
                int n_v, const double vox_origin[3],
                const double vox_size[3], int nx, int ny, int nz,
                void *stream);

/* Same projectors on a geometry descriptor (the instrument's virtual
   forward_project / backward_project, Core/src/instruments.hpp:61-69). */
int ccpi_forward_project(const ccpi_geometry *g, float *pixels,
                         const float *voxels, void *stream);
int ccpi_backward_project(const ccpi_geometry *g, const float *pixels,
                          float *voxels, void *stream);

/* ---- iteration bodies --------------------------------------------------
   reconstruction_alg::reconstruct (Core/src/algorithms.hpp:13-21).
   b: measured sinogram (device); x: volume (device, in/out). */

/* CGLS, cgls_3d (Core/src/Algorithms/cgls.cpp:26-157).  b is overwritten
   with the final residual and x receives x + CGLS(b), as the reference.
   norms (host, may be NULL) receives ||A^T r|| after each iteration. */
int ccpi_cgls(const ccpi_geometry *g, float *b, float *x, int iterations,
              double *norms, int flags, void *stream);

/* SIRT (Core/src/Algorithms/sirt.cpp:15-80); x is the starting point. */
int ccpi_sirt(const ccpi_geometry *g, const float *b, float *x,
              int iterations, int flags, void *stream);

/* MLEM (Core/src/Algorithms/mlem.cpp:15-80); x is reset to 1 first.
   flags: 0 (the reference's float sums) or CCPI_ACCUM_FP64: the
   projections' sums - A x, the ratio backprojection A^T (b / A x) and the
   sensitivity A^T 1 - accumulate in double and are rounded to float once
   (projector weights and element-wise steps stay float). */
#define CCPI_ACCUM_FP64      4
int ccpi_mlem(const ccpi_geometry *g, const float *b, float *x,
              int iterations, int flags, void *stream);

/* ---- regularised CGLS ---------------------------------------------------
   Regularisation terms CCPi::tikhonov_regularize / CCPi::tv_regularize
   (Core/src/regularize.hpp:6-9; tikhonov.cpp:4-32, total_v.cpp:21-195):
   out = L(a) for a [nx][ny][nz] device volume.  Neighbours past a face are
   mirrored, so every extent must be >= 2 (CCPI_E_ARG otherwise).  a and
   out must not overlap. */
#define CCPI_REG_TIKHONOV    0
#define CCPI_REG_TV          1
int ccpi_regularize(int kind, const float *a, float *out, int nx, int ny,
                    int nz, void *stream);

/* cgls_tikhonov / cgls_tv_reg (Core/src/Algorithms/cgls.hpp:103-131,
   cgls.cpp:343-435) with regularisation parameter lambda.  Same buffer
   contract as ccpi_cgls: b becomes the residual, x receives x + the
   iterations' updates, norms[i] = ||A^T r|| after iteration i. */
int ccpi_cgls_reg(const ccpi_geometry *g, float *b, float *x, int iterations,
                  int kind, double lambda, double *norms, int flags,
                  void *stream);

/* ---- ingest -------------------------------------------------------------
   Nikon_XTek::find_centre(v_slice) (Core/src/Readers/xtek.cpp:671-758):
   centre-of-rotation search on row v_slice of the device sinogram
   pixels [n_angles][n_h][n_v] of a cone geometry.  *y_centre receives the
   shift the reference then adds to source_y and to every h pixel
   (xtek.cpp:755-756); the geometry itself is not modified.  Fails with
   CCPI_E_GEOMETRY where the reference reports "Failure in XTek find
   centre". */
int ccpi_xtek_find_centre(const ccpi_geometry *g, const float *pixels,
                          int v_slice, double *y_centre, void *stream);

/* ---- CGLS session: the same body split into launch-only steps ----------
   No counterpart in the reference (its loop is internal to
   cgls_base::reconstruct, cgls.cpp:54-82).  For callers that interleave
   iterations with their own work or time them; iterate() enqueues work on
   the session's stream without any host synchronisation.
   CCPI_TIMING brackets every projector launch with HIP events. */
#define CCPI_TIMING          2

typedef struct ccpi_cgls_state ccpi_cgls_state;

int ccpi_cgls_create(const ccpi_geometry *g, float *b, float *x,
                     int max_iterations, int flags, void *stream,
                     ccpi_cgls_state **state);
int ccpi_cgls_iterate(ccpi_cgls_state *state, int n);
/* Host sync; norms[0..done) = ||A^T r|| per iteration, *done = count. */
int ccpi_cgls_norms(ccpi_cgls_state *state, double *norms, int *done);
/* Host sync; summed kernel milliseconds of the forward projections and of
   the backprojections since creation (CCPI_TIMING), and their counts. */
int ccpi_cgls_timings(ccpi_cgls_state *state, double *fp_ms, int *fp_count,
                      double *bp_ms, int *bp_count);
int ccpi_cgls_destroy(ccpi_cgls_state *state);

/* Number of nonzero system-matrix entries of a geometry's projector (the
   voxel operands one forward projection gathers), counted on the device. */
int ccpi_projector_nnz(const ccpi_geometry *g, unsigned long long *nnz,
                       void *stream);

/* ---- batched parallel-beam CGLS ----------------------------------------
   No counterpart in the reference (one problem per call there).  Runs
   `count` independent parallel-beam CGLS problems of uneven sizes in one
   persistent launch per step, captured in a hipGraph.  b[i] / x[i] are
   device pointers per problem; norms (host, may be NULL) is
   [count][iterations]. */
int ccpi_pb_cgls_batched(const ccpi_geometry *geoms, int count,
                         float *const *b, float *const *x, int iterations,
                         double *norms, int flags, void *stream);

/* BP accumulation layout the library picks for a cone geometry, computed
   on the host only (no device work, usable without a GPU): *row_stride =
   detector rows per lane (S), *phased = 1 for the one-array form whose
   lanes update in *phases rounds, *rows_bound = the bound on detector rows
   per voxel slice over all columns that the choice guarantees (lanes of one
   round are >= that many rows apart, or alone in their round).  No
   counterpart in the reference: its b2D accumulates each voxel column on
   one host thread (c2D.cpp:1016-1046). */
int ccpi_cone_bp_layout(const ccpi_geometry *g, int *row_stride, int *phased,
                        int *phases, int *rows_bound);

/* ---- utilities --------------------------------------------------------- */
const char *ccpi_error_string(int status);
/* Detail of the last failing call on this host thread (empty after a
   successful call). */
const char *ccpi_last_error(void);
int ccpi_version(void);
/* Drop cached per-geometry plans (device tables) on the current device. */
void ccpi_plan_cache_clear(void);
/* Device bytes held by cached plans on the current device. */
long long ccpi_plan_cache_bytes(void);

#ifdef __cplusplus
}
#endif
#endif /* CCPI_HIP_H */
