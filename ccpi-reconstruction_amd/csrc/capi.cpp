// C ABI of libccpi_hip.so (declared in include/ccpi_hip.h).  Every entry
// point converts exceptions to status codes; nothing C++ crosses it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "ccpi_hip.h"
#include "common.hpp"
#include "plan.hpp"
#include "solvers.hpp"

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F &&f) {
  try {
    f();
    g_last_error.clear();
    return CCPI_OK;
  } catch (const ccpi::Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return CCPI_E_NOMEM;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return CCPI_E_HIP;
  }
}

hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

int ccpi_pb_f2D(const double *h_pixels, const double *v_pixels, const double *angles,
                int n_angles, int n_h, int n_v, const double vox_origin[3],
                const double vox_size[3], int nx, int ny, int nz, float *pixels,
                const float *voxels, void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(pixels && voxels, CCPI_E_ARG, "null data pointer");
    const ccpi::Geometry G = ccpi::Geometry::parallel(h_pixels, v_pixels, angles, n_angles,
                                                      n_h, n_v, vox_origin, vox_size, nx, ny, nz);
    ccpi::Projector P(G, S(stream));
    P.fp(voxels, pixels, S(stream));
  });
}

int ccpi_pb_f2D_noalloc(const double *h_pixels, const double *v_pixels, const double *angles,
                        int n_angles, int n_h, int n_v, const double vox_origin[3],
                        const double vox_size[3], int nx, int ny, int nz, float *pixels,
                        const float *voxels, void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(pixels && voxels, CCPI_E_ARG, "null data pointer");
    const ccpi::Geometry G = ccpi::Geometry::parallel(h_pixels, v_pixels, angles, n_angles,
                                                      n_h, n_v, vox_origin, vox_size, nx, ny, nz);
    ccpi::par_fp_noalloc(G, voxels, pixels, S(stream));
  });
}

int ccpi_pb_b2D(const double *h_pixels, const double *v_pixels, const double *angles,
                const float *pixels, float *voxels, int n_angles, int n_h, int n_v,
                const double vox_origin[3], const double vox_size[3], int nx, int ny, int nz,
                void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(pixels && voxels, CCPI_E_ARG, "null data pointer");
    const ccpi::Geometry G = ccpi::Geometry::parallel(h_pixels, v_pixels, angles, n_angles,
                                                      n_h, n_v, vox_origin, vox_size, nx, ny, nz);
    ccpi::Projector P(G, S(stream));
    P.bp(pixels, voxels, S(stream));
  });
}

int ccpi_cb_f2D(double source_x, double source_y, double source_z, double detector_x,
                const double *h_pixels, const double *v_pixels, const double *angles,
                float *pixels, const float *voxels, int n_angles, int n_h, int n_v,
                const double vox_origin[3], const double vox_size[3], int nx, int ny, int nz,
                void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(pixels && voxels, CCPI_E_ARG, "null data pointer");
    const ccpi::Geometry G =
        ccpi::Geometry::cone(source_x, source_y, source_z, detector_x, h_pixels, v_pixels,
                             angles, n_angles, n_h, n_v, vox_origin, vox_size, nx, ny, nz);
    ccpi::Projector P(G, S(stream));
    P.fp(voxels, pixels, S(stream));
  });
}

int ccpi_cb_b2D(double source_x, double source_y, double source_z, double detector_x,
                const double *h_pixels, const double *v_pixels, const double *angles,
                const float *pixels, float *voxels, int n_angles, int n_h, int n_v,
                const double vox_origin[3], const double vox_size[3], int nx, int ny, int nz,
                void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(pixels && voxels, CCPI_E_ARG, "null data pointer");
    const ccpi::Geometry G =
        ccpi::Geometry::cone(source_x, source_y, source_z, detector_x, h_pixels, v_pixels,
                             angles, n_angles, n_h, n_v, vox_origin, vox_size, nx, ny, nz);
    ccpi::Projector P(G, S(stream));
    P.bp(pixels, voxels, S(stream));
  });
}

int ccpi_forward_project(const ccpi_geometry *g, float *pixels, const float *voxels,
                         void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(pixels && voxels, CCPI_E_ARG, "null data pointer");
    ccpi::Projector P(ccpi::Geometry::from(g), S(stream));
    P.fp(voxels, pixels, S(stream));
  });
}

int ccpi_backward_project(const ccpi_geometry *g, const float *pixels, float *voxels,
                          void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(pixels && voxels, CCPI_E_ARG, "null data pointer");
    ccpi::Projector P(ccpi::Geometry::from(g), S(stream));
    P.bp(pixels, voxels, S(stream));
  });
}

int ccpi_cgls(const ccpi_geometry *g, float *b, float *x, int iterations, double *norms,
              int flags, void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(b && x, CCPI_E_ARG, "null data pointer");
    CCPI_REQUIRE(iterations >= 0, CCPI_E_ARG, "negative iteration count");
    ccpi::Projector P(ccpi::Geometry::from(g), S(stream));
    ccpi::Workspace W;
    ccpi::run_cgls(P, b, x, iterations, norms, (flags & CCPI_REDUCE_FP32) != 0, W,
                   S(stream));
    CCPI_HIP(hipStreamSynchronize(S(stream)));
  });
}

int ccpi_regularize(int kind, const float *a, float *out, int nx, int ny, int nz,
                    void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(a && out, CCPI_E_ARG, "null data pointer");
    CCPI_REQUIRE(a != out, CCPI_E_ARG, "in-place regularisation is not supported");
    CCPI_REQUIRE(kind == CCPI_REG_TIKHONOV || kind == CCPI_REG_TV, CCPI_E_ARG,
                 "unknown regularisation kind");
    ccpi::check_reg_dims(nx, ny, nz);
    ccpi::DeviceBuffer scr;
    const size_t nscr = ccpi::reg_scratch_floats(kind, (size_t)nx * ny * nz);
    if (nscr) scr.alloc(nscr * sizeof(float));
    ccpi::reg_term(kind, a, out, nscr ? scr.as<float>() : nullptr, nx, ny, nz, S(stream));
    CCPI_HIP(hipStreamSynchronize(S(stream)));
  });
}

int ccpi_cgls_reg(const ccpi_geometry *g, float *b, float *x, int iterations, int kind,
                  double lambda, double *norms, int flags, void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(b && x, CCPI_E_ARG, "null data pointer");
    CCPI_REQUIRE(iterations >= 0, CCPI_E_ARG, "negative iteration count");
    CCPI_REQUIRE(kind == CCPI_REG_TIKHONOV || kind == CCPI_REG_TV, CCPI_E_ARG,
                 "unknown regularisation kind");
    ccpi::Projector P(ccpi::Geometry::from(g), S(stream));
    ccpi::Workspace W;
    ccpi::run_cgls_reg(P, b, x, iterations, kind, lambda, norms,
                       (flags & CCPI_REDUCE_FP32) != 0, W, S(stream));
  });
}

int ccpi_xtek_find_centre(const ccpi_geometry *g, const float *pixels, int v_slice,
                          double *y_centre, void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(pixels && y_centre, CCPI_E_ARG, "null data pointer");
    const ccpi::Geometry G = ccpi::Geometry::from(g);
    *y_centre = ccpi::xtek_find_centre(G, pixels, v_slice, S(stream));
  });
}

int ccpi_sirt(const ccpi_geometry *g, const float *b, float *x, int iterations, int flags,
              void *stream) {
  (void)flags;
  return guarded([&] {
    CCPI_REQUIRE(b && x, CCPI_E_ARG, "null data pointer");
    CCPI_REQUIRE(iterations >= 0, CCPI_E_ARG, "negative iteration count");
    ccpi::Projector P(ccpi::Geometry::from(g), S(stream));
    ccpi::Workspace W;
    ccpi::run_sirt(P, b, x, iterations, W, S(stream));
    CCPI_HIP(hipStreamSynchronize(S(stream)));
  });
}

int ccpi_mlem(const ccpi_geometry *g, const float *b, float *x, int iterations, int flags,
              void *stream) {
  return guarded([&] {
    CCPI_REQUIRE((flags & ~CCPI_ACCUM_FP64) == 0, CCPI_E_ARG, "unknown MLEM flag");
    CCPI_REQUIRE(b && x, CCPI_E_ARG, "null data pointer");
    CCPI_REQUIRE(iterations >= 0, CCPI_E_ARG, "negative iteration count");
    ccpi::Projector P(ccpi::Geometry::from(g), S(stream));
    ccpi::Workspace W;
    ccpi::run_mlem(P, b, x, iterations, W, S(stream), (flags & CCPI_ACCUM_FP64) != 0);
    CCPI_HIP(hipStreamSynchronize(S(stream)));
  });
}

}  // extern "C"

// Session state behind the opaque ccpi_cgls_state handle.
struct ccpi_cgls_state {
  ccpi::Projector P;
  ccpi::Workspace W;
  ccpi::EventLog log;
  ccpi::CglsRun run;
  hipStream_t st;
  int max_iters;
  int done = 0;
  ccpi_cgls_state(const ccpi::Geometry &G, float *b, float *x, int n, int flags,
                  hipStream_t s)
      : P(G, s), run{P, W, b, x, n, (flags & CCPI_REDUCE_FP32) != 0}, st(s), max_iters(n) {
    if (flags & CCPI_TIMING) run.log = &log;
  }
};

extern "C" {

int ccpi_cgls_create(const ccpi_geometry *g, float *b, float *x, int max_iterations,
                     int flags, void *stream, ccpi_cgls_state **state) {
  return guarded([&] {
    CCPI_REQUIRE(b && x && state, CCPI_E_ARG, "null pointer");
    CCPI_REQUIRE(max_iterations >= 0, CCPI_E_ARG, "negative iteration count");
    *state = nullptr;
    auto *s = new ccpi_cgls_state(ccpi::Geometry::from(g), b, x, max_iterations, flags,
                                  S(stream));
    try {
      s->run.start(s->st);
    } catch (...) {
      delete s;
      throw;
    }
    *state = s;
  });
}

int ccpi_cgls_iterate(ccpi_cgls_state *s, int n) {
  return guarded([&] {
    CCPI_REQUIRE(s && n >= 0, CCPI_E_ARG, "bad session call");
    CCPI_REQUIRE(s->done + n <= s->max_iters, CCPI_E_LIMIT,
                 "more iterations than the session was created for");
    for (int q = 0; q < n; q++) s->run.iterate(s->done++, s->st);
  });
}

int ccpi_cgls_norms(ccpi_cgls_state *s, double *norms, int *done) {
  return guarded([&] {
    CCPI_REQUIRE(s, CCPI_E_ARG, "null session");
    if (norms && s->done > 0) {
      CCPI_HIP(hipMemcpyAsync(norms, s->run.norms_dev.p, (size_t)s->done * sizeof(double),
                              hipMemcpyDeviceToHost, s->st));
    }
    CCPI_HIP(hipStreamSynchronize(s->st));
    if (done) *done = s->done;
  });
}

int ccpi_cgls_timings(ccpi_cgls_state *s, double *fp_ms, int *fp_count, double *bp_ms,
                      int *bp_count) {
  return guarded([&] {
    CCPI_REQUIRE(s, CCPI_E_ARG, "null session");
    CCPI_HIP(hipStreamSynchronize(s->st));
    if (fp_ms) *fp_ms = ccpi::EventLog::total_ms(s->log.fp);
    if (bp_ms) *bp_ms = ccpi::EventLog::total_ms(s->log.bp);
    if (fp_count) *fp_count = (int)(s->log.fp.size() / 2);
    if (bp_count) *bp_count = (int)(s->log.bp.size() / 2);
  });
}

int ccpi_cgls_destroy(ccpi_cgls_state *s) {
  return guarded([&] {
    if (!s) return;
    (void)hipStreamSynchronize(s->st);
    delete s;
  });
}

int ccpi_projector_nnz(const ccpi_geometry *g, unsigned long long *nnz, void *stream) {
  return guarded([&] {
    CCPI_REQUIRE(nnz, CCPI_E_ARG, "null output");
    ccpi::Projector P(ccpi::Geometry::from(g), S(stream));
    *nnz = ccpi::projector_nnz(P, S(stream));
  });
}

int ccpi_pb_cgls_batched(const ccpi_geometry *geoms, int count, float *const *b,
                         float *const *x, int iterations, double *norms, int flags,
                         void *stream) {
  (void)flags;
  return guarded([&] {
    CCPI_REQUIRE(geoms && b && x && count >= 0, CCPI_E_ARG, "bad batch arguments");
    CCPI_REQUIRE(iterations >= 0, CCPI_E_ARG, "negative iteration count");
    std::vector<ccpi::Geometry> G;
    for (int i = 0; i < count; i++) {
      CCPI_REQUIRE(geoms[i].kind == CCPI_PARALLEL, CCPI_E_ARG,
                   "batched CGLS takes parallel-beam problems");
      CCPI_REQUIRE(b[i] && x[i], CCPI_E_ARG, "null data pointer");
      G.push_back(ccpi::Geometry::from(&geoms[i]));
    }
    ccpi::run_pb_cgls_batched(G, b, x, iterations, norms, S(stream));
  });
}

int ccpi_cone_bp_layout(const ccpi_geometry *g, int *row_stride, int *phased, int *phases,
                        int *rows_bound) {
  return guarded([&] {
    CCPI_REQUIRE(row_stride && phased && phases && rows_bound, CCPI_E_ARG, "null output");
    const ccpi::Geometry G = ccpi::Geometry::from(g);
    CCPI_REQUIRE(G.kind == CCPI_CONE, CCPI_E_ARG, "BP layout query takes a cone geometry");
    const ccpi::BpLayout L = ccpi::cone_bp_layout(G);
    *row_stride = L.stride;
    *phased = L.phased;
    *phases = L.phases;
    *rows_bound = L.want;
  });
}

const char *ccpi_error_string(int status) {
  switch (status) {
    case CCPI_OK: return "ok";
    case CCPI_E_ARG: return "invalid argument";
    case CCPI_E_GEOMETRY: return "geometry not representable by the projector model";
    case CCPI_E_HIP: return "HIP runtime error";
    case CCPI_E_NOMEM: return "device allocation failed";
    case CCPI_E_LIMIT: return "size beyond a kernel limit";
    default: return "unknown status";
  }
}

const char *ccpi_last_error(void) { return g_last_error.c_str(); }

int ccpi_version(void) { return 100; }

void ccpi_plan_cache_clear(void) { ccpi::plan_cache_clear(); }

long long ccpi_plan_cache_bytes(void) { return ccpi::plan_cache_bytes(); }

}  // extern "C"
