// Per-geometry device plans and their cache.
//
// A plan holds everything about one projector instance that does not
// depend on the data: per-angle / per-row / per-slice tables and, for the
// cone beam, the xy walk of every ray.  Plans are built once and reused by
// every projector call on the same geometry (the reference recomputes all
// of it inside every f2D / b2D call, c2D.cpp:480-516, p2D.cpp:477-495).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace ccpi {

// Owned copy of a ccpi_geometry.
struct Geometry {
  int kind = CCPI_PARALLEL;
  int na = 0, nh = 0, nv = 0, nx = 0, ny = 0, nz = 0;
  std::vector<double> angles, h_pix, v_pix;
  double source[3] = {0, 0, 0};
  double det_x = 0;
  double origin[3] = {0, 0, 0};
  double vsize[3] = {0, 0, 0};

  static Geometry from(const ccpi_geometry *g);
  static Geometry parallel(const double *h, const double *v, const double *ang,
                           int na, int nh, int nv, const double o[3],
                           const double s[3], int nx, int ny, int nz);
  static Geometry cone(double sx, double sy, double sz, double dx,
                       const double *h, const double *v, const double *ang,
                       int na, int nh, int nv, const double o[3],
                       const double s[3], int nx, int ny, int nz);
  // Throws Error(CCPI_E_ARG / CCPI_E_GEOMETRY) for inputs the model cannot
  // represent.
  void validate() const;
  std::string key() const;
  size_t n_vox() const { return (size_t)nx * ny * nz; }
  size_t n_pix() const { return (size_t)na * nh * nv; }
};

// Owning device allocation.
struct DeviceBuffer {
  void *p = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  ~DeviceBuffer() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    if (n <= bytes && p) return;
    reset();
    CCPI_HIP(hipMalloc(&p, n ? n : 4));
    bytes = n;
  }
  template <class T>
  void upload(const std::vector<T> &v, hipStream_t st) {
    alloc(v.size() * sizeof(T));
    if (!v.empty())
      CCPI_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T),
                              hipMemcpyHostToDevice, st));
  }
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

// BP scratch of the calls that bring none of their own: the cone BP's
// scaled sinogram (y * d_conv, the reference's limited_memory = false
// copy, c2D.cpp:1105-1114) or the parallel BP's angle-group partials.
// One per plan, shared by every such caller of the geometry:
// a caller holds `mu` while it enqueues the scaling and the BP, and waits
// on the device for `done` (recorded after the previous caller's BP), so
// calls from different streams or host threads run one after another on
// the device and never overwrite a copy another BP still reads.
struct SharedScratch {
  mutable std::mutex mu;
  DeviceBuffer buf;
  hipEvent_t done = nullptr;
  SharedScratch() = default;
  SharedScratch(const SharedScratch &) = delete;
  SharedScratch &operator=(const SharedScratch &) = delete;
  ~SharedScratch() {
    if (done) (void)hipEventDestroy(done);
  }
  size_t bytes() const {
    std::lock_guard<std::mutex> lk(mu);
    return buf.bytes;
  }
};

struct ConePlan {
  Geometry g;
  ConeArgs args{};
  DeviceBuffer atab, hpix, hrange, dz, iz, vz, dconv, offs, rays, ents, eai, alim;
  DeviceBuffer bpcol, bprec;  // backprojection lists
  SharedScratch scratch;      // BP scratch of callers without a workspace
  uint64_t n_entries = 0;
  size_t bytes() const {
    return atab.bytes + hpix.bytes + hrange.bytes + dz.bytes + iz.bytes +
           vz.bytes + dconv.bytes + offs.bytes + rays.bytes + ents.bytes + eai.bytes + alim.bytes +
           bpcol.bytes + bprec.bytes + scratch.bytes();
  }
};

// Launch shape of the tiled parallel BP (detector rows staged in LDS per
// 16 x 16 column tile and block of angles); tiled = 0 keeps the
// column-per-lane kernel (rows per slice other than 1, or
// footprints too wide for LDS).
struct ParBpShape {
  int tiled = 0;
  int slices = 16;   // slices per lane (8 or 16)
  int hmax = 0;      // detector columns staged per angle
  int ab = 8;        // angles per staged block
  int groups = 1;    // angle groups summed through scratch partials
};

struct ParPlan {
  Geometry g;
  ParArgs args{};
  ParBpShape bp{};
  DeviceBuffer atab, hpix, map, rlo, rhi;
  DeviceBuffer woff, went;  // walk table (args.woff / args.went)
  SharedScratch scratch;    // BP partials of callers without a workspace
  size_t bytes() const {
    return atab.bytes + hpix.bytes + map.bytes + rlo.bytes + rhi.bytes + woff.bytes +
           went.bytes + scratch.bytes();
  }
};

// Walk table of a parallel plan: built for FPs of fewer than
// kParWalkTableWaves wavefronts (4 detector rows per lane) when its bound
// (8 bytes per cell, at most nx + ny cells per ray, plus the ray offsets)
// is within kParWalkTableBytes.
constexpr size_t kParWalkTableBytes = size_t(4) << 30;
constexpr long long kParWalkTableWaves = 8192;
inline size_t par_walk_table_bound(const Geometry &g) {
  const size_t rays = (size_t)g.na * g.nh;
  return rays * 8 * (size_t)(g.nx + g.ny) + (rays + 1) * 8;
}
inline long long par_fp_waves(const Geometry &g) {
  return (long long)g.na * g.nh * ((g.nv + 3) / 4) / 64;
}
inline bool par_walk_table_wanted(const Geometry &g) {
  return par_walk_table_bound(g) <= kParWalkTableBytes && par_fp_waves(g) < kParWalkTableWaves;
}

std::shared_ptr<ConePlan> cone_plan(const Geometry &g, hipStream_t st);
std::shared_ptr<ParPlan> par_plan(const Geometry &g, hipStream_t st);
void plan_cache_clear();
long long plan_cache_bytes();

// Accumulation layout of the cone BP for a geometry (host only): row stride
// S, whether it is the phased one-array form and its lane phases, the bound
// on detector rows per slice it was chosen for, the padded row length, and
// the rows per slice at walk parameter 1 (voxel z size / smallest pitch).
struct BpLayout {
  int stride = 1, unique = 1, main_safe = 0, phased = 0, phases = 1, nvp = 0, want = 1;
  double rps1 = 1.0;
};
BpLayout cone_bp_layout(const Geometry &g);

// launchers (cone.hip, parallel.hip).  f64: accumulate the projector sums
// in double (the FP64-accumulation option of the solvers); the default is
// the reference's float sums.  acc_in: add to the output (the reference's
// f2D / b2D) or overwrite it.  norm: null, or *_norm_parts(P) doubles that
// receive FP64 partial sums of the squares of the output (its ||.||^2 is
// their sum, in index order; the epilogue norms of common.hpp).
void cone_build_plan(ConePlan &P, hipStream_t st);
void cone_fp(const ConePlan &P, const float *vol, float *sino, hipStream_t st,
             bool f64 = false, bool acc_in = true, double *norm = nullptr);
size_t cone_fp_norm_parts(const ConePlan &P);
// dpix holds na * nh rows of args.nvp floats (cone_dpix_floats)
size_t cone_dpix_floats(const ConePlan &P);
void cone_scale(const ConePlan &P, const float *pix, float *dpix, hipStream_t st);
void cone_bp_scaled(const ConePlan &P, const float *dpix, float *vol, hipStream_t st,
                    bool f64 = false, bool acc_in = true, double *norm = nullptr);
size_t cone_bp_norm_parts(const ConePlan &P);

void par_build_plan(ParPlan &P, hipStream_t st);
// Parallel FP without any device allocation: the geometry's tables go into
// a fixed file-scope device array of kNoAllocScratchBytes (parallel.hip).
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
size_t par_noalloc_bytes(const Geometry &g);
void par_fp_noalloc(const Geometry &g, const float *vol, float *sino, hipStream_t st);
void par_fp(const ParPlan &P, const float *vol, float *sino, hipStream_t st, bool f64 = false,
            bool acc_in = true, double *norm = nullptr);
size_t par_fp_norm_parts(const ParPlan &P);
// scratch: par_bp_scratch_floats(P) floats (partials of the angle groups),
// unused when that is 0
void par_bp(const ParPlan &P, const float *sino, float *vol, float *scratch, hipStream_t st,
            bool f64 = false, bool acc_in = true, double *norm = nullptr);
size_t par_bp_scratch_floats(const ParPlan &P);
size_t par_bp_norm_parts(const ParPlan &P);
// nonzero system-matrix entries (gathered voxel operands of one FP)
void cone_nnz(const ConePlan &P, unsigned long long *count, hipStream_t st);
void par_nnz(const ParPlan &P, unsigned long long *count, hipStream_t st);

// Projector pair on one geometry with device-resident plan; BP leaves the
// sinogram untouched.
struct Projector {
  std::shared_ptr<ConePlan> cone;
  std::shared_ptr<ParPlan> par;
  explicit Projector(const Geometry &g, hipStream_t st);
  // f64: projector sums accumulated in double (the solvers' FP64-
  // accumulation option); default: the reference's float sums.  acc_in /
  // norm: as the launchers above (norm: fp_norm_parts() / bp_norm_parts()
  // doubles)
  void fp(const float *vol, float *sino, hipStream_t st, bool f64 = false, bool acc_in = true,
          double *norm = nullptr) const;
  // scratch: bp_scratch_floats() floats owned by the caller and used by
  // no other call meanwhile (a solver's workspace); nullptr takes the
  // plan's shared scratch
  void bp(const float *sino, float *vol, hipStream_t st, float *scratch = nullptr,
          bool f64 = false, bool acc_in = true, double *norm = nullptr) const;
  // BP of an already scaled sinogram in the padded BP layout (cone only):
  // a solver's fused update kernels write it
  void bp_scaled(const float *dpix, float *vol, hipStream_t st, bool acc_in = true,
                 double *norm = nullptr) const;
  size_t bp_scratch_floats() const;
  size_t fp_norm_parts() const;
  size_t bp_norm_parts() const;
  const Geometry &geometry() const { return cone ? cone->g : par->g; }
};

}  // namespace ccpi
