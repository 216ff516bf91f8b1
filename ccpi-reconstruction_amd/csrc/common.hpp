// Shared definitions of the MI355X projector library: status handling,
// plan layouts shared between host launchers and gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "ccpi_hip.h"

namespace ccpi {

// Carries a CCPI_E_* status through host code; converted to a return code
// at the C ABI so no exception crosses it.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define CCPI_HIP(call)                                                      \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      throw ::ccpi::Error(e_ == hipErrorOutOfMemory ? CCPI_E_NOMEM          \
                                                    : CCPI_E_HIP,           \
                          std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CCPI_REQUIRE(cond, code, msg)                                       \
  do {                                                                      \
    if (!(cond)) throw ::ccpi::Error((code), (msg));                        \
  } while (0)

// Model tolerance of every branch decision (p2D.cpp:20, c2D.cpp:15).
constexpr double kEps = 1.1920928955078125e-07;  // FLT_EPSILON

// ---- cone beam plan layout ------------------------------------------------

// One source->pixel ray of the xy walk.  Entry m of the ray (m = 0..n-1)
// holds the walk parameter alpha_m and, for m >= 1, the xy cell whose
// segment is [alpha_{m-1}, alpha_m] (c2D.cpp:193-467 layout).
struct alignas(32) ConeRay {
  uint64_t off;   // first entry in the entry table
  int n;          // number of entries (0: ray misses / outside footprint)
  int ncom;       // first segment where an edge row can leave the volume
  int min_v;      // lowest row still inside at the start of the last segment
  int max_v;      // highest such row
  int pad[2];
};

struct ConeEnt {
  uint32_t col;   // i * ny + j
  float alpha;
};

// Backprojection list of one voxel column: its (angle, h) rays in angle-
// then-h order, and the column's row-set bounds (c2D.cpp:657-676).
struct alignas(32) BpCol {
  uint64_t off;   // first record in the record table
  int n;          // number of records
  int min_v;      // rows >= min_v stay above the bottom for every ray
  int max_v;      // rows < max_v stay below the top for every ray
  float amin;     // smallest walk parameter at which a ray enters the column
  int pad[2];
};

// One record (16 bytes): x = ray (a * nh + h), y / z = alpha interval of
// the ray inside the column (float bits), w = vlo | vhi << 16, the rows
// whose ray is still inside the volume's z range at the interval start.
using BpRec = uint4;

// Kernel argument block of the cone projectors (by value).
struct ConeArgs {
  int na, nh, nv, nx, ny, nz;
  int mid;               // first detector row above the source
  int bp_stride;         // BP row distance of adjacent lanes
  int bp_unique;         // stride guarantees distinct slices per instruction
  int bp_main_safe;      // source inside the volume's z range
  int bp_phased;         // BP accumulates in one array, in lane phases
  int bp_phases;         // lane phases (lanes of one phase on distinct slices)
  float bp_rps1;         // voxel z size / smallest row pitch: rows per slice at
                         // walk parameter 1
  int nvp;               // row stride of the BP's scaled sinogram (padded)
  float zb, zt;          // volume bottom / top relative to source z
  const float *alim;     // [nv] walk parameter where row v leaves the volume's
                         // z range (>= 2 where it never does)
  float pzbz, pzbz1, inv_dz;
  float pzdv, z1, znm;   // BP safe-row bounds (c2D.cpp:993-1000)
  double vx, vy, dx, dy; // volume xy origin and voxel size
  double sx, sy, det_x, L;
  double ipix, hpix0;    // detector column index map
  const double *atab;    // [na][8]: c, s, p1x, p1y, c*det_x, s*det_x, L/c, L/s
  const double *h_pix;   // [nh]
  const int *hrange;     // [na][2] footprint of the whole volume
  const float *dz;       // [nv] v_pix - source_z
  const float *iz;       // [nv] 1 / dz
  const float *vz;       // [nz+1] slice planes - source_z
  const float *dconv;    // [nh][nv] source -> pixel distance
};

// ---- parallel beam plan layout ---------------------------------------------

struct ParArgs {
  int na, nh, nv, nx, ny, nz;
  int map_type;          // 1, 2, 4 (rows per slice) or 0 (general)
  double ox, oy, dx, dy; // volume xy origin and voxel size
  double D;              // detector line distance 2 * max|origin|
  double dconv;          // float(3 D) as double: walk start bound
  double ihp, hp0;       // detector column index map
  const double *atab;    // [na][5]: c, s, y_offset, 1/y_offset, length
  const double *h_pix;   // [nh]
  const int *map;        // [nv] row -> slice
  const int *rlo;        // [nz] first row of slice k
  const int *rhi;        // [nz] one past its last row
  // walk table (null: the FP walks each ray itself): ray r's cells are
  // entries woff[r] .. woff[r + 1] - 1 of went, {x * ny + y, float bits of
  // the xy length}, in walk order
  const unsigned long long *woff;
  const uint2 *went;
};

// ---- projector epilogue norms ----------------------------------------------
// A projector given a `part` array leaves FP64 partial sums of the squares
// of the values it wrote there (a fixed number per launch, each from a fixed
// set of threads in a fixed order), so a solver's reduction over `part` is
// deterministic and no pass re-reads the projector's output for its norm.

// Sum of v over the 64 lanes of a wavefront (every lane gets it).
__device__ inline double wave_sum64(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Partial of a 256-thread workgroup: the lanes' values summed per wavefront,
// then the four wavefronts in order into part[idx].  Every thread of the
// workgroup calls it (it synchronises).
__device__ inline void block_part256(double v, double *part, size_t idx) {
  __shared__ double sh[4];
  v = wave_sum64(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[idx] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

}  // namespace ccpi
