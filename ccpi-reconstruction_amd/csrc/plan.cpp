// Geometry validation and the per-device plan cache.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <list>
#include <map>
#include <condition_variable>
#include <mutex>
#include <set>

namespace ccpi {

static std::vector<double> copy_arr(const double *p, int n) {
  return std::vector<double>(p, p + n);
}

Geometry Geometry::from(const ccpi_geometry *g) {
  CCPI_REQUIRE(g, CCPI_E_ARG, "null geometry");
  CCPI_REQUIRE(g->angles && g->h_pixels && g->v_pixels, CCPI_E_ARG,
               "null geometry array");
  CCPI_REQUIRE(g->n_angles > 0 && g->n_h > 0 && g->n_v > 0, CCPI_E_ARG,
               "non-positive detector size");
  Geometry G;
  G.kind = g->kind;
  G.na = g->n_angles; G.nh = g->n_h; G.nv = g->n_v;
  G.nx = g->nx; G.ny = g->ny; G.nz = g->nz;
  G.angles = copy_arr(g->angles, g->n_angles);
  G.h_pix = copy_arr(g->h_pixels, g->n_h);
  G.v_pix = copy_arr(g->v_pixels, g->n_v);
  for (int q = 0; q < 3; q++) {
    G.source[q] = g->source[q];
    G.origin[q] = g->origin[q];
    G.vsize[q] = g->voxel_size[q];
  }
  G.det_x = g->detector_x;
  G.validate();
  return G;
}

Geometry Geometry::parallel(const double *h, const double *v, const double *ang,
                            int na, int nh, int nv, const double o[3],
                            const double s[3], int nx, int ny, int nz) {
  CCPI_REQUIRE(h && v && ang && o && s, CCPI_E_ARG, "null geometry array");
  CCPI_REQUIRE(na > 0 && nh > 0 && nv > 0, CCPI_E_ARG, "non-positive detector size");
  Geometry G;
  G.kind = CCPI_PARALLEL;
  G.na = na; G.nh = nh; G.nv = nv; G.nx = nx; G.ny = ny; G.nz = nz;
  G.angles = copy_arr(ang, na);
  G.h_pix = copy_arr(h, nh);
  G.v_pix = copy_arr(v, nv);
  for (int q = 0; q < 3; q++) { G.origin[q] = o[q]; G.vsize[q] = s[q]; }
  G.validate();
  return G;
}

Geometry Geometry::cone(double sx, double sy, double sz, double dx,
                        const double *h, const double *v, const double *ang,
                        int na, int nh, int nv, const double o[3],
                        const double s[3], int nx, int ny, int nz) {
  Geometry G = parallel(h, v, ang, na, nh, nv, o, s, nx, ny, nz);
  G.kind = CCPI_CONE;
  G.source[0] = sx; G.source[1] = sy; G.source[2] = sz;
  G.det_x = dx;
  G.validate();
  return G;
}

void Geometry::validate() const {
  CCPI_REQUIRE(kind == CCPI_PARALLEL || kind == CCPI_CONE, CCPI_E_ARG, "bad kind");
  CCPI_REQUIRE(nx > 0 && ny > 0 && nz > 0, CCPI_E_ARG, "non-positive volume size");
  CCPI_REQUIRE((long long)nx * ny < (1LL << 31), CCPI_E_LIMIT, "nx*ny too large");
  for (int q = 0; q < 3; q++)
    CCPI_REQUIRE(vsize[q] > 0 && std::isfinite(vsize[q]) && std::isfinite(origin[q]),
                 CCPI_E_GEOMETRY, "voxel size must be positive and finite");
  // the detector index map uses the first pixel step (p2D.cpp:493-494,
  // c2D.cpp:518-520)
  CCPI_REQUIRE(nh >= 2, CCPI_E_GEOMETRY, "need at least two detector columns");
  for (int h = 1; h < nh; h++)
    CCPI_REQUIRE(h_pix[h] > h_pix[h - 1], CCPI_E_GEOMETRY,
                 "h pixel positions must increase");
  for (int v = 1; v < nv; v++)
    CCPI_REQUIRE(v_pix[v] > v_pix[v - 1], CCPI_E_GEOMETRY,
                 "v pixel positions must increase");
  for (double a : angles) CCPI_REQUIRE(std::isfinite(a), CCPI_E_GEOMETRY, "bad angle");
  if (kind == CCPI_CONE) {
    // b2D divides by nz - 1 (c2D.cpp:808) and uses v_pix[1] - v_pix[0]
    CCPI_REQUIRE(nz >= 2, CCPI_E_GEOMETRY, "cone beam needs nz >= 2");
    CCPI_REQUIRE(nv >= 2, CCPI_E_GEOMETRY, "cone beam needs two detector rows");
    // a row at the source height has no defined slope (c2D.cpp:487-492)
    for (int v = 0; v < nv; v++)
      CCPI_REQUIRE(std::fabs(v_pix[v] - source[2]) >= kEps, CCPI_E_GEOMETRY,
                   "detector row at the source height");
    // the walk starts at the source, which must lie outside the volume's
    // xy footprint at every angle: outside its circumscribed circle
    double R = 0.0;   // farthest volume corner from the rotation axis
    for (int cx = 0; cx < 2; cx++)
      for (int cy = 0; cy < 2; cy++)
        R = std::max(R, std::hypot(origin[0] + cx * nx * vsize[0],
                                   origin[1] + cy * ny * vsize[1]));
    CCPI_REQUIRE(std::hypot(source[0], source[1]) > R, CCPI_E_GEOMETRY,
                 "source inside the volume's rotation circle");
    CCPI_REQUIRE(det_x > source[0], CCPI_E_GEOMETRY,
                 "detector must be on the far side of the source");
    // the slice logic of f2D / b2D (c2D.cpp:17-191, 657-749) takes the
    // source within the volume's z extent: half a slice beyond it the
    // reference's sums are undefined (ray sums of order 1e35, varying run
    // to run, on an 8x8x4 volume whose exact chords are at most 10)
    CCPI_REQUIRE(source[2] >= origin[2] && source[2] <= origin[2] + (double)nz * vsize[2],
                 CCPI_E_GEOMETRY, "source outside the volume's z range");
  }
}

std::string Geometry::key() const {
  std::string k;
  auto put = [&](const void *p, size_t n) { k.append((const char *)p, n); };
  int hdr[7] = {kind, na, nh, nv, nx, ny, nz};
  put(hdr, sizeof hdr);
  put(angles.data(), angles.size() * 8);
  put(h_pix.data(), h_pix.size() * 8);
  put(v_pix.data(), v_pix.size() * 8);
  put(source, sizeof source);
  put(&det_x, sizeof det_x);
  put(origin, sizeof origin);
  put(vsize, sizeof vsize);
  return k;
}

// ---- cache ---------------------------------------------------------------

namespace {
struct Entry {
  std::shared_ptr<ConePlan> cone;
  std::shared_ptr<ParPlan> par;
  size_t bytes() const { return cone ? cone->bytes() : par->bytes(); }
};
// The mutex guards the list and the set of keys being built; plans are
// built outside it, so lookups of cached geometries never wait for a build
// (a 512^3 cone plan takes seconds).  A second caller of a geometry that is
// being built waits for that build instead of starting its own.
struct Cache {
  std::mutex mu;
  std::condition_variable built;
  std::set<std::string> building;
  std::list<std::pair<std::string, Entry>> lru;  // front = most recent
};
Cache &cache() {
  static Cache c;
  return c;
}
constexpr size_t kMaxPlans = 8;
// plans beyond this total (the least recently used first) are dropped; the
// newest plan always stays
constexpr size_t kMaxPlanBytes = size_t(64) << 30;

std::string dev_key(const Geometry &g) {
  int dev = 0;
  CCPI_HIP(hipGetDevice(&dev));
  return std::to_string(dev) + ":" + g.key();
}

Entry *lookup(Cache &c, const std::string &k) {
  for (auto it = c.lru.begin(); it != c.lru.end(); ++it)
    if (it->first == k) {
      c.lru.splice(c.lru.begin(), c.lru, it);
      return &c.lru.front().second;
    }
  return nullptr;
}

size_t total_bytes(const Cache &c) {
  size_t b = 0;
  for (auto &q : c.lru) b += q.second.bytes();
  return b;
}

// Device bytes a plan of g will hold, from the table sizes: the cone walk
// lists about one entry per crossed xy cell of each ray and the BP lists
// about two detector columns per angle and voxel column.  Used to make room
// before a build, so the cache's peak is not the old plans plus the new one.
size_t plan_estimate(const Geometry &g) {
  if (g.kind != CCPI_CONE) {
    return (size_t)g.na * 48 + (size_t)(g.nh + g.nv + 2 * g.nz) * 8 +
           (par_walk_table_wanted(g) ? par_walk_table_bound(g) : 0);
  }
  const size_t rays = (size_t)g.na * g.nh, cols = (size_t)g.nx * g.ny;
  return rays * (48 + 12 * (size_t)(g.nx + g.ny)) + cols * (32 + 32 * (size_t)g.na);
}

// Drops least recently used plans until `incoming` more bytes fit; callers
// still holding a dropped plan keep it alive until they release it.
void make_room(Cache &c, size_t incoming) {
  while (!c.lru.empty() &&
         (c.lru.size() + 1 > kMaxPlans || total_bytes(c) + incoming > kMaxPlanBytes))
    c.lru.pop_back();
}

template <class Build>
Entry find_or_build(const Geometry &g, Build build) {
  Cache &c = cache();
  const std::string k = dev_key(g);
  std::unique_lock<std::mutex> lk(c.mu);
  for (;;) {
    if (Entry *e = lookup(c, k)) return *e;
    if (!c.building.count(k)) break;
    c.built.wait(lk);
  }
  c.building.insert(k);
  make_room(c, plan_estimate(g));
  lk.unlock();
  Entry e;
  try {
    e = build();
  } catch (...) {
    lk.lock();
    c.building.erase(k);
    c.built.notify_all();
    throw;
  }
  lk.lock();
  c.building.erase(k);
  make_room(c, e.bytes());
  c.lru.emplace_front(k, e);
  c.built.notify_all();
  return e;
}
}  // namespace

std::shared_ptr<ConePlan> cone_plan(const Geometry &g, hipStream_t st) {
  return find_or_build(g, [&] {
           auto P = std::make_shared<ConePlan>();
           P->g = g;
           cone_build_plan(*P, st);
           return Entry{P, nullptr};
         })
      .cone;
}

std::shared_ptr<ParPlan> par_plan(const Geometry &g, hipStream_t st) {
  return find_or_build(g, [&] {
           auto P = std::make_shared<ParPlan>();
           P->g = g;
           par_build_plan(*P, st);
           return Entry{nullptr, P};
         })
      .par;
}

void plan_cache_clear() {
  Cache &c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  c.lru.clear();
}

long long plan_cache_bytes() {
  Cache &c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  return (long long)total_bytes(c);
}

// ---- projector pair -------------------------------------------------------

Projector::Projector(const Geometry &g, hipStream_t st) {
  if (g.kind == CCPI_CONE)
    cone = cone_plan(g, st);
  else
    par = par_plan(g, st);
}

void Projector::fp(const float *vol, float *sino, hipStream_t st, bool f64, bool acc_in,
                   double *norm) const {
  if (cone)
    cone_fp(*cone, vol, sino, st, f64, acc_in, norm);
  else
    par_fp(*par, vol, sino, st, f64, acc_in, norm);
}

size_t Projector::bp_scratch_floats() const {
  return cone ? cone_dpix_floats(*cone) : par_bp_scratch_floats(*par);
}

size_t Projector::fp_norm_parts() const {
  return cone ? cone_fp_norm_parts(*cone) : par_fp_norm_parts(*par);
}

size_t Projector::bp_norm_parts() const {
  return cone ? cone_bp_norm_parts(*cone) : par_bp_norm_parts(*par);
}

void Projector::bp_scaled(const float *dpix, float *vol, hipStream_t st, bool acc_in,
                          double *norm) const {
  CCPI_REQUIRE(cone, CCPI_E_ARG, "pre-scaled backprojection is a cone-beam form");
  cone_bp_scaled(*cone, dpix, vol, st, false, acc_in, norm);
}

void Projector::bp(const float *sino, float *vol, hipStream_t st, float *scratch, bool f64,
                   bool acc_in, double *norm) const {
  auto run = [&](float *scr) {
    if (cone) {
      cone_scale(*cone, sino, scr, st);
      cone_bp_scaled(*cone, scr, vol, st, f64, acc_in, norm);
    } else {
      par_bp(*par, sino, vol, scr, st, f64, acc_in, norm);
    }
  };
  const size_t need = bp_scratch_floats();
  if (scratch || need == 0) {
    run(scratch);
    return;
  }
  SharedScratch &S = const_cast<SharedScratch &>(cone ? cone->scratch : par->scratch);
  std::lock_guard<std::mutex> lk(S.mu);
  if (!S.done) CCPI_HIP(hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
  if (!S.buf.p) {
    S.buf.alloc(need * sizeof(float));
  } else {
    // the previous caller's BP (any stream) has finished with the buffer
    CCPI_HIP(hipStreamWaitEvent(st, S.done, 0));
  }
  run(S.buf.as<float>());
  CCPI_HIP(hipEventRecord(S.done, st));
}

}  // namespace ccpi
