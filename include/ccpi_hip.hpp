/*
 * ccpi_hip.hpp - header-only C++ facade over the C ABI (ccpi_hip.h) with
 * the shape of the reference's instrument / algorithm classes:
 *   CCPi::instrument::forward_project / backward_project
 *       (Core/src/instruments.hpp:61-69)  -> ccpi::hip::parallel_beam,
 *                                            ccpi::hip::cone_beam
 *   CCPi::reconstruction_alg::reconstruct (Core/src/algorithms.hpp:13-21)
 *                                         -> ccpi::hip::cgls_3d, sirt, mlem
 * Data pointers are device pointers; geometry is host memory.  Failures
 * throw ccpi::hip::error carrying the CCPI_E_* status.
 */
#ifndef CCPI_HIP_HPP
#define CCPI_HIP_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "ccpi_hip.h"

namespace ccpi {
namespace hip {

struct error : std::runtime_error {
  int status;
  explicit error(int s)
      : std::runtime_error(std::string("ccpi: ") + ccpi_error_string(s) + ": " +
                           ccpi_last_error()),
        status(s) {}
};

inline void check(int s) {
  if (s != CCPI_OK) throw error(s);
}

// Instrument state shared by both beam types (instruments.hpp:130-143).
class instrument {
 public:
  virtual ~instrument() = default;
  void set_phi(const std::vector<double> &radians) { phi_ = radians; }
  void set_h_pixels(const std::vector<double> &mm) { h_ = mm; }
  void set_v_pixels(const std::vector<double> &mm) { v_ = mm; }
  int get_num_angles() const { return (int)phi_.size(); }
  int get_num_h_pixels() const { return (int)h_.size(); }
  int get_num_v_pixels() const { return (int)v_.size(); }
  const std::vector<double> &get_phi() const { return phi_; }

  // pixels: device [n_angles][n_h][n_v]; voxels: device [nx][ny][nz]
  void forward_project(float *pixels, const float *voxels, const double origin[3],
                       const double width[3], int nx, int ny, int nz,
                       void *stream = nullptr) const {
    const ccpi_geometry g = geometry(origin, width, nx, ny, nz);
    check(ccpi_forward_project(&g, pixels, voxels, stream));
  }
  void backward_project(const float *pixels, float *voxels, const double origin[3],
                        const double width[3], int nx, int ny, int nz,
                        void *stream = nullptr) const {
    const ccpi_geometry g = geometry(origin, width, nx, ny, nz);
    check(ccpi_backward_project(&g, pixels, voxels, stream));
  }
  virtual ccpi_geometry geometry(const double origin[3], const double width[3], int nx,
                                 int ny, int nz) const = 0;

 protected:
  ccpi_geometry base(int kind, const double origin[3], const double width[3], int nx,
                     int ny, int nz) const {
    ccpi_geometry g{};
    g.kind = kind;
    g.n_angles = get_num_angles();
    g.n_h = get_num_h_pixels();
    g.n_v = get_num_v_pixels();
    g.angles = phi_.data();
    g.h_pixels = h_.data();
    g.v_pixels = v_.data();
    for (int q = 0; q < 3; q++) {
      g.origin[q] = origin[q];
      g.voxel_size[q] = width[q];
    }
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    return g;
  }
  std::vector<double> phi_, h_, v_;
};

class parallel_beam : public instrument {
 public:
  ccpi_geometry geometry(const double origin[3], const double width[3], int nx, int ny,
                         int nz) const override {
    return base(CCPI_PARALLEL, origin, width, nx, ny, nz);
  }
};

class cone_beam : public instrument {
 public:
  void set_source(double x, double y, double z) {
    src_[0] = x;
    src_[1] = y;
    src_[2] = z;
  }
  void set_detector(double x) { det_x_ = x; }
  ccpi_geometry geometry(const double origin[3], const double width[3], int nx, int ny,
                         int nz) const override {
    ccpi_geometry g = base(CCPI_CONE, origin, width, nx, ny, nz);
    for (int q = 0; q < 3; q++) g.source[q] = src_[q];
    g.detector_x = det_x_;
    return g;
  }

 private:
  double src_[3] = {0, 0, 0};
  double det_x_ = 0;
};

// reconstruction_alg: reconstruct(device, voxels, origin, voxel_size) with the
// instrument's measured sinogram passed explicitly (the reference keeps it
// inside the instrument, instruments.hpp:111).
class reconstruction_alg {
 public:
  explicit reconstruction_alg(int iterations, int flags = 0)
      : iterations_(iterations), flags_(flags) {}
  virtual ~reconstruction_alg() = default;
  virtual bool reconstruct(const instrument &device, float *pixels, float *voxels,
                           const double origin[3], const double voxel_size[3], int nx,
                           int ny, int nz, void *stream = nullptr) = 0;

 protected:
  int iterations_, flags_;
};

// cgls_3d (cgls.cpp:26-157): `pixels` is overwritten with the residual.
class cgls_3d : public reconstruction_alg {
 public:
  using reconstruction_alg::reconstruction_alg;
  bool reconstruct(const instrument &device, float *pixels, float *voxels,
                   const double origin[3], const double voxel_size[3], int nx, int ny,
                   int nz, void *stream = nullptr) override {
    const ccpi_geometry g = device.geometry(origin, voxel_size, nx, ny, nz);
    norms_.assign(iterations_, 0.0);
    check(ccpi_cgls(&g, pixels, voxels, iterations_, norms_.data(), flags_, stream));
    return true;
  }
  // ||A^T r|| per iteration (convergence_data, cgls.cpp:20-24)
  const std::vector<double> &convergence_data() const { return norms_; }

 private:
  std::vector<double> norms_;
};

// cgls_regularize / cgls_tikhonov / cgls_tv_reg (cgls.hpp:103-131,
// cgls.cpp:343-435): the CGLS loop with lambda L(.) added to the normal
// equations' voxel term and to x each iteration.
class cgls_regularize : public reconstruction_alg {
 public:
  cgls_regularize(int iterations, double param, int kind, int flags = 0)
      : reconstruction_alg(iterations, flags), param_(param), kind_(kind) {}
  bool reconstruct(const instrument &device, float *pixels, float *voxels,
                   const double origin[3], const double voxel_size[3], int nx, int ny,
                   int nz, void *stream = nullptr) override {
    const ccpi_geometry g = device.geometry(origin, voxel_size, nx, ny, nz);
    norms_.assign(iterations_, 0.0);
    check(ccpi_cgls_reg(&g, pixels, voxels, iterations_, kind_, param_, norms_.data(),
                        flags_, stream));
    return true;
  }
  const std::vector<double> &convergence_data() const { return norms_; }

 private:
  double param_;
  int kind_;
  std::vector<double> norms_;
};

class cgls_tikhonov : public cgls_regularize {
 public:
  cgls_tikhonov(int iterations, double param, int flags = 0)
      : cgls_regularize(iterations, param, CCPI_REG_TIKHONOV, flags) {}
};

class cgls_tv_reg : public cgls_regularize {
 public:
  cgls_tv_reg(int iterations, double param, int flags = 0)
      : cgls_regularize(iterations, param, CCPI_REG_TV, flags) {}
};

class sirt : public reconstruction_alg {
 public:
  using reconstruction_alg::reconstruction_alg;
  bool reconstruct(const instrument &device, float *pixels, float *voxels,
                   const double origin[3], const double voxel_size[3], int nx, int ny,
                   int nz, void *stream = nullptr) override {
    const ccpi_geometry g = device.geometry(origin, voxel_size, nx, ny, nz);
    check(ccpi_sirt(&g, pixels, voxels, iterations_, flags_, stream));
    return true;
  }
};

class mlem : public reconstruction_alg {
 public:
  using reconstruction_alg::reconstruction_alg;
  bool reconstruct(const instrument &device, float *pixels, float *voxels,
                   const double origin[3], const double voxel_size[3], int nx, int ny,
                   int nz, void *stream = nullptr) override {
    const ccpi_geometry g = device.geometry(origin, voxel_size, nx, ny, nz);
    check(ccpi_mlem(&g, pixels, voxels, iterations_, flags_, stream));
    return true;
  }
};

}  // namespace hip
}  // namespace ccpi

#endif  // CCPI_HIP_HPP
