// skyplace::DensityGradient and skyplace::PoissonSolver -- the reference's
// C++ classes (skyplace/src/density_gradient/DensityGradient.h:23-152,
// PoissonSolver.h:9-68) over the C ABI of skyline_density.h.  Same method
// names and device-pointer arguments; the geometry comes from a DensGeom
// (see dens_prepare) instead of std::shared_ptr<SkyPlaceDB>, and failures
// throw std::runtime_error instead of asserting.
#pragma once

#include <stdexcept>
#include <string>

#include "../skyline_density.h"

namespace skyplace {

inline void dens_check(int status) {
  if (status != DENS_OK) throw std::runtime_error(std::string("density: ") + dens_last_error());
}

class PoissonSolver {
 public:
  PoissonSolver(int numBinX, int numBinY) { dens_check(poisson_create(numBinX, numBinY, &h_)); }
  ~PoissonSolver() { poisson_destroy(h_); }
  PoissonSolver(const PoissonSolver&) = delete;
  PoissonSolver& operator=(const PoissonSolver&) = delete;

  // PoissonSolver::solvePoisson (PoissonSolver.cu:224-302); device arrays
  void solvePoisson(const float* binDensity, float* potential, float* electroForceX, float* electroForceY,
                    int precision = DENS_PREC_SPLIT_BF16, void* stream = nullptr) {
    dens_check(poisson_solve(h_, binDensity, potential, electroForceX, electroForceY, precision, stream));
  }

 private:
  PoissonHandle* h_ = nullptr;
};

class DensityGradient {
 public:
  explicit DensityGradient(const DensGeom& geom) : nbx_(geom.num_bin_x), nby_(geom.num_bin_y) {
    dens_check(dens_create(&geom, &h_));
  }
  ~DensityGradient() { dens_destroy(h_); }
  DensityGradient(const DensityGradient&) = delete;
  DensityGradient& operator=(const DensityGradient&) = delete;

  // DensityGradient::computeGrad (DensityGradient.cu:302-403); device arrays
  void computeGrad(float* densityGradX, float* densityGradY, const float* cellCx, const float* cellCy,
                   void* stream = nullptr) {
    dens_check(dens_compute_grad(h_, cellCx, cellCy, densityGradX, densityGradY, nullptr, stream));
  }

  int numBinX() const { return nbx_; }
  int numBinY() const { return nby_; }
  float overflow() const {
    float v = 0.f;
    dens_check(dens_get_overflow(h_, &v));
    return v;
  }
  const float* getDevicePotential() const { return dens_get_potential(h_); }
  const float* getDeviceBinDensity() const { return dens_get_bin_density(h_); }
  const float* getDevicePreconditioner() const { return dens_get_preconditioner(h_); }

 private:
  DensHandle* h_ = nullptr;
  int nbx_, nby_;
};

}  // namespace skyplace
