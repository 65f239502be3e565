// skyplace::SkyPlace (include/skyline/SkyPlace.hpp) over the C ABI of
// skyline_gp.h: the setters of SkyPlace.cpp:35-181 and run().
#include <cstdio>

#include "../../include/skyline/SkyPlace.hpp"

namespace skyplace {

SkyPlace::SkyPlace(const LgDesign* design) : design_(design) { gp_default_params(&params_); }

int SkyPlace::run(int32_t* out_lx, int32_t* out_ly) {
  return gp_place_design(design_, &params_, out_lx, out_ly, &stats_, nullptr);
}

void SkyPlace::setPlotMode(const std::string&) {}

void SkyPlace::setSolver(const std::string& s) {
  if (s == "Adam" || s == "adam")
    params_.solver = GP_SOLVER_ADAM;
  else if (s == "Nesterov" || s == "nesterov")
    params_.solver = GP_SOLVER_NESTEROV;
  else
    std::printf("Undefined SolverType %s\nArgument is ignored and Nesterov will be used...\n", s.c_str());
}

void SkyPlace::setInitialPlaceMethod(const std::string& m) {
  if (m != "random")
    std::printf("[WARN] Initialization method %s is not available\n"
                "[WARN] Argument is ignored and random initialization will be used...\n", m.c_str());
}

void SkyPlace::setInitialPlaceOnly() { params_.initial_place_only = 1; }
void SkyPlace::setTargetOverflow(float v) { params_.target_overflow = v; }
void SkyPlace::setInitLambda(float v) { params_.init_lambda = v; }
void SkyPlace::setInitGammaInv(float v) { params_.init_gamma_inv_coef = v; }
void SkyPlace::setMaxPhiCoef(float v) { params_.max_phi_coef = v; }
void SkyPlace::setRefHpwl(float v) { params_.reference_hpwl = v; }
void SkyPlace::setAdamAlpha(float v) { params_.adam_alpha = v; }
void SkyPlace::setAdamBeta1(float v) { params_.adam_beta1 = v; }
void SkyPlace::setAdamBeta2(float v) { params_.adam_beta2 = v; }

// SkyPlace::setTargetDensity (SkyPlace.cpp:98-114)
void SkyPlace::setTargetDensity(float density) {
  const float pct = density * 100.0f;
  if (pct > 100.0f || pct <= 0.0f) {
    std::printf("Invalid target density (%4.2f%%). %4.2f%% will be used.\n", pct, 100.0f);
    params_.target_density = 1.0f;
  } else {
    params_.target_density = density;
  }
}

}  // namespace skyplace
