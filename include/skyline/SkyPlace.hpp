// skyplace::SkyPlace -- the reference's public C++ class of the global placer
// (src/skyline/skyplace/include/skyplace/SkyPlace.h:19-85) over the C ABI of
// skyline_gp.h.  Same setter names and meanings; the design comes from an
// in-memory LgDesign instead of db::dbDatabase, and run() writes the placed
// lower-left corners into caller arrays (the reference writes them back into
// its database with exportDB) and returns a status code instead of exiting.
#pragma once

#include <cstdint>
#include <string>

#include "../skyline_gp.h"

namespace skyplace {

class SkyPlace {
 public:
  explicit SkyPlace(const LgDesign* design);

  // SkyPlace::run (SkyPlace.cpp:183-213): out_lx / out_ly [num_insts]
  int run(int32_t* out_lx, int32_t* out_ly);

  void setPlotMode(const std::string& plot_path);  // accepted; no plots are drawn
  void setSolver(const std::string& solver);        // "Nesterov" / "nesterov" / "Adam" / "adam"
  void setInitialPlaceMethod(const std::string& method);  // "random" ("sdp" is not built: warning, random)
  void setInitialPlaceOnly();

  void setTargetOverflow(float val);
  void setInitLambda(float val);
  void setInitGammaInv(float val);
  void setMaxPhiCoef(float val);
  void setRefHpwl(float val);
  void setAdamAlpha(float val);
  void setAdamBeta1(float val);
  void setAdamBeta2(float val);
  void setTargetDensity(float density);  // out of (0, 1] -> 1.0, as the reference

  const GpParams& params() const { return params_; }
  GpParams& params() { return params_; }
  const GpStats& stats() const { return stats_; }

 private:
  const LgDesign* design_;
  GpParams params_;
  GpStats stats_{};
};

}  // namespace skyplace
