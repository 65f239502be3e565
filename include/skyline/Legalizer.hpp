// legalizer::Legalizer -- the reference's public C++ class
// (src/skyline/legalizer/include/legalizer/Legalizer.h:66-97) over the C ABI
// of skyline_lg.h.  Same setter names and meanings; the design comes from an
// in-memory LgDesign instead of db::dbDatabase, and run() reports a status
// code and fills an LgResult instead of printing and exiting.
#pragma once

#include <string>
#include <string_view>

#include "../skyline_lg.h"

namespace legalizer {

class Legalizer {
 public:
  explicit Legalizer(const LgDesign* design);

  int run(LgResult* result);  // Legalizer::run (Legalizer.cpp:197-228)

  // Legalizer.h:74-75: the ICCAD 2017 placement.constraints file (turns on
  // the contest metric in the statistics) and the modified-ISPD2015 cell
  // size file (needs LgDesign::inst_name); an empty string clears them
  void setICCAD17Constraint(const std::string& file);
  void setSizeByFile(const std::string& file);
  void setTechPenalty(float val);
  void setStandardAdmm(bool val);
  void setQpRefine(bool val);
  void setCoeffToInitRho(float val);
  void setCoeffToInitDual(float val);
  void setMaxAdmmIter(int iter);
  void setFlagICCAD2017(bool val);
  void setSortingAlgorithm(std::string_view algo);
  void setSortingPolicy(std::string_view policy);
  void setAdaptiveSearch(bool val);
  void setMinIterToCallAdaptiveSearch(int iter);
  void setOvfRatioToCallAdaptiveSearch(float ratio);
  void setCoeffToNeglectDisplace(float val);
  void setRandomSorting(bool val);
  void setRandomSortingSeed(int seed);

  const LgParams& params() const { return params_; }
  LgParams& params() { return params_; }

 private:
  const LgDesign* design_;
  LgParams params_;
  std::string constraint_, size_file_;
};

}  // namespace legalizer
