// HipModelSpecifics.h -- AbstractModelSpecifics implemented on the MI355X by
// calling libccdhip's C ABI.  This is the class the reference's factory would
// construct in its reserved GPU slot (src/cyclops/engine/AbstractModelSpecifics.cpp:32-35).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/ccdhip.h"
#include "AbstractModelSpecifics.h"

namespace ccdamd {

// Shared ownership of one resident data handle: clones (CV threads) reuse it.
struct SharedData {
  ccdh_data_t handle = nullptr;
  bool owned = false;
  ~SharedData() {
    if (owned && handle) ccdh_data_destroy(handle);
  }
};

class HipModelSpecifics : public AbstractModelSpecifics {
 public:
  HipModelSpecifics(std::shared_ptr<SharedData> data, int model, int K, int J);
  ~HipModelSpecifics() override;

  void initialize(int iN, int iK, int iJ, const void*, double*, double*, double*, double*,
                  std::vector<std::vector<int>*>*, const int*, double*, double*, double*,
                  double*, const double*) override;
  void setWeights(double* inWeights, double* cenWeights, bool useCrossValidation) override;
  void computeGradientAndHessian(int index, double* ogradient, double* ohessian,
                                 bool useWeights) override;
  void computeMMGradientAndHessian(std::vector<GradientHessian>&, const std::vector<bool>&,
                                   bool) override;
  void computeNumeratorForGradient(int index, bool useWeights) override;
  void computeFisherInformation(int, int, double*, bool) override;
  void updateXBeta(double realDelta, int index, bool useWeights) override;
  void computeXBeta(double* beta, bool useWeights) override;
  void computeRemainingStatistics(bool useWeights) override;
  void computeFixedTermsInLogLikelihood(bool useCrossValidation) override;
  void computeFixedTermsInGradientAndHessian(bool useCrossValidation) override;
  double getLogLikelihood(bool useCrossValidation) override;
  double getPredictiveLogLikelihood(double* weights) override;
  void getPredictiveEstimates(double* y, double* weights) override;
  double getGradientObjective(bool useCrossValidation) override;
  void printTiming() override {}
  AbstractModelSpecifics* clone() const override;
  const std::vector<double> getXBeta() override;
  const std::vector<double> getXBetaSave() override { return xBetaSave; }
  void saveXBeta() override { xBetaSave = getXBeta(); }
  void zeroXBeta() override;
  void axpyXBeta(const double beta, const int j) override;

  ccdh_fit_t handle() const { return fit; }

 private:
  static void check(int rc);
  std::shared_ptr<SharedData> data;
  int model, K, J;
  ccdh_fit_t fit = nullptr;
  std::vector<double> xBetaSave;
};

}  // namespace ccdamd
