// AbstractModelSpecifics.h -- the engine interface the CCD optimizer calls,
// with the method names, argument meanings and call order of Cyclops'
// bsccs::AbstractModelSpecifics (src/cyclops/engine/AbstractModelSpecifics.h:36-229).
//
// This standalone build compiles HipModelSpecifics against this header; inside
// the reference tree the same class derives from the reference's own
// AbstractModelSpecifics instead (the signatures below are the reference's).
// Only the virtuals on the CCD hot path are declared; the MM and Fisher-
// information entry points are kept so the vtable shape matches and report
// "not supported" the way the reference's ModelSpecifics reports unsupported
// operations (by throwing).
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

namespace ccdamd {

typedef std::pair<double, double> GradientHessian;

class AbstractModelSpecifics {
 public:
  virtual ~AbstractModelSpecifics() {}

  // AbstractModelSpecifics.h:46-62 (the CCD passes N, K, J and null pointers)
  virtual void initialize(int iN, int iK, int iJ, const void* iXi, double* iNumerPid,
                          double* iNumerPid2, double* iDenomPid, double* iXjY,
                          std::vector<std::vector<int>*>* iSparseIndices, const int* iPid,
                          double* iOffsExpXBeta, double* iXBeta, double* iOffs, double* iBeta,
                          const double* iY) = 0;
  virtual void setWeights(double* inWeights, double* cenWeights, bool useCrossValidation) = 0;
  virtual void computeGradientAndHessian(int index, double* ogradient, double* ohessian,
                                         bool useWeights) = 0;
  virtual void computeMMGradientAndHessian(std::vector<GradientHessian>& gh,
                                           const std::vector<bool>& fixBeta, bool useWeights) = 0;
  virtual void computeNumeratorForGradient(int index, bool useWeights) = 0;
  virtual void computeFisherInformation(int indexOne, int indexTwo, double* oinfo,
                                        bool useWeights) = 0;
  virtual void updateXBeta(double realDelta, int index, bool useWeights) = 0;
  virtual void computeXBeta(double* beta, bool useWeights) = 0;
  virtual void computeRemainingStatistics(bool useWeights) = 0;
  virtual void computeFixedTermsInLogLikelihood(bool useCrossValidation) = 0;
  virtual void computeFixedTermsInGradientAndHessian(bool useCrossValidation) = 0;
  virtual double getLogLikelihood(bool useCrossValidation) = 0;
  virtual double getPredictiveLogLikelihood(double* weights) = 0;
  virtual void getPredictiveEstimates(double* y, double* weights) = 0;
  virtual double getGradientObjective(bool useCrossValidation) = 0;
  virtual void makeDirty() {}
  virtual void printTiming() = 0;
  virtual AbstractModelSpecifics* clone() const = 0;
  virtual const std::vector<double> getXBeta() = 0;
  virtual const std::vector<double> getXBetaSave() = 0;
  virtual void saveXBeta() = 0;
  virtual void zeroXBeta() = 0;
  virtual void axpyXBeta(const double beta, const int j) = 0;
};

}  // namespace ccdamd
