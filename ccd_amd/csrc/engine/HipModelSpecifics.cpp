// HipModelSpecifics.cpp -- every virtual forwards to one libccdhip C ABI call.
// Status codes become exceptions here (the reference's ModelSpecifics throws for
// unsupported operations, ModelSpecifics.h:417,1105,1211; the caller maps them to
// ErrorHandler::throwError).
#include "HipModelSpecifics.h"

#include <cmath>

namespace ccdamd {

void HipModelSpecifics::check(int rc) {
  if (rc != CCDH_OK) throw std::runtime_error(std::string("ccdhip: ") + ccdh_last_error());
}

HipModelSpecifics::HipModelSpecifics(std::shared_ptr<SharedData> d, int m, int k, int j)
    : data(std::move(d)), model(m), K(k), J(j) {
  check(ccdh_fit_create(data->handle, &fit));
}

HipModelSpecifics::~HipModelSpecifics() {
  if (fit) ccdh_fit_destroy(fit);
}

void HipModelSpecifics::initialize(int, int, int, const void*, double*, double*, double*,
                                   double*, std::vector<std::vector<int>*>*, const int*, double*,
                                   double*, double*, double*, const double*) {
  // sizes and the Cox risk-set grouping are set up by ccdh_fit_create
}

void HipModelSpecifics::setWeights(double* inWeights, double*, bool useCrossValidation) {
  check(ccdh_set_weights(fit, useCrossValidation ? inWeights : nullptr, useCrossValidation));
}

void HipModelSpecifics::computeGradientAndHessian(int index, double* ogradient,
                                                  double* ohessian, bool useWeights) {
  check(ccdh_grad_hess(fit, index, useWeights, ogradient, ohessian));
}

void HipModelSpecifics::computeMMGradientAndHessian(std::vector<GradientHessian>& gh,
                                                    const std::vector<bool>& fixBeta,
                                                    bool useWeights) {
  std::vector<uint8_t> fixed(J);
  for (int j = 0; j < J; ++j) fixed[j] = fixBeta[j] ? 1 : 0;
  std::vector<double> flat(2 * (size_t)J);
  check(ccdh_mm_grad_hess(fit, fixed.data(), useWeights, flat.data()));
  gh.resize(J);
  for (int j = 0; j < J; ++j) gh[j] = GradientHessian(flat[2 * j], flat[2 * j + 1]);
}

void HipModelSpecifics::computeNumeratorForGradient(int index, bool useWeights) {
  check(ccdh_numerator(fit, index, useWeights));
}

void HipModelSpecifics::computeFisherInformation(int, int, double*, bool) {
  throw std::logic_error("Fisher information is not implemented by the HIP engine");
}

void HipModelSpecifics::updateXBeta(double realDelta, int index, bool useWeights) {
  check(ccdh_update_xbeta(fit, realDelta, index, useWeights));
}

void HipModelSpecifics::computeXBeta(double* beta, bool) {
  check(ccdh_compute_xbeta(fit, beta));
}

void HipModelSpecifics::computeRemainingStatistics(bool useWeights) {
  check(ccdh_remaining_stats(fit, useWeights));
}

void HipModelSpecifics::computeFixedTermsInLogLikelihood(bool useCrossValidation) {
  check(ccdh_fixed_terms_loglik(fit, useCrossValidation));
}

void HipModelSpecifics::computeFixedTermsInGradientAndHessian(bool useCrossValidation) {
  check(ccdh_fixed_terms_grad_hess(fit, useCrossValidation));
}

double HipModelSpecifics::getLogLikelihood(bool useCrossValidation) {
  double v = 0.0;
  check(ccdh_loglik(fit, useCrossValidation, &v));
  return v;
}

double HipModelSpecifics::getPredictiveLogLikelihood(double* weights) {
  double v = 0.0;
  check(ccdh_pred_loglik(fit, weights, &v));
  return v;
}

// predictEstimate per model (ModelSpecifics.h:1011-1014 LR, 1592-1594 Poisson)
void HipModelSpecifics::getPredictiveEstimates(double* y, double* weights) {
  const std::vector<double> xb = getXBeta();
  for (int k = 0; k < K; ++k) {
    if (weights && !weights[k]) continue;
    double v = xb[k];
    if (model == CCDH_LOGISTIC) {
      const double t = std::exp(v);
      v = t / (t + 1.0);
    } else if (model == CCDH_POISSON) {
      v = std::exp(v);
    }
    y[k] = v;
  }
}

double HipModelSpecifics::getGradientObjective(bool useCrossValidation) {
  double v = 0.0;
  check(ccdh_gradient_objective(fit, useCrossValidation, &v));
  return v;
}

AbstractModelSpecifics* HipModelSpecifics::clone() const {
  return new HipModelSpecifics(data, model, K, J);
}

const std::vector<double> HipModelSpecifics::getXBeta() {
  std::vector<double> xb(K);
  check(ccdh_get_xbeta(fit, xb.data()));
  return xb;
}

void HipModelSpecifics::zeroXBeta() { check(ccdh_zero_xbeta(fit)); }

void HipModelSpecifics::axpyXBeta(const double beta, const int j) {
  check(ccdh_axpy_xbeta(fit, beta, j));
}

}  // namespace ccdamd
