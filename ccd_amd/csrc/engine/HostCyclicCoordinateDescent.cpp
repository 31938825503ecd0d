// HostCyclicCoordinateDescent.cpp -- the host CCD optimizer for single fits
// (the "plain" GPU engine).  It drives any AbstractModelSpecifics one
// coordinate at a time exactly as Cyclops' CyclicCoordinateDescent does:
//   update()  ............ CyclicCoordinateDescent.cpp:586-615 (ILLCONDITIONED retry, bound/10)
//   findMode() ........... :909-1216 (lazy statistics, resetBounds, cycle, check)
//   ccdUpdateBeta() ...... :1394-1413 (H < 0 -> (0, 0), prior step)
//   applyBounds() ........ :1503-1520
//   performCheckConvergence() :839-906 with criterion :1498-1501
//   getNewPredictiveLogLikelihood() :319-355
//   MM algorithm: cycle :974-1005, check :1026-1048, mmUpdateAllBeta() :1354-1391
// With HipModelSpecifics underneath, every coordinate is one gradient/Hessian
// launch, one 16-byte device-to-host copy for the host prior step, and one
// update launch: the reference's per-coordinate host round trip, kept on purpose
// as the single-fit baseline and the drop-in path.
#include <chrono>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../ccd_device.h"
#include "HipModelSpecifics.h"

namespace ccdamd {

int setError(int code, const std::string& msg);

// The MM check's error (CyclicCoordinateDescent.cpp:1036-1042)
struct NonIncreasing : std::runtime_error {
  NonIncreasing() : std::runtime_error("Non-increasing!") {}
};

class HostCyclicCoordinateDescent {
 public:
  HostCyclicCoordinateDescent(AbstractModelSpecifics& ms, int J, bool hasOffset,
                              const ccdh_fit_spec& spec, int K, bool mm = false)
      : ms(ms), J(J), K(K), hasOffset(hasOffset), mm(mm), spec(spec) {
    beta.assign(spec.beta0, spec.beta0 + J);
    fixBeta.assign(J, false);
    for (int j = 0; j < J; ++j) fixBeta[j] = spec.fixed[j] != 0;
    bound.assign(J, spec.initial_bound);
    useCV = spec.weights != nullptr;
    if (useCV) weights.assign(spec.weights, spec.weights + K);
  }

  void update() {
    initialBound = spec.initial_bound;
    int count = 0;
    bool done = false;
    while (!done) {
      findMode();
      ++count;
      if (lastReturnFlag == CCDH_ILLCONDITIONED && count < spec.max_bound_count) {
        initialBound /= 10.0;
        resetBeta();
      } else {
        done = true;
      }
    }
    boundCount = count;
  }

  double logLikelihood() {
    checkAllLazyFlags();
    return ms.getLogLikelihood(useCV);
  }

  double logPrior() const {
    double s = 0.0;
    for (int j = 0; j < J; ++j) s += priorLogDensity(spec.prior_type[j], spec.prior_var[j], beta[j]);
    return s;
  }

  double predictiveLogLikelihood(const double* w) {
    computeXBeta();
    ms.computeRemainingStatistics(useCV);
    sufficientStatisticsKnown = true;
    return ms.getPredictiveLogLikelihood(const_cast<double*>(w));
  }

  std::vector<double> beta;
  int lastIterationCount = 0, lastReturnFlag = CCDH_SUCCESS, boundCount = 0;
  int iterationReached = 0;  // the iteration an MM fit stopped at with NonIncreasing

 private:
  void computeXBeta() {
    ms.zeroXBeta();
    for (int j = 0; j < J; ++j) ms.axpyXBeta(beta[j], j);
  }

  void resetBeta() {
    for (int j = hasOffset ? 1 : 0; j < J; ++j) beta[j] = 0.0;
    computeXBeta();
    sufficientStatisticsKnown = false;
  }

  void computeNEventsEtc() {
    ms.setWeights(useCV ? weights.data() : nullptr, nullptr, useCV);
    ms.computeFixedTermsInLogLikelihood(useCV);
    ms.computeFixedTermsInGradientAndHessian(useCV);
    validWeights = true;
  }

  void checkAllLazyFlags() {
    if (!xBetaKnown) {
      computeXBeta();
      xBetaKnown = true;
      sufficientStatisticsKnown = false;
    }
    if (!validWeights) computeNEventsEtc();
    if (!sufficientStatisticsKnown) {
      ms.computeRemainingStatistics(useCV);
      sufficientStatisticsKnown = true;
    }
  }

  double objective() {
    switch (spec.convergence_type) {
      case CCDH_CONV_GRADIENT:
        return ms.getGradientObjective(useCV);
      case CCDH_CONV_MITTAL:
        return logLikelihood();
      case CCDH_CONV_LANGE:
        return logLikelihood() + logPrior();
      case CCDH_CONV_ONE_STEP:
        return 0.0;
    }
    throw std::invalid_argument("Invalid convergence type");
  }

  void findMode() {
    if (!validWeights) computeNEventsEtc();
    if (!xBetaKnown) {
      computeXBeta();
      xBetaKnown = true;
      sufficientStatisticsKnown = false;
    }
    if (!sufficientStatisticsKnown) {
      ms.computeRemainingStatistics(useCV);
      sufficientStatisticsKnown = true;
    }
    for (int j = 0; j < J; ++j) bound[j] = initialBound;
    int iteration = 0;
    double lastObj = objective();
    double lastLogPosterior = 0.0;
    bool done = false;
    while (!done) {
      if (mm) {
        mmUpdateAllBeta();
        for (int j = 0; j < J; ++j)
          if (!fixBeta[j] && allDelta[j] != 0.0) beta[j] += allDelta[j];
        ms.computeXBeta(beta.data(), useCV);
        ms.computeRemainingStatistics(useCV);
        sufficientStatisticsKnown = true;
        iteration++;
        const double logPosterior = logLikelihood() + logPrior();
        if (iteration > 1 && logPosterior - lastLogPosterior < 0.0) {
          iterationReached = iteration;
          throw NonIncreasing();
        }
        lastLogPosterior = logPosterior;
      }
      for (int j = 0; !mm && j < J; ++j) {
        if (fixBeta[j]) continue;
        // ccdUpdateBeta
        ms.computeNumeratorForGradient(j, useCV);
        double g = 0.0, h = 0.0;
        ms.computeGradientAndHessian(j, &g, &h, useCV);
        if (h < 0.0) {
          g = 0.0;
          h = 0.0;
        }
        double delta = priorDelta(spec.prior_type[j], spec.prior_var[j], g, h, beta[j]);
        delta = applyBounds(delta, &bound[j]);
        if (delta != 0.0) {
          beta[j] += delta;
          ms.updateXBeta(delta, j, useCV);
        }
      }
      if (!mm) iteration++;
      // performCheckConvergence
      double conv;
      bool ill = false;
      const double obj = objective();
      if (obj != obj) {
        conv = 0.0;
        ill = true;
      } else {
        conv = std::fabs(obj - lastObj) / (std::fabs(obj) + 1.0);
      }
      lastObj = obj;
      if (spec.tolerance > 0 && conv < spec.tolerance) {
        lastReturnFlag = ill ? CCDH_ILLCONDITIONED : CCDH_SUCCESS;
        done = true;
      } else if (iteration == spec.max_iterations) {
        lastReturnFlag = CCDH_MAX_ITERATIONS;
        done = true;
      }
    }
    lastIterationCount = iteration;
  }

  // every free coordinate's prior step from one set of statistics; no bounds
  void mmUpdateAllBeta() {
    ms.computeMMGradientAndHessian(gh, fixBeta, useCV);
    allDelta.assign(J, 0.0);
    for (int j = 0; j < J; ++j) {
      if (fixBeta[j]) continue;
      double g = gh[j].first, h = gh[j].second;
      if (h < 0.0) {
        g = 0.0;
        h = 0.0;
      }
      allDelta[j] = priorDelta(spec.prior_type[j], spec.prior_var[j], g, h, beta[j]);
    }
  }

  AbstractModelSpecifics& ms;
  int J, K;
  bool hasOffset, mm;
  std::vector<GradientHessian> gh;
  std::vector<double> allDelta;
  const ccdh_fit_spec& spec;
  std::vector<bool> fixBeta;
  std::vector<double> bound, weights;
  double initialBound = 2.0;
  bool useCV = false, validWeights = false, xBetaKnown = false, sufficientStatisticsKnown = false;
};

}  // namespace ccdamd

using namespace ccdamd;

static thread_local ccdh_mm_stats g_mmStats{};  // the last MM fit on this host thread

static int fitHost(ccdh_data_t data, const ccdh_fit_spec* spec, double* beta_out,
                   ccdh_fit_result* result, bool mm) {
  if (!data || !spec || !beta_out || !result) return setError(CCDH_E_INVALID, "null argument");
  if (!spec->prior_type || !spec->prior_var || !spec->fixed || !spec->beta0)
    return setError(CCDH_E_INVALID, "fit spec arrays required");
  int model = 0, K = 0, J = 0, hasOffset = 0;
  ccdh_data_shape(data, &model, &K, &J, &hasOffset);
  try {
    auto t0 = std::chrono::steady_clock::now();
    auto shared = std::make_shared<SharedData>();
    shared->handle = data;
    HipModelSpecifics ms(shared, model, K, J);
    HostCyclicCoordinateDescent ccd(ms, J, hasOffset != 0, *spec, K, mm);
    auto keepStats = [&]() {
      ccdh_mm_stats st{};
      if (mm && ccdh_fit_mm_stats(ms.handle(), &st) == CCDH_OK) g_mmStats = st;
    };
    try {
      ccd.update();
    } catch (const NonIncreasing& e) {
      for (int j = 0; j < J; ++j) beta_out[j] = ccd.beta[j];
      result->iterations = ccd.iterationReached;
      keepStats();
      return setError(CCDH_E_FIT, e.what());
    }
    keepStats();
    for (int j = 0; j < J; ++j) beta_out[j] = ccd.beta[j];
    result->log_likelihood = ccd.logLikelihood();
    result->log_prior = ccd.logPrior();
    result->iterations = ccd.lastIterationCount;
    result->return_flag = ccd.lastReturnFlag;
    result->bound_count = ccd.boundCount;
    result->pred_log_likelihood = 0.0;
    if (spec->heldout) result->pred_log_likelihood = ccd.predictiveLogLikelihood(spec->heldout);
    auto t1 = std::chrono::steady_clock::now();
    result->seconds = std::chrono::duration<double>(t1 - t0).count();
  } catch (const std::exception& e) {
    return setError(CCDH_E_DEVICE, e.what());
  }
  return CCDH_OK;
}

extern "C" int ccdh_fit_plain(ccdh_data_t data, const ccdh_fit_spec* spec, double* beta_out,
                              ccdh_fit_result* result) {
  return fitHost(data, spec, beta_out, result, false);
}

extern "C" int ccdh_fit_mm(ccdh_data_t data, const ccdh_fit_spec* spec, double* beta_out,
                           ccdh_fit_result* result) {
  if (data) {
    int model = 0, K = 0, J = 0, hasOffset = 0;
    ccdh_data_shape(data, &model, &K, &J, &hasOffset);
    if (model == CCDH_POISSON)  // the reference throws (ModelSpecifics.h:1546-1552)
      return setError(CCDH_E_INVALID, "MM algorithm: not model-specific (Poisson)");
    // FP32 row state puts ~1e-7 relative noise into the log posterior, which trips
    // the strict non-increasing check long before the fit converges
    if (ccdh_data_precision(data) != CCDH_FP64)
      return setError(CCDH_E_INVALID, "MM algorithm: FP64 data only");
  }
  return fitHost(data, spec, beta_out, result, true);
}

extern "C" int ccdh_mm_last_stats(ccdh_mm_stats* out) {
  if (!out) return setError(CCDH_E_INVALID, "null argument");
  *out = g_mmStats;
  return CCDH_OK;
}
