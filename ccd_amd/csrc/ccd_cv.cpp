// ccd_cv.cpp -- cross-validation fold weights and the variance grid, as the
// batched engine's host-side batch builder uses them.
//   CrossValidationSelector (drivers/CrossValidationSelector.cpp:21-116,138-143):
//     fold sizes N/fold (+1 for the first N%fold folds), one std::shuffle of the
//     unit permutation per repetition with std::mt19937 seeded by `seed`
//     (reseed per grid point, :60-66), held-out units get weight 0;
//     BY_PID holds out whole strata, BY_ROW single rows.
//   GridSearchCrossValidationDriver::computeGridPoint (GridSearchCrossValidationDriver.cpp:38-48):
//     log-uniform grid between lower and upper.
//   BootstrapSelector (drivers/BootstrapSelector.cpp:17-95) with BootstrapDriver::drive
//     (drivers/BootstrapDriver.cpp:47-71): resampling weights of bootstrap replicates.
#include <algorithm>
#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "../../include/ccdhip.h"

namespace ccdamd {
int setError(int code, const std::string& msg);
}

extern "C" int ccdh_cv_fold_weights(const int32_t* ids, int K, int by_pid, int n_fold, long seed,
                                    int task, double* weights_out) {
  if (!ids || !weights_out || K <= 0 || n_fold <= 0 || task < 0)
    return ccdamd::setError(CCDH_E_INVALID, "bad cross-validation arguments");
  int N = K;
  if (by_pid) {
    N = 0;
    for (int k = 0; k < K; ++k) {
      if (ids[k] < 0) return ccdamd::setError(CCDH_E_INVALID, "negative id");
      N = std::max(N, ids[k] + 1);
    }
  }
  std::vector<int> start(n_fold + 1);
  const int fraction = N / n_fold, extra = N - fraction * n_fold;
  int index = 0;
  for (int i = 0; i < n_fold; ++i) {
    start[i] = index;
    index += fraction + (i < extra ? 1 : 0);
  }
  start[n_fold] = N;
  std::vector<int> perm(N);
  for (int i = 0; i < N; ++i) perm[i] = i;
  std::mt19937 prng;
  prng.seed(seed);
  for (int i = 0; i <= task; ++i)
    if (i % n_fold == 0) std::shuffle(perm.begin(), perm.end(), prng);
  const int fold = task % n_fold;
  std::fill(weights_out, weights_out + K, 1.0);
  if (by_pid) {
    std::vector<char> held(N, 0);
    for (int p = start[fold]; p < start[fold + 1]; ++p) held[perm[p]] = 1;
    for (int k = 0; k < K; ++k)
      if (held[ids[k]]) weights_out[k] = 0.0;
  } else {
    for (int p = start[fold]; p < start[fold + 1]; ++p) weights_out[perm[p]] = 0.0;
  }
  return CCDH_OK;
}

// A sample draws as many pids as are included, uniformly with replacement
// (std::uniform_int_distribution over the included pids, std::mt19937 seeded
// by `seed`); a row weighs the number of times its pid was drawn.  The
// selector's constructor draws one sample and the driver one more before each
// replicate, so replicate r is the (r+2)-th sample of the stream.  Included
// pids: those i < n_exclude with exclude[i] == 0, or every pid without exclude.
extern "C" int ccdh_bootstrap_weights(const int32_t* ids, int K, long seed, int replicate,
                                      const double* exclude, int n_exclude, double* weights_out) {
  if (!ids || !weights_out || K <= 0 || replicate < 0 || (exclude && n_exclude < 0))
    return ccdamd::setError(CCDH_E_INVALID, "bad bootstrap arguments");
  int N = 0;
  for (int k = 0; k < K; ++k) {
    if (ids[k] < 0) return ccdamd::setError(CCDH_E_INVALID, "negative id");
    N = std::max(N, ids[k] + 1);
  }
  std::vector<int> included;
  if (exclude) {
    for (int i = 0; i < n_exclude; ++i)
      if (exclude[i] == 0.0) included.push_back(i);
  } else {
    included.resize(N);
    for (int i = 0; i < N; ++i) included[i] = i;
  }
  if (included.empty()) return ccdamd::setError(CCDH_E_INVALID, "no pid to draw");
  const int nIn = (int)included.size();
  std::vector<int> count(std::max(N, n_exclude), 0);
  std::mt19937 prng;
  prng.seed(seed);
  for (int sample = 0; sample <= replicate + 1; ++sample) {
    std::uniform_int_distribution<int> uniform(0, nIn - 1);
    std::fill(count.begin(), count.end(), 0);
    for (int i = 0; i < nIn; ++i) count[included[uniform(prng)]]++;
  }
  for (int k = 0; k < K; ++k) weights_out[k] = ids[k] < (int)count.size() ? count[ids[k]] : 0.0;
  return CCDH_OK;
}

extern "C" double ccdh_cv_grid_point(double lower, double upper, int size, int step) {
  if (size == 1) return upper;
  const double stepSize = (std::log(upper) - std::log(lower)) / (size - 1);
  return std::exp(std::log(lower) + step * stepSize);
}
