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
//   UniModalSearch (utils/HParSearch.h:30-65, utils/HParSearch.cpp:22-120): the next
//     prior variance of auto-search cross-validation.
#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
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
  if (seed == -99)  // the selector's clock seed would differ from task to task
    return ccdamd::setError(CCDH_E_INVALID, "seed -99 (clock seed): pass an explicit seed");
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
  // Seed -1 makes the selector deterministic (AbstractSelector.cpp:32-48): the
  // permutation stays the identity (permute, CrossValidationSelector.cpp:140-144).
  std::mt19937 prng;
  prng.seed(seed);
  for (int i = 0; i <= task; ++i)
    if (i % n_fold == 0 && seed != -1) std::shuffle(perm.begin(), perm.end(), prng);
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
  if (seed == -99)
    return ccdamd::setError(CCDH_E_INVALID, "seed -99 (clock seed): pass an explicit seed");
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
  // Seed -1 leaves the engine default-constructed (AbstractSelector.cpp:32-48).
  std::mt19937 prng;
  if (seed != -1) prng.seed(seed);
  for (int sample = 0; sample <= replicate + 1; ++sample) {
    std::uniform_int_distribution<int> uniform(0, nIn - 1);
    std::fill(count.begin(), count.end(), 0);
    for (int i = 0; i < nIn; ++i) count[included[uniform(prng)]]++;
  }
  for (int k = 0; k < K; ++k) weights_out[k] = ids[k] < (int)count.size() ? count[ids[k]] : 0.0;
  return CCDH_OK;
}

namespace {

// Weighted least squares of y on (1, log x, log^2 x) with weights 1/stddev
// (100/smallest positive stddev where a stddev is 0; all 1 when none is
// positive), solved by Gaussian elimination with complete pivoting.
void quadrLogFit(const std::map<double, std::pair<double, double>>& pts, double c[3]) {
  double minSd = std::numeric_limits<double>::max();
  for (const auto& p : pts)
    if (0 < p.second.second && p.second.second < minSd) minSd = p.second.second;
  const double zeroAdjust = minSd < std::numeric_limits<double>::max() ? 100.0 / minSd : 1.0;
  double A[3][3] = {{0}}, b[3] = {0};
  for (const auto& p : pts) {
    const double w = p.second.second > 0 ? 1.0 / p.second.second : zeroAdjust;
    const double lx = std::log(p.first);
    const double row[3] = {1.0, lx, lx * lx};
    for (int r = 0; r < 3; ++r) {
      for (int q = 0; q < 3; ++q) A[r][q] += w * row[r] * row[q];
      b[r] += w * row[r] * p.second.first;
    }
  }
  int perm[3] = {0, 1, 2};  // column order after pivoting
  for (int k = 0; k < 3; ++k) {
    int pr = k, pc = k;
    for (int r = k; r < 3; ++r)
      for (int q = k; q < 3; ++q)
        if (std::fabs(A[r][q]) > std::fabs(A[pr][pc])) {
          pr = r;
          pc = q;
        }
    for (int q = 0; q < 3; ++q) std::swap(A[k][q], A[pr][q]);
    std::swap(b[k], b[pr]);
    for (int r = 0; r < 3; ++r) std::swap(A[r][k], A[r][pc]);
    std::swap(perm[k], perm[pc]);
    for (int r = k + 1; r < 3; ++r) {
      const double m = A[k][k] != 0.0 ? A[r][k] / A[k][k] : 0.0;
      for (int q = k; q < 3; ++q) A[r][q] -= m * A[k][q];
      b[r] -= m * b[k];
    }
  }
  double z[3];
  for (int k = 2; k >= 0; --k) {
    double t = b[k];
    for (int q = k + 1; q < 3; ++q) t -= A[k][q] * z[q];
    z[k] = A[k][k] != 0.0 ? t / A[k][k] : 0.0;
  }
  for (int k = 0; k < 3; ++k) c[perm[k]] = z[k];
}

}  // namespace

// The search keeps every tried (x, mean, stddev) in a map by x (a repeated x
// replaces its entry) and the best entry (the first, then any later one with a
// larger mean); with 0 points it proposes 1, with 1 it steps by stdstep (up
// when x < first_cut, else down), with 2 it steps beyond the better end, with
// 3+ it steps beyond an end that is best, else proposes the argmax of the
// quadratic fit in log x and stops when that gains less than stop_by_y
// (relative) or moves less than stop_by_x in log x.
extern "C" int ccdh_search_step(const double* xs, const double* ys, const double* sds, int n,
                                double stdstep, double stop_by_y, double stop_by_x,
                                double first_cut, int* cont, double* next, double* expected) {
  if (n < 0 || (n > 0 && (!xs || !ys)) || !cont || !next || !expected)
    return ccdamd::setError(CCDH_E_INVALID, "bad search arguments");
  std::map<double, std::pair<double, double>> pts;
  double bestX = 0.0;
  for (int i = 0; i < n; ++i) {
    pts[xs[i]] = {ys[i], sds ? sds[i] : 0.0};
    if (pts.size() == 1 || ys[i] > pts[bestX].first) bestX = xs[i];
  }
  *cont = 1;
  *next = 0.0;
  *expected = 0.0;
  if (pts.empty()) {
    *next = 1.0;
  } else if (pts.size() == 1) {
    const double x = pts.begin()->first;
    *next = x < first_cut ? x * stdstep : x / stdstep;
  } else if (pts.size() == 2) {
    if (pts.begin()->second.first > pts.rbegin()->second.first)
      *next = pts.begin()->first / stdstep;
    else
      *next = pts.rbegin()->first * stdstep;
  } else {
    const double bestY = pts[bestX].first;
    if (pts.begin()->first == bestX) {
      *next = bestX / stdstep;
      if (!(bestX > std::numeric_limits<double>::denorm_min())) {
        *cont = 0;
        *expected = bestY;
      }
    } else if (pts.rbegin()->first == bestX) {
      *next = bestX * stdstep;
      if (!(bestX < std::numeric_limits<double>::infinity())) {
        *cont = 0;
        *expected = bestY;
      }
    } else {
      double c[3];
      quadrLogFit(pts, c);
      const double logArgmax = -c[1] / c[2] / 2;
      const double expMax = -c[1] * c[1] / c[2] / 4 + c[0];
      if (bestY == 0.0) *cont = 0;
      else if ((expMax - bestY) / std::fabs(bestY) < stop_by_y) *cont = 0;
      else if (std::fabs(logArgmax - std::log(bestX)) < stop_by_x) *cont = 0;
      *next = std::exp(logArgmax);
      *expected = expMax;
    }
  }
  return CCDH_OK;
}

// QuadrLogFit's coefficients (c0 + c1 log x + c2 log^2 x), for tests and callers.
extern "C" int ccdh_search_quadratic(const double* xs, const double* ys, const double* sds, int n,
                                     double* c_out) {
  if (n < 3 || !xs || !ys || !c_out) return ccdamd::setError(CCDH_E_INVALID, "bad fit arguments");
  std::map<double, std::pair<double, double>> pts;
  for (int i = 0; i < n; ++i) pts[xs[i]] = {ys[i], sds ? sds[i] : 0.0};
  quadrLogFit(pts, c_out);
  return CCDH_OK;
}

extern "C" double ccdh_cv_grid_point(double lower, double upper, int size, int step) {
  if (size == 1) return upper;
  const double stepSize = (std::log(upper) - std::log(lower)) / (size - 1);
  return std::exp(std::log(lower) + step * stepSize);
}
