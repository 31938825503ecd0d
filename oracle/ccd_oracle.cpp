// ccd_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A clean-room FP64 CPU restatement of the Cyclops cyclic-coordinate-descent
// (CCD) hot path, used as the parity checker for the HIP engine and as the
// timed CPU baseline.  It is never linked into, loaded by or called from the
// product library (ccd_amd/libccdhip.so); only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg use it.
//
// Reference behaviour restated (file:line into the Cyclops sources):
//   findMode / cycle / check ........ src/cyclops/CyclicCoordinateDescent.cpp:909-1216
//   update (ill-conditioned retry) .. CyclicCoordinateDescent.cpp:586-615
//   ccdUpdateBeta (H<0 -> 0) ........ CyclicCoordinateDescent.cpp:1394-1413
//   applyBounds ..................... CyclicCoordinateDescent.cpp:1503-1520
//   performCheckConvergence ......... CyclicCoordinateDescent.cpp:839-906, criterion :1498-1501
//   computeXBeta (column axpy) ...... CyclicCoordinateDescent.cpp:1452-1470; engine/ModelSpecifics.hpp:393-423
//   setWeights ...................... engine/ModelSpecifics.hpp:427-484
//   computeXjY ...................... engine/ModelSpecifics.hpp:486-524
//   fixed LL terms (Poisson) ........ engine/ModelSpecifics.hpp:570-587; ModelSpecifics.h:1596-1601
//   getLogLikelihood ................ engine/ModelSpecifics.hpp:605-650
//   getPredictiveLogLikelihood ...... engine/ModelSpecifics.hpp:652-690
//   computeGradientAndHessianImpl ... engine/ModelSpecifics.hpp:891-1201
//   computeNumeratorForGradient ..... engine/ModelSpecifics.hpp:1353-1464
//   updateXBetaImpl ................. engine/ModelSpecifics.hpp:1524-1586
//   computeRemainingStatisticsImpl .. engine/ModelSpecifics.hpp:1588-1612
//   computeAccumlatedDenominator .... engine/ModelSpecifics.hpp:1670-1694
//   setPidForAccumulation ........... engine/ModelSpecifics.hpp:1809-1868
//   setupSparseIndices .............. engine/ModelSpecifics.hpp:1870-1892
//   model increments ................ engine/ModelSpecifics.h:540-568 (LR), 820-837 (CLR),
//                                     1185-1202 (Breslow Cox), 1508-1543 (Poisson)
//   priors .......................... priors/CovariatePrior.h:186-188, 221-292, 404-415
//   MM algorithm .................... CyclicCoordinateDescent.cpp:958-1005 (cycle), 1026-1048
//                                     (check), 1354-1391 (mmUpdateAllBeta);
//                                     engine/ModelSpecifics.hpp:776-889 (MM gradient/Hessian),
//                                     313-364 (computeXBeta over the transpose);
//                                     CompressedDataMatrix.cpp:151-208 (transpose)
//   CV fold weights ................. drivers/CrossValidationSelector.cpp:21-116, 138-143
//   grid points ..................... drivers/GridSearchCrossValidationDriver.cpp:38-48
//
// Parity pins for each part are listed in the tests (tests/test_oracle.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <set>
#include <map>
#include <vector>

namespace {

enum Fmt { DENSE = 0, SPARSE = 1, INDICATOR = 2, INTERCEPT = 3 };
enum Model { M_POISSON = 2, M_LOGISTIC = 3, M_CLR = 4, M_COX = 8 };
enum Prior { P_NONE = 0, P_LAPLACE = 1, P_NORMAL = 2 };
enum Conv { C_GRADIENT = 0, C_LANGE = 1, C_MITTAL = 2, C_ONE_STEP = 3 };
enum Flag { SUCCESS = 0, FAIL = 1, MAX_ITERATIONS = 2, ILLCONDITIONED = 3 };

struct Data {
  int K, J;
  const int64_t* colPtr;
  const int32_t* rows;
  const double* vals;
  const int8_t* fmt;
  const double* y;
  const double* offs;     // event time for Cox; unused otherwise
  const int32_t* strata;  // stratum per row (CLR, Cox); may be null
  int model;
};

inline bool isIndicatorFmt(int f) { return f == INDICATOR || f == INTERCEPT; }
inline bool isSparseFmt(int f) { return f == INDICATOR || f == SPARSE; }

// One model's sufficient statistics, restating ModelSpecifics<Model,double>.
struct Engine {
  const Data& d;
  int K, J, N;
  bool grouped, cox;
  std::vector<double> xb, oeb, denom, numer, numer2, accDenom;
  std::vector<double> kw, nw, xjy;
  std::vector<int> pidOrig, pid, accReset;
  std::vector<std::vector<int>> sparseIdx;  // empty + flag for dense
  std::vector<char> sparseIsNull;
  double fixedLL = 0.0;

  explicit Engine(const Data& data) : d(data), K(data.K), J(data.J) {
    grouped = d.model == M_CLR;
    cox = d.model == M_COX;
    pidOrig.resize(K);
    if (grouped || cox) {
      for (int k = 0; k < K; ++k) pidOrig[k] = d.strata ? d.strata[k] : 0;
    } else {
      for (int k = 0; k < K; ++k) pidOrig[k] = k;
    }
    if (grouped) {
      N = 0;
      for (int k = 0; k < K; ++k) N = std::max(N, pidOrig[k] + 1);
    } else {
      N = K;
    }
    pid = pidOrig;
    xb.assign(K, 0.0);
    oeb.assign(K, 0.0);
    xjy.assign(J, 0.0);
    if (cox) {
      setPidForAccumulation(nullptr);
    } else {
      setupSparseIndices(N);
    }
    size_t aligned = (size_t)((N + 1 + 15) / 16) * 16;
    denom.assign(aligned, 0.0);
    numer.assign(aligned, 0.0);
    numer2.assign(aligned, 0.0);
  }

  int group(int k) const { return (grouped || cox) ? pid[k] : k; }

  template <class F>
  void forColumn(int j, F f) const {
    const int fmt = d.fmt[j];
    for (int64_t p = d.colPtr[j]; p < d.colPtr[j + 1]; ++p) {
      const int k = (fmt == DENSE || fmt == INTERCEPT) ? (int)(p - d.colPtr[j]) : d.rows[p];
      const double x = (fmt == SPARSE || fmt == DENSE) ? d.vals[p] : 1.0;
      f(k, x);
    }
  }

  // ModelSpecifics.hpp:1809-1868
  void setPidForAccumulation(const double* w) {
    pid = pidOrig;
    accReset.clear();
    const int ignore = -1;
    int index = 0;
    while (w != nullptr && index < K && w[index] == 0.0) {
      pid[index] = ignore;
      index++;
    }
    int lastPid = pid[index];
    double lastTime = d.offs[index];
    double lastEvent = d.y[index];
    int p = pid[index] = 0;
    for (int k = index + 1; k < K; ++k) {
      if (w == nullptr || w[k] != 0.0) {
        int nextPid = pid[k];
        if (nextPid != lastPid) {
          p++;
          accReset.push_back(p);
          lastPid = nextPid;
        } else {
          if (lastEvent == 1.0 && lastTime == d.offs[k] && lastEvent == d.y[k]) {
            // tied event: shares the denominator of the previous row
          } else {
            p++;
          }
        }
        lastTime = d.offs[k];
        lastEvent = d.y[k];
        pid[k] = p;
      } else {
        pid[k] = ignore;
      }
    }
    p++;
    accReset.push_back(p);
    N = p;
    if (w != nullptr) {
      for (int k = 0; k < K; ++k)
        if (pid[k] == ignore) pid[k] = N;
    }
    setupSparseIndices(N);
    size_t aligned = (size_t)((N + 1 + 15) / 16) * 16;
    if (denom.size() < aligned) {
      denom.resize(aligned, 0.0);
      numer.resize(aligned, 0.0);
      numer2.resize(aligned, 0.0);
    }
  }

  // ModelSpecifics.hpp:1870-1892
  void setupSparseIndices(int max) {
    sparseIdx.assign(J, std::vector<int>());
    sparseIsNull.assign(J, 0);
    for (int j = 0; j < J; ++j) {
      if (d.fmt[j] == DENSE || d.fmt[j] == INTERCEPT) {
        sparseIsNull[j] = 1;
      } else {
        std::set<int> unique;
        for (int64_t p = d.colPtr[j]; p < d.colPtr[j + 1]; ++p) {
          const int k = d.rows[p];
          const int i = (cox || grouped) ? pid[k] : k;
          if (i < max) unique.insert(i);
        }
        sparseIdx[j].assign(unique.begin(), unique.end());
      }
    }
  }

  double observationCount(double y) const {
    if (grouped || cox) return y;
    return 1.0;
  }

  // ModelSpecifics.hpp:427-484 (+ setPidForAccumulation for Cox)
  void setWeights(const double* w, bool useCV) {
    kw.assign(K, 1.0);
    if (useCV)
      for (int k = 0; k < K; ++k) kw[k] = w[k];
    if (cox) setPidForAccumulation(useCV ? w : nullptr);
    nw.assign(N + 1, 0.0);
    for (int k = 0; k < K; ++k) nw[group(k)] += observationCount(d.y[k]) * kw[k];
  }

  // ModelSpecifics.hpp:486-524
  void computeXjY(bool useCV) {
    for (int j = 0; j < J; ++j) {
      double s = 0.0;
      forColumn(j, [&](int k, double x) {
        const double yy = cox ? observationCount(d.y[k]) : d.y[k];
        s += useCV ? x * yy * kw[k] : x * yy;
      });
      xjy[j] = s;
    }
  }

  // ModelSpecifics.hpp:570-587, ModelSpecifics.h:1596-1601
  void computeFixedTermsInLogLikelihood(bool useCV) {
    fixedLL = 0.0;
    if (d.model != M_POISSON) return;
    for (int k = 0; k < K; ++k) {
      double t = 0.0;
      for (int i = 2; i <= (int)d.y[k]; ++i) t += -std::log((double)i);
      fixedLL += useCV ? t * kw[k] : t;
    }
  }

  void zeroXBeta() { std::fill(xb.begin(), xb.end(), 0.0); }

  void axpyXBeta(double beta, int j) {
    if (beta != 0.0) forColumn(j, [&](int k, double x) { xb[k] += beta * x; });
  }

  double denomNull() const { return d.model == M_LOGISTIC ? 1.0 : 0.0; }

  // ModelSpecifics.hpp:1670-1694
  void computeAccumulatedDenominator() {
    if (!cox) return;
    if (accDenom.size() != (size_t)(N + 1)) accDenom.assign(N + 1, 0.0);
    double total = 0.0;
    size_t r = 0;
    for (int i = 0; i < N; ++i) {
      if (r < accReset.size() && accReset[r] == i) {
        total = 0.0;
        ++r;
      }
      total += denom[i];
      accDenom[i] = total;
    }
  }

  // ModelSpecifics.hpp:1588-1612
  void computeRemainingStatistics(bool useWeights) {
    std::fill(denom.begin(), denom.begin() + N, denomNull());
    for (int k = 0; k < K; ++k) {
      oeb[k] = std::exp(xb[k]);
      const double v = (cox && useWeights) ? kw[k] * oeb[k] : oeb[k];
      denom[group(k)] += v;
    }
    computeAccumulatedDenominator();
  }

  // ModelSpecifics.hpp:1353-1464 (Cox only)
  void computeNumeratorForGradient(int j, bool useWeights) {
    if (!cox) return;
    const int fmt = d.fmt[j];
    if (fmt == INDICATOR) {
      for (int i : sparseIdx[j]) numer[i] = 0.0;
    } else if (fmt == SPARSE) {
      for (int i : sparseIdx[j]) {
        numer[i] = 0.0;
        numer2[i] = 0.0;
      }
    } else {
      std::fill(numer.begin(), numer.begin() + N, 0.0);
      std::fill(numer2.begin(), numer2.begin() + N, 0.0);
    }
    const bool ind = isIndicatorFmt(fmt);
    forColumn(j, [&](int k, double x) {
      const double w = useWeights ? kw[k] : 1.0;
      numer[pid[k]] += useWeights ? w * (oeb[k] * x) : oeb[k] * x;
      if (!ind) numer2[pid[k]] += useWeights ? w * (oeb[k] * x * x) : oeb[k] * x * x;
    });
  }

  // ModelSpecifics.hpp:891-1201 with the increments of ModelSpecifics.h
  void gradHess(int j, bool useWeights, double* og, double* oh) {
    const int64_t nnz = d.colPtr[j + 1] - d.colPtr[j];
    if (nnz == 0) {
      *og = 0.0;
      *oh = 0.0;
      return;
    }
    const int fmt = d.fmt[j];
    const bool ind = isIndicatorFmt(fmt);
    double g = 0.0, h = 0.0;
    if (cox) {
      const bool dense = sparseIsNull[j];
      if (dense || !sparseIdx[j].empty()) {
        auto incr = [&](double num, double num2, int i) {
          const double t = num / accDenom[i];
          const double gg = nw[i] * t;
          g += gg;
          if (ind)
            h += gg * (1.0 - t);
          else
            h += nw[i] * (num2 / accDenom[i] - t * t);
        };
        double acc = 0.0, acc2 = 0.0;
        if (dense) {
          size_t r = 0;
          while (accReset[r] < 0) ++r;
          for (int i = 0; i < N; ++i) {
            if (accReset[r] <= i) {
              acc = 0.0;
              acc2 = 0.0;
              ++r;
            }
            acc += numer[i];
            acc2 += numer2[i];
            incr(acc, acc2, i);
          }
        } else {
          const std::vector<int>& idx = sparseIdx[j];
          size_t r = 0;
          while (accReset[r] < idx[0]) ++r;
          for (size_t s = 0; s < idx.size(); ++s) {
            int i = idx[s];
            if (accReset[r] <= i) {
              acc = 0.0;
              acc2 = 0.0;
              ++r;
            }
            acc += numer[i];
            acc2 += ind ? 0.0 : numer2[i];
            incr(acc, acc2, i);
            const int next = (s + 1 < idx.size()) ? idx[s + 1] : N;
            for (++i; i < next; ++i) {
              if (accReset[r] <= i) {
                acc = 0.0;
                acc2 = 0.0;
                ++r;
              }
              incr(acc, acc2, i);
            }
          }
        }
      }
    } else if (!grouped) {
      const bool lr = d.model == M_LOGISTIC;
      forColumn(j, [&](int k, double x) {
        const double n1 = oeb[k] * x;
        const double n2 = ind ? 0.0 : oeb[k] * x * x;
        const double w = nw[k];
        if (lr) {
          const double t = n1 / denom[k];
          g += useWeights ? w * t : t;
          double hh = ind ? t * (1.0 - t) : (n2 / denom[k] - t * t);
          h += useWeights ? w * hh : hh;
        } else {
          if (ind) {
            const double v = useWeights ? w * n1 : n1;
            g += v;
            h += v;
          } else {
            g += useWeights ? w * n1 : n1;
            h += useWeights ? w * n2 : n2;
          }
        }
      });
    } else {
      // grouped (conditional logistic): per-stratum sums over a sorted column
      double n1 = 0.0, n2 = 0.0;
      int key = -1;
      bool first = true;
      auto flush = [&]() {
        const double t = n1 / denom[key];
        const double gg = nw[key] * t;
        g += gg;
        if (ind)
          h += gg * (1.0 - t);
        else
          h += nw[key] * (n2 / denom[key] - t * t);
      };
      forColumn(j, [&](int k, double x) {
        const int kk = pid[k];
        if (first) {
          key = kk;
          first = false;
        } else if (kk != key) {
          flush();
          n1 = 0.0;
          n2 = 0.0;
          key = kk;
        }
        n1 += oeb[k] * x;
        n2 += ind ? 0.0 : oeb[k] * x * x;
      });
      flush();
    }
    g -= xjy[j];
    *og = g;
    *oh = h;
  }

  // ModelSpecifics.hpp:1524-1586
  void updateXBeta(double delta, int j, bool useWeights) {
    forColumn(j, [&](int k, double x) {
      xb[k] += delta * x;
      if (cox) {
        const double oldE = useWeights ? kw[k] * oeb[k] : oeb[k];
        oeb[k] = std::exp(xb[k]);
        const double newE = useWeights ? kw[k] * oeb[k] : oeb[k];
        denom[pid[k]] += newE - oldE;
      } else {
        const double oldE = oeb[k];
        const double newE = oeb[k] = std::exp(xb[k]);
        denom[group(k)] += newE - oldE;
      }
    });
    computeAccumulatedDenominator();
  }

  // ModelSpecifics.hpp:605-650
  double logLikelihood(bool useCV) const {
    double ll = 0.0;
    for (int k = 0; k < K; ++k) {
      const int yi = (int)d.y[k];
      double c = cox ? (yi != 1 ? 0.0 : xb[k]) : yi * xb[k];
      ll += useCV ? c * kw[k] : c;
    }
    if (cox) {
      for (int i = 0; i < N; ++i) ll -= nw[i] * std::log(accDenom[i]);
    } else if (d.model == M_POISSON) {
      for (int i = 0; i < N; ++i) ll -= nw[i] * denom[i];
    } else {
      for (int i = 0; i < N; ++i) ll -= nw[i] * std::log(denom[i]);
    }
    if (d.model == M_POISSON) ll += fixedLL;
    return ll;
  }

  // ModelSpecifics.hpp:103-120
  double gradientObjective(bool useCV) const {
    double c = 0.0;
    for (int k = 0; k < K; ++k) c += useCV ? xb[k] * d.y[k] * kw[k] : xb[k] * d.y[k];
    return c;
  }

  // Row-major copy of X as CompressedDataMatrix::transpose builds it
  // (CompressedDataMatrix.cpp:151-208; add_data CompressedDataMatrix.h:159-188).
  // Every row gets one format: INDICATOR when the matrix has INDICATOR columns
  // and no DENSE one, so an entry keeps only its position (a SPARSE value counts
  // as 1); otherwise the values themselves.  Zero values are dropped (SPARSE /
  // INDICATOR add_data); a DENSE row keeps them, which adds exact zeros to the
  // row sums below and so changes nothing for finite coefficients.
  std::vector<std::vector<std::pair<int, double>>> xt;
  void buildTranspose() {
    bool anyDense = false, anyIndicator = false;
    for (int j = 0; j < J; ++j) {
      anyDense = anyDense || d.fmt[j] == DENSE;
      anyIndicator = anyIndicator || d.fmt[j] == INDICATOR;
    }
    const bool positionsOnly = anyIndicator && !anyDense;
    xt.assign(K, {});
    for (int j = 0; j < J; ++j) {
      forColumn(j, [&](int k, double x) {
        if (x == 0.0) return;
        xt[k].push_back({j, positionsOnly ? 1.0 : x});
      });
    }
  }

  // ModelSpecifics::computeXBeta(beta) (ModelSpecifics.hpp:313-364): per row, the
  // sum over its entries in column order.
  void mmComputeXBeta(const double* beta) {
    if (xt.empty()) buildTranspose();
    for (int k = 0; k < K; ++k) {
      double sum = 0.0;
      for (const auto& e : xt[k]) sum += e.second * beta[e.first];
      xb[k] = sum;
    }
  }

  // computeMMGradientAndHessianImpl (ModelSpecifics.hpp:848-889) with the
  // increments of ModelSpecifics.h:500-517 (Survival: conditional logistic and
  // Cox) and 571-584 (logistic): weight nW and denominator of the row's group, no
  // 1-t factor.  For Cox the group is the row's pid and the denominator the pid's
  // own sum (denomPid, not the accumulated risk set); a zero-weight row's pid is N,
  // whose event weight is 0 and whose denominator slot keeps what an earlier
  // statistics pass left there (0 in a fresh engine: the term is then 0/0).
  void mmGradHess(int j, double* og, double* oh) const {
    const bool ind = isIndicatorFmt(d.fmt[j]);
    double g = 0.0, h = 0.0;
    forColumn(j, [&](int k, double x) {
      const int i = group(k);
      if (ind) {
        g += nw[i] * oeb[k] / denom[i];
        h += nw[i] * oeb[k] / denom[i];
      } else {
        g += nw[i] * oeb[k] * x / denom[i];
        h += nw[i] * oeb[k] * x * x / denom[i];
      }
    });
    g -= xjy[j];
    *og = g;
    *oh = h;
  }

  // ModelSpecifics.hpp:652-690 (the CCD recomputes xb and statistics first)
  double predictiveLogLikelihood(const double* w) {
    std::vector<double> saved;
    if (cox) {
      saved = kw;
      setWeights(w, true);
      computeRemainingStatistics(true);
    }
    double ll = 0.0;
    for (int k = 0; k < K; ++k) {
      const double y = d.y[k];
      switch (d.model) {
        case M_POISSON:
          ll += (y * xb[k] - std::exp(xb[k])) * w[k];
          break;
        case M_LOGISTIC:
          ll += y * w[k] * (xb[k] - std::log(denom[k]));
          break;
        case M_CLR:
          ll += y * w[k] * (xb[k] - std::log(denom[pid[k]]));
          break;
        case M_COX:
          ll += w[k] == 0.0 ? 0.0 : y * w[k] * (xb[k] - std::log(accDenom[pid[k]]));
          break;
      }
    }
    if (cox) {
      setWeights(saved.data(), true);
      computeRemainingStatistics(true);
    }
    return ll;
  }
};

// priors/CovariatePrior.h:186-188, 240-273, 410-415
double priorDelta(int type, double var, double g, double h, double beta) {
  if (type == P_NONE) return -(g / h);
  if (type == P_NORMAL) return -(g + (beta / var)) / (h + (1.0 / var));
  const double lambda = std::sqrt(2.0 / var);
  auto sign = [](double x) { return x == 0 ? 0 : (x < 0 ? -1 : 1); };
  double delta = 0.0;
  const double neg = -(g - lambda) / h;
  const double pos = -(g + lambda) / h;
  const int s = sign(beta);
  if (s == 0) {
    if (neg < 0)
      delta = neg;
    else if (pos > 0)
      delta = pos;
    else
      delta = 0;
  } else {
    delta = s < 0 ? neg : pos;
    if (sign(beta + delta) != s) delta = -beta;
  }
  return delta;
}

// priors/CovariatePrior.h:221-225, 404-408
double priorLogDensity(int type, double var, double beta) {
  if (type == P_NONE) return 0.0;
  if (type == P_NORMAL) return -0.5 * std::log(2.0 * M_PI * var) - 0.5 * beta * beta / var;
  const double lambda = std::sqrt(2.0 / var);
  return std::log(0.5 * lambda) - lambda * std::abs(beta);
}

template <class T>
const T& stdmax(const T& a, const T& b) {
  return (a < b) ? b : a;
}

}  // namespace

extern "C" {

typedef struct {
  int K, J;
  const int64_t* col_ptr;
  const int32_t* rows;
  const double* vals;
  const int8_t* fmt;
  const double* y;
  const double* offs;
  const int32_t* strata;
  int model;
} orc_data;

typedef struct {
  double tolerance;
  int max_iterations;
  int convergence_type;
  double initial_bound;
  int max_bound_count;
  int has_offset;  // column 0 is an offset covariate (kept by resetBeta)
  int algorithm;   // 0: cyclic coordinate descent, 1: MM (AlgorithmType, Types.h)
} orc_control;

typedef struct {
  double log_likelihood;
  double log_prior;
  int iterations;
  int return_flag;
  int bound_count;
  double pred_log_likelihood;  // filled when held-out weights are supplied
} orc_result;

// Full fit, restating CyclicCoordinateDescent::update -> findMode.
//   beta      in: starting coefficients (offset column carries 1); out: estimate
//   fixed     per-column fixBeta flags (offset column fixed)
//   weights   nullable; non-null means useCrossValidation=true
//   heldout   nullable; when non-null, the predictive log likelihood of these weights
//             is evaluated after the fit (CV scoring)
// Returns 0; with the MM algorithm -1 for a model without MM increments (Poisson:
// ModelSpecifics.h:1546-1552 throws) and -2 when the log posterior decreased
// ("Non-increasing!", CyclicCoordinateDescent.cpp:1033-1046).
int orc_fit(const orc_data* od, const int* prior_type, const double* prior_var,
            const uint8_t* fixed, double* beta, const double* weights, const double* heldout,
            const orc_control* ctl, orc_result* out) {
  Data d{od->K, od->J, od->col_ptr, od->rows, od->vals, od->fmt, od->y, od->offs, od->strata,
         od->model};
  const int J = d.J;
  Engine e(d);
  const bool useCV = weights != nullptr;
  const bool hasOffset = ctl->has_offset != 0;  // CyclicCoordinateDescent.cpp:256-264
  const bool mm = ctl->algorithm == 1;
  if (mm && d.model == M_POISSON) return -1;
  std::vector<double> allDelta(J, 0.0);

  std::vector<double> b(beta, beta + J), bound(J, ctl->initial_bound);
  double initialBound = ctl->initial_bound;
  bool validWeights = false, xbKnown = false, ssKnown = false;
  int flag = SUCCESS, iterations = 0, count = 0;

  auto computeXBeta = [&]() {
    e.zeroXBeta();
    for (int j = 0; j < J; ++j) e.axpyXBeta(b[j], j);
  };
  auto logPrior = [&]() {
    double s = 0.0;
    for (int j = 0; j < J; ++j) s += priorLogDensity(prior_type[j], prior_var[j], b[j]);
    return s;
  };
  auto objective = [&]() {
    switch (ctl->convergence_type) {
      case C_GRADIENT:
        return e.gradientObjective(useCV);
      case C_MITTAL:
        return e.logLikelihood(useCV);
      case C_LANGE:
        return e.logLikelihood(useCV) + logPrior();
      default:
        return 0.0;
    }
  };

  bool done = false;
  while (!done) {
    // findMode (CyclicCoordinateDescent.cpp:909-1216)
    if (!validWeights) {
      e.setWeights(weights, useCV);
      e.computeFixedTermsInLogLikelihood(useCV);
      e.computeXjY(useCV);
      validWeights = true;
    }
    if (!xbKnown) {
      computeXBeta();
      xbKnown = true;
      ssKnown = false;
    }
    if (!ssKnown) {
      e.computeRemainingStatistics(true);
      ssKnown = true;
    }
    std::fill(bound.begin(), bound.end(), initialBound);
    int iteration = 0;
    double lastObj = objective();
    double lastLogPosterior = 0.0;
    bool fmDone = false;
    while (!fmDone) {
      if (mm) {
        // mmUpdateAllBeta: every free coordinate's step from one set of statistics
        for (int j = 0; j < J; ++j) {
          allDelta[j] = 0.0;
          if (fixed[j]) continue;
          double g, h;
          e.mmGradHess(j, &g, &h);
          if (h < 0.0) {
            g = 0.0;
            h = 0.0;
          }
          allDelta[j] = priorDelta(prior_type[j], prior_var[j], g, h, b[j]);
        }
        for (int j = 0; j < J; ++j)
          if (!fixed[j] && allDelta[j] != 0.0) b[j] += allDelta[j];
        e.mmComputeXBeta(b.data());
        e.computeRemainingStatistics(true);
        iteration++;
        const double logPosterior = e.logLikelihood(useCV) + logPrior();
        if (iteration > 1 && logPosterior - lastLogPosterior < 0.0) {
          for (int j = 0; j < J; ++j) beta[j] = b[j];
          out->iterations = iteration;
          return -2;
        }
        lastLogPosterior = logPosterior;
      }
      for (int j = 0; !mm && j < J; ++j) {
        if (fixed[j]) continue;
        e.computeNumeratorForGradient(j, useCV);
        double g, h;
        e.gradHess(j, useCV, &g, &h);
        if (h < 0.0) {
          g = 0.0;
          h = 0.0;
        }
        double delta = priorDelta(prior_type[j], prior_var[j], g, h, b[j]);
        // applyBounds
        if (delta < -bound[j])
          delta = -bound[j];
        else if (delta > bound[j])
          delta = bound[j];
        double inter = stdmax(std::abs(delta) * 2, bound[j] / 2);
        inter = stdmax(inter, 1E-3);
        bound[j] = inter;
        if (delta != 0.0) {
          b[j] += delta;
          e.updateXBeta(delta, j, useCV);
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
        conv = std::abs(obj - lastObj) / (std::abs(obj) + 1.0);
      }
      lastObj = obj;
      if (ctl->tolerance > 0 && conv < ctl->tolerance) {
        flag = ill ? ILLCONDITIONED : SUCCESS;
        fmDone = true;
      } else if (iteration == ctl->max_iterations) {
        flag = MAX_ITERATIONS;
        fmDone = true;
      }
    }
    iterations = iteration;
    ++count;
    if (flag == ILLCONDITIONED && count < ctl->max_bound_count) {
      initialBound /= 10.0;
      for (int j = hasOffset ? 1 : 0; j < J; ++j) b[j] = 0.0;
      computeXBeta();
      ssKnown = false;
    } else {
      done = true;
    }
  }
  for (int j = 0; j < J; ++j) beta[j] = b[j];
  out->log_likelihood = e.logLikelihood(useCV);
  out->log_prior = logPrior();
  out->iterations = iterations;
  out->return_flag = flag;
  out->bound_count = count;
  out->pred_log_likelihood = 0.0;
  if (heldout != nullptr) {
    computeXBeta();
    e.computeRemainingStatistics(true);
    out->pred_log_likelihood = e.predictiveLogLikelihood(heldout);
  }
  return 0;
}

// Log likelihood, gradient and Hessian of coordinate j at a given beta (fresh
// statistics).  Used by the finite-difference / brute-force checks.
int orc_eval(const orc_data* od, const double* beta, const double* weights, int j,
             double* loglik, double* grad, double* hess) {
  Data d{od->K, od->J, od->col_ptr, od->rows, od->vals, od->fmt, od->y, od->offs, od->strata,
         od->model};
  Engine e(d);
  const bool useCV = weights != nullptr;
  e.setWeights(weights, useCV);
  e.computeFixedTermsInLogLikelihood(useCV);
  e.computeXjY(useCV);
  e.zeroXBeta();
  for (int c = 0; c < d.J; ++c) e.axpyXBeta(beta[c], c);
  e.computeRemainingStatistics(true);
  *loglik = e.logLikelihood(useCV);
  if (j >= 0) {
    e.computeNumeratorForGradient(j, useCV);
    e.gradHess(j, useCV, grad, hess);
  }
  return 0;
}

// Log likelihood and the gradient/Hessian of each coordinate in cols[0..ncols)
// at one beta, from one set of fresh statistics (orc_eval for many columns).
int orc_eval_cols(const orc_data* od, const double* beta, const double* weights,
                  const int32_t* cols, int ncols, double* loglik, double* grad, double* hess) {
  Data d{od->K, od->J, od->col_ptr, od->rows, od->vals, od->fmt, od->y, od->offs, od->strata,
         od->model};
  Engine e(d);
  const bool useCV = weights != nullptr;
  e.setWeights(weights, useCV);
  e.computeFixedTermsInLogLikelihood(useCV);
  e.computeXjY(useCV);
  e.zeroXBeta();
  for (int c = 0; c < d.J; ++c) e.axpyXBeta(beta[c], c);
  e.computeRemainingStatistics(true);
  *loglik = e.logLikelihood(useCV);
  for (int i = 0; i < ncols; ++i) {
    e.computeNumeratorForGradient(cols[i], useCV);
    e.gradHess(cols[i], useCV, grad + i, hess + i);
  }
  return 0;
}

// MM gradient and Hessian of every column at a given beta (fresh statistics;
// fixed columns included).
int orc_mm_eval(const orc_data* od, const double* beta, const double* weights, double* g,
                double* h) {
  Data d{od->K, od->J, od->col_ptr, od->rows, od->vals, od->fmt, od->y, od->offs, od->strata,
         od->model};
  if (d.model == M_POISSON) return -1;
  Engine e(d);
  const bool useCV = weights != nullptr;
  e.setWeights(weights, useCV);
  e.computeXjY(useCV);
  e.zeroXBeta();
  for (int j = 0; j < d.J; ++j) e.axpyXBeta(beta[j], j);
  e.computeRemainingStatistics(true);
  for (int j = 0; j < d.J; ++j) e.mmGradHess(j, &g[j], &h[j]);
  return 0;
}

// xBeta as ModelSpecifics::computeXBeta(beta) forms it from the transpose.
int orc_mm_xbeta(const orc_data* od, const double* beta, double* xb_out) {
  Data d{od->K, od->J, od->col_ptr, od->rows, od->vals, od->fmt, od->y, od->offs, od->strata,
         od->model};
  Engine e(d);
  e.mmComputeXBeta(beta);
  for (int k = 0; k < d.K; ++k) xb_out[k] = e.xb[k];
  return 0;
}

double orc_prior_delta(int type, double var, double g, double h, double beta) {
  return priorDelta(type, var, g, h, beta);
}

double orc_prior_log_density(int type, double var, double beta) {
  return priorLogDensity(type, var, beta);
}

// Cox risk-set grouping (setPidForAccumulation): writes pid[K], returns N and
// fills accReset (capacity K+1) with *nreset entries.
int orc_cox_pids(const orc_data* od, const double* weights, int32_t* pid_out, int32_t* reset_out,
                 int* nreset) {
  Data d{od->K, od->J, od->col_ptr, od->rows, od->vals, od->fmt, od->y, od->offs, od->strata,
         od->model};
  Engine e(d);
  e.setPidForAccumulation(weights);
  for (int k = 0; k < d.K; ++k) pid_out[k] = e.pid[k];
  for (size_t r = 0; r < e.accReset.size(); ++r) reset_out[r] = e.accReset[r];
  *nreset = (int)e.accReset.size();
  return e.N;
}

// CrossValidationSelector fold weights: ids are the exchangeable unit per row
// (stratum for BY_PID, row index for BY_ROW).  The selector is reseeded, then
// permuted once per repetition (rep = task / fold); weights of fold `fold`'s
// members are zeroed.  CrossValidationSelector.cpp:21-116,138-143.
int orc_cv_weights(const int32_t* ids, int K, int by_pid, int nfold, long seed, int task,
                   double* weights_out) {
  int N = 0;
  if (by_pid) {
    for (int k = 0; k < K; ++k) N = std::max(N, ids[k] + 1);
  } else {
    N = K;
  }
  std::vector<int> start;
  int index = 0, fraction = N / nfold, extra = N - fraction * nfold;
  for (int i = 0; i < nfold; ++i) {
    start.push_back(index);
    index += fraction;
    if (i < extra) index++;
  }
  start.push_back(N);
  // AbstractSelector.cpp:32-48: seed -1 makes the selector deterministic (permute
  // does not shuffle, CrossValidationSelector.cpp:140-144); reseed (:60-66) seeds
  // the engine whatever the seed.
  std::mt19937 prng;
  prng.seed(seed);
  std::vector<int> perm(N);
  for (int i = 0; i < N; ++i) perm[i] = i;
  for (int i = 0; i <= task; ++i)
    if (i % nfold == 0 && seed != -1) std::shuffle(perm.begin(), perm.end(), prng);
  const int fold = task % nfold;
  for (int k = 0; k < K; ++k) weights_out[k] = 1.0;
  if (by_pid) {
    std::set<int> ex(perm.begin() + start[fold], perm.begin() + start[fold + 1]);
    for (int k = 0; k < K; ++k)
      if (ex.count(ids[k])) weights_out[k] = 0.0;
  } else {
    for (int p = start[fold]; p < start[fold + 1]; ++p) weights_out[perm[p]] = 0.0;
  }
  return 0;
}

// BootstrapSelector (drivers/BootstrapSelector.cpp:17-95): the selector keeps
// the drawn pids in a std::multiset and a row weighs the multiset count of its
// pid; the constructor permutes once and BootstrapDriver::drive
// (drivers/BootstrapDriver.cpp:56-60) permutes again before each replicate.
int orc_bootstrap_weights(const int32_t* ids, int K, long seed, int replicate,
                          const double* exclude, int n_exclude, double* weights_out) {
  int N = 0;
  for (int k = 0; k < K; ++k) N = std::max(N, ids[k] + 1);
  std::vector<int> indicesIncluded;
  if (exclude) {
    for (int i = 0; i < n_exclude; ++i)
      if (exclude[i] == 0) indicesIncluded.push_back(i);
  } else {
    for (int i = 0; i < N; ++i) indicesIncluded.push_back(i);
  }
  // AbstractSelector.cpp:32-48: seed -1 leaves the engine default-constructed.
  std::mt19937 prng;
  if (seed != -1) prng.seed(seed);
  std::multiset<int> selected;
  auto permute = [&]() {
    selected.clear();
    const int nNew = (int)indicesIncluded.size();
    std::uniform_int_distribution<int> uniform(0, nNew - 1);
    for (int i = 0; i < nNew; i++) selected.insert(indicesIncluded[uniform(prng)]);
  };
  permute();  // constructor
  for (int r = 0; r <= replicate; ++r) permute();  // one per replicate
  for (int k = 0; k < K; ++k) weights_out[k] = (double)selected.count(ids[k]);
  return 0;
}

// UniModalSearch (utils/HParSearch.h:30-65, utils/HParSearch.cpp:22-120) as a
// small class: tried() keeps a map x -> (mean, stddev) and the best entry;
// step() proposes the next x.  The quadratic fit in log x solves its 3x3 normal
// equations by Cramer's rule.
struct OrcSearch {
  std::map<double, std::pair<double, double>> pts;
  std::map<double, std::pair<double, double>>::iterator best;
  void tried(double x, double y, double sd) {
    pts[x] = std::make_pair(y, sd);
    if (pts.size() == 1) best = pts.begin();
    else if (y > best->second.first) best = pts.find(x);
  }
  static double det3(const double m[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
           m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  }
  void fit(double c[3]) const {
    double minSd = std::numeric_limits<double>::max();
    for (auto& p : pts)
      if (p.second.second > 0 && p.second.second < minSd) minSd = p.second.second;
    const double adj = minSd < std::numeric_limits<double>::max() ? 100.0 / minSd : 1.0;
    double M[3][3] = {{0}}, v[3] = {0};
    for (auto& p : pts) {
      const double w = p.second.second > 0 ? 1.0 / p.second.second : adj;
      const double l = std::log(p.first), f[3] = {1.0, l, l * l};
      for (int a = 0; a < 3; ++a) {
        v[a] += w * f[a] * p.second.first;
        for (int b = 0; b < 3; ++b) M[a][b] += w * f[a] * f[b];
      }
    }
    const double D = det3(M);
    for (int k = 0; k < 3; ++k) {
      double Mk[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Mk[a][b] = b == k ? v[a] : M[a][b];
      c[k] = det3(Mk) / D;
    }
  }
  void step(double stdstep, double stopY, double stopX, double firstCut, int* cont, double* next,
            double* expected) const {
    *cont = 1;
    *next = 0.0;
    *expected = 0.0;
    switch (pts.size()) {
      case 0: *next = 1.0; break;
      case 1: {
        const double x = pts.begin()->first;
        *next = x < firstCut ? x * stdstep : x / stdstep;
        break;
      }
      case 2:
        *next = pts.begin()->second.first > pts.rbegin()->second.first ? pts.begin()->first / stdstep
                                                                       : pts.rbegin()->first * stdstep;
        break;
      default:
        if (pts.begin()->first == best->first) {
          *next = best->first / stdstep;
          if (!(best->first > std::numeric_limits<double>::denorm_min())) {
            *cont = 0;
            *expected = best->second.first;
          }
        } else if (pts.rbegin()->first == best->first) {
          *next = best->first * stdstep;
          if (!(best->first < std::numeric_limits<double>::infinity())) {
            *cont = 0;
            *expected = best->second.first;
          }
        } else {
          double c[3];
          fit(c);
          const double la = -c[1] / c[2] / 2, em = -c[1] * c[1] / c[2] / 4 + c[0];
          const double mv = best->second.first;
          if (mv == 0) *cont = 0;
          else if ((em - mv) / std::fabs(mv) < stopY) *cont = 0;
          else if (std::fabs(la - std::log(best->first)) < stopX) *cont = 0;
          *next = std::exp(la);
          *expected = em;
        }
    }
  }
};

int orc_search_step(const double* xs, const double* ys, const double* sds, int n, double stdstep,
                    double stopY, double stopX, double firstCut, int* cont, double* next,
                    double* expected) {
  OrcSearch s;
  for (int i = 0; i < n; ++i) s.tried(xs[i], ys[i], sds ? sds[i] : 0.0);
  s.step(stdstep, stopY, stopX, firstCut, cont, next, expected);
  return 0;
}

int orc_search_quadratic(const double* xs, const double* ys, const double* sds, int n, double* c) {
  OrcSearch s;
  for (int i = 0; i < n; ++i) s.tried(xs[i], ys[i], sds ? sds[i] : 0.0);
  s.fit(c);
  return 0;
}

// GridSearchCrossValidationDriver::computeGridPoint (log-uniform grid)
double orc_grid_point(double lower, double upper, int size, int step) {
  if (size == 1) return upper;
  const double stepSize = (std::log(upper) - std::log(lower)) / (size - 1);
  return std::exp(std::log(lower) + step * stepSize);
}

}  // extern "C"
