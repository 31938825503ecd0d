// ccd_engine.hip -- resident data, per-fit state and the per-coordinate HIP
// kernels behind the AbstractModelSpecifics-shaped C ABI (include/ccdhip.h).
//
// Layout in HBM (one data handle, shared read-only by every fit over it):
//   colPtr int64[J+1] | rows int32[nnz] | vals f64[nnz] (+ f32 copy, mixed) | fmt int8[J]
//   y f64[K] | strata int32[K] | grouped models: per-column runs of equal stratum
//   (segPtr int64[J+1], segOff int32[nseg], segKey int32[nseg]) and stratum row
//   starts int32[S+1].
// Per fit: xBeta and exp(xBeta) [K] in FP64 (or FP32 on the mixed path), weights
// f64[K], per-group denominators / event weights f64[N+1] (FP64 on both paths),
// XjY f64[J]; Cox adds pid int32[K], the first row of each pid int32[N+1], the
// stratum-reset flag per pid and the numerator vectors f64[N+1].
//
// Reference counterparts (src/cyclops/engine/ModelSpecifics.hpp):
//   k_axpy ................. axpy / axpyXBeta :384-423
//   k_remaining_* .......... computeRemainingStatisticsImpl :1588-1612
//   k_gh_indep ............. computeGradientAndHessianImpl, independent rows :1022-1038
//   k_gh_clr ............... grouped branch :1101-1179 (one thread per stratum run)
//   k_cox_numer ............ computeNumeratorForGradient :1353-1464
//   k_cox_tile_* ........... cumulative branch :904-962 and computeAccumlatedDenominator
//                            :1670-1694 as a reduce-then-scan over pid tiles
//   k_update_* ............. updateXBetaImpl :1524-1586
//   k_rows_reduce .......... getLogLikelihood :605-650, getGradientObjective :103-120,
//                            getPredictiveLogLikelihood :652-690
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ccd_device.h"
#include "ccd_host.h"

using namespace ccdamd;

// ------------------------------------------------------------------ errors
namespace ccdamd {
thread_local std::string g_lastError;

int setError(int code, const std::string& msg) {
  g_lastError = msg;
  return code;
}
}  // namespace ccdamd

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return setError(CCDH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

namespace {

constexpr int kThreads = 256;
constexpr int kMaxGridBlocks = 2048;
constexpr int kCoxItems = 8;
constexpr int kCoxTile = kThreads * kCoxItems;

int gridFor(int64_t n, int perThread = 4) {
  int64_t b = (n + (int64_t)kThreads * perThread - 1) / ((int64_t)kThreads * perThread);
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, kMaxGridBlocks));
}

// ------------------------------------------------------------------ kernels
template <typename R>
__global__ void k_fill(R* a, int64_t n, R v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// xBeta += beta * X_j
template <typename R>
__global__ void k_axpy(DevX d, R* xb, int j, int fmt, double beta) {
  const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
  const R rb = (R)beta;
  for (int64_t p = b + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < e;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int k = colRow(d, fmt, p, b);
    xb[k] += rb * colValue<R>(d, fmt, p);
  }
}

// offsExpXBeta for every row; independent models start each row's denominator at
// the model's null value plus it (1 + oeb for logistic, oeb for Poisson;
// computeRemainingStatisticsImpl, ModelSpecifics.hpp:1588-1612); den may be null
template <typename R>
__global__ void k_exp_rows(int K, const R* xb, R* oeb, R* den, R nullValue) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < K;
       k += (int64_t)gridDim.x * blockDim.x) {
    const R o = rexp(xb[k]);
    oeb[k] = o;
    if (den) den[k] = nullValue + o;
  }
}

// grouped: one thread per stratum, rows summed in row order (as the reference)
template <typename R>
__global__ void k_remaining_clr(DevX d, const R* xb, R* oeb, double* denom) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < d.nStrata; s += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = d.stratStart[s]; k < d.stratStart[s + 1]; ++k) {
      const R o = rexp(xb[k]);
      oeb[k] = o;
      acc += (double)o;
    }
    denom[s] = acc;
  }
}

// Cox: one thread per pid; rows of a pid lie in [pidFirst[p], pidFirst[p+1]),
// interleaved only with ignored (zero-weight) rows whose pid is N.
template <typename R>
__global__ void k_remaining_cox(int N, const int32_t* pid, const int32_t* pidFirst, const R* oeb,
                                const double* kw, int useW, double* denom) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = pidFirst[p]; k < pidFirst[p + 1]; ++k) {
      if (pid[k] != p) continue;
      acc += useW ? kw[k] * (double)oeb[k] : (double)oeb[k];
    }
    denom[p] = acc;
  }
}

// per-group event weights nW[g] = sum y*w over the group's rows (setWeights)
__global__ void k_nw_clr(DevX d, const double* kw, double* nw) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < d.nStrata; s += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = d.stratStart[s]; k < d.stratStart[s + 1]; ++k) acc += d.y[k] * kw[k];
    nw[s] = acc;
  }
}

__global__ void k_nw_cox(int N, const double* y, const int32_t* pid, const int32_t* pidFirst,
                         const double* kw, double* nw) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = pidFirst[p]; k < pidFirst[p + 1]; ++k)
      if (pid[k] == p) acc += y[k] * kw[k];
    nw[p] = acc;
  }
}

// XjY = sum_k x_kj * y_k (* w_k): one block per column
__global__ void k_xjy(DevX d, const double* kw, int useW, double* xjy) {
  const int j = blockIdx.x;
  const int fmt = d.fmt[j];
  const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
  double s[1] = {0.0};
  for (int64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
    const int k = colRow(d, fmt, p, b);
    const double x = colValue<double>(d, fmt, p);
    s[0] += useW ? x * d.y[k] * kw[k] : x * d.y[k];
  }
  blockSum<1>(s);
  if (threadIdx.x == 0) xjy[j] = s[0];
}

// gradient / Hessian of one coordinate, independent rows (LR, Poisson)
template <int MODEL, typename R>
__global__ void k_gh_indep(DevX d, const R* oeb, const R* rden, const double* kw, int j, int fmt,
                           int useW, const double* xjy, double* partials, unsigned* ticket,
                           double* out) {
  const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
  const bool ind = isIndicator(fmt);
  double v[2] = {0.0, 0.0};
  for (int64_t p = b + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < e;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int k = colRow(d, fmt, p, b);
    const R x = colValue<R>(d, fmt, p);
    const R o = oeb[k];
    const R n1 = o * x;
    const double w = useW ? kw[k] : 1.0;
    if (MODEL == M_LOGISTIC) {
      const R den = rden[k];
      const R t = n1 / den;
      const R hh = ind ? t * ((R)1 - t) : ((n1 * x) / den - t * t);
      v[0] += useW ? w * (double)t : (double)t;
      v[1] += useW ? w * (double)hh : (double)hh;
    } else {  // Poisson
      const R n2 = ind ? n1 : n1 * x;
      v[0] += useW ? w * (double)n1 : (double)n1;
      v[1] += useW ? w * (double)n2 : (double)n2;
    }
  }
  blockSum<2>(v);
  if (gridFinish<2>(v, partials, ticket, out) && threadIdx.x == 0) out[0] -= xjy[j];
}

// grouped (conditional logistic): one thread per run of equal stratum in column j
template <typename R>
__global__ void k_gh_clr(DevX d, const R* oeb, const double* denom, const double* nw, int j,
                         int fmt, const double* xjy, double* partials, unsigned* ticket,
                         double* out) {
  const int64_t b = d.colPtr[j];
  const int nj = (int)(d.colPtr[j + 1] - b);
  const int64_t s0 = d.segPtr[j], s1 = d.segPtr[j + 1];
  const bool ind = isIndicator(fmt);
  double v[2] = {0.0, 0.0};
  for (int64_t s = s0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < s1;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int q0 = d.segOff[s];
    const int q1 = (s + 1 < s1) ? d.segOff[s + 1] : nj;
    double n1 = 0.0, n2 = 0.0;
    for (int q = q0; q < q1; ++q) {
      const int64_t p = b + q;
      const int k = colRow(d, fmt, p, b);
      const R x = colValue<R>(d, fmt, p);
      const R o = oeb[k];
      n1 += (double)(o * x);
      if (!ind) n2 += (double)(o * x * x);
    }
    const int key = d.segKey[s];
    const double t = n1 / denom[key];
    const double g = nw[key] * t;
    v[0] += g;
    v[1] += ind ? g * (1.0 - t) : nw[key] * (n2 / denom[key] - t * t);
  }
  blockSum<2>(v);
  if (gridFinish<2>(v, partials, ticket, out) && threadIdx.x == 0) out[0] -= xjy[j];
}

// Cox: is position p the head of a run of equal pid (ignoring zero-weight rows)?
template <typename R>
__device__ __forceinline__ bool coxRunHead(const DevX& d, int fmt, int64_t b, int64_t p,
                                           const int32_t* pid, int N, int* myPid) {
  const int pp = pid[colRow(d, fmt, p, b)];
  *myPid = pp;
  if (pp >= N) return false;
  for (int64_t q = p - 1; q >= b; --q) {
    const int pq = pid[colRow(d, fmt, q, b)];
    if (pq < N) return pq != pp;
  }
  return true;
}

// computeNumeratorForGradient: numer[p] = sum over the run of w*x*oeb
template <typename R>
__global__ void k_cox_numer(DevX d, const R* oeb, const double* kw, const int32_t* pid, int N,
                            int j, int fmt, int useW, double* numer, double* numer2) {
  const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
  const bool ind = isIndicator(fmt);
  for (int64_t p = b + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < e;
       p += (int64_t)gridDim.x * blockDim.x) {
    int pp;
    if (!coxRunHead<R>(d, fmt, b, p, pid, N, &pp)) continue;
    double n1 = 0.0, n2 = 0.0;
    for (int64_t q = p; q < e; ++q) {
      const int k = colRow(d, fmt, q, b);
      const int pq = pid[k];
      if (pq >= N) continue;
      if (pq != pp) break;
      const double x = (double)colValue<R>(d, fmt, q);
      const double o = (double)oeb[k];
      n1 += useW ? kw[k] * (o * x) : o * x;
      if (!ind) n2 += useW ? kw[k] * (o * x * x) : o * x * x;
    }
    numer[pp] = n1;
    if (!ind) numer2[pp] = n2;
  }
}

// Cox tile pass 1: segmented aggregate of (denom, numer, numer2) per pid tile
__global__ void k_cox_tile_agg(int N, const double* denom, const double* numer,
                               const double* numer2, const uint8_t* reset, int nv,
                               Seg<3>* agg) {
  const int base = blockIdx.x * kCoxTile + threadIdx.x * kCoxItems;
  Seg<3> a;
  a.f = 0;
  a.v[0] = a.v[1] = a.v[2] = 0.0;
#pragma unroll
  for (int it = 0; it < kCoxItems; ++it) {
    const int i = base + it;
    if (i < N) {
      Seg<3> x;
      x.f = reset[i];
      x.v[0] = denom[i];
      x.v[1] = nv > 1 ? numer[i] : 0.0;
      x.v[2] = nv > 2 ? numer2[i] : 0.0;
      a = segCombine(a, x);
    }
  }
  Seg<3> total;
  (void)blockSegScanExcl<3>(a, total);
  if (threadIdx.x == 0) agg[blockIdx.x] = total;
}

// Cox tile pass 2: exclusive segmented scan over tile aggregates (one block)
__global__ void k_cox_tile_scan(int T, const Seg<3>* agg, Seg<3>* carry) {
  const int per = (T + blockDim.x - 1) / blockDim.x;
  const int t0 = threadIdx.x * per;
  const int t1 = min(T, t0 + per);
  Seg<3> a;
  a.f = 0;
  a.v[0] = a.v[1] = a.v[2] = 0.0;
  for (int t = t0; t < t1; ++t) a = segCombine(a, agg[t]);
  Seg<3> total;
  Seg<3> run = blockSegScanExcl<3>(a, total);
  for (int t = t0; t < t1; ++t) {
    carry[t] = run;
    run = segCombine(run, agg[t]);
  }
}

// Cox tile pass 3: per-pid accumulated denominator / numerators and their use.
//   mode 0: gradient + Hessian (cumulative branch), clears numer/numer2 after use
//   mode 1: sum nW * log(accDenom)   (getLogLikelihood denominator term)
//   mode 2: store accDenom into accOut
__global__ void k_cox_tile_use(int N, int mode, int fmtIndicator, const double* denom,
                               double* numer, double* numer2, const uint8_t* reset,
                               const double* nw, const Seg<3>* carry, double* accOut, int j,
                               const double* xjy, double* partials, unsigned* ticket,
                               double* out) {
  const int base = blockIdx.x * kCoxTile + threadIdx.x * kCoxItems;
  const int nv = mode == 0 ? (fmtIndicator ? 2 : 3) : 1;
  Seg<3> items[kCoxItems];
  Seg<3> a;
  a.f = 0;
  a.v[0] = a.v[1] = a.v[2] = 0.0;
#pragma unroll
  for (int it = 0; it < kCoxItems; ++it) {
    const int i = base + it;
    Seg<3> x;
    x.f = 0;
    x.v[0] = x.v[1] = x.v[2] = 0.0;
    if (i < N) {
      x.f = reset[i];
      x.v[0] = denom[i];
      if (nv > 1) x.v[1] = numer[i];
      if (nv > 2) x.v[2] = numer2[i];
    }
    items[it] = x;
    a = segCombine(a, x);
  }
  Seg<3> total;
  Seg<3> run = segCombine(carry[blockIdx.x], blockSegScanExcl<3>(a, total));
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int it = 0; it < kCoxItems; ++it) {
    const int i = base + it;
    if (i >= N) break;
    run = segCombine(run, items[it]);
    const double accD = run.v[0];
    if (mode == 0) {
      const double t = run.v[1] / accD;
      const double g = nw[i] * t;
      v[0] += g;
      v[1] += fmtIndicator ? g * (1.0 - t) : nw[i] * (run.v[2] / accD - t * t);
      if (items[it].v[1] != 0.0) numer[i] = 0.0;
      if (nv > 2 && items[it].v[2] != 0.0) numer2[i] = 0.0;
    } else if (mode == 1) {
      v[0] += nw[i] * log(accD);
    } else {
      accOut[i] = accD;
    }
  }
  if (mode == 2) return;
  blockSum<2>(v);
  if (gridFinish<2>(v, partials, ticket, out) && threadIdx.x == 0 && mode == 0) out[0] -= xjy[j];
}

// updateXBeta, independent rows
template <typename R>
__global__ void k_update_indep(DevX d, R* xb, R* oeb, R* rden, int j, int fmt, double delta) {
  const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
  const R rd = (R)delta;
  for (int64_t p = b + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < e;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int k = colRow(d, fmt, p, b);
    const R nx = xb[k] + rd * colValue<R>(d, fmt, p);
    xb[k] = nx;
    const R o = rexp(nx);
    rden[k] += o - oeb[k];  // the row's denominator moves by the change
    oeb[k] = o;
  }
}

// updateXBeta, grouped: one thread per stratum run, denominator changes added in
// row order onto the stratum denominator (the reference's order)
template <typename R>
__global__ void k_update_clr(DevX d, R* xb, R* oeb, double* denom, int j, int fmt, double delta) {
  const int64_t b = d.colPtr[j];
  const int nj = (int)(d.colPtr[j + 1] - b);
  const int64_t s0 = d.segPtr[j], s1 = d.segPtr[j + 1];
  const R rd = (R)delta;
  for (int64_t s = s0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < s1;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int q0 = d.segOff[s];
    const int q1 = (s + 1 < s1) ? d.segOff[s + 1] : nj;
    const int key = d.segKey[s];
    double den = denom[key];
    for (int q = q0; q < q1; ++q) {
      const int64_t p = b + q;
      const int k = colRow(d, fmt, p, b);
      const R nx = xb[k] + rd * colValue<R>(d, fmt, p);
      xb[k] = nx;
      const R o = rexp(nx);
      den += (double)o - (double)oeb[k];
      oeb[k] = o;
    }
    denom[key] = den;
  }
}

// updateXBeta, Cox: rows of a pid run update denomPid[pid] with weighted changes
template <typename R>
__global__ void k_update_cox(DevX d, R* xb, R* oeb, double* denom, const double* kw,
                             const int32_t* pid, int N, int j, int fmt, int useW, double delta) {
  const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
  const R rd = (R)delta;
  for (int64_t p = b + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < e;
       p += (int64_t)gridDim.x * blockDim.x) {
    int pp;
    if (!coxRunHead<R>(d, fmt, b, p, pid, N, &pp)) {
      if (pp >= N) {  // ignored row: xBeta still moves, no denominator
        const int k = colRow(d, fmt, p, b);
        const R nx = xb[k] + rd * colValue<R>(d, fmt, p);
        xb[k] = nx;
        oeb[k] = rexp(nx);
      }
      continue;
    }
    double den = denom[pp];
    for (int64_t q = p; q < e; ++q) {
      const int k = colRow(d, fmt, q, b);
      const int pq = pid[k];
      if (pq >= N) continue;
      if (pq != pp) break;
      const R nx = xb[k] + rd * colValue<R>(d, fmt, q);
      xb[k] = nx;
      const double w = useW ? kw[k] : 1.0;
      const double oldE = useW ? w * (double)oeb[k] : (double)oeb[k];
      const R o = rexp(nx);
      oeb[k] = o;
      const double newE = useW ? w * (double)o : (double)o;
      den += newE - oldE;
    }
    denom[pp] = den;
  }
}

enum RowMode : int {
  RR_GRADOBJ = 0,
  RR_LL_LR = 1,
  RR_LL_POISSON = 2,
  RR_LL_NUM = 3,       // grouped: numerator term only
  RR_LL_NUM_SURV = 4,  // Cox numerator term
  RR_PRED_LR = 5,
  RR_PRED_POISSON = 6,
  RR_PRED_GROUPED = 7, // CLR (group = stratum) and Cox (group = pid, acc denominators)
  RR_PRED_COX = 8,
};

// Row-level sums over K rows: out[0], out[1] (see RowMode)
template <typename R>
__global__ void k_rows_reduce(int mode, int K, const R* xb, const R* rden, const double* y,
                              const double* kw, int useW, const double* w2,
                              const int32_t* group, const double* gden, double* partials,
                              unsigned* ticket, double* out) {
  double v[2] = {0.0, 0.0};
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < K;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double x = (double)xb[k];
    const double yy = y[k];
    const double w = useW ? kw[k] : 1.0;
    switch (mode) {
      case RR_GRADOBJ:
        v[0] += useW ? x * yy * w : x * yy;
        break;
      case RR_LL_LR: {
        const int yi = (int)yy;
        v[0] += useW ? (yi * x) * w : yi * x;
        v[1] += w * log((double)rden[k]);
        break;
      }
      case RR_LL_POISSON: {
        const int yi = (int)yy;
        v[0] += useW ? (yi * x) * w : yi * x;
        v[1] += w * (double)rden[k];
        break;
      }
      case RR_LL_NUM: {
        const int yi = (int)yy;
        v[0] += useW ? (yi * x) * w : yi * x;
        break;
      }
      case RR_LL_NUM_SURV: {
        const int yi = (int)yy;
        const double c = yi != 1 ? 0.0 : x;
        v[0] += useW ? c * w : c;
        break;
      }
      case RR_PRED_LR:
        v[0] += yy * w2[k] * (x - log((double)rden[k]));
        break;
      case RR_PRED_POISSON:
        v[0] += (yy * x - exp(x)) * w2[k];
        break;
      case RR_PRED_GROUPED:
        v[0] += yy * w2[k] * (x - log(gden[group[k]]));
        break;
      case RR_PRED_COX:
        v[0] += w2[k] == 0.0 ? 0.0 : yy * w2[k] * (x - log(gden[group[k]]));
        break;
    }
  }
  blockSum<2>(v);
  gridFinish<2>(v, partials, ticket, out);
}

// sum over groups of nW * log(denominator) (grouped log likelihood)
__global__ void k_groups_log(int N, const double* nw, const double* denom, double* partials,
                             unsigned* ticket, double* out) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    v[0] += nw[i] * log(denom[i]);
  blockSum<2>(v);
  gridFinish<2>(v, partials, ticket, out);
}

// ------------------------------------------- transpose and the MM step
// The row-major copy follows CompressedDataMatrix::transpose
// (CompressedDataMatrix.cpp:151-208): with INDICATOR columns and no DENSE one
// every row is an INDICATOR row, whose entries count 1 whatever the column's
// value; otherwise entries keep their values.  Zero values are left out (an
// exact zero added to a row sum changes nothing for finite coefficients).

// entries per row; one workgroup per column (grid-stride over columns)
__global__ void k_t_count(DevX d, int32_t* cnt) {
  for (int j = blockIdx.x; j < d.J; j += gridDim.x) {
    const int fmt = d.fmt[j];
    const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
    for (int64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
      if (colValue<double>(d, fmt, p) == 0.0) continue;
      atomicAdd(&cnt[colRow(d, fmt, p, b)], 1);
    }
  }
}

// exclusive scan of cnt[0..K) into ptr[0..K], one workgroup: each thread sums a
// contiguous run, the run totals are scanned in LDS, then each run is written
__global__ void k_scan_rows(int K, const int32_t* cnt, int64_t* ptr) {
  __shared__ int64_t part[kMaxBlockThreads];
  const int T = blockDim.x, t = threadIdx.x;
  const int64_t per = ((int64_t)K + T - 1) / T;
  const int64_t b = std::min<int64_t>(K, t * per), e = std::min<int64_t>(K, b + per);
  int64_t s = 0;
  for (int64_t i = b; i < e; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t acc = 0;
    for (int i = 0; i < T; ++i) {
      const int64_t v = part[i];
      part[i] = acc;
      acc += v;
    }
    ptr[K] = acc;
  }
  __syncthreads();
  int64_t acc = part[t];
  for (int64_t i = b; i < e; ++i) {
    ptr[i] = acc;
    acc += cnt[i];
  }
}

// scatter every entry to its row (slot order within a row is arbitrary here)
__global__ void k_t_fill(DevX d, unsigned long long* cursor, int32_t* tCol, double* tVal) {
  for (int j = blockIdx.x; j < d.J; j += gridDim.x) {
    const int fmt = d.fmt[j];
    const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
    for (int64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
      const double v = colValue<double>(d, fmt, p);
      if (v == 0.0) continue;
      const unsigned long long slot = atomicAdd(&cursor[colRow(d, fmt, p, b)], 1ull);
      tCol[slot] = j;
      if (tVal) tVal[slot] = v;
    }
  }
}

// sort each row's entries by column (rows are short: insertion sort per thread)
__global__ void k_t_sort(int K, const int64_t* ptr, int32_t* tCol, double* tVal) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < K;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = ptr[k], e = ptr[k + 1];
    for (int64_t q = b + 1; q < e; ++q) {
      const int32_t c = tCol[q];
      const double v = tVal ? tVal[q] : 0.0;
      int64_t r = q - 1;
      while (r >= b && tCol[r] > c) {
        tCol[r + 1] = tCol[r];
        if (tVal) tVal[r + 1] = tVal[r];
        --r;
      }
      tCol[r + 1] = c;
      if (tVal) tVal[r + 1] = v;
    }
  }
}

// ModelSpecifics::computeXBetaImpl (ModelSpecifics.hpp:346-364): each row's sum
// over its entries in column order.  A workgroup takes kThreads consecutive rows,
// whose entries are contiguous; it stages them through LDS in tiles with
// coalesced loads, and each thread adds its own row's part of every tile in
// order, so a row is still summed entry by entry in column order.
constexpr int kXbTile = 2048;
template <typename R>
__global__ __launch_bounds__(kThreads) void k_xbeta_rows(int K, const int64_t* ptr,
                                                         const int32_t* tCol, const double* tVal,
                                                         const double* beta, R* xb) {
  __shared__ double sv[kXbTile], sb[kXbTile];
  for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x; k0 < K;
       k0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = k0 + threadIdx.x;
    const int64_t e0 = ptr[k0], e1 = ptr[std::min<int64_t>(K, k0 + blockDim.x)];
    const int64_t rb = k < K ? ptr[k] : e1, re = k < K ? ptr[k + 1] : e1;
    double sum = 0.0;
    for (int64_t t0 = e0; t0 < e1; t0 += kXbTile) {
      const int64_t t1 = std::min<int64_t>(e1, t0 + kXbTile);
      __syncthreads();  // the previous tile is consumed
      for (int64_t q = t0 + threadIdx.x; q < t1; q += blockDim.x) {
        sv[q - t0] = tVal ? tVal[q] : 1.0;
        sb[q - t0] = beta[tCol[q]];
      }
      __syncthreads();
      const int64_t a = std::max(rb, t0), b = std::min(re, t1);
      for (int64_t q = a; q < b; ++q) sum += sv[q - t0] * sb[q - t0];
    }
    if (k < K) xb[k] = (R)sum;
  }
}

// MM increments (ModelSpecifics.h:500-517, 571-584) factor into a per-row value
// r_k = nW_g * oeb_k / denom_g over the row's group g (logistic: nW_k = w_k and
// denom_k = 1 + oeb_k; conditional logistic: the stratum; Cox: the row's pid,
// whose denominator is the pid's own sum -- a tie group of events or one row --
// as computeMMGradientAndHessianImpl reads denomPid, not the accumulated risk
// set; zero-weight rows read slot N, whose event weight is 0); a column's
// gradient is sum r_k x_k and its Hessian sum r_k x_k^2.
template <typename R>
__global__ void k_mm_rows(int model, int K, const R* oeb, const double* kw, const double* nw,
                          const double* denom, const int32_t* group, double* r) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < K;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double o = (double)oeb[k];
    if (model == M_LOGISTIC) {
      r[k] = kw[k] * o / (1.0 + o);
    } else {
      const int s = group[k];
      r[k] = nw[s] * o / denom[s];
    }
  }
}

constexpr int kMmChunk = 2048;  // nonzeros of one column chunk (one wave)

// one wave per column chunk: partial (gradient, Hessian) sums
__global__ void k_mm_chunks(DevX d, const double* r, const int32_t* chunk0,
                            const int32_t* chunkCol, int nChunks, const uint8_t* fixed,
                            double* part) {
  const int lane = threadIdx.x & (kWave - 1);
  const int c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (c >= nChunks) return;  // uniform over the wave
  const int j = chunkCol[c];
  double g = 0.0, h = 0.0;
  if (!fixed[j]) {
    const int fmt = d.fmt[j];
    const int64_t cb = d.colPtr[j];
    const int64_t b = cb + (int64_t)(c - chunk0[j]) * kMmChunk;
    const int64_t e = std::min<int64_t>(d.colPtr[j + 1], b + kMmChunk);
    if (isIndicator(fmt)) {
      for (int64_t p = b + lane; p < e; p += kWave) g += r[colRow(d, fmt, p, cb)];
      h = g;
    } else {
      for (int64_t p = b + lane; p < e; p += kWave) {
        const double t = r[colRow(d, fmt, p, cb)], x = d.vals[p];
        g += t * x;
        h += t * x * x;
      }
    }
  }
  g = waveSum(g);
  h = waveSum(h);
  if (lane == 0) {
    part[2 * c] = g;
    part[2 * c + 1] = h;
  }
}

// one wave per column: its chunk sums (lane-strided, then a fixed wave tree),
// minus XjY; fixed columns (0, 0)
__global__ void k_mm_finish(int J, const int32_t* chunk0, const double* part, const double* xjy,
                            const uint8_t* fixed, double* gh) {
  const int lane = threadIdx.x & (kWave - 1);
  const int waves = blockDim.x / kWave;
  for (int j = blockIdx.x * waves + threadIdx.x / kWave; j < J; j += gridDim.x * waves) {
    const bool fx = fixed[j] != 0;
    double g = 0.0, h = 0.0;
    if (!fx) {
      for (int c = chunk0[j] + lane; c < chunk0[j + 1]; c += kWave) {
        g += part[2 * c];
        h += part[2 * c + 1];
      }
    }
    g = waveSum(g);
    h = waveSum(h);
    if (lane == 0) {
      gh[2 * j] = fx ? 0.0 : g - xjy[j];
      gh[2 * j + 1] = fx ? 0.0 : h;
    }
  }
}

}  // namespace

namespace ccdamd {
void dataDestroy(ccdh_data* d) {
  if (!d) return;
  void* ptrs[] = {d->colPtr, d->rows, d->vals, d->valsF, d->fmt, d->y, d->time, d->strata,
                  d->segPtr, d->segOff, d->segKey, d->stratStart, d->nzKey, d->shardRow,
                  d->colShard, d->blk};
  if (!d->placed)
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  void* lazy[] = {d->tRowPtr, d->tCol, d->tVal, d->mmChunk0, d->mmChunkCol};
  for (void* p : lazy)
    if (p) (void)hipFree(p);
  delete d;
}
}  // namespace ccdamd

// ================================================================ host side
DevX ccdh_data::view() const {
  DevX x;
  x.K = K;
  x.J = J;
  x.nStrata = nStrata;
  x.colPtr = colPtr;
  x.rows = rows;
  x.vals = vals;
  x.valsF = valsF;
  x.fmt = fmt;
  x.y = y;
  x.time = time;
  x.strata = strata;
  x.segPtr = segPtr;
  x.segOff = segOff;
  x.segKey = segKey;
  x.stratStart = stratStart;
  x.nzKey = nzKey;
  x.nShards = kShards;
  x.shardRow = shardRow;
  x.colShard = colShard;
  x.blk = blk;
  return x;
}

namespace ccdamd {

// Device arrays come from hipMalloc, or are carved from `place` (a block the
// caller owns, e.g. a served request's arena block).
template <typename T>
static int devAlloc(T** p, size_t n, int64_t* acc, Bump* place = nullptr) {
  *p = nullptr;
  if (n == 0) n = 1;
  if (place) {
    *p = (T*)place->take(n * sizeof(T));
    if (!*p) return setError(CCDH_E_NOMEM, "data does not fit its device block");
  } else {
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess)
      return setError(CCDH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  if (acc) *acc += (int64_t)(n * sizeof(T));
  return 0;
}

template <typename T>
static int devUpload(T** p, const T* h, size_t n, int64_t* acc, Bump* place) {
  int rc = devAlloc(p, n, acc, place);
  if (rc) return rc;
  if (n && h) HIP_TRY(hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Device bytes dataCreate will carve for this description when placed in a
// block (an upper bound: stratum runs are counted as one per nonzero).
int64_t dataFootprint(const ccdh_data_desc* desc) {
  const int64_t K = desc->K, J = desc->J, nnz = desc->col_ptr[J];
  bool needVals = false;
  for (int j = 0; j < J; ++j) needVals |= desc->fmt[j] == CCDH_SPARSE || desc->fmt[j] == CCDH_DENSE;
  int64_t nStrata = 1;
  if (desc->strata)
    for (int64_t k = 0; k < K; ++k) nStrata = std::max<int64_t>(nStrata, desc->strata[k] + 1);
  auto r = [](int64_t b) { return (int64_t)Arena::round((size_t)std::max<int64_t>(b, 1)); };
  int64_t t = r((J + 1) * 8) + r(nnz * 4) + r(J) + r(K * 8) + r(K * 4);
  if (needVals) t += r(nnz * 8) + (desc->precision == CCDH_MIXED ? r(nnz * 4) : 0);
  if (desc->model == CCDH_CONDITIONAL_LOGISTIC)
    t += r(nnz * 4) + r((J + 1) * 8) + 2 * r(nnz * 4) + r((nStrata + 1) * 4);
  t += r((kShards + 1) * 4) + r(J * (int64_t)(kShards + 1) * 4);
  if (desc->model != CCDH_COX) t += r(J);
  if (desc->model == CCDH_COX) t += r(K * 8);  // times
  return t;
}

// Column blocks (see DevX::blk): greedy from each block start, a column joins
// while it is short (1..kBlockNnz nonzeros, not dense) and touches no row
// (CLR: stratum) of the block's earlier columns, up to kBlockCols.  Updating a
// column then cannot change another block column's gradient, so the block's
// columns may share one gather, reduction and exchange and still give the
// sequential results.
static void buildBlocks(const ccdh_data_desc* desc, const std::vector<int32_t>& strata,
                        std::vector<int8_t>& blk) {
  const int J = desc->J, K = desc->K;
  const bool clr = desc->model == CCDH_CONDITIONAL_LOGISTIC;
  blk.assign(J, 1);
  int nUnit = K;
  if (clr)
    for (int k = 0; k < K; ++k) nUnit = std::max(nUnit, strata[k] + 1);
  std::vector<int32_t> stamp(nUnit, -1);
  auto isShort = [&](int c) {
    const int64_t n = desc->col_ptr[c + 1] - desc->col_ptr[c];
    return desc->fmt[c] != CCDH_DENSE && desc->fmt[c] != CCDH_INTERCEPT && n > 0 && n <= kBlockNnz;
  };
  auto unit = [&](int64_t p) { return clr ? strata[desc->rows[p]] : desc->rows[p]; };
  for (int j = 0; j < J;) {
    int B = 0;
    for (int c = j; c < J && B < kBlockCols && isShort(c); ++c) {
      bool clash = false;
      for (int64_t p = desc->col_ptr[c]; p < desc->col_ptr[c + 1] && !clash; ++p)
        clash = stamp[unit(p)] == j;
      if (clash) break;
      for (int64_t p = desc->col_ptr[c]; p < desc->col_ptr[c + 1]; ++p) stamp[unit(p)] = j;
      ++B;
    }
    if (B < 2) {
      ++j;
      continue;
    }
    blk[j] = (int8_t)B;
    for (int c = j + 1; c < j + B; ++c) blk[c] = 0;
    j += B;
  }
}

// Row shards for workgroup gangs: kShards near-equal row ranges whose
// boundaries start a stratum (CLR) or a tie group of events (Cox: consecutive
// events of one stratum at one time share a risk-set group under any weights),
// and per column the offset of each shard's first nonzero.
static void buildShards(const ccdh_data_desc* desc, const std::vector<int32_t>& strata,
                        std::vector<int32_t>& shardRow, std::vector<int32_t>& colShard) {
  const int K = desc->K, J = desc->J, S = kShards;
  auto startsGroup = [&](int k) {
    if (k <= 0 || k >= K) return true;
    if (desc->model == CCDH_CONDITIONAL_LOGISTIC) return strata[k] != strata[k - 1];
    if (desc->model == CCDH_COX)
      return strata[k] != strata[k - 1] ||
             !(desc->y[k] == 1.0 && desc->y[k - 1] == 1.0 && desc->offs[k] == desc->offs[k - 1]);
    return true;
  };
  shardRow.assign(S + 1, K);
  shardRow[0] = 0;
  for (int s = 1; s < S; ++s) {
    int k = std::max((int)((int64_t)K * s / S), shardRow[s - 1]);
    while (k < K && !startsGroup(k)) ++k;
    shardRow[s] = k;
  }
  colShard.assign((size_t)J * (S + 1), 0);
  for (int j = 0; j < J; ++j) {
    const int64_t b = desc->col_ptr[j], e = desc->col_ptr[j + 1];
    int32_t* out = &colShard[(size_t)j * (S + 1)];
    const int f = desc->fmt[j];
    if (f == CCDH_DENSE || f == CCDH_INTERCEPT) {
      for (int s = 0; s <= S; ++s) out[s] = shardRow[s];
      continue;
    }
    int64_t p = b;
    for (int s = 0; s <= S; ++s) {
      while (p < e && desc->rows[p] < shardRow[s]) ++p;
      out[s] = (int32_t)(p - b);
    }
  }
}

int dataCreate(const ccdh_data_desc* desc, ccdh_data** out, Bump* place) {
  if (!desc || !out) return setError(CCDH_E_INVALID, "null argument");
  const int model = desc->model;
  if (model != CCDH_POISSON && model != CCDH_LOGISTIC && model != CCDH_CONDITIONAL_LOGISTIC &&
      model != CCDH_COX)
    return setError(CCDH_E_INVALID, "unsupported model type " + std::to_string(model));
  if (desc->precision != CCDH_FP64 && desc->precision != CCDH_MIXED)
    return setError(CCDH_E_INVALID, "unsupported precision");
  if (desc->K <= 0 || desc->J <= 0 || !desc->col_ptr || !desc->fmt || !desc->y)
    return setError(CCDH_E_INVALID, "empty problem");
  const int K = desc->K, J = desc->J;
  const int64_t nnz = desc->col_ptr[J];
  if (nnz >= (int64_t)INT32_MAX) return setError(CCDH_E_INVALID, "nnz must be < 2^31");
  bool needVals = false;
  for (int j = 0; j < J; ++j) {
    const int f = desc->fmt[j];
    if (f < 0 || f > 3) return setError(CCDH_E_INVALID, "bad format tag");
    const int64_t n = desc->col_ptr[j + 1] - desc->col_ptr[j];
    if (n < 0) return setError(CCDH_E_INVALID, "col_ptr not monotone");
    if ((f == CCDH_DENSE || f == CCDH_INTERCEPT) && n != K)
      return setError(CCDH_E_INVALID, "dense/intercept column must span K rows");
    if (f == CCDH_SPARSE || f == CCDH_DENSE) needVals = true;
    if ((f == CCDH_SPARSE || f == CCDH_INDICATOR)) {
      for (int64_t p = desc->col_ptr[j]; p < desc->col_ptr[j + 1]; ++p) {
        const int r = desc->rows[p];
        if (r < 0 || r >= K) return setError(CCDH_E_INVALID, "row index out of range");
        if (p > desc->col_ptr[j] && desc->rows[p - 1] >= r)
          return setError(CCDH_E_INVALID, "row indices must increase within a column");
      }
    }
  }
  if (needVals && !desc->vals) return setError(CCDH_E_INVALID, "values required");
  if (model == CCDH_COX && !desc->offs) return setError(CCDH_E_INVALID, "Cox needs times");
  if (desc->has_offset && desc->fmt[0] != CCDH_DENSE)
    return setError(CCDH_E_INVALID, "offset must be a DENSE column 0");

  auto* d = new ccdh_data();
  d->placed = place != nullptr;
  d->model = model;
  d->precision = desc->precision;
  d->K = K;
  d->J = J;
  d->nnz = nnz;
  d->hasOffset = desc->has_offset;
  d->hColPtr.assign(desc->col_ptr, desc->col_ptr + J + 1);
  d->hFmt.assign(desc->fmt, desc->fmt + J);
  d->hY.assign(desc->y, desc->y + K);
  d->hOffs.assign(K, 0.0);
  if (desc->offs) d->hOffs.assign(desc->offs, desc->offs + K);
  d->hStrata.assign(K, 0);
  if (desc->strata) d->hStrata.assign(desc->strata, desc->strata + K);
  int nStrata = 0;
  for (int k = 0; k < K; ++k) {
    if (d->hStrata[k] < 0) {
      delete d;
      return setError(CCDH_E_INVALID, "negative stratum id");
    }
    if (k && d->hStrata[k] < d->hStrata[k - 1] &&
        (model == CCDH_CONDITIONAL_LOGISTIC || model == CCDH_COX)) {
      delete d;
      return setError(CCDH_E_INVALID, "rows must be sorted by stratum");
    }
    nStrata = std::max(nStrata, d->hStrata[k] + 1);
  }
  d->nStrata = nStrata;
  // Poisson fixed term per row: -log(y!) as the loop of ModelSpecifics.h:1596-1601
  if (model == CCDH_POISSON) {
    d->hLogFact.resize(K);
    for (int k = 0; k < K; ++k) {
      double t = 0.0;
      for (int i = 2; i <= (int)d->hY[k]; ++i) t += -std::log((double)i);
      d->hLogFact[k] = t;
    }
  }

  int64_t& acc = d->devBytes;
  int rc = 0;
  rc = rc ? rc : devUpload(&d->colPtr, desc->col_ptr, J + 1, &acc, place);
  std::vector<int32_t> rowsFull;
  const int32_t* rowsSrc = desc->rows;
  rc = rc ? rc : devUpload(&d->rows, rowsSrc, (size_t)nnz, &acc, place);
  if (!rc && needVals) {
    rc = devUpload(&d->vals, desc->vals, (size_t)nnz, &acc, place);
    if (!rc && d->precision == CCDH_MIXED) {
      std::vector<float> vf(desc->vals, desc->vals + nnz);
      rc = devUpload(&d->valsF, vf.data(), (size_t)nnz, &acc, place);
    }
  }
  rc = rc ? rc : devUpload(&d->fmt, desc->fmt, J, &acc, place);
  rc = rc ? rc : devUpload(&d->y, desc->y, K, &acc, place);
  if (model == CCDH_COX) rc = rc ? rc : devUpload(&d->time, desc->offs, K, &acc, place);
  rc = rc ? rc : devUpload(&d->strata, d->hStrata.data(), K, &acc, place);
  if (!rc && model == CCDH_CONDITIONAL_LOGISTIC) {
    // runs of equal stratum inside every column (rows are sorted by stratum)
    std::vector<int64_t> segPtr(J + 1, 0);
    std::vector<int32_t> segOff, segKey;
    segOff.reserve(nnz);
    segKey.reserve(nnz);
    for (int j = 0; j < J; ++j) {
      const int64_t b = desc->col_ptr[j], e = desc->col_ptr[j + 1];
      const int f = desc->fmt[j];
      int last = -1;
      for (int64_t p = b; p < e; ++p) {
        const int k = (f == CCDH_DENSE || f == CCDH_INTERCEPT) ? (int)(p - b) : desc->rows[p];
        const int s = d->hStrata[k];
        if (s != last) {
          segOff.push_back((int32_t)(p - b));
          segKey.push_back(s);
          last = s;
        }
      }
      segPtr[j + 1] = (int64_t)segOff.size();
    }
    // stratum of every nonzero, stored beside the row index for one-hop gathers
    {
      std::vector<int32_t> key((size_t)nnz);
      for (int j = 0; j < J; ++j) {
        const int64_t b = desc->col_ptr[j], e = desc->col_ptr[j + 1];
        const int f = desc->fmt[j];
        for (int64_t p = b; p < e; ++p) {
          const int k = (f == CCDH_DENSE || f == CCDH_INTERCEPT) ? (int)(p - b) : desc->rows[p];
          key[p] = d->hStrata[k];
        }
      }
      rc = devUpload(&d->nzKey, key.data(), key.size(), &acc, place);
      if (rc) {
        dataDestroy(d);
        return rc;
      }
    }
    std::vector<int32_t> ss(nStrata + 1, 0);
    for (int k = 0; k < K; ++k) ss[d->hStrata[k] + 1]++;
    for (int s = 0; s < nStrata; ++s) ss[s + 1] += ss[s];
    d->hSegPtr = segPtr;
    rc = devUpload(&d->segPtr, segPtr.data(), J + 1, &acc, place);
    rc = rc ? rc : devUpload(&d->segOff, segOff.data(), segOff.size(), &acc, place);
    rc = rc ? rc : devUpload(&d->segKey, segKey.data(), segKey.size(), &acc, place);
    rc = rc ? rc : devUpload(&d->stratStart, ss.data(), ss.size(), &acc, place);
    d->nSeg = (int64_t)segOff.size();
  }
  if (!rc) {
    std::vector<int32_t> colShard;
    buildShards(desc, d->hStrata, d->hShardRow, colShard);
    rc = devUpload(&d->shardRow, d->hShardRow.data(), d->hShardRow.size(), &acc, place);
    rc = rc ? rc : devUpload(&d->colShard, colShard.data(), colShard.size(), &acc, place);
    if (!rc && model != CCDH_COX) {
      std::vector<int8_t> blk;
      buildBlocks(desc, d->hStrata, blk);
      rc = devUpload(&d->blk, blk.data(), blk.size(), &acc, place);
    }
  }
  if (rc) {
    dataDestroy(d);
    return rc;
  }
  *out = d;
  return 0;
}



// ---------------------------------------------------------------- fit state
static int ensureGroups(ccdh_fit* f, int N) {
  if (N + 1 <= f->nCap) return 0;
  const int cap = N + 1 + 64;
  double* bufs[4] = {f->denom, f->nw, f->numer, f->numer2};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (f->resetFlag) (void)hipFree(f->resetFlag);
  if (f->pidFirst) (void)hipFree(f->pidFirst);
  if (f->tileAgg) (void)hipFree(f->tileAgg);
  if (f->tileCarry) (void)hipFree(f->tileCarry);
  if (f->accTmp) (void)hipFree(f->accTmp);
  f->denom = f->nw = f->numer = f->numer2 = f->accTmp = nullptr;
  f->resetFlag = nullptr;
  f->pidFirst = nullptr;
  f->tileAgg = f->tileCarry = nullptr;
  int rc = 0;
  rc = rc ? rc : devAlloc(&f->denom, cap, nullptr);
  rc = rc ? rc : devAlloc(&f->nw, cap, nullptr);
  if (f->d->model == CCDH_COX) {
    rc = rc ? rc : devAlloc(&f->numer, cap, nullptr);
    rc = rc ? rc : devAlloc(&f->numer2, cap, nullptr);
    rc = rc ? rc : devAlloc(&f->accTmp, cap, nullptr);
    rc = rc ? rc : devAlloc(&f->resetFlag, cap, nullptr);
    rc = rc ? rc : devAlloc(&f->pidFirst, cap, nullptr);
    const int T = (cap + kCoxTile - 1) / kCoxTile;
    rc = rc ? rc : devAlloc(&f->tileAgg, T, nullptr);
    rc = rc ? rc : devAlloc(&f->tileCarry, T, nullptr);
    if (!rc) {
      HIP_TRY(hipMemsetAsync(f->numer, 0, cap * sizeof(double), f->stream));
      HIP_TRY(hipMemsetAsync(f->numer2, 0, cap * sizeof(double), f->stream));
    }
  }
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(f->denom, 0, cap * sizeof(double), f->stream));
  HIP_TRY(hipMemsetAsync(f->nw, 0, cap * sizeof(double), f->stream));
  f->nCap = cap;
  return 0;
}

// Cox risk-set grouping on the host (ModelSpecifics.hpp:1809-1868): rows are
// sorted by (stratum, -time, y); consecutive tied events share a pid, a new
// stratum starts a new pid and a reset point, zero-weight rows get pid N.
void coxPids(const ccdh_data* d, const double* w, std::vector<int32_t>& pid, int* Nout,
             std::vector<uint8_t>& reset, std::vector<int32_t>& pidFirst) {
  const int K = d->K;
  pid.assign(d->hStrata.begin(), d->hStrata.end());
  std::vector<int> accReset;
  const int ignore = -1;
  int index = 0;
  while (w != nullptr && index < K && w[index] == 0.0) {
    pid[index] = ignore;
    index++;
  }
  int lastPid = pid[index];
  double lastTime = d->hOffs[index];
  double lastEvent = d->hY[index];
  int p = pid[index] = 0;
  for (int k = index + 1; k < K; ++k) {
    if (w == nullptr || w[k] != 0.0) {
      const int nextPid = pid[k];
      if (nextPid != lastPid) {
        p++;
        accReset.push_back(p);
        lastPid = nextPid;
      } else if (!(lastEvent == 1.0 && lastTime == d->hOffs[k] && lastEvent == d->hY[k])) {
        p++;
      }
      lastTime = d->hOffs[k];
      lastEvent = d->hY[k];
      pid[k] = p;
    } else {
      pid[k] = ignore;
    }
  }
  p++;
  const int N = p;
  for (int k = 0; k < K; ++k)
    if (pid[k] == ignore) pid[k] = N;
  reset.assign(N + 1, 0);
  for (int r : accReset)
    if (r < N) reset[r] = 1;
  pidFirst.assign(N + 1, K);
  for (int k = K - 1; k >= 0; --k)
    if (pid[k] < N) pidFirst[pid[k]] = k;
  for (int i = N - 1; i >= 0; --i) pidFirst[i] = std::min(pidFirst[i], pidFirst[i + 1]);
  *Nout = N;
}

int fitCreate(ccdh_data* d, ccdh_fit** out) {
  auto* f = new ccdh_fit();
  f->d = d;
  f->mixed = d->precision == CCDH_MIXED;
  const size_t rs = f->mixed ? sizeof(float) : sizeof(double);
  int rc = 0;
  HIP_TRY(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
  HIP_TRY(hipMalloc(&f->xb, (size_t)d->K * rs));
  HIP_TRY(hipMalloc(&f->oeb, (size_t)d->K * rs));
  HIP_TRY(hipMemsetAsync(f->xb, 0, (size_t)d->K * rs, f->stream));
  HIP_TRY(hipMemsetAsync(f->oeb, 0, (size_t)d->K * rs, f->stream));
  if (d->model == CCDH_LOGISTIC || d->model == CCDH_POISSON) {
    HIP_TRY(hipMalloc(&f->rden, (size_t)d->K * rs));
    HIP_TRY(hipMemsetAsync(f->rden, 0, (size_t)d->K * rs, f->stream));
  }
  rc = rc ? rc : devAlloc(&f->kw, d->K, nullptr);
  rc = rc ? rc : devAlloc(&f->w2, d->K, nullptr);
  rc = rc ? rc : devAlloc(&f->xjy, d->J, nullptr);
  rc = rc ? rc : devAlloc(&f->partials, (size_t)kMaxGridBlocks * 2, nullptr);
  rc = rc ? rc : devAlloc(&f->ticket, 4, nullptr);
  rc = rc ? rc : devAlloc(&f->dres, 8, nullptr);
  if (rc) {
    fitDestroy(f);
    return rc;
  }
  HIP_TRY(hipMemsetAsync(f->ticket, 0, 4 * sizeof(unsigned), f->stream));
  HIP_TRY(hipMemsetAsync(f->xjy, 0, d->J * sizeof(double), f->stream));
  HIP_TRY(hipHostMalloc((void**)&f->hres, 8 * sizeof(double), hipHostMallocDefault));
  if (d->model == CCDH_COX) {
    rc = devAlloc(&f->pid, d->K, nullptr);
    if (rc) {
      fitDestroy(f);
      return rc;
    }
  }
  // initialize(): N groups and Cox pids without weights (ModelSpecifics.hpp:1899-1952)
  f->N = d->model == CCDH_CONDITIONAL_LOGISTIC ? d->nStrata : d->K;
  f->hKw.assign(d->K, 1.0);
  if (d->model == CCDH_COX) {
    rc = fitSetPids(f, nullptr);
  } else if (d->model == CCDH_CONDITIONAL_LOGISTIC) {
    rc = ensureGroups(f, f->N);
  }
  if (!rc) rc = launchFill(f, f->kw, d->K, 1.0);
  if (rc) {
    fitDestroy(f);
    return rc;
  }
  *out = f;
  return 0;
}

void fitDestroy(ccdh_fit* f) {
  if (!f) return;
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  void* ptrs[] = {f->xb,  f->oeb,  f->rden,  f->kw,      f->w2,      f->xjy,     f->partials,
                  f->ticket, f->dres,   f->denom,   f->nw,      f->numer,   f->numer2,
                  f->accTmp, f->pid,    f->pidFirst, f->resetFlag, f->tileAgg, f->tileCarry,
                  f->mmR,    f->mmPart, f->mmGH,     f->mmBeta,    f->mmFixed};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (f->hres) (void)hipHostFree(f->hres);
  for (hipEvent_t e : f->mmEv)
    if (e) (void)hipEventDestroy(e);
  if (f->stream) (void)hipStreamDestroy(f->stream);
  delete f;
}

int launchFill(ccdh_fit* f, double* p, int64_t n, double v) {
  k_fill<double><<<gridFor(n), kThreads, 0, f->stream>>>(p, n, v);
  HIP_TRY(hipGetLastError());
  return 0;
}

int fitSetPids(ccdh_fit* f, const double* w) {
  std::vector<int32_t> pid, pidFirst;
  std::vector<uint8_t> reset;
  int N = 0;
  coxPids(f->d, w, pid, &N, reset, pidFirst);
  int rc = ensureGroups(f, N);
  if (rc) return rc;
  f->N = N;
  HIP_TRY(hipMemcpyAsync(f->pid, pid.data(), pid.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                         f->stream));
  HIP_TRY(hipMemcpyAsync(f->pidFirst, pidFirst.data(), pidFirst.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, f->stream));
  HIP_TRY(hipMemcpyAsync(f->resetFlag, reset.data(), reset.size(), hipMemcpyHostToDevice,
                         f->stream));
  HIP_TRY(hipStreamSynchronize(f->stream));
  return 0;
}

static int fetch(ccdh_fit* f, int n) {
  HIP_TRY(hipMemcpyAsync(f->hres, f->dres, n * sizeof(double), hipMemcpyDeviceToHost, f->stream));
  HIP_TRY(hipStreamSynchronize(f->stream));
  return 0;
}

#define LAUNCH_R(kern, grid, ...)                                                   \
  do {                                                                              \
    if (f->mixed) {                                                                 \
      using T_ = float;                                                             \
      kern<T_><<<(grid), kThreads, 0, f->stream>>>(__VA_ARGS__);                    \
    } else {                                                                        \
      using T_ = double;                                                            \
      kern<T_><<<(grid), kThreads, 0, f->stream>>>(__VA_ARGS__);                    \
    }                                                                               \
    HIP_TRY(hipGetLastError());                                                     \
  } while (0)

#define XB(T) ((T*)f->xb)
#define OEB(T) ((T*)f->oeb)
#define RDEN(T) ((T*)f->rden)

int setWeights(ccdh_fit* f, const double* w, int useCV) {
  ccdh_data* d = f->d;
  if (useCV && !w) return setError(CCDH_E_INVALID, "weights required with use_cv");
  if (useCV) {
    f->hKw.assign(w, w + d->K);
    HIP_TRY(hipMemcpyAsync(f->kw, w, d->K * sizeof(double), hipMemcpyHostToDevice, f->stream));
  } else {
    f->hKw.assign(d->K, 1.0);
    int rc = launchFill(f, f->kw, d->K, 1.0);
    if (rc) return rc;
  }
  if (d->model == CCDH_COX) {
    int rc = fitSetPids(f, useCV ? w : nullptr);
    if (rc) return rc;
    k_nw_cox<<<gridFor(f->N, 1), kThreads, 0, f->stream>>>(f->N, d->y, f->pid, f->pidFirst,
                                                            f->kw, f->nw);
    HIP_TRY(hipGetLastError());
    // the slot of the zero-weight rows (pid N) holds no events, as setWeights fills
    // all N + 1 event weights before summing (ModelSpecifics.hpp:447-455)
    rc = launchFill(f, f->nw + f->N, 1, 0.0);
    if (rc) return rc;
  } else if (d->model == CCDH_CONDITIONAL_LOGISTIC) {
    k_nw_clr<<<gridFor(d->nStrata, 1), kThreads, 0, f->stream>>>(d->view(), f->kw, f->nw);
    HIP_TRY(hipGetLastError());
  }
  f->weightsSet = true;
  return 0;
}

int fixedTermsLogLik(ccdh_fit* f, int useCV) {
  f->fixedLL = 0.0;
  if (f->d->model != CCDH_POISSON) return 0;
  double s = 0.0;
  for (int k = 0; k < f->d->K; ++k)
    s += useCV ? f->d->hLogFact[k] * f->hKw[k] : f->d->hLogFact[k];
  f->fixedLL = s;
  return 0;
}

int fixedTermsGradHess(ccdh_fit* f, int useCV) {
  k_xjy<<<f->d->J, kThreads, 0, f->stream>>>(f->d->view(), f->kw, useCV, f->xjy);
  HIP_TRY(hipGetLastError());
  return 0;
}

int zeroXBeta(ccdh_fit* f) {
  HIP_TRY(hipMemsetAsync(f->xb, 0, (size_t)f->d->K * (f->mixed ? 4 : 8), f->stream));
  return 0;
}

int axpyXBeta(ccdh_fit* f, double beta, int j) {
  if (j < 0 || j >= f->d->J) return setError(CCDH_E_INVALID, "column out of range");
  if (beta == 0.0) return 0;
  const int64_t n = f->d->hColPtr[j + 1] - f->d->hColPtr[j];
  if (n == 0) return 0;
  const int fmt = f->d->hFmt[j];
  if (f->mixed)
    k_axpy<float><<<gridFor(n), kThreads, 0, f->stream>>>(f->d->view(), XB(float), j, fmt, beta);
  else
    k_axpy<double><<<gridFor(n), kThreads, 0, f->stream>>>(f->d->view(), XB(double), j, fmt, beta);
  HIP_TRY(hipGetLastError());
  return 0;
}

int remainingStats(ccdh_fit* f, int useW) {
  ccdh_data* d = f->d;
  const int K = d->K;
  if (d->model == CCDH_CONDITIONAL_LOGISTIC) {
    LAUNCH_R(k_remaining_clr, gridFor(d->nStrata, 1), d->view(), XB(T_), OEB(T_), f->denom);
  } else {
    LAUNCH_R(k_exp_rows, gridFor(K), K, XB(T_), OEB(T_), RDEN(T_),
             (T_)(d->model == CCDH_LOGISTIC ? 1 : 0));
    if (d->model == CCDH_COX) {
      LAUNCH_R(k_remaining_cox, gridFor(f->N, 1), f->N, f->pid, f->pidFirst, OEB(T_), f->kw, useW,
               f->denom);
    }
  }
  f->statsKnown = true;
  return 0;
}

int numerator(ccdh_fit* f, int j, int useW) {
  ccdh_data* d = f->d;
  if (d->model != CCDH_COX) return 0;
  const int64_t n = d->hColPtr[j + 1] - d->hColPtr[j];
  if (n == 0) return 0;
  const int fmt = d->hFmt[j];
  LAUNCH_R(k_cox_numer, gridFor(n), d->view(), OEB(T_), f->kw, f->pid, f->N, j, fmt, useW,
           f->numer, f->numer2);
  return 0;
}

// runs the three Cox tile passes; mode as k_cox_tile_use
static int coxTiles(ccdh_fit* f, int mode, int j, int fmtIndicator) {
  const int N = f->N;
  const int T = (N + kCoxTile - 1) / kCoxTile;
  const int nv = mode == 0 ? (fmtIndicator ? 2 : 3) : 1;
  k_cox_tile_agg<<<T, kThreads, 0, f->stream>>>(N, f->denom, f->numer, f->numer2, f->resetFlag,
                                                 nv, f->tileAgg);
  HIP_TRY(hipGetLastError());
  k_cox_tile_scan<<<1, 1024, 0, f->stream>>>(T, f->tileAgg, f->tileCarry);
  HIP_TRY(hipGetLastError());
  k_cox_tile_use<<<T, kThreads, 0, f->stream>>>(N, mode, fmtIndicator, f->denom, f->numer,
                                                 f->numer2, f->resetFlag, f->nw, f->tileCarry,
                                                 f->accTmp, j, f->xjy, f->partials, f->ticket,
                                                 f->dres);
  HIP_TRY(hipGetLastError());
  return 0;
}

int gradHessLaunch(ccdh_fit* f, int j, int useW) {
  ccdh_data* d = f->d;
  const int64_t n = d->hColPtr[j + 1] - d->hColPtr[j];
  const int fmt = d->hFmt[j];
  if (n == 0) {
    HIP_TRY(hipMemsetAsync(f->dres, 0, 2 * sizeof(double), f->stream));
    return 0;
  }
  switch (d->model) {
    case CCDH_LOGISTIC:
      if (f->mixed)
        k_gh_indep<M_LOGISTIC, float><<<gridFor(n), kThreads, 0, f->stream>>>(
            d->view(), OEB(float), RDEN(float), f->kw, j, fmt, useW, f->xjy, f->partials, f->ticket, f->dres);
      else
        k_gh_indep<M_LOGISTIC, double><<<gridFor(n), kThreads, 0, f->stream>>>(
            d->view(), OEB(double), RDEN(double), f->kw, j, fmt, useW, f->xjy, f->partials, f->ticket, f->dres);
      break;
    case CCDH_POISSON:
      if (f->mixed)
        k_gh_indep<M_POISSON, float><<<gridFor(n), kThreads, 0, f->stream>>>(
            d->view(), OEB(float), RDEN(float), f->kw, j, fmt, useW, f->xjy, f->partials, f->ticket, f->dres);
      else
        k_gh_indep<M_POISSON, double><<<gridFor(n), kThreads, 0, f->stream>>>(
            d->view(), OEB(double), RDEN(double), f->kw, j, fmt, useW, f->xjy, f->partials, f->ticket, f->dres);
      break;
    case CCDH_CONDITIONAL_LOGISTIC: {
      const int64_t ns = d->hSegCount(j);
      LAUNCH_R(k_gh_clr, gridFor(ns, 2), d->view(), OEB(T_), f->denom, f->nw, j, fmt, f->xjy,
               f->partials, f->ticket, f->dres);
      break;
    }
    case CCDH_COX: {
      int rc = coxTiles(f, 0, j, (fmt == CCDH_INDICATOR || fmt == CCDH_INTERCEPT) ? 1 : 0);
      if (rc) return rc;
      break;
    }
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

int gradHess(ccdh_fit* f, int j, int useW, double* g, double* h) {
  int rc = gradHessLaunch(f, j, useW);
  if (rc) return rc;
  rc = fetch(f, 2);
  if (rc) return rc;
  *g = f->hres[0];
  *h = f->hres[1];
  return 0;
}

int updateXBeta(ccdh_fit* f, double delta, int j, int useW) {
  ccdh_data* d = f->d;
  const int64_t n = d->hColPtr[j + 1] - d->hColPtr[j];
  if (n == 0) return 0;
  const int fmt = d->hFmt[j];
  switch (d->model) {
    case CCDH_LOGISTIC:
    case CCDH_POISSON:
      LAUNCH_R(k_update_indep, gridFor(n), d->view(), XB(T_), OEB(T_), RDEN(T_), j, fmt, delta);
      break;
    case CCDH_CONDITIONAL_LOGISTIC:
      LAUNCH_R(k_update_clr, gridFor(d->hSegCount(j), 2), d->view(), XB(T_), OEB(T_), f->denom, j,
               fmt, delta);
      break;
    case CCDH_COX:
      LAUNCH_R(k_update_cox, gridFor(n), d->view(), XB(T_), OEB(T_), f->denom, f->kw, f->pid,
               f->N, j, fmt, useW, delta);
      break;
  }
  return 0;
}

static int rowsReduce(ccdh_fit* f, int mode, int useW, const double* w2, const int32_t* group,
                      const double* gden) {
  const int K = f->d->K;
  LAUNCH_R(k_rows_reduce, gridFor(K), mode, K, XB(T_), RDEN(T_), f->d->y, f->kw, useW, w2, group,
           gden, f->partials, f->ticket, f->dres);
  return fetch(f, 2);
}

int logLik(ccdh_fit* f, int useCV, double* out) {
  ccdh_data* d = f->d;
  int rc = 0;
  switch (d->model) {
    case CCDH_LOGISTIC:
    case CCDH_POISSON: {
      rc = rowsReduce(f, d->model == CCDH_LOGISTIC ? RR_LL_LR : RR_LL_POISSON, useCV, nullptr,
                      nullptr, nullptr);
      if (rc) return rc;
      *out = f->hres[0] - f->hres[1] + (d->model == CCDH_POISSON ? f->fixedLL : 0.0);
      return 0;
    }
    case CCDH_CONDITIONAL_LOGISTIC: {
      // both reductions, then one read-back: the numerator in slot 0, the
      // groups' sum in slot 2 (the second kernel starts after the first ends,
      // so they can share the partials and the ticket)
      const int K = d->K;
      LAUNCH_R(k_rows_reduce, gridFor(K), RR_LL_NUM, K, XB(T_), RDEN(T_), d->y, f->kw, useCV,
               nullptr, nullptr, nullptr, f->partials, f->ticket, f->dres);
      k_groups_log<<<gridFor(f->N), kThreads, 0, f->stream>>>(f->N, f->nw, f->denom, f->partials,
                                                              f->ticket, f->dres + 2);
      HIP_TRY(hipGetLastError());
      rc = fetch(f, 3);
      if (rc) return rc;
      *out = f->hres[0] - f->hres[2];
      return 0;
    }
    case CCDH_COX: {
      rc = rowsReduce(f, RR_LL_NUM_SURV, useCV, nullptr, nullptr, nullptr);
      if (rc) return rc;
      const double num = f->hres[0];
      rc = coxTiles(f, 1, 0, 1);
      if (rc) return rc;
      rc = fetch(f, 1);
      if (rc) return rc;
      *out = num - f->hres[0];
      return 0;
    }
  }
  return setError(CCDH_E_INVALID, "model");
}

int gradientObjective(ccdh_fit* f, int useCV, double* out) {
  int rc = rowsReduce(f, RR_GRADOBJ, useCV, nullptr, nullptr, nullptr);
  if (rc) return rc;
  *out = f->hres[0];
  return 0;
}

int predLogLik(ccdh_fit* f, const double* w, double* out) {
  ccdh_data* d = f->d;
  if (!w) return setError(CCDH_E_INVALID, "weights required");
  HIP_TRY(hipMemcpyAsync(f->w2, w, d->K * sizeof(double), hipMemcpyHostToDevice, f->stream));
  int rc = 0;
  switch (d->model) {
    case CCDH_LOGISTIC:
      rc = rowsReduce(f, RR_PRED_LR, 0, f->w2, nullptr, nullptr);
      break;
    case CCDH_POISSON:
      rc = rowsReduce(f, RR_PRED_POISSON, 0, f->w2, nullptr, nullptr);
      break;
    case CCDH_CONDITIONAL_LOGISTIC:
      rc = rowsReduce(f, RR_PRED_GROUPED, 0, f->w2, d->strata, f->denom);
      break;
    case CCDH_COX: {
      // ModelSpecifics.hpp:656-687: regroup with the held-out weights, score, restore
      std::vector<double> saved = f->hKw;
      rc = setWeights(f, w, 1);
      rc = rc ? rc : remainingStats(f, 1);
      rc = rc ? rc : coxTiles(f, 2, 0, 1);
      rc = rc ? rc : rowsReduce(f, RR_PRED_COX, 0, f->w2, f->pid, f->accTmp);
      const double v = f->hres[0];
      rc = rc ? rc : setWeights(f, saved.data(), 1);
      rc = rc ? rc : remainingStats(f, 1);
      if (rc) return rc;
      *out = v;
      HIP_TRY(hipStreamSynchronize(f->stream));
      return 0;
    }
  }
  if (rc) return rc;
  *out = f->hres[0];
  return 0;
}

int getXBeta(ccdh_fit* f, double* out) {
  const int K = f->d->K;
  if (f->mixed) {
    std::vector<float> tmp(K);
    HIP_TRY(hipMemcpyAsync(tmp.data(), f->xb, K * sizeof(float), hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipStreamSynchronize(f->stream));
    for (int k = 0; k < K; ++k) out[k] = tmp[k];
  } else {
    HIP_TRY(hipMemcpyAsync(out, f->xb, K * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipStreamSynchronize(f->stream));
  }
  return 0;
}

// Row-major copy of X and the MM column chunks, built once per data set on the
// device: entries per row, an exclusive scan, a scatter, a per-row sort.
static int ensureTranspose(ccdh_data* d, hipStream_t s) {
  std::lock_guard<std::mutex> lock(d->lazyMu);
  if (d->tRowPtr) return 0;
  const int K = d->K, J = d->J;
  bool anyDense = false, anyIndicator = false;
  for (int j = 0; j < J; ++j) {
    anyDense = anyDense || d->hFmt[j] == CCDH_DENSE;
    anyIndicator = anyIndicator || d->hFmt[j] == CCDH_INDICATOR;
  }
  const bool positionsOnly = anyIndicator && !anyDense;
  int64_t* ptr = nullptr;
  int32_t *cnt = nullptr, *col = nullptr, *c0 = nullptr, *cc = nullptr;
  unsigned long long* cursor = nullptr;
  double* val = nullptr;
  auto fail = [&](hipError_t e, const char* what) {
    void* all[] = {ptr, cnt, col, c0, cc, cursor, val};
    for (void* p : all)
      if (p) (void)hipFree(p);
    return setError(CCDH_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  };
#define T_TRY(expr)                            \
  do {                                         \
    hipError_t e_ = (expr);                    \
    if (e_ != hipSuccess) return fail(e_, #expr); \
  } while (0)
  T_TRY(hipMalloc(&ptr, (size_t)(K + 1) * sizeof(int64_t)));
  T_TRY(hipMalloc(&cnt, (size_t)K * sizeof(int32_t)));
  T_TRY(hipMemsetAsync(cnt, 0, (size_t)K * sizeof(int32_t), s));
  const int colGrid = std::max(1, std::min(J, 65535));
  k_t_count<<<colGrid, kThreads, 0, s>>>(d->view(), cnt);
  T_TRY(hipGetLastError());
  k_scan_rows<<<1, kMaxBlockThreads, 0, s>>>(K, cnt, ptr);
  T_TRY(hipGetLastError());
  int64_t total = 0;
  T_TRY(hipMemcpyAsync(&total, ptr + K, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  T_TRY(hipStreamSynchronize(s));
  T_TRY(hipMalloc(&col, (size_t)std::max<int64_t>(total, 1) * sizeof(int32_t)));
  if (!positionsOnly) T_TRY(hipMalloc(&val, (size_t)std::max<int64_t>(total, 1) * sizeof(double)));
  T_TRY(hipMalloc(&cursor, (size_t)K * sizeof(unsigned long long)));
  T_TRY(hipMemcpyAsync(cursor, ptr, (size_t)K * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
  k_t_fill<<<colGrid, kThreads, 0, s>>>(d->view(), cursor, col, val);
  T_TRY(hipGetLastError());
  k_t_sort<<<gridFor(K, 1), kThreads, 0, s>>>(K, ptr, col, val);
  T_TRY(hipGetLastError());
  std::vector<int32_t> h0(J + 1, 0), hc;
  for (int j = 0; j < J; ++j) {
    const int64_t n = d->hColPtr[j + 1] - d->hColPtr[j];
    const int nc = (int)((n + kMmChunk - 1) / kMmChunk);
    h0[j + 1] = h0[j] + nc;
    for (int i = 0; i < nc; ++i) hc.push_back(j);
  }
  T_TRY(hipMalloc(&c0, (size_t)(J + 1) * sizeof(int32_t)));
  T_TRY(hipMalloc(&cc, std::max<size_t>(hc.size(), 1) * sizeof(int32_t)));
  T_TRY(hipMemcpyAsync(c0, h0.data(), h0.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  if (!hc.empty())
    T_TRY(hipMemcpyAsync(cc, hc.data(), hc.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  T_TRY(hipStreamSynchronize(s));
#undef T_TRY
  (void)hipFree(cnt);
  (void)hipFree(cursor);
  d->tRowPtr = ptr;
  d->tCol = col;
  d->tVal = val;
  d->mmChunk0 = c0;
  d->mmChunkCol = cc;
  d->nChunks = (int)hc.size();
  d->devBytes += (int64_t)(K + 1) * 8 + total * (positionsOnly ? 4 : 12) +
                 (int64_t)(J + 1 + hc.size()) * 4;
  return 0;
}

static int ensureMm(ccdh_fit* f) {
  ccdh_data* d = f->d;
  int rc = ensureTranspose(d, f->stream);
  for (int i = 0; !rc && i < 4; ++i)
    if (!f->mmEv[i]) {
      hipError_t e = hipEventCreate(&f->mmEv[i]);
      if (e != hipSuccess) rc = setError(CCDH_E_DEVICE, "hipEventCreate");
    }
  if (!rc && !f->mmPart) rc = devAlloc(&f->mmPart, 2 * (size_t)std::max(d->nChunks, 1), nullptr);
  if (!rc && !f->mmGH) rc = devAlloc(&f->mmGH, 2 * (size_t)d->J, nullptr);
  if (!rc && !f->mmBeta) rc = devAlloc(&f->mmBeta, (size_t)d->J, nullptr);
  if (!rc && !f->mmFixed) rc = devAlloc(&f->mmFixed, (size_t)d->J, nullptr);
  if (!rc && !f->mmR) rc = devAlloc(&f->mmR, (size_t)d->K, nullptr);
  return rc;
}

// adds the last xBeta launch's event time once it has run
static int readXbEvents(ccdh_fit* f) {
  if (!f->mmXbPending) return 0;
  float ms = 0.0f;
  HIP_TRY(hipEventSynchronize(f->mmEv[3]));
  HIP_TRY(hipEventElapsedTime(&ms, f->mmEv[2], f->mmEv[3]));
  f->mmStats.xbeta_launches++;
  f->mmStats.xbeta_ms += ms;
  f->mmXbPending = false;
  return 0;
}

int mmStats(ccdh_fit* f, ccdh_mm_stats* out) {
  int rc = readXbEvents(f);
  if (rc) return rc;
  ccdh_data* d = f->d;
  *out = f->mmStats;
  out->nnz = d->hColPtr[d->J];
  out->entries = 0;
  if (d->tRowPtr) HIP_TRY(hipMemcpy(&out->entries, d->tRowPtr + d->K, 8, hipMemcpyDeviceToHost));
  out->rows = d->K;
  out->values = 0;
  for (int j = 0; j < d->J; ++j)
    if (d->hFmt[j] == CCDH_SPARSE || d->hFmt[j] == CCDH_DENSE) out->values = 1;
  return 0;
}

// ModelSpecifics::computeXBeta(beta) (ModelSpecifics.hpp:313-364): xBeta from
// the row-major copy, one launch
int computeXBetaRows(ccdh_fit* f, const double* beta) {
  if (!beta) return setError(CCDH_E_INVALID, "beta required");
  int rc = ensureMm(f);
  if (rc) return rc;
  ccdh_data* d = f->d;
  HIP_TRY(hipMemcpyAsync(f->mmBeta, beta, d->J * sizeof(double), hipMemcpyHostToDevice,
                         f->stream));
  rc = readXbEvents(f);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(f->mmEv[2], f->stream));
  LAUNCH_R(k_xbeta_rows, gridFor(d->K, 1), d->K, d->tRowPtr, d->tCol, d->tVal, f->mmBeta,
           XB(T_));
  HIP_TRY(hipEventRecord(f->mmEv[3], f->stream));
  f->mmXbPending = true;
  f->statsKnown = false;
  return 0;
}

// computeMMGradientAndHessian (ModelSpecifics.hpp:776-889) for every column:
// per-row factors, one wave per column chunk, then chunk sums per column
int mmGradHess(ccdh_fit* f, const uint8_t* fixed, double* gh) {
  ccdh_data* d = f->d;
  if (d->model == CCDH_POISSON)  // ModelSpecifics.h:1546-1552
    return setError(CCDH_E_INVALID, "MM gradient: not model-specific (Poisson)");
  if (!fixed || !gh) return setError(CCDH_E_INVALID, "null argument");
  if (!f->weightsSet || !f->statsKnown)
    return setError(CCDH_E_STATE, "MM gradient before weights and statistics");
  int rc = ensureMm(f);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(f->mmFixed, fixed, d->J, hipMemcpyHostToDevice, f->stream));
  LAUNCH_R(k_mm_rows, gridFor(d->K), d->model, d->K, OEB(T_), f->kw, f->nw, f->denom,
           d->model == CCDH_COX ? (const int32_t*)f->pid : d->strata, f->mmR);
  const int wavesPerBlock = kThreads / kWave;
  if (d->nChunks > 0) {
    HIP_TRY(hipEventRecord(f->mmEv[0], f->stream));
    k_mm_chunks<<<(d->nChunks + wavesPerBlock - 1) / wavesPerBlock, kThreads, 0, f->stream>>>(
        d->view(), f->mmR, d->mmChunk0, d->mmChunkCol, d->nChunks, f->mmFixed, f->mmPart);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(f->mmEv[1], f->stream));
  }
  const int finishBlocks = std::min((d->J + wavesPerBlock - 1) / wavesPerBlock, 16384);
  k_mm_finish<<<finishBlocks, kThreads, 0, f->stream>>>(d->J, d->mmChunk0, f->mmPart, f->xjy,
                                                        f->mmFixed, f->mmGH);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(gh, f->mmGH, 2 * (size_t)d->J * sizeof(double), hipMemcpyDeviceToHost,
                         f->stream));
  HIP_TRY(hipStreamSynchronize(f->stream));
  if (d->nChunks > 0) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, f->mmEv[0], f->mmEv[1]));
    f->mmStats.chunk_launches++;
    f->mmStats.chunk_ms += ms;
  }
  return readXbEvents(f);
}

}  // namespace ccdamd

// =================================================================== C ABI
using namespace ccdamd;

#define GUARD(fit)                                                         \
  if (!(fit)) return setError(CCDH_E_INVALID, "null handle");

extern "C" {

const char* ccdh_last_error(void) { return g_lastError.c_str(); }

int ccdh_device_info(char* name, size_t len, int* cu_count, double* hbm_gb) {
  hipDeviceProp_t prop;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (name && len) {
    std::snprintf(name, len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (hbm_gb) *hbm_gb = prop.totalGlobalMem / 1e9;
  return 0;
}

int ccdh_data_create(const ccdh_data_desc* desc, ccdh_data_t* out) {
  try {
    return dataCreate(desc, out);
  } catch (const std::exception& e) {
    return setError(CCDH_E_INVALID, e.what());
  }
}

int ccdh_data_destroy(ccdh_data_t data) {
  dataDestroy(data);
  return 0;
}

int64_t ccdh_data_device_bytes(ccdh_data_t data) { return data ? data->devBytes : 0; }

int ccdh_data_precision(ccdh_data_t data) { return data ? data->precision : -1; }

int ccdh_data_shape(ccdh_data_t data, int* model, int* K, int* J, int* has_offset) {
  if (!data) return setError(CCDH_E_INVALID, "null handle");
  if (model) *model = data->model;
  if (K) *K = data->K;
  if (J) *J = data->J;
  if (has_offset) *has_offset = data->hasOffset;
  return 0;
}

int ccdh_fit_create(ccdh_data_t data, ccdh_fit_t* out) {
  if (!data || !out) return setError(CCDH_E_INVALID, "null argument");
  try {
    return fitCreate(data, out);
  } catch (const std::exception& e) {
    return setError(CCDH_E_INVALID, e.what());
  }
}

int ccdh_fit_clone(ccdh_fit_t fit, ccdh_fit_t* out) {
  GUARD(fit);
  return ccdh_fit_create(fit->d, out);
}

int ccdh_fit_destroy(ccdh_fit_t fit) {
  fitDestroy(fit);
  return 0;
}

int ccdh_set_weights(ccdh_fit_t fit, const double* weights, int use_cv) {
  GUARD(fit);
  return setWeights(fit, weights, use_cv);
}

int ccdh_fixed_terms_loglik(ccdh_fit_t fit, int use_cv) {
  GUARD(fit);
  return fixedTermsLogLik(fit, use_cv);
}

int ccdh_fixed_terms_grad_hess(ccdh_fit_t fit, int use_cv) {
  GUARD(fit);
  return fixedTermsGradHess(fit, use_cv);
}

int ccdh_data_get_csc(ccdh_data_t data, int64_t* col_ptr, int32_t* rows, double* vals,
                      int8_t* fmt, double* y, double* offs, int32_t* strata) {
  if (!data) return setError(CCDH_E_INVALID, "null handle");
  const int K = data->K, J = data->J;
  const int64_t nnz = data->hColPtr[J];
  if (col_ptr) std::copy(data->hColPtr.begin(), data->hColPtr.end(), col_ptr);
  if (fmt) std::copy(data->hFmt.begin(), data->hFmt.end(), fmt);
  if (y) std::copy(data->hY.begin(), data->hY.end(), y);
  if (offs) std::copy(data->hOffs.begin(), data->hOffs.end(), offs);
  if (strata) std::copy(data->hStrata.begin(), data->hStrata.end(), strata);
  if (rows && nnz > 0) HIP_TRY(hipMemcpy(rows, data->rows, nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (vals && nnz > 0) {
    if (data->vals)
      HIP_TRY(hipMemcpy(vals, data->vals, nnz * sizeof(double), hipMemcpyDeviceToHost));
    else
      std::fill(vals, vals + nnz, 1.0);
  }
  (void)K;
  return 0;
}

int ccdh_zero_xbeta(ccdh_fit_t fit) {
  GUARD(fit);
  return zeroXBeta(fit);
}

int ccdh_axpy_xbeta(ccdh_fit_t fit, double beta, int j) {
  GUARD(fit);
  return axpyXBeta(fit, beta, j);
}

int ccdh_compute_xbeta(ccdh_fit_t fit, const double* beta) {
  GUARD(fit);
  return computeXBetaRows(fit, beta);
}

int ccdh_fit_mm_stats(ccdh_fit_t fit, ccdh_mm_stats* out) {
  GUARD(fit);
  if (!out) return setError(CCDH_E_INVALID, "null argument");
  return mmStats(fit, out);
}

int ccdh_mm_grad_hess(ccdh_fit_t fit, const uint8_t* fixed, int use_weights, double* gh_out) {
  GUARD(fit);
  (void)use_weights;  // the MM increments weigh by nW whatever the flag
  return mmGradHess(fit, fixed, gh_out);
}

int ccdh_remaining_stats(ccdh_fit_t fit, int use_weights) {
  GUARD(fit);
  return remainingStats(fit, use_weights);
}

int ccdh_numerator(ccdh_fit_t fit, int j, int use_weights) {
  GUARD(fit);
  if (j < 0 || j >= fit->d->J) return setError(CCDH_E_INVALID, "column out of range");
  return numerator(fit, j, use_weights);
}

int ccdh_grad_hess(ccdh_fit_t fit, int j, int use_weights, double* gradient, double* hessian) {
  GUARD(fit);
  if (j < 0 || j >= fit->d->J) return setError(CCDH_E_INVALID, "column out of range");
  if (!fit->statsKnown) return setError(CCDH_E_STATE, "statistics not computed");
  return gradHess(fit, j, use_weights, gradient, hessian);
}

int ccdh_update_xbeta(ccdh_fit_t fit, double delta, int j, int use_weights) {
  GUARD(fit);
  if (j < 0 || j >= fit->d->J) return setError(CCDH_E_INVALID, "column out of range");
  return updateXBeta(fit, delta, j, use_weights);
}

int ccdh_loglik(ccdh_fit_t fit, int use_cv, double* out) {
  GUARD(fit);
  return logLik(fit, use_cv, out);
}

int ccdh_gradient_objective(ccdh_fit_t fit, int use_cv, double* out) {
  GUARD(fit);
  return gradientObjective(fit, use_cv, out);
}

int ccdh_pred_loglik(ccdh_fit_t fit, const double* weights, double* out) {
  GUARD(fit);
  return predLogLik(fit, weights, out);
}

int ccdh_get_xbeta(ccdh_fit_t fit, double* out) {
  GUARD(fit);
  return getXBeta(fit, out);
}

}  // extern "C"
