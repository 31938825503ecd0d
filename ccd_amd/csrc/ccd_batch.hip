// ccd_batch.hip -- device-resident batched CCD: many fits over one resident
// data set, each fit's whole coordinate sweep run inside one workgroup.
//
// Replaces, for batches of fits, the host loop of CyclicCoordinateDescent::findMode
// (CyclicCoordinateDescent.cpp:909-1216) around the per-coordinate engine calls:
// the gradient/Hessian gather (ModelSpecifics.hpp:891-1201), the prior step
// (priors/CovariatePrior.h:186-415), applyBounds (:1503-1520), the fused
// xBeta / exp(xBeta) / denominator update (ModelSpecifics.hpp:1524-1586) and the
// convergence check (CyclicCoordinateDescent.cpp:839-906) all run on the GPU.
// The batches are AbstractCrossValidationDriver's fold x grid-point fits
// (drivers/AbstractCrossValidationDriver.cpp:131-248) or served requests.
//
// Design (MI355X): one workgroup owns one fit for a whole sweep; the column's
// nonzeros (or stratum runs) are spread over the workgroup's lanes, the
// gradient and Hessian are reduced in FP64 with wave64 shuffles + LDS, lane 0
// takes the prior step, and the same lanes apply the update.  No grid-wide
// barrier and no host round trip per coordinate.  Per-fit state is contiguous
// per fit, so a workgroup streams its own rows; X (indices, values, runs) is
// read by every workgroup and stays hot in L2/MALL.  Reductions are in a fixed
// order, so a fit's result does not depend on what else is in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ccd_device.h"
#include "ccd_host.h"

using namespace ccdamd;

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return setError(CCDH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

namespace {

constexpr int kItems = 4;  // Cox pid items per lane per scan chunk

struct FitCtl {
  double tolerance;
  double initialBound;
  int maxIterations;
  int convergenceType;
  int useW;
  int pad;
};

struct FitStatus {
  double lastObj;
  double logLik;
  double predLogLik;
  int iteration;
  int flag;
  int done;
  int N;  // groups of this fit (Cox: tie groups under its weights)
  // work counters (thread 0), read by the host for the bandwidth figure
  long long gatherNnz, gatherRuns, updateNnz, updateRuns, scanPids, objRows;
  // wall-clock ticks (thread 0) spent per phase: single-pass columns, long columns, objective
  long long tSmall, tLarge, tObj, nLarge;
};

template <typename R>
struct BatchArgs {
  DevX d;
  int model, K, J, nFits;
  size_t strideK, strideN, strideJ;
  R* xb;
  R* oeb;
  double* kw;
  double* heldout;  // optional [nFits * K]
  double* denom;
  double* nw;
  double* numer;
  double* numer2;
  double* xjy;
  double* beta;
  double* bound;
  double* pvar;
  int8_t* ptype;
  uint8_t* fixed;
  int32_t* pid;
  int32_t* pidFirst;
  uint8_t* reset;
  const FitCtl* ctl;
  FitStatus* st;
  const int32_t* active;  // fit index per workgroup
  // per-fit records of long columns (stride strideRec)
  size_t strideRec;
  int32_t* recRow;
  int32_t* recKey;
  uint8_t* recFlags;
  R* recO;
  R* recXo;
  double* recDen;
};

// ------------------------------------------------------------ gather pieces


template <typename R>
__device__ __forceinline__ bool coxHead(const DevX& d, int fmt, int64_t b, int64_t p,
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

// Cox inside one workgroup: numerators by run heads, then one pass over all
// pids with a block-level segmented scan carrying across chunks.
template <typename R>
__device__ void ghCox(const BatchArgs<R>& a, int j, int fmt, const R* oeb, const double* kw,
                      int useW, const int32_t* pid, int N, const double* denom, double* numer,
                      double* numer2, const uint8_t* reset, const double* nw, double (&v)[2]) {
  const DevX& d = a.d;
  const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
  const bool ind = isIndicator(fmt);
  for (int64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
    int pp;
    if (!coxHead<R>(d, fmt, b, p, pid, N, &pp)) continue;
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
  __syncthreads();
  Seg<3> carry;
  carry.f = 0;
  carry.v[0] = carry.v[1] = carry.v[2] = 0.0;
  const int chunk = blockDim.x * kItems;
  for (int c0 = 0; c0 < N; c0 += chunk) {
    const int base = c0 + threadIdx.x * kItems;
    Seg<3> items[kItems];
    Seg<3> agg;
    agg.f = 0;
    agg.v[0] = agg.v[1] = agg.v[2] = 0.0;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      const int i = base + it;
      Seg<3> x;
      x.f = 0;
      x.v[0] = x.v[1] = x.v[2] = 0.0;
      if (i < N) {
        x.f = reset[i];
        x.v[0] = denom[i];
        x.v[1] = numer[i];
        x.v[2] = ind ? 0.0 : numer2[i];
      }
      items[it] = x;
      agg = segCombine(agg, x);
    }
    Seg<3> total;
    Seg<3> run = segCombine(carry, blockSegScanExcl<3>(agg, total));
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
      const int i = base + it;
      if (i >= N) break;
      run = segCombine(run, items[it]);
      const double accD = run.v[0];
      const double t = run.v[1] / accD;
      const double g = nw[i] * t;
      v[0] += g;
      v[1] += ind ? g * (1.0 - t) : nw[i] * (run.v[2] / accD - t * t);
      if (items[it].v[1] != 0.0) numer[i] = 0.0;
      if (!ind && items[it].v[2] != 0.0) numer2[i] = 0.0;
    }
    carry = segCombine(carry, total);
  }
}

// ------------------------------------------------------------ column passes
// A column is walked in passes of blockDim*IT nonzeros with the striped
// assignment p = pass start + it*blockDim + lane, so each index load of a wave
// reads 64 consecutive entries.  Grouped models: runs of equal stratum are
// contiguous in a column (rows are sorted by stratum); the lane at a run's last
// nonzero owns the stratum.  Runs of one nonzero (almost all) are summed in
// registers; columns with longer runs stage per-nonzero values in LDS, the owner
// sums its run there in row order, and a run that continues into the next pass
// is carried.  Each row belongs to one run, so lanes never share a row or a
// stratum.  Columns of at most blockDim*kIT nonzeros keep indices, exp(xBeta),
// xBeta and stratum terms in registers from gather to update; longer columns
// write one record per nonzero in the gather pass and the update pass reads the
// records back in one coalesced hop.
constexpr int kIT = 4;
constexpr int kITL = 8;
constexpr int kPassMax = 512 * kITL;

struct PassLds {
  double v1[kPassMax];
  double v2[kPassMax];
  unsigned char start[kPassMax];
  double c1, c2, n1, n2;
  int cValid, nValid;
};

template <int IT, typename R>
struct Lanes {
  int row[IT];
  int key[IT];
  bool own[IT];    // valid nonzero that ends its run (always, for independent rows)
  bool start[IT];  // valid nonzero that starts its run (always, for independent rows)
  R x[IT], o[IT], xo[IT];
  double den[IT], wn[IT], w[IT];
};

template <int MODEL, int IT, typename R>
__device__ __forceinline__ void passLoad(const DevX& d, int fmt, int64_t b, int64_t e, int64_t c0,
                                         const R* oeb, const R* xb, const double* denom,
                                         const double* nw, const double* kw, int useW,
                                         Lanes<IT, R>& L) {
  int kp[IT], kn[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int64_t p = c0 + (int64_t)it * blockDim.x + threadIdx.x;
    const bool valid = p < e;
    L.row[it] = valid ? colRow(d, fmt, p, b) : -1;
    L.x[it] = valid ? colValue<R>(d, fmt, p) : (R)0;
    if (MODEL == M_CLR) {
      L.key[it] = valid ? d.nzKey[p] : -1;
      kp[it] = (valid && p > b) ? d.nzKey[p - 1] : -2;
      kn[it] = (valid && p + 1 < e) ? d.nzKey[p + 1] : -3;
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int r = L.row[it];
    const bool valid = r >= 0;
    L.own[it] = MODEL == M_CLR ? (valid && L.key[it] != kn[it]) : valid;
    L.start[it] = MODEL == M_CLR ? (valid && L.key[it] != kp[it]) : valid;
    L.o[it] = valid ? oeb[r] : (R)0;
    L.xo[it] = valid ? xb[r] : (R)0;
    L.w[it] = (MODEL != M_CLR && useW && valid) ? kw[r] : 1.0;
    L.den[it] = (MODEL == M_CLR && L.own[it]) ? denom[L.key[it]] : 1.0;
    L.wn[it] = (MODEL == M_CLR && L.own[it]) ? nw[L.key[it]] : 0.0;
  }
}

// Run totals of NV per-nonzero values for the owners of one pass (see above).
// `staged` is uniform: false when every run of the column has one nonzero.
template <int IT, int NV>
__device__ __forceinline__ void runSums(PassLds& S, int nLocal, const bool (&own)[IT],
                                        const bool (&start)[IT], double (&val)[IT][NV],
                                        bool staged) {
  if (!staged) return;  // every owner's run is its own nonzero
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int li = it * blockDim.x + threadIdx.x;
    if (li < nLocal) {
      S.v1[li] = val[it][0];
      if (NV > 1) S.v2[li] = val[it][1];
      S.start[li] = start[it] ? 1 : 0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int li = it * blockDim.x + threadIdx.x;
    const bool last = li == nLocal - 1;
    if (li >= nLocal || !(own[it] || last)) continue;
    if (own[it] && start[it]) {
      if (last) S.nValid = 0;
      continue;
    }
    int k = li;
    while (k > 0 && !S.start[k]) --k;
    const bool before = !S.start[k];  // run began in an earlier pass
    double s1 = (before && S.cValid) ? S.c1 : 0.0;
    double s2 = (before && S.cValid) ? S.c2 : 0.0;
    for (int q = k; q <= li; ++q) {
      s1 += S.v1[q];
      if (NV > 1) s2 += S.v2[q];
    }
    if (own[it]) {
      val[it][0] = s1;
      if (NV > 1) val[it][1] = s2;
      if (last) S.nValid = 0;
    } else {  // last nonzero of the pass, run continues
      S.n1 = s1;
      S.n2 = s2;
      S.nValid = 1;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    S.c1 = S.n1;
    S.c2 = S.n2;
    S.cValid = S.nValid;
  }
}

template <int MODEL, int IT, typename R>
__device__ __forceinline__ void passGather(int fmt, PassLds& S, int nLocal, bool staged,
                                           const Lanes<IT, R>& L, int useW, double (&v)[2]) {
  const bool ind = isIndicator(fmt);
  if (MODEL == M_CLR) {
    double val[IT][2];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const R m = L.o[it] * L.x[it];
      val[it][0] = L.row[it] >= 0 ? (double)m : 0.0;
      val[it][1] = (L.row[it] >= 0 && !ind) ? (double)(m * L.x[it]) : 0.0;
    }
    runSums<IT, 2>(S, nLocal, L.own, L.start, val, staged);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      if (!L.own[it]) continue;
      const double t = val[it][0] / L.den[it];
      const double g = L.wn[it] * t;
      v[0] += g;
      v[1] += ind ? g * (1.0 - t) : L.wn[it] * (val[it][1] / L.den[it] - t * t);
    }
  } else {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      if (!L.own[it]) continue;
      const R n1 = L.o[it] * L.x[it];
      if (MODEL == M_LOGISTIC) {
        const R dd = (R)1 + L.o[it];
        const R t = n1 / dd;
        const R hh = ind ? t * ((R)1 - t) : ((n1 * L.x[it]) / dd - t * t);
        v[0] += useW ? L.w[it] * (double)t : (double)t;
        v[1] += useW ? L.w[it] * (double)hh : (double)hh;
      } else {
        const R n2 = ind ? n1 : n1 * L.x[it];
        v[0] += useW ? L.w[it] * (double)n1 : (double)n1;
        v[1] += useW ? L.w[it] * (double)n2 : (double)n2;
      }
    }
  }
}

template <int MODEL, int IT, typename R>
__device__ __forceinline__ void passUpdate(PassLds& S, int nLocal, bool staged, R* xb, R* oeb,
                                           double* denom, const Lanes<IT, R>& L, double delta) {
  const R rd = (R)delta;
  double val[IT][1];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    val[it][0] = 0.0;
    const int r = L.row[it];
    if (r < 0) continue;
    const R nx = L.xo[it] + rd * L.x[it];
    const R no = rexp(nx);
    xb[r] = nx;
    oeb[r] = no;
    val[it][0] = (double)no - (double)L.o[it];
  }
  if (MODEL == M_CLR) {
    runSums<IT, 1>(S, nLocal, L.own, L.start, val, staged);
#pragma unroll
    for (int it = 0; it < IT; ++it)
      if (L.own[it]) denom[L.key[it]] = L.den[it] + val[it][0];
  }
}

// gather-pass records of long columns, read back by the update pass
template <typename R>
struct Records {
  int32_t* row;
  int32_t* key;
  uint8_t* flags;  // bit 0 own, bit 1 start
  R* o;
  R* xo;
  double* den;
};

template <int MODEL, int IT, typename R>
__device__ __forceinline__ void storeRecords(const Records<R>& rec, int64_t off,
                                             const Lanes<IT, R>& L) {
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    if (L.row[it] < 0) continue;
    const int64_t q = off + (int64_t)it * blockDim.x + threadIdx.x;
    rec.row[q] = L.row[it];
    rec.o[q] = L.o[it];
    rec.xo[q] = L.xo[it];
    if (MODEL == M_CLR) {
      rec.key[q] = L.key[it];
      rec.flags[q] = (L.own[it] ? 1 : 0) | (L.start[it] ? 2 : 0);
      rec.den[q] = L.den[it];
    }
  }
}

template <int MODEL, int IT, typename R>
__device__ __forceinline__ void loadRecords(const DevX& d, int fmt, int64_t e, int64_t c0,
                                            int64_t off, const Records<R>& rec, Lanes<IT, R>& L) {
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int64_t p = c0 + (int64_t)it * blockDim.x + threadIdx.x;
    const int64_t q = off + (int64_t)it * blockDim.x + threadIdx.x;
    const bool valid = p < e;
    L.row[it] = valid ? rec.row[q] : -1;
    L.x[it] = valid ? colValue<R>(d, fmt, p) : (R)0;
    L.o[it] = valid ? rec.o[q] : (R)0;
    L.xo[it] = valid ? rec.xo[q] : (R)0;
    if (MODEL == M_CLR) {
      const int fl = valid ? rec.flags[q] : 0;
      L.key[it] = valid ? rec.key[q] : -1;
      L.own[it] = fl & 1;
      L.start[it] = fl & 2;
      L.den[it] = valid ? rec.den[q] : 1.0;
    } else {
      L.own[it] = valid;
      L.start[it] = valid;
    }
  }
}

// One full coordinate cycle of a fit (independent rows or grouped).  Every lane
// takes the (deterministic) prior step from the block-reduced gradient and
// Hessian, so no broadcast barrier is needed; a barrier follows an update only.
template <int MODEL, typename R>
__device__ void cycleFit(const BatchArgs<R>& a, int f, const FitCtl& c, R* xb, R* oeb,
                         const double* kw, double* denom, const double* nw, double* beta,
                         double* bound, const double* xjy, const double* pvar,
                         const int8_t* ptype, const uint8_t* fixed, const Records<R>& rec) {
  const DevX& d = a.d;
  __shared__ PassLds S;
  const int64_t CH = (int64_t)blockDim.x * kIT;
  const int64_t CHL = (int64_t)blockDim.x * kITL;
  const int J = a.J;
  const int useW = c.useW;
  long long cG = 0, cGR = 0, cU = 0, cUR = 0, tS = 0, tL = 0, nL = 0;
  for (int j = 0; j < J; ++j) {
    if (fixed[j]) continue;
    const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
    const int64_t nj = e - b;
    const int fmt = d.fmt[j];
    const double bj = beta[j], pv = pvar[j], xy = xjy[j];
    double bnd = bound[j];
    const int pt = ptype[j];
    const long long runs = MODEL == M_CLR ? (long long)(d.segPtr[j + 1] - d.segPtr[j]) : nj;
    const bool staged = MODEL == M_CLR && runs < nj;
    if (threadIdx.x == 0) S.cValid = 0;
    const long long t0 = wall_clock64();
    double delta;
    if (nj <= CH) {
      Lanes<kIT, R> L;
      passLoad<MODEL, kIT, R>(d, fmt, b, e, b, oeb, xb, denom, nw, kw, useW, L);
      double v[2] = {0.0, 0.0};
      passGather<MODEL, kIT, R>(fmt, S, (int)nj, staged, L, useW, v);
      blockSum<2>(v);
      double g = nj > 0 ? v[0] - xy : 0.0, h = nj > 0 ? v[1] : 0.0;
      if (h < 0.0) {
        g = 0.0;
        h = 0.0;
      }
      delta = applyBounds(priorDelta(pt, pv, g, h, bj), &bnd);
      if (threadIdx.x == 0) {
        bound[j] = bnd;
        if (delta != 0.0) beta[j] = bj + delta;
      }
      if (delta != 0.0 && nj > 0) {
        passUpdate<MODEL, kIT, R>(S, (int)nj, staged, xb, oeb, denom, L, delta);
        __syncthreads();
      }
    } else {
      double v[2] = {0.0, 0.0};
      for (int64_t c0 = b; c0 < e; c0 += CHL) {
        Lanes<kITL, R> L;
        passLoad<MODEL, kITL, R>(d, fmt, b, e, c0, oeb, xb, denom, nw, kw, useW, L);
        storeRecords<MODEL, kITL, R>(rec, c0 - b, L);
        passGather<MODEL, kITL, R>(fmt, S, (int)min(CHL, e - c0), staged, L, useW, v);
      }
      blockSum<2>(v);
      double g = v[0] - xy, h = v[1];
      if (h < 0.0) {
        g = 0.0;
        h = 0.0;
      }
      delta = applyBounds(priorDelta(pt, pv, g, h, bj), &bnd);
      if (threadIdx.x == 0) {
        bound[j] = bnd;
        if (delta != 0.0) beta[j] = bj + delta;
        S.cValid = 0;
      }
      if (delta != 0.0) {
        __syncthreads();  // carry reset visible; records were written by these lanes
        for (int64_t c0 = b; c0 < e; c0 += CHL) {
          Lanes<kITL, R> L;
          loadRecords<MODEL, kITL, R>(d, fmt, e, c0, c0 - b, rec, L);
          passUpdate<MODEL, kITL, R>(S, (int)min(CHL, e - c0), staged, xb, oeb, denom, L, delta);
        }
        __syncthreads();
      }
    }
    const long long t1 = wall_clock64();
    if (nj > CH) {
      tL += t1 - t0;
      nL += 1;
    } else {
      tS += t1 - t0;
    }
    cG += nj;
    cGR += MODEL == M_CLR ? runs : 0;
    if (delta != 0.0) {
      cU += nj;
      cUR += MODEL == M_CLR ? runs : 0;
    }
  }
  if (threadIdx.x == 0) {
    FitStatus& cs = a.st[f];
    cs.tSmall += tS;
    cs.tLarge += tL;
    cs.nLarge += nL;
    cs.gatherNnz += cG;
    cs.gatherRuns += cGR;
    cs.updateNnz += cU;
    cs.updateRuns += cUR;
  }
}

// ------------------------------------------------------------ update pieces


template <typename R>
__device__ void updCox(const BatchArgs<R>& a, int j, int fmt, R* xb, R* oeb, double* denom,
                       const double* kw, int useW, const int32_t* pid, int N, double delta) {
  const DevX& d = a.d;
  const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
  const R rd = (R)delta;
  for (int64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
    int pp;
    if (!coxHead<R>(d, fmt, b, p, pid, N, &pp)) {
      if (pp >= N) {
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

// ------------------------------------------------------------ whole-fit pieces
// log likelihood of one fit (ModelSpecifics.hpp:605-650); accumulated
// denominators for Cox come from an in-block scan
template <int MODEL, typename R>
__device__ double fitLogLik(const BatchArgs<R>& a, const R* xb, const R* oeb, const double* kw,
                            int useW, const double* denom, const double* nw, const uint8_t* reset,
                            int N) {
  const DevX& d = a.d;
  double v[2] = {0.0, 0.0};
  for (int k = threadIdx.x; k < a.K; k += blockDim.x) {
    const double x = (double)xb[k];
    const int yi = (int)d.y[k];
    const double w = useW ? kw[k] : 1.0;
    double c = (MODEL == M_COX) ? (yi != 1 ? 0.0 : x) : yi * x;
    v[0] += useW ? c * w : c;
    if (MODEL == M_LOGISTIC) v[1] += w * log(1.0 + (double)oeb[k]);
    if (MODEL == M_POISSON) v[1] += w * (double)oeb[k];
  }
  if (MODEL == M_CLR) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) v[1] += nw[i] * log(denom[i]);
  }
  if (MODEL == M_COX) {
    Seg<3> carry;
    carry.f = 0;
    carry.v[0] = carry.v[1] = carry.v[2] = 0.0;
    const int chunk = blockDim.x * kItems;
    for (int c0 = 0; c0 < N; c0 += chunk) {
      const int base = c0 + threadIdx.x * kItems;
      Seg<3> items[kItems];
      Seg<3> agg;
      agg.f = 0;
      agg.v[0] = agg.v[1] = agg.v[2] = 0.0;
#pragma unroll
      for (int it = 0; it < kItems; ++it) {
        const int i = base + it;
        Seg<3> x;
        x.f = 0;
        x.v[0] = x.v[1] = x.v[2] = 0.0;
        if (i < N) {
          x.f = reset[i];
          x.v[0] = denom[i];
        }
        items[it] = x;
        agg = segCombine(agg, x);
      }
      Seg<3> total;
      Seg<3> run = segCombine(carry, blockSegScanExcl<3>(agg, total));
#pragma unroll
      for (int it = 0; it < kItems; ++it) {
        const int i = base + it;
        if (i >= N) break;
        run = segCombine(run, items[it]);
        v[1] += nw[i] * log(run.v[0]);
      }
      carry = segCombine(carry, total);
    }
  }
  blockSum<2>(v);
  return v[0] - v[1];
}

template <typename R>
__device__ double fitGradObjective(const BatchArgs<R>& a, const R* xb, const double* kw,
                                   int useW) {
  double v[1] = {0.0};
  for (int k = threadIdx.x; k < a.K; k += blockDim.x) {
    const double x = (double)xb[k];
    v[0] += useW ? x * a.d.y[k] * kw[k] : x * a.d.y[k];
  }
  blockSum<1>(v);
  return v[0];
}

template <int MODEL, typename R>
__device__ double fitObjective(const BatchArgs<R>& a, int f, double fixedLL) {
  const FitCtl c = a.ctl[f];
  R* xb = a.xb + f * a.strideK;
  R* oeb = a.oeb + f * a.strideK;
  const double* kw = a.kw + f * a.strideK;
  if (c.convergenceType == 0) return fitGradObjective<R>(a, xb, kw, c.useW);
  if (c.convergenceType == 3) return 0.0;
  const double ll =
      fitLogLik<MODEL, R>(a, xb, oeb, kw, c.useW, a.denom + f * a.strideN, a.nw + f * a.strideN,
                          a.reset + f * a.strideN, a.st[f].N) +
      fixedLL;
  if (c.convergenceType == 2) return ll;
  double v[1] = {0.0};
  for (int j = threadIdx.x; j < a.J; j += blockDim.x)
    v[0] += priorLogDensity(a.ptype[f * a.strideJ + j], a.pvar[f * a.strideJ + j],
                            a.beta[f * a.strideJ + j]);
  blockSum<1>(v);
  return ll + v[0];
}

// ------------------------------------------------------------------ kernels
// Per-fit set-up: event weights, XjY, xBeta from the starting coefficients,
// exp / denominators, bounds and the starting objective (findMode's lazy part).
template <int MODEL, typename R>
__global__ void k_batch_init(BatchArgs<R> a, const double* fixedLL) {
  const int f = a.active[blockIdx.x];
  const DevX& d = a.d;
  const FitCtl c = a.ctl[f];
  R* xb = a.xb + f * a.strideK;
  R* oeb = a.oeb + f * a.strideK;
  const double* kw = a.kw + f * a.strideK;
  double* denom = a.denom + f * a.strideN;
  double* nw = a.nw + f * a.strideN;
  const int N = a.st[f].N;
  const int32_t* pid = a.pid ? a.pid + f * a.strideK : nullptr;
  const int32_t* pidFirst = a.pidFirst ? a.pidFirst + f * a.strideN : nullptr;
  // nW per group
  if (MODEL == M_CLR) {
    for (int s = threadIdx.x; s < d.nStrata; s += blockDim.x) {
      double acc = 0.0;
      for (int k = d.stratStart[s]; k < d.stratStart[s + 1]; ++k) acc += d.y[k] * kw[k];
      nw[s] = acc;
    }
  } else if (MODEL == M_COX) {
    for (int p = threadIdx.x; p < N; p += blockDim.x) {
      double acc = 0.0;
      for (int k = pidFirst[p]; k < pidFirst[p + 1]; ++k)
        if (pid[k] == p) acc += d.y[k] * kw[k];
      nw[p] = acc;
    }
  }
  // XjY
  for (int j = 0; j < a.J; ++j) {
    const int fmt = d.fmt[j];
    const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
    double s[1] = {0.0};
    for (int64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
      const int k = colRow(d, fmt, p, b);
      const double x = colValue<double>(d, fmt, p);
      s[0] += c.useW ? x * d.y[k] * kw[k] : x * d.y[k];
    }
    blockSum<1>(s);
    if (threadIdx.x == 0) a.xjy[f * a.strideJ + j] = s[0];
  }
  // xBeta = sum_j beta_j X_j in column order
  for (int k = threadIdx.x; k < a.K; k += blockDim.x) xb[k] = (R)0;
  __syncthreads();
  for (int j = 0; j < a.J; ++j) {
    const double bj = a.beta[f * a.strideJ + j];
    if (bj == 0.0) continue;
    const int fmt = d.fmt[j];
    const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
    const R rb = (R)bj;
    for (int64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
      const int k = colRow(d, fmt, p, b);
      xb[k] += rb * colValue<R>(d, fmt, p);
    }
    __syncthreads();
  }
  // exp and denominators
  if (MODEL == M_CLR) {
    for (int s = threadIdx.x; s < d.nStrata; s += blockDim.x) {
      double acc = 0.0;
      for (int k = d.stratStart[s]; k < d.stratStart[s + 1]; ++k) {
        const R o = rexp(xb[k]);
        oeb[k] = o;
        acc += (double)o;
      }
      denom[s] = acc;
    }
  } else {
    for (int k = threadIdx.x; k < a.K; k += blockDim.x) oeb[k] = rexp(xb[k]);
    if (MODEL == M_COX) {
      __syncthreads();
      for (int p = threadIdx.x; p < N; p += blockDim.x) {
        double acc = 0.0;
        for (int k = pidFirst[p]; k < pidFirst[p + 1]; ++k) {
          if (pid[k] != p) continue;
          acc += c.useW ? kw[k] * (double)oeb[k] : (double)oeb[k];
        }
        denom[p] = acc;
      }
    }
  }
  for (int j = threadIdx.x; j < a.J; j += blockDim.x) a.bound[f * a.strideJ + j] = c.initialBound;
  __syncthreads();
  const double obj = fitObjective<MODEL, R>(a, f, fixedLL[f]);
  if (threadIdx.x == 0) {
    a.st[f].lastObj = obj;
    a.st[f].iteration = 0;
    a.st[f].done = 0;
    a.st[f].flag = CCDH_SUCCESS;
  }
}

// Up to `sweeps` complete CCD cycles plus the convergence check after each.
template <int MODEL, typename R, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_batch_sweep(BatchArgs<R> a, const double* fixedLL,
                                                        int sweeps) {
  const int f = a.active[blockIdx.x];
  if (a.st[f].done) return;
  const DevX& d = a.d;
  const FitCtl c = a.ctl[f];
  R* xb = a.xb + f * a.strideK;
  R* oeb = a.oeb + f * a.strideK;
  const double* kw = a.kw + f * a.strideK;
  double* denom = a.denom + f * a.strideN;
  const double* nw = a.nw + f * a.strideN;
  double* beta = a.beta + f * a.strideJ;
  double* bound = a.bound + f * a.strideJ;
  const double* xjy = a.xjy + f * a.strideJ;
  const double* pvar = a.pvar + f * a.strideJ;
  const int8_t* ptype = a.ptype + f * a.strideJ;
  const uint8_t* fixed = a.fixed + f * a.strideJ;
  const int N = a.st[f].N;
  const int32_t* pid = MODEL == M_COX ? a.pid + f * a.strideK : nullptr;
  double* numer = MODEL == M_COX ? a.numer + f * a.strideN : nullptr;
  double* numer2 = MODEL == M_COX ? a.numer2 + f * a.strideN : nullptr;
  const uint8_t* reset = MODEL == M_COX ? a.reset + f * a.strideN : nullptr;
  __shared__ double sDelta;
  for (int sw = 0; sw < sweeps; ++sw) {
    if (MODEL != M_COX) {
      Records<R> rec;
      rec.row = a.recRow + f * a.strideRec;
      rec.key = a.recKey ? a.recKey + f * a.strideRec : nullptr;
      rec.flags = a.recFlags ? a.recFlags + f * a.strideRec : nullptr;
      rec.o = a.recO + f * a.strideRec;
      rec.xo = a.recXo + f * a.strideRec;
      rec.den = a.recDen ? a.recDen + f * a.strideRec : nullptr;
      cycleFit<MODEL, R>(a, f, c, xb, oeb, kw, denom, nw, beta, bound, xjy, pvar, ptype, fixed,
                         rec);
    } else
    for (int j = 0; j < a.J; ++j) {
      if (fixed[j]) continue;
      const int64_t nj = d.colPtr[j + 1] - d.colPtr[j];
      const int fmt = d.fmt[j];
      double v[2] = {0.0, 0.0};
      if (nj > 0) {
        ghCox<R>(a, j, fmt, oeb, kw, c.useW, pid, N, denom, numer, numer2, reset, nw, v);
        blockSum<2>(v);
      }
      if (threadIdx.x == 0) {
        FitStatus& cs = a.st[f];
        const long long runs = MODEL == M_CLR ? (long long)(d.segPtr[j + 1] - d.segPtr[j]) : 0;
        cs.gatherNnz += nj;
        cs.gatherRuns += runs;
        if (MODEL == M_COX && nj > 0) cs.scanPids += N;
        double g = nj > 0 ? v[0] - xjy[j] : 0.0;
        double h = nj > 0 ? v[1] : 0.0;
        if (h < 0.0) {
          g = 0.0;
          h = 0.0;
        }
        double delta = priorDelta(ptype[j], pvar[j], g, h, beta[j]);
        delta = applyBounds(delta, &bound[j]);
        if (delta != 0.0) {
          beta[j] += delta;
          cs.updateNnz += nj;
          cs.updateRuns += runs;
        }
        sDelta = delta;
      }
      __syncthreads();
      const double delta = sDelta;
      if (delta != 0.0) {
        updCox<R>(a, j, fmt, xb, oeb, denom, kw, c.useW, pid, N, delta);
      }
      __syncthreads();
    }
    // performCheckConvergence
    const long long tq0 = wall_clock64();
    const double obj = fitObjective<MODEL, R>(a, f, fixedLL[f]);
    if (threadIdx.x == 0) a.st[f].tObj += wall_clock64() - tq0;
    __shared__ int sDone;
    if (threadIdx.x == 0) {
      FitStatus& s = a.st[f];
      s.iteration += 1;
      s.objRows += (c.convergenceType == 0 ? a.K : (c.convergenceType == 3 ? 0 : a.K + N));
      double conv;
      bool ill = false;
      if (obj != obj) {
        conv = 0.0;
        ill = true;
      } else {
        conv = fabs(obj - s.lastObj) / (fabs(obj) + 1.0);
      }
      s.lastObj = obj;
      int done = 0;
      if (c.tolerance > 0 && conv < c.tolerance) {
        s.flag = ill ? CCDH_ILLCONDITIONED : CCDH_SUCCESS;
        done = 1;
      } else if (s.iteration == c.maxIterations) {
        s.flag = CCDH_MAX_ITERATIONS;
        done = 1;
      }
      s.done = done;
      sDone = done;
    }
    __syncthreads();
    if (sDone) return;
  }
}

// Final log likelihood and, when held-out weights are given, the predictive
// log likelihood (ModelSpecifics.hpp:652-690; LR / Poisson / CLR forms).
template <int MODEL, typename R>
__global__ void k_batch_finish(BatchArgs<R> a, const double* fixedLL) {
  const int f = a.active[blockIdx.x];
  const FitCtl c = a.ctl[f];
  R* xb = a.xb + f * a.strideK;
  R* oeb = a.oeb + f * a.strideK;
  const double* kw = a.kw + f * a.strideK;
  const double* denom = a.denom + f * a.strideN;
  const double ll = fitLogLik<MODEL, R>(a, xb, oeb, kw, c.useW, denom, a.nw + f * a.strideN,
                                        a.reset ? a.reset + f * a.strideN : nullptr, a.st[f].N);
  double v[1] = {0.0};
  if (a.heldout && MODEL != M_COX) {
    const double* w2 = a.heldout + f * a.strideK;
    for (int k = threadIdx.x; k < a.K; k += blockDim.x) {
      const double x = (double)xb[k];
      const double y = a.d.y[k];
      if (MODEL == M_LOGISTIC) v[0] += y * w2[k] * (x - log(1.0 + (double)oeb[k]));
      if (MODEL == M_POISSON) v[0] += (y * x - exp(x)) * w2[k];
      if (MODEL == M_CLR) v[0] += y * w2[k] * (x - log(denom[a.d.strata[k]]));
    }
  }
  blockSum<1>(v);
  if (threadIdx.x == 0) {
    a.st[f].logLik = ll + (MODEL == M_POISSON ? fixedLL[f] : 0.0);
    a.st[f].predLogLik = v[0];
  }
}

}  // namespace

// ================================================================ host side
namespace {

thread_local ccdh_batch_stats g_stats;

size_t roundUp(size_t n, size_t m) { return (n + m - 1) / m * m; }

template <typename T>
int allocZero(T** p, size_t n, hipStream_t s, std::vector<void*>& owned) {
  *p = nullptr;
  if (n == 0) return 0;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return setError(CCDH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  owned.push_back(*p);
  HIP_TRY(hipMemsetAsync(*p, 0, n * sizeof(T), s));
  return 0;
}

struct Owned {
  std::vector<void*> ptrs;
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  FitStatus* hst = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~Owned() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : ptrs) (void)hipFree(p);
    if (hst) (void)hipHostFree(hst);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

template <int MODEL, typename R>
int runBatch(ccdh_data* d, int F, const ccdh_fit_spec* specs, double* betasOut,
             ccdh_fit_result* results, const ccdh_batch_options* opt) {
  const auto t0 = std::chrono::steady_clock::now();
  const int K = d->K, J = d->J;
  int threads = (opt && opt->threads) ? opt->threads : 256;
  threads = std::max(64, std::min(1024, threads / 64 * 64));
  const int sweepsPerSync = (opt && opt->sweeps_per_sync > 0) ? opt->sweeps_per_sync : 4;
  const bool useGraph = opt ? opt->use_graph != 0 : true;

  Owned o;
  HIP_TRY(hipStreamCreateWithFlags(&o.stream, hipStreamNonBlocking));
  hipStream_t s = o.stream;

  // groups per fit
  std::vector<int> Nf(F, MODEL == M_CLR ? d->nStrata : K);
  std::vector<std::vector<int32_t>> pids, pidFirsts;
  std::vector<std::vector<uint8_t>> resets;
  if (MODEL == M_COX) {
    pids.resize(F);
    pidFirsts.resize(F);
    resets.resize(F);
    for (int f = 0; f < F; ++f)
      coxPids(d, specs[f].weights, pids[f], &Nf[f], resets[f], pidFirsts[f]);
  }
  int maxN = 1;
  for (int f = 0; f < F; ++f) maxN = std::max(maxN, Nf[f]);
  const size_t sK = roundUp(K, 64), sJ = roundUp(J, 64);
  const size_t sN = (MODEL == M_CLR || MODEL == M_COX) ? roundUp(maxN + 1, 64) : 1;
  bool anyHeld = false;
  for (int f = 0; f < F; ++f) anyHeld |= specs[f].heldout != nullptr;

  BatchArgs<R> a;
  std::memset(&a, 0, sizeof(a));
  a.d = d->view();
  a.model = MODEL;
  a.K = K;
  a.J = J;
  a.nFits = F;
  a.strideK = sK;
  a.strideN = sN;
  a.strideJ = sJ;
  int rc = 0;
  auto& ow = o.ptrs;
  rc = rc ? rc : allocZero(&a.xb, F * sK, s, ow);
  rc = rc ? rc : allocZero(&a.oeb, F * sK, s, ow);
  rc = rc ? rc : allocZero(&a.kw, F * sK, s, ow);
  if (anyHeld && MODEL != M_COX) rc = rc ? rc : allocZero(&a.heldout, F * sK, s, ow);
  if (MODEL == M_CLR || MODEL == M_COX) {
    rc = rc ? rc : allocZero(&a.denom, F * sN, s, ow);
    rc = rc ? rc : allocZero(&a.nw, F * sN, s, ow);
  }
  if (MODEL == M_COX) {
    rc = rc ? rc : allocZero(&a.numer, F * sN, s, ow);
    rc = rc ? rc : allocZero(&a.numer2, F * sN, s, ow);
    rc = rc ? rc : allocZero(&a.pid, F * sK, s, ow);
    rc = rc ? rc : allocZero(&a.pidFirst, F * sN, s, ow);
    rc = rc ? rc : allocZero(&a.reset, F * sN, s, ow);
  }
  if (MODEL != M_COX) {
    int64_t maxLen = 1;
    for (int j = 0; j < J; ++j) maxLen = std::max(maxLen, d->hColPtr[j + 1] - d->hColPtr[j]);
    const size_t sR = roundUp((size_t)maxLen, 64);
    a.strideRec = sR;
    rc = rc ? rc : allocZero(&a.recRow, F * sR, s, ow);
    rc = rc ? rc : allocZero(&a.recO, F * sR, s, ow);
    rc = rc ? rc : allocZero(&a.recXo, F * sR, s, ow);
    if (MODEL == M_CLR) {
      rc = rc ? rc : allocZero(&a.recKey, F * sR, s, ow);
      rc = rc ? rc : allocZero(&a.recFlags, F * sR, s, ow);
      rc = rc ? rc : allocZero(&a.recDen, F * sR, s, ow);
    }
  }
  rc = rc ? rc : allocZero(&a.xjy, F * sJ, s, ow);
  rc = rc ? rc : allocZero(&a.beta, F * sJ, s, ow);
  rc = rc ? rc : allocZero(&a.bound, F * sJ, s, ow);
  rc = rc ? rc : allocZero(&a.pvar, F * sJ, s, ow);
  rc = rc ? rc : allocZero(&a.ptype, F * sJ, s, ow);
  rc = rc ? rc : allocZero(&a.fixed, F * sJ, s, ow);
  FitCtl* dctl = nullptr;
  FitStatus* dst = nullptr;
  int32_t* dactive = nullptr;
  double* dfixedLL = nullptr;
  rc = rc ? rc : allocZero(&dctl, F, s, ow);
  rc = rc ? rc : allocZero(&dst, F, s, ow);
  rc = rc ? rc : allocZero(&dactive, F, s, ow);
  rc = rc ? rc : allocZero(&dfixedLL, F, s, ow);
  if (rc) return rc;
  a.ctl = dctl;
  a.st = dst;
  a.active = dactive;
  HIP_TRY(hipHostMalloc((void**)&o.hst, F * sizeof(FitStatus), hipHostMallocDefault));

  // per-fit host inputs
  std::vector<FitCtl> ctl(F);
  std::vector<FitStatus> st(F);
  std::vector<double> fixedLL(F, 0.0), hb(sJ), hv(sJ);
  std::vector<int8_t> ht(sJ);
  std::vector<uint8_t> hf(sJ);
  std::vector<int32_t> active(F);
  std::vector<int> boundCount(F, 1);
  std::vector<double> initialBound(F);
  for (int f = 0; f < F; ++f) {
    const ccdh_fit_spec& sp = specs[f];
    ctl[f] = FitCtl{sp.tolerance, sp.initial_bound, sp.max_iterations, sp.convergence_type,
                    sp.weights != nullptr, 0};
    initialBound[f] = sp.initial_bound;
    std::memset(&st[f], 0, sizeof(FitStatus));
    st[f].N = Nf[f];
    active[f] = f;
    for (int j = 0; j < J; ++j) {
      hb[j] = sp.beta0[j];
      hv[j] = sp.prior_var[j];
      ht[j] = (int8_t)sp.prior_type[j];
      hf[j] = sp.fixed[j];
    }
    HIP_TRY(hipMemcpyAsync(a.beta + f * sJ, hb.data(), J * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(a.pvar + f * sJ, hv.data(), J * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(a.ptype + f * sJ, ht.data(), J, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(a.fixed + f * sJ, hf.data(), J, hipMemcpyHostToDevice, s));
    if (sp.weights) {
      HIP_TRY(hipMemcpyAsync(a.kw + f * sK, sp.weights, K * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
      std::vector<double> ones(K, 1.0);
      HIP_TRY(hipMemcpyAsync(a.kw + f * sK, ones.data(), K * sizeof(double), hipMemcpyHostToDevice, s));
      HIP_TRY(hipStreamSynchronize(s));
    }
    if (sp.heldout && a.heldout)
      HIP_TRY(hipMemcpyAsync(a.heldout + f * sK, sp.heldout, K * sizeof(double), hipMemcpyHostToDevice, s));
    if (MODEL == M_POISSON) {
      double t = 0.0;
      for (int k = 0; k < K; ++k)
        t += sp.weights ? d->hLogFact[k] * sp.weights[k] : d->hLogFact[k];
      fixedLL[f] = t;
    }
    if (MODEL == M_COX) {
      HIP_TRY(hipMemcpyAsync(a.pid + f * sK, pids[f].data(), K * sizeof(int32_t), hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(a.pidFirst + f * sN, pidFirsts[f].data(), (Nf[f] + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(a.reset + f * sN, resets[f].data(), Nf[f] + 1, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));  // host vectors are reused per fit
  }
  HIP_TRY(hipMemcpyAsync(dctl, ctl.data(), F * sizeof(FitCtl), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst, st.data(), F * sizeof(FitStatus), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dactive, active.data(), F * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dfixedLL, fixedLL.data(), F * sizeof(double), hipMemcpyHostToDevice, s));

  k_batch_init<MODEL, R><<<F, threads, 0, s>>>(a, dfixedLL);
  HIP_TRY(hipGetLastError());

  // sweep launches, optionally replayed from a captured graph
  auto enqueueSweeps = [&]() -> int {
    if (threads <= 256)
      k_batch_sweep<MODEL, R, 256><<<F, 256, 0, s>>>(a, dfixedLL, sweepsPerSync);
    else
      k_batch_sweep<MODEL, R, 512><<<F, 512, 0, s>>>(a, dfixedLL, sweepsPerSync);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(o.hst, dst, F * sizeof(FitStatus), hipMemcpyDeviceToHost, s));
    return 0;
  };
  if (useGraph) {
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = enqueueSweeps();
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(s, &g);
    if (rc) return rc;
    if (ce != hipSuccess) return setError(CCDH_E_DEVICE, "graph capture failed");
    o.graph = g;
    HIP_TRY(hipGraphInstantiate(&o.exec, o.graph, nullptr, nullptr, 0));
  }
  std::vector<double> doneAt(F, 0.0);
  std::vector<int> lastIter(F, 0);
  std::memset(&g_stats, 0, sizeof(g_stats));
  HIP_TRY(hipEventCreate(&o.ev0));
  HIP_TRY(hipEventCreate(&o.ev1));
  int guard = 0;
  for (;;) {
    HIP_TRY(hipEventRecord(o.ev0, s));
    if (useGraph) {
      HIP_TRY(hipGraphLaunch(o.exec, s));
    } else {
      rc = enqueueSweeps();
      if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(o.ev1, s));
    HIP_TRY(hipStreamSynchronize(s));
    {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, o.ev0, o.ev1));
      g_stats.kernel_ms += ms;
      g_stats.launches += 1;
      int rounds = 0;
      for (int f = 0; f < F; ++f) {
        rounds = std::max(rounds, o.hst[f].iteration - lastIter[f]);
        lastIter[f] = o.hst[f].iteration;
      }
      g_stats.sweep_rounds += rounds;
    }
    const double now = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    bool allDone = true, retry = false;
    for (int f = 0; f < F; ++f) {
      FitStatus& h = o.hst[f];
      if (!h.done) {
        allDone = false;
        continue;
      }
      if (doneAt[f] == 0.0) doneAt[f] = now;
      // CyclicCoordinateDescent::update: ill-conditioned -> bound / 10, cold restart
      if (h.flag == CCDH_ILLCONDITIONED && boundCount[f] < specs[f].max_bound_count) {
        boundCount[f]++;
        initialBound[f] /= 10.0;
        ctl[f].initialBound = initialBound[f];
        for (int j = 0; j < J; ++j)
          hb[j] = (d->hasOffset && j == 0) ? specs[f].beta0[0] : 0.0;
        HIP_TRY(hipMemcpyAsync(a.beta + f * sJ, hb.data(), J * sizeof(double), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(dctl + f, &ctl[f], sizeof(FitCtl), hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        doneAt[f] = 0.0;
        lastIter[f] = 0;
        retry = true;
        active[0] = f;
        HIP_TRY(hipMemcpyAsync(dactive, active.data(), sizeof(int32_t), hipMemcpyHostToDevice, s));
        k_batch_init<MODEL, R><<<1, threads, 0, s>>>(a, dfixedLL);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
      }
    }
    if (retry) {
      for (int f = 0; f < F; ++f) active[f] = f;
      HIP_TRY(hipMemcpyAsync(dactive, active.data(), F * sizeof(int32_t), hipMemcpyHostToDevice, s));
      continue;
    }
    if (allDone) break;
    if (++guard > 100000000) return setError(CCDH_E_STATE, "batch did not terminate");
  }
  k_batch_finish<MODEL, R><<<F, threads, 0, s>>>(a, dfixedLL);
  HIP_TRY(hipGetLastError());
  std::vector<FitStatus> fin(F);
  HIP_TRY(hipMemcpyAsync(fin.data(), dst, F * sizeof(FitStatus), hipMemcpyDeviceToHost, s));
  for (int f = 0; f < F; ++f)
    HIP_TRY(hipMemcpyAsync(betasOut + (size_t)f * J, a.beta + f * sJ, J * sizeof(double),
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int f = 0; f < F; ++f) {
    g_stats.sweeps += fin[f].iteration;
    g_stats.gather_nnz += fin[f].gatherNnz;
    g_stats.gather_runs += fin[f].gatherRuns;
    g_stats.update_nnz += fin[f].updateNnz;
    g_stats.update_runs += fin[f].updateRuns;
    g_stats.scan_pids += fin[f].scanPids;
    g_stats.objective_rows += fin[f].objRows;
    g_stats.ticks_single += fin[f].tSmall;
    g_stats.ticks_long += fin[f].tLarge;
    g_stats.ticks_objective += fin[f].tObj;
    g_stats.long_columns += fin[f].nLarge;
  }
  g_stats.wall_ms = total * 1e3;
  for (int f = 0; f < F; ++f) {
    ccdh_fit_result& r = results[f];
    r.log_likelihood = fin[f].logLik;
    double lp = 0.0;
    for (int j = 0; j < J; ++j)
      lp += priorLogDensity(specs[f].prior_type[j], specs[f].prior_var[j], betasOut[(size_t)f * J + j]);
    r.log_prior = lp;
    r.iterations = fin[f].iteration;
    r.return_flag = fin[f].flag;
    r.bound_count = boundCount[f];
    r.pred_log_likelihood = fin[f].predLogLik;
    r.seconds = doneAt[f] > 0 ? doneAt[f] : total;
  }
  // Cox held-out scoring regroups the risk sets under the held-out weights:
  // run it on the per-coordinate engine with the batch's coefficients.
  if (MODEL == M_COX && anyHeld) {
    for (int f = 0; f < F; ++f) {
      if (!specs[f].heldout) continue;
      ccdh_fit* fh = nullptr;
      rc = fitCreate(d, &fh);
      if (rc) return rc;
      const int useW = specs[f].weights != nullptr;
      rc = setWeights(fh, specs[f].weights, useW);
      if (!rc) {
        rc = zeroXBeta(fh);
        for (int j = 0; !rc && j < J; ++j) rc = axpyXBeta(fh, betasOut[(size_t)f * J + j], j);
      }
      rc = rc ? rc : remainingStats(fh, useW);
      double v = 0.0;
      rc = rc ? rc : predLogLik(fh, specs[f].heldout, &v);
      fitDestroy(fh);
      if (rc) return rc;
      results[f].pred_log_likelihood = v;
    }
  }
  return 0;
}

}  // namespace

extern "C" int ccdh_batch_last_stats(ccdh_batch_stats* out) {
  if (!out) return setError(CCDH_E_INVALID, "null argument");
  *out = g_stats;
  return 0;
}

extern "C" int ccdh_batch_fit(ccdh_data_t data, int n_fits, const ccdh_fit_spec* specs,
                              double* betas_out, ccdh_fit_result* results,
                              const ccdh_batch_options* options) {
  if (!data || !specs || !betas_out || !results || n_fits <= 0)
    return setError(CCDH_E_INVALID, "null argument");
  for (int f = 0; f < n_fits; ++f) {
    const ccdh_fit_spec& sp = specs[f];
    if (!sp.prior_type || !sp.prior_var || !sp.fixed || !sp.beta0)
      return setError(CCDH_E_INVALID, "fit spec arrays required");
    if (sp.convergence_type < 0 || sp.convergence_type > 3)
      return setError(CCDH_E_INVALID, "unsupported convergence type");
  }
  const bool mixed = data->precision == CCDH_MIXED;
  try {
    switch (data->model) {
      case CCDH_LOGISTIC:
        return mixed ? runBatch<M_LOGISTIC, float>(data, n_fits, specs, betas_out, results, options)
                     : runBatch<M_LOGISTIC, double>(data, n_fits, specs, betas_out, results, options);
      case CCDH_POISSON:
        return mixed ? runBatch<M_POISSON, float>(data, n_fits, specs, betas_out, results, options)
                     : runBatch<M_POISSON, double>(data, n_fits, specs, betas_out, results, options);
      case CCDH_CONDITIONAL_LOGISTIC:
        return mixed ? runBatch<M_CLR, float>(data, n_fits, specs, betas_out, results, options)
                     : runBatch<M_CLR, double>(data, n_fits, specs, betas_out, results, options);
      case CCDH_COX:
        return mixed ? runBatch<M_COX, float>(data, n_fits, specs, betas_out, results, options)
                     : runBatch<M_COX, double>(data, n_fits, specs, betas_out, results, options);
    }
  } catch (const std::exception& e) {
    return setError(CCDH_E_DEVICE, e.what());
  }
  return setError(CCDH_E_INVALID, "unsupported model");
}
