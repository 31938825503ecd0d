// ccd_batch.hip -- device-resident CCD over fit slots: many fits in flight, each
// fit's whole coordinate sweep run inside one workgroup.
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
// Design (MI355X): one workgroup owns one slot for a whole sweep; the column's
// nonzeros (or stratum runs) are spread over the workgroup's lanes, the
// gradient and Hessian are reduced in FP64 with wave64 shuffles + LDS, every
// lane takes the prior step, and the same lanes apply the update.  No
// grid-wide barrier and no host round trip per coordinate.  A slot's state is
// one contiguous block, so a workgroup streams its own rows; X (indices,
// values, runs) is read by every workgroup over it and stays hot in L2/MALL.
// Reductions are in a fixed order, so a fit's result does not depend on what
// else is in the launch.  The sweep kernel dispatches on each slot's model, so
// slots of different models and data sets share one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "ccd_device.h"
#include "ccd_host.h"
#include "ccd_slot.h"

using namespace ccdamd;

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return setError(CCDH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

namespace {

constexpr int kIT = 4;           // nonzeros per lane of a single-pass column
constexpr int kITLDefault = 8;  // long-column pass width (nonzeros per lane); 4 is the alternative

// ------------------------------------------------------------ Cox column passes
// Numerators and updates of a Cox column part, in passes of blockDim * kIT nonzeros.
// Entry li = it * blockDim + lane of a pass is nonzero c0 + li of the part (so
// each index load of a wave reads 64 consecutive entries), with its pid (key; N
// for a zero-weight row, whose values are 0 and which belongs to no run) and its
// values.  Rows are sorted by pid, so a pid's entries form one run of the part,
// with zero-weight entries possibly inside.  The run's last entry (no later
// entry of the pass carries a pid, or the next one carries another) accumulates
// the run forward in column order, as the reference does row by row --
// numerators from 0 (MODE 0), a pid's denominator from its current value
// (MODE 1) -- and writes the result.  A run that may go on past the pass is
// carried: its running values pass to the next pass, where the run's last entry
// continues them, or thread 0 writes them out when the pass starts with another
// pid.  `key` is an LDS array of the pass size; S.cValid[0] = 0 before the
// first pass of a part.
template <int MODE, typename R, class LDS>
__device__ __forceinline__ void coxPassRuns(const Slot<R>& s, LDS& S, int* key, int n, bool final,
                                            bool two, const int (&k)[kIT], const double (&v1)[kIT],
                                            const double (&v2)[kIT]) {
  const int N = s.N;
  auto write = [&](int pid, double a1, double a2) {
    if (MODE == 0) {
      s.numer[pid] = a1;
      if (two) s.numer2[pid] = a2;
    } else {
      s.grp[pid].den = a1;
    }
  };
#pragma unroll
  for (int it = 0; it < kIT; ++it) {
    const int li = it * blockDim.x + threadIdx.x;
    if (li < n) {
      key[li] = k[it];
      S.v1[li] = v1[it];
      if (two) S.v2[li] = v2[it];
    }
  }
  __syncthreads();
  const bool cin = S.cValid[0] != 0;
  const int ck = S.cKey;
  if (threadIdx.x == 0) {
    int q = 0;
    while (q < n && key[q] >= N) ++q;
    // the carried run ended with the last pass: the pass starts with another pid
    // (or the part ends here without one)
    if (cin && (q < n ? key[q] != ck : final)) write(ck, S.c1[0], S.c2[0]);
    if (q == n && !final) {  // no pid in this pass: the carried run goes on
      S.n1[0] = S.c1[0];
      S.n2[0] = S.c2[0];
      S.nKey = ck;
      S.nValid[0] = S.cValid[0];
    } else if (final) {
      S.nValid[0] = 0;
    }  // otherwise the pass's last pid entry sets the carry below
  }
#pragma unroll
  for (int it = 0; it < kIT; ++it) {
    const int li = it * blockDim.x + threadIdx.x;
    const int p = k[it];
    if (li >= n || p >= N) continue;
    int q = li + 1;
    while (q < n && key[q] >= N) ++q;
    if (q < n && key[q] == p) continue;  // not the run's last entry in this pass
    int s0 = li;
    bool toStart = true;  // the run reaches back to the pass's first entry
    for (int r = li - 1; r >= 0; --r) {
      const int kr = key[r];
      if (kr < N && kr != p) {
        toStart = false;
        break;
      }
      s0 = r;
    }
    double a1, a2 = 0.0;
    if (toStart && cin && ck == p) {
      a1 = S.c1[0];
      a2 = S.c2[0];
    } else {
      a1 = MODE == 0 ? 0.0 : s.grp[p].den;
    }
    for (int r = s0; r <= li; ++r) {
      a1 += S.v1[r];
      if (two) a2 += S.v2[r];
    }
    if (q == n && !final) {  // the run may go on in the next pass
      S.n1[0] = a1;
      S.n2[0] = a2;
      S.nKey = p;
      S.nValid[0] = 1;
    } else {
      write(p, a1, a2);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    S.c1[0] = S.n1[0];
    S.c2[0] = S.n2[0];
    S.cKey = S.nKey;
    S.cValid[0] = S.nValid[0];
  }
}

// Numerators of the pids a column part [b, e) touches (computeNumeratorForGradient,
// ModelSpecifics.hpp:1353-1464), in passes of blockDim * kIT nonzeros.
template <typename R, class LDS>
__device__ __forceinline__ void coxNumeratorsLds(const Slot<R>& s, LDS& S, int* key, int fmt,
                                                 int64_t cb, int64_t b, int64_t e, int useW) {
  const DevX& d = s.d;
  const int N = s.N;
  const bool ind = isIndicator(fmt);
  const int64_t P = (int64_t)blockDim.x * kIT;
  if (threadIdx.x == 0) S.cValid[0] = 0;
  for (int64_t c0 = b; c0 < e; c0 += P) {
    const int n = (int)min(P, e - c0);
    int k[kIT];
    double v1[kIT], v2[kIT];
#pragma unroll
    for (int it = 0; it < kIT; ++it) {
      const int li = it * blockDim.x + threadIdx.x;
      k[it] = N;
      v1[it] = v2[it] = 0.0;
      if (li < n) {
        const int64_t p = c0 + li;
        const int r = colRow(d, fmt, p, cb);
        const int pp = s.pid[r];
        k[it] = pp;
        if (pp < N) {
          const double x = (double)colValue<R>(d, fmt, p);
          const double o = (double)s.row[r].e;
          v1[it] = useW ? s.kw[r] * (o * x) : o * x;
          if (!ind) v2[it] = useW ? s.kw[r] * (o * x * x) : o * x * x;
        }
      }
    }
    coxPassRuns<0, R>(s, S, key, n, c0 + P >= e, !ind, k, v1, v2);
  }
}

// xBeta, exp(xBeta) of the rows of a column part [b, e) and the denominators of
// their pids (updateXBeta, ModelSpecifics.hpp:1524-1586), in passes as above.
template <typename R, class LDS>
__device__ __forceinline__ void updCoxLds(const Slot<R>& s, LDS& S, int* key, int fmt, int64_t cb,
                                          int64_t b, int64_t e, int useW, double delta) {
  const DevX& d = s.d;
  const int N = s.N;
  const R rd = (R)delta;
  const int64_t P = (int64_t)blockDim.x * kIT;
  if (threadIdx.x == 0) S.cValid[0] = 0;
  for (int64_t c0 = b; c0 < e; c0 += P) {
    const int n = (int)min(P, e - c0);
    int k[kIT];
    double v1[kIT], v2[kIT];
#pragma unroll
    for (int it = 0; it < kIT; ++it) {
      const int li = it * blockDim.x + threadIdx.x;
      k[it] = N;
      v1[it] = v2[it] = 0.0;
      if (li < n) {
        const int64_t p = c0 + li;
        const int r = colRow(d, fmt, p, cb);
        const int pp = s.pid[r];
        const RowSt<R> old = s.row[r];
        const R nx = old.x + rd * colValue<R>(d, fmt, p);
        const R o = rexp(nx);
        s.row[r] = RowSt<R>{nx, o};
        k[it] = pp;
        if (pp < N) {
          const double w = useW ? s.kw[r] : 1.0;
          const double oldE = useW ? w * (double)old.e : (double)old.e;
          const double newE = useW ? w * (double)o : (double)o;
          v1[it] = newE - oldE;
        }
      }
    }
    coxPassRuns<1, R>(s, S, key, n, c0 + P >= e, false, k, v1, v2);
  }
}

// Device ticks per nonzero of this workgroup's last column of more than one
// pass (the slot's lead carries it from launch to launch): thread 0 keeps it,
// and it prices the next such column when the quantum is checked.
__device__ __forceinline__ float& longTicksPerNnz() {
  __shared__ float sTpn;
  return sTpn;
}

// Ticks that a column of `nnz` nonzeros (this rank's part) is expected to take,
// 0 for one that fits a single pass (thread 0).
__device__ __forceinline__ long long aheadTicks(int64_t nnz, int64_t onePass) {
  return nnz > onePass ? (long long)(longTicksPerNnz() * (float)nnz) : 0;
}

// True when the launch's time quantum has run out, or would run out within
// `nnzAhead` nonzeros of a column of more than one pass (block-uniform).  A
// launch lasts until its last slot stops, so a workgroup that stops a column
// early idles alone while one that overruns holds every other workgroup.
__device__ __forceinline__ bool quantumOver(long long deadline, int64_t nnzAhead = 0,
                                            int64_t onePass = 0) {
  __shared__ int sOver;
  if (deadline <= 0) return false;
  __syncthreads();
  if (threadIdx.x == 0) sOver = wall_clock64() + aheadTicks(nnzAhead, onePass) >= deadline;
  __syncthreads();
  return sOver != 0;
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

template <int NMAX>
struct PassLds {
  static constexpr int kItems = NMAX;  // workgroup size x pass width
  double v1[NMAX];
  double v2[NMAX];
  unsigned char start[NMAX];
  double c1[kMaxBundle], c2[kMaxBundle], n1[kMaxBundle], n2[kMaxBundle];
  int cValid[kMaxBundle], nValid[kMaxBundle];
  int cKey, nKey;  // Cox: pid of the run carried into / out of a pass
};

// One lane's items of a pass.  In a slot of F fits (a bundle) lane u of a pass
// carries nonzero c0 + (u >> log2 F) for fit u & (F-1): the F fits of one
// nonzero sit in adjacent lanes and read the F adjacent entries of the
// interleaved row and group state (entry r*F + f), one line for all of them.
template <int IT, typename R>
struct Lanes {
  int row[IT];
  int key[IT];
  bool own[IT];    // valid nonzero that ends its run (always, for independent rows)
  bool start[IT];  // valid nonzero that starts its run (always, for independent rows)
  R x[IT], o[IT], xo[IT];
  double den[IT], wn[IT], w[IT];
};

// Loads one pass of nonzeros [c0, ...) of a column part ending at e (column
// start b) for the lanes whose fit runs (on).
template <int MODEL, int IT, typename R>
__device__ __forceinline__ void passLoad(const Slot<R>& s, int fmt, int64_t b, int64_t e,
                                         int64_t c0, int useW, bool on, Lanes<IT, R>& L) {
  const DevX& d = s.d;
  const int F = s.F, lF = s.lF;
  const int f = threadIdx.x & (F - 1);
  int kp[IT], kn[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int64_t p = c0 + (((int64_t)it * blockDim.x + threadIdx.x) >> lF);
    const bool valid = on && p < e;
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
    const size_t rf = (size_t)r * F + f;
    const RowSt<R> rs = valid ? s.row[rf] : RowSt<R>{(R)0, (R)0};
    L.o[it] = rs.e;
    L.xo[it] = rs.x;
    L.w[it] = (MODEL != M_CLR && useW && valid) ? s.kw[rf] : 1.0;
    const GrpSt gs = (MODEL == M_CLR && L.own[it]) ? s.grp[(size_t)L.key[it] * F + f]
                                                   : GrpSt{1.0, 0.0};
    L.den[it] = MODEL == M_CLR ? gs.den : (valid ? (double)s.rden[rf] : 1.0);
    L.wn[it] = gs.nw;
  }
}

// Run totals of NV per-nonzero values for the owners of one pass (see above).
// `staged` is uniform: false when every run of the column has one nonzero.
// nLocal counts the pass's (nonzero, fit) pairs; a fit's pairs are F apart, and
// each fit keeps its own carry for a run that continues into the next pass.
template <int IT, int NV, class LDS>
__device__ __forceinline__ void runSums(LDS& S, int nLocal, const bool (&own)[IT],
                                        const bool (&start)[IT], double (&val)[IT][NV],
                                        bool staged, int F, bool on) {
  if (!staged) return;  // every owner's run is its own nonzero
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int li = it * blockDim.x + threadIdx.x;
    if (li < nLocal) {
      S.v1[li] = val[it][0];
      if constexpr (NV > 1) S.v2[li] = val[it][1];
      S.start[li] = start[it] ? 1 : 0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int li = it * blockDim.x + threadIdx.x;
    const int f = li & (F - 1);
    const bool last = li >= nLocal - F;  // the pass's last nonzero
    if (!on || li >= nLocal || !(own[it] || last)) continue;
    if (own[it] && start[it]) {
      if (last) S.nValid[f] = 0;
      continue;
    }
    int k = li;
    while (k >= F && !S.start[k]) k -= F;
    const bool before = !S.start[k];  // run began in an earlier pass
    double s1 = (before && S.cValid[f]) ? S.c1[f] : 0.0;
    double s2 = (before && S.cValid[f]) ? S.c2[f] : 0.0;
    for (int q = k; q <= li; q += F) {
      s1 += S.v1[q];
      if constexpr (NV > 1) s2 += S.v2[q];
    }
    if (own[it]) {
      val[it][0] = s1;
      if constexpr (NV > 1) val[it][1] = s2;
      if (last) S.nValid[f] = 0;
    } else {  // last nonzero of the pass, run continues
      S.n1[f] = s1;
      S.n2[f] = s2;
      S.nValid[f] = 1;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < F) {
    S.c1[threadIdx.x] = S.n1[threadIdx.x];
    S.c2[threadIdx.x] = S.n2[threadIdx.x];
    S.cValid[threadIdx.x] = S.nValid[threadIdx.x];
  }
}

template <int MODEL, int IT, typename R, class LDS>
__device__ __forceinline__ void passGather(int fmt, LDS& S, int nLocal, bool staged,
                                           const Lanes<IT, R>& L, int useW, int F, bool on,
                                           double (&v)[2]) {
  const bool ind = isIndicator(fmt);
  if (MODEL == M_CLR) {
    double val[IT][2];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const R m = L.o[it] * L.x[it];
      val[it][0] = L.row[it] >= 0 ? (double)m : 0.0;
      val[it][1] = (L.row[it] >= 0 && !ind) ? (double)(m * L.x[it]) : 0.0;
    }
    runSums<IT, 2>(S, nLocal, L.own, L.start, val, staged, F, on);
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
        const R dd = (R)L.den[it];
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

// Update of one pass by this lane's fit's step `delta` (0: the fit stays).
template <int MODEL, int IT, typename R, class LDS>
__device__ __forceinline__ void passUpdate(const Slot<R>& s, LDS& S, int nLocal, bool staged,
                                           const Lanes<IT, R>& L, double delta, bool on) {
  const int F = s.F;
  const int f = threadIdx.x & (F - 1);
  const R rd = (R)delta;
  double val[IT][1];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    val[it][0] = 0.0;
    const int r = L.row[it];
    if (r < 0 || delta == 0.0) continue;
    const R nx = L.xo[it] + rd * L.x[it];
    const R no = rexp(nx);
    s.row[(size_t)r * F + f] = RowSt<R>{nx, no};
    // the row's denominator moves by the change (independent rows)
    if (MODEL != M_CLR) s.rden[(size_t)r * F + f] = (R)L.den[it] + (no - L.o[it]);
    val[it][0] = (double)no - (double)L.o[it];
  }
  if (MODEL == M_CLR) {
    runSums<IT, 1>(S, nLocal, L.own, L.start, val, staged, F, on && delta != 0.0);
#pragma unroll
    for (int it = 0; it < IT; ++it)
      if (L.own[it] && delta != 0.0) s.grp[(size_t)L.key[it] * F + f].den = L.den[it] + val[it][0];
  }
}

// Records of one pass at pair offset `off` of the column part.
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
    rec.den[q] = L.den[it];
    if (MODEL == M_CLR) {
      rec.key[q] = L.key[it];
      rec.flags[q] = (L.own[it] ? 1 : 0) | (L.start[it] ? 2 : 0);
    }
  }
}

template <int MODEL, int IT, typename R>
__device__ __forceinline__ void loadRecords(const DevX& d, int fmt, int64_t e, int64_t c0,
                                            int64_t off, const Records<R>& rec, bool on, int lF,
                                            Lanes<IT, R>& L) {
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int64_t p = c0 + (((int64_t)it * blockDim.x + threadIdx.x) >> lF);
    const int64_t q = off + (int64_t)it * blockDim.x + threadIdx.x;
    const bool valid = on && p < e;
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
      L.den[it] = valid ? rec.den[q] : 1.0;
    }
  }
}

// Per-fit totals of NV values when lane u carries fit u & (F-1): every lane gets
// the totals of its own fit.  F = 1 sums in blockSum's order (waveSum).
template <int NV>
__device__ __forceinline__ void blockSumFits(double (&v)[NV], int F) {
  __shared__ double sh[NV][kMaxBlockThreads / kWave][kMaxBundle];
  __shared__ double tot[NV][kMaxBundle];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  if (F == 1) {  // uniform
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = waveSum(v[i]);
  } else {
#pragma unroll
    for (int i = 0; i < NV; ++i)
      for (int off = kWave / 2; off >= F; off >>= 1) v[i] += __shfl_down(v[i], off, kWave);
  }
  if (lane < F) {
#pragma unroll
    for (int i = 0; i < NV; ++i) sh[i][wid][lane] = v[i];
  }
  __syncthreads();
  // thread i * F + f adds value i of fit f over the waves, in wave order
  if ((int)threadIdx.x < NV * F) {
    const int i = (int)threadIdx.x / F, ff = (int)threadIdx.x & (F - 1);
    tot[i][ff] = sumWaves(&sh[i][0][ff], kMaxBundle, nw);
  }
  __syncthreads();
  const int f = threadIdx.x & (F - 1);
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = tot[i][f];
  __syncthreads();
}

// ------------------------------------------------------------ gangs
// XCD (accelerator complex die) this wave runs on.
__device__ __forceinline__ int xccId() {
  unsigned v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return (int)(v & 0xfu);
}

__global__ void k_xcc_probe(int* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = xccId();
}

// A slot may be run by a gang of G workgroups (G a power of two <= kShards).
// Rank r owns the rows of shards [r*S/G, (r+1)*S/G) -- their xBeta, exp(xBeta),
// weights and (CLR) strata -- and walks only its part of every column.  Per
// coordinate the ranks exchange their block-reduced (gradient, Hessian) through
// the slot's exchange buffer and all sum them in rank order, so every rank takes
// the same prior step and updates its own rows; no rank reads another's rows.
// The coefficient, its bound and the end-of-quantum decision belong to the lead
// rank (rank 0): only it reads and writes them in memory during a gang launch,
// and it publishes them in the same exchange.  A gang runs inside one XCD.
// Exchanged values travel as agent-scope atomics: a rank publishes its values
// by returning atomic swaps and increments the slot's arrival counter only
// after the swaps have returned (a plain or atomic store may be acknowledged
// before it is performed, and a waiter could then see the arrival first).
// Spins are bounded: a rank that waits longer than kSpinLimit polls (or the
// slot's own limit) gives up and the launch reports the failure; the host then
// restarts the slot's fits cold on one workgroup.
struct Gang {
  int rank, G;
  int rowLo, rowHi;  // rows owned by this rank
  int s0, s1;        // shard range
  unsigned seq;      // exchanges done in this launch
  long long spin;    // polls before a wait gives up
  Xch* x;
  bool ok;
};

constexpr long long kSpinLimit = 1ll << 24;

__device__ __forceinline__ double loadShared(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Publish by atomic swap: its return means the value has been performed at the
// coherence point (a store may be acknowledged before that), so an arrival
// issued after the returns cannot overtake it.
__device__ __forceinline__ double publishShared(double* p, double v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum NS block totals over the gang (in rank order) and receive NL values from
// the lead rank; every thread of every rank gets the same results.  `sum` holds
// this rank's totals (the same in every thread), `lead` the lead rank's values
// (read in thread 0 of rank 0 only).
template <int NS, int NL>
__device__ __forceinline__ bool gangExchange(Gang& g, double (&sum)[NS], double (&lead)[NL]) {
  static_assert(NS <= kXchSum && NL <= kXchLead, "exchange too wide");
  __shared__ double sOut[2][kXchSum + kXchLead];
  __shared__ int sOk[2];
  const int par = (int)(g.seq & 1u);
  if (threadIdx.x == 0) {
    double* mine = &g.x->v[par][g.rank][0];
    double chk = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) chk += publishShared(mine + i, sum[i]);
    if (g.rank == 0) {
#pragma unroll
      for (int i = 0; i < NL; ++i) chk += publishShared(mine + kXchSum + i, lead[i]);
    }
    // consume the swaps' results, so they are returning atomics
    asm volatile("; swap results %0" : : "v"(chk));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);  // ... and have returned
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_fetch_add(&g.x->cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = (unsigned)g.G * (g.seq + 1u);
    long long spins = 0;
    int ok = 1;
    while (__hip_atomic_load(&g.x->cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > g.spin) {
        ok = 0;
        break;
      }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    sOk[par] = ok;
  }
  __syncthreads();
  const bool ok = sOk[par] != 0;
  if (threadIdx.x < kWave) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      double t = (ok && (int)threadIdx.x < g.G) ? loadShared(&g.x->v[par][threadIdx.x][i]) : 0.0;
      t = waveSum(t);
      if (threadIdx.x == 0) sOut[par][i] = t;
    }
    if ((int)threadIdx.x < NL)
      sOut[par][kXchSum + threadIdx.x] = ok ? loadShared(&g.x->v[par][0][kXchSum + threadIdx.x]) : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NS; ++i) sum[i] = sOut[par][i];
#pragma unroll
  for (int i = 0; i < NL; ++i) lead[i] = sOut[par][kXchSum + i];
  g.seq++;
  g.ok = g.ok && ok;
  return ok;
}

// gangExchange with run-time widths through LDS arrays: on entry xs[0..ns) and
// xl[0..nl) hold this rank's values (written by any threads of the block); on
// return xs holds the sums over the gang in rank order and xl the lead rank's
// values, for every thread.
__device__ __forceinline__ bool gangExchangeN(Gang& g, double* xs, int ns, double* xl, int nl) {
  __shared__ int sOk[2];
  const int par = (int)(g.seq & 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    double* mine = &g.x->v[par][g.rank][0];
    double chk = 0.0;
    for (int i = 0; i < ns; ++i) chk += publishShared(mine + i, xs[i]);
    if (g.rank == 0)
      for (int i = 0; i < nl; ++i) chk += publishShared(mine + kXchSum + i, xl[i]);
    asm volatile("; swap results %0" : : "v"(chk));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_fetch_add(&g.x->cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = (unsigned)g.G * (g.seq + 1u);
    long long spins = 0;
    int ok = 1;
    while (__hip_atomic_load(&g.x->cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > g.spin) {
        ok = 0;
        break;
      }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    sOk[par] = ok;
  }
  __syncthreads();
  const bool ok = sOk[par] != 0;
  if (threadIdx.x < kWave) {
    // a chunk's loads all in flight before its first sum; chunks of 8 values
    // (one chunk unless a bundle exchanges more), so that the exchange holds
    // few registers beside the lane state live across it
    constexpr int kChunk = 8;
    const bool mine = ok && (int)threadIdx.x < g.G;
    const double l = (ok && (int)threadIdx.x < nl) ? loadShared(&g.x->v[par][0][kXchSum + threadIdx.x]) : 0.0;
    for (int c0 = 0; c0 < ns; c0 += kChunk) {
      double t[kChunk];
#pragma unroll
      for (int i = 0; i < kChunk; ++i)
        t[i] = (mine && c0 + i < ns) ? loadShared(&g.x->v[par][threadIdx.x][c0 + i]) : 0.0;
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        if (c0 + i >= ns) break;
        const double r = waveSum(t[i]);
        if (threadIdx.x == 0) xs[c0 + i] = r;
      }
    }
    if ((int)threadIdx.x < nl) xl[threadIdx.x] = l;
  }
  __syncthreads();
  g.seq++;
  g.ok = g.ok && ok;
  return ok;
}

// A solo-run boundary of a gang (see cycleFit): every rank's row and group
// writes before it are visible to every rank's reads after it -- an agent-scope
// release before arriving and an acquire after leaving -- and the lead's two
// values xl[0..1] (its quantum decision and where it stopped) reach every rank.
__device__ __forceinline__ bool gangSoloSync(Gang& g, double* xl) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  double none[1] = {0.0};
  const bool ok = gangExchangeN(g, none, 0, xl, 2);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok;
}

// Column metadata of a window of kMetaW coordinates, staged in LDS by one
// coalesced load per window: the sweep's per-column reads of the fixed flag, the
// block length, the format, the column's extent (this rank's part of it in a
// gang) and its stratum-run count then come from LDS instead of a chain of
// dependent global loads ahead of each column's gather.
constexpr int kMetaW = 256;
struct ColMeta {
  int64_t cb[kMetaW + 1];  // column starts (window-relative index)
  int32_t b[kMetaW], e[kMetaW];  // this rank's part, relative to the column start
  int32_t runs[kMetaW];          // stratum runs (grouped models)
  uint8_t fixed[kMetaW], blk[kMetaW], fmt[kMetaW];
  int w0;
};

template <int MODEL, typename R>
__device__ __forceinline__ void loadColMeta(ColMeta& M, const Slot<R>& s, int w0, const Gang& gg) {
  const DevX& d = s.d;
  const int J = s.J;
  __syncthreads();  // the previous window's readers are done
  for (int t = threadIdx.x; t <= kMetaW; t += blockDim.x) {
    const int j = w0 + t;
    if (j <= J) M.cb[t] = d.colPtr[j];
    if (t < kMetaW && j < J) {
      M.fixed[t] = s.fixed[j];
      M.blk[t] = d.blk != nullptr ? d.blk[j] : 1;
      M.fmt[t] = d.fmt[j];
      M.runs[t] = MODEL == M_CLR ? (int32_t)(d.segPtr[j + 1] - d.segPtr[j]) : 0;
      if (gg.G > 1) {
        const int32_t* off = d.colShard + (size_t)j * (d.nShards + 1);
        M.b[t] = off[gg.s0];
        M.e[t] = off[gg.s1];
      } else {
        M.b[t] = 0;
        M.e[t] = -1;  // the whole column
      }
    }
  }
  if (threadIdx.x == 0) M.w0 = w0;
  __syncthreads();
}

// The first coordinate from k that ends a lead-alone run -- a free column of
// more than `soloMax` nonzeros -- or J: the ranks other than the lead skip the
// run's columns with one search of each metadata window instead of a visit per
// column (block-uniform; leaves the window that holds the result in M).
template <int MODEL, typename R>
__device__ __forceinline__ int soloRunEnd(ColMeta& M, const Slot<R>& s, int k, int64_t soloMax,
                                          const Gang& gg) {
  __shared__ int sEnd;
  const int J = s.J;
  while (k < J) {
    if (k + kBlockCols > M.w0 + kMetaW) loadColMeta<MODEL, R>(M, s, k, gg);
    if (threadIdx.x == 0) sEnd = J;
    __syncthreads();
    const int lim = min(kMetaW, J - M.w0);
    for (int t = k - M.w0 + (int)threadIdx.x; t < lim; t += blockDim.x)
      if (!M.fixed[t] && M.cb[t + 1] - M.cb[t] > soloMax) atomicMin(&sEnd, M.w0 + t);
    __syncthreads();
    const int e = sEnd;
    __syncthreads();  // every thread has read sEnd before it is set again
    if (e < J) return e;
    k = M.w0 + lim;
  }
  return J;
}

// One coordinate cycle of a slot's fits (independent rows or grouped) over this
// rank's rows, from coordinate j0; `act` has bit f set for each fit still
// running (the others' lanes load and store nothing).  Every lane takes its
// fit's (deterministic) prior step from the reduced gradient and Hessian, so no
// broadcast barrier is needed; a barrier follows an update only.  Returns the
// coordinate to resume at when the quantum ran out (J when the cycle
// completed), or -1 when a gang exchange failed.
template <int MODEL, typename R, int ITL, class LDS>
__device__ __forceinline__ int cycleFit(const Slot<R>& s, FitStatus* st, LDS& S, int j0,
                                        long long deadline, Gang& gg, unsigned act) {
  const DevX& d = s.d;
  const int F = s.F, lF = s.lF;
  const int f = threadIdx.x & (F - 1);
  const bool on = (act >> f) & 1u;
  const int nAct = __popc(act);
  const int64_t CH = ((int64_t)blockDim.x * kIT) >> lF;    // nonzeros of a single pass
  const int64_t CHL = ((int64_t)blockDim.x * ITL) >> lF;  // nonzeros of a long-column pass
  const int J = s.J;
  const int useW = s.ctl.useW;
  const bool multi = gg.G > 1;
  const bool lead = gg.rank == 0;
  __shared__ double xs[kXchSum], xl[kXchLead];
  // work counters and phase ticks of the cycle (gathered and updated nonzeros
  // and runs; device ticks of single-pass and multi-pass columns, multi-pass
  // visits), kept in LDS by thread 0 rather than in every lane's registers
  enum { kCG, kCGR, kCU, kCUR, kTS, kTL, kNL, kNCtr };
  __shared__ long long sCtr[kNCtr];
  if (threadIdx.x == 0)
    for (int i = 0; i < kNCtr; ++i) sCtr[i] = 0;
  auto count = [&](int i, long long v) {
    if (threadIdx.x == 0) sCtr[i] += v;
  };
  int jEnd = J;
  // Records are indexed by position in the column, and a rank's positions move
  // from column to column: a rank may already store coordinate k+1's records
  // while another still reads coordinate k's.  A gang alternates two buffers
  // (no rank reaches coordinate k+2 before every rank has finished k).
  unsigned recSel = 0;
  __shared__ ColMeta M;
  loadColMeta<MODEL, R>(M, s, j0, gg);
  // Solo runs (gangs): consecutive short columns -- one pass of this workgroup at
  // most -- are gathered, stepped and updated by the lead rank alone over the
  // whole column while the other ranks skip ahead, so they cost no exchange.
  // Entering and leaving a run the ranks meet (gangSoloSync), which orders every
  // rank's row writes before the meeting before every read after it.  A lead
  // whose quantum runs out inside a run stops there and tells the others when
  // they meet it at the run's end.
  const int64_t soloMax = (multi && s.ctl.solo >= 0) ? CH : -1;
  bool inSolo = false, soloStop = false;
  auto soloSync = [&](int jStop) -> bool {
    if (threadIdx.x == 0 && lead) {
      xl[0] = soloStop ? 1.0 : 0.0;
      xl[1] = (double)jStop;
    }
    return gangSoloSync(gg, xl);
  };
  for (int j = j0; j < J; ++j) {
    if (j + kBlockCols > M.w0 + kMetaW) loadColMeta<MODEL, R>(M, s, j, gg);  // uniform
    const int mj = j - M.w0;
    // The quantum is checked every fourth coordinate, and before every column
    // of more than one pass (milliseconds for a workgroup alone), with that
    // column's expected time: the launch lasts until its last slot stops.
    const int64_t nCol = M.fixed[mj] ? 0 : M.cb[mj + 1] - M.cb[mj];
    if ((!multi || (inSolo && lead)) && j > j0 && ((j & 3) == 0 || nCol > CH) &&
        quantumOver(deadline, nCol, CH)) {
      soloStop = multi;
      jEnd = j;
      break;
    }
    if (M.fixed[mj]) continue;
    if (multi) {
      const bool soloCol = M.cb[mj + 1] - M.cb[mj] <= soloMax;
      if (soloCol != inSolo) {
        if (!soloSync(j)) return -1;
        const bool stopped = inSolo && xl[0] != 0.0;
        inSolo = soloCol;
        if (stopped) {  // the lead's quantum ran out inside the run just left
          inSolo = false;
          jEnd = (int)xl[1];
          break;
        }
      }
      if (inSolo && !lead) {  // on to the column that ends the run
        j = soloRunEnd<MODEL, R>(M, s, j + 1, soloMax, gg) - 1;
        continue;
      }
    }
    const bool part = multi && !inSolo;  // this rank's part of the column, one exchange
    // A block of short columns with pairwise disjoint rows (strata), swept by
    // one workgroup (a fit alone, or a gang's lead in a lead-alone run): each
    // column is gathered on its own, then one block reduction serves the whole
    // block, then each column takes its step and its update.  No column's
    // update changes another's gradient, so this is the sequential result.
    // Block columns are short, so in a gang they run on the lead alone; with
    // lead-alone runs off, a gang sweeps them one at a time with an exchange
    // each (keeping the exchange out of this path keeps its registers low).
    const int B = (F == 1 && MODEL != M_COX && d.blk != nullptr && !part) ? M.blk[mj] : 1;
    if (B > 1) {
      // nonzeros per lane of a block column: one pass of the workgroup covers it
      constexpr int kLanes = LDS::kItems / ITL;
      constexpr int kBlkIT = (kBlockNnz + kLanes - 1) / kLanes;
      static_assert(kBlkIT * kLanes >= kBlockNnz, "a block column must fit one pass");
      const long long t0 = wall_clock64();
      Lanes<kBlkIT, R> L[kBlockCols];
      double v[2 * kBlockCols], bj[kBlockCols], bnd[kBlockCols], delta[kBlockCols];
      int nAll[kBlockCols], nMine[kBlockCols], runsC[kBlockCols];  // <= kBlockNnz each
      bool live[kBlockCols], stg[kBlockCols];
#pragma unroll
      for (int c = 0; c < kBlockCols; ++c) {
        const int jj = j + c;
        v[2 * c] = v[2 * c + 1] = 0.0;
        bj[c] = bnd[c] = delta[c] = 0.0;
        nAll[c] = nMine[c] = 0;
        runsC[c] = 0;
        stg[c] = false;
        live[c] = c < B && !M.fixed[mj + c];  // uniform
        if (!live[c]) continue;
        const int64_t cb = M.cb[mj + c], ce = M.cb[mj + c + 1];
        const int64_t b = cb, e = ce;
        nAll[c] = (int)(ce - cb);
        nMine[c] = (int)(e - b);
        const int fmtc = M.fmt[mj + c];
        bj[c] = s.beta[jj];
        bnd[c] = s.bound[jj];
        runsC[c] = MODEL == M_CLR ? M.runs[mj + c] : nAll[c];
        stg[c] = MODEL == M_CLR && runsC[c] < nAll[c];
        if (threadIdx.x == 0) S.cValid[0] = 0;
        passLoad<MODEL, kBlkIT, R>(s, fmtc, cb, e, b, useW, true, L[c]);
        double vv[2] = {0.0, 0.0};
        passGather<MODEL, kBlkIT, R>(fmtc, S, (int)nMine[c], stg[c], L[c], useW, 1, true, vv);
        v[2 * c] = vv[0];
        v[2 * c + 1] = vv[1];
      }
      blockSumFits<2 * kBlockCols>(v, 1);
      bool moved = false;
#pragma unroll
      for (int c = 0; c < kBlockCols; ++c) {
        if (!live[c]) continue;
        const int jj = j + c;
        double g = nAll[c] > 0 ? v[2 * c] - s.xjy[jj] : 0.0, h = nAll[c] > 0 ? v[2 * c + 1] : 0.0;
        if (h < 0.0) {
          g = 0.0;
          h = 0.0;
        }
        delta[c] = applyBounds(priorDelta(s.ptype[jj], s.pvar[jj], g, h, bj[c]), &bnd[c]);
        if (threadIdx.x == 0 && lead) {
          s.bound[jj] = bnd[c];
          if (delta[c] != 0.0) s.beta[jj] = bj[c] + delta[c];
        }
        moved = moved || delta[c] != 0.0;
        count(kCG, nMine[c]);
        count(kCGR, MODEL == M_CLR && lead ? runsC[c] : 0);
      }
#pragma unroll
      for (int c = 0; c < kBlockCols; ++c) {
        if (!live[c] || delta[c] == 0.0 || nMine[c] == 0) continue;
        if (threadIdx.x == 0) S.cValid[0] = 0;
        passUpdate<MODEL, kBlkIT, R>(s, S, (int)nMine[c], stg[c], L[c], delta[c], true);
        count(kCU, nMine[c]);
        count(kCUR, MODEL == M_CLR && lead ? runsC[c] : 0);
      }
      if (moved) __syncthreads();
      count(kTS, wall_clock64() - t0);
      j += B - 1;  // the quantum ends at the check on top of the loop
      continue;
    }
    Records<R> rec = s.rec;
    if (part && (recSel++ & 1u)) {
      rec.row += s.recStride;
      rec.o += s.recStride;
      rec.xo += s.recStride;
      rec.den += s.recStride;
      if (MODEL == M_CLR) {
        rec.key += s.recStride;
        rec.flags += s.recStride;
      }
    }
    const int64_t cb = M.cb[mj], ce = M.cb[mj + 1];
    int64_t b = cb, e = ce;
    if (part) {
      b = cb + M.b[mj];
      e = cb + M.e[mj];
    }
    const int64_t nj = e - b;      // this rank's nonzeros
    const int64_t njAll = ce - cb;  // the column's
    const int fmt = M.fmt[mj];
    // this lane's fit; in a gang only the lead rank's reads count (it publishes them)
    const size_t jf = (size_t)j * F + f;
    double bj = s.beta[jf];
    double bnd = s.bound[jf];
    const double pv = s.pvar[jf], xy = s.xjy[jf];
    const int pt = s.ptype[jf];
    const long long runs = MODEL == M_CLR ? (long long)M.runs[mj] : njAll;
    const bool staged = MODEL == M_CLR && runs < njAll;
    if ((int)threadIdx.x < F) S.cValid[threadIdx.x] = 0;
    const long long t0 = wall_clock64();
    double v[2] = {0.0, 0.0};
    bool stop = false;
    // the exchange also carries the lead rank's end-of-quantum decision
    auto reduce = [&]() -> bool {
      blockSumFits<2>(v, F);
      if (!part) return true;
      if ((int)threadIdx.x < F) {  // lane f holds fit f's totals and coefficient state
        xs[2 * threadIdx.x] = v[0];
        xs[2 * threadIdx.x + 1] = v[1];
        xl[2 * threadIdx.x] = bj;
        xl[2 * threadIdx.x + 1] = bnd;
      }
      if (threadIdx.x == 0) {  // this rank's part of the next column, when the window holds it
        const int64_t nNext = (mj + 1 < kMetaW && j + 1 < J && !M.fixed[mj + 1])
                                  ? (int64_t)M.e[mj + 1] - M.b[mj + 1]
                                  : 0;
        xl[2 * F] = (deadline > 0 && wall_clock64() + aheadTicks(nNext, CH) >= deadline) ? 1.0 : 0.0;
      }
      if (!gangExchangeN(gg, xs, 2 * F, xl, 2 * F + 1)) return false;
      v[0] = xs[2 * f];
      v[1] = xs[2 * f + 1];
      bj = xl[2 * f];
      bnd = xl[2 * f + 1];
      stop = xl[2 * F] != 0.0;
      return true;
    };
    auto step = [&]() -> double {
      if (!on) return 0.0;
      double g = njAll > 0 ? v[0] - xy : 0.0, h = njAll > 0 ? v[1] : 0.0;
      if (h < 0.0) {
        g = 0.0;
        h = 0.0;
      }
      const double dl = applyBounds(priorDelta(pt, pv, g, h, bj), &bnd);
      if ((int)threadIdx.x < F && lead) {  // thread f writes fit f's
        s.bound[jf] = bnd;
        if (dl != 0.0) s.beta[jf] = bj + dl;
      }
      return dl;
    };
    // fits that take a step (all lanes of a fit agree)
    auto moving = [&](double dl) -> int {
      if (F == 1) return dl != 0.0 ? 1 : 0;
      return __syncthreads_count(dl != 0.0) / ((int)blockDim.x / F);
    };
    double delta;
    int mv;
    if (nj <= CH) {
      Lanes<kIT, R> L;
      passLoad<MODEL, kIT, R>(s, fmt, cb, e, b, useW, on, L);
      passGather<MODEL, kIT, R>(fmt, S, (int)(nj << lF), staged, L, useW, F, on, v);
      if (!reduce()) return -1;
      delta = step();
      mv = moving(delta);
      if (mv > 0 && nj > 0) {
        passUpdate<MODEL, kIT, R>(s, S, (int)(nj << lF), staged, L, delta, on);
        __syncthreads();
      }
    } else {
      for (int64_t c0 = b; c0 < e; c0 += CHL) {
        Lanes<ITL, R> L;
        passLoad<MODEL, ITL, R>(s, fmt, cb, e, c0, useW, on, L);
        storeRecords<MODEL, ITL, R>(rec, (c0 - cb) << lF, L);
        passGather<MODEL, ITL, R>(fmt, S, (int)(min(CHL, e - c0) << lF), staged, L, useW, F, on, v);
      }
      if (!reduce()) return -1;
      delta = step();
      if ((int)threadIdx.x < F) S.cValid[threadIdx.x] = 0;
      mv = moving(delta);
      if (mv > 0) {
        __syncthreads();  // carry reset visible; records were written by these lanes
        for (int64_t c0 = b; c0 < e; c0 += CHL) {
          Lanes<ITL, R> L;
          loadRecords<MODEL, ITL, R>(d, fmt, e, c0, (c0 - cb) << lF, rec, on, lF, L);
          passUpdate<MODEL, ITL, R>(s, S, (int)(min(CHL, e - c0) << lF), staged, L, delta, on);
        }
        __syncthreads();
      }
    }
    const long long t1 = wall_clock64();
    if (nj > CH) {
      count(kTL, t1 - t0);
      count(kNL, 1);
      if (threadIdx.x == 0) {  // the last column's rate, decayed slowly towards a faster one
        float& tpn = longTicksPerNnz();
        tpn = fmaxf((float)(t1 - t0) / (float)nj, 0.75f * tpn);
      }
    } else {
      count(kTS, t1 - t0);
    }
    count(kCG, nj * nAct);
    count(kCGR, MODEL == M_CLR && lead ? runs * nAct : 0);
    count(kCU, nj * mv);
    count(kCUR, MODEL == M_CLR && lead ? runs * mv : 0);
    if (stop) {
      jEnd = j + 1;
      break;
    }
  }
  if (multi && inSolo) {  // every rank leaves the last run together
    if (!soloSync(jEnd)) return -1;
    if (xl[0] != 0.0) jEnd = (int)xl[1];
  }
  if (threadIdx.x == 0) {
    FitStatus& cs = st[0];
    atomicAdd((unsigned long long*)&cs.tSmall, (unsigned long long)sCtr[kTS]);
    atomicAdd((unsigned long long*)&cs.tLarge, (unsigned long long)sCtr[kTL]);
    atomicAdd((unsigned long long*)&cs.nLarge, (unsigned long long)sCtr[kNL]);
    atomicAdd((unsigned long long*)&cs.gatherNnz, (unsigned long long)sCtr[kCG]);
    atomicAdd((unsigned long long*)&cs.gatherRuns, (unsigned long long)sCtr[kCGR]);
    atomicAdd((unsigned long long*)&cs.updateNnz, (unsigned long long)sCtr[kCU]);
    atomicAdd((unsigned long long*)&cs.updateRuns, (unsigned long long)sCtr[kCUR]);
  }
  return jEnd;
}

// ------------------------------------------------------------ Cox risk-set scans
template <int NV>
__device__ __forceinline__ Seg<NV> segZero() {
  Seg<NV> z;
  z.f = 0;
#pragma unroll
  for (int i = 0; i < NV; ++i) z.v[i] = 0.0;
  return z;
}

// Carry into this rank of a gang-wide segmented scan: every rank publishes the
// segmented total of its pid range and receives the combination, in rank
// order, of the totals of the ranks before it (the identity on rank 0).  Same
// publication protocol as gangExchange; the two share the exchange rounds.
template <int NV>
__device__ __forceinline__ bool gangCarry(Gang& g, const Seg<NV>& total, Seg<NV>& carry) {
  static_assert(NV + 1 <= kXchSum, "exchange too wide");
  __shared__ Seg<NV> sCarry[2];
  __shared__ int sOk[2];
  const int par = (int)(g.seq & 1u);
  if (threadIdx.x == 0) {
    double* mine = &g.x->v[par][g.rank][0];
    double chk = 0.0;
#pragma unroll
    for (int i = 0; i < NV; ++i) chk += publishShared(mine + i, total.v[i]);
    chk += publishShared(mine + NV, (double)total.f);
    asm volatile("; swap results %0" : : "v"(chk));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_fetch_add(&g.x->cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = (unsigned)g.G * (g.seq + 1u);
    long long spins = 0;
    int ok = 1;
    while (__hip_atomic_load(&g.x->cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > g.spin) {
        ok = 0;
        break;
      }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    sOk[par] = ok;
  }
  __syncthreads();
  const bool ok = sOk[par] != 0;
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    Seg<NV> x = segZero<NV>();
    if (ok && lane < g.rank) {
      const double* src = &g.x->v[par][lane][0];
#pragma unroll
      for (int i = 0; i < NV; ++i) x.v[i] = loadShared(src + i);
      x.f = loadShared(src + NV) != 0.0 ? 1 : 0;
    }
    // inclusive scan over lanes in lane (= rank) order
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      Seg<NV> o = segShflUp(x, off);
      if (lane >= off) x = segCombine(o, x);
    }
    if (lane == kWave - 1) sCarry[par] = x;  // ranks >= g.rank contributed the identity
  }
  __syncthreads();
  carry = sCarry[par];
  g.seq++;
  g.ok = g.ok && ok;
  return ok;
}

// Thread t's contiguous block of the pids [pLo, pHi).
__device__ __forceinline__ void pidBlock(int pLo, int pHi, int* a, int* z) {
  const int n = pHi - pLo;
  const int per = (n + (int)blockDim.x - 1) / (int)blockDim.x;
  *a = pLo + min(n, (int)threadIdx.x * per);
  *z = pLo + min(n, ((int)threadIdx.x + 1) * per);
}

// ---- risk-set scans over a rank's pids, one contiguous stretch per wave
// Wave w of the workgroup owns the pids [a, z) of the rank's range (whole steps
// of kWave * kScanItems pids); it walks them step by step, lane l taking the
// kScanItems consecutive pids q0 + l * kScanItems + [0, kScanItems), so one
// step's loads of a wave cover one contiguous stretch of every pid array.  A
// lane folds its pids in order, a wave scan over the lanes' folds orders the
// step, and the step's total carries into the next step.
constexpr int kScanItems = 4;

__device__ __forceinline__ void waveRange(int pLo, int pHi, int* a, int* z) {
  const int nw = ((int)blockDim.x + kWave - 1) / kWave;
  const int wid = (int)threadIdx.x / kWave;
  const int n = pHi - pLo;
  const int step = kWave * kScanItems;
  const int per = ((n + nw - 1) / nw + step - 1) / step * step;
  *a = pLo + min(n, wid * per);
  *z = pLo + min(n, (wid + 1) * per);
}

// Each wave's total -> the combination of the earlier waves' totals (this wave's
// carry) and of all of them (`total`), in wave order.
template <int NV>
__device__ __forceinline__ Seg<NV> waveCarries(const Seg<NV>& waveTotal, Seg<NV>& total) {
  __shared__ Seg<NV> wsh[kMaxBlockThreads / kWave];
  const int wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) wsh[wid] = waveTotal;
  __syncthreads();
  Seg<NV> carry = segZero<NV>();
  for (int w = 0; w < wid; ++w) carry = segCombine(carry, wsh[w]);
  Seg<NV> tot = carry;
  for (int w = wid; w < nw; ++w) tot = segCombine(tot, wsh[w]);
  total = tot;
  __syncthreads();
  return carry;
}

// Segmented sums without segments: a fit whose rows form one stratum has no
// reset after its first pid, and its scans add plainly.
template <int NV, bool SEGS>
__device__ __forceinline__ Seg<NV> scanCombine(const Seg<NV>& a, const Seg<NV>& b) {
  if constexpr (SEGS) {
    return segCombine(a, b);
  } else {
    Seg<NV> r;
    r.f = 0;
#pragma unroll
    for (int i = 0; i < NV; ++i) r.v[i] = a.v[i] + b.v[i];
    return r;
  }
}

// DPP moves between lanes: a lane with no source lane, or in a row the mask
// leaves out, receives 0 (a Seg of zeros with no reset is the identity of the
// combines above).  Every lane of the wave must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dppD(double v) {
  return __builtin_amdgcn_update_dpp(0.0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ int dppI(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}

template <int NV, bool SEGS, int CTRL, int ROWS>
__device__ __forceinline__ void dppScanStep(Seg<NV>& x) {
  Seg<NV> o;
#pragma unroll
  for (int i = 0; i < NV; ++i) o.v[i] = dppD<CTRL, ROWS>(x.v[i]);
  o.f = SEGS ? dppI<CTRL, ROWS>(x.f) : 0;
  x = scanCombine<NV, SEGS>(o, x);
}

// Inclusive (segmented) scan over the wave's lanes in lane order: combining
// with the lanes 1, 2, 4 and 8 back inside each 16-lane row gives every lane its
// row prefix; row_bcast:15 brings row 0's total into row 1 and row 2's into row
// 3, row_bcast:31 brings lane 31's (rows 0-1) into rows 2 and 3.
template <int NV, bool SEGS>
__device__ __forceinline__ Seg<NV> waveScanInclS(Seg<NV> x) {
  dppScanStep<NV, SEGS, 0x111, 0xf>(x);  // row_shr:1
  dppScanStep<NV, SEGS, 0x112, 0xf>(x);  // row_shr:2
  dppScanStep<NV, SEGS, 0x114, 0xf>(x);  // row_shr:4
  dppScanStep<NV, SEGS, 0x118, 0xf>(x);  // row_shr:8
  dppScanStep<NV, SEGS, 0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
  dppScanStep<NV, SEGS, 0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

template <int NV>
__device__ __forceinline__ Seg<NV> waveScanIncl(Seg<NV> x) {
  return waveScanInclS<NV, true>(x);
}

// Lane i - 1's value in lane i (wave_shr:1), the identity in lane 0.
template <int NV>
__device__ __forceinline__ Seg<NV> segPrevLane(const Seg<NV>& a) {
  Seg<NV> r;
#pragma unroll
  for (int i = 0; i < NV; ++i) r.v[i] = dppD<0x138, 0xf>(a.v[i]);
  r.f = dppI<0x138, 0xf>(a.f);
  return r;
}

// Lane 63's value, to every lane.
template <int NV>
__device__ __forceinline__ Seg<NV> segLastLane(const Seg<NV>& a) {
  Seg<NV> r;
#pragma unroll
  for (int i = 0; i < NV; ++i) r.v[i] = lastLane(a.v[i]);
  r.f = __builtin_amdgcn_readlane(a.f, kWave - 1);
  return r;
}

// Gradient and Hessian terms of one coordinate over this rank's pids [pLo, pHi)
// (computeGradientAndHessianImpl's cumulative branch, ModelSpecifics.hpp:904-1020,
// with BreslowTiedCoxProportionalHazards::incrementGradientAndHessian,
// ModelSpecifics.h:1185-1202).  The accumulated denominator and numerators are a
// segmented prefix sum over pids that restarts at each stratum
// (computeAccumlatedDenominator, ModelSpecifics.hpp:1670-1694).  Pass 1 folds
// each wave's stretch into its segmented total; the waves' totals give each
// wave its carry and the rank's total, and a gang scan across ranks the rank's
// carry; pass 2 walks the stretch again with the running sums, adding nW*t and
// the Hessian term to v for the pids with events (nW = 0 adds exactly 0) and
// clearing the numerators for the next coordinate.  NV = 2 (indicator column:
// denominator, numerator) or 3 (with the second numerator); SEGS = false when the
// fit has one stratum (no reset after its first pid).
//
// A range of at most blockDim * kScanSmall pids (risk-set pieces of tied event
// times, or a rank's share of them) is loaded once: thread t holds the pids
// [t*c, (t+1)*c) of it in registers, folds them, takes its carry from one wave
// scan of the thread folds, the waves' totals (and the gang carry) and walks
// them again from registers -- one round of loads instead of two passes of
// them, and one wave scan for all of them (a wave scan per row of blockDim pids
// with coalesced loads instead measured 1.5x slower: its shuffles cost more than
// the strided loads save).
constexpr int kScanSmall = 8;

template <typename R, int NV, bool SEGS>
__device__ __forceinline__ bool coxScanSmall(const Slot<R>& s, int pLo, int pHi, Gang& gg,
                                             double (&v)[2]) {
  const GrpSt* grp = s.grp;
  double* numer = s.numer;
  double* numer2 = s.numer2;
  const int lane = threadIdx.x & (kWave - 1);
  const int n = pHi - pLo;
  const int c = (n + (int)blockDim.x - 1) / (int)blockDim.x;  // <= kScanSmall
  const int a = pLo + min(n, (int)threadIdx.x * c);
  const int z = pLo + min(n, ((int)threadIdx.x + 1) * c);
  Seg<NV> it[kScanSmall];
  double wn[kScanSmall];
#pragma unroll
  for (int u = 0; u < kScanSmall; ++u) {
    it[u] = segZero<NV>();
    wn[u] = 0.0;
    const int i = a + u;
    if (i < z) {
      const GrpSt gs = grp[i];
      it[u].f = SEGS ? s.reset[i] : 0;
      it[u].v[0] = gs.den;
      wn[u] = gs.nw;
      it[u].v[1] = numer[i];
      if constexpr (NV > 2) it[u].v[2] = numer2[i];
    }
  }
  Seg<NV> l = segZero<NV>();
#pragma unroll
  for (int u = 0; u < kScanSmall; ++u) l = scanCombine<NV, SEGS>(l, it[u]);
  const Seg<NV> inc = waveScanInclS<NV, SEGS>(l);
  Seg<NV> r = segPrevLane(inc);
  if (lane == 0) r = segZero<NV>();
  Seg<NV> total;
  Seg<NV> run = waveCarries<NV>(segLastLane(inc), total);
  if (gg.G > 1) {
    Seg<NV> carry;
    if (!gangCarry<NV>(gg, total, carry)) return false;
    run = scanCombine<NV, SEGS>(carry, run);
  }
  r = scanCombine<NV, SEGS>(run, r);
#pragma unroll
  for (int u = 0; u < kScanSmall; ++u) {
    const int i = a + u;
    if (i >= z) break;
    r = scanCombine<NV, SEGS>(r, it[u]);
    if (wn[u] != 0.0) {
      const double accD = r.v[0];
      const double t = r.v[1] / accD;
      const double g = wn[u] * t;
      v[0] += g;
      if constexpr (NV > 2) v[1] += wn[u] * (r.v[2] / accD - t * t);
      else v[1] += g * (1.0 - t);
    }
    if (it[u].v[1] != 0.0) numer[i] = 0.0;
    if constexpr (NV > 2) {
      if (it[u].v[2] != 0.0) numer2[i] = 0.0;
    }
  }
  return true;
}

template <typename R, int NV, bool SEGS>
__device__ __forceinline__ bool coxScanT(const Slot<R>& s, int pLo, int pHi, Gang& gg,
                                         double (&v)[2]) {
  if (pHi - pLo <= (int)blockDim.x * kScanSmall)  // uniform
    return coxScanSmall<R, NV, SEGS>(s, pLo, pHi, gg, v);
  const GrpSt* grp = s.grp;
  double* numer = s.numer;
  double* numer2 = s.numer2;
  const uint8_t* reset = s.reset;
  const int lane = threadIdx.x & (kWave - 1);
  constexpr int kStep = kWave * kScanItems;
  auto load = [&](int i, Seg<NV>& x) {
    x.f = SEGS ? reset[i] : 0;
    x.v[0] = grp[i].den;
    x.v[1] = numer[i];
    if constexpr (NV > 2) x.v[2] = numer2[i];
  };
  int a, z;
  waveRange(pLo, pHi, &a, &z);
  Seg<NV> wagg = segZero<NV>();
  for (int q0 = a; q0 < z; q0 += kStep) {
    const int i0 = q0 + lane * kScanItems;
    Seg<NV> l = segZero<NV>();
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
      if (i0 + u < z) {
        Seg<NV> x;
        load(i0 + u, x);
        l = scanCombine<NV, SEGS>(l, x);
      }
    }
    wagg = scanCombine<NV, SEGS>(wagg, segLastLane(waveScanInclS<NV, SEGS>(l)));
  }
  Seg<NV> total;
  Seg<NV> run = waveCarries<NV>(wagg, total);
  if (gg.G > 1) {
    Seg<NV> carry;
    if (!gangCarry<NV>(gg, total, carry)) return false;
    run = scanCombine<NV, SEGS>(carry, run);
  }
  for (int q0 = a; q0 < z; q0 += kStep) {
    const int i0 = q0 + lane * kScanItems;
    Seg<NV> it[kScanItems];
    Seg<NV> l = segZero<NV>();
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
      it[u] = segZero<NV>();
      if (i0 + u < z) load(i0 + u, it[u]);
      l = scanCombine<NV, SEGS>(l, it[u]);
    }
    const Seg<NV> inc = waveScanInclS<NV, SEGS>(l);
    Seg<NV> r = segPrevLane(inc);
    if (lane == 0) r = segZero<NV>();
    r = scanCombine<NV, SEGS>(run, r);
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
      const int i = i0 + u;
      if (i >= z) break;
      r = scanCombine<NV, SEGS>(r, it[u]);
      const double wn = grp[i].nw;
      if (wn != 0.0) {
        const double accD = r.v[0];
        const double t = r.v[1] / accD;
        const double g = wn * t;
        v[0] += g;
        if constexpr (NV > 2) v[1] += wn * (r.v[2] / accD - t * t);
        else v[1] += g * (1.0 - t);
      }
      if (it[u].v[1] != 0.0) numer[i] = 0.0;
      if constexpr (NV > 2) {
        if (it[u].v[2] != 0.0) numer2[i] = 0.0;
      }
    }
    run = scanCombine<NV, SEGS>(run, segLastLane(inc));
  }
  return true;
}

template <typename R>
__device__ __forceinline__ bool coxScan(const Slot<R>& s, int pLo, int pHi, bool ind, Gang& gg,
                                        double (&v)[2]) {
  const bool segs = s.d.nStrata > 1;
  if (ind) return segs ? coxScanT<R, 2, true>(s, pLo, pHi, gg, v) : coxScanT<R, 2, false>(s, pLo, pHi, gg, v);
  return segs ? coxScanT<R, 3, true>(s, pLo, pHi, gg, v) : coxScanT<R, 3, false>(s, pLo, pHi, gg, v);
}

// Sum of nW_i * log(accumulated denominator_i) over this rank's pids (the Cox
// term of getLogLikelihood, ModelSpecifics.hpp:624-628), by the same two-pass
// scan; the per-thread part is returned (not block-summed).  gg == nullptr:
// the whole fit in one workgroup.  On a failed gang exchange gg->ok turns false.
template <typename R>
__device__ __forceinline__ double coxLogDen(const Slot<R>& s, int pLo, int pHi, Gang* gg) {
  const GrpSt* grp = s.grp;
  const uint8_t* reset = s.reset;
  const int lane = threadIdx.x & (kWave - 1);
  constexpr int kStep = kWave * kScanItems;
  int a, z;
  waveRange(pLo, pHi, &a, &z);
  Seg<1> wagg = segZero<1>();
  for (int q0 = a; q0 < z; q0 += kStep) {
    const int i0 = q0 + lane * kScanItems;
    Seg<1> l = segZero<1>();
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
      const int i = i0 + u;
      if (i < z) {
        Seg<1> x;
        x.f = reset[i];
        x.v[0] = grp[i].den;
        l = segCombine(l, x);
      }
    }
    wagg = segCombine(wagg, segLastLane(waveScanIncl(l)));
  }
  Seg<1> total;
  Seg<1> run = waveCarries<1>(wagg, total);
  if (gg != nullptr && gg->G > 1) {
    Seg<1> carry;
    if (!gangCarry<1>(*gg, total, carry)) return 0.0;
    run = segCombine(carry, run);
  }
  double acc = 0.0;
  for (int q0 = a; q0 < z; q0 += kStep) {
    const int i0 = q0 + lane * kScanItems;
    Seg<1> it[kScanItems];
    double wn[kScanItems];
    Seg<1> l = segZero<1>();
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
      const int i = i0 + u;
      it[u] = segZero<1>();
      wn[u] = 0.0;
      if (i < z) {
        const GrpSt gs = grp[i];
        it[u].f = reset[i];
        it[u].v[0] = gs.den;
        wn[u] = gs.nw;
      }
      l = segCombine(l, it[u]);
    }
    const Seg<1> inc = waveScanIncl(l);
    Seg<1> r = segPrevLane(inc);
    if (lane == 0) r = segZero<1>();
    r = segCombine(run, r);
#pragma unroll
    for (int u = 0; u < kScanItems; ++u) {
      if (i0 + u >= z) break;
      r = segCombine(r, it[u]);
      if (wn[u] != 0.0) acc += wn[u] * log(r.v[0]);
    }
    run = segCombine(run, segLastLane(inc));
  }
  return acc;
}

// One coordinate cycle of a Cox fit over this rank's part, from coordinate j0:
// numerators of its rows, the risk-set scan of its pids (carry across the
// gang), one exchange of (gradient, Hessian), the prior step, the update of its
// rows and pids.  Rank r owns the rows of its shards and the pids of those rows
// (shards start at tie groups, so no pid spans two ranks).  Returns as cycleFit.
//
// Solo runs (gangs over at most kCoxSoloPids pids, as risk-set pieces of tied
// event times give): consecutive short columns are run by the lead rank alone --
// numerators of the whole column, the scan of every pid, the step and the update
// of every row the column touches -- while the other ranks skip ahead, so such
// a column costs no exchange and no carry.  Entering and leaving a run, and every
// kSoloSpan coordinates inside one, the ranks meet (gangSoloSync), which orders
// every rank's writes before the meeting before every read after it and keeps
// the waiting ranks' bounded waits short.  The sums are the same sums in the
// same order as on one workgroup, so a column's step does not depend on G.
constexpr int kCoxSoloPids = 16384;
constexpr int kSoloSpan = 2048;

template <typename R, int ITL, class LDS>
__device__ __forceinline__ int cycleCox(const Slot<R>& s, FitStatus& cs, LDS& S, int j0,
                                        long long deadline, Gang& gg) {
  static_assert(ITL >= kIT, "a Cox pass stages blockDim * kIT nonzeros in LDS");
  __shared__ int sKey[LDS::kItems];
  const int useW = s.ctl.useW;
  const bool multi = gg.G > 1;
  const bool lead = gg.rank == 0;
  const int pLo = multi ? s.pidShard[gg.s0] : 0;
  const int pHi = multi ? s.pidShard[gg.s1] : s.N;
  Gang one = gg;  // the lead alone: whole columns, every pid, no exchange
  one.G = 1;
  const int64_t soloMax =
      (multi && s.N <= kCoxSoloPids && s.ctl.solo >= 0) ? (int64_t)blockDim.x * kIT : -1;
  long long cG = 0, cU = 0, cP = 0;
  long long tN = 0, tS = 0, tU = 0;  // device ticks: numerators, scan, exchange + update
  int jEnd = s.J;
  __shared__ double xl[kXchLead];
  __shared__ ColMeta M;
  loadColMeta<M_COX, R>(M, s, j0, gg);
  bool inSolo = false, soloStop = false;
  auto soloSync = [&](int jStop) -> bool {
    if (threadIdx.x == 0 && lead) {
      xl[0] = soloStop ? 1.0 : 0.0;
      xl[1] = (double)jStop;
    }
    return gangSoloSync(gg, xl);
  };
  for (int j = j0; j < s.J; ++j) {
    if (j + 1 > M.w0 + kMetaW) loadColMeta<M_COX, R>(M, s, j, gg);  // uniform
    const int mj = j - M.w0;
    const int64_t cb = M.cb[mj], ce = M.cb[mj + 1];
    // every fourth coordinate, and before every column of more than one pass
    if ((!multi || (inSolo && lead)) && j > j0 &&
        ((j & 3) == 0 || ce - cb > (int64_t)blockDim.x * kIT) && quantumOver(deadline)) {
      soloStop = multi;
      jEnd = j;
      break;
    }
    if (M.fixed[mj]) continue;
    if (multi) {
      const bool soloCol = ce - cb <= soloMax;
      // every rank sees the same columns, so every rank meets at the same j
      if (soloCol != inSolo || (inSolo && (j % kSoloSpan) == 0)) {
        if (!soloSync(j)) return -1;
        const bool stopped = inSolo && xl[0] != 0.0;
        inSolo = soloCol;
        if (stopped) {  // the lead's quantum ran out inside the run just left
          inSolo = false;
          jEnd = (int)xl[1];
          break;
        }
      }
      if (inSolo && !lead) continue;
    }
    const bool part = multi && !inSolo;
    int64_t b = cb, e = ce;
    if (part) {
      b = cb + M.b[mj];
      e = cb + M.e[mj];
    }
    const int qLo = part ? pLo : 0, qHi = part ? pHi : s.N;
    Gang& gx = part ? gg : one;
    const int64_t njAll = ce - cb;
    const int fmt = M.fmt[mj];
    // in a gang, only the lead rank's reads count (it publishes them)
    double bj = s.beta[j];
    double bnd = s.bound[j];
    const double pv = s.pvar[j], xy = s.xjy[j];
    const int pt = s.ptype[j];
    double v[2] = {0.0, 0.0};
    bool stop = false;
    const long long t0 = wall_clock64();
    long long t1 = t0, t2 = t0;
    if (njAll > 0) {
      coxNumeratorsLds<R>(s, S, sKey, fmt, cb, b, e, useW);
      __syncthreads();
      t1 = wall_clock64();
      if (!coxScan<R>(s, qLo, qHi, isIndicator(fmt), gx, v)) return -1;
      t2 = wall_clock64();
      cP += qHi - qLo;
    }
    blockSum<2>(v);
    if (part) {
      double ld[3] = {bj, bnd, (deadline > 0 && wall_clock64() >= deadline) ? 1.0 : 0.0};
      if (!gangExchange<2, 3>(gg, v, ld)) return -1;
      bj = ld[0];
      bnd = ld[1];
      stop = ld[2] != 0.0;
    }
    double g = njAll > 0 ? v[0] - xy : 0.0, h = njAll > 0 ? v[1] : 0.0;
    if (h < 0.0) {
      g = 0.0;
      h = 0.0;
    }
    const double delta = applyBounds(priorDelta(pt, pv, g, h, bj), &bnd);
    if (threadIdx.x == 0 && lead) {
      s.bound[j] = bnd;
      if (delta != 0.0) s.beta[j] = bj + delta;
    }
    cG += e - b;
    if (delta != 0.0) {
      updCoxLds<R>(s, S, sKey, fmt, cb, b, e, useW, delta);
      cU += e - b;
    }
    __syncthreads();
    tN += t1 - t0;
    tS += t2 - t1;
    tU += wall_clock64() - t2;
    if (stop) {
      jEnd = j + 1;
      break;
    }
  }
  if (multi && inSolo) {  // every rank leaves the last run together
    if (!soloSync(jEnd)) return -1;
    if (xl[0] != 0.0) jEnd = (int)xl[1];
  }
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long*)&cs.gatherNnz, (unsigned long long)cG);
    atomicAdd((unsigned long long*)&cs.updateNnz, (unsigned long long)cU);
    atomicAdd((unsigned long long*)&cs.scanPids, (unsigned long long)cP);
    if (lead) {
      atomicAdd((unsigned long long*)&cs.tSmall, (unsigned long long)tN);
      atomicAdd((unsigned long long*)&cs.tLarge, (unsigned long long)tS);
      atomicAdd((unsigned long long*)&cs.nLarge, (unsigned long long)tU);
    }
  }
  return jEnd;
}

// ------------------------------------------------------------ whole-fit pieces
// Rows [rowLo, rowHi) and groups [grpLo, grpHi) a workgroup sums over.  In a
// bundle every loop walks (row, fit) or (group, fit) pairs, so lane u serves fit
// u & (F-1) and the sums come back per fit (blockSumFits).
struct Part {
  int rowLo, rowHi, grpLo, grpHi;
};

template <typename R>
__device__ __forceinline__ Part wholeFit(const Slot<R>& s) {
  return Part{0, s.K, 0, s.N};
}

// log likelihood terms of this lane's fit over a part (ModelSpecifics.hpp:605-650);
// Cox (one fit per slot): the accumulated denominators of the part's pids
// [grpLo, grpHi) come from a segmented scan carried across the gang `gg`
// (nullptr: one workgroup)
template <int MODEL, typename R>
__device__ __forceinline__ double fitLogLik(const Slot<R>& s, const Part& pt, Gang* gg = nullptr) {
  const DevX& d = s.d;
  const int useW = s.ctl.useW;
  const int F = s.F, lF = s.lF;
  double v[2] = {0.0, 0.0};
  for (int64_t u = ((int64_t)pt.rowLo << lF) + threadIdx.x; u < ((int64_t)pt.rowHi << lF);
       u += blockDim.x) {
    const int k = (int)(u >> lF);
    const RowSt<R> rs = s.row[u];
    const double x = (double)rs.x;
    const int yi = (int)d.y[k];
    const double w = useW ? s.kw[u] : 1.0;
    double c = (MODEL == M_COX) ? (yi != 1 ? 0.0 : x) : yi * x;
    v[0] += useW ? c * w : c;
    if (MODEL == M_LOGISTIC) v[1] += w * log((double)s.rden[u]);
    if (MODEL == M_POISSON) v[1] += w * (double)s.rden[u];
  }
  if (MODEL == M_CLR) {
    for (int64_t u = ((int64_t)pt.grpLo << lF) + threadIdx.x; u < ((int64_t)pt.grpHi << lF);
         u += blockDim.x) {
      const GrpSt gs = s.grp[u];
      if (gs.nw != 0.0) v[1] += gs.nw * log(gs.den);
    }
  }
  if (MODEL == M_COX) v[1] += coxLogDen<R>(s, pt.grpLo, pt.grpHi, gg);
  blockSumFits<2>(v, F);
  return v[0] - v[1];
}

template <typename R>
__device__ __forceinline__ double fitGradObjective(const Slot<R>& s, const Part& pt) {
  const int useW = s.ctl.useW;
  const int lF = s.lF;
  double v[1] = {0.0};
  for (int64_t u = ((int64_t)pt.rowLo << lF) + threadIdx.x; u < ((int64_t)pt.rowHi << lF);
       u += blockDim.x) {
    const int k = (int)(u >> lF);
    const double x = (double)s.row[u].x;
    v[0] += useW ? x * s.d.y[k] * s.kw[u] : x * s.d.y[k];
  }
  blockSumFits<1>(v, s.F);
  return v[0];
}

template <typename R>
__device__ __forceinline__ double priorPart(const Slot<R>& s) {
  double v[1] = {0.0};
  for (int64_t u = threadIdx.x; u < ((int64_t)s.J << s.lF); u += blockDim.x)
    v[0] += priorLogDensity(s.ptype[u], s.pvar[u], s.beta[u]);
  blockSumFits<1>(v, s.F);
  return v[0];
}

// Convergence objective (getObjectiveFunction, CyclicCoordinateDescent.cpp:541-558)
// of this lane's fit (single workgroup).
template <int MODEL, typename R>
__device__ __forceinline__ double fitObjective(const Slot<R>& s) {
  const int ct = s.ctl.convergenceType;
  const Part all = wholeFit(s);
  if (ct == 0) return fitGradObjective<R>(s, all);
  if (ct == 3) return 0.0;
  const double ll = fitLogLik<MODEL, R>(s, all) + s.fixedLL[threadIdx.x & (s.F - 1)];
  if (ct == 2) return ll;
  return ll + priorPart<R>(s);
}

// The same objective summed over a gang; `stop` carries the lead rank's
// end-of-quantum decision.  Returns false when the exchange failed.
template <int MODEL, typename R>
__device__ __forceinline__ bool gangObjective(const Slot<R>& s, Gang& gg, long long deadline,
                                              double* obj, bool* stop) {
  const int ct = s.ctl.convergenceType;
  const int F = s.F;
  const int f = threadIdx.x & (F - 1);
  __shared__ double xs[kXchSum], xl[kXchLead];
  Part pt{gg.rowLo, gg.rowHi, 0, 0};
  if (MODEL == M_CLR && gg.rowHi > gg.rowLo) {
    pt.grpLo = s.d.strata[gg.rowLo];
    pt.grpHi = s.d.strata[gg.rowHi - 1] + 1;
  }
  if (MODEL == M_COX) {
    pt.grpLo = s.pidShard[gg.s0];
    pt.grpHi = s.pidShard[gg.s1];
  }
  double v = 0.0;
  if (ct == 0) v = fitGradObjective<R>(s, pt);
  if (ct == 1 || ct == 2) {
    v = fitLogLik<MODEL, R>(s, pt, &gg);
    if (!gg.ok) return false;
  }
  // the coefficients belong to the lead rank: it adds the log prior
  const double lp = (ct == 1 && gg.rank == 0) ? priorPart<R>(s) : 0.0;
  if ((int)threadIdx.x < F) {
    xs[threadIdx.x] = v;
    xl[threadIdx.x] = lp;
  }
  if (threadIdx.x == 0) {  // with this rank's part of the next cycle's first column
    int64_t n0 = 0;
    if (MODEL != M_COX && s.J > 0 && !s.fixed[0])
      n0 = (int64_t)s.d.colShard[gg.s1] - s.d.colShard[gg.s0];
    const int64_t onePass = ((int64_t)blockDim.x * kIT) >> s.lF;
    xl[F] = (deadline > 0 && wall_clock64() + aheadTicks(n0, onePass) >= deadline) ? 1.0 : 0.0;
  }
  if (!gangExchangeN(gg, xs, F, xl, F + 1)) return false;
  *stop = xl[F] != 0.0;
  const double fl = s.fixedLL[f];
  if (ct == 0) *obj = xs[f];
  else if (ct == 3) *obj = 0.0;
  else if (ct == 2) *obj = xs[f] + fl;
  else *obj = xs[f] + fl + xl[f];
  return true;
}

// Per-fit set-up (findMode's lazy part), one rank of G over the row shards
// [r*S/G, (r+1)*S/G), for the fits of the slot in `mask`: event weights of the
// rank's groups, XjY of every G-th coordinate, xBeta of its rows from the
// starting coefficients (column order, as computeXBeta), exp and the
// denominators of its groups, bounds of every G-th coordinate.  Groups and rows
// of a rank are its own (shards start at strata / tie groups), so ranks need no
// exchange and the result does not depend on G.
template <int MODEL, typename R>
__device__ __forceinline__ void initPart(const Slot<R>& s, bool full, int rank, int G,
                                         unsigned mask) {
  const DevX& d = s.d;
  const int useW = s.ctl.useW;
  const int F = s.F, lF = s.lF;
  const int f = threadIdx.x & (F - 1);
  const bool mine = (mask >> f) & 1u;
  RowSt<R>* rowSt = s.row;
  const double* kw = s.kw;
  const int s0 = rank * d.nShards / G, s1 = (rank + 1) * d.nShards / G;
  const int rowLo = d.shardRow[s0], rowHi = d.shardRow[s1];
  // the rank's groups tile [0, groups): from its first row's group to the next rank's
  int gLo = 0, gHi = 0;
  if (MODEL == M_CLR) {
    auto first = [&](int sh) { return sh == 0 ? 0 : (d.shardRow[sh] < s.K ? d.strata[d.shardRow[sh]] : d.nStrata); };
    gLo = first(s0);
    gHi = s1 == d.nShards ? d.nStrata : first(s1);
  } else if (MODEL == M_COX) {
    gLo = s.pidShard[s0];
    gHi = s.pidShard[s1];
  }
  // nW and XjY depend on the weights only: a restart keeps them
  if (full && MODEL == M_CLR) {
    for (int64_t u = ((int64_t)gLo << lF) + threadIdx.x; u < ((int64_t)gHi << lF); u += blockDim.x) {
      if (!mine) continue;
      const int g = (int)(u >> lF);
      double acc = 0.0;
      for (int k = d.stratStart[g]; k < d.stratStart[g + 1]; ++k) acc += d.y[k] * kw[(size_t)k * F + f];
      s.grp[u].nw = acc;
    }
  } else if (full && MODEL == M_COX) {
    for (int p = gLo + threadIdx.x; p < gHi; p += blockDim.x) {
      double acc = 0.0;
      for (int k = s.pidFirst[p]; k < s.pidFirst[p + 1]; ++k)
        if (s.pid[k] == p) acc += d.y[k] * kw[k];
      s.grp[p].nw = acc;
    }
  }
  for (int j = rank; full && j < s.J; j += G) {
    const int fmt = d.fmt[j];
    const int64_t b = d.colPtr[j], e = d.colPtr[j + 1];
    double t[1] = {0.0};
    for (int64_t u = threadIdx.x; u < ((e - b) << lF); u += blockDim.x) {
      const int64_t p = b + (u >> lF);
      const int k = colRow(d, fmt, p, b);
      const double x = colValue<double>(d, fmt, p);
      t[0] += useW ? x * d.y[k] * kw[(size_t)k * F + f] : x * d.y[k];
    }
    blockSumFits<1>(t, F);
    if ((int)threadIdx.x < F && mine) s.xjy[(size_t)j * F + f] = t[0];
  }
  // xBeta = sum_j beta_j X_j in column order, over the rank's part of each column
  for (int64_t u = ((int64_t)rowLo << lF) + threadIdx.x; u < ((int64_t)rowHi << lF); u += blockDim.x)
    if (mine) rowSt[u].x = (R)0;
  __syncthreads();
  for (int j = 0; j < s.J; ++j) {
    bool any = false;
    for (int i = 0; i < F; ++i) any |= ((mask >> i) & 1u) && s.beta[(size_t)j * F + i] != 0.0;
    if (!any) continue;
    const double bj = s.beta[(size_t)j * F + f];
    const int fmt = d.fmt[j];
    const int64_t cb = d.colPtr[j];
    const int32_t* off = d.colShard + (size_t)j * (d.nShards + 1);
    const int64_t b = cb + off[s0], e = cb + off[s1];
    const R rb = (R)bj;
    if (mine && bj != 0.0) {
      for (int64_t u = threadIdx.x; u < ((e - b) << lF); u += blockDim.x) {
        const int64_t p = b + (u >> lF);
        const int k = colRow(d, fmt, p, cb);
        rowSt[(size_t)k * F + f].x += rb * colValue<R>(d, fmt, p);
      }
    }
    __syncthreads();
  }
  // exp and denominators
  if (MODEL == M_CLR) {
    for (int64_t u = ((int64_t)gLo << lF) + threadIdx.x; u < ((int64_t)gHi << lF); u += blockDim.x) {
      if (!mine) continue;
      const int g = (int)(u >> lF);
      double acc = 0.0;
      for (int k = d.stratStart[g]; k < d.stratStart[g + 1]; ++k) {
        const size_t kf = (size_t)k * F + f;
        const R o = rexp(rowSt[kf].x);
        rowSt[kf].e = o;
        acc += (double)o;
      }
      s.grp[u].den = acc;
    }
  } else {
    for (int64_t u = ((int64_t)rowLo << lF) + threadIdx.x; u < ((int64_t)rowHi << lF); u += blockDim.x) {
      if (!mine) continue;
      const R o = rexp(rowSt[u].x);
      rowSt[u].e = o;
      // the model's null value plus exp(xBeta) (ModelSpecifics.hpp:1588-1612)
      if (MODEL != M_COX) s.rden[u] = (MODEL == M_LOGISTIC ? (R)1 : (R)0) + o;
    }
    if (MODEL == M_COX) {
      __syncthreads();
      for (int p = gLo + threadIdx.x; p < gHi; p += blockDim.x) {
        double acc = 0.0;
        for (int k = s.pidFirst[p]; k < s.pidFirst[p + 1]; ++k) {
          if (s.pid[k] != p) continue;
          acc += useW ? kw[k] * (double)rowSt[k].e : (double)rowSt[k].e;
        }
        s.grp[p].den = acc;
      }
    }
  }
  for (int64_t u = (int64_t)rank * blockDim.x + threadIdx.x; u < ((int64_t)s.J << lF);
       u += (int64_t)G * blockDim.x)
    if (mine) s.bound[u] = s.initBound[f];
}

// The starting objective and status of the set-up fits in `mask` (one workgroup).
template <int MODEL, typename R>
__device__ __forceinline__ void initObjective(const Slot<R>& s, FitStatus* st, unsigned mask) {
  const double obj = fitObjective<MODEL, R>(s);
  const int f = threadIdx.x;
  if (f < s.F && ((mask >> f) & 1u)) {
    st[f].lastObj = obj;
    st[f].iteration = 0;
    st[f].done = 0;
    st[f].flag = CCDH_SUCCESS;
  }
  if (threadIdx.x == 0) st[0].jNext = 0;
}

// Up to `sweeps` complete CCD cycles of the slot's running fits plus the
// convergence check of each after every cycle; with a deadline, the slot stops
// at a coordinate boundary once it has passed and the next launch resumes the
// sweep there.  The fits of a slot move in lock step; one that converges stops
// loading and updating while the others go on.  Every rank of a gang computes
// the same objectives and decisions; the lead rank records them (thread f, fit f).
template <int MODEL, typename R, int ITL, class LDS>
__device__ __forceinline__ void sweepSlot(const Slot<R>& s, FitStatus* st, LDS& S, int sweeps,
                                          long long deadline, Gang& gg) {
  const FitCtl c = s.ctl;
  const int F = s.F;
  const int f = threadIdx.x & (F - 1);
  const bool multi = gg.G > 1;
  const bool rec = (int)threadIdx.x < F && gg.rank == 0;
  __shared__ unsigned sAct;
  unsigned act = 0;
  for (int i = 0; i < F; ++i) act |= (st[i].done ? 0u : 1u) << i;
  int j0 = st[0].jNext;
  int iteration = st[f].iteration;
  double lastObj = st[f].lastObj;
  for (int sw = 0; sw < sweeps; ++sw) {
    const int jEnd = MODEL == M_COX ? cycleCox<R, ITL>(s, st[0], S, j0, deadline, gg)
                                    : cycleFit<MODEL, R, ITL>(s, st, S, j0, deadline, gg, act);
    __syncthreads();
    if (jEnd < 0) {
      if (threadIdx.x == 0) st[0].fault = 1;
      return;
    }
    if (jEnd < s.J) {
      if (threadIdx.x == 0 && gg.rank == 0) st[0].jNext = jEnd;
      return;
    }
    j0 = 0;
    // performCheckConvergence
    const long long tq0 = wall_clock64();
    double obj;
    bool stop = false;
    if (multi) {
      if (!gangObjective<MODEL, R>(s, gg, deadline, &obj, &stop)) {
        if (threadIdx.x == 0) st[0].fault = 1;
        return;
      }
    } else {
      obj = fitObjective<MODEL, R>(s);
    }
    const bool on = (act >> f) & 1u;
    int done = 0, flag = CCDH_SUCCESS;
    if (on) {
      iteration += 1;
      double conv;
      bool ill = false;
      if (obj != obj) {
        conv = 0.0;
        ill = true;
      } else {
        conv = fabs(obj - lastObj) / (fabs(obj) + 1.0);
      }
      lastObj = obj;
      if (c.tolerance > 0 && conv < c.tolerance) {
        flag = ill ? CCDH_ILLCONDITIONED : CCDH_SUCCESS;
        done = 1;
      } else if (iteration == c.maxIterations) {
        flag = CCDH_MAX_ITERATIONS;
        done = 1;
      }
    }
    if (rec && on) {
      if (f == 0) st[0].tObj += wall_clock64() - tq0;
      st[f].objRows += (c.convergenceType == 0 ? s.K : (c.convergenceType == 3 ? 0 : s.K + s.N));
      st[f].iteration = iteration;
      st[f].lastObj = obj;
      if (done) {
        st[f].flag = flag;
        st[f].done = 1;
      }
    }
    if (threadIdx.x == 0 && gg.rank == 0) st[0].jNext = 0;
    // fits still running (lane f < F speaks for fit f)
    if (threadIdx.x == 0) sAct = act;
    __syncthreads();
    if ((int)threadIdx.x < F && done) atomicAnd(&sAct, ~(1u << threadIdx.x));
    __syncthreads();
    act = sAct;
    __syncthreads();
    if (act == 0) return;
    // the next cycle opens with column 0
    const int64_t n0 =
        (MODEL != M_COX && s.J > 0 && !s.fixed[0]) ? s.d.colPtr[1] - s.d.colPtr[0] : 0;
    if (multi ? stop : quantumOver(deadline, n0, ((int64_t)blockDim.x * kIT) >> s.lF)) return;
  }
}

// Final log likelihood and, when held-out weights are given, the predictive
// log likelihood (ModelSpecifics.hpp:652-690; LR / Poisson / CLR forms) of
// every fit of the slot.
template <int MODEL, typename R>
__device__ __forceinline__ void finishSlot(const Slot<R>& s, FitStatus* st) {
  const int F = s.F, lF = s.lF;
  const int f = threadIdx.x & (F - 1);
  const double ll = fitLogLik<MODEL, R>(s, wholeFit(s));
  double v[1] = {0.0};
  if (s.heldout && MODEL != M_COX) {
    const double* w2 = s.heldout;
    for (int64_t u = threadIdx.x; u < ((int64_t)s.K << lF); u += blockDim.x) {
      const int k = (int)(u >> lF);
      const RowSt<R> rs = s.row[u];
      const double x = (double)rs.x;
      const double y = s.d.y[k];
      if (MODEL == M_LOGISTIC) v[0] += y * w2[u] * (x - log((double)s.rden[u]));
      if (MODEL == M_POISSON) v[0] += (y * x - exp(x)) * w2[u];
      if (MODEL == M_CLR) v[0] += y * w2[u] * (x - log(s.grp[(size_t)s.d.strata[k] * F + f].den));
    }
  }
  blockSumFits<1>(v, F);
  if ((int)threadIdx.x < F) {
    st[f].logLik = ll + (MODEL == M_POISSON ? s.fixedLL[f] : 0.0);
    st[f].predLogLik = v[0];
  }
}

// ------------------------------------------------------------------ kernels
// Each kernel runs one workgroup (or a gang, or set-up ranks) per listed slot
// and dispatches on the slot's model (uniform per workgroup).  A slot's fits
// have their statuses at st[s.fit0 .. s.fit0 + s.F).

// `ranks` workgroups per listed slot, each setting up its shards, for the fits in `mask`.
template <typename R>
__global__ void k_slot_init(const Slot<R>* slots, const int32_t* active, int full, int ranks,
                            unsigned mask) {
  const int f = active[blockIdx.x / ranks];
  const int rank = blockIdx.x % ranks;
  const Slot<R>& s = slots[f];
  switch (s.model) {
    case M_LOGISTIC: initPart<M_LOGISTIC, R>(s, full != 0, rank, ranks, mask); break;
    case M_POISSON: initPart<M_POISSON, R>(s, full != 0, rank, ranks, mask); break;
    case M_CLR: initPart<M_CLR, R>(s, full != 0, rank, ranks, mask); break;
    case M_COX: initPart<M_COX, R>(s, full != 0, rank, ranks, mask); break;
  }
}

template <typename R>
__global__ void k_slot_objective(const Slot<R>* slots, FitStatus* st, const int32_t* active,
                                 unsigned mask) {
  const Slot<R>& s = slots[active[blockIdx.x]];
  FitStatus* fs = st + s.fit0;
  switch (s.model) {
    case M_LOGISTIC: initObjective<M_LOGISTIC, R>(s, fs, mask); break;
    case M_POISSON: initObjective<M_POISSON, R>(s, fs, mask); break;
    case M_CLR: initObjective<M_CLR, R>(s, fs, mask); break;
    case M_COX: initObjective<M_COX, R>(s, fs, mask); break;
  }
}

// ONLY: -1 dispatches on each slot's model (launches that mix models: the
// server); a model id builds that model's path alone (a batch over one data
// set), which needs fewer registers than all four together.
template <typename R, int BLOCK, int ITL, int ONLY>
__global__ void __launch_bounds__(BLOCK)
    k_slot_sweep(const Slot<R>* slots, FitStatus* st, const GangEntry* map, Xch* xch,
                 int sweeps, long long quantum) {
  __shared__ PassLds<BLOCK * ITL> S;
  __shared__ long long sStart;
  const GangEntry ge = map[blockIdx.x];
  if (ge.slot < 0) return;
  const Slot<R>& s = slots[ge.slot];
  FitStatus* fs = st + s.fit0;
  bool running = false;
  for (int i = 0; i < s.F; ++i) running |= fs[i].done == 0;
  if (!running) return;
  if (threadIdx.x == 0) {
    sStart = wall_clock64();
    longTicksPerNnz() = fs[0].longTpn;
  }
  __syncthreads();
  const long long deadline = quantum > 0 ? sStart + quantum : 0;
  Gang gg;
  gg.rank = ge.rank;
  gg.G = ge.G;
  gg.seq = 0;
  gg.spin = s.ctl.spinLimit > 0 ? s.ctl.spinLimit : kSpinLimit;
  gg.ok = true;
  gg.x = ge.G > 1 ? xch + ge.slot : nullptr;
  if (ge.G > 1) {
    // a gang shares one XCD (one L2): check before touching any state
    const double x = (double)xccId();
    double v[2] = {x, x * x};
    double ld[1] = {0.0};
    if (!gangExchange<2, 1>(gg, v, ld)) {
      if (threadIdx.x == 0) fs[0].fault = 1;
      return;
    }
    if (v[0] * v[0] != v[1] * ge.G) {
      if (threadIdx.x == 0) fs[0].fault = 2;
      return;
    }
  }
  if (ge.G > 1) {
    gg.s0 = ge.rank * s.d.nShards / ge.G;
    gg.s1 = (ge.rank + 1) * s.d.nShards / ge.G;
    gg.rowLo = s.d.shardRow[gg.s0];
    gg.rowHi = s.d.shardRow[gg.s1];
  } else {
    gg.s0 = 0;
    gg.s1 = s.d.nShards;
    gg.rowLo = 0;
    gg.rowHi = s.K;
  }
  if constexpr (ONLY >= 0) {
    sweepSlot<ONLY, R, ITL>(s, fs, S, sweeps, deadline, gg);
    if (threadIdx.x == 0 && gg.rank == 0) fs[0].longTpn = longTicksPerNnz();
  } else {
    switch (s.model) {
      case M_LOGISTIC: sweepSlot<M_LOGISTIC, R, ITL>(s, fs, S, sweeps, deadline, gg); break;
      case M_POISSON: sweepSlot<M_POISSON, R, ITL>(s, fs, S, sweeps, deadline, gg); break;
      case M_CLR: sweepSlot<M_CLR, R, ITL>(s, fs, S, sweeps, deadline, gg); break;
      case M_COX: sweepSlot<M_COX, R, ITL>(s, fs, S, sweeps, deadline, gg); break;
    }
    if (threadIdx.x == 0 && gg.rank == 0) fs[0].longTpn = longTicksPerNnz();
  }
}

// Risk-set groups of a Cox fit under its weights, built on the device
// (setPidForAccumulation, ModelSpecifics.hpp:1809-1868): walking the rows of
// non-zero weight in their (stratum, -time, y) order, a row starts a new group
// (pid) unless it is an event at the same time as the previous such row of the
// same stratum, and a new stratum also marks a reset; zero-weight rows get pid
// N.  One workgroup per listed slot, three passes over per-thread contiguous
// row blocks: the last weighted row of each block (an exclusive max-scan gives
// every thread the weighted row before its block), the groups each block
// starts (an exclusive sum-scan gives its first pid), then pid, first row and
// reset flag of every group and cleared numerators.  Also writes the slot's N
// and the first pid of each row shard (the Cox gang ranks' boundaries).
template <typename R>
__global__ void k_cox_groups(Slot<R>* slots, FitStatus* st, const int32_t* active) {
  Slot<R>& s = slots[active[blockIdx.x]];
  const DevX& d = s.d;
  const int K = s.K;
  const double* w = s.kw;
  __shared__ int sLast[kMaxBlockThreads];
  __shared__ int sCnt[kMaxBlockThreads];
  __shared__ int sN;
  int a, z;
  pidBlock(0, K, &a, &z);
  int last = -1;
  for (int k = a; k < z; ++k)
    if (w[k] != 0.0) last = k;
  sLast[threadIdx.x] = last;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = -1;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      const int v = sLast[t];
      sLast[t] = run;
      run = max(run, v);
    }
  }
  __syncthreads();
  const int carry = sLast[threadIdx.x];
  // the row shard holding row k (shards start at tie groups)
  auto shardOf = [&](int k) {
    int lo = 0, hi = d.nShards;  // shardRow[lo] <= k < shardRow[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (d.shardRow[mid] <= k) lo = mid;
      else hi = mid;
    }
    return lo;
  };
  const bool pieces = s.coxPieces != 0;
  auto starts = [&](int prev, int k) {
    const bool newPid = prev < 0 || d.strata[k] != d.strata[prev] ||
                        !(d.y[prev] == 1.0 && d.time[prev] == d.time[k] && d.y[k] == 1.0);
    if (!pieces || !newPid) return newPid;
    // a piece ends with an event group, at a stratum's end or at a shard's end
    return prev < 0 || d.strata[k] != d.strata[prev] || d.y[prev] == 1.0 ||
           shardOf(prev) != shardOf(k);
  };
  int cnt = 0;
  int prev = carry;
  for (int k = a; k < z; ++k) {
    if (w[k] == 0.0) continue;
    cnt += starts(prev, k) ? 1 : 0;
    prev = k;
  }
  sCnt[threadIdx.x] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      const int v = sCnt[t];
      sCnt[t] = run;
      run += v;
    }
    sN = run;
  }
  __syncthreads();
  const int N = sN;
  int p = sCnt[threadIdx.x] - 1;
  prev = carry;
  for (int k = a; k < z; ++k) {
    if (w[k] == 0.0) {
      s.pid[k] = N;
      continue;
    }
    if (starts(prev, k)) {
      ++p;
      s.pidFirst[p] = k;
      s.reset[p] = (prev >= 0 && d.strata[k] != d.strata[prev]) ? 1 : 0;
      s.numer[p] = 0.0;
      s.numer2[p] = 0.0;
    }
    s.pid[k] = p;
    prev = k;
  }
  if (threadIdx.x == 0) {
    s.pidFirst[N] = K;
    s.reset[N] = 0;
    s.numer[N] = 0.0;
    s.numer2[N] = 0.0;
    s.N = N;
    st[s.fit0].N = N;
  }
  __syncthreads();
  // the first pid whose first row is at or after each shard's first row
  for (int sh = threadIdx.x; sh <= d.nShards; sh += blockDim.x) {
    int v = N;
    if (sh == 0) {
      v = 0;
    } else if (sh < d.nShards) {
      const int row = d.shardRow[sh];
      int lo = 0, hi = N;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (s.pidFirst[mid] < row) lo = mid + 1;
        else hi = mid;
      }
      v = lo;
    }
    s.pidShard[sh] = v;
  }
}

// Zero the exchange counters of the gangs of the next launch.
__global__ void k_gang_reset(const GangEntry* map, int n, Xch* xch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && map[i].slot >= 0 && map[i].rank == 0 && map[i].G > 1) xch[map[i].slot].cnt = 0;
}

template <typename R>
__global__ void k_slot_finish(const Slot<R>* slots, FitStatus* st, const int32_t* active) {
  const Slot<R>& s = slots[active[blockIdx.x]];
  FitStatus* fs = st + s.fit0;
  switch (s.model) {
    case M_LOGISTIC: finishSlot<M_LOGISTIC, R>(s, fs); break;
    case M_POISSON: finishSlot<M_POISSON, R>(s, fs); break;
    case M_CLR: finishSlot<M_CLR, R>(s, fs); break;
    case M_COX: finishSlot<M_COX, R>(s, fs); break;
  }
}

__global__ void k_fill_d(double* p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int clampThreads(int threads) { return threads <= 256 ? 256 : 512; }

}  // namespace

// ================================================================ slot host side
namespace ccdamd {

int slotLaunchInit(int /*model*/, bool mixed, const void* slots, FitStatus* st,
                   const int32_t* active, int n, int threads, hipStream_t s, bool full, int ranks,
                   unsigned mask) {
  if (n <= 0) return 0;
  threads = clampThreads(threads);
  ranks = std::max(1, std::min(ranks, kShards));
  if (mixed) {
    k_slot_init<float><<<n * ranks, threads, 0, s>>>((const Slot<float>*)slots, active,
                                                     full ? 1 : 0, ranks, mask);
    k_slot_objective<float><<<n, threads, 0, s>>>((const Slot<float>*)slots, st, active, mask);
  } else {
    k_slot_init<double><<<n * ranks, threads, 0, s>>>((const Slot<double>*)slots, active,
                                                      full ? 1 : 0, ranks, mask);
    k_slot_objective<double><<<n, threads, 0, s>>>((const Slot<double>*)slots, st, active, mask);
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

int slotLaunchCoxGroups(bool mixed, void* slots, FitStatus* st, const int32_t* active, int n,
                        int threads, hipStream_t s) {
  if (n <= 0) return 0;
  threads = clampThreads(threads);
  if (mixed)
    k_cox_groups<float><<<n, threads, 0, s>>>((Slot<float>*)slots, st, active);
  else
    k_cox_groups<double><<<n, threads, 0, s>>>((Slot<double>*)slots, st, active);
  HIP_TRY(hipGetLastError());
  return 0;
}

int initRanks(int64_t nnz, int K) {
  int r = 1;
  while (r * 2 <= kShards && nnz / (r * 2) >= (1 << 17) && K / (r * 2) >= (1 << 12)) r *= 2;
  return r;
}

// Sweep kernel variants: (256 lanes, pass width 8 or 4), (512 lanes, width 8).
template <typename R, int ONLY>
static const void* sweepKernelOf(int threads, int itl) {
  const bool t512 = clampThreads(threads) == 512;
  if constexpr (ONLY == M_COX) {
    return t512 ? (const void*)k_slot_sweep<R, 512, 4, M_COX> : (const void*)k_slot_sweep<R, 256, 4, M_COX>;
  } else if constexpr (ONLY >= 0) {
    if (t512 && itl == 4) return (const void*)k_slot_sweep<R, 512, 4, ONLY>;
    if (!t512 && itl == 8) return (const void*)k_slot_sweep<R, 256, 8, ONLY>;
    return nullptr;
  } else {
    if (t512) return itl == 4 ? (const void*)k_slot_sweep<R, 512, 4, -1> : (const void*)k_slot_sweep<R, 512, 8, -1>;
    return itl == 4 ? (const void*)k_slot_sweep<R, 256, 4, -1> : (const void*)k_slot_sweep<R, 256, 8, -1>;
  }
}

// The sweep kernel for a launch whose slots are all of model `only` (-1: any
// mix of models); a shape without a one-model build takes the general kernel.
template <typename R>
static const void* sweepKernelR(int threads, int itl, int only) {
  const void* fn = nullptr;
  switch (only) {
    case M_LOGISTIC: fn = sweepKernelOf<R, M_LOGISTIC>(threads, itl); break;
    case M_POISSON: fn = sweepKernelOf<R, M_POISSON>(threads, itl); break;
    case M_CLR: fn = sweepKernelOf<R, M_CLR>(threads, itl); break;
    case M_COX: fn = sweepKernelOf<R, M_COX>(threads, itl); break;
    default: break;
  }
  return fn ? fn : sweepKernelOf<R, -1>(threads, itl);
}

static const void* sweepKernel(bool mixed, int threads, int itl, int only = -1) {
  return mixed ? sweepKernelR<float>(threads, itl, only) : sweepKernelR<double>(threads, itl, only);
}

int slotLaunchSweeps(bool mixed, const void* slots, FitStatus* st, const GangEntry* map, int n,
                     Xch* xch, bool gangs, int threads, int itl, int sweeps,
                     long long quantumTicks, hipStream_t s, int only) {
  if (n <= 0) return 0;
  if (gangs) {
    k_gang_reset<<<(n + 255) / 256, 256, 0, s>>>(map, n, xch);
    HIP_TRY(hipGetLastError());
  }
  long long q = quantumTicks;
  const void* fn = sweepKernel(mixed, threads, itl, only);
  const int block = clampThreads(threads);
  void* args[] = {(void*)&slots, (void*)&st, (void*)&map, (void*)&xch, (void*)&sweeps, (void*)&q};
  if (gangs) {
    // every workgroup of a gang must be resident while its peers wait for it
    HIP_TRY(hipLaunchCooperativeKernel(fn, dim3(n), dim3(block), args, 0, s));
  } else {
    HIP_TRY(hipLaunchKernel(fn, dim3(n), dim3(block), args, 0, s));
  }
  return 0;
}


static int probeXcds() {
  const int n = 512;
  int* d = nullptr;
  if (hipMalloc((void**)&d, n * sizeof(int)) != hipSuccess) return 0;
  std::vector<int> h(n, -1);
  hipStream_t s = nullptr;
  int nx = 0;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess) {
    k_xcc_probe<<<n, 64, 0, s>>>(d);
    if (hipMemcpyAsync(h.data(), d, n * sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipStreamSynchronize(s) == hipSuccess) {
      for (int v : h) nx = std::max(nx, v + 1);
      bool roundRobin = nx > 0;
      for (int i = 0; i < n && roundRobin; ++i) roundRobin = h[i] == i % nx;
      if (!roundRobin) nx = 0;
    }
    (void)hipStreamDestroy(s);
  }
  (void)hipFree(d);
  return nx;
}

int xcdCount() {
  static const int n = probeXcds();  // once per process (thread-safe initialisation)
  return n;
}

bool layoutGangs(const std::vector<GangEntry>& gangs, int capacity, int nX,
                 std::vector<GangEntry>& map) {
  map.clear();
  if (nX <= 0) return false;
  const int perBin = capacity / nX;
  std::vector<int> order(gangs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return gangs[a].G > gangs[b].G; });
  std::vector<int> fill(nX, 0);
  std::vector<std::vector<GangEntry>> bins(nX);
  for (int i : order) {
    const GangEntry& g = gangs[i];
    int best = -1;
    for (int b = 0; b < nX; ++b)
      if (fill[b] + g.G <= perBin && (best < 0 || fill[b] < fill[best])) best = b;
    if (best < 0) return false;
    for (int r = 0; r < g.G; ++r) bins[best].push_back(GangEntry{g.slot, r, g.G, 0});
    fill[best] += g.G;
  }
  const int depth = *std::max_element(fill.begin(), fill.end());
  map.assign((size_t)depth * nX, GangEntry{-1, 0, 1, 0});
  for (int b = 0; b < nX; ++b)
    for (int k = 0; k < fill[b]; ++k) map[(size_t)k * nX + b] = bins[b][k];
  return true;
}

static int deviceCUs() {
  static const int n = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    return cus;
  }();
  return n;
}

int gangSize(int n, int capacity, int64_t nnz, int model, int maxGang, int64_t minNnzPerRank) {
  if (n <= 0) return 1;
  const int nX = xcdCount();
  if (nX <= 0) return 1;
  // a rank per CU: two ranks sharing a CU gain less than the exchange costs
  const int budget = std::min(capacity, deviceCUs() > 0 ? deviceCUs() : capacity);
  int G = 1;
  const int cap = std::min(std::min(kShards, budget / nX), std::max(1, maxGang));
  while (G * 2 <= cap && (int64_t)n * G * 2 <= budget && nnz / (G * 2) >= minNnzPerRank) G *= 2;
  return G;
}

void assignGangs(std::vector<GangEntry>& want, const std::vector<double>& priority,
                 const std::vector<int64_t>& nnz, const std::vector<int>& model, int capacity,
                 int maxGang, int64_t minNnzPerRank, bool tiers) {
  const int n = (int)want.size();
  if (n == 0) return;
  const int nX = xcdCount();
  const int budget = std::min(capacity, deviceCUs() > 0 ? deviceCUs() : capacity);
  const int cap = nX > 0 ? std::min(std::min(kShards, budget / nX), std::max(1, maxGang)) : 1;
  int used = 0;
  for (int i = 0; i < n; ++i) {
    want[i].G = gangSize(n, capacity, nnz[i], model[i], maxGang, minNnzPerRank);
    used += want[i].G;
  }
  // spare CUs go to the highest-priority fits, one doubling at a time
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return priority[a] > priority[b]; });
  if (tiers) {
    // fits of equal priority form a tier; the first tier doubles together as
    // far as it can, then the next (fits expected to need the most sweeps get
    // their gangs from the start, and a tier's fits keep equal gangs)
    for (size_t t0 = 0; t0 < order.size();) {
      size_t t1 = t0 + 1;
      while (t1 < order.size() && priority[order[t1]] == priority[order[t0]]) ++t1;
      for (;;) {
        int need = 0;
        bool ok = nX > 0;
        for (size_t k = t0; k < t1 && ok; ++k) {
          const int i = order[k];
          const int G = want[i].G;
          ok = G * 2 <= cap && nnz[i] / (G * 2) >= minNnzPerRank;
          need += G;
        }
        if (!ok || used + need > budget) break;
        for (size_t k = t0; k < t1; ++k) want[order[k]].G *= 2;
        used += need;
      }
      t0 = t1;
    }
    return;
  }
  bool grew = true;
  while (grew && nX > 0) {
    grew = false;
    for (int i : order) {
      const int G = want[i].G;
      if (G * 2 > cap || nnz[i] / (G * 2) < minNnzPerRank) continue;
      if (used + G > budget) continue;
      want[i].G = G * 2;
      used += G;
      grew = true;
    }
  }
}

int64_t slotWork(const ccdh_data* d, int N) {
  // a Cox rank also scans its pids twice per coordinate: count them as
  // nonzeros, so a gang grows with the risk-set length as with the column
  // length (a short scan split over many ranks costs more in exchanges than it saves)
  return d->model == CCDH_COX ? d->nnz + 2 * (int64_t)N : d->nnz;
}

int slotGangCapacity(bool mixed, int threads, int itl) {
  const int block = clampThreads(threads);
  const void* fn = sweepKernel(mixed, threads, itl);
  int perCU = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, block, 0) != hipSuccess) return 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return perCU * cus;
}

int slotLaunchFinish(int /*model*/, bool mixed, const void* slots, FitStatus* st,
                     const int32_t* active, int n, int threads, hipStream_t s) {
  if (n <= 0) return 0;
  threads = clampThreads(threads);
  if (mixed)
    k_slot_finish<float><<<n, threads, 0, s>>>((const Slot<float>*)slots, st, active);
  else
    k_slot_finish<double><<<n, threads, 0, s>>>((const Slot<double>*)slots, st, active);
  HIP_TRY(hipGetLastError());
  return 0;
}

namespace {
constexpr size_t kNone = ~(size_t)0;
size_t roundUp(size_t n, size_t m) { return (n + m - 1) / m * m; }
}  // namespace

SlotLayout slotLayout(int model, int K, int J, int nCap, int64_t maxLen, bool heldout, size_t rs,
                      int F) {
  SlotLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += roundUp(bytes > 0 ? bytes : 1, 256);
    return at;
  };
  const bool grouped = model == CCDH_CONDITIONAL_LOGISTIC || model == CCDH_COX;
  const bool cox = model == CCDH_COX;
  const size_t nG = grouped ? (size_t)nCap + 1 : 1;
  const size_t nR = cox ? 1 : (size_t)std::max<int64_t>(maxLen, 1);
  const size_t Fz = (size_t)std::max(1, F);
  L.row = take((size_t)K * 2 * rs * Fz);
  const bool indep = model == CCDH_LOGISTIC || model == CCDH_POISSON;
  L.rden = indep ? take((size_t)K * rs * Fz) : kNone;
  L.kw = take((size_t)K * 8 * Fz);
  L.heldout = heldout && !cox ? take((size_t)K * 8 * Fz) : kNone;
  L.grp = grouped ? take(nG * sizeof(GrpSt) * Fz) : kNone;
  L.numer = cox ? take(nG * 8) : kNone;
  L.numer2 = cox ? take(nG * 8) : kNone;
  L.pid = cox ? take((size_t)K * 4) : kNone;
  L.pidFirst = cox ? take(nG * 4) : kNone;
  L.reset = cox ? take(nG) : kNone;
  L.pidShard = cox ? take((kShards + 1) * 4) : kNone;
  L.xjy = take((size_t)J * 8 * Fz);
  L.beta = take((size_t)J * 8 * Fz);
  L.bound = take((size_t)J * 8 * Fz);
  L.pvar = take((size_t)J * 8 * Fz);
  L.ptype = take((size_t)J * Fz);
  L.fixed = take((size_t)J);
  // two record buffers of one record per (nonzero, fit): a gang alternates
  // them by coordinate (see cycleFit)
  L.recStride = (int64_t)(nR * Fz);
  const size_t nRF = nR * Fz;
  L.recRow = cox ? kNone : take(2 * nRF * 4);
  L.recO = cox ? kNone : take(2 * nRF * rs);
  L.recXo = cox ? kNone : take(2 * nRF * rs);
  const bool clr = model == CCDH_CONDITIONAL_LOGISTIC;
  L.recKey = clr ? take(2 * nRF * 4) : kNone;
  L.recFlags = clr ? take(2 * nRF) : kNone;
  L.recDen = (clr || indep) ? take(2 * nRF * 8) : kNone;
  L.bytes = off;
  return L;
}

int slotGroups(const ccdh_data* d, const double* /*w*/) {
  if (d->model == CCDH_CONDITIONAL_LOGISTIC) return d->nStrata;
  return d->K;  // Cox: at most K groups (built on the device, k_cox_groups)
}

int64_t maxColumnLength(const ccdh_data* d) {
  int64_t m = 1;
  for (int j = 0; j < d->J; ++j) m = std::max(m, d->hColPtr[j + 1] - d->hColPtr[j]);
  return m;
}

template <typename T>
static T* at(char* block, size_t off) {
  return off == kNone ? nullptr : (T*)(block + off);
}

template <typename R>
int slotPrepare(const ccdh_data* d, const ccdh_fit_spec* sps, int F, char* block,
                const SlotLayout& L, Slot<R>* out, FitStatus* st, hipStream_t s,
                const double* sharedKw) {
  const int K = d->K, J = d->J;
  if (F < 1 || F > kMaxBundle || (F & (F - 1)) != 0 || (F > 1 && d->model == CCDH_COX))
    return setError(CCDH_E_INVALID, "a slot holds 1, 2, 4 or 8 fits (Cox: 1)");
  const ccdh_fit_spec& sp = sps[0];
  Slot<R> x;
  std::memset(&x, 0, sizeof(x));
  x.d = d->view();
  x.model = d->model;
  x.K = K;
  x.J = J;
  x.F = F;
  x.lF = F == 8 ? 3 : (F == 4 ? 2 : (F == 2 ? 1 : 0));
  x.fit0 = 0;
  x.row = at<RowSt<R>>(block, L.row);
  x.rden = at<R>(block, L.rden);
  x.kw = at<double>(block, L.kw);
  x.heldout = sp.heldout ? at<double>(block, L.heldout) : nullptr;
  x.grp = at<GrpSt>(block, L.grp);
  x.numer = at<double>(block, L.numer);
  x.numer2 = at<double>(block, L.numer2);
  x.xjy = at<double>(block, L.xjy);
  x.beta = at<double>(block, L.beta);
  x.bound = at<double>(block, L.bound);
  x.pvar = at<double>(block, L.pvar);
  x.ptype = at<int8_t>(block, L.ptype);
  x.fixed = at<uint8_t>(block, L.fixed);
  x.pid = at<int32_t>(block, L.pid);
  x.pidFirst = at<int32_t>(block, L.pidFirst);
  x.reset = at<uint8_t>(block, L.reset);
  x.pidShard = at<int32_t>(block, L.pidShard);
  x.coxPieces = 1;
  x.rec.row = at<int32_t>(block, L.recRow);
  x.rec.key = at<int32_t>(block, L.recKey);
  x.rec.flags = at<uint8_t>(block, L.recFlags);
  x.rec.o = at<R>(block, L.recO);
  x.rec.xo = at<R>(block, L.recXo);
  x.rec.den = at<double>(block, L.recDen);
  x.recStride = L.recStride;
  x.ctl = FitCtl{sp.tolerance, sp.initial_bound, sp.max_iterations, sp.convergence_type,
                 sp.weights != nullptr, 0};
  // groups (Cox: K until k_cox_groups has built them under the fit's weights)
  const int N = d->model == CCDH_CONDITIONAL_LOGISTIC ? d->nStrata : K;
  x.N = N;
  // per fit: fixed term of the Poisson log likelihood (sum of w * -log(y!)) and starting bound
  for (int f = 0; f < F; ++f) {
    x.initBound[f] = sps[f].initial_bound;
    double t = 0.0;
    if (d->model == CCDH_POISSON)
      for (int k = 0; k < K; ++k)
        t += sps[f].weights ? d->hLogFact[k] * sps[f].weights[k] : d->hLogFact[k];
    x.fixedLL[f] = t;
  }
  // interleaved uploads (entry i of fit f at i*F + f)
  std::vector<double> hb((size_t)J * F), hv((size_t)J * F);
  std::vector<int8_t> ht((size_t)J * F);
  for (int f = 0; f < F; ++f)
    for (int j = 0; j < J; ++j) {
      hb[(size_t)j * F + f] = sps[f].beta0[j];
      hv[(size_t)j * F + f] = sps[f].prior_var[j];
      ht[(size_t)j * F + f] = (int8_t)sps[f].prior_type[j];
    }
  HIP_TRY(hipMemcpyAsync(x.beta, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(x.pvar, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(x.ptype, ht.data(), ht.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(x.fixed, sp.fixed, J, hipMemcpyHostToDevice, s));
  std::vector<double> hw;
  if (sp.weights && F == 1 && sharedKw) {
    x.kw = const_cast<double*>(sharedKw);  // read-only on the device
  } else if (sp.weights) {
    if (F == 1) {
      HIP_TRY(hipMemcpyAsync(x.kw, sp.weights, K * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
      hw.resize((size_t)K * F);
      for (int f = 0; f < F; ++f)
        for (int k = 0; k < K; ++k) hw[(size_t)k * F + f] = sps[f].weights[k];
      HIP_TRY(hipMemcpyAsync(x.kw, hw.data(), hw.size() * sizeof(double), hipMemcpyHostToDevice, s));
    }
  } else {
    const int64_t n = (int64_t)K * F;
    k_fill_d<<<(int)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, s>>>(x.kw, n, 1.0);
    HIP_TRY(hipGetLastError());
  }
  std::vector<double> hh;
  if (x.heldout) {
    if (F == 1) {
      HIP_TRY(hipMemcpyAsync(x.heldout, sp.heldout, K * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
      hh.resize((size_t)K * F);
      for (int f = 0; f < F; ++f)
        for (int k = 0; k < K; ++k) hh[(size_t)k * F + f] = sps[f].heldout[k];
      HIP_TRY(hipMemcpyAsync(x.heldout, hh.data(), hh.size() * sizeof(double),
                             hipMemcpyHostToDevice, s));
    }
  }
  HIP_TRY(hipStreamSynchronize(s));  // host staging goes out of scope
  *out = x;
  std::memset(st, 0, (size_t)F * sizeof(FitStatus));
  for (int f = 0; f < F; ++f) st[f].N = N;
  return 0;
}

template int slotPrepare<double>(const ccdh_data*, const ccdh_fit_spec*, int, char*,
                                 const SlotLayout&, Slot<double>*, FitStatus*, hipStream_t,
                                 const double*);
template int slotPrepare<float>(const ccdh_data*, const ccdh_fit_spec*, int, char*,
                                const SlotLayout&, Slot<float>*, FitStatus*, hipStream_t,
                                const double*);

}  // namespace ccdamd

// ================================================================ batch host side
namespace {

thread_local ccdh_batch_stats g_stats;

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

template <typename T>
int devAllocOwned(T** p, size_t n, std::vector<void*>& owned) {
  *p = nullptr;
  hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return setError(CCDH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  owned.push_back(*p);
  return 0;
}

// ---------------------------------------------------------------- autotuner
// Candidates for a workload: long-column pass width (8 or 4 nonzeros per lane)
// and whether fits may run on gangs.  Each candidate runs one launch of a fixed
// device-time quantum over every fit from its starting point; the rate is the
// coordinates completed per millisecond of kernel time.  The fits are then put
// back to their starting state.  The choice is cached per workload signature.
struct TuneKey {
  int model, K, J, F, bundle, precision, threads, maxGang;
  int64_t nnz;
  bool operator<(const TuneKey& o) const {
    return std::tie(model, K, J, F, bundle, precision, threads, maxGang, nnz) <
           std::tie(o.model, o.K, o.J, o.F, o.bundle, o.precision, o.threads, o.maxGang, o.nnz);
  }
};
std::mutex g_tuneMu;
struct TuneChoice {
  int threads, itl, gang;
};
std::map<TuneKey, TuneChoice> g_tuned;

constexpr long long kTrialTicks = 2500000;  // 25 ms of device time per candidate

// One quantum-long launch from the starting point; rate = fit-coordinates per ms.
int timeTrial(bool mixed, const void* dslots, FitStatus* dst, const GangEntry* dmap, int mapN,
              Xch* dxch, bool gangs, int threads, int itl, int nFit, int J,
              const std::vector<int>& fit0, const std::vector<int>& fitsOf, Owned& o,
              hipStream_t s, double* rate, int model, float* trialMs) {
  HIP_TRY(hipEventRecord(o.ev0, s));
  int rc = slotLaunchSweeps(mixed, dslots, dst, dmap, mapN, dxch, gangs, threads, itl, 1 << 30,
                            kTrialTicks, s, model);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(o.ev1, s));
  HIP_TRY(hipMemcpyAsync(o.hst, dst, nFit * sizeof(FitStatus), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, o.ev0, o.ev1));
  *trialMs = ms;
  double coords = 0.0;
  for (size_t b = 0; b < fit0.size(); ++b) {
    const FitStatus& h = o.hst[fit0[b]];
    if (h.fault) return setError(CCDH_E_DEVICE, "gang exchange failed during tuning");
    coords += ((double)h.iteration * J + h.jNext) * fitsOf[b];
  }
  *rate = ms > 0.f ? coords / ms : 0.0;
  return 0;
}

// Launch shape (workgroup size x pass width) x gang policy timed on this
// workload (one quantum per candidate, back to the starting point after each);
// the choice is cached per workload.  With the workgroup size left open the
// shapes are 256 x 8 and 512 x 4 (two waves per SIMD, the same pass width);
// a caller's own size keeps it and varies the pass width.  When the last
// candidate timed wins, its launch is kept as the batch's first (*kept).
template <class Trial, class Reset>
int tuneBatch(const TuneKey& key, int threads, bool autoThreads, int maxGang, Trial trial,
              Reset reset, TuneChoice* out, bool* kept) {
  *kept = false;
  const auto t0 = std::chrono::steady_clock::now();
  {
    std::lock_guard<std::mutex> g(g_tuneMu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      *out = it->second;
      return 0;
    }
  }
  std::vector<TuneChoice> cands;
  if (autoThreads) {  // the shape that usually wins last, so that its launch is kept
    cands = {{256, 8, maxGang}};
    if (maxGang > 1) cands.push_back({256, 8, 1});
    cands.push_back({512, 4, maxGang});
  } else {
    const int T = clampThreads(threads);
    cands = {{T, 8, maxGang}};
    if (maxGang > 1) cands.push_back({T, 8, 1});
    if (T == 256) cands.push_back({T, 4, maxGang});
  }
  double best = -1.0;
  TuneChoice pick = cands[0];
  for (size_t c = 0; c < cands.size(); ++c) {
    double rate = 0.0;
    int rc = trial(cands[c], &rate);
    if (rc) return rc;
    g_stats.tune_rates[c] = rate;
    if (rate > best) {
      best = rate;
      pick = cands[c];
      if (c + 1 == cands.size()) {  // the fits go on from this launch
        *kept = true;
        break;
      }
    }
    rc = reset();
    if (rc) return rc;
  }
  *out = pick;
  {
    std::lock_guard<std::mutex> g(g_tuneMu);
    g_tuned[key] = pick;
  }
  g_stats.tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// Fits that may share a slot: same fixed coordinates and the same controls.
bool bundleable(const ccdh_fit_spec& a, const ccdh_fit_spec& b, int J) {
  return a.tolerance == b.tolerance && a.max_iterations == b.max_iterations &&
         a.convergence_type == b.convergence_type && a.initial_bound == b.initial_bound &&
         a.max_bound_count == b.max_bound_count && (a.weights == nullptr) == (b.weights == nullptr) &&
         (a.heldout == nullptr) == (b.heldout == nullptr) && std::memcmp(a.fixed, b.fixed, J) == 0;
}

template <typename R>
int runBatch(ccdh_data* d, int nFit, const ccdh_fit_spec* specs, double* betasOut,
             ccdh_fit_result* results, const ccdh_batch_options* opt) {
  const auto t0 = std::chrono::steady_clock::now();
  const int J = d->J;
  const bool mixed = std::is_same<R, float>::value;
  const int model = d->model;
  const bool autoThreads = !(opt && opt->threads);
  int threads = autoThreads ? 256 : opt->threads;
  threads = std::max(64, std::min(1024, threads / 64 * 64));
  const int maxGang = (opt && opt->max_gang > 0) ? opt->max_gang : kShards;
  const int64_t minNnz = (opt && opt->gang_min_nnz > 0) ? opt->gang_min_nnz : 32768;
  int itl = (opt && opt->pass_items == 4) ? 4 : kITLDefault;
  const bool autotune = opt && opt->autotune && model != CCDH_COX;
  // graph replay of a fixed launch when gangs are off
  const bool useGraph = (opt ? opt->use_graph != 0 : true) && maxGang == 1 && !autotune;
  // Launch length: a fixed number of sweeps without gangs (every fit moves in
  // step, so the launch sequence replays); with gangs a time quantum, so a fit
  // on a large gang keeps sweeping instead of waiting for the one-workgroup fits
  // of the same launch, and gang sizes are re-dealt between launches.
  const bool timed = maxGang > 1;
  const int sweepsPerSync = (opt && opt->sweeps_per_sync > 0) ? opt->sweeps_per_sync
                                                              : (timed ? (1 << 20) : 4);
  const double quantumMs = (opt && opt->quantum_ms > 0.0) ? opt->quantum_ms : 200.0;
  const long long quantumTicks = timed ? (long long)(quantumMs * 1e5) : 0;  // 100 MHz clock
  int bundleMax = (opt && opt->bundle > 0) ? opt->bundle : 1;
  if (model == CCDH_COX) bundleMax = 1;
  if (bundleMax != 1 && bundleMax != 2 && bundleMax != 4 && bundleMax != 8)
    return setError(CCDH_E_INVALID, "bundle must be 1, 2, 4 or 8");

  // slots: runs of consecutive fits that may share one, split into powers of two
  std::vector<int> fit0, fitsOf;
  for (int a = 0; a < nFit;) {
    int e = a + 1;
    while (e < nFit && e - a < bundleMax && bundleable(specs[a], specs[e], J)) ++e;
    for (int left = e - a; left > 0;) {
      int b = 1;
      while (b * 2 <= left) b *= 2;
      fit0.push_back(a);
      fitsOf.push_back(b);
      a += b;
      left -= b;
    }
  }
  const int nSlot = (int)fit0.size();
  int Fmax = 1;
  for (int b : fitsOf) Fmax = std::max(Fmax, b);
  std::vector<int> slotOf(nFit);
  for (int b = 0; b < nSlot; ++b)
    for (int i = 0; i < fitsOf[b]; ++i) slotOf[fit0[b] + i] = b;

  Owned o;
  HIP_TRY(hipStreamCreateWithFlags(&o.stream, hipStreamNonBlocking));
  hipStream_t s = o.stream;

  // one block per slot, all with the layout of the largest group count and bundle
  int maxN = 1;
  bool anyHeld = false;
  for (int f = 0; f < nFit; ++f) {
    maxN = std::max(maxN, slotGroups(d, specs[f].weights));
    anyHeld |= specs[f].heldout != nullptr;
  }
  const SlotLayout L = slotLayout(model, d->K, J, maxN, maxColumnLength(d), anyHeld, sizeof(R), Fmax);
  // co-resident workgroups of the launch shapes this batch may use
  int capacity = std::min(slotGangCapacity(mixed, threads, 8), slotGangCapacity(mixed, threads, 4));
  if (autoThreads && opt && opt->autotune) capacity = std::min(capacity, slotGangCapacity(mixed, 512, 4));
  if (capacity <= 0) return setError(CCDH_E_DEVICE, "sweep kernel occupancy query failed");
  (void)xcdCount();
  char* arena = nullptr;
  Slot<R>* dslots = nullptr;
  FitStatus* dst = nullptr;
  int32_t* dactive = nullptr;
  GangEntry* dmap = nullptr;
  Xch* dxch = nullptr;
  const int mapCap = std::max(nSlot, capacity);
  int rc = 0;
  rc = rc ? rc : devAllocOwned(&arena, (size_t)nSlot * L.bytes, o.ptrs);
  rc = rc ? rc : devAllocOwned(&dslots, nSlot, o.ptrs);
  rc = rc ? rc : devAllocOwned(&dst, nFit, o.ptrs);
  rc = rc ? rc : devAllocOwned(&dactive, nSlot, o.ptrs);
  rc = rc ? rc : devAllocOwned(&dmap, mapCap, o.ptrs);
  rc = rc ? rc : devAllocOwned(&dxch, nSlot, o.ptrs);
  if (rc) return rc;
  HIP_TRY(hipHostMalloc((void**)&o.hst, nFit * sizeof(FitStatus), hipHostMallocDefault));

  std::vector<Slot<R>> slots(nSlot);
  std::vector<FitStatus> st(nFit);
  std::vector<int32_t> active(nSlot);
  std::vector<int> boundCount(nFit, 1);
  // Slots of one fit whose specs pass the same host weights array (a CV grid's
  // fits of one fold, given one array per fold) read the first such slot's
  // device copy of it and take over its weight-only set-up (X'Y, group weight
  // sums) instead of repeating the upload and the pass over every nonzero.
  std::vector<int> shareOf(nSlot, -1);
  {
    std::map<const double*, int> firstOf;
    for (int b = 0; b < nSlot; ++b) {
      const double* w = specs[fit0[b]].weights;
      if (fitsOf[b] != 1 || w == nullptr) continue;
      const auto it = firstOf.find(w);
      if (it == firstOf.end())
        firstOf.emplace(w, b);
      else
        shareOf[b] = it->second;
    }
  }
  for (int b = 0; b < nSlot; ++b) {
    rc = slotPrepare<R>(d, specs + fit0[b], fitsOf[b], arena + (size_t)b * L.bytes, L, &slots[b],
                        &st[fit0[b]], s, shareOf[b] >= 0 ? slots[shareOf[b]].kw : nullptr);
    if (rc) return rc;
    slots[b].fit0 = fit0[b];
    slots[b].ctl.solo = (opt && opt->solo < 0) ? -1 : 0;
    slots[b].ctl.spinLimit = (opt && opt->spin_limit > 0) ? opt->spin_limit : 0;
    active[b] = b;
  }
  HIP_TRY(hipMemcpyAsync(dslots, slots.data(), nSlot * sizeof(Slot<R>), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dst, st.data(), nFit * sizeof(FitStatus), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(dactive, active.data(), nSlot * sizeof(int32_t), hipMemcpyHostToDevice, s));
  if (model == CCDH_COX) {  // risk-set groups of every fit under its weights, then their counts
    rc = slotLaunchCoxGroups(mixed, dslots, dst, dactive, nSlot, threads, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(st.data(), dst, nFit * sizeof(FitStatus), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int b = 0; b < nSlot; ++b) slots[b].N = st[fit0[b]].N;
  }
  const int setupRanks = initRanks(d->nnz, d->K);
  const bool grouped = model == CCDH_CONDITIONAL_LOGISTIC || model == CCDH_COX;
  std::vector<int32_t> setupOrder;  // slots with their own set-up, then those sharing one
  for (int b = 0; b < nSlot; ++b) {
    if (shareOf[b] >= 0 && slots[b].N != slots[shareOf[b]].N) shareOf[b] = -1;
    if (shareOf[b] < 0) setupOrder.push_back(b);
  }
  const int nOwn = (int)setupOrder.size();
  for (int b = 0; b < nSlot; ++b)
    if (shareOf[b] >= 0) setupOrder.push_back(b);
  if (nOwn == nSlot) {
    rc = slotLaunchInit(model, mixed, dslots, dst, dactive, nSlot, threads, s, true, setupRanks);
    if (rc) return rc;
  } else {
    HIP_TRY(hipMemcpyAsync(dactive, setupOrder.data(), nSlot * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
    rc = slotLaunchInit(model, mixed, dslots, dst, dactive, nOwn, threads, s, true, setupRanks);
    if (rc) return rc;
    // X'Y and the group weight sums (the denominators beside them are rebuilt
    // by the set-up of the slot's own rows), then the rest of the set-up
    for (int k = nOwn; k < nSlot; ++k) {
      const int b = setupOrder[k], a = shareOf[b];
      HIP_TRY(hipMemcpyAsync(slots[b].xjy, slots[a].xjy, (size_t)J * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
      if (grouped)
        HIP_TRY(hipMemcpyAsync(slots[b].grp, slots[a].grp, (size_t)slots[a].N * sizeof(GrpSt),
                               hipMemcpyDeviceToDevice, s));
    }
    rc = slotLaunchInit(model, mixed, dslots, dst, dactive + nOwn, nSlot - nOwn, threads, s, false,
                        setupRanks);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(dactive, active.data(), nSlot * sizeof(int32_t), hipMemcpyHostToDevice,
                           s));
  }
  // work per sweep of each slot (its fits; Cox: its own tie-group count) for the gang sizes
  std::vector<int64_t> work(nSlot);
  for (int b = 0; b < nSlot; ++b) work[b] = slotWork(d, st[fit0[b]].N) * fitsOf[b];

  // Launch map: every slot with a running fit, each on a gang sized by how many
  // slots are left (the tail of slow fits spreads over the idle CUs).  Slots
  // expected to need the most sweeps first: the weakest prior (largest mean
  // variance over the free coordinates) of their fits.
  std::vector<double> priority(nSlot, 0.0);
  for (int f = 0; f < nFit; ++f) {
    double sum = 0.0;
    int cnt = 0;
    for (int j = 0; j < J; ++j)
      if (!specs[f].fixed[j] && specs[f].prior_type[j] != CCDH_PRIOR_NONE) {
        sum += specs[f].prior_var[j];
        ++cnt;
      }
    const double strength = cnt ? sum / cnt : 1e300;
    priority[slotOf[f]] = std::max(priority[slotOf[f]], strength);
  }
  std::vector<GangEntry> map;
  int mapN = 0;
  bool mapGangs = false;
  std::vector<char> running(nSlot, 1);
  int gangCap = maxGang;
  int curFits = nFit, curGang = 1;
  std::vector<char> oneGroup(nSlot, 0);  // slots held to one workgroup
  auto buildMap = [&]() -> int {
    std::vector<GangEntry> gangs;
    std::vector<double> prio;
    std::vector<int64_t> fw;
    for (int b = 0; b < nSlot; ++b)
      if (running[b] || useGraph) {
        gangs.push_back(GangEntry{b, 0, 1, 0});
        prio.push_back(priority[b]);
        fw.push_back(work[b]);
      }
    if (!useGraph && gangCap > 1)
      assignGangs(gangs, prio, fw, std::vector<int>(gangs.size(), model), capacity, gangCap,
                  minNnz, true);
    for (GangEntry& g : gangs)
      if (oneGroup[g.slot]) g.G = 1;  // restarted after a gang wait gave up
    int G = 1;
    for (const GangEntry& g : gangs) G = std::max(G, g.G);
    if (G > 1 && !layoutGangs(gangs, capacity, xcdCount(), map)) G = 1;
    if (G == 1) {
      map.clear();
      for (const GangEntry& g : gangs) map.push_back(GangEntry{g.slot, 0, 1, 0});
    }
    mapN = (int)map.size();
    mapGangs = G > 1;
    curGang = G;
    HIP_TRY(hipMemcpyAsync(dmap, map.data(), mapN * sizeof(GangEntry), hipMemcpyHostToDevice, s));
    return 0;
  };
  // sweep launches, optionally replayed from a captured graph
  auto enqueueSweeps = [&]() -> int {
    int r = slotLaunchSweeps(mixed, dslots, dst, dmap, mapN, dxch, mapGangs, threads, itl,
                             sweepsPerSync, quantumTicks, s, model);
    if (r) return r;
    HIP_TRY(hipMemcpyAsync(o.hst, dst, nFit * sizeof(FitStatus), hipMemcpyDeviceToHost, s));
    return 0;
  };
  // coefficients of the fits of slot b in `mask` back to b0(f) (interleaved in
  // the slot: the others are read back and kept)
  auto uploadStart = [&](int b, unsigned mask, auto b0) -> int {
    const int Fb = fitsOf[b];
    std::vector<double> hb((size_t)J * Fb);
    if (mask != (1u << Fb) - 1u) {
      HIP_TRY(hipMemcpyAsync(hb.data(), slots[b].beta, hb.size() * sizeof(double),
                             hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
    }
    for (int i = 0; i < Fb; ++i) {
      if (!((mask >> i) & 1u)) continue;
      const double* src = b0(fit0[b] + i);
      for (int j = 0; j < J; ++j) hb[(size_t)j * Fb + i] = src[j];
    }
    HIP_TRY(hipMemcpyAsync(slots[b].beta, hb.data(), hb.size() * sizeof(double),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));  // hb goes out of scope
    return 0;
  };
  std::memset(&g_stats, 0, sizeof(g_stats));
  g_stats.shared_setups = nSlot - nOwn;
  g_stats.tuned_pass_items = itl;
  g_stats.tuned_max_gang = gangCap;
  g_stats.tuned_threads = clampThreads(threads);
  if (!o.ev0) HIP_TRY(hipEventCreate(&o.ev0));
  if (!o.ev1) HIP_TRY(hipEventCreate(&o.ev1));
  if (autotune) {
    const TuneKey key{d->model, d->K, d->J, nFit, bundleMax, d->precision,
                      autoThreads ? 0 : clampThreads(threads), maxGang, d->nnz};
    TuneChoice choice{};
    // the starting point (every slot's state and every fit's status after
    // set-up), copied aside before the first trial and copied back after each;
    // without room for the copy, the coefficients are uploaded again and the
    // set-up kernel re-run instead.  A sweep writes three ranges of a slot's
    // block: the row state, the group state (Cox: and its numerators), the
    // coefficients and bounds; the rest is read-only or scratch (records).
    const size_t arenaBytes = (size_t)nSlot * L.bytes;
    struct Range {
      size_t at, end;
    };
    std::vector<Range> mutableRanges = {{L.row, L.kw}, {L.beta, L.pvar}};
    if (L.grp != kNone) mutableRanges.push_back({L.grp, model == CCDH_COX ? L.pid : L.xjy});
    auto copySlots = [&](char* dstBase, const char* srcBase) -> int {
      for (const Range& r : mutableRanges)
        HIP_TRY(hipMemcpy2DAsync(dstBase + r.at, L.bytes, srcBase + r.at, L.bytes, r.end - r.at,
                                 nSlot, hipMemcpyDeviceToDevice, s));
      return 0;
    };
    char* snapArena = nullptr;
    FitStatus* snapSt = nullptr;
    bool snapTried = false;
    float lastTrialMs = 0.f;
    bool keptTrial = false;
    rc = tuneBatch(
        key, threads, autoThreads, maxGang,
        [&](const TuneChoice& cand, double* rate) -> int {
          threads = cand.threads;
          itl = cand.itl;
          gangCap = cand.gang;
          if (!snapTried) {
            snapTried = true;
            if (hipMalloc((void**)&snapArena, arenaBytes) == hipSuccess &&
                hipMalloc((void**)&snapSt, nFit * sizeof(FitStatus)) == hipSuccess) {
              o.ptrs.push_back(snapArena);
              o.ptrs.push_back(snapSt);
              const int cr = copySlots(snapArena, arena);
              if (cr) return cr;
              HIP_TRY(hipMemcpyAsync(snapSt, dst, nFit * sizeof(FitStatus),
                                     hipMemcpyDeviceToDevice, s));
            } else {
              (void)hipGetLastError();
              if (snapArena) (void)hipFree(snapArena);
              snapArena = nullptr;
              snapSt = nullptr;
            }
          }
          int r = buildMap();
          if (r) return r;
          return timeTrial(mixed, dslots, dst, dmap, mapN, dxch, mapGangs, threads, itl, nFit, J,
                           fit0, fitsOf, o, s, rate, model, &lastTrialMs);
        },
        [&]() -> int {
          if (snapSt) {
            const int cr = copySlots(arena, snapArena);
            if (cr) return cr;
            HIP_TRY(hipMemcpyAsync(dst, snapSt, nFit * sizeof(FitStatus), hipMemcpyDeviceToDevice,
                                   s));
            HIP_TRY(hipMemcpyAsync(dactive, active.data(), nSlot * sizeof(int32_t),
                                   hipMemcpyHostToDevice, s));
            HIP_TRY(hipStreamSynchronize(s));
            return 0;
          }
          // back to the starting point: coefficients, statuses, then the set-up kernel
          for (int b = 0; b < nSlot; ++b) {
            const int r = uploadStart(b, (1u << fitsOf[b]) - 1u,
                                      [&](int f) { return specs[f].beta0; });
            if (r) return r;
          }
          HIP_TRY(hipMemcpyAsync(dst, st.data(), nFit * sizeof(FitStatus), hipMemcpyHostToDevice, s));
          HIP_TRY(hipMemcpyAsync(dactive, active.data(), nSlot * sizeof(int32_t),
                                 hipMemcpyHostToDevice, s));
          const int r = slotLaunchInit(model, mixed, dslots, dst, dactive, nSlot, threads, s, false,
                                       setupRanks);
          if (r) return r;
          HIP_TRY(hipStreamSynchronize(s));
          return 0;
        },
        &choice, &keptTrial);
    if (rc) return rc;
    threads = choice.threads;
    itl = choice.itl;
    gangCap = choice.gang;
    if (keptTrial) {  // the winner's trial launch is the batch's first
      g_stats.kernel_ms += lastTrialMs;
      g_stats.launch_ms[0] = lastTrialMs;
      g_stats.launch_fits[0] = curFits;
      g_stats.launch_gang[0] = curGang;
      g_stats.launches = 1;
    }
    g_stats.tuned_pass_items = itl;
    g_stats.tuned_max_gang = gangCap;
    g_stats.tuned_threads = clampThreads(threads);
  }
  rc = buildMap();
  if (rc) return rc;
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
  std::vector<double> doneAt(nFit, 0.0);
  std::vector<int> lastIter(nFit, 0);
  std::vector<char> pendingRestart(nFit, 0);
  int guard = 0;
  for (;;) {
    if (!useGraph) {
      rc = buildMap();
      if (rc) return rc;
    }
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
      if (g_stats.launches < 128) {
        g_stats.launch_ms[g_stats.launches] = ms;
        g_stats.launch_fits[g_stats.launches] = curFits;
        g_stats.launch_gang[g_stats.launches] = curGang;
      }
      g_stats.launches += 1;
      int rounds = 0;
      for (int f = 0; f < nFit; ++f) {
        rounds = std::max(rounds, o.hst[f].iteration - lastIter[f]);
        lastIter[f] = o.hst[f].iteration;
      }
      g_stats.sweep_rounds += rounds;
    }
    const double now = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    bool misplaced = false;
    for (int b = 0; b < nSlot; ++b) {
      FitStatus& h = o.hst[fit0[b]];
      if (h.fault == 1) {
        // A rank of the slot's gang gave up waiting for its peers (other kernels
        // holding CUs can keep a rank from being scheduled): the ranks may have
        // left the fits' state half updated, so every fit of the slot starts over
        // from its starting coefficients with its current bound, on one
        // workgroup, where no rank waits for another.
        oneGroup[b] = 1;
        g_stats.gang_restarts += 1;
        h.fault = 0;
        HIP_TRY(hipMemcpyAsync(&dst[fit0[b]].fault, &h.fault, sizeof(int), hipMemcpyHostToDevice, s));
        const unsigned all = (1u << fitsOf[b]) - 1u;
        rc = uploadStart(b, all, [&](int f) { return specs[f].beta0; });
        if (rc) return rc;
        active[0] = b;
        HIP_TRY(hipMemcpyAsync(dactive, active.data(), sizeof(int32_t), hipMemcpyHostToDevice, s));
        rc = slotLaunchInit(model, mixed, dslots, dst, dactive, 1, threads, s, false, setupRanks,
                            all);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(o.hst + fit0[b], dst + fit0[b], fitsOf[b] * sizeof(FitStatus),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < fitsOf[b]; ++i) {
          const int f = fit0[b] + i;
          pendingRestart[f] = 0;
          doneAt[f] = 0.0;
          lastIter[f] = 0;
        }
        for (int i = 0; i < nSlot; ++i) active[i] = i;
        HIP_TRY(hipMemcpyAsync(dactive, active.data(), nSlot * sizeof(int32_t), hipMemcpyHostToDevice, s));
        continue;
      }
      if (h.fault == 2) {  // a gang spanned XCDs: it backed off untouched
        misplaced = true;
        h.fault = 0;
        HIP_TRY(hipMemcpyAsync(&dst[fit0[b]].fault, &h.fault, sizeof(int), hipMemcpyHostToDevice, s));
      }
    }
    if (misplaced) {
      gangCap = 1;
      HIP_TRY(hipStreamSynchronize(s));
    }
    // CyclicCoordinateDescent::update: ill-conditioned -> bound / 10, cold
    // restart.  A fit of a bundle restarts from coordinate 0 once the other
    // fits of its slot have finished (they move in lock step).
    curFits = 0;
    bool allDone = true;
    for (int f = 0; f < nFit; ++f) {
      const FitStatus& h = o.hst[f];
      if (!h.done) {
        ++curFits;
        continue;
      }
      if (doneAt[f] == 0.0) doneAt[f] = now;
      if (h.flag == CCDH_ILLCONDITIONED && boundCount[f] < specs[f].max_bound_count)
        pendingRestart[f] = 1;
    }
    for (int b = 0; b < nSlot; ++b) {
      bool busy = false, pend = false;
      unsigned mask = 0;
      for (int i = 0; i < fitsOf[b]; ++i) {
        const int f = fit0[b] + i;
        busy |= !o.hst[f].done;
        if (pendingRestart[f]) {
          pend = true;
          mask |= 1u << i;
        }
      }
      if (pend && !busy) {
        std::vector<double> zero(J, 0.0);  // cold restart (the offset keeps its 1)
        std::vector<std::vector<double>> cold(fitsOf[b]);
        for (int i = 0; i < fitsOf[b]; ++i) {
          const int f = fit0[b] + i;
          if (!((mask >> i) & 1u)) continue;
          pendingRestart[f] = 0;
          boundCount[f]++;
          slots[b].initBound[i] /= 10.0;
          cold[i] = zero;
          if (d->hasOffset) cold[i][0] = specs[f].beta0[0];
          doneAt[f] = 0.0;
          lastIter[f] = 0;
        }
        rc = uploadStart(b, mask, [&](int f) { return cold[f - fit0[b]].data(); });
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(dslots + b, &slots[b], sizeof(Slot<R>), hipMemcpyHostToDevice, s));
        active[0] = b;
        HIP_TRY(hipMemcpyAsync(dactive, active.data(), sizeof(int32_t), hipMemcpyHostToDevice, s));
        rc = slotLaunchInit(model, mixed, dslots, dst, dactive, 1, threads, s, false, setupRanks,
                            mask);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < fitsOf[b]; ++i)
          if ((mask >> i) & 1u) o.hst[fit0[b] + i].done = 0;
        busy = true;
      }
      running[b] = busy;
      allDone = allDone && !busy && !pend;
    }
    for (int b = 0; b < nSlot; ++b) active[b] = b;
    HIP_TRY(hipMemcpyAsync(dactive, active.data(), nSlot * sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (allDone) break;
    if (++guard > 100000000) return setError(CCDH_E_STATE, "batch did not terminate");
  }
  rc = slotLaunchFinish(model, mixed, dslots, dst, dactive, nSlot, threads, s);
  if (rc) return rc;
  std::vector<FitStatus> fin(nFit);
  HIP_TRY(hipMemcpyAsync(fin.data(), dst, nFit * sizeof(FitStatus), hipMemcpyDeviceToHost, s));
  std::vector<double> hb((size_t)J * Fmax);
  for (int b = 0; b < nSlot; ++b) {
    const int Fb = fitsOf[b];
    double* dstB = Fb == 1 ? betasOut + (size_t)fit0[b] * J : hb.data();
    HIP_TRY(hipMemcpyAsync(dstB, slots[b].beta, (size_t)J * Fb * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    if (Fb == 1) continue;
    HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < Fb; ++i)
      for (int j = 0; j < J; ++j) betasOut[(size_t)(fit0[b] + i) * J + j] = hb[(size_t)j * Fb + i];
  }
  HIP_TRY(hipStreamSynchronize(s));
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int f = 0; f < nFit; ++f) {
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
  g_stats.bundles = nSlot;
  for (int f = 0; f < nFit; ++f) {
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
  if (model == CCDH_COX && anyHeld) {
    for (int f = 0; f < nFit; ++f) {
      if (!specs[f].heldout) continue;
      double v = 0.0;
      rc = coxHeldOut(d, specs[f].weights, specs[f].heldout, betasOut + (size_t)f * J, &v);
      if (rc) return rc;
      results[f].pred_log_likelihood = v;
    }
  }
  return 0;
}

}  // namespace

namespace ccdamd {

// Predictive log likelihood of a Cox fit on held-out weights, by the
// per-coordinate engine (the risk sets are regrouped under the held-out weights).
int coxHeldOut(ccdh_data* d, const double* weights, const double* heldout, const double* beta,
               double* out) {
  ccdh_fit* fh = nullptr;
  int rc = fitCreate(d, &fh);
  if (rc) return rc;
  const int useW = weights != nullptr;
  rc = setWeights(fh, weights, useW);
  if (!rc) {
    rc = zeroXBeta(fh);
    for (int j = 0; !rc && j < d->J; ++j) rc = axpyXBeta(fh, beta[j], j);
  }
  rc = rc ? rc : remainingStats(fh, useW);
  rc = rc ? rc : predLogLik(fh, heldout, out);
  fitDestroy(fh);
  return rc;
}

int validateSpec(const ccdh_data* d, const ccdh_fit_spec& sp) {
  if (!sp.prior_type || !sp.prior_var || !sp.fixed || !sp.beta0)
    return setError(CCDH_E_INVALID, "fit spec arrays required");
  if (sp.convergence_type < 0 || sp.convergence_type > 3)
    return setError(CCDH_E_INVALID, "unsupported convergence type");
  for (int j = 0; j < d->J; ++j)
    if (sp.prior_type[j] < 0 || sp.prior_type[j] > 2)
      return setError(CCDH_E_INVALID, "unsupported prior type");
  return 0;
}

}  // namespace ccdamd

namespace {

template <typename R>
int coxGroupsOnDevice(ccdh_data* d, const double* weights, int32_t* pidOut, uint8_t* resetOut,
                      int32_t* pidFirstOut, int32_t* pidShardOut, int* nOut) {
  const int K = d->K, J = d->J;
  std::vector<int32_t> pt(J, CCDH_PRIOR_NONE);
  std::vector<double> pv(J, 1.0), b0(J, 0.0);
  std::vector<uint8_t> fx(J, 0);
  ccdh_fit_spec sp{pt.data(), pv.data(), fx.data(), b0.data(), weights, nullptr, 1e-6, 1, 0, 2.0, 1};
  Owned o;
  HIP_TRY(hipStreamCreateWithFlags(&o.stream, hipStreamNonBlocking));
  const SlotLayout L = slotLayout(CCDH_COX, K, J, K, 1, false, sizeof(R), 1);
  char* block = nullptr;
  Slot<R>* dslot = nullptr;
  FitStatus* dst = nullptr;
  int32_t* dact = nullptr;
  int rc = devAllocOwned(&block, L.bytes, o.ptrs);
  rc = rc ? rc : devAllocOwned(&dslot, 1, o.ptrs);
  rc = rc ? rc : devAllocOwned(&dst, 1, o.ptrs);
  rc = rc ? rc : devAllocOwned(&dact, 1, o.ptrs);
  if (rc) return rc;
  Slot<R> x;
  FitStatus st;
  rc = slotPrepare<R>(d, &sp, 1, block, L, &x, &st, o.stream);
  if (rc) return rc;
  x.coxPieces = 0;  // the tie groups themselves (pids)
  const int32_t zero = 0;
  HIP_TRY(hipMemcpyAsync(dslot, &x, sizeof(x), hipMemcpyHostToDevice, o.stream));
  HIP_TRY(hipMemcpyAsync(dst, &st, sizeof(st), hipMemcpyHostToDevice, o.stream));
  HIP_TRY(hipMemcpyAsync(dact, &zero, sizeof(zero), hipMemcpyHostToDevice, o.stream));
  rc = slotLaunchCoxGroups(std::is_same<R, float>::value, dslot, dst, dact, 1, 256, o.stream);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(&st, dst, sizeof(st), hipMemcpyDeviceToHost, o.stream));
  HIP_TRY(hipStreamSynchronize(o.stream));
  const int N = st.N;
  HIP_TRY(hipMemcpyAsync(pidOut, x.pid, (size_t)K * sizeof(int32_t), hipMemcpyDeviceToHost, o.stream));
  if (resetOut) HIP_TRY(hipMemcpyAsync(resetOut, x.reset, N + 1, hipMemcpyDeviceToHost, o.stream));
  if (pidFirstOut)
    HIP_TRY(hipMemcpyAsync(pidFirstOut, x.pidFirst, (size_t)(N + 1) * sizeof(int32_t),
                           hipMemcpyDeviceToHost, o.stream));
  if (pidShardOut)
    HIP_TRY(hipMemcpyAsync(pidShardOut, x.pidShard, (kShards + 1) * sizeof(int32_t),
                           hipMemcpyDeviceToHost, o.stream));
  HIP_TRY(hipStreamSynchronize(o.stream));
  *nOut = N;
  return 0;
}

// ------------------------------------------- Cox grouping without allocation
// ccdh_cox_groups_noalloc stages everything ccdh_cox_groups allocates per call
// (the slot descriptor, the fit status, the weights and the grouping arrays the
// kernel writes) in one file-scope device array instead.  Each GPU holds its own
// copy for the life of the process; its address on a device comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr int kNoAllocMaxRows = 1 << 20;

struct NoAllocLayout {
  size_t slot, st, active, kw, pid, pidFirst, reset, numer, numer2, pidShard, bytes;
};

constexpr size_t naAlign(size_t v) { return (v + 255) & ~(size_t)255; }

constexpr NoAllocLayout noAllocLayout(size_t K) {
  NoAllocLayout L{};
  size_t off = 0;
  const size_t slotBytes =
      sizeof(Slot<double>) > sizeof(Slot<float>) ? sizeof(Slot<double>) : sizeof(Slot<float>);
  L.slot = off; off = naAlign(off + slotBytes);
  L.st = off; off = naAlign(off + sizeof(FitStatus));
  L.active = off; off = naAlign(off + sizeof(int32_t));
  L.kw = off; off = naAlign(off + K * sizeof(double));
  L.pid = off; off = naAlign(off + K * sizeof(int32_t));
  L.pidFirst = off; off = naAlign(off + (K + 1) * sizeof(int32_t));
  L.reset = off; off = naAlign(off + (K + 1));
  L.numer = off; off = naAlign(off + (K + 1) * sizeof(double));
  L.numer2 = off; off = naAlign(off + (K + 1) * sizeof(double));
  L.pidShard = off; off = naAlign(off + (kShards + 1) * sizeof(int32_t));
  L.bytes = off;
  return L;
}

constexpr size_t kNoAllocBytes = noAllocLayout(kNoAllocMaxRows).bytes;  // about 33 MiB

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

__global__ void k_na_fill(double* a, int n, double v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

std::mutex g_naMu;
std::map<int, char*> g_naScratch;  // device -> address of its copy of `scratch`

// The address of this device's copy of `scratch` (the current device's), looked
// up on its first use; the caller holds g_naMu.
int noAllocScratch(char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  auto it = g_naScratch.find(dev);
  if (it == g_naScratch.end()) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    it = g_naScratch.emplace(dev, (char*)ptr).first;
  }
  *out = it->second;
  return 0;
}

template <typename R>
int coxGroupsNoAlloc(ccdh_data* d, const double* weights, int32_t* pidOut, uint8_t* resetOut,
                     int32_t* pidFirstOut, int32_t* pidShardOut, int* nOut) {
  const int K = d->K;
  std::lock_guard<std::mutex> lock(g_naMu);  // calls in this process take turns
  char* base = nullptr;
  int rc = noAllocScratch(&base);
  if (rc) return rc;
  const NoAllocLayout L = noAllocLayout(kNoAllocMaxRows);
  Slot<R> x;
  std::memset(&x, 0, sizeof(x));
  x.d = d->view();
  x.model = d->model;
  x.K = K;
  x.J = d->J;
  x.N = K;
  x.F = 1;
  x.lF = 0;
  x.fit0 = 0;
  x.kw = (double*)(base + L.kw);
  x.pid = (int32_t*)(base + L.pid);
  x.pidFirst = (int32_t*)(base + L.pidFirst);
  x.reset = (uint8_t*)(base + L.reset);
  x.numer = (double*)(base + L.numer);
  x.numer2 = (double*)(base + L.numer2);
  x.pidShard = (int32_t*)(base + L.pidShard);
  FitStatus st;
  std::memset(&st, 0, sizeof(st));
  const int32_t zero = 0;
  // the null stream: ordered after every earlier call of this thread on the device
  hipStream_t s = nullptr;
  HIP_TRY(hipMemcpyAsync(base + L.slot, &x, sizeof(x), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(base + L.st, &st, sizeof(st), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(base + L.active, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
  if (weights) {
    HIP_TRY(hipMemcpyAsync(x.kw, weights, (size_t)K * sizeof(double), hipMemcpyHostToDevice, s));
  } else {
    k_na_fill<<<std::min((K + 255) / 256, 1024), 256, 0, s>>>(x.kw, K, 1.0);
    HIP_TRY(hipGetLastError());
  }
  rc = slotLaunchCoxGroups(std::is_same<R, float>::value, base + L.slot, (FitStatus*)(base + L.st),
                           (const int32_t*)(base + L.active), 1, 256, s);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(&st, base + L.st, sizeof(st), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const int N = st.N;
  HIP_TRY(hipMemcpyAsync(pidOut, x.pid, (size_t)K * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (resetOut) HIP_TRY(hipMemcpyAsync(resetOut, x.reset, N + 1, hipMemcpyDeviceToHost, s));
  if (pidFirstOut)
    HIP_TRY(hipMemcpyAsync(pidFirstOut, x.pidFirst, (size_t)(N + 1) * sizeof(int32_t),
                           hipMemcpyDeviceToHost, s));
  if (pidShardOut)
    HIP_TRY(hipMemcpyAsync(pidShardOut, x.pidShard, (kShards + 1) * sizeof(int32_t),
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *nOut = N;
  return 0;
}

}  // namespace

extern "C" int ccdh_cox_groups(ccdh_data_t data, const double* weights, int32_t* pid_out,
                               uint8_t* reset_out, int32_t* pid_first_out,
                               int32_t* shard_pid_out, int* n_out) {
  if (!data || !pid_out || !n_out) return setError(CCDH_E_INVALID, "null argument");
  if (data->model != CCDH_COX) return setError(CCDH_E_INVALID, "not a Cox data set");
  try {
    return data->precision == CCDH_MIXED
               ? coxGroupsOnDevice<float>(data, weights, pid_out, reset_out, pid_first_out,
                                          shard_pid_out, n_out)
               : coxGroupsOnDevice<double>(data, weights, pid_out, reset_out, pid_first_out,
                                           shard_pid_out, n_out);
  } catch (const std::exception& e) {
    return setError(CCDH_E_DEVICE, e.what());
  }
}

extern "C" int ccdh_cox_groups_noalloc(ccdh_data_t data, const double* weights,
                                       int32_t* pid_out, uint8_t* reset_out,
                                       int32_t* pid_first_out, int32_t* shard_pid_out,
                                       int* n_out) {
  if (!data || !pid_out || !n_out) return setError(CCDH_E_INVALID, "null argument");
  if (data->model != CCDH_COX) return setError(CCDH_E_INVALID, "not a Cox data set");
  if (data->K > kNoAllocMaxRows)
    return setError(CCDH_E_INVALID, "ccdh_cox_groups_noalloc: " + std::to_string(data->K) +
                                        " rows, at most " + std::to_string(kNoAllocMaxRows));
  try {
    return data->precision == CCDH_MIXED
               ? coxGroupsNoAlloc<float>(data, weights, pid_out, reset_out, pid_first_out,
                                         shard_pid_out, n_out)
               : coxGroupsNoAlloc<double>(data, weights, pid_out, reset_out, pid_first_out,
                                          shard_pid_out, n_out);
  } catch (const std::exception& e) {
    return setError(CCDH_E_DEVICE, e.what());
  }
}

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
    const int rc = validateSpec(data, specs[f]);
    if (rc) return rc;
  }
  try {
    return data->precision == CCDH_MIXED
               ? runBatch<float>(data, n_fits, specs, betas_out, results, options)
               : runBatch<double>(data, n_fits, specs, betas_out, results, options);
  } catch (const std::exception& e) {
    return setError(CCDH_E_DEVICE, e.what());
  }
}
