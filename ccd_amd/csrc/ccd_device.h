// ccd_device.h -- device-side building blocks shared by the per-coordinate
// (plain) engine and the device-resident batched engine.
//
// Everything here is written for gfx950: 64-lane wavefronts, FP64 shuffles for
// reductions, LDS for the cross-wave step.  Reductions are deterministic: the
// partition of work over threads and blocks is fixed by the launch shape and
// partial results are combined in a fixed tree order, so a fit produces the same
// bits alone and inside a batch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccdamd {

enum : int { F_DENSE = 0, F_SPARSE = 1, F_INDICATOR = 2, F_INTERCEPT = 3 };
enum : int { M_POISSON = 2, M_LOGISTIC = 3, M_CLR = 4, M_COX = 8 };
enum : int { P_NONE = 0, P_LAPLACE = 1, P_NORMAL = 2 };

constexpr int kWave = 64;
constexpr int kMaxBlockThreads = 1024;

// Read-only, device-resident data shared by every fit over it.
struct DevX {
  int K, J;
  int nStrata;              // CLR strata / Cox strata (stratum ids 0..nStrata-1)
  const int64_t* colPtr;    // [J+1]
  const int32_t* rows;      // [nnz]
  const double* vals;       // [nnz] FP64 values (nullptr if none)
  const float* valsF;       // [nnz] FP32 copy (mixed precision only)
  const int8_t* fmt;        // [J]
  const double* y;          // [K]
  const double* time;       // [K] Cox event / censoring times (nullptr otherwise)
  const int32_t* strata;    // [K]
  // grouped models: runs of equal stratum inside each column
  const int64_t* segPtr;    // [J+1] run range of column j
  const int32_t* segOff;    // [nseg] run start, relative to colPtr[j]
  const int32_t* segKey;    // [nseg] stratum of the run
  const int32_t* stratStart;// [nStrata+1] first row of each stratum
  const int32_t* nzKey;     // [nnz] stratum of each nonzero's row (grouped models)
  // row shards for workgroup gangs: shard s holds rows [shardRow[s], shardRow[s+1]),
  // boundaries on stratum (CLR) / tie-group (Cox) starts; column j's nonzeros of
  // shard s start at colPtr[j] + colShard[j * (nShards + 1) + s]
  int nShards;
  const int32_t* shardRow;  // [nShards+1]
  const int32_t* colShard;  // [J * (nShards+1)]
  // column blocks: blk[j] = B > 1 starts a block of B consecutive short columns
  // whose rows (CLR: strata) are pairwise disjoint; 1 otherwise, 0 inside a
  // block; nullptr for Cox
  const int8_t* blk;        // [J]
};

constexpr int kShards = 64;      // row shards per data set (largest gang)
constexpr int kBlockCols = 3;    // columns of one block at most
constexpr int kBlockNnz = 512;   // nonzeros of a block column at most

template <typename R>
__device__ __forceinline__ R colValue(const DevX& d, int fmt, int64_t p);

template <>
__device__ __forceinline__ double colValue<double>(const DevX& d, int fmt, int64_t p) {
  return (fmt == F_SPARSE || fmt == F_DENSE) ? d.vals[p] : 1.0;
}
template <>
__device__ __forceinline__ float colValue<float>(const DevX& d, int fmt, int64_t p) {
  return (fmt == F_SPARSE || fmt == F_DENSE) ? d.valsF[p] : 1.0f;
}

__device__ __forceinline__ int colRow(const DevX& d, int fmt, int64_t p, int64_t start) {
  return (fmt == F_DENSE || fmt == F_INTERCEPT) ? (int)(p - start) : d.rows[p];
}

__device__ __forceinline__ bool isIndicator(int fmt) {
  return fmt == F_INDICATOR || fmt == F_INTERCEPT;
}

__device__ __forceinline__ double rexp(double x) { return exp(x); }
__device__ __forceinline__ float rexp(float x) { return expf(x); }

// ---------------------------------------------------------------- reductions
// Lane 63's value, to every lane (two 32-bit reads of the lane).
__device__ __forceinline__ double lastLane(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, kWave - 1);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), kWave - 1);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sum over the 64 lanes of a wave, returned to every lane; every lane must be
// active.  DPP row shifts by 1, 2, 4 and 8 leave each 16-lane row's inclusive
// sums in the row (lane 15 holds the row's total), row_bcast:15 adds row 0's
// total to row 1 and row 2's to row 3, row_bcast:31 adds lane 31 (rows 0-1) to
// rows 2-3, so lane 63 ends with the total; it is read back as two 32-bit
// halves.  A fixed order of additions, and register moves instead of LDS
// permutes.
__device__ __forceinline__ double waveSum(double v) {
  v += __builtin_amdgcn_update_dpp(0.0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0.0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0.0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0.0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0.0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0.0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return lastLane(v);
}

// p[0] + p[stride] + ... + p[(nw - 1) * stride] from LDS, added in that order
// (nw <= kMaxBlockThreads / kWave); the loads are unrolled so that they are all
// in flight before the first add.
__device__ __forceinline__ double sumWaves(const double* p, int stride, int nw) {
  constexpr int kMaxW = kMaxBlockThreads / kWave;
  double part[kMaxW];
#pragma unroll
  for (int w = 0; w < kMaxW; ++w) part[w] = w < nw ? p[w * stride] : 0.0;
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kMaxW; ++w)
    if (w < nw) t += part[w];
  return t;
}

// Sum NV values over the block; the totals are returned to every thread.
template <int NV>
__device__ __forceinline__ void blockSum(double (&v)[NV]) {
  __shared__ double sh[NV][kMaxBlockThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = waveSum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) sh[i][wid] = v[i];
  }
  __syncthreads();
  // the waves' totals in wave order, every load issued before the first add
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = sumWaves(&sh[i][0], 1, nw);
  __syncthreads();
}

// Grid-wide deterministic finish: every block stores its totals, the last block
// to arrive sums them in block order and writes out[].  Returns true in the
// last block (all its threads).  `ticket` must be zero on entry; it is reset.
template <int NV>
__device__ __forceinline__ bool gridFinish(double (&v)[NV], double* partials, unsigned* ticket,
                                           double* out) {
  __shared__ int amLast;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) partials[(size_t)blockIdx.x * NV + i] = v[i];
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    amLast = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!amLast) return false;
  __threadfence();
  double s[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) s[i] = 0.0;
  const volatile double* vp = partials;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] += vp[(size_t)b * NV + i];
  }
  blockSum<NV>(s);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) out[i] = s[i];
    *ticket = 0u;
  }
  return true;
}

// ---------------------------------------------------------- segmented scans
// A segmented-sum element: values plus "a segment starts inside".  The
// combination (a then b) = (a.f | b.f, b.f ? b.v : a.v + b.v) is associative.
template <int NV>
struct Seg {
  double v[NV];
  int f;
};

template <int NV>
__device__ __forceinline__ Seg<NV> segCombine(const Seg<NV>& a, const Seg<NV>& b) {
  Seg<NV> r;
  r.f = a.f | b.f;
#pragma unroll
  for (int i = 0; i < NV; ++i) r.v[i] = b.f ? b.v[i] : a.v[i] + b.v[i];
  return r;
}

template <int NV>
__device__ __forceinline__ Seg<NV> segShflUp(const Seg<NV>& a, int off) {
  Seg<NV> r;
#pragma unroll
  for (int i = 0; i < NV; ++i) r.v[i] = __shfl_up(a.v[i], off, kWave);
  r.f = __shfl_up(a.f, off, kWave);
  return r;
}

// Exclusive segmented scan of one element per thread, in thread order.  Returns
// the exclusive prefix for this thread; `total` receives the block aggregate.
template <int NV>
__device__ __forceinline__ Seg<NV> blockSegScanExcl(const Seg<NV>& x, Seg<NV>& total) {
  __shared__ Seg<NV> wsh[kMaxBlockThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = (blockDim.x + kWave - 1) / kWave;
  Seg<NV> inc = x;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    Seg<NV> o = segShflUp(inc, off);
    if (lane >= off) inc = segCombine(o, inc);
  }
  if (lane == kWave - 1) wsh[wid] = inc;
  __syncthreads();
  // carry into this wave: combination of all earlier waves' aggregates
  Seg<NV> carry;
  carry.f = 0;
#pragma unroll
  for (int i = 0; i < NV; ++i) carry.v[i] = 0.0;
  for (int w = 0; w < wid; ++w) carry = segCombine(carry, wsh[w]);
  Seg<NV> tot = carry;
  for (int w = wid; w < nw; ++w) tot = segCombine(tot, wsh[w]);
  total = tot;
  // exclusive within wave
  Seg<NV> ex = segShflUp(inc, 1);
  if (lane == 0) {
    ex.f = 0;
#pragma unroll
    for (int i = 0; i < NV; ++i) ex.v[i] = 0.0;
  }
  __syncthreads();
  return segCombine(carry, ex);
}

// ------------------------------------------------------------------- priors
// priors/CovariatePrior.h:186-188 (none), 240-273 (Laplace), 410-415 (Normal)
__host__ __device__ inline double priorDelta(int type, double var, double g, double h,
                                             double beta) {
  if (type == P_NONE) return -(g / h);
  if (type == P_NORMAL) return -(g + (beta / var)) / (h + (1.0 / var));
  const double lambda = sqrt(2.0 / var);
  const double neg = -(g - lambda) / h;
  const double pos = -(g + lambda) / h;
  const int s = beta == 0 ? 0 : (beta < 0 ? -1 : 1);
  double delta;
  if (s == 0) {
    delta = (neg < 0) ? neg : ((pos > 0) ? pos : 0.0);
  } else {
    delta = s < 0 ? neg : pos;
    const double nb = beta + delta;
    const int s2 = nb == 0 ? 0 : (nb < 0 ? -1 : 1);
    if (s2 != s) delta = -beta;
  }
  return delta;
}

// priors/CovariatePrior.h:221-225 (Laplace), 404-408 (Normal)
__host__ __device__ inline double priorLogDensity(int type, double var, double beta) {
  if (type == P_NONE) return 0.0;
  if (type == P_NORMAL)
    return -0.5 * log(2.0 * 3.14159265358979323846 * var) - 0.5 * beta * beta / var;
  const double lambda = sqrt(2.0 / var);
  return log(0.5 * lambda) - lambda * fabs(beta);
}

// CyclicCoordinateDescent::applyBounds (CyclicCoordinateDescent.cpp:1503-1520);
// max() keeps std::max's (a < b) ? b : a so NaN propagates the same way.
__host__ __device__ inline double applyBounds(double delta, double* bound) {
  const double b = *bound;
  if (delta < -b)
    delta = -b;
  else if (delta > b)
    delta = b;
  const double a2 = fabs(delta) * 2;
  double inter = (a2 < b / 2) ? b / 2 : a2;
  inter = (inter < 1E-3) ? 1E-3 : inter;
  *bound = inter;
  return delta;
}

}  // namespace ccdamd
