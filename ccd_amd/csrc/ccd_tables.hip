// ccd_tables.hip -- a data set's derived per-column tables, built on the device
// from its resident compressed columns (the tables ccdh_data_create builds on the
// host for a caller's arrays):
//   * conditional logistic: the stratum of every nonzero (nzKey), the runs of
//     equal stratum inside every column (segPtr / segOff / segKey: a flag at each
//     run head, one scan, one compaction) and each stratum's first row
//     (stratStart), as the reference's per-stratum sums walk them
//     (ModelSpecifics.hpp:891-1020 over the sorted rows);
//   * row shards for workgroup gangs: 64 near-equal row ranges whose boundaries
//     start a stratum (CLR) or a tie group of events (Cox), found by one
//     workgroup per boundary, and per column the offset of each shard's first
//     nonzero (one binary search per column and shard);
//   * column blocks (not Cox): for every column the length of the block that
//     would start there (one wave per column stages the units -- rows, or CLR
//     strata -- of up to three short columns in LDS and binary-searches the
//     next column's units in them), then the greedy walk from column 0 as a
//     prefix composition of per-column maps on "distance to the next block
//     start" (one workgroup);
//   * the FP32 copy of the values on the mixed path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>
#include <vector>

#include "ccd_host.h"
#include "ccd_prim.h"

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return setError(CCDH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

namespace ccdamd {
namespace {

constexpr int kT = 256;

int gridOf(int64_t n) { return prim::gridFor(n); }

__device__ inline bool denseFmt(int f) { return f == CCDH_DENSE || f == CCDH_INTERCEPT; }

// the column holding nonzero p: the last j with colPtr[j] <= p
__device__ inline int columnOf(const int64_t* colPtr, int J, int64_t p) {
  int lo = 0, hi = J;  // colPtr[lo] <= p < colPtr[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (colPtr[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// CLR: stratum of each nonzero, its offset in its column, and a flag at each run
// of equal stratum inside a column; flag[nnz] = 0 closes the scan
__global__ void k_nz_strata(int J, int64_t nnz, const int64_t* colPtr, const int32_t* rows,
                            const int8_t* fmt, const int32_t* strata, int32_t* nzKey,
                            int32_t* off, uint8_t* flag) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p <= nnz;
       p += (int64_t)gridDim.x * blockDim.x) {
    if (p == nnz) {
      flag[p] = 0;
      continue;
    }
    const int j = columnOf(colPtr, J, p);
    const int64_t b = colPtr[j];
    const bool dense = denseFmt(fmt[j]);
    const int k = dense ? (int)(p - b) : rows[p];
    const int s = strata[k];
    nzKey[p] = s;
    off[p] = (int32_t)(p - b);
    bool head = p == b;
    if (!head) {
      const int kp = dense ? (int)(p - 1 - b) : rows[p - 1];
      head = strata[kp] != s;
    }
    flag[p] = head ? 1 : 0;
  }
}

// segPtr[j] = runs before column j (pos is the exclusive count of run heads)
__global__ void k_seg_ptr(int J, const int64_t* colPtr, const int32_t* pos, int64_t* segPtr) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j <= J; j += gridDim.x * blockDim.x)
    segPtr[j] = pos[colPtr[j]];
}

// stratStart[s] = first row of stratum s (rows sorted by stratum); [nStrata] = K
__global__ void k_strat_start(int K, int nStrata, const int32_t* strata, int32_t* ss) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < K; k += gridDim.x * blockDim.x) {
    const int prev = k == 0 ? -1 : strata[k - 1];
    for (int s = prev + 1; s <= strata[k]; ++s) ss[s] = k;
    if (k == K - 1)
      for (int s = strata[k] + 1; s <= nStrata; ++s) ss[s] = K;
  }
}

// does row k start a group that a shard boundary may not split
__device__ inline bool startsGroup(int model, int K, const int32_t* strata, const double* y,
                                   const double* time, int64_t k) {
  if (k <= 0 || k >= K) return true;
  if (model == CCDH_CONDITIONAL_LOGISTIC) return strata[k] != strata[k - 1];
  if (model == CCDH_COX)
    return strata[k] != strata[k - 1] ||
           !(y[k] == 1.0 && y[k - 1] == 1.0 && time[k] == time[k - 1]);
  return true;
}

// one workgroup per shard boundary s (1..kShards-1): the first group start at or
// after row K*s/kShards
__global__ void k_shard_starts(int model, int K, const int32_t* strata, const double* y,
                               const double* time, int32_t* g) {
  __shared__ int found;
  const int s = blockIdx.x + 1;
  if (threadIdx.x == 0) found = K;
  __syncthreads();
  const int64_t start = (int64_t)K * s / kShards;
  for (int64_t base = start; base < K; base += blockDim.x) {
    const int64_t k = base + threadIdx.x;
    const bool st = k < K && startsGroup(model, K, strata, y, time, k);
    if (st) atomicMin(&found, (int)k);
    if (__syncthreads_or(st)) break;
  }
  __syncthreads();
  if (threadIdx.x == 0) g[s] = found;
}

// shardRow = running maximum of the boundaries (a long group pushes the next ones)
__global__ void k_shard_rows(int K, const int32_t* g, int32_t* shardRow) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int32_t prev = 0;
  shardRow[0] = 0;
  for (int s = 1; s < kShards; ++s) {
    prev = max(prev, g[s]);
    shardRow[s] = prev;
  }
  shardRow[kShards] = K;
}

// colShard[j * (kShards + 1) + s] = nonzeros of column j in rows below shardRow[s]
__global__ void k_col_shards(int J, const int64_t* colPtr, const int32_t* rows, const int8_t* fmt,
                             const int32_t* shardRow, int32_t* colShard) {
  const int64_t total = (int64_t)J * (kShards + 1);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(i / (kShards + 1)), s = (int)(i % (kShards + 1));
    const int32_t r = shardRow[s];
    if (denseFmt(fmt[j])) {
      colShard[i] = r;
      continue;
    }
    const int64_t b = colPtr[j];
    int64_t lo = b, hi = colPtr[j + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (rows[mid] < r)
        lo = mid + 1;
      else
        hi = mid;
    }
    colShard[i] = (int32_t)(lo - b);
  }
}

// Length of the block that would start at column j: columns j, j+1, ... join
// while short (1..kBlockNnz nonzeros, not dense), up to kBlockCols, and while
// none of their units (rows; CLR: strata) occurs in an earlier column of the
// block.  Units of a column are sorted (rows ascending, rows sorted by stratum),
// so the earlier columns' units are staged in LDS and searched.
constexpr int kLenWaves = 4;
__global__ __launch_bounds__(kLenWaves * 64) void k_block_len(int J, const int64_t* colPtr,
                                                              const int32_t* rows,
                                                              const int8_t* fmt,
                                                              const int32_t* strata,
                                                              int8_t* len) {
  __shared__ int32_t units[kLenWaves][kBlockCols - 1][kBlockNnz];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = blockIdx.x * kLenWaves + wave;
  if (j >= J) return;  // whole waves leave; no workgroup barrier below
  int cnt[kBlockCols - 1] = {};
  int B = 0;
  for (int c = j; c < J && B < kBlockCols; ++c) {
    const int64_t b = colPtr[c], n = colPtr[c + 1] - b;
    if (denseFmt(fmt[c]) || n <= 0 || n > kBlockNnz) break;
    bool clash = false;
    for (int p = lane; p < n; p += 64) {
      const int u = strata ? strata[rows[b + p]] : rows[b + p];
      for (int q = 0; q < B && q < kBlockCols - 1; ++q) {
        int lo = 0, hi = cnt[q];
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (units[wave][q][mid] < u)
            lo = mid + 1;
          else
            hi = mid;
        }
        clash = clash || (lo < cnt[q] && units[wave][q][lo] == u);
      }
    }
    if (__ballot(clash) != 0ull) break;
    if (B < kBlockCols - 1) {
      for (int p = lane; p < n; p += 64)
        units[wave][B][p] = strata ? strata[rows[b + p]] : rows[b + p];
      cnt[B] = (int)n;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    ++B;
  }
  if (lane == 0) len[j] = (int8_t)B;
}

// The greedy walk over the block lengths.  State = distance from column j to the
// next block start (0..kBlockCols-1); column j maps it as d -> step(j) - 1 when
// d == 0 (step = len >= 2 ? len : 1) and d -> d - 1 otherwise.  Each thread
// composes the maps of its chunk of columns (all four entry states at once), one
// thread chains the chunks from state 0 at column 0, and every thread replays
// its chunk writing blk (block start: its length, or 1 when alone; inside a
// block: 0).
__global__ __launch_bounds__(kT) void k_block_walk(int J, const int8_t* len, int8_t* blk) {
  __shared__ uint8_t fn[kT][kBlockCols];
  __shared__ uint8_t entry[kT];
  const int t = threadIdx.x;
  const int chunk = (J + kT - 1) / kT;
  const int j0 = min(J, t * chunk), j1 = min(J, j0 + chunk);
  uint8_t d[kBlockCols];
  for (int e = 0; e < kBlockCols; ++e) d[e] = (uint8_t)e;
  for (int j = j0; j < j1; ++j) {
    const int L = len[j];
    const int step = L >= 2 ? L : 1;
    for (int e = 0; e < kBlockCols; ++e) d[e] = d[e] == 0 ? (uint8_t)(step - 1) : (uint8_t)(d[e] - 1);
  }
  for (int e = 0; e < kBlockCols; ++e) fn[t][e] = d[e];
  __syncthreads();
  if (t == 0) {
    uint8_t st = 0;
    for (int u = 0; u < kT; ++u) {
      entry[u] = st;
      st = fn[u][st];
    }
  }
  __syncthreads();
  uint8_t st = entry[t];
  for (int j = j0; j < j1; ++j) {
    const int L = len[j];
    const int step = L >= 2 ? L : 1;
    if (st == 0) {
      blk[j] = (int8_t)(L >= 2 ? L : 1);
      st = (uint8_t)(step - 1);
    } else {
      blk[j] = 0;
      st = (uint8_t)(st - 1);
    }
  }
}

__global__ void k_to_f32(int64_t n, const double* a, float* b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

template <typename T>
int dalloc(ccdh_data* d, T** p, int64_t n) {
  *p = nullptr;
  hipError_t e = hipMalloc((void**)p, (size_t)std::max<int64_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return setError(CCDH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  d->devBytes += (int64_t)std::max<int64_t>(n, 1) * (int64_t)sizeof(T);
  return 0;
}

}  // namespace

int buildTablesDevice(ccdh_data* d, hipStream_t s) {
  const int K = d->K, J = d->J;
  const int64_t nnz = d->nnz;
  int rc = 0;
  if (d->precision == CCDH_MIXED && d->vals) {
    if ((rc = dalloc(d, &d->valsF, nnz))) return rc;
    if (nnz > 0) {
      k_to_f32<<<gridOf(nnz), kT, 0, s>>>(nnz, d->vals, d->valsF);
      HIP_TRY(hipGetLastError());
    }
  }
  if (d->model == CCDH_CONDITIONAL_LOGISTIC) {
    if ((rc = dalloc(d, &d->nzKey, nnz))) return rc;
    prim::Scratch tmp;
    HIP_TRY(tmp.need((size_t)nnz * 4 + (size_t)(nnz + 1)));
    int32_t* off = (int32_t*)tmp.p;
    uint8_t* flag = (uint8_t*)(off + nnz);
    k_nz_strata<<<gridOf(nnz + 1), kT, 0, s>>>(J, nnz, d->colPtr, d->rows, d->fmt, d->strata,
                                                d->nzKey, off, flag);
    HIP_TRY(hipGetLastError());
    prim::Compactor cmp;
    int nSeg = 0;
    HIP_TRY(cmp.plan(flag, nnz + 1, s, &nSeg));
    d->nSeg = nSeg;
    if ((rc = dalloc(d, &d->segOff, nSeg))) return rc;
    if ((rc = dalloc(d, &d->segKey, nSeg))) return rc;
    if ((rc = dalloc(d, &d->segPtr, (int64_t)J + 1))) return rc;
    HIP_TRY(cmp.apply(off, flag, nnz, d->segOff, s));
    HIP_TRY(cmp.apply(d->nzKey, flag, nnz, d->segKey, s));
    k_seg_ptr<<<gridOf(J + 1), kT, 0, s>>>(J, d->colPtr, cmp.pos, d->segPtr);
    HIP_TRY(hipGetLastError());
    if ((rc = dalloc(d, &d->stratStart, (int64_t)d->nStrata + 1))) return rc;
    k_strat_start<<<gridOf(K), kT, 0, s>>>(K, d->nStrata, d->strata, d->stratStart);
    HIP_TRY(hipGetLastError());
    d->hSegPtr.resize((size_t)J + 1);
    HIP_TRY(hipMemcpyAsync(d->hSegPtr.data(), d->segPtr, (size_t)(J + 1) * 8,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  {
    if ((rc = dalloc(d, &d->shardRow, kShards + 1))) return rc;
    if ((rc = dalloc(d, &d->colShard, (int64_t)J * (kShards + 1)))) return rc;
    prim::Scratch tmp;
    HIP_TRY(tmp.need((kShards + 1) * sizeof(int32_t)));
    int32_t* g = (int32_t*)tmp.p;
    k_shard_starts<<<kShards - 1, kT, 0, s>>>(d->model, K, d->strata, d->y, d->time, g);
    k_shard_rows<<<1, 64, 0, s>>>(K, g, d->shardRow);
    k_col_shards<<<gridOf((int64_t)J * (kShards + 1)), kT, 0, s>>>(J, d->colPtr, d->rows, d->fmt,
                                                                 d->shardRow, d->colShard);
    HIP_TRY(hipGetLastError());
    d->hShardRow.resize(kShards + 1);
    HIP_TRY(hipMemcpyAsync(d->hShardRow.data(), d->shardRow, (kShards + 1) * sizeof(int32_t),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  if (d->model != CCDH_COX) {
    if ((rc = dalloc(d, &d->blk, J))) return rc;
    prim::Scratch tmp;
    HIP_TRY(tmp.need((size_t)J));
    int8_t* len = (int8_t*)tmp.p;
    k_block_len<<<(J + kLenWaves - 1) / kLenWaves, kLenWaves * 64, 0, s>>>(
        J, d->colPtr, d->rows, d->fmt,
        d->model == CCDH_CONDITIONAL_LOGISTIC ? d->strata : nullptr, len);
    k_block_walk<<<1, kT, 0, s>>>(J, len, d->blk);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
  }
  return 0;
}

}  // namespace ccdamd

// Introspection of the derived tables (no reference counterpart): sizes first
// (any array may be NULL), then the arrays.
extern "C" int ccdh_data_get_tables(ccdh_data_t data, int64_t* n_seg, int* n_strata,
                                    int64_t* seg_ptr, int32_t* seg_off, int32_t* seg_key,
                                    int32_t* nz_key, int32_t* strat_start, int32_t* shard_row,
                                    int32_t* col_shard, int8_t* blk) {
  using namespace ccdamd;
  if (!data) return setError(CCDH_E_INVALID, "null handle");
  if (n_seg) *n_seg = data->nSeg;
  if (n_strata) *n_strata = data->nStrata;
  const int J = data->J;
  const bool clr = data->model == CCDH_CONDITIONAL_LOGISTIC;
  auto get = [&](void* dst, const void* src, size_t bytes) -> int {
    if (!dst || !src || bytes == 0) return 0;
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
  };
  int rc = 0;
  if (clr) {
    rc = rc ? rc : get(seg_ptr, data->segPtr, (size_t)(J + 1) * 8);
    rc = rc ? rc : get(seg_off, data->segOff, (size_t)data->nSeg * 4);
    rc = rc ? rc : get(seg_key, data->segKey, (size_t)data->nSeg * 4);
    rc = rc ? rc : get(nz_key, data->nzKey, (size_t)data->nnz * 4);
    rc = rc ? rc : get(strat_start, data->stratStart, (size_t)(data->nStrata + 1) * 4);
  }
  rc = rc ? rc : get(shard_row, data->shardRow, (kShards + 1) * 4);
  rc = rc ? rc : get(col_shard, data->colShard, (size_t)J * (kShards + 1) * 4);
  if (data->model != CCDH_COX) rc = rc ? rc : get(blk, data->blk, (size_t)J);
  return rc;
}
