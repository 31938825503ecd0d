// ccd_ingest.hip -- long-format data ingestion on the device.
//
// Cyclops' R front end sorts the outcome rows and the covariate entries and its
// loaders map row ids to rows and covariate ids to columns on the host
// (R/NewDataConversion.R:95-235, ModelData.cpp:69-146 loadY, 148-243
// loadMultipleX, and for Poisson finalizeSqlCyclopsData(useOffsetCovariate = -1):
// RcppModelData.cpp:482-556, ModelData.cpp:609-652).  Here the same steps run as
// device radix sorts and scans:
//   1. the row order: stable sorts by row id, then (Cox) y, time descending,
//      then (CLR, Cox) stratum -- least significant key first;
//   2. stratum numbers by change points in that order (loadY's lastStratumMap);
//   3. the row map: sorted row ids -> row, one binary search per entry; entries
//      with unknown row ids are left out (checkRowIds);
//   4. the column map: unique sorted covariate ids, one binary search per entry;
//   5. one sort of the entries by (column, row), a check for repeated entries,
//      zero values left out, column starts by boundary detection, and a column
//      is SPARSE when a value other than 0 and 1 occurs (loadMultipleX's upcast).
// The compressed columns, outcomes and strata then pass through ccdh_data_create,
// which builds the per-column tables (runs, shards, blocks) as for any input.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ccd_host.h"

#define HIP_CHECK(expr)                                                                \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return setError(CCDH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace ccdamd {
namespace {

constexpr int kT = 256;

int blocksFor(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kT - 1) / kT, 65535)); }

template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)); }
};

__global__ void k_iota(int n, int32_t* a) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = i;
}

template <typename T>
__global__ void k_gather(int n, const T* src, const int32_t* idx, T* dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

// 1 where the sorted stratum id differs from the previous row's
__global__ void k_change_flags(int n, const int64_t* s, int32_t* f) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    f[i] = (i > 0 && s[i] != s[i - 1]) ? 1 : 0;
}

template <typename T>
__global__ void k_adjacent_equal(int n, const T* a, int* found) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x + 1; i < n; i += gridDim.x * blockDim.x)
    if (a[i] == a[i - 1]) atomicOr(found, 1);
}

// first index in a[0..n) whose value is >= v
__device__ inline int lowerBound(const int64_t* a, int n, int64_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// row of each entry (-1: row id not among the outcomes) and its validity flag
__global__ void k_entry_rows(int m, const int64_t* entryRowId, const int64_t* sortedIds,
                             const int32_t* rowOf, int n, int32_t* row, uint8_t* valid) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const int64_t id = entryRowId[e];
    const int p = lowerBound(sortedIds, n, id);
    const bool ok = p < n && sortedIds[p] == id;
    row[e] = ok ? rowOf[p] : -1;
    valid[e] = ok ? 1 : 0;
  }
}

// (column << 32 | row) per entry; entries without a row sort last
__global__ void k_entry_keys(int m, const int64_t* cov, const int32_t* row, const int64_t* uniq,
                             int nCols, const double* val, uint64_t* key, double* v) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const int r = row[e];
    key[e] = r < 0 ? ~0ull
                   : ((uint64_t)lowerBound(uniq, nCols, cov[e]) << 32) | (uint64_t)(uint32_t)r;
    v[e] = val ? val[e] : 1.0;
  }
}

__global__ void k_nonzero_flags(int m, const double* v, uint8_t* f) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
    f[e] = v[e] != 0.0 ? 1 : 0;
}

// colPtr[c] = first entry of column c (entries sorted by column); colPtr[nCols] = nnz
__global__ void k_col_bounds(int nnz, const uint64_t* key, int nCols, int64_t* colPtr) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
    const int c = (int)(key[i] >> 32);
    const int prev = i == 0 ? -1 : (int)(key[i - 1] >> 32);
    for (int cc = prev + 1; cc <= c; ++cc) colPtr[cc] = i;
    if (i == nnz - 1)
      for (int cc = c + 1; cc <= nCols; ++cc) colPtr[cc] = nnz;
  }
}

// rows and values of the sorted entries; a value other than 1 makes its column SPARSE
__global__ void k_split(int nnz, const uint64_t* key, const double* v, int32_t* rows, double* vals,
                        int* sparse) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += gridDim.x * blockDim.x) {
    rows[i] = (int32_t)(key[i] & 0xffffffffull);
    vals[i] = v[i];
    if (v[i] != 1.0) sparse[key[i] >> 32] = 1;
  }
}

}  // namespace

int dataCreateLong(const ccdh_long_desc* in, ccdh_data** out) {
  if (!in || !out) return setError(CCDH_E_INVALID, "null argument");
  const int model = in->model;
  if (model != CCDH_LOGISTIC && model != CCDH_POISSON && model != CCDH_CONDITIONAL_LOGISTIC &&
      model != CCDH_COX)
    return setError(CCDH_E_INVALID, "model");
  if (in->n_rows <= 0 || in->n_rows >= (1ll << 31) || in->n_entries < 0 ||
      in->n_entries >= (1ll << 31))
    return setError(CCDH_E_INVALID, "row or entry count out of range");
  if (!in->row_id || !in->y || (in->n_entries > 0 && (!in->covariate_id || !in->entry_row_id)))
    return setError(CCDH_E_INVALID, "outcome and entry arrays required");
  const int n = (int)in->n_rows, m = (int)in->n_entries;
  const bool cox = model == CCDH_COX, clr = model == CCDH_CONDITIONAL_LOGISTIC;
  const bool poisson = model == CCDH_POISSON;
  if (clr && !in->stratum_id) return setError(CCDH_E_INVALID, "stratum ids required");
  if ((cox || poisson) && !in->time) return setError(CCDH_E_INVALID, "time required");
  if (poisson)  // R/NewDataConversion.R:113-115
    for (int i = 0; i < n; ++i)
      if (!(in->time[i] > 0.0)) return setError(CCDH_E_INVALID, "time cannot be non-positive");
  if (cox) {  // R/NewDataConversion.R:154-156
    std::vector<double> seen;
    for (int i = 0; i < n && seen.size() <= 2; ++i)
      if (std::find(seen.begin(), seen.end(), in->y[i]) == seen.end()) seen.push_back(in->y[i]);
    if (seen.size() > 2) return setError(CCDH_E_INVALID, "Cox model only accepts one outcome type");
  }
  const bool strata = clr || (cox && in->stratum_id);

  hipStream_t s = nullptr;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{s};
  DevBuf<char> temp;
  size_t tempBytes = 0;
  // runs a hipcub call: size query, (re)allocation, then the call
  auto cub = [&](auto&& call) -> hipError_t {
    size_t need = 0;
    hipError_t e = call((void*)nullptr, need);
    if (e != hipSuccess) return e;
    if (need > tempBytes) {
      if (temp.p) (void)hipFree(temp.p);
      temp.p = nullptr;
      e = temp.alloc(need);
      if (e != hipSuccess) return e;
      tempBytes = need;
    }
    return call((void*)temp.p, tempBytes);
  };

  // ---- 1. row order
  DevBuf<int64_t> dRowId, dStratum, k64a, k64b;
  DevBuf<double> dY, dTime, kda, kdb;
  DevBuf<int32_t> perm, perm2;
  HIP_CHECK(dRowId.alloc(n));
  HIP_CHECK(dY.alloc(n));
  HIP_CHECK(k64a.alloc(n));
  HIP_CHECK(k64b.alloc(n));
  HIP_CHECK(kda.alloc(n));
  HIP_CHECK(kdb.alloc(n));
  HIP_CHECK(perm.alloc(n));
  HIP_CHECK(perm2.alloc(n));
  HIP_CHECK(hipMemcpyAsync(dRowId.p, in->row_id, (size_t)n * 8, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dY.p, in->y, (size_t)n * 8, hipMemcpyHostToDevice, s));
  if (in->time) {
    HIP_CHECK(dTime.alloc(n));
    HIP_CHECK(hipMemcpyAsync(dTime.p, in->time, (size_t)n * 8, hipMemcpyHostToDevice, s));
  }
  if (strata) {
    HIP_CHECK(dStratum.alloc(n));
    HIP_CHECK(hipMemcpyAsync(dStratum.p, in->stratum_id, (size_t)n * 8, hipMemcpyHostToDevice, s));
  }
  k_iota<<<blocksFor(n), kT, 0, s>>>(n, perm.p);
  HIP_CHECK(hipGetLastError());
  // one stable pass: keys gathered through the current order, sorted with it
  auto passI64 = [&](const int64_t* src) -> hipError_t {
    k_gather<int64_t><<<blocksFor(n), kT, 0, s>>>(n, src, perm.p, k64a.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = cub([&](void* t, size_t& b) {
      return hipcub::DeviceRadixSort::SortPairs(t, b, k64a.p, k64b.p, perm.p, perm2.p, n, 0, 64, s);
    });
    std::swap(perm.p, perm2.p);
    return e;
  };
  auto passF64 = [&](const double* src, bool descending) -> hipError_t {
    k_gather<double><<<blocksFor(n), kT, 0, s>>>(n, src, perm.p, kda.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = cub([&](void* t, size_t& b) {
      return descending ? hipcub::DeviceRadixSort::SortPairsDescending(t, b, kda.p, kdb.p, perm.p,
                                                                       perm2.p, n, 0, 64, s)
                        : hipcub::DeviceRadixSort::SortPairs(t, b, kda.p, kdb.p, perm.p, perm2.p,
                                                             n, 0, 64, s);
    });
    std::swap(perm.p, perm2.p);
    return e;
  };
  HIP_CHECK(passI64(dRowId.p));
  if (cox) {
    HIP_CHECK(passF64(dY.p, false));
    HIP_CHECK(passF64(dTime.p, true));
  }
  if (strata) HIP_CHECK(passI64(dStratum.p));

  // ---- 2. outcomes in row order, stratum numbers
  std::vector<double> hY(n), hTime;
  std::vector<int32_t> hStrata;
  k_gather<double><<<blocksFor(n), kT, 0, s>>>(n, dY.p, perm.p, kda.p);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(hY.data(), kda.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  if (in->time && (cox || poisson)) {
    hTime.resize(n);
    k_gather<double><<<blocksFor(n), kT, 0, s>>>(n, dTime.p, perm.p, kdb.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(hTime.data(), kdb.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  }
  if (strata) {
    DevBuf<int32_t> flags, pid;
    HIP_CHECK(flags.alloc(n));
    HIP_CHECK(pid.alloc(n));
    k_gather<int64_t><<<blocksFor(n), kT, 0, s>>>(n, dStratum.p, perm.p, k64a.p);
    k_change_flags<<<blocksFor(n), kT, 0, s>>>(n, k64a.p, flags.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(cub([&](void* t, size_t& b) {
      return hipcub::DeviceScan::InclusiveSum(t, b, flags.p, pid.p, n, s);
    }));
    hStrata.resize(n);
    HIP_CHECK(hipMemcpyAsync(hStrata.data(), pid.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }

  // ---- 3. row map: row ids in row order, sorted, with their rows
  DevBuf<int32_t> rowOf;
  DevBuf<int> found;
  HIP_CHECK(rowOf.alloc(n));
  HIP_CHECK(found.alloc(1));
  k_gather<int64_t><<<blocksFor(n), kT, 0, s>>>(n, dRowId.p, perm.p, k64a.p);
  k_iota<<<blocksFor(n), kT, 0, s>>>(n, perm2.p);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(cub([&](void* t, size_t& b) {
    return hipcub::DeviceRadixSort::SortPairs(t, b, k64a.p, k64b.p, perm2.p, rowOf.p, n, 0, 64, s);
  }));
  HIP_CHECK(hipMemsetAsync(found.p, 0, sizeof(int), s));
  k_adjacent_equal<int64_t><<<blocksFor(n), kT, 0, s>>>(n, k64b.p, found.p);
  HIP_CHECK(hipGetLastError());
  int dup = 0;
  HIP_CHECK(hipMemcpyAsync(&dup, found.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (dup) return setError(CCDH_E_INVALID, "repeated outcome row id");

  // ---- 4./5. entries
  std::vector<int64_t> hColPtr(1, 0);
  std::vector<int32_t> hRows;
  std::vector<double> hVals;
  std::vector<int> hSparse;
  int nCols = 0;
  if (m > 0) {
    DevBuf<int64_t> dCov, dEntryRow, covValid, covSorted, uniq;
    DevBuf<double> dVal, v, v2;
    DevBuf<int32_t> row;
    DevBuf<uint8_t> valid;
    DevBuf<uint64_t> key, key2;
    DevBuf<int> nSel;
    HIP_CHECK(dCov.alloc(m));
    HIP_CHECK(dEntryRow.alloc(m));
    HIP_CHECK(row.alloc(m));
    HIP_CHECK(valid.alloc(m));
    HIP_CHECK(nSel.alloc(1));
    HIP_CHECK(hipMemcpyAsync(dCov.p, in->covariate_id, (size_t)m * 8, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dEntryRow.p, in->entry_row_id, (size_t)m * 8, hipMemcpyHostToDevice, s));
    if (in->value) {
      HIP_CHECK(dVal.alloc(m));
      HIP_CHECK(hipMemcpyAsync(dVal.p, in->value, (size_t)m * 8, hipMemcpyHostToDevice, s));
    }
    k_entry_rows<<<blocksFor(m), kT, 0, s>>>(m, dEntryRow.p, k64b.p, rowOf.p, n, row.p, valid.p);
    HIP_CHECK(hipGetLastError());
    // column ids present among entries with a row: sorted, unique
    HIP_CHECK(covValid.alloc(m));
    HIP_CHECK(covSorted.alloc(m));
    HIP_CHECK(uniq.alloc(m));
    HIP_CHECK(cub([&](void* t, size_t& b) {
      return hipcub::DeviceSelect::Flagged(t, b, dCov.p, valid.p, covValid.p, nSel.p, m, s);
    }));
    int nValid = 0;
    HIP_CHECK(hipMemcpyAsync(&nValid, nSel.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (nValid > 0) {
      HIP_CHECK(cub([&](void* t, size_t& b) {
        return hipcub::DeviceRadixSort::SortKeys(t, b, covValid.p, covSorted.p, nValid, 0, 64, s);
      }));
      HIP_CHECK(cub([&](void* t, size_t& b) {
        return hipcub::DeviceSelect::Unique(t, b, covSorted.p, uniq.p, nSel.p, nValid, s);
      }));
      HIP_CHECK(hipMemcpyAsync(&nCols, nSel.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      // entries sorted by (column, row); those without a row last
      HIP_CHECK(key.alloc(m));
      HIP_CHECK(key2.alloc(m));
      HIP_CHECK(v.alloc(m));
      HIP_CHECK(v2.alloc(m));
      k_entry_keys<<<blocksFor(m), kT, 0, s>>>(m, dCov.p, row.p, uniq.p, nCols, dVal.p, key.p, v.p);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(cub([&](void* t, size_t& b) {
        return hipcub::DeviceRadixSort::SortPairs(t, b, key.p, key2.p, v.p, v2.p, m, 0, 64, s);
      }));
      // loadMultipleX: a repeated (row, column) entry is an error, zero values included
      HIP_CHECK(hipMemsetAsync(found.p, 0, sizeof(int), s));
      k_adjacent_equal<uint64_t><<<blocksFor(nValid), kT, 0, s>>>(nValid, key2.p, found.p);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(&dup, found.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (dup) return setError(CCDH_E_INVALID, "Repeated row-column entry");
      // zero values are not stored
      k_nonzero_flags<<<blocksFor(nValid), kT, 0, s>>>(nValid, v2.p, valid.p);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(cub([&](void* t, size_t& b) {
        return hipcub::DeviceSelect::Flagged(t, b, key2.p, valid.p, key.p, nSel.p, nValid, s);
      }));
      HIP_CHECK(cub([&](void* t, size_t& b) {
        return hipcub::DeviceSelect::Flagged(t, b, v2.p, valid.p, v.p, nSel.p, nValid, s);
      }));
      int nnz = 0;
      HIP_CHECK(hipMemcpyAsync(&nnz, nSel.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      DevBuf<int64_t> colPtr;
      DevBuf<int32_t> rows;
      DevBuf<double> vals;
      DevBuf<int> sparse;
      HIP_CHECK(colPtr.alloc(nCols + 1));
      HIP_CHECK(rows.alloc(nnz));
      HIP_CHECK(vals.alloc(nnz));
      HIP_CHECK(sparse.alloc(nCols));
      HIP_CHECK(hipMemsetAsync(colPtr.p, 0, (size_t)(nCols + 1) * 8, s));
      HIP_CHECK(hipMemsetAsync(sparse.p, 0, (size_t)nCols * sizeof(int), s));
      if (nnz > 0) {
        k_col_bounds<<<blocksFor(nnz), kT, 0, s>>>(nnz, key.p, nCols, colPtr.p);
        k_split<<<blocksFor(nnz), kT, 0, s>>>(nnz, key.p, v.p, rows.p, vals.p, sparse.p);
        HIP_CHECK(hipGetLastError());
      }
      hColPtr.resize(nCols + 1);
      hRows.resize(nnz);
      hSparse.resize(nCols);
      HIP_CHECK(hipMemcpyAsync(hColPtr.data(), colPtr.p, (size_t)(nCols + 1) * 8,
                               hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipMemcpyAsync(hSparse.data(), sparse.p, (size_t)nCols * sizeof(int),
                               hipMemcpyDeviceToHost, s));
      if (nnz > 0)
        HIP_CHECK(hipMemcpyAsync(hRows.data(), rows.p, (size_t)nnz * 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      // values only when a column keeps them (an all-INDICATOR design needs none)
      if (nnz > 0 && std::find(hSparse.begin(), hSparse.end(), 1) != hSparse.end()) {
        hVals.resize(nnz);
        HIP_CHECK(hipMemcpyAsync(hVals.data(), vals.p, (size_t)nnz * 8, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
      }
    }
  }

  // ---- the data set: [log time offset (Poisson)] [intercept (LR, Poisson)] covariates
  const bool intercept = in->add_intercept && (model == CCDH_LOGISTIC || poisson);
  const int lead = (poisson ? 1 : 0) + (intercept ? 1 : 0);
  const int J = lead + nCols;
  const int64_t nnzCov = (int64_t)hRows.size();
  const bool withVals = poisson || !hVals.empty();  // a DENSE offset or a SPARSE column
  std::vector<int64_t> colPtr(J + 1, 0);
  std::vector<int32_t> rows((size_t)lead * n + nnzCov);
  std::vector<double> vals(withVals ? rows.size() : 0);
  std::vector<int8_t> fmt(J);
  int64_t at = 0;
  for (int c = 0; c < lead; ++c) {
    const bool offset = poisson && c == 0;
    fmt[c] = offset ? CCDH_DENSE : CCDH_INTERCEPT;
    for (int k = 0; k < n; ++k) {
      rows[at + k] = k;
      if (withVals) vals[at + k] = offset ? std::log(hTime[k]) : 1.0;
    }
    at += n;
    colPtr[c + 1] = at;
  }
  if (nnzCov > 0) {
    std::memcpy(rows.data() + at, hRows.data(), nnzCov * sizeof(int32_t));
    if (!hVals.empty())
      std::memcpy(vals.data() + at, hVals.data(), nnzCov * sizeof(double));
    else if (withVals)
      std::fill(vals.begin() + at, vals.end(), 1.0);
  }
  for (int c = 0; c < nCols; ++c) {
    fmt[lead + c] = hSparse[c] ? CCDH_SPARSE : CCDH_INDICATOR;
    colPtr[lead + c + 1] = at + hColPtr[c + 1];
  }
  ccdh_data_desc desc{};
  desc.model = model;
  desc.precision = in->precision;
  desc.K = n;
  desc.J = J;
  desc.col_ptr = colPtr.data();
  desc.rows = rows.data();
  desc.vals = withVals ? vals.data() : nullptr;
  desc.fmt = fmt.data();
  desc.y = hY.data();
  desc.offs = cox ? hTime.data() : nullptr;
  desc.strata = strata ? hStrata.data() : nullptr;
  desc.has_offset = poisson ? 1 : 0;
  return dataCreate(&desc, out);
}

}  // namespace ccdamd

extern "C" int ccdh_data_create_long(const ccdh_long_desc* desc, ccdh_data_t* out) {
  return ccdamd::dataCreateLong(desc, out);
}
