// ccd_ingest.hip -- long-format data ingestion on the device.
//
// Cyclops' R front end sorts the outcome rows and the covariate entries and its
// loaders map row ids to rows and covariate ids to columns on the host
// (R/NewDataConversion.R:95-235, ModelData.cpp:69-146 loadY, 148-243
// loadMultipleX, and for Poisson finalizeSqlCyclopsData(useOffsetCovariate = -1):
// RcppModelData.cpp:482-556, ModelData.cpp:609-652).  Here the same steps run as
// device radix sorts, scans and compactions (ccd_prim.h):
//   1. the row order: stable sorts by row id, then (Cox) y, time descending,
//      then (CLR, Cox) stratum -- least significant key first;
//   2. stratum numbers by change points in that order (loadY's lastStratumMap);
//   3. the row map: sorted row ids -> row, one binary search per entry; entries
//      with unknown row ids are left out (checkRowIds);
//   4. the column map: unique sorted covariate ids, one binary search per entry;
//   5. one sort of the entries by (column, row), a check for repeated entries,
//      zero values left out, column starts by boundary detection, and a column
//      is SPARSE when a value other than 0 and 1 occurs (loadMultipleX's upcast).
// The data set's arrays are assembled on the device (the LR / Poisson intercept
// and the Poisson log-time offset column first) and its per-column tables are
// built there (ccd_tables.hip); only the O(rows + columns) outcome, stratum and
// column-start vectors come back to the host, which keeps copies of them.  The
// offset column's log(time) and the Poisson -log(y!) terms are taken on the
// host, as the reference takes them, so an ingested data set is bit-identical to
// one built from the same columns by ccdh_data_create.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "ccd_host.h"
#include "ccd_prim.h"

#define HIP_CHECK(expr)                                                                \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return setError(CCDH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace ccdamd {
namespace {

constexpr int kT = 256;

int blocksFor(int64_t n) { return prim::gridFor(n); }

template <typename T>
struct DevBuf {
  T* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)); }
};

template <typename T>
void swapBuf(DevBuf<T>& a, DevBuf<T>& b) {
  std::swap(a.p, b.p);
}

__global__ void k_iota(int n, int32_t* a) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = i;
}

template <typename T>
__global__ void k_gather(int n, const T* src, const int32_t* idx, T* dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

// sort keys of a pass: the source gathered through the current order
__global__ void k_key_i64(int64_t n, const int64_t* src, const int32_t* idx, uint64_t* key) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    key[i] = prim::keyOfI64(src[idx ? idx[i] : i]);
}
__global__ void k_key_f64(int n, const double* src, const int32_t* idx, bool descending,
                          uint64_t* key) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    key[i] = prim::keyOfF64(src[idx[i]], descending);
}

// 1 where the sorted stratum id differs from the previous row's
__global__ void k_change_flags(int n, const uint64_t* s, uint8_t* f) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    f[i] = (i > 0 && s[i] != s[i - 1]) ? 1 : 0;
}

// stratum number = change points at or before the row
__global__ void k_stratum_numbers(int n, const uint8_t* f, const int32_t* pos, int32_t* strata) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    strata[i] = pos[i] + f[i];
}

template <typename T>
__global__ void k_adjacent_equal(int64_t n, const T* a, int* found) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (a[i] == a[i - 1]) atomicOr(found, 1);
}

// 1 at the first of each run of equal keys
__global__ void k_unique_flags(int64_t n, const uint64_t* a, uint8_t* f) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    f[i] = (i == 0 || a[i] != a[i - 1]) ? 1 : 0;
}

// first index in a[0..n) whose value is >= v
__device__ inline int lowerBound(const uint64_t* a, int n, uint64_t v) {
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

// row of each entry (-1: row id not among the outcomes), its validity flag and
// its covariate id as a sort key
__global__ void k_entry_rows(int m, const int64_t* entryRowId, const uint64_t* sortedIds,
                             const int32_t* rowOf, int n, const int64_t* cov, int32_t* row,
                             uint8_t* valid, uint64_t* covKey) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const uint64_t id = prim::keyOfI64(entryRowId[e]);
    const int p = lowerBound(sortedIds, n, id);
    const bool ok = p < n && sortedIds[p] == id;
    row[e] = ok ? rowOf[p] : -1;
    valid[e] = ok ? 1 : 0;
    covKey[e] = prim::keyOfI64(cov[e]);
  }
}

// (column << 32 | row) per entry; entries without a row sort last
__global__ void k_entry_keys(int m, const uint64_t* covKey, const int32_t* row,
                             const uint64_t* uniq, int nCols, const double* val, uint64_t* key,
                             double* v) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const int r = row[e];
    key[e] = r < 0 ? ~0ull
                   : ((uint64_t)lowerBound(uniq, nCols, covKey[e]) << 32) | (uint64_t)(uint32_t)r;
    v[e] = val ? val[e] : 1.0;
  }
}

__global__ void k_nonzero_flags(int m, const double* v, uint8_t* f) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
    f[e] = v[e] != 0.0 ? 1 : 0;
}

// colPtr[c] = c * n for the lead columns; colPtr[lead + c] = lead * n + first
// entry of covariate column c (entries sorted by column); colPtr[J] = nonzeros
__global__ void k_col_bounds(int64_t nnz, const uint64_t* key, int nCols, int lead, int n,
                             int64_t* colPtr) {
  const int64_t base = (int64_t)lead * n;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t c = tid; c < lead; c += stride) colPtr[c] = c * n;
  if (nnz == 0) {
    for (int64_t c = tid; c <= nCols; c += stride) colPtr[lead + c] = base;
    return;
  }
  for (int64_t i = tid; i < nnz; i += stride) {
    const int c = (int)(key[i] >> 32);
    const int prev = i == 0 ? -1 : (int)(key[i - 1] >> 32);
    for (int cc = prev + 1; cc <= c; ++cc) colPtr[lead + cc] = base + i;
    if (i == nnz - 1)
      for (int cc = c + 1; cc <= nCols; ++cc) colPtr[lead + cc] = base + nnz;
  }
}

// rows of the lead columns (their positions) and of the covariate entries; a
// value other than 1 makes a column SPARSE
__global__ void k_rows_and_sparse(int64_t nnz, const uint64_t* key, const double* v, int lead,
                                  int n, int32_t* rows, int* sparse) {
  const int64_t leadN = (int64_t)lead * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < leadN + nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < leadN) {
      rows[i] = (int32_t)(i % n);
      continue;
    }
    const int64_t e = i - leadN;
    rows[i] = (int32_t)(key[e] & 0xffffffffull);
    if (v[e] != 1.0) sparse[key[e] >> 32] = 1;
  }
}

__global__ void k_formats(int J, int lead, bool offset, const int* sparse, int8_t* fmt) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < J; c += gridDim.x * blockDim.x)
    fmt[c] = c < lead ? (int8_t)(offset && c == 0 ? CCDH_DENSE : CCDH_INTERCEPT)
                      : (int8_t)(sparse[c - lead] ? CCDH_SPARSE : CCDH_INDICATOR);
}

// values: [offset column from the host] [intercept 1.0] [entry values]
__global__ void k_values(int64_t nnz, const double* v, int lead, bool offset, int n, double* vals) {
  const int64_t leadN = (int64_t)lead * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < leadN + nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < leadN) {
      if (!(offset && i < n)) vals[i] = 1.0;
      continue;
    }
    vals[i] = v[i - leadN];
  }
}

}  // namespace

int dataCreateLong(const ccdh_long_desc* in, ccdh_data** out) {
  if (!in || !out) return setError(CCDH_E_INVALID, "null argument");
  const int model = in->model;
  if (model != CCDH_LOGISTIC && model != CCDH_POISSON && model != CCDH_CONDITIONAL_LOGISTIC &&
      model != CCDH_COX)
    return setError(CCDH_E_INVALID, "model");
  if (in->precision != CCDH_FP64 && in->precision != CCDH_MIXED)
    return setError(CCDH_E_INVALID, "unsupported precision");
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
  const bool intercept = in->add_intercept && (model == CCDH_LOGISTIC || poisson);
  const int lead = (poisson ? 1 : 0) + (intercept ? 1 : 0);

  hipStream_t s = nullptr;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{s};
  // the data set under construction, destroyed unless handed out
  struct DataGuard {
    ccdh_data* d = new ccdh_data();
    ~DataGuard() {
      if (d) dataDestroy(d);
    }
  } dg;
  ccdh_data* d = dg.d;
  d->model = model;
  d->precision = in->precision;
  d->K = n;
  d->hasOffset = poisson ? 1 : 0;
  auto dalloc = [&](auto** p, size_t count) -> hipError_t {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(*p)>>;
    hipError_t e = hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) d->devBytes += (int64_t)(std::max<size_t>(count, 1) * sizeof(T));
    return e;
  };
  prim::Scratch sortMem;
  int dup = 0;
  DevBuf<int> found;
  HIP_CHECK(found.alloc(1));

  // ---- 1. row order
  DevBuf<int64_t> dRowId, dStratum;
  DevBuf<double> dY, dTime;
  DevBuf<uint64_t> ka, kb;
  DevBuf<int32_t> perm, perm2;
  HIP_CHECK(dRowId.alloc(n));
  HIP_CHECK(dY.alloc(n));
  HIP_CHECK(ka.alloc(n));
  HIP_CHECK(kb.alloc(n));
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
  auto pass = [&](bool alt) {
    if (alt) {
      swapBuf(perm, perm2);
      swapBuf(ka, kb);
    }
  };
  auto passI64 = [&](const int64_t* src) -> hipError_t {
    k_key_i64<<<blocksFor(n), kT, 0, s>>>(n, src, perm.p, ka.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    bool alt = false;
    e = prim::radixSortPairs<int32_t>(ka.p, kb.p, perm.p, perm2.p, n, sortMem, s, &alt);
    pass(alt);
    return e;
  };
  auto passF64 = [&](const double* src, bool descending) -> hipError_t {
    k_key_f64<<<blocksFor(n), kT, 0, s>>>(n, src, perm.p, descending, ka.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    bool alt = false;
    e = prim::radixSortPairs<int32_t>(ka.p, kb.p, perm.p, perm2.p, n, sortMem, s, &alt);
    pass(alt);
    return e;
  };
  HIP_CHECK(passI64(dRowId.p));
  if (cox) {
    HIP_CHECK(passF64(dY.p, false));
    HIP_CHECK(passF64(dTime.p, true));
  }
  if (strata) HIP_CHECK(passI64(dStratum.p));

  // ---- 2. outcomes in row order, stratum numbers
  HIP_CHECK(dalloc(&d->y, n));
  k_gather<double><<<blocksFor(n), kT, 0, s>>>(n, dY.p, perm.p, d->y);
  HIP_CHECK(hipGetLastError());
  d->hY.resize(n);
  HIP_CHECK(hipMemcpyAsync(d->hY.data(), d->y, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  std::vector<double> hTime;
  if (cox || poisson) {
    DevBuf<double> t;
    double* tp = nullptr;
    if (cox) {
      HIP_CHECK(dalloc(&d->time, n));
      tp = d->time;
    } else {
      HIP_CHECK(t.alloc(n));
      tp = t.p;
    }
    k_gather<double><<<blocksFor(n), kT, 0, s>>>(n, dTime.p, perm.p, tp);
    HIP_CHECK(hipGetLastError());
    hTime.resize(n);
    HIP_CHECK(hipMemcpyAsync(hTime.data(), tp, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  HIP_CHECK(dalloc(&d->strata, n));
  if (strata) {
    DevBuf<uint8_t> flags;
    HIP_CHECK(flags.alloc(n));
    k_key_i64<<<blocksFor(n), kT, 0, s>>>(n, dStratum.p, perm.p, ka.p);
    k_change_flags<<<blocksFor(n), kT, 0, s>>>(n, ka.p, flags.p);
    HIP_CHECK(hipGetLastError());
    prim::Compactor cmp;
    int nChanges = 0;
    HIP_CHECK(cmp.plan(flags.p, n, s, &nChanges));
    k_stratum_numbers<<<blocksFor(n), kT, 0, s>>>(n, flags.p, cmp.pos, d->strata);
    HIP_CHECK(hipGetLastError());
    d->nStrata = nChanges + 1;
    d->hStrata.resize(n);
    HIP_CHECK(hipMemcpyAsync(d->hStrata.data(), d->strata, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } else {
    prim::k_fill<int32_t><<<blocksFor(n), kT, 0, s>>>(d->strata, n, 0);
    HIP_CHECK(hipGetLastError());
    d->nStrata = 1;
    d->hStrata.assign(n, 0);
  }
  d->hOffs = cox ? hTime : std::vector<double>(n, 0.0);

  // ---- 3. row map: row ids in row order, sorted, with their rows
  DevBuf<int32_t> rowOf;
  HIP_CHECK(rowOf.alloc(n));
  k_key_i64<<<blocksFor(n), kT, 0, s>>>(n, dRowId.p, perm.p, ka.p);
  k_iota<<<blocksFor(n), kT, 0, s>>>(n, rowOf.p);
  HIP_CHECK(hipGetLastError());
  {
    bool alt = false;
    HIP_CHECK(prim::radixSortPairs<int32_t>(ka.p, kb.p, rowOf.p, perm2.p, n, sortMem, s, &alt));
    if (alt) {
      swapBuf(ka, kb);
      swapBuf(rowOf, perm2);
    }
  }
  const uint64_t* sortedIds = ka.p;
  HIP_CHECK(hipMemsetAsync(found.p, 0, sizeof(int), s));
  k_adjacent_equal<uint64_t><<<blocksFor(n), kT, 0, s>>>(n, sortedIds, found.p);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(&dup, found.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (dup) return setError(CCDH_E_INVALID, "repeated outcome row id");

  // ---- 4./5. entries
  int nCols = 0;
  int64_t nnz = 0;
  DevBuf<uint64_t> key, key2;
  DevBuf<double> v, v2;
  if (m > 0) {
    DevBuf<int64_t> dCov, dEntryRow;
    DevBuf<double> dVal;
    DevBuf<int32_t> row;
    DevBuf<uint8_t> valid;
    DevBuf<uint64_t> covKey, covA, covB;
    HIP_CHECK(dCov.alloc(m));
    HIP_CHECK(dEntryRow.alloc(m));
    HIP_CHECK(row.alloc(m));
    HIP_CHECK(valid.alloc(m));
    HIP_CHECK(covKey.alloc(m));
    HIP_CHECK(hipMemcpyAsync(dCov.p, in->covariate_id, (size_t)m * 8, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dEntryRow.p, in->entry_row_id, (size_t)m * 8, hipMemcpyHostToDevice, s));
    if (in->value) {
      HIP_CHECK(dVal.alloc(m));
      HIP_CHECK(hipMemcpyAsync(dVal.p, in->value, (size_t)m * 8, hipMemcpyHostToDevice, s));
    }
    k_entry_rows<<<blocksFor(m), kT, 0, s>>>(m, dEntryRow.p, sortedIds, rowOf.p, n, dCov.p, row.p,
                                             valid.p, covKey.p);
    HIP_CHECK(hipGetLastError());
    // column ids present among entries with a row: sorted, unique
    prim::Compactor cmp;
    int nValid = 0;
    HIP_CHECK(cmp.plan(valid.p, m, s, &nValid));
    if (nValid > 0) {
      HIP_CHECK(covA.alloc(nValid));
      HIP_CHECK(covB.alloc(nValid));
      HIP_CHECK(cmp.apply(covKey.p, valid.p, m, covA.p, s));
      bool alt = false;
      HIP_CHECK(prim::radixSortPairs<int32_t>(covA.p, covB.p, nullptr, nullptr, nValid, sortMem, s,
                                              &alt));
      if (alt) swapBuf(covA, covB);
      DevBuf<uint8_t> first;
      HIP_CHECK(first.alloc(nValid));
      k_unique_flags<<<blocksFor(nValid), kT, 0, s>>>(nValid, covA.p, first.p);
      HIP_CHECK(hipGetLastError());
      prim::Compactor ucmp;
      HIP_CHECK(ucmp.plan(first.p, nValid, s, &nCols));
      HIP_CHECK(ucmp.apply(covA.p, first.p, nValid, covB.p, s));
      const uint64_t* uniq = covB.p;
      // entries sorted by (column, row); those without a row last
      HIP_CHECK(key.alloc(m));
      HIP_CHECK(key2.alloc(m));
      HIP_CHECK(v.alloc(m));
      HIP_CHECK(v2.alloc(m));
      k_entry_keys<<<blocksFor(m), kT, 0, s>>>(m, covKey.p, row.p, uniq, nCols, dVal.p, key.p, v.p);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(prim::radixSortPairs<double>(key.p, key2.p, v.p, v2.p, m, sortMem, s, &alt));
      if (alt) {
        swapBuf(key, key2);
        swapBuf(v, v2);
      }
      // loadMultipleX: a repeated (row, column) entry is an error, zero values included
      HIP_CHECK(hipMemsetAsync(found.p, 0, sizeof(int), s));
      k_adjacent_equal<uint64_t><<<blocksFor(nValid), kT, 0, s>>>(nValid, key.p, found.p);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(&dup, found.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (dup) return setError(CCDH_E_INVALID, "Repeated row-column entry");
      // zero values are not stored
      k_nonzero_flags<<<blocksFor(nValid), kT, 0, s>>>(nValid, v.p, valid.p);
      HIP_CHECK(hipGetLastError());
      prim::Compactor zcmp;
      int kept = 0;
      HIP_CHECK(zcmp.plan(valid.p, nValid, s, &kept));
      HIP_CHECK(zcmp.apply(key.p, valid.p, nValid, key2.p, s));
      HIP_CHECK(zcmp.apply(v.p, valid.p, nValid, v2.p, s));
      swapBuf(key, key2);
      swapBuf(v, v2);
      nnz = kept;
      HIP_CHECK(hipStreamSynchronize(s));
    }
  }

  // ---- the data set: [log time offset (Poisson)] [intercept (LR, Poisson)] covariates
  const int J = lead + nCols;
  if (J <= 0) return setError(CCDH_E_INVALID, "empty problem");
  const int64_t total = (int64_t)lead * n + nnz;
  if (total >= (int64_t)INT32_MAX) return setError(CCDH_E_INVALID, "nnz must be < 2^31");
  d->J = J;
  d->nnz = total;
  DevBuf<int> sparse;
  HIP_CHECK(sparse.alloc(std::max(nCols, 1)));
  HIP_CHECK(hipMemsetAsync(sparse.p, 0, (size_t)std::max(nCols, 1) * sizeof(int), s));
  HIP_CHECK(dalloc(&d->colPtr, (size_t)J + 1));
  HIP_CHECK(dalloc(&d->rows, (size_t)total));
  HIP_CHECK(dalloc(&d->fmt, (size_t)J));
  k_col_bounds<<<blocksFor(std::max<int64_t>(nnz, J + 1)), kT, 0, s>>>(nnz, key.p, nCols, lead, n,
                                                                       d->colPtr);
  k_rows_and_sparse<<<blocksFor(total), kT, 0, s>>>(nnz, key.p, v.p, lead, n, d->rows, sparse.p);
  k_formats<<<blocksFor(J), kT, 0, s>>>(J, lead, poisson, sparse.p, d->fmt);
  HIP_CHECK(hipGetLastError());
  d->hColPtr.resize((size_t)J + 1);
  d->hFmt.resize(J);
  HIP_CHECK(hipMemcpyAsync(d->hColPtr.data(), d->colPtr, (size_t)(J + 1) * 8,
                           hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(d->hFmt.data(), d->fmt, (size_t)J, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  // values when a column keeps them: the DENSE offset or a SPARSE column
  const bool withVals =
      std::find(d->hFmt.begin(), d->hFmt.end(), (int8_t)CCDH_SPARSE) != d->hFmt.end() || poisson;
  if (withVals) {
    HIP_CHECK(dalloc(&d->vals, (size_t)total));
    std::vector<double> logTime;
    if (poisson) {
      logTime.resize(n);
      for (int k = 0; k < n; ++k) logTime[k] = std::log(hTime[k]);
      HIP_CHECK(hipMemcpyAsync(d->vals, logTime.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
    }
    k_values<<<blocksFor(total), kT, 0, s>>>(nnz, v.p, lead, poisson, n, d->vals);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
  }
  // Poisson fixed term per row: -log(y!) as the loop of ModelSpecifics.h:1596-1601
  if (poisson) {
    d->hLogFact.resize(n);
    for (int k = 0; k < n; ++k) {
      double t = 0.0;
      for (int i = 2; i <= (int)d->hY[k]; ++i) t += -std::log((double)i);
      d->hLogFact[k] = t;
    }
  }
  int rc = buildTablesDevice(d, s);
  if (rc) return rc;
  HIP_CHECK(hipStreamSynchronize(s));
  *out = d;
  dg.d = nullptr;
  return 0;
}

}  // namespace ccdamd

extern "C" int ccdh_data_create_long(const ccdh_long_desc* desc, ccdh_data_t* out) {
  return ccdamd::dataCreateLong(desc, out);
}
