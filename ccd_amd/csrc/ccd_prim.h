// ccd_prim.h -- device sort, scan and compaction primitives for the data-set
// builds (long-format ingestion, per-column tables).
//
//   radixSortPairs   stable least-significant-digit radix sort of unsigned 64-bit
//                    keys with an optional payload, 8-bit digits; one pass over
//                    the keys first counts all eight digits, and a digit that
//                    every key shares is skipped;
//   scanExclusive    exclusive prefix sum (reduce, scan of the tile sums, then
//                    scan with offsets), tiles of 2,048 items;
//   Compactor        stream compaction by a flag array (one scan, then one
//                    scatter per array compacted with the same flags).
//
// Each tile is 256 threads x 8 items.  Stability of the sort: a wave owns 512
// consecutive keys of its tile in eight chunks of 64; a key's rank among equal
// digits is (keys of the digit in earlier tiles) + (in earlier waves of its tile)
// + (in earlier chunks of its wave) + (in earlier lanes of its chunk, from a
// peer mask of eight ballots).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace ccdamd {
namespace prim {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kWaves = kThreads / 64;
constexpr int kBins = 256;

inline int tilesFor(int64_t n) { return (int)std::max<int64_t>(1, (n + kTile - 1) / kTile); }
inline int gridFor(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 65535));
}

// Scratch device memory of one build, grown on demand.
struct Scratch {
  void* p = nullptr;
  size_t bytes = 0;
  Scratch() = default;
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
  hipError_t need(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
};

// ---------------------------------------------------------------- key maps
// order-preserving maps to unsigned 64-bit keys
__host__ __device__ inline uint64_t keyOfI64(int64_t v) {
  return (uint64_t)v ^ 0x8000000000000000ull;
}
// -0.0 and +0.0 map to one key (they compare equal)
__host__ __device__ inline uint64_t keyOfF64(double v, bool descending) {
  if (v == 0.0) v = 0.0;
  uint64_t b = __builtin_bit_cast(uint64_t, v);
  b = (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
  return descending ? ~b : b;
}

template <typename T>
__global__ void k_fill(T* a, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// ---------------------------------------------------------------- scan
template <typename T>
__device__ inline T blockExclusive(T v, T* lds, T* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const T s = lds[w];
    if (w < wave) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

// per tile sum
template <typename T, typename In>
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const In* in, int64_t n, T* sums) {
  __shared__ T lds[kWaves];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  T s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i)
    if (base + i < n) s += (T)in[base + i];
  T tot;
  (void)blockExclusive<T>(s, lds, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// exclusive scan of each tile plus its offset (offs may be null: one tile);
// in and out may be the same array
template <typename T, typename In>
__global__ __launch_bounds__(kThreads) void k_scan_tiles(const In* in, T* out, int64_t n,
                                                         const T* offs) {
  __shared__ T lds[kWaves];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  T item[kItems];
  T s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    item[i] = base + i < n ? (T)in[base + i] : (T)0;
    s += item[i];
  }
  T tot;
  T run = blockExclusive<T>(s, lds, &tot) + (offs ? offs[blockIdx.x] : (T)0);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += item[i];
  }
}

// scratch elements of T a scan of n items needs
inline int64_t scanScratch(int64_t n) {
  int64_t t = 0;
  for (int64_t m = n; m > kTile;) {
    m = tilesFor(m);
    t += m;
  }
  return t;
}

// out[i] = in[0] + ... + in[i-1]; tmp holds scanScratch(n) elements of T
template <typename T, typename In>
hipError_t scanExclusive(const In* in, T* out, int64_t n, T* tmp, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n <= kTile) {
    k_scan_tiles<T, In><<<1, kThreads, 0, s>>>(in, out, n, nullptr);
    return hipGetLastError();
  }
  const int tiles = tilesFor(n);
  k_scan_reduce<T, In><<<tiles, kThreads, 0, s>>>(in, n, tmp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = scanExclusive<T, T>(tmp, tmp, tiles, tmp + tiles, s);
  if (e != hipSuccess) return e;
  k_scan_tiles<T, In><<<tiles, kThreads, 0, s>>>(in, out, n, tmp);
  return hipGetLastError();
}

// ---------------------------------------------------------------- compaction
template <typename T>
__global__ void k_compact(const T* in, const uint8_t* flag, const int32_t* pos, int64_t n, T* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (flag[i]) out[pos[i]] = in[i];
}

__global__ void k_flag_total(const uint8_t* flag, const int32_t* pos, int64_t n,
                                    int32_t* total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = n > 0 ? pos[n - 1] + (flag[n - 1] ? 1 : 0) : 0;
}

// positions of the flagged items: pos[i] (exclusive count), *count on the host
struct Compactor {
  int32_t* pos = nullptr;
  int32_t* total = nullptr;
  Scratch mem;
  hipError_t plan(const uint8_t* flag, int64_t n, hipStream_t s, int* count) {
    const int64_t tmpN = scanScratch(n);
    hipError_t e = mem.need((size_t)(std::max<int64_t>(n, 1) + 1 + tmpN) * sizeof(int32_t));
    if (e != hipSuccess) return e;
    pos = (int32_t*)mem.p;
    total = pos + std::max<int64_t>(n, 1);
    e = scanExclusive<int32_t, uint8_t>(flag, pos, n, total + 1, s);
    if (e != hipSuccess) return e;
    k_flag_total<<<1, 64, 0, s>>>(flag, pos, n, total);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    int c = 0;
    if ((e = hipMemcpyAsync(&c, total, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    *count = c;
    return hipSuccess;
  }
  template <typename T>
  hipError_t apply(const T* in, const uint8_t* flag, int64_t n, T* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_compact<T><<<gridFor(n), kThreads, 0, s>>>(in, flag, pos, n, out);
    return hipGetLastError();
  }
};

// ---------------------------------------------------------------- radix sort
// counts of all eight digits of every key
__global__ __launch_bounds__(kThreads) void k_radix_hist_all(const uint64_t* keys,
                                                                     int64_t n, uint32_t* hist) {
  __shared__ uint32_t h[8 * kBins];
  for (int i = threadIdx.x; i < 8 * kBins; i += kThreads) h[i] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const uint64_t k = keys[i];
#pragma unroll
    for (int b = 0; b < 8; ++b) atomicAdd(&h[b * kBins + (int)((k >> (8 * b)) & 255)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 8 * kBins; i += kThreads)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// counts[d * tiles + t]: keys of tile t with digit d
__global__ __launch_bounds__(kThreads) void k_radix_tile_counts(const uint64_t* keys,
                                                                        int64_t n, int shift,
                                                                        int tiles,
                                                                        uint32_t* counts) {
  __shared__ uint32_t c[kBins];
  c[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int i = threadIdx.x; i < kTile; i += kThreads)
    if (base + i < n) atomicAdd(&c[(int)((keys[base + i] >> shift) & 255)], 1u);
  __syncthreads();
  counts[(int64_t)threadIdx.x * tiles + blockIdx.x] = c[threadIdx.x];
}

// stable scatter of one digit pass; offs = exclusive scan of the tile counts
template <typename V>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(const uint64_t* keys, const V* vals,
                                                            int64_t n, int shift, int tiles,
                                                            const uint32_t* offs,
                                                            uint64_t* keysOut, V* valsOut) {
  __shared__ uint32_t waveCnt[kWaves][kBins];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) waveCnt[w][threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)wave * (64 * kItems);
  const uint64_t below = (1ull << lane) - 1ull;
  uint64_t k[kItems];
  uint32_t rank[kItems];
  int dig[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int64_t idx = base + (int64_t)i * 64 + lane;
    const bool ok = idx < n;
    k[i] = ok ? keys[idx] : 0ull;
    const int d = (int)((k[i] >> shift) & 255);
    dig[i] = d;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t before = ok ? waveCnt[wave][d] : 0u;
    rank[i] = before + (uint32_t)__popcll(peers & below);
    // the lowest lane of each digit group adds the group's count
    if (ok && (peers & below) == 0) waveCnt[wave][d] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  {
    const int d = threadIdx.x;
    uint32_t run = offs[(int64_t)d * tiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t c = waveCnt[w][d];
      waveCnt[w][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int64_t idx = base + (int64_t)i * 64 + lane;
    if (idx < n) {
      const uint32_t p = waveCnt[wave][dig[i]] + rank[i];
      keysOut[p] = k[i];
      if (vals) valsOut[p] = vals[idx];
    }
  }
}

// Stable ascending sort of n keys (and vals, which may be null) between the two
// buffers of each pair; *inAlt tells which holds the result.
template <typename V>
hipError_t radixSortPairs(uint64_t* keys, uint64_t* keysAlt, V* vals, V* valsAlt, int64_t n,
                          Scratch& scratch, hipStream_t s, bool* inAlt) {
  *inAlt = false;
  if (n <= 1) return hipSuccess;
  if (n >= (1ll << 31)) return hipErrorInvalidValue;
  const int tiles = tilesFor(n);
  const int64_t nCounts = (int64_t)kBins * tiles;
  const size_t bytes = (size_t)(8 * kBins + nCounts + scanScratch(nCounts)) * sizeof(uint32_t);
  hipError_t e = scratch.need(bytes);
  if (e != hipSuccess) return e;
  uint32_t* hist = (uint32_t*)scratch.p;
  uint32_t* counts = hist + 8 * kBins;
  uint32_t* tmp = counts + nCounts;
  k_fill<uint32_t><<<8, kThreads, 0, s>>>(hist, 8 * kBins, 0u);
  k_radix_hist_all<<<(int)std::min<int64_t>(tiles, 1024), kThreads, 0, s>>>(keys, n, hist);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  uint32_t h[8 * kBins];
  if ((e = hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  uint64_t* kIn = keys;
  uint64_t* kOut = keysAlt;
  V* vIn = vals;
  V* vOut = valsAlt;
  for (int b = 0; b < 8; ++b) {
    bool trivial = false;
    for (int d = 0; d < kBins && !trivial; ++d) trivial = h[b * kBins + d] == (uint32_t)n;
    if (trivial) continue;
    k_radix_tile_counts<<<tiles, kThreads, 0, s>>>(kIn, n, 8 * b, tiles, counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = scanExclusive<uint32_t, uint32_t>(counts, counts, nCounts, tmp, s)) != hipSuccess)
      return e;
    k_radix_scatter<V><<<tiles, kThreads, 0, s>>>(kIn, vIn, n, 8 * b, tiles, counts, kOut, vOut);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    std::swap(kIn, kOut);
    std::swap(vIn, vOut);
    *inAlt = !*inAlt;
  }
  return hipSuccess;
}

}  // namespace
}  // namespace prim
}  // namespace ccdamd
