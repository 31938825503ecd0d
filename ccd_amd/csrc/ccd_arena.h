// ccd_arena.h -- device memory for served fits.
//
// The server takes one device allocation (its HBM budget) up front and hands
// out one contiguous block per admitted request: the request's data (X, outcome,
// index structures) and its fit state.  Admission is the success of that block
// allocation, so requests are admitted by HBM footprint, and the steady state
// makes no hipMalloc/hipFree calls (hipFree waits for the device, which would
// stall the running sweeps).  Blocks are first-fit with coalescing on release.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>

namespace ccdamd {

class Arena {
 public:
  static constexpr size_t kAlign = 256;
  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;

  // Adopt [base, base+bytes) (owned by the caller).
  void reset(char* base, size_t bytes) {
    std::lock_guard<std::mutex> g(mu_);
    base_ = base;
    cap_ = bytes / kAlign * kAlign;
    free_.clear();
    live_.clear();
    used_ = peak_ = 0;
    if (cap_) free_[0] = cap_;
  }

  // nullptr when no free range of that size exists
  void* alloc(size_t bytes) {
    bytes = round(bytes);
    std::lock_guard<std::mutex> g(mu_);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes) free_[off + bytes] = len - bytes;
      live_[off] = bytes;
      used_ += bytes;
      if (used_ > peak_) peak_ = used_;
      return base_ + off;
    }
    return nullptr;
  }

  void release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    const size_t off = (size_t)((char*)p - base_);
    auto lv = live_.find(off);
    if (lv == live_.end()) return;
    size_t start = off, len = lv->second;
    used_ -= len;
    live_.erase(lv);
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first == start + len) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    free_[start] = len;
  }

  size_t capacity() const { return cap_; }
  size_t used() {
    std::lock_guard<std::mutex> g(mu_);
    return used_;
  }
  size_t peak() {
    std::lock_guard<std::mutex> g(mu_);
    return peak_;
  }
  static size_t round(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

 private:
  std::mutex mu_;
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  std::map<size_t, size_t> live_;  // offset -> length
};

// Sequential carving of one block (a request's data arrays).
struct Bump {
  char* p = nullptr;
  size_t left = 0;
  void* take(size_t bytes) {
    bytes = Arena::round(bytes > 0 ? bytes : 1);
    if (bytes > left) return nullptr;
    void* r = p;
    p += bytes;
    left -= bytes;
    return r;
  }
};

}  // namespace ccdamd
