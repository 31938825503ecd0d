// Randomised check of the serving arena (first fit, coalescing): every live
// block lies inside the arena, no two overlap, used() equals the sum of live
// blocks, and after releasing everything one block of the full size fits.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../ccd_amd/csrc/ccd_arena.h"

int main() {
  using ccdamd::Arena;
  const size_t cap = 1 << 20;
  std::vector<char> mem(cap + Arena::kAlign);
  Arena a;
  a.reset(mem.data(), cap);
  std::mt19937 rng(7);
  struct Blk { char* p; size_t n; };
  std::vector<Blk> live;
  for (int step = 0; step < 20000; ++step) {
    if (live.empty() || rng() % 3 != 0) {
      const size_t n = 1 + rng() % 40000;
      char* p = (char*)a.alloc(n);
      if (!p) continue;
      const size_t r = Arena::round(n);
      if (p < mem.data() || p + r > mem.data() + cap) { std::puts("out of arena"); return 1; }
      if ((size_t)(p - mem.data()) % Arena::kAlign) { std::puts("misaligned"); return 1; }
      for (const Blk& b : live)
        if (p < b.p + b.n && b.p < p + r) { std::puts("overlap"); return 1; }
      live.push_back({p, r});
    } else {
      const size_t i = rng() % live.size();
      a.release(live[i].p);
      live.erase(live.begin() + i);
    }
    size_t sum = 0;
    for (const Blk& b : live) sum += b.n;
    if (sum != a.used()) { std::puts("used() mismatch"); return 1; }
  }
  for (const Blk& b : live) a.release(b.p);
  if (a.used() != 0 || !a.alloc(cap)) { std::puts("not coalesced"); return 1; }
  if (a.alloc(1)) { std::puts("over-allocated"); return 1; }
  std::puts("ok");
  return 0;
}
