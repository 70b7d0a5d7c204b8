// CPU unit test of the served form's admission pieces (range allocator, request sizing, FIFO
// admission). Built and run by tests/test_serve_core.py; prints "ok" and exits 0 on success.
#include <cassert>
#include <cstdio>
#include <deque>
#include <random>
#include <vector>

#include "../../tim_raytracing_amd/csrc/device/serve_core.h"

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  rtq::RangeAllocator a;
  a.reset(1 << 20);
  CHECK(a.capacity() == (1u << 20));
  uint64_t o1, o2, o3;
  CHECK(a.alloc(1000, &o1) && o1 == 0);
  CHECK(a.alloc(256, &o2) && o2 == 1024);
  CHECK(a.alloc(1 << 19, &o3));
  CHECK(a.used() == 1024 + 256 + (1 << 19));
  uint64_t big;
  CHECK(!a.alloc(1 << 20, &big));  // does not fit while others are live
  a.release(o2, 256);
  a.release(o1, 1000);
  a.release(o3, 1 << 19);
  CHECK(a.used() == 0 && a.fragments() == 1);  // fully coalesced
  CHECK(a.alloc(1 << 20, &big) && big == 0);
  a.release(big, 1 << 20);

  // random alloc/free never overlaps and always coalesces back
  std::mt19937 rng(3);
  std::vector<std::pair<uint64_t, uint64_t>> live;
  for (int it = 0; it < 20000; ++it) {
    if (!live.empty() && (rng() % 2)) {
      size_t k = rng() % live.size();
      a.release(live[k].first, live[k].second);
      live.erase(live.begin() + k);
    } else {
      uint64_t sz = 1 + rng() % 50000, off;
      if (a.alloc(sz, &off)) {
        for (auto& l : live) CHECK(off + rtq::RangeAllocator::round(sz) <= l.first || l.first + rtq::RangeAllocator::round(l.second) <= off);
        CHECK(off + sz <= a.capacity());
        live.push_back({off, sz});
      }
    }
  }
  for (auto& l : live) a.release(l.first, l.second);
  CHECK(a.used() == 0 && a.fragments() == 1);

  // request sizes: 32 + 16 + 4 bytes per closest-hit ray, 32 + 4 per any-hit ray, 256-aligned parts
  CHECK(rtq::requestBytes(8, true) == 256 * 3);
  CHECK(rtq::requestBytes(1000, true) == 32000 + 16128 + 4096);
  CHECK(rtq::requestBytes(1000, false) == 32000 + 4096);

  // FIFO admission stops at the first request that does not fit
  std::deque<int> q = {100, 200, 5000, 10};
  int budget = 1000;
  std::vector<int> admitted;
  rtq::admitFifo(q, [&](int r) {
    if (r > budget) return false;
    budget -= r;
    admitted.push_back(r);
    return true;
  });
  CHECK(admitted.size() == 2 && q.size() == 2 && q.front() == 5000);
  std::puts("ok");
  return 0;
}
