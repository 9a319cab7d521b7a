// Host-side check of the slab sub-allocator (mls-mpm_amd/csrc/mpm_arena.h)
// with slabs from aligned_alloc: alignment, no overlap among live blocks,
// best fit, merging back to whole slabs, slabs of their own for large
// requests, the cap, source failures, and foreign pointers ignored.
// Prints "ok" and exits 0, or names the failed check and exits 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "mpm_arena.h"

namespace {

struct MallocSlabs {
  static constexpr int kNoMemory = 2;
  int *gets;
  int *fail_after;  // fail the get after this many (negative: never)
  int get(size_t bytes, void **p) {
    if (*fail_after >= 0 && *gets >= *fail_after) return 7;
    ++*gets;
    *p = std::aligned_alloc(256, bytes);
    return *p ? 0 : 7;
  }
  void put(void *p) { std::free(p); }
};

int failures = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::fprintf(stderr, "failed: %s (line %d)\n", #c, __LINE__); \
      failures++;                                                 \
    }                                                             \
  } while (0)

using Arena = mpm::SlabArena<MallocSlabs>;

struct Live {
  char *p;
  size_t n;
};

bool overlaps(const std::vector<Live> &live) {
  for (size_t i = 0; i < live.size(); i++)
    for (size_t j = i + 1; j < live.size(); j++) {
      const char *a = live[i].p, *b = live[j].p;
      if (a < b + live[j].n && b < a + live[i].n) return true;
    }
  return false;
}

}  // namespace

int main() {
  int gets = 0, never = -1;
  const size_t slab = 1 << 20;
  {
    // random traffic: alignment, no overlap, bytes accounted
    Arena ar(MallocSlabs{&gets, &never}, slab);
    std::mt19937 rng(5);
    std::vector<Live> live;
    for (int it = 0; it < 4000; it++) {
      if (live.empty() || rng() % 3) {
        size_t n = 1 + rng() % 300000;
        if (rng() % 50 == 0) n = slab + rng() % slab;  // a slab of its own
        void *p = nullptr;
        CHECK(ar.alloc(n, &p) == 0);
        CHECK(((uintptr_t)p & 255) == 0);
        std::memset(p, it & 255, n);  // the whole block is usable
        live.push_back({(char *)p, (n + 255) & ~(size_t)255});
      } else {
        const size_t k = rng() % live.size();
        ar.release(live[k].p);
        live.erase(live.begin() + k);
      }
      if (it % 200 == 0) CHECK(!overlaps(live));
    }
    CHECK(!overlaps(live));
    size_t used = 0;
    for (const Live &l : live) used += l.n;
    CHECK(ar.used_bytes() == used);
    for (const Live &l : live) ar.release(l.p);
    // every slab merged back into one free block
    CHECK(ar.used_bytes() == 0);
    CHECK(ar.free_blocks() == ar.slabs());
  }
  {
    // best fit, and a freed block merging with both neighbours
    Arena ar(MallocSlabs{&gets, &never}, slab);
    void *a, *b, *c, *d, *e;
    CHECK(ar.alloc(4096, &a) == 0);
    CHECK(ar.alloc(1024, &b) == 0);
    CHECK(ar.alloc(8192, &c) == 0);
    CHECK(ar.alloc(1024, &d) == 0);
    CHECK(ar.alloc(512, &e) == 0);
    ar.release(c);  // an 8192 hole
    ar.release(a);  // a 4096 hole
    void *f;
    CHECK(ar.alloc(3000, &f) == 0);
    CHECK(f == a);  // the smaller hole that fits
    void *g;
    CHECK(ar.alloc(5000, &g) == 0);
    CHECK(g == c);
    ar.release(f);
    ar.release(g);
    ar.release(b);  // a, b, c merge into one block
    void *h;
    CHECK(ar.alloc(4096 + 1024 + 8192, &h) == 0);
    CHECK(h == a);
    ar.release(h);
    ar.release(d);
    ar.release(e);
    CHECK(ar.free_blocks() == 1 && ar.slabs() == 1);
    ar.release((char *)a + 256);  // not a block start: ignored
    ar.release(nullptr);
    CHECK(ar.used_bytes() == 0 && ar.free_blocks() == 1);
  }
  {
    // the cap, and a failing source
    int g2 = 0, fail = -1;
    Arena ar(MallocSlabs{&g2, &fail}, slab, 2 * slab);
    void *a, *b, *c;
    CHECK(ar.alloc(slab, &a) == 0);
    CHECK(ar.alloc(slab, &b) == 0);
    CHECK(ar.alloc(256, &c) == MallocSlabs::kNoMemory);
    ar.release(a);
    CHECK(ar.alloc(256, &c) == 0);  // reuses the freed slab
    CHECK(ar.reserved() == 2 * slab);
    int g3 = 0, fail0 = 0;
    Arena bad(MallocSlabs{&g3, &fail0}, slab);
    void *z = nullptr;
    CHECK(bad.alloc(100, &z) == 7);
    CHECK(bad.reserved() == 0 && bad.slabs() == 0);
  }
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
