// mpm_arena.h -- slab sub-allocator (host code, no HIP types).
//
// Memory handed out from slabs that stay allocated until the arena is
// destroyed.  A Source provides the slabs:
//   int get(size_t bytes, void **p)   0 on success, an error code otherwise
//   void put(void *p)                 returns a slab
//   static constexpr int kNoMemory    the code alloc returns at the cap
// Blocks are multiples of 256 bytes at 256-byte aligned addresses (slabs
// must be 256-byte aligned); a freed block merges with the free neighbours
// of its own slab; an allocation takes the smallest free block that holds it
// and, when none does, a new slab of max(slab_bytes, request).  A cap (0:
// none) bounds the slabs' total.  Thread-safe.
#pragma once
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>
#include <mutex>
#include <vector>

namespace mpm {

template <class Source>
class SlabArena {
 public:
  SlabArena(Source src, size_t slab_bytes, size_t cap = 0)
      : src_(src), slab_bytes_(slab_bytes), cap_(cap) {}
  ~SlabArena() {
    for (void *b : slabs_) src_.put(b);
  }
  SlabArena(const SlabArena &) = delete;
  SlabArena &operator=(const SlabArena &) = delete;

  static constexpr size_t kAlign = 256;

  int alloc(size_t bytes, void **p) {
    bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
    if (!bytes) bytes = kAlign;
    std::lock_guard<std::mutex> lk(mu_);
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second.size >= bytes && (best == free_.end() || it->second.size < best->second.size))
        best = it;
    if (best == free_.end()) {
      const size_t sz = bytes > slab_bytes_ ? bytes : slab_bytes_;
      if (cap_ && reserved_ + sz > cap_) return Source::kNoMemory;
      void *b = nullptr;
      const int e = src_.get(sz, &b);
      if (e) return e;
      slabs_.push_back(b);
      reserved_ += sz;
      best = free_.emplace((char *)b, Block{sz, (uint32_t)(slabs_.size() - 1)}).first;
    }
    char *a = best->first;
    const Block blk = best->second;
    free_.erase(best);
    if (blk.size > bytes) free_.emplace(a + bytes, Block{blk.size - bytes, blk.slab});
    used_.emplace(a, Block{bytes, blk.slab});
    used_bytes_ += bytes;
    *p = a;
    return 0;
  }

  // Pointers that this arena did not hand out (or handed out and took back)
  // are ignored.
  void release(void *p) {
    std::lock_guard<std::mutex> lk(mu_);
    auto u = used_.find((char *)p);
    if (u == used_.end()) return;
    char *a = u->first;
    Block blk = u->second;
    used_.erase(u);
    used_bytes_ -= blk.size;
    auto next = free_.find(a + blk.size);
    if (next != free_.end() && next->second.slab == blk.slab) {
      blk.size += next->second.size;
      free_.erase(next);
    }
    auto it = free_.emplace(a, blk).first;
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->second.slab == blk.slab && prev->first + prev->second.size == a) {
        prev->second.size += blk.size;
        free_.erase(it);
      }
    }
  }

  size_t reserved() const {
    std::lock_guard<std::mutex> lk(mu_);
    return reserved_;
  }
  size_t used_bytes() const {
    std::lock_guard<std::mutex> lk(mu_);
    return used_bytes_;
  }
  size_t free_blocks() const {
    std::lock_guard<std::mutex> lk(mu_);
    return free_.size();
  }
  size_t slabs() const {
    std::lock_guard<std::mutex> lk(mu_);
    return slabs_.size();
  }

 private:
  struct Block {
    size_t size;
    uint32_t slab;
  };
  Source src_;
  size_t slab_bytes_;
  size_t cap_;
  size_t reserved_ = 0, used_bytes_ = 0;
  mutable std::mutex mu_;
  std::vector<void *> slabs_;
  std::map<char *, Block> free_, used_;
};

}  // namespace mpm
