// edk_blockcache.h -- per-device cache of freed device blocks (host logic only).
//
// Allocations are rounded up to size classes (1/8 of the enclosing power of
// two above 64 MiB, at most 12.5 % extra), so a block freed by one job serves
// the next job of a similar sector.  Blocks stay cached for the life of the
// process; an allocation that finds the device full releases cached blocks,
// the largest first, until the released bytes cover it, retries, and
// releases the rest of the cache only if that still fails.  The raw allocator
// is a template parameter so the same logic runs over hipMalloc in the
// library and over a host heap in the stress test (tests/cache_stress.cpp):
//   struct Raw { static int alloc(void **p, size_t bytes);   // RAW_OK / RAW_OOM / RAW_ERR
//                static void free(void *p); };
//
// Invariants, each checked where it could break:
//   - a block is either handed out or cached, never both (free_set mirrors
//     the cached lists; a second release of a cached block is counted as a
//     double free and ignored, so the block is never handed out twice);
//   - only cached blocks are released to the raw allocator; a handed-out block
//     stays owned until its holder gives it back;
//   - every release of cached blocks is counted (and reported on stderr by
//     the library).
#pragma once
#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <vector>

enum { RAW_OK = 0, RAW_OOM = 1, RAW_ERR = 2 };

struct BlockCacheStats {
  int64_t cached_bytes = 0;    // bytes held in the cached lists
  int64_t owned_blocks = 0;    // blocks allocated through the cache and not yet released to the raw allocator
  int64_t releases = 0;        // cache releases after an allocation failure (largest blocks first, or all)
  int64_t released_bytes = 0;  // bytes they returned
  int64_t double_frees = 0;    // releases of a block that was already cached (ignored)
  int64_t foreign_frees = 0;   // releases of a block the cache never handed out (passed to the raw allocator)
};

inline size_t edk_size_class(size_t b) {
  if (b < ((size_t)64 << 20)) return (b + 255) & ~(size_t)255;
  const int e = 63 - __builtin_clzll((unsigned long long)b);
  const size_t g = (size_t)1 << (e - 3);
  return (b + g - 1) & ~(g - 1);
}

template <class Raw>
struct BlockCacheT {
  std::mutex mu;
  std::map<size_t, std::vector<void *>> free_blocks;
  std::unordered_set<void *> free_set;       // the cached blocks
  std::unordered_map<void *, size_t> owned;  // every block the cache handed out or holds -> its size class
  BlockCacheStats st;

  // A block of at least `bytes`; RAW_OOM when the raw allocator is full even
  // after the cache was released.  *released (optional) gets the bytes that
  // release returned (0: none).
  int alloc(void **p, size_t bytes, size_t *released = nullptr) {
    if (released) *released = 0;
    if (bytes == 0) bytes = 16;
    const size_t b = edk_size_class(bytes);
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = free_blocks.find(b);
      if (it != free_blocks.end() && !it->second.empty()) {
        *p = it->second.back();
        it->second.pop_back();
        free_set.erase(*p);
        st.cached_bytes -= (int64_t)b;
        return RAW_OK;
      }
    }
    int e = Raw::alloc(p, b);
    if (e == RAW_OOM) {
      size_t r = release_largest(b);
      if (r) e = Raw::alloc(p, b);
      if (e == RAW_OOM) {
        r += release_all();
        e = Raw::alloc(p, b);
      }
      if (released) *released = r;
    }
    if (e == RAW_OK) {
      std::lock_guard<std::mutex> lk(mu);
      owned[*p] = b;
      st.owned_blocks = (int64_t)owned.size();
    }
    return e;
  }

  // Gives a block back (its last use must have completed).
  void release(void *p) {
    if (!p) return;
    std::unique_lock<std::mutex> lk(mu);
    if (free_set.count(p)) {
      st.double_frees++;
      return;
    }
    auto it = owned.find(p);
    if (it == owned.end()) {
      st.foreign_frees++;
      lk.unlock();
      Raw::free(p);
      return;
    }
    free_blocks[it->second].push_back(p);
    free_set.insert(p);
    st.cached_bytes += (int64_t)it->second;
  }

  // Returns cached blocks to the raw allocator, largest size class first,
  // until at least `need` bytes are released (or the cache is empty); counts
  // as one release.  The bytes released.
  size_t release_largest(size_t need) {
    std::lock_guard<std::mutex> lk(mu);
    size_t bytes = 0;
    for (auto it = free_blocks.rbegin(); it != free_blocks.rend() && bytes < need; ++it) {
      std::vector<void *> &v = it->second;
      while (!v.empty() && bytes < need) {
        void *q = v.back();
        v.pop_back();
        Raw::free(q);
        owned.erase(q);
        free_set.erase(q);
        bytes += it->first;
        st.cached_bytes -= (int64_t)it->first;
      }
    }
    if (bytes) {
      st.owned_blocks = (int64_t)owned.size();
      st.releases++;
      st.released_bytes += (int64_t)bytes;
    }
    return bytes;
  }

  // Returns every cached block to the raw allocator; handed-out blocks are
  // untouched.  The bytes released.
  size_t release_all() {
    std::lock_guard<std::mutex> lk(mu);
    size_t bytes = 0;
    for (auto &kv : free_blocks)
      for (void *q : kv.second) {
        Raw::free(q);
        owned.erase(q);
        bytes += kv.first;
      }
    free_blocks.clear();
    free_set.clear();
    st.cached_bytes = 0;
    st.owned_blocks = (int64_t)owned.size();
    st.releases++;
    st.released_bytes += (int64_t)bytes;
    return bytes;
  }

  BlockCacheStats stats() {
    std::lock_guard<std::mutex> lk(mu);
    return st;
  }
};
