// cache_stress.cpp -- the device block cache (edk_blockcache.h) over a host
// heap with a byte budget, driven by threads that allocate, hold, stamp and
// give back blocks while the budget keeps forcing cache releases.
//
// Checked on every step:
//   - a block is never held by two threads at once;
//   - a held block is never returned to the raw allocator (it stays alive
//     and keeps its holder's stamp until given back);
//   - the raw allocator never frees a block it does not hold;
// and at the end: a second release of a cached block is refused (counted,
// never handed out twice), a block the cache did not hand out is passed
// through, and releasing everything leaves no raw block alive.
//
// usage: cache_stress THREADS ITERATIONS BUDGET_MB   (exit status 0: all hold)
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "edk_blockcache.h"

static std::atomic<int> g_errors{0};
static void fail(const char *what) {
  if (g_errors.fetch_add(1) < 10) fprintf(stderr, "FAIL: %s\n", what);
}

static std::mutex g_heap_mu;
static size_t g_budget = 0, g_used = 0;
static std::unordered_map<void *, size_t> g_live;  // raw blocks alive -> bytes
static std::atomic<long> g_raw_oom{0};

struct HeapRaw {
  static int alloc(void **p, size_t b) {
    std::lock_guard<std::mutex> lk(g_heap_mu);
    if (g_used + b > g_budget) {
      g_raw_oom++;
      return RAW_OOM;
    }
    *p = malloc(b);
    if (!*p) return RAW_ERR;
    g_used += b;
    g_live[*p] = b;
    return RAW_OK;
  }
  static void free(void *p) {
    std::lock_guard<std::mutex> lk(g_heap_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {
      fail("raw free of a block that is not alive");
      return;
    }
    g_used -= it->second;
    g_live.erase(it);
    ::free(p);
  }
};

static bool alive(void *p) {
  std::lock_guard<std::mutex> lk(g_heap_mu);
  return g_live.count(p) != 0;
}

static BlockCacheT<HeapRaw> g_cache;
static std::mutex g_hold_mu;
static std::unordered_map<void *, int> g_holder;  // held block -> thread

struct Held {
  void *p;
  uint64_t stamp;
};

static void worker(int id, long iters) {
  std::mt19937_64 rng(1234 + id);
  std::vector<Held> held;
  // sizes around a few classes, below and above the 64 MiB rounding threshold
  const size_t sizes[] = {1000, 4096, 70000, 1 << 20, (3 << 20) + 17, 9 << 20};
  for (long it = 0; it < iters; it++) {
    const bool take = held.empty() || (held.size() < 6 && (rng() & 1));
    if (take) {
      void *p = nullptr;
      const size_t b = sizes[rng() % 6] + (rng() % 512);
      if (g_cache.alloc(&p, b) != RAW_OK) continue;  // full even after a release: try again later
      {
        std::lock_guard<std::mutex> lk(g_hold_mu);
        if (!g_holder.emplace(p, id).second) fail("block handed out while another thread holds it");
      }
      if (!alive(p)) fail("handed-out block is not alive");
      const uint64_t stamp = ((uint64_t)id << 40) | (uint64_t)it;
      memcpy(p, &stamp, sizeof(stamp));
      held.push_back({p, stamp});
    } else {
      const size_t k = rng() % held.size();
      Held h = held[k];
      held[k] = held.back();
      held.pop_back();
      uint64_t st = 0;
      if (!alive(h.p)) fail("held block was returned to the raw allocator");
      else {
        memcpy(&st, h.p, sizeof(st));
        if (st != h.stamp) fail("held block was overwritten by another holder");
      }
      {
        std::lock_guard<std::mutex> lk(g_hold_mu);
        g_holder.erase(h.p);
      }
      g_cache.release(h.p);
    }
  }
  for (Held &h : held) {
    {
      std::lock_guard<std::mutex> lk(g_hold_mu);
      g_holder.erase(h.p);
    }
    g_cache.release(h.p);
  }
}

int main(int argc, char **argv) {
  const int nt = argc > 1 ? atoi(argv[1]) : 8;
  const long iters = argc > 2 ? atol(argv[2]) : 20000;
  g_budget = (size_t)(argc > 3 ? atol(argv[3]) : 48) << 20;
  std::vector<std::thread> th;
  for (int i = 0; i < nt; i++) th.emplace_back(worker, i, iters);
  for (auto &t : th) t.join();
  BlockCacheStats st = g_cache.stats();
  if (st.double_frees || st.foreign_frees) fail("unexpected double or foreign release under threads");
  if (st.releases == 0) fail("the budget never forced a cache release (test too weak)");
  const long releases_threads = (long)st.releases;

  // a second release of a cached block is refused and never hands it out twice
  void *a = nullptr, *b = nullptr, *c = nullptr;
  if (g_cache.alloc(&a, 5000) != RAW_OK) fail("alloc a");
  g_cache.release(a);
  g_cache.release(a);
  st = g_cache.stats();
  if (st.double_frees != 1) fail("double release not counted");
  if (g_cache.alloc(&b, 5000) != RAW_OK || g_cache.alloc(&c, 5000) != RAW_OK) fail("alloc b, c");
  if (b == c) fail("one block handed out twice after a double release");
  g_cache.release(b);
  g_cache.release(c);
  // a block the cache never handed out goes straight to the raw allocator
  void *f = nullptr;
  if (HeapRaw::alloc(&f, 4096) != RAW_OK) fail("raw alloc");
  g_cache.release(f);
  if (alive(f)) fail("foreign block was not passed to the raw allocator");
  if (g_cache.stats().foreign_frees != 1) fail("foreign release not counted");
  // everything given back: releasing the cache leaves nothing alive
  g_cache.release_all();
  st = g_cache.stats();
  {
    std::lock_guard<std::mutex> lk(g_heap_mu);
    if (!g_live.empty() || g_used != 0) fail("raw blocks alive after everything was released");
  }
  if (st.owned_blocks != 0 || st.cached_bytes != 0) fail("cache still owns blocks after release_all");
  printf("threads %d iterations %ld: cache releases %ld, raw allocation failures %ld, errors %d\n", nt, iters,
         releases_threads, (long)g_raw_oom, g_errors.load());
  return g_errors.load() ? 1 : 0;
}
