/*
 * Process-wide cache of device allocations (declarations and policy in
 * pool.h).
 */
#include "pool.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>

#include "common.h"

namespace hodlr {
namespace {

constexpr size_t kLarge = size_t(16) << 20;  // blocks from 16 MiB are rounded to 2 MiB
constexpr size_t kSlack = size_t(2) << 20;

struct Device {
  std::multimap<size_t, void*> cached;  // size -> block
  long long live = 0, cached_bytes = 0, hits = 0, misses = 0;
  double malloc_s = 0, free_s = 0;
  long long cap = -1;  // cached bytes kept at most (-1: not read yet)
};

struct Live {
  int dev;
  size_t bytes;
};

struct State {
  std::mutex mu;
  std::map<int, Device> devs;
  std::unordered_map<void*, Live> live;
  bool enabled = std::getenv("HODLR_NO_MEMORY_CACHE") == nullptr;
};

// Never destroyed: the HIP runtime may be torn down before static
// destructors run, and the driver reclaims the memory at process exit.
State& state() {
  static State* s = new State();
  return *s;
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t round_size(size_t b) {
  if (b >= kLarge) return (b + kSlack - 1) / kSlack * kSlack;
  return (b + 511) / 512 * 512;
}

int current_device() {
  int d = 0;
  HODLR_HIP_CHECK(hipGetDevice(&d));
  return d;
}

void free_block(Device& d, void* p) {
  const double t = now();
  (void)hipFree(p);
  d.free_s += now() - t;
}

void trim(Device& d) {
  for (auto& kv : d.cached) free_block(d, kv.second);
  d.cached.clear();
  d.cached_bytes = 0;
}

// every device's cache, each after its own device is idle
void trim_all(State& s) {
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto& kv : s.devs) {
    if (kv.second.cached.empty()) continue;
    (void)hipSetDevice(kv.first);
    (void)hipDeviceSynchronize();
    trim(kv.second);
  }
  (void)hipSetDevice(cur);
}

}  // namespace

void* pool_alloc(size_t bytes) {
  if (bytes == 0) return nullptr;
  State& s = state();
  const size_t want = round_size(bytes);
  const int dev = current_device();
  std::lock_guard<std::mutex> lk(s.mu);
  Device& d = s.devs[dev];
  if (d.cap < 0) {
    size_t fr = 0, tot = 0;
    d.cap = hipMemGetInfo(&fr, &tot) == hipSuccess ? (long long)(tot / 2) : 0;
  }
  auto it = d.cached.lower_bound(want);
  if (it != d.cached.end() && it->first <= want + want / 4 + kSlack) {
    void* p = it->second;
    const size_t sz = it->first;
    d.cached.erase(it);
    d.cached_bytes -= (long long)sz;
    d.live += (long long)sz;
    d.hits++;
    s.live[p] = {dev, sz};
    return p;
  }
  void* p = nullptr;
  double t = now();
  hipError_t e = hipMalloc(&p, want);
  d.malloc_s += now() - t;
  if (e != hipSuccess && !d.cached.empty()) {  // give the cache back and try again
    (void)hipGetLastError();
    trim(d);
    t = now();
    e = hipMalloc(&p, want);
    d.malloc_s += now() - t;
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw HipError("hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  d.live += (long long)want;
  d.misses++;
  s.live[p] = {dev, want};
  return p;
}

void pool_release(void* p) {
  if (!p) return;
  State& s = state();
  std::lock_guard<std::mutex> lk(s.mu);
  auto it = s.live.find(p);
  if (it == s.live.end()) {  // not from the cache
    (void)hipFree(p);
    return;
  }
  const Live l = it->second;
  s.live.erase(it);
  Device& d = s.devs[l.dev];
  d.live -= (long long)l.bytes;
  if (!s.enabled || d.cached_bytes + (long long)l.bytes > d.cap) {
    free_block(d, p);
    return;
  }
  // queued work of every stream may still use the block (hipFree would wait
  // for it as well)
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (cur != l.dev) (void)hipSetDevice(l.dev);
  (void)hipDeviceSynchronize();
  if (cur != l.dev) (void)hipSetDevice(cur);
  d.cached.emplace(l.bytes, p);
  d.cached_bytes += (long long)l.bytes;
}

void pool_trim() {
  State& s = state();
  const int dev = current_device();
  std::lock_guard<std::mutex> lk(s.mu);
  (void)hipDeviceSynchronize();
  trim(s.devs[dev]);
}

PoolStats pool_stats() {
  State& s = state();
  const int dev = current_device();
  std::lock_guard<std::mutex> lk(s.mu);
  const Device& d = s.devs[dev];
  return {d.live, d.cached_bytes, d.hits, d.misses, d.malloc_s, d.free_s};
}

void pool_set_enabled(bool on) {
  State& s = state();
  std::lock_guard<std::mutex> lk(s.mu);
  s.enabled = on;
  if (!on) trim_all(s);
}

}  // namespace hodlr
