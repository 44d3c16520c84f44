/*
 * Process-wide cache of device allocations.
 *
 * Handles allocate their panels, factors and workspaces in multi-GB pieces
 * and release them when they are destroyed or grow; a serving loop that
 * builds one handle per request would otherwise pay hipMalloc/hipFree for
 * every piece, every time.  Released blocks are kept per device and handed
 * out again for requests they fit (at most 25% + 2 MiB larger), so repeated
 * handles of one shape allocate nothing after the first.
 *
 * Ordering: a block is cached only after hipDeviceSynchronize, so no queued
 * kernel of any stream still uses it when another handle receives it (the
 * same guarantee hipFree gives).  When hipMalloc fails, the cache of that
 * device is freed and the allocation tried once more.  Cached bytes above a
 * cap (half of the device's memory) are freed at release.
 */
#ifndef DENSE_HODLR_POOL_H
#define DENSE_HODLR_POOL_H

#include <cstddef>

namespace hodlr {

struct PoolStats {
  long long live_bytes;    // handed out and not yet released
  long long cached_bytes;  // released, kept for reuse
  long long hits;          // requests served from the cache
  long long misses;        // requests that called hipMalloc
  double malloc_s;         // host seconds spent in hipMalloc
  double free_s;           // host seconds spent in hipFree
};

void* pool_alloc(size_t bytes);  // throws HipError when the device is out of memory
void pool_release(void* p);
void pool_trim();                // hipFree every cached block of the current device
PoolStats pool_stats();          // current device
void pool_set_enabled(bool on);  // off: release frees at once (the cache is trimmed)

}  // namespace hodlr

#endif
