// Cached device allocations.
//
// Cell, preconditioner and solver set-up allocate a few large buffers (about
// 17 GB for a 256^3 six-load mixed solve).  hipMalloc/hipFree of such blocks
// costs milliseconds each, so freed blocks are kept per (device, size) and
// handed back to the next request of exactly that size: repeated solves of one
// grid, the common pattern, allocate from the driver once.  This plays the
// role the PyTorch caching allocator plays for the reference's buffers.
//
// A block is returned to the cache only after its owner synchronised the
// stream that used it (the destroy entry points do), so a reused block never
// has pending work.  When hipMalloc runs out of memory the cache of that
// device is released and the request retried once.
#include "vxl_internal.h"

#include <map>
#include <mutex>
#include <unordered_map>

namespace vxl {
namespace {

struct Block {
  int device;
  size_t bytes;
};

std::mutex g_mu;
std::unordered_map<void*, Block> g_live;
std::multimap<std::pair<int, size_t>, void*> g_free;
size_t g_cached = 0;

void release_device(int device) {  // g_mu held
  for (auto it = g_free.begin(); it != g_free.end();) {
    if (it->first.first == device) {
      int cur = 0;
      (void)hipGetDevice(&cur);
      (void)hipSetDevice(device);
      (void)hipFree(it->second);
      (void)hipSetDevice(cur);
      g_cached -= it->first.second;
      it = g_free.erase(it);
    } else {
      ++it;
    }
  }
}

}  // namespace

hipError_t dev_malloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 1;
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_free.find({device, bytes});
  if (it != g_free.end()) {
    *p = it->second;
    g_free.erase(it);
    g_cached -= bytes;
    g_live[*p] = {device, bytes};
    return hipSuccess;
  }
  e = hipMalloc(p, bytes);
  if (e == hipErrorOutOfMemory && g_cached > 0) {
    (void)hipGetLastError();
    release_device(device);
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    return e;
  }
  g_live[*p] = {device, bytes};
  return hipSuccess;
}

void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_live.find(p);
  if (it == g_live.end()) {
    (void)hipFree(p);  // not from dev_malloc
    return;
  }
  g_free.insert({{it->second.device, it->second.bytes}, p});
  g_cached += it->second.bytes;
  g_live.erase(it);
}

size_t cached_bytes() {
  std::lock_guard<std::mutex> lock(g_mu);
  return g_cached;
}

void release_cached() {
  std::lock_guard<std::mutex> lock(g_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto& kv : g_free) {
    (void)hipSetDevice(kv.first.first);
    (void)hipFree(kv.second);
  }
  (void)hipSetDevice(cur);
  g_free.clear();
  g_cached = 0;
}

}  // namespace vxl
