// gdp_devmem.cpp - the device block cache (gdp_devmem.h).
#include "gdp_devmem.h"

#include <map>
#include <mutex>
#include <unordered_map>

namespace gdp {

namespace {

constexpr size_t kCacheCap = size_t(4) << 30;

struct DevicePool {
    std::multimap<size_t, void *> free_blocks;  // size -> block
    std::unordered_map<void *, size_t> size_of;  // every live or cached block of the pool
    size_t cached = 0;
};

std::mutex g_mu;
std::unordered_map<int, DevicePool> g_pools;

void flush(DevicePool &pool) {
    for (auto &kv : pool.free_blocks) {
        (void)hipFree(kv.second);
        pool.size_of.erase(kv.second);
    }
    pool.free_blocks.clear();
    pool.cached = 0;
}

}  // namespace

hipError_t cached_malloc(void **p, size_t bytes) {
    *p = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (bytes == 0) bytes = 1;
    bytes = (bytes + 255) / 256 * 256;
    std::lock_guard<std::mutex> lk(g_mu);
    DevicePool &pool = g_pools[dev];
    auto it = pool.free_blocks.lower_bound(bytes);
    if (it != pool.free_blocks.end() && it->first <= bytes + bytes / 4) {
        *p = it->second;
        pool.cached -= it->first;
        pool.free_blocks.erase(it);
        return hipSuccess;
    }
    e = hipMalloc(p, bytes);
    if (e != hipSuccess && !pool.free_blocks.empty()) {
        (void)hipGetLastError();
        flush(pool);
        e = hipMalloc(p, bytes);
    }
    if (e == hipSuccess) pool.size_of[*p] = bytes;
    return e;
}

hipError_t cached_free(void *p) {
    if (!p) return hipSuccess;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    // no kernel or copy may still use the block when it is handed out again
    e = hipDeviceSynchronize();
    std::lock_guard<std::mutex> lk(g_mu);
    DevicePool &pool = g_pools[dev];
    auto it = pool.size_of.find(p);
    if (it == pool.size_of.end() || e != hipSuccess) {
        // not from this pool (or the device failed): the runtime's free
        if (it != pool.size_of.end()) pool.size_of.erase(it);
        const hipError_t f = hipFree(p);
        return e != hipSuccess ? e : f;
    }
    const size_t bytes = it->second;
    if (pool.cached + bytes > kCacheCap) {
        pool.size_of.erase(it);
        return hipFree(p);
    }
    pool.free_blocks.emplace(bytes, p);
    pool.cached += bytes;
    return hipSuccess;
}

size_t cached_bytes() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_pools[dev].cached;
}

}  // namespace gdp
