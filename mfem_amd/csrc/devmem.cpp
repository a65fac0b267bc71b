// devmem.cpp -- device blocks behind DevBuf, with an optional block cache.
//
// hipFree synchronizes the whole device.  A server that releases one
// request's objects (a few dozen buffers) while other requests' solves run
// on their own streams therefore waits, at every free, for all the work
// queued on the device.  With the cache on, freed blocks are kept per device
// and handed to later allocations of the same size class; memory goes back
// to the driver only when the cache would exceed its limit, when hipMalloc
// fails (the cache is emptied and the allocation tried again) or when the
// cache is turned off.  MFEM's counterpart is its MemoryManager's pooled
// device memory types (general/mem_manager.hpp:52-55, DEVICE_UMPIRE).
#include "internal.hpp"

#include <iterator>
#include <map>
#include <mutex>

namespace h1pa {
namespace {

constexpr int kDevices = 64;

struct BlockCache {
    std::multimap<size_t, void *> blocks;  // size -> block
    size_t bytes = 0;
};

std::mutex g_mu;
int64_t g_limit = 0;  // bytes kept per device at most; 0: cache off
BlockCache g_cache[kDevices];

// Size class: powers of two up to 1 MiB, above that multiples of 1/16 of the
// next power of two (less than 12.5% larger than asked).
size_t size_class(size_t bytes) {
    if (bytes <= (size_t(1) << 20)) {
        size_t b = 256;
        while (b < bytes) b <<= 1;
        return b;
    }
    size_t pw = size_t(1) << 20;
    while (pw < bytes) pw <<= 1;
    const size_t step = pw / 16;
    return (bytes + step - 1) / step * step;
}

void flush_locked(BlockCache &c) {
    for (auto &kv : c.blocks) (void)hipFree(kv.second);
    c.blocks.clear();
    c.bytes = 0;
}

int current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevices) {
        (void)hipGetLastError();
        dev = 0;
    }
    return dev;
}

}  // namespace

void *device_alloc(size_t bytes, size_t &cap, int &dev) {
    dev = current_device();
    void *p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (g_limit > 0) {
            BlockCache &c = g_cache[dev];
            const size_t want = size_class(bytes);
            auto it = c.blocks.lower_bound(want);
            if (it != c.blocks.end() && it->first <= want + want / 4) {
                cap = it->first;
                p = it->second;
                c.blocks.erase(it);
                c.bytes -= cap;
                return p;
            }
            cap = want;
        } else {
            cap = bytes;
        }
    }
    if (hipMalloc(&p, cap) != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lock(g_mu);
            flush_locked(g_cache[dev]);
        }
        if (hipMalloc(&p, cap) != hipSuccess) {
            (void)hipGetLastError();
            throw Error{H1PA_ENOMEM, "hipMalloc of " + std::to_string(cap) + " bytes failed"};
        }
    }
    return p;
}

void device_free(void *p, size_t cap, int dev) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (g_limit > 0 && dev >= 0 && dev < kDevices) {
            BlockCache &c = g_cache[dev];
            if ((int64_t)(c.bytes + cap) <= g_limit) {
                c.blocks.emplace(cap, p);
                c.bytes += cap;
                return;
            }
        }
    }
    (void)hipFree(p);
}

void device_cache_set(int64_t max_bytes) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_limit = max_bytes;
    for (auto &c : g_cache) {  // keep at most the new limit (largest blocks go first)
        while (!c.blocks.empty() && (int64_t)c.bytes > max_bytes) {
            auto it = std::prev(c.blocks.end());
            (void)hipFree(it->second);
            c.bytes -= it->first;
            c.blocks.erase(it);
        }
    }
}

int64_t device_cache_bytes() {
    std::lock_guard<std::mutex> lock(g_mu);
    int64_t b = 0;
    for (auto &c : g_cache) b += (int64_t)c.bytes;
    return b;
}

}  // namespace h1pa
