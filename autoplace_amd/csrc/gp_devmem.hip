// Device memory of the library's own buffers: a cache of released blocks per device.
//
// A placement run creates and destroys engines, plans and clusterings inside its clock; hipFree waits for the device
// and unmaps the block's pages, which takes milliseconds for the buffers of a multi-million-cell design. Released
// blocks are kept instead, by size class, and handed out again to requests of the same class. Classes: multiples of
// 256 bytes up to 64 KiB; above, the next of eight equal steps between two powers of two (at most 1/8 more than the
// request). A release synchronises the device first, as hipFree does, so no work in flight still uses a block that
// is handed out again; a caller that has just synchronised the device releases without a second wait. When
// hipMalloc runs out of memory the cache is emptied and the request tried once more. With AP_DEVMEM_POISON=1 in the
// environment every block handed out is first filled with 0xff bytes (a check that no buffer is read before it is
// written, which fresh pages, zero-filled, could hide).
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "gp_common.h"

namespace ap {

namespace {

struct Block {
    int device;
    size_t bytes;   // the class
};

struct DevMem {
    std::mutex mu;
    std::unordered_map<void*, Block> live;             // handed out
    std::multimap<std::pair<int, size_t>, void*> idle;   // (device, class) -> released blocks
    size_t idle_bytes = 0;
};

bool poison() {
    static const bool on = [] {
        const char* v = std::getenv("AP_DEVMEM_POISON");
        return v && std::strcmp(v, "0") != 0 && v[0] != 0;
    }();
    return on;
}

DevMem& devmem() {
    static DevMem* m = new DevMem();   // never destroyed: the runtime may be gone at exit
    return *m;
}

size_t size_class(size_t bytes) {
    if (bytes == 0) bytes = 1;
    if (bytes <= (size_t)65536) return (bytes + 255) & ~(size_t)255;
    size_t top = (size_t)1 << (63 - __builtin_clzll((unsigned long long)bytes));   // 2^floor(log2 bytes)
    const size_t step = top >> 3;
    return (bytes + step - 1) / step * step;
}

// the idle blocks back to HIP (the caller holds the lock)
void release_idle(DevMem& m) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto& kv : m.idle) {
        (void)hipSetDevice(kv.first.first);
        (void)hipFree(kv.second);
    }
    (void)hipSetDevice(cur);
    m.idle.clear();
    m.idle_bytes = 0;
}

void give_back(void* p, bool synced) {
    if (!p) return;
    DevMem& m = devmem();
    Block b;
    {
        std::lock_guard<std::mutex> g(m.mu);
        auto it = m.live.find(p);
        if (it == m.live.end()) {          // not from dev_malloc
            (void)hipFree(p);
            return;
        }
        b = it->second;
        m.live.erase(it);
    }
    if (!synced) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (cur != b.device) (void)hipSetDevice(b.device);
        (void)hipDeviceSynchronize();
        if (cur != b.device) (void)hipSetDevice(cur);
    }
    std::lock_guard<std::mutex> g(m.mu);
    m.idle.emplace(std::make_pair(b.device, b.bytes), p);
    m.idle_bytes += b.bytes;
}

}  // namespace

hipError_t dev_malloc_block(void** p, size_t bytes);

hipError_t dev_malloc(void** p, size_t bytes) {
    const hipError_t e = dev_malloc_block(p, bytes);
    if (e == hipSuccess && poison()) {
        const hipError_t f = hipMemset(*p, 0xff, size_class(bytes));
        if (f != hipSuccess) return f;
        return hipDeviceSynchronize();
    }
    return e;
}

hipError_t dev_malloc_block(void** p, size_t bytes) {
    *p = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t cls = size_class(bytes);
    DevMem& m = devmem();
    {
        std::lock_guard<std::mutex> g(m.mu);
        auto it = m.idle.find(std::make_pair(dev, cls));
        if (it != m.idle.end()) {
            *p = it->second;
            m.idle.erase(it);
            m.idle_bytes -= cls;
            m.live.emplace(*p, Block{dev, cls});
            return hipSuccess;
        }
    }
    e = hipMalloc(p, cls);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> g(m.mu);
            release_idle(m);
        }
        e = hipMalloc(p, cls);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> g(m.mu);
    m.live.emplace(*p, Block{dev, cls});
    return hipSuccess;
}

void dev_free(void* p) { give_back(p, false); }

void dev_free_synced(void* p) { give_back(p, true); }

}  // namespace ap

extern "C" {

long long ap_cached_device_bytes(void) {
    ap::DevMem& m = ap::devmem();
    std::lock_guard<std::mutex> g(m.mu);
    return (long long)m.idle_bytes;
}

int ap_release_cached_memory(void) {
    ap::DevMem& m = ap::devmem();
    std::lock_guard<std::mutex> g(m.mu);
    ap::release_idle(m);
    return AP_OK;
}

}  // extern "C"
