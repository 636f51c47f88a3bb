// Device memory of the library's own tables and scratch, and the host->device
// copy of freshly built tables.
//
// An AMR regrid rebuilds levels and interface registers every few steps, and
// a serving loop builds a level per request.  hipMalloc costs tens of
// microseconds per call and hipFree waits for the whole device, so the
// library keeps freed blocks per device in size classes (four per power of
// two, at most 25 % above the request) and hands them out again; the role
// AMReX's caching arena plays for the reference's MultiFabs and tables.
// Blocks go back to HIP only in ocfd_release_cached_memory.
//
// Tables are copied through one pinned host buffer per device: a copy from
// pageable memory is staged by the runtime and waits for the kernels already
// queued on the device, which would hold a regrid's host work until the
// previous level's steps had drained.
#include <initializer_list>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ocfd_internal.hpp"

namespace ocfd {
namespace {

struct Block {
    int dev;
    size_t bytes;
};

struct Pool {
    std::mutex mu;
    std::map<int, std::multimap<size_t, void*>> free_blocks;   // device -> size class -> block
    std::unordered_map<void*, Block> live;
    std::map<int, std::pair<void*, size_t>> pinned;           // device -> staging buffer, capacity
    size_t cached = 0;
};

// never destroyed: blocks must not be handed to HIP after its own teardown
Pool& pool()
{
    static Pool* p = new Pool;
    return *p;
}

size_t size_class(size_t b)
{
    if (b <= 4096) return 4096;
    const int e = 63 - __builtin_clzll((unsigned long long)(b - 1));   // 2^e < b <= 2^(e+1)
    const size_t q = (size_t)1 << (e - 2);
    return (b + q - 1) / q * q;
}

int release_all(Pool& P)
{
    // the caller holds P.mu
    int rc = OCFD_OK;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return OCFD_EHIP;
    for (auto& [dev, fl] : P.free_blocks) {
        if (fl.empty()) continue;
        if (hipSetDevice(dev) != hipSuccess) { rc = OCFD_EHIP; continue; }
        for (auto& [sz, p] : fl) {
            if (hipFree(p) != hipSuccess) rc = OCFD_EHIP;
            P.cached -= sz;
        }
        fl.clear();
    }
    if (hipSetDevice(cur) != hipSuccess) rc = OCFD_EHIP;
    return rc;
}

}  // namespace

bool dev_alloc(void** out, size_t bytes)
{
    *out = nullptr;
    if (bytes == 0) return true;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    const size_t c = size_class(bytes);
    Pool& P = pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto& fl = P.free_blocks[dev];
        auto it = fl.find(c);
        if (it != fl.end()) {
            *out = it->second;
            fl.erase(it);
            P.cached -= c;
            P.live[*out] = Block{dev, c};
            return true;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, c) != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(P.mu);   // cached blocks of other sizes back to HIP, then once more
        release_all(P);
        if (hipMalloc(&p, c) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
    }
    std::lock_guard<std::mutex> lk(P.mu);
    P.live[p] = Block{dev, c};
    *out = p;
    return true;
}

void dev_free(std::initializer_list<void*> ps)
{
    Pool& P = pool();
    std::vector<std::pair<void*, Block>> got;
    {
        std::lock_guard<std::mutex> lk(P.mu);
        for (void* p : ps) {
            if (!p) continue;
            auto it = P.live.find(p);
            if (it == P.live.end()) continue;
            got.emplace_back(p, it->second);
            P.live.erase(it);
        }
    }
    if (got.empty()) return;
    // kernels on any stream may still read the blocks: wait for their device,
    // as hipFree does, before a block can be handed out again
    int cur = 0;
    (void)hipGetDevice(&cur);
    int synced = -1;
    for (auto& [p, b] : got)
        if (b.dev != synced) {
            if (b.dev != cur) (void)hipSetDevice(b.dev);
            (void)hipDeviceSynchronize();
            synced = b.dev;
        }
    if (synced != cur) (void)hipSetDevice(cur);
    std::lock_guard<std::mutex> lk(P.mu);
    for (auto& [p, b] : got) {
        P.free_blocks[b.dev].emplace(b.bytes, p);
        P.cached += b.bytes;
    }
}

int upload_bytes(void* dst, const void* src, size_t n)
{
    if (n == 0) return OCFD_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { set_error("upload: no device"); return OCFD_EHIP; }
    hipStream_t s = upload_stream();
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);   // one staging buffer per device, one copy at a time
    auto& pin = P.pinned[dev];
    if (pin.second < n) {
        if (pin.first) (void)hipHostFree(pin.first);
        size_t cap = std::max<size_t>(n, std::max<size_t>(2 * pin.second, (size_t)1 << 20));
        void* h = nullptr;
        if (hipHostMalloc(&h, cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            pin = {nullptr, 0};
            set_error("upload: out of pinned host memory");
            return OCFD_ENOMEM;
        }
        pin = {h, cap};
    }
    std::memcpy(pin.first, src, n);
    if (hipMemcpyAsync(dst, pin.first, n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_error("upload: copy of the tables to the device failed");
        return OCFD_EHIP;
    }
    return OCFD_OK;
}

}  // namespace ocfd

extern "C" long ocfd_cached_device_bytes(void)
{
    ocfd::Pool& P = ocfd::pool();
    std::lock_guard<std::mutex> lk(P.mu);
    return (long)P.cached;
}

extern "C" int ocfd_release_cached_memory(void)
{
    ocfd::Pool& P = ocfd::pool();
    std::lock_guard<std::mutex> lk(P.mu);
    const int rc = ocfd::release_all(P);
    if (rc) ocfd::set_error("ocfd_release_cached_memory: hipFree failed");
    return rc;
}
