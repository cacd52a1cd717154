// host_pool.hip -- recycled pinned host blocks for frames in flight and results.
//
// Host-input calls copy each frame to the device and each result back. With
// the frame (a served request) and the result (sift_result_t) in pinned
// memory, both go by one DMA, without the copy through the context's pinned
// staging that pageable buffers need. Blocks come in power-of-two size classes
// from 64 KiB and are kept for reuse once released (sift_result_free, or the
// end of a served request); the pool holds at most kPoolLimit bytes, in use or
// free (SIFT_HIP_HOST_POOL_BYTES overrides the limit; 0 turns the pool off).
// Past that, pinned_alloc returns nullptr and callers use pageable memory and
// the staging path.
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include <mutex>
#include <unordered_map>
#include <vector>

#include "sift_context.h"

namespace sh {

namespace {

constexpr int kMinShift = 16;                    // smallest block: 64 KiB
constexpr int kClasses = 13;                     // largest block: 256 MiB
constexpr size_t kPoolLimit = (size_t)2 << 30;   // pinned bytes held in all

struct Pool {
    std::mutex mu;
    std::vector<void*> spare[kClasses];  // released blocks per class
    std::unordered_map<void*, int> live;  // block in use -> its class
    size_t held = 0;                      // bytes of all blocks, live and spare
};

// Never destroyed: blocks stay valid until the process ends, also for results
// released by static destructors.
Pool& pool() {
    static Pool* p = new Pool();
    return *p;
}

// read on every allocation, so that a process can lower it (tests of the
// pageable path)
size_t pool_limit() {
    const char* e = getenv("SIFT_HIP_HOST_POOL_BYTES");
    return e ? (size_t)strtoull(e, nullptr, 10) : kPoolLimit;
}

// Gives every spare block back to the system.
void drop_spares(Pool& P) {
    for (int k = 0; k < kClasses; k++) {
        for (void* q : P.spare[k]) {
            (void)hipHostFree(q);
            P.held -= (size_t)1 << (k + kMinShift);
        }
        P.spare[k].clear();
    }
}

int class_of(size_t bytes) {
    int c = 0;
    while (c < kClasses && ((size_t)1 << (c + kMinShift)) < bytes) c++;
    return c;
}

}  // namespace

void* pinned_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    const int c = class_of(bytes);
    if (c >= kClasses) return nullptr;
    const size_t sz = (size_t)1 << (c + kMinShift);
    const size_t limit = pool_limit();
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    if (P.held > limit) drop_spares(P);  // the limit was lowered
    void* p = nullptr;
    if (!P.spare[c].empty()) {
        p = P.spare[c].back();
        P.spare[c].pop_back();
    } else {
        if (P.held + sz > limit) {
            drop_spares(P);  // the spare blocks of the other classes go first
            if (P.held + sz > limit) return nullptr;
        }
        if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) return nullptr;
        P.held += sz;
    }
    P.live[p] = c;
    return p;
}

bool pinned_release(void* p) {
    if (!p) return false;
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(p);
    if (it == P.live.end()) return false;
    P.spare[it->second].push_back(p);
    P.live.erase(it);
    return true;
}

}  // namespace sh
