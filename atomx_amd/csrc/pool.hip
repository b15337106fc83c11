// Caching device allocator for the engine's own workspaces.
//
// Plans are created and destroyed per request in a serving loop; hipMalloc /
// hipFree of their dozen buffers would dominate small requests.  Freed blocks
// are kept per (device, size class) and reused; on allocation failure the
// cache is released and the allocation retried once.  Size classes: powers
// of two up to 2 MiB, multiples of 2 MiB above (served by a cached block up
// to 1/8 larger).
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "ax_common.h"

namespace ax {

namespace {

struct Block {
    size_t bytes;
    int device;
};

std::mutex g_mu;
std::multimap<std::pair<int, size_t>, void *> g_free;  // (device, class) -> block
std::unordered_map<void *, Block> g_live;
size_t g_cached = 0;

size_t size_class(size_t n) {
    if (n <= 256) return 256;
    const size_t mb2 = size_t(2) << 20;
    if (n >= mb2) return (n + mb2 - 1) / mb2 * mb2;
    size_t c = 256;
    while (c < n) c <<= 1;
    return c;
}

void release_cached_locked() {
    for (auto &kv : g_free) (void)hipFree(kv.second);
    g_free.clear();
    g_cached = 0;
}

}  // namespace

hipError_t pool_alloc(void **ptr, size_t bytes) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    const size_t cls = size_class(bytes == 0 ? 1 : bytes);
    std::lock_guard<std::mutex> lk(g_mu);
    // exact class, or for the 2 MiB-multiple classes the smallest cached
    // block at most 1/8 larger (the sizes of one-shot plans vary from set to
    // set; a miss costs a hipMalloc); a block keeps its own class
    auto it = g_free.lower_bound({dev, cls});
    const size_t slack = cls >= (size_t(2) << 20) ? cls / 8 : 0;
    if (it != g_free.end() && it->first.first == dev && it->first.second <= cls + slack) {
        const size_t got = it->first.second;
        *ptr = it->second;
        g_free.erase(it);
        g_cached -= got;
        g_live[*ptr] = Block{got, dev};
        return hipSuccess;
    }
    hipError_t e = hipMalloc(ptr, cls);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        release_cached_locked();
        e = hipMalloc(ptr, cls);
    }
    if (e != hipSuccess) return e;
    g_live[*ptr] = Block{cls, dev};
    return hipSuccess;
}

void pool_free(void *ptr) {
    if (!ptr) return;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(ptr);
    if (it == g_live.end()) {
        (void)hipFree(ptr);
        return;
    }
    g_free.insert({{it->second.device, it->second.bytes}, ptr});
    g_cached += it->second.bytes;
    g_live.erase(it);
}

size_t pool_class(size_t bytes) { return size_class(bytes == 0 ? 1 : bytes); }

// ---------------------------------------------------------------------------
// Pinned host staging for set-up uploads.  A buffer is handed out, filled by
// the host, copied asynchronously, and given back with the stream position
// (an event) after which the copy has read it; it is reused only once that
// event has completed, so set-up never waits on the host.
// ---------------------------------------------------------------------------
namespace {

struct Staging {
    void *ptr;
    size_t bytes;
    hipEvent_t ready;  // recorded after the last copy out of the buffer
    bool busy;
    void *dev;           // device mirror (null until a caller asks for one)
    int dev_device;      // the device that holds it
    hipStream_t stream;  // the stream its releases are recorded on
};

std::mutex g_smu;
std::vector<Staging> g_staging;

}  // namespace

hipError_t staging_acquire(size_t bytes, hipStream_t stream, void **ptr, int *slot, void **dev) {
    const size_t cls = size_class(bytes == 0 ? 1 : bytes);
    int device = 0;
    if (dev) (void)hipGetDevice(&device);
    std::lock_guard<std::mutex> lk(g_smu);
    auto mirror = [&](Staging &st) -> hipError_t {
        if (!dev) return hipSuccess;
        if (!st.dev) {
            const hipError_t e = hipMalloc(&st.dev, cls);
            if (e != hipSuccess) {
                st.dev = nullptr;
                return e;
            }
            st.dev_device = device;
        }
        *dev = st.dev;
        return hipSuccess;
    };
    for (size_t i = 0; i < g_staging.size(); ++i) {
        Staging &st = g_staging[i];
        if (st.busy || st.bytes != cls || st.stream != stream || !st.ptr) continue;
        if (dev && st.dev && st.dev_device != device) continue;
        const hipError_t q = hipEventQuery(st.ready);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return q;
        const hipError_t e = mirror(st);
        if (e != hipSuccess) return e;
        st.busy = true;
        *ptr = st.ptr;
        *slot = static_cast<int>(i);
        return hipSuccess;
    }
    Staging st{nullptr, cls, nullptr, true, nullptr, device, stream};
    hipError_t e = hipHostMalloc(&st.ptr, cls, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    e = hipEventCreateWithFlags(&st.ready, hipEventDisableTiming);
    if (e == hipSuccess) e = mirror(st);
    if (e != hipSuccess) {
        if (st.ready) (void)hipEventDestroy(st.ready);
        (void)hipHostFree(st.ptr);
        return e;
    }
    size_t at = 0;  // an entry freed by staging_forget_stream, else a new one
    while (at < g_staging.size() && g_staging[at].ptr) ++at;
    if (at == g_staging.size())
        g_staging.push_back(st);
    else
        g_staging[at] = st;
    *ptr = st.ptr;
    *slot = static_cast<int>(at);
    return hipSuccess;
}

hipError_t staging_release(int slot) {
    std::lock_guard<std::mutex> lk(g_smu);
    Staging &st = g_staging[static_cast<size_t>(slot)];
    const hipError_t e = hipEventRecord(st.ready, st.stream);
    st.busy = false;
    return e;
}

void staging_forget_stream(hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g_smu);
    for (Staging &st : g_staging) {
        if (st.busy || st.stream != stream || !st.ptr) continue;
        (void)hipEventSynchronize(st.ready);
        (void)hipEventDestroy(st.ready);
        (void)hipHostFree(st.ptr);
        if (st.dev) (void)hipFree(st.dev);
        st = Staging{nullptr, 0, nullptr, false, nullptr, 0, nullptr};  // slot index stays
    }
}

hipStream_t setup_stream() {
    // one non-blocking stream per (host thread, device) for plan set-up
    // copies, so that creating a plan never waits on work queued by others
    static thread_local hipStream_t streams[64] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) return nullptr;
    if (!streams[dev] && hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess)
        streams[dev] = nullptr;
    return streams[dev];
}

}  // namespace ax

extern "C" {

int64_t ax_pool_cached_bytes(void) {
    std::lock_guard<std::mutex> lk(ax::g_mu);
    return static_cast<int64_t>(ax::g_cached);
}

ax_status ax_pool_trim(void) {
    AX_CHECK_HIP(hipDeviceSynchronize());
    std::lock_guard<std::mutex> lk(ax::g_mu);
    ax::release_cached_locked();
    return AX_OK;
}

}  // extern "C"
