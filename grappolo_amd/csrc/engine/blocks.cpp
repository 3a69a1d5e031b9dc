// Device block cache behind DBuf (declared in common.hpp).
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"

namespace glv {
namespace {

struct Key {
    int dev;
    hipStream_t s;
    size_t cap;
    bool operator<(const Key &o) const { return std::tie(dev, s, cap) < std::tie(o.dev, o.s, o.cap); }
};
std::mutex g_blocks_mu;
std::map<Key, std::vector<void *>> g_free;
size_t g_cached = 0;                          // bytes on the free lists
constexpr size_t kCacheLimit = size_t(32) << 30;  // beyond this, released blocks go back to the pool

size_t size_class(size_t bytes) {
    if (bytes <= 256) return 256;
    size_t top = 1;
    while (top < bytes) top <<= 1;                 // power of two >= bytes
    const size_t step = std::max<size_t>(256, top >> 4);  // 1/8 of the power of two below
    return (bytes + step - 1) / step * step;
}

int current_dev() {
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

}  // namespace

void *block_get(size_t bytes, hipStream_t s, size_t *cap) {
    const size_t c = size_class(bytes);
    const Key k{current_dev(), s, c};
    {
        std::lock_guard<std::mutex> lk(g_blocks_mu);
        auto it = g_free.find(k);
        if (it != g_free.end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            g_cached -= c;
            *cap = c;
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMallocAsync(&p, c, s);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        block_trim();
        e = hipMallocAsync(&p, c, s);
    }
    GLV_HIP(e);
    *cap = c;
    return p;
}

// Blocks are cached only under the engine's own streams, which live until
// shutdown: a caller's stream may be destroyed while blocks wait on its free
// list, so a block released on it goes back to the pool at once (the stream
// is alive while the call that used it runs).
void block_put(void *p, size_t cap, hipStream_t s) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_blocks_mu);
    if (!is_engine_stream(s) || g_cached + cap > kCacheLimit) {
        (void)hipFreeAsync(p, s);
        return;
    }
    g_free[Key{current_dev(), s, cap}].push_back(p);
    g_cached += cap;
}

namespace {
// frees the cached blocks of device dev (every device when dev < 0)
void trim_locked(int dev) {
    for (auto it = g_free.begin(); it != g_free.end();) {
        if (dev < 0 || it->first.dev == dev) {
            int was = -1;
            if (dev < 0) {
                (void)hipGetDevice(&was);
                (void)hipSetDevice(it->first.dev);
            }
            for (void *p : it->second) (void)hipFreeAsync(p, it->first.s);
            g_cached -= it->first.cap * it->second.size();
            (void)hipStreamSynchronize(it->first.s);
            if (was >= 0) (void)hipSetDevice(was);
            it = g_free.erase(it);
        } else {
            ++it;
        }
    }
}

// pinned staging buffers of Readback, one per host thread; a thread's buffer
// is valid while its generation equals g_stage_gen (shutdown frees them all)
std::mutex g_stage_mu;
std::vector<char *> g_stages;
unsigned g_stage_gen = 1;
}  // namespace

void block_trim() {
    std::lock_guard<std::mutex> lk(g_blocks_mu);
    trim_locked(current_dev());
}

void blocks_shutdown() {
    {
        std::lock_guard<std::mutex> lk(g_blocks_mu);
        trim_locked(-1);
    }
    std::lock_guard<std::mutex> lk(g_stage_mu);
    for (char *p : g_stages) (void)hipHostFree(p);
    g_stages.clear();
    g_stage_gen++;
}

void Readback::run() {
    if (items_.empty()) return;
    thread_local char *stage = nullptr;  // pinned, grows; kept until shutdown
    thread_local size_t stage_cap = 0;
    thread_local unsigned stage_gen = 0;
    size_t total = 0;
    for (const Item &it : items_) total += (it.bytes + 15) / 16 * 16;
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        if (stage_gen != g_stage_gen) {  // freed by a shutdown
            stage = nullptr;
            stage_cap = 0;
            stage_gen = g_stage_gen;
        }
        if (total > stage_cap) {
            if (stage) {
                (void)hipHostFree(stage);
                g_stages.erase(std::find(g_stages.begin(), g_stages.end(), stage));
            }
            stage = nullptr;
            stage_cap = 0;
            const size_t want = std::max<size_t>(total, 4096);
            GLV_HIP(hipHostMalloc((void **)&stage, want, hipHostMallocDefault));
            g_stages.push_back(stage);
            stage_cap = want;
        }
    }
    size_t off = 0;
    for (const Item &it : items_) {
        if (it.bytes) GLV_HIP(hipMemcpyAsync(stage + off, it.src, it.bytes, hipMemcpyDeviceToHost, s_));
        off += (it.bytes + 15) / 16 * 16;
    }
    GLV_HIP(hipStreamSynchronize(s_));
    off = 0;
    for (const Item &it : items_) {
        if (it.bytes) std::memcpy(it.dst, stage + off, it.bytes);
        off += (it.bytes + 15) / 16 * 16;
    }
    items_.clear();
}

}  // namespace glv
