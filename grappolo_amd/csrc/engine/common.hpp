// common.hpp -- shared types and helpers of the HIP engine (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "glv_status.h"

namespace glv {

using i64 = int64_t;
using u64 = uint64_t;

constexpr int kWave = 64;         // CDNA wavefront width
constexpr int kBlock = 256;       // default workgroup size (4 waves)
constexpr int kMaxColors = 4096;  // Coloring/coloringUtils.h:37 (MaxDegree)
constexpr i64 kEmpty = -1;        // empty hash slot key

// Degree bins of the sweep (see sweep kernels in kernels.hip).
//  0: isolated        -> no sweep
//  1..5: deg <= 4/8/16/32/64 -> sub-wave groups of G lanes, one vertex each
//  6: deg <= 256      -> one wavefront per vertex, LDS hash of 512 slots
//  7: deg <= 2048     -> one workgroup per vertex, LDS hash of 4096 slots
//  8: deg <= kGiantDeg -> one workgroup per vertex, hash table in HBM
//  9: larger (giant)  -> kGiantBlocks workgroups per vertex, chunks pre-summed
//                        in LDS, one HBM hash table per vertex of the class
constexpr int kNumBins = 10;
constexpr i64 kGiantDeg = 16384;
constexpr int kGiantBlocks = 64;
constexpr int kBinGroup[6] = {0, 4, 8, 16, 32, 64};
constexpr i64 kBinMaxDeg[kNumBins] = {0, 4, 8, 16, 32, 64, 256, 2048, kGiantDeg, INT64_MAX};
constexpr int kWaveHashSlots = 512;
constexpr int kBlockHashSlots = 4096;
// Graphs of at most this many vertices sweep their high-degree bins with a
// dense LDS accumulator indexed by community id: 15360 doubles (120 KB), its
// touched bitmap (1.9 KB) and a uint16 list of touched ids (30 KB) fit the
// 160 KB of LDS of a gfx950 CU.
constexpr i64 kDenseMaxNV = 15360;
constexpr int kDenseBlocks = 1024;
// grid of the persistent sweep: at most one workgroup per CU of the dense
// form, and at most kPersistMaxBlocksHash for the hash form, whose
// workgroups each keep a dense row of the graph's communities in HBM
constexpr i64 kPersistMaxBlocks = 256;
constexpr i64 kPersistMaxBlocksHash = 64;
// The persistent sweep serves graphs of at most this many vertices whose
// color classes hold at most kPersistMaxPerColor vertices on average (coarse
// phases: many small classes, where per-color launches would be latency-bound).
constexpr i64 kPersistMaxNV = 1 << 18;
constexpr i64 kPersistMaxPerColor = 256;
constexpr int kBarrierWords = 16 + 16 * (int)kPersistMaxBlocks;  // abort word + one 64-byte flag line per workgroup
// The coloring of a graph of more than 16384 vertices takes a core (its
// highest degrees, coloured by one workgroup) only when the core starts at
// this degree or above (hubs, whose priority chains make the rounds deep);
// the default of the tunable launch shapes.
constexpr i64 kJpCoreMinDeg = 1024;
// touched communities of one vertex, listed in LDS: at most min(degree, nv)
constexpr int kDenseListMax = kDenseMaxNV;

// Degree ranges of the sweep bins as used at run time: bin b takes degrees
// in (max_deg[b-1], max_deg[b]].  The upper ends of bins 1..6 may be lowered
// (a bin may become empty) by the autotuner, within the capacity of their
// kernels (kBinMaxDeg); bins 7..9 end where kBinMaxDeg says.
struct BinConfig {
    i64 max_deg[kNumBins];
};
inline BinConfig default_bins() {
    BinConfig c;
    for (int b = 0; b < kNumBins; b++) c.max_deg[b] = kBinMaxDeg[b];
    return c;
}
// true when c is a legal configuration
inline bool valid_bins(const BinConfig &c) {
    if (c.max_deg[0] != 0) return false;
    for (int b = 1; b < kNumBins; b++) {
        if (c.max_deg[b] < c.max_deg[b - 1] || c.max_deg[b] > kBinMaxDeg[b]) return false;
        if (b >= 7 && c.max_deg[b] != kBinMaxDeg[b]) return false;
    }
    return true;
}
__host__ __device__ inline int degree_bin_of(i64 d, const BinConfig &c) {
    if (d == 0) return 0;
    for (int b = 1; b < kNumBins - 1; b++)
        if (d <= c.max_deg[b]) return b;
    return kNumBins - 1;
}

__host__ __device__ inline int degree_bin(i64 d) {
    if (d == 0) return 0;
    if (d <= 4) return 1;
    if (d <= 8) return 2;
    if (d <= 16) return 3;
    if (d <= 32) return 4;
    if (d <= 64) return 5;
    if (d <= 256) return 6;
    if (d <= 2048) return 7;
    if (d <= kGiantDeg) return 8;
    return 9;
}

// Diagnostics on stderr when the environment sets GLV_TRACE (which launch
// path each stage took, refused persistent launches); silent otherwise.
bool trace_on();
void trace(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Error carried from the HIP layer up to the C ABI.
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg);

#define GLV_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            ::glv::fail(e_ == hipErrorOutOfMemory ? GLV_E_OOM : GLV_E_DEVICE,                  \
                        std::string(#call) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// Stream-ordered reuse of device blocks.  A released block goes back to the
// free list of its size class, device and stream and serves the next request
// of that class there (stream order makes the reuse safe without a
// synchronisation); blocks come from the device's default pool the first
// time.  Size classes round up to 1/8 of a power of two (at least 256 bytes).
// block_get returns the block and its class size in *cap; block_put takes
// both back.  block_trim frees the cached blocks (after an out-of-memory).
void *block_get(size_t bytes, hipStream_t s, size_t *cap);
void block_put(void *p, size_t cap, hipStream_t s);
void block_trim();
// frees every cached block on every device and the pinned staging buffers
void blocks_shutdown();
// true for a stream the engine created (those live until shutdown)
bool is_engine_stream(hipStream_t s);

// Device-to-host reads gathered behind one synchronisation: the copies go
// through a pinned staging buffer of the calling thread (into pageable
// memory every copy would wait on its own), then to their destinations.
class Readback {
  public:
    explicit Readback(hipStream_t s) : s_(s) {}
    void add(void *dst, const void *src, size_t bytes) { items_.push_back(Item{dst, src, bytes}); }
    void run();  // enqueue every copy, synchronise the stream, copy out
  private:
    struct Item {
        void *dst;
        const void *src;
        size_t bytes;
    };
    hipStream_t s_;
    std::vector<Item> items_;
};

// Stream-ordered device buffer on the block cache, so per-phase scratch
// reaches neither the driver nor the pool's bookkeeping after warm-up.  A
// pointer taken out of a DBuf (p set to null) leaves the cache for good and
// is freed with hipFreeAsync.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    size_t cap = 0;
    hipStream_t s = nullptr;
    DBuf() = default;
    DBuf(size_t count, hipStream_t stream) { alloc(count, stream); }
    void alloc(size_t count, hipStream_t stream) {
        release();
        s = stream;
        n = count;
        if (count) p = static_cast<T *>(block_get(count * sizeof(T), stream, &cap));
    }
    void release() {
        if (p) block_put(p, cap, s);
        p = nullptr;
        n = 0;
        cap = 0;
    }
    ~DBuf() { release(); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), cap(o.cap), s(o.s) {
        o.p = nullptr;
        o.n = 0;
        o.cap = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            cap = o.cap;
            s = o.s;
            o.p = nullptr;
            o.n = 0;
            o.cap = 0;
        }
        return *this;
    }
    T *get() const { return p; }
};

// Device-resident CSR graph (structure of arrays; the reference's `head`
// field is not stored).
struct DevGraph {
    i64 nv = 0, nnz = 0, num_edges = 0;
    const i64 *off = nullptr;
    const i64 *tails = nullptr;
    const double *w = nullptr;
    // storage when the engine owns the graph (coarsened levels)
    DBuf<i64> off_buf, tails_buf;
    DBuf<double> w_buf;
    void adopt() {
        off = off_buf.p;
        tails = tails_buf.p;
        w = w_buf.p;
    }
};

inline unsigned grid_for(i64 n, int block = kBlock, i64 cap = 1 << 20) {
    i64 g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace glv
