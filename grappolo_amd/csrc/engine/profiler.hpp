// profiler.hpp -- optional per-kernel-class timing with HIP events on the
// engine's stream.  Off by default; when on, every instrumented launch is
// bracketed by two events and its algorithmic byte count is recorded.
#pragma once

#include <vector>

#include "common.hpp"

namespace glv {

enum KClass {
    K_SWEEP_B1 = 0,  // .. K_SWEEP_B1 + 7: sweep kernels of degree bins 1..8
    K_MOD = 8,       // modularity edge pass (all bins)
    K_APPLY = 9,     // color-step / Jacobi commits
    K_COLOR = 10,    // one Jones-Plassmann round
    K_COARSEN = 11,  // whole coarsening (sorts included)
    K_RENUMBER = 12,
    K_PERSIST = 13,  // one persistent launch per colored iteration of a small graph
    K_COLOR_TAIL = 14,  // sequential first-fit of the vertices JP rounds left
    K_SWEEP_GIANT = 15,  // sweep of degree bin 9 (giant vertices, several workgroups each)
    K_WIDE = 16,         // one whole-GPU persistent launch per colored iteration of a large graph
    K_NUM = 17
};
inline int sweep_class(int bin) { return bin <= 8 ? K_SWEEP_B1 + bin - 1 : K_SWEEP_GIANT; }

struct KStat {
    double ms = 0;
    double bytes = 0;
    int64_t launches = 0;
};

class Profiler {
public:
    static Profiler &get();
    bool on() const { return on_; }
    void enable(bool v);
    void reset();
    void begin(int cls, hipStream_t s);
    void end(int cls, double bytes, hipStream_t s);
    // waits for all recorded events and folds them into the totals
    void collect();
    // turns timing off and destroys every event
    void shutdown();
    KStat stat(int cls) const { return stats_[cls]; }

private:
    struct Rec {
        int cls;
        double bytes;
        hipEvent_t a, b;
    };
    bool on_ = false;
    std::vector<Rec> pending_;
    std::vector<hipEvent_t> free_;
    hipEvent_t open_[K_NUM] = {};
    KStat stats_[K_NUM];
    hipEvent_t take();
    // folds the oldest `count` records into the totals
    void fold(size_t count);
    static constexpr size_t kMaxPending = 1 << 14;
};

}  // namespace glv
