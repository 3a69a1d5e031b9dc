// profiler.cpp -- see profiler.hpp.
#include "profiler.hpp"

namespace glv {

Profiler &Profiler::get() {
    static Profiler p;
    return p;
}

void Profiler::enable(bool v) {
    if (!v) collect();
    on_ = v;
}

void Profiler::shutdown() {
    on_ = false;
    for (auto &r : pending_) {
        (void)hipEventSynchronize(r.b);
        free_.push_back(r.a);
        free_.push_back(r.b);
    }
    pending_.clear();
    for (auto &e : open_) {
        if (e) free_.push_back(e);
        e = nullptr;
    }
    for (hipEvent_t e : free_) (void)hipEventDestroy(e);
    free_.clear();
}

void Profiler::reset() {
    collect();
    for (auto &s : stats_) s = KStat();
}

hipEvent_t Profiler::take() {
    if (!free_.empty()) {
        hipEvent_t e = free_.back();
        free_.pop_back();
        return e;
    }
    hipEvent_t e;
    GLV_HIP(hipEventCreate(&e));
    return e;
}

void Profiler::begin(int cls, hipStream_t s) {
    if (!on_) return;
    hipEvent_t a = take();
    GLV_HIP(hipEventRecord(a, s));
    open_[cls] = a;
}

void Profiler::end(int cls, double bytes, hipStream_t s) {
    if (!on_ || !open_[cls]) return;
    hipEvent_t b = take();
    GLV_HIP(hipEventRecord(b, s));
    pending_.push_back(Rec{cls, bytes, open_[cls], b});
    open_[cls] = nullptr;
}

void Profiler::collect() {
    for (auto &r : pending_) {
        GLV_HIP(hipEventSynchronize(r.b));
        float ms = 0;
        GLV_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        stats_[r.cls].ms += ms;
        stats_[r.cls].bytes += r.bytes;
        stats_[r.cls].launches += 1;
        free_.push_back(r.a);
        free_.push_back(r.b);
    }
    pending_.clear();
}

}  // namespace glv
