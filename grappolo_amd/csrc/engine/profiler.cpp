// profiler.cpp -- see profiler.hpp.
#include "profiler.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>

namespace glv {

bool trace_on() {
    static const bool on = std::getenv("GLV_TRACE") != nullptr;
    return on;
}

void trace(const char *fmt, ...) {
    if (!trace_on()) return;
    va_list ap;
    va_start(ap, fmt);
    std::fputs("[glv] ", stderr);
    std::vfprintf(stderr, fmt, ap);
    std::fputc('\n', stderr);
    va_end(ap);
}

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
    // long runs record hundreds of thousands of launches: fold the older half
    // of the records once many are pending (those events completed long ago,
    // so the wait is short) instead of holding every event until the end
    if (pending_.size() >= kMaxPending) fold(pending_.size() / 2);
}

void Profiler::collect() { fold(pending_.size()); }

void Profiler::fold(size_t count) {
    for (size_t i = 0; i < count; i++) {
        Rec &r = pending_[i];
        GLV_HIP(hipEventSynchronize(r.b));
        float ms = 0;
        GLV_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        stats_[r.cls].ms += ms;
        stats_[r.cls].bytes += r.bytes;
        stats_[r.cls].launches += 1;
        free_.push_back(r.a);
        free_.push_back(r.b);
    }
    pending_.erase(pending_.begin(), pending_.begin() + count);
}

}  // namespace glv
