// autotune.cpp -- choice of the sweep's degree-bin ranges and of the launch
// shapes by measurement.
//
// Launch shapes (Tuning): the grid of the persistent coloring rounds and the
// degree at which the coloring's core starts are timed on the coloring of
// the caller's graph; the grid of the persistent sweep (coarse phases) and of
// the whole-GPU sweep (large phases) on whole colored runs of it.  One knob
// at a time, each candidate from the best so far; a change is kept when it
// saves at least 2 %.  No knob changes a partition on integer weights (on
// real weights a sweep grid changes the order in which modularity partials
// are added, a last-bit difference).
//
// Bins:
// The sub-wave group kernels (bins 1-5), the wavefront hash (bin 6) and the
// workgroup hash (bin 7) overlap in what they can serve: a vertex of degree
// 6 runs in a group of 8 lanes, or of 16, or 32, or 64.  Which mapping is
// fastest depends on the degree distribution, so the tuner times one plain
// sweep over every vertex of the caller's graph per candidate and keeps the
// fastest: starting from the default ranges, it tries emptying each of bins
// 1..6 in turn (its degrees move to the next bin) and keeps a change when it
// saves at least 2 %.
#include <algorithm>
#include <chrono>
#include <limits>

#include "kernels.hpp"
#include "louvain.hpp"
#include "wide.hpp"

namespace glv {

double Engine::time_sweep(const DevGraph &g, const BinConfig &bins, int reps) {
    const i64 nv = g.nv;
    const BinConfig saved = bins_;
    bins_ = bins;
    PhaseSetup P;
    setup(g, nullptr, 1, P);
    bins_ = saved;
    launch_init_comm(P.k.p, nv, P.a.p, P.n.p, s_);
    DBuf<i64> curr(nv, s_), tgt(nv, s_);
    launch_iota(curr.p, nv, s_);
    i64 nparts = kRedBlocks;
    for (int b = 1; b < kNumBins; b++) nparts += sweep_blocks(b, P.bin_hi(b) - P.bin_lo(b), nv);
    DBuf<double> part(nparts, s_);
    DBuf<i64> gkeys;
    DBuf<double> gvals;
    i64 gslots = 0;
    if (P.maxdeg_global > 0) {
        gslots = 1;
        while (gslots < 2 * P.maxdeg_global) gslots <<= 1;
        const i64 nb = std::min<i64>(kGlobalHashBlocks, P.bin_hi(8) - P.bin_lo(8));
        gkeys.alloc(nb * gslots, s_);
        gvals.alloc(nb * gslots, s_);
    }
    SweepArgs A;
    A.off = g.off;
    A.tails = g.tails;
    A.w = g.w;
    A.C = curr.p;
    A.k = P.k.p;
    A.a = P.a.p;
    A.n = P.n.p;
    A.constant = P.constant;
    A.tgt = tgt.p;
    A.gkeys = gkeys.p;
    A.gvals = gvals.p;
    A.gslots = gslots;
    A.nv = nv;
    hipEvent_t e0, e1;
    GLV_HIP(hipEventCreate(&e0));
    GLV_HIP(hipEventCreate(&e1));
    double best = std::numeric_limits<double>::infinity();
    for (int r = 0; r < reps; r++) {
        GLV_HIP(hipEventRecord(e0, s_));
        i64 off = 0;
        for (int b = 1; b < kNumBins; b++) {
            const i64 lo = P.bin_lo(b), cnt = P.bin_hi(b) - lo;
            if (cnt <= 0) continue;
            A.verts = P.perm.p + lo;
            A.count = cnt;
            A.eii_part = part.p + off;
            if (b == 9) giant_args(P, 0, A);
            launch_sweep(b, true, A, s_);
            off += sweep_blocks(b, cnt, nv);
        }
        GLV_HIP(hipEventRecord(e1, s_));
        GLV_HIP(hipEventSynchronize(e1));
        float ms = 0;
        GLV_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, (double)ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

BinConfig autotune_bins(Engine &E, const DevGraph &g, int reps, double *ms_default, double *ms_tuned) {
    BinConfig best = default_bins();
    E.time_sweep(g, best, 1);  // warm caches and the allocator
    double tb = E.time_sweep(g, best, reps);
    if (ms_default) *ms_default = tb;
    for (int b = 1; b <= 6; b++) {
        BinConfig c = best;
        c.max_deg[b] = c.max_deg[b - 1];
        if (!valid_bins(c)) continue;
        const double t = E.time_sweep(g, c, reps);
        if (t < 0.98 * tb) {
            best = c;
            tb = t;
        }
    }
    if (ms_tuned) *ms_tuned = tb;
    return best;
}

namespace {
double host_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

Tuning autotune_launch(Engine &E, const DevGraph &g, const RunParams &p, int reps, bool runs,
                       std::vector<TuneTrial> *log) {
    const Tuning saved = E.tuning();
    Tuning best = saved;
    reps = std::max(reps, 1);
    DBuf<int32_t> colors(std::max<i64>(g.nv, 1), E.stream());
    auto time_color = [&](const Tuning &t) {
        E.set_tuning(t);
        E.color(g, colors.p);  // warm
        double b = std::numeric_limits<double>::infinity();
        for (int r = 0; r < reps; r++) {
            E.sync();
            const double t0 = host_ms();
            E.color(g, colors.p);
            E.sync();
            b = std::min(b, host_ms() - t0);
        }
        if (log) log->push_back({t, 0, b});
        return b;
    };
    double tb = time_color(best);
    auto try_color = [&](Tuning c) {
        if (!valid_tuning(c)) return;
        const double t = time_color(c);
        if (t < 0.98 * tb) {
            best = c;
            tb = t;
        }
    };
    for (int jb : {64, 128, 256}) {
        Tuning c = best;
        if (jb == c.jp_blocks) continue;
        c.jp_blocks = jb;
        try_color(c);
    }
    for (i64 cd : {i64(256), i64(1024), i64(4096)}) {
        Tuning c = best;
        if (cd == c.core_min_deg) continue;
        c.core_min_deg = cd;
        try_color(c);
    }
    if (runs && g.nnz > 0) {
        DBuf<i64> C(g.nv, E.stream());
        auto time_run = [&](const Tuning &t) {
            E.set_tuning(t);
            double b = std::numeric_limits<double>::infinity();
            for (int r = 0; r < reps; r++) {
                RunResult rr;
                E.sync();
                const double t0 = host_ms();
                E.run_colored(g, C.p, p, rr);
                E.sync();
                b = std::min(b, host_ms() - t0);
            }
            if (log) log->push_back({t, 1, b});
            return b;
        };
        {
            RunResult rr;
            E.set_tuning(best);
            E.run_colored(g, C.p, p, rr);  // warm
        }
        tb = time_run(best);
        auto try_run = [&](Tuning c) {
            if (!valid_tuning(c)) return;
            const double t = time_run(c);
            if (t < 0.98 * tb) {
                best = c;
                tb = t;
            }
        };
        for (int pb : {16, 32, 64, 128, 256}) {
            Tuning c = best;
            if (pb == c.persist_blocks) continue;
            c.persist_blocks = pb;
            try_run(c);
        }
        const int cus = wide_blocks();
        for (int wb : {0, cus / 2, (3 * cus) / 4}) {
            Tuning c = best;
            if (wb == c.wide_blocks || (wb == cus && c.wide_blocks == 0) || cus <= 0) continue;
            c.wide_blocks = wb;
            try_run(c);
        }
    }
    E.set_tuning(saved);
    return best;
}

}  // namespace glv
