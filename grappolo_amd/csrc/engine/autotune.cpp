// autotune.cpp -- choice of the sweep's degree-bin ranges by measurement.
//
// The sub-wave group kernels (bins 1-5), the wavefront hash (bin 6) and the
// workgroup hash (bin 7) overlap in what they can serve: a vertex of degree
// 6 runs in a group of 8 lanes, or of 16, or 32, or 64.  Which mapping is
// fastest depends on the degree distribution, so the tuner times one plain
// sweep over every vertex of the caller's graph per candidate and keeps the
// fastest: starting from the default ranges, it tries emptying each of bins
// 1..6 in turn (its degrees move to the next bin) and keeps a change when it
// saves at least 2 %.
#include <algorithm>
#include <limits>

#include "kernels.hpp"
#include "louvain.hpp"

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

}  // namespace glv
