// louvain.cpp -- host orchestration of the engine (see louvain.hpp).
#include "louvain.hpp"
#include "coloring_modes.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

#include "active_set.hpp"
#include "kernels.hpp"
#include "profiler.hpp"
#include "wide.hpp"

namespace glv {

[[noreturn]] void fail(int code, const std::string &msg) { throw Error{code, msg}; }

BinConfig &current_bins() {
    static BinConfig c = default_bins();
    return c;
}

Tuning &current_tuning() {
    static Tuning t;
    return t;
}
bool valid_tuning(const Tuning &t) {
    return t.jp_blocks >= 1 && t.jp_blocks <= kJpPersistMaxBlocks && t.core_min_deg >= 1 &&
           t.persist_blocks >= 1 && t.persist_blocks <= kPersistMaxBlocks && t.wide_blocks >= 0;
}

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int bits_for(i64 x) {  // smallest b with 2^b > x
    int b = 0;
    while (b < 63 && ((i64)1 << b) <= x) b++;
    return b;
}

// Iterations of one phase never exceed this; the reference has no bound and
// a phase whose modularity keeps rising by at least `thresh` per iteration
// could only run longer than this with thresh below ~1e-5.
constexpr int kMaxPhaseIters = 100000;

}  // namespace

template <class T>
T Engine::fetch(const T *dptr) {
    T h;
    Readback rb(s_);
    rb.add(&h, dptr, sizeof(T));
    rb.run();
    return h;
}

void Engine::sync() { GLV_HIP(hipStreamSynchronize(s_)); }

void Engine::setup(const DevGraph &g, const int32_t *colors, int ncol, PhaseSetup &P) {
    const i64 nv = g.nv;
    P.ncol = ncol;
    const i64 nkeys = (i64)kNumBins * ncol;
    P.k.alloc(nv, s_);
    P.a.alloc(nv, s_);
    P.n.alloc(nv, s_);
    P.perm.alloc(nv, s_);
    DBuf<uint32_t> keys(nv, s_);
    DBuf<int32_t> err(1, s_);
    P.dptr.alloc(nkeys + 1, s_);
    DBuf<i64> pnnz(nkeys, s_);
    i64 *start = P.dptr.p;
    launch_fill_i32(err.p, 1, 0, s_);
    launch_bin_keys(g.off, nv, colors, ncol, keys.p, P.perm.p, err.p, s_, bins_);
    sort_pairs_u32_i64(keys.p, P.perm.p, nv, bits_for(nkeys), s_);
    launch_slices(keys.p, P.perm.p, g.off, nv, nkeys, start, pnnz.p, s_);
    P.ptr.assign(nkeys + 1, 0);
    P.pnnz.assign(nkeys, 0);
    int32_t e = 0;
    Readback rb(s_);
    rb.add(&e, err.p, sizeof(e));
    if (nv > 0) {
        rb.add(P.ptr.data(), start, (nkeys + 1) * sizeof(i64));
        rb.add(P.pnnz.data(), pnnz.p, nkeys * sizeof(i64));
    }
    rb.run();
    if (e) fail(GLV_E_INVALID_ARG, "color values out of range [0, nColors)");
    if (trace_on()) {
        for (int b = 1; b < kNumBins; b++) {
            i64 ents = 0;
            for (int c = 0; c < ncol; c++) ents += P.pnnz[(size_t)b * ncol + c];
            trace("  bin %d: %ld vertices, %ld entries", b, (long)(P.bin_hi(b) - P.bin_lo(b)), (long)ents);
        }
    }
    // degrees and 1/(2m)
    const i64 big_lo = P.bin_lo(8), big_hi = P.bin_hi(9);
    launch_degree(g, P.perm.p + big_lo, big_hi - big_lo, P.k.p, s_);
    DBuf<double> part(kRedBlocks, s_), res(1, s_);
    launch_sum(P.k.p, nv, part.p, res.p, s_);
    P.total_weight = fetch(res.p);
    P.constant = (double)1 / P.total_weight;
    P.maxdeg_global = 0;
    if (P.bin_hi(8) > big_lo) {
        DBuf<i64> md(1, s_);
        launch_fill_i64(md.p, 1, 0, s_);
        launch_max_degree(g.off, P.perm.p + big_lo, P.bin_hi(8) - big_lo, md.p, s_);
        P.maxdeg_global = fetch(md.p);
    }
    // giant vertices: class-relative table offsets and the table memory
    const i64 g_lo = P.bin_lo(9), ng = P.bin_hi(9) - g_lo;
    P.gt_seg.assign(ncol, 0);
    if (ng > 0 && nv > kDenseMaxNV) {
        DBuf<i64> sz(ng, s_);
        launch_giant_table_sizes(g.off, P.perm.p + g_lo, ng, sz.p, s_);
        std::vector<i64> hsz(ng);
        GLV_HIP(hipMemcpyAsync(hsz.data(), sz.p, ng * sizeof(i64), hipMemcpyDeviceToHost, s_));
        GLV_HIP(hipStreamSynchronize(s_));
        std::vector<i64> tab;
        tab.reserve(ng + ncol);
        i64 most = 0;
        for (int c = 0; c < ncol; c++) {
            P.gt_seg[c] = (i64)tab.size();
            i64 at = 0;
            tab.push_back(0);
            for (i64 i = P.lo(9, c) - g_lo; i < P.hi(9, c) - g_lo; i++) {
                at += hsz[i];
                tab.push_back(at);
            }
            most = std::max(most, at);
        }
        if (trace_on()) {
            i64 mx = 0;
            for (i64 x : hsz) mx = std::max(mx, x);
            trace("  giants: %ld, largest table %ld slots, tables of one class at most %ld slots", (long)ng,
                  (long)mx, (long)most);
        }
        P.gt_tab.alloc(tab.size(), s_);
        GLV_HIP(hipMemcpyAsync(P.gt_tab.p, tab.data(), tab.size() * sizeof(i64), hipMemcpyHostToDevice, s_));
        P.gt_keys.alloc(most, s_);
        P.gt_vals.alloc(most, s_);
        launch_fill_i64(P.gt_keys.p, most, kEmpty, s_);
        launch_fill_f64(P.gt_vals.p, most, 0.0, s_);
        P.gt_sc.alloc(2 * ng, s_);
        launch_fill_f64(P.gt_sc.p, 2 * ng, 0.0, s_);
        P.gt_rg.alloc(ng * kGiantBlocks, s_);
        P.gt_ri.alloc(ng * kGiantBlocks, s_);
        P.gt_rt.alloc(2 * ng * kGiantBlocks, s_);
        GLV_HIP(hipStreamSynchronize(s_));  // `tab` is pageable host memory
    }
}

// Sweep arguments of the giant vertices of color class c (bin 9).
void giant_args(const PhaseSetup &P, int c, SweepArgs &A) {
    if (P.gt_tab.p == nullptr) return;
    const i64 qb = P.lo(9, c) - P.bin_lo(9);
    A.gt_tab = P.gt_tab.p + P.gt_seg[c];
    A.gt_keys = P.gt_keys.p;
    A.gt_vals = P.gt_vals.p;
    A.gt_sc = P.gt_sc.p + 2 * qb;
    A.gt_rg = P.gt_rg.p + qb * kGiantBlocks;
    A.gt_ri = P.gt_ri.p + qb * kGiantBlocks;
    A.gt_rt = P.gt_rt.p + 2 * qb * kGiantBlocks;
}

// ----------------------------------------------------------------- coloring
// Jones-Plassmann with degree-first priorities and a sequential tail (the
// engine's own order; the reference's coloring, coloringDistanceOne.cpp:50-233,
// is speculative and thread-dependent).  The core (the highest degrees) is
// coloured first by one workgroup, the rest by frontier rounds; depths give
// the tail (kernels.hpp, coloring).
int Engine::color(const DevGraph &g, int32_t *colors, int *rounds_out) {
    const i64 nv = g.nv;
    if (nv == 0) return -1;
    launch_fill_i32(colors, nv, -1, s_);
    const int pnb = (int)std::min<i64>({(i64)jp_persistent_blocks(), (i64)tune_.jp_blocks,
                                         std::max<i64>(4, nv / 2048)});
    DBuf<i64> ring(10, s_);
    DBuf<unsigned> bar(kJpBarrierWords, s_);
    DBuf<int32_t> cnt(nv, s_), dep(nv, s_), err(1, s_), maxc(1, s_);
    DBuf<u64> prio(nv, s_);
    DBuf<uint8_t> code(std::max<i64>(g.nnz, 1), s_);
    launch_jp_prio(g, prio.p, s_);
    launch_fill_i32(err.p, 1, 0, s_);
    launch_fill_i32(maxc.p, 1, -1, s_);
    DBuf<i64> sA(nv, s_), sB(nv, s_), lA(nv, s_), lB(nv, s_), gA(nv, s_), gB(nv, s_), counts(8, s_);
    JpLists cur{sA.p, lA.p, gA.p, counts.p, prio.p, code.p}, next{sB.p, lB.p, gB.p, counts.p + 4, prio.p, code.p};
    cur.dep = next.dep = dep.p;

    // the core: every vertex of a small graph; on a larger one the vertices
    // of the largest degrees, at most kJpCoreMax, when they reach the tuned
    // core degree (kJpCoreMinDeg by default)
    u64 dcore = kJpNoCore;
    i64 K = 0;
    if (nv <= kJpCoreMax) {
        dcore = 0;
        K = nv;
    } else {
        DBuf<unsigned long long> hist(kJpDegBuckets, s_);
        launch_fill_i64((i64 *)hist.p, kJpDegBuckets, 0, s_);
        launch_jp_degree_hist(g, hist.p, s_);
        std::vector<unsigned long long> h(kJpDegBuckets);
        Readback rb(s_);
        rb.add(h.data(), hist.p, kJpDegBuckets * sizeof(unsigned long long));
        rb.run();
        i64 above = 0;
        for (int b = kJpDegBuckets - 1; b > 0; b--) {
            if (above + (i64)h[b] > kJpCoreMax) break;
            above += (i64)h[b];
            const i64 lo = jp_bucket_lo(b);
            if (lo >= tune_.core_min_deg && lo < (i64(1) << 23) && above > 0) {
                dcore = (u64)lo;
                K = above;
            }
        }
    }
    DBuf<i64> hlist, loff;
    DBuf<int32_t> core_cnt;
    if (K > 0) {
        hlist.alloc(K, s_);
        loff.alloc(K + 1, s_);
        core_cnt.alloc(2 * K, s_);
        launch_jp_core_list(g, dcore, gA.p, gB.p, hlist.p, cnt.p, s_);
        cur.dcore = next.dcore = dcore;
        cur.core_cnt = next.core_cnt = core_cnt.p;
    }
    // counts, codes and the first frontier of the non-core vertices; core
    // neighbour counts of the core vertices
    i64 hc[3] = {0, 0, 0};  // sizes of the current frontier's small, large and giant lists
    launch_fill_i64(counts.p, 8, 0, s_);
    Profiler::get().begin(K_COLOR, s_);
    // heavy and giant lists in the next round's buffers
    launch_jp_count(g, cnt.p, cur, next.small, cur.counts + 3, next.large, next.counts + 3, s_);
    // algorithmic bytes of the whole coloring, credited here once: per directed
    // entry 17 in the count (tail, the neighbour's priority, the entry's code)
    // and 13 when its vertex is coloured (tail, code, the neighbour's colour or
    // count), per vertex 20 (offsets, priority, count); the tail's first-fit
    // over at most kJpTail vertices is left out
    Profiler::get().end(K_COLOR, 30.0 * g.nnz + 20.0 * nv, s_);
    i64 core_entries = 0;
    if (K > 0) {
        DBuf<i64> sizes(K + 1, s_);  // (the giant lists hold the count's frontier)
        launch_jp_core_sizes(core_cnt.p, K, sizes.p, s_);
        scan_exclusive_i64(sizes.p, loff.p, K + 1, s_);
    }
    {
        Readback rb(s_);
        rb.add(hc, cur.counts, sizeof(hc));
        if (K > 0) rb.add(&core_entries, loff.p + K, sizeof(i64));
        rb.run();
    }
    i64 colored = 0;
    if (K > 0) {
        DBuf<uint16_t> ladj(std::max<i64>(core_entries, 1), s_);
        Profiler::get().begin(K_COLOR, s_);
        launch_jp_core_fill(g, code.p, cnt.p, hlist.p, K, loff.p, core_cnt.p, ladj.p, s_);
        // a core that is the whole graph stops where the tail starts
        launch_jp_core(hlist.p, K, loff.p, core_cnt.p, ladj.p, colors, dep.p, err.p, K == nv ? (int)kJpTail : -1,
                       s_);
        Profiler::get().end(K_COLOR, 0.0, s_);
        colored = K;
    }

    // Rounds of the rest in one persistent launch while the frontier does not
    // fill the GPU: continue from the frontier in cur (hc, no giant list).
    // Returns 0 when the launch was refused, 1 when every vertex is coloured,
    // 2 when a large frontier handed it back (cur, hc and colored then
    // describe the state).
    int prounds = 0;
    auto persistent = [&]() -> int {
        i64 h[10] = {hc[0], hc[1], 0, 0, 0, 0, 0, 0, 0, 0};
        GLV_HIP(hipMemcpyAsync(ring.p, h, sizeof(h), hipMemcpyHostToDevice, s_));
        launch_fill_i32((int32_t *)bar.p, kJpBarrierWords, 0, s_);
        Profiler::get().begin(K_COLOR, s_);
        const bool ok = pnb > 0 && launch_jp_persistent(g, colors, cnt.p, cur.small, next.small, cur.large,
                                                        next.large, ring.p, bar.p, err.p, colored, prio.p, code.p,
                                                        dep.p, kJpPersistExit, pnb, s_);
        Profiler::get().end(K_COLOR, 0.0, s_);
        unsigned abort_flag = 0;
        Readback rb(s_);
        rb.add(h, ring.p, sizeof(h));
        rb.add(&abort_flag, bar.p, sizeof(unsigned));
        rb.run();
        if (!ok) return 0;
        if (abort_flag) fail(GLV_E_DEVICE, "grid barrier of the persistent coloring timed out");
        const int r = (int)h[6];
        prounds += r;
        colored = h[8];
        if (h[7] == 0) return 1;
        hc[0] = h[2 * (r % 3)];
        hc[1] = h[2 * (r % 3) + 1];
        hc[2] = 0;
        if (r % 2 == 1) std::swap(cur, next);  // the frontier sits in the second pair of lists
        return 2;
    };
    {
        DBuf<uint32_t> gmark;
        while (hc[0] + hc[1] + hc[2] > 0) {
            if (prounds >= (1 << 24)) fail(GLV_E_DEVICE, "coloring did not converge");
            const i64 won = hc[0] + hc[1] + hc[2];
            if (won < kJpPersistFrontier) {
                // hand the following rounds to one persistent launch: giant
                // frontier vertices join the large list
                if (hc[2] > 0)
                    GLV_HIP(hipMemcpyAsync(cur.large + hc[1], cur.giant, hc[2] * sizeof(i64),
                                           hipMemcpyDeviceToDevice, s_));
                hc[1] += hc[2];
                hc[2] = 0;
                const int st = persistent();
                if (st == 1) break;
                if (st == 2) continue;
            }
            const i64 words = hc[2] * (kMaxColors / 32 + 1);
            if (hc[2] > 0 && (i64)gmark.n < words) gmark.alloc(words, s_);
            launch_fill_i64(next.counts, 4, 0, s_);
            Profiler::get().begin(K_COLOR, s_);
            launch_jp_frontier(g, colors, cnt.p, cur, hc, next, gmark.p, err.p, s_);
            Profiler::get().end(K_COLOR, 0.0, s_);
            Readback rb(s_);
            rb.add(hc, next.counts, sizeof(hc));
            rb.run();
            colored += won;
            std::swap(cur, next);
            prounds++;
        }
    }
    if (colored != nv) fail(GLV_E_DEVICE, "coloring left vertices unreached (is the adjacency symmetric?)");

    // the tail: the smallest round r >= 1 after which at most kJpTail
    // vertices remain, i.e. with at most kJpTail vertices deeper than r
    int32_t maxd = 0;
    {
        DBuf<int32_t> md(1, s_);
        launch_fill_i32(md.p, 1, 0, s_);
        launch_jp_depth_max(dep.p, nv, md.p, s_);
        maxd = fetch(md.p);
    }
    int32_t rounds = maxd;
    i64 deeper = 0;  // vertices deeper than rounds
    {
        DBuf<unsigned> dh(kJpDepthWindow, s_);
        std::vector<unsigned> h(kJpDepthWindow);
        bool done = false;
        for (int32_t top = maxd; top >= 1 && !done; top -= kJpDepthWindow) {
            launch_fill_i32((int32_t *)dh.p, kJpDepthWindow, 0, s_);
            launch_jp_depth_hist(dep.p, nv, top, dh.p, s_);
            Readback rb(s_);
            rb.add(h.data(), dh.p, kJpDepthWindow * sizeof(unsigned));
            rb.run();
            for (int i = 0; i < kJpDepthWindow && !done; i++) {
                const int32_t d = top - i;  // moving the cut from d to d - 1 adds the vertices of depth d
                if (d - 1 < 1 || deeper + (i64)h[i] > kJpTail) {
                    done = true;
                } else {
                    deeper += h[i];
                    rounds = d - 1;
                }
            }
        }
    }
    if (deeper > 0) {
        launch_jp_depth_cut(dep.p, nv, rounds, colors, s_);
        DBuf<uint32_t> base(deeper * (kMaxColors / 32), s_);
        DBuf<u64> adj(deeper * (kJpTail / 64), s_);
        Profiler::get().begin(K_COLOR_TAIL, s_);
        launch_color_tail(g, colors, next.small, next.large, next.giant, deeper, cnt.p, base.p, adj.p, err.p, s_);
        Profiler::get().end(K_COLOR_TAIL, 0.0, s_);
    }
    glv::trace("coloring nv %ld: core %ld (degree >= %ld), %d rounds after it, depth %d, %d rounds, tail %ld",
               (long)nv, (long)K, K > 0 ? (long)dcore : -1L, prounds, maxd, rounds, (long)deeper);
    launch_jp_finish(colors, nv, maxc.p, s_);
    const int32_t e = fetch(err.p);
    if (e & 1) fail(GLV_E_TOO_MANY_COLORS, "distance-1 coloring needs more than 4096 colors");
    if (e & 4) fail(GLV_E_DEVICE, "core coloring did not complete (is the adjacency symmetric?)");
    if (e) fail(GLV_E_DEVICE, "coloring tail overflow");
    if (rounds_out) *rounds_out = rounds;
    return fetch(maxc.p);
}

i64 Engine::count_conflicts(const DevGraph &g, const int32_t *colors) {
    DBuf<i64> c(1, s_);
    launch_fill_i64(c.p, 1, 0, s_);
    launch_color_conflicts(g, colors, c.p, s_);
    return fetch(c.p) / 2;
}

// --------------------------------------------------------------- modularity
// e_xx over all vertices (per-bin passes into block partials), sum of a_c^2.
double Engine::colored_modularity(const DevGraph &g, const PhaseSetup &P, const i64 *C, double *part,
                                  double *res, double *exx) {
    // algorithmic bytes: tail, weight and C[tail] per entry; C[v] per vertex; a_c per community
    Profiler::get().begin(K_MOD, s_);
    i64 off = 0;
    for (int b = 1; b < kNumBins; b++) {
        const i64 lo = P.bin_lo(b), cnt = P.bin_hi(b) - lo;
        if (cnt <= 0) continue;
        launch_mod_pass(b, g, C, P.perm.p + lo, cnt, part + off, s_);
        off += mod_blocks(b, cnt);
    }
    if (off == 0) launch_fill_f64(res, 1, 0.0, s_);
    else launch_sum_partials(part, off, res, s_);
    launch_sumsq(P.a.p, g.nv, part + off, res + 1, s_);
    Profiler::get().end(K_MOD, 24.0 * g.nnz + 16.0 * g.nv, s_);
    double h[2];
    Readback rb(s_);
    rb.add(h, res, sizeof(h));
    rb.run();
    if (exx) *exx = h[0];
    const double c = P.constant;
    return h[0] * (double)c - h[1] * (double)c * (double)c;
}

static i64 mod_part_count(const PhaseSetup &P) {
    i64 n = 0;
    for (int b = 1; b < kNumBins; b++) n += mod_blocks(b, P.bin_hi(b) - P.bin_lo(b));
    return n + kRedBlocks;
}

// ----------------------------------------------------------- colored phase
// algoLouvainWithDistOneColoring, Coloring/parallelLouvainWithColoring.cpp:47-286.
// With a learned active set (las), see LasParams.
double Engine::colored_phase(const DevGraph &g, i64 *C, const int32_t *colors, int ncol, double Lower,
                             double thresh, int *iters, std::vector<double> *trace, LasParams *las,
                             int phase_index, bool valid_colors) {
    const i64 nv = g.nv;
    PhaseSetup P;
    setup(g, colors, ncol, P);
    launch_init_labels(g.off, nv, C, s_);
    launch_init_comm(P.k.p, nv, P.a.p, P.n.p, s_);
    DBuf<i64> tgt(nv, s_);
    DBuf<double> part(mod_part_count(P), s_), res(2, s_);

    std::vector<char> color_nonempty(ncol, 0);
    i64 max_class = 0;
    for (int c = 0; c < ncol; c++) {
        i64 size = 0;
        for (int b = 1; b < kNumBins; b++) size += P.hi(b, c) - P.lo(b, c);
        color_nonempty[c] = size > 0;
        max_class = std::max(max_class, size);
    }
    // coarse graphs (small, or many small color classes): one persistent
    // launch per iteration (all colors + modularity), grid sized to the color
    // classes: every extra workgroup only lengthens the barriers
    int pblocks = 0;
    if (nv <= kPersistMaxNV) {
        int nonempty = 0;
        for (int c = 0; c < ncol; c++) nonempty += color_nonempty[c];
        const i64 per_class = (nv + std::max(nonempty, 1) - 1) / std::max(nonempty, 1);
        if (nv <= kDenseMaxNV || per_class <= kPersistMaxPerColor) {
            pblocks = persistent_blocks();
            // dense graphs: the grid's two halves take the color steps in turn
            // (an even count; twice the mean class size per half, or a
            // workgroup per vertex of the largest class, whose step is a
            // chain of ceil(size / half) vertices otherwise)
            const bool dense = nv <= kDenseMaxNV;
            const i64 want = dense ? std::max(4 * per_class, 2 * max_class) : 2 * per_class;
            if (pblocks > 0) {
                pblocks = (int)std::min<i64>({(i64)pblocks, (i64)tune_.persist_blocks, std::max<i64>(4, want),
                                              dense ? kPersistMaxBlocks : kPersistMaxBlocksHash});
                if (dense) pblocks = std::max(2, pblocks & ~1);
            }
        }
    }
    // large graphs: the whole-GPU persistent sweep (wide.hpp), plain form only
    const bool las_mode = las && !(las->mode == 2 && nv <= kDenseMaxNV) && las->mode != 0;
    int wblocks = 0;
    {
        int nonempty = 0;
        for (int c = 0; c < ncol; c++) nonempty += color_nonempty[c];
        // classes of more than kWideMaxPerColor vertices on average fill the GPU
        // with per-color launches of the binned kernels
        const bool small_classes = nonempty > 0 && nv / nonempty <= kWideMaxPerColor;
        if (!las_mode && pblocks == 0 && nv > kPersistMaxNV && small_classes && wide_enabled()) {
            wblocks = wide_blocks();
            if (tune_.wide_blocks > 0) wblocks = std::min(wblocks, tune_.wide_blocks);
        }
    }
    bool use_wide = wblocks > 0;
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
    A.C = C;
    A.k = P.k.p;
    A.a = P.a.p;
    A.n = P.n.p;
    A.constant = P.constant;
    A.tgt = tgt.p;
    A.eii_part = nullptr;
    A.gkeys = gkeys.p;
    A.gvals = gvals.p;
    A.gslots = gslots;
    A.nv = nv;
    DBuf<double> dacc(pblocks > 0 && nv > kDenseMaxNV ? (size_t)pblocks * nv : 0, s_);
    if (dacc.n) {
        launch_fill_f64(dacc.p, (i64)dacc.n, 0.0, s_);
        A.dacc = dacc.p;
    }
    DBuf<unsigned> bar(pblocks > 0 ? kBarrierWords : 0, s_);
    DBuf<double> ppart(pblocks > 0 ? 2 * pblocks : 0, s_);
    if (use_wide && (nv >= (i64(1) << 31) || g.nnz >= (i64(1) << 40))) use_wide = false;  // 32-bit ids
    const bool dbl = pblocks > 0 || use_wide;
    DBuf<i64> cperm(dbl ? nv : 0, s_), cstart(dbl ? ncol + 1 : 0, s_);
    // second state buffers of the persistent sweep (one grid barrier per color step)
    DBuf<i64> pC1(pblocks > 0 ? nv : 0, s_), pn1(pblocks > 0 ? nv : 0, s_);
    DBuf<double> pa1(pblocks > 0 ? nv : 0, s_);
    DBuf<uint16_t> pl16(pblocks >= 2 && nv <= kDenseMaxNV ? 2 * nv : 0, s_);  // the dense form's 16-bit labels
    if (dbl) launch_color_major(P.perm.p, colors, nv, ncol, cperm.p, cstart.p, s_);
    bool persistent = pblocks > 0;
    WidePlan WP;
    WideArgs W{};
    DBuf<unsigned> wbar(use_wide ? 16 + 16 * (size_t)wblocks : 0, s_);
    DBuf<double> wpart(use_wide ? 2 * (size_t)wblocks : 0, s_);
    DBuf<i64> wmoves(use_wide ? 1 : 0, s_);
    DBuf<CommInfo> wm0(use_wide ? nv : 0, s_), wm1(use_wide ? nv : 0, s_);
    if (use_wide) use_wide = build_wide_plan(g, P, wblocks, WP, s_);
    // the wide sweep's 32-bit labels (two buffers) and neighbour ids; while
    // it runs, the phase's labels live in wc0 (copied back to C when a full
    // modularity pass or the end of the phase needs them)
    DBuf<int32_t> wc0(use_wide ? nv : 0, s_), wc1(use_wide ? nv : 0, s_), wtails(use_wide ? g.nnz : 0, s_);
    bool labels32 = false;
    if (use_wide) {
        launch_narrow_ids(C, nv, wc0.p, s_);
        launch_narrow_ids(g.tails, g.nnz, wtails.p, s_);
        labels32 = true;
        launch_fill_i64(wmoves.p, 1, 0, s_);
        W.A = A;
        W.C0 = wc0.p;
        W.C1 = wc1.p;
        W.tails = wtails.p;
        W.a = P.a.p;
        W.n = P.n.p;
        W.m0 = wm0.p;
        W.m1 = wm1.p;
        W.perm = P.perm.p;
        W.cperm = cperm.p;
        W.cstart = cstart.p;
        W.witems = WP.witems.p;
        W.wstart = WP.wstart.p;
        W.vitems = WP.vitems.p;
        W.vstart = WP.vstart.p;
        W.giants = WP.giants.p;
        W.ngiants = WP.ngiants;
        W.gpk = WP.gpk.p;
        W.gpv = WP.gpv.p;
        W.gcnt = WP.gcnt.p;
        W.gstate = WP.gstate.p;
        W.gsc = WP.gsc.p;
        W.gcand = WP.gcand.p;
        W.gcid = WP.gcid.p;
        W.gparts = WP.gparts;
        W.gout = WP.gout.p;
        W.ncol = ncol;
        W.bar = wbar.p;
        W.part = wpart.p;
        W.moves = wmoves.p;
    }
    glv::trace("colored phase nv %ld, %d colors: %s (%d workgroups)", (long)nv, ncol,
               use_wide ? "wide persistent sweep" : persistent ? "persistent sweep" : "per-color launches",
               use_wide ? wblocks : pblocks);

    // learned active set: recorded features, sweeps since the last move,
    // active flags and the compacted sweep lists
    // Pruning skips the phases of the dense persistent sweep: there a color
    // step costs its heaviest vertex and a grid barrier, so sweeping fewer
    // vertices saves nothing while the re-check costs an iteration; such a
    // phase runs as the plain form (and counts as full sweeps).
    const bool las_off = las && las->mode == 2 && nv <= kDenseMaxNV;
    const int mode = las && !las_off ? las->mode : 0;
    const bool prune = mode == 2;
    DBuf<float4> feat(mode ? nv : 0, s_);
    DBuf<uint8_t> stale(mode ? nv : 0, s_), active(prune ? nv : 0, s_);
    DBuf<i64> flags(prune ? nv : 0, s_), fpos(prune ? nv : 0, s_), list2(prune ? nv : 0, s_);
    const i64 nkeys = (i64)kNumBins * ncol;
    DBuf<i64> start2(prune ? std::max<i64>(nkeys, ncol) + 1 : 0, s_);
    DBuf<i64> lcount(mode ? 2 : 0, s_);
    // movers' internal-weight deltas: pruned sweeps, and the sweeps after the
    // first with a valid coloring and no giant vertex (whose decision does not
    // keep the winner's weight)
    bool has_giant = false;
    for (int c = 0; c < ncol; c++) has_giant = has_giant || P.hi(9, c) > P.lo(9, c);
    const bool exact_incr = valid_colors && !has_giant && !prune;
    const bool track = prune || exact_incr;
    DBuf<double> dein(track ? nv : 0, s_), dsum(track ? 1 : 0, s_), dpart(track ? kApplyMaxBlocks : 0, s_);
    std::vector<i64> ptr2;
    LasView V;
    if (mode) {
        launch_fill_i32((int32_t *)feat.p, 4 * nv, 0, s_);
        GLV_HIP(hipMemsetAsync(stale.p, 0, nv, s_));
        A.feat = feat.p;
        V.off = g.off;
        V.feat = feat.p;
        V.stale = stale.p;
        V.colors = colors;
        V.ncol = ncol;
        V.C = C;
        V.n = P.n.p;
        V.phase = phase_index;
        V.nv = nv;
    }
    i64 nonisolated = 0;
    for (int b = 1; b < kNumBins; b++) nonisolated += P.bin_hi(b) - P.bin_lo(b);

    double prevMod = Lower, currMod = -1;
    double exx = 0;    // internal weight of the last exact modularity pass
    bool full = true;  // this iteration sweeps every vertex
    int rechecks = 0;  // full iterations that followed a pruned one
    int numItrs = 0;
    while (true) {
        numItrs++;
        // trace capture: features of the previous sweep, labelled after this one
        i64 tlo = 0, thi = 0;
        if (mode == 1 && numItrs >= 2 && las->tfeats && las->count < las->cap) {
            tlo = las->count;
            GLV_HIP(hipMemcpyAsync(lcount.p, &tlo, sizeof(i64), hipMemcpyHostToDevice, s_));
            const u64 seed = las->seed + 0x100000000ULL * (u64)phase_index + 0x10000ULL * (u64)numItrs;
            launch_las_capture(V, seed, las->rate, las->tfeats, las->tverts, lcount.p, las->cap, s_);
            las->count = fetch(lcount.p);
            thi = std::min(las->count, las->cap);
        }
        const i64 *plist = persistent ? cperm.p : P.perm.p;
        const i64 *pstart = cstart.p;
        const std::vector<i64> *hptr = &P.ptr;
        if (prune && !full) {
            GLV_HIP(hipMemsetAsync(lcount.p, 0, 2 * sizeof(i64), s_));
            launch_las_infer(V, las->P, las->logit_tau, active.p, lcount.p, s_);
            if (persistent) {
                launch_compact_active(cperm.p, nv, active.p, cstart.p, ncol, flags.p, fpos.p, list2.p,
                                      start2.p, s_);
                i64 hc[2];
                Readback rb(s_);
                rb.add(hc, lcount.p, sizeof(hc));
                rb.run();
                las->swept_vertices += hc[0];
                las->swept_entries += hc[1];
                pstart = start2.p;
            } else {
                launch_compact_active(P.perm.p, nv, active.p, P.dptr.p, nkeys, flags.p, fpos.p, list2.p,
                                      start2.p, s_);
                ptr2.resize(nkeys + 1);
                i64 hc[2];
                Readback rb(s_);
                rb.add(ptr2.data(), start2.p, (nkeys + 1) * sizeof(i64));
                rb.add(hc, lcount.p, sizeof(hc));
                rb.run();
                las->swept_vertices += hc[0];
                las->swept_entries += hc[1];
                hptr = &ptr2;
                GLV_HIP(hipMemsetAsync(dsum.p, 0, sizeof(double), s_));
            }
            plist = list2.p;
            las->pruned_sweeps++;
        } else if (prune) {
            las->swept_vertices += nonisolated;
            las->swept_entries += g.nnz;
            las->full_sweeps++;
        }
        if (mode) launch_las_tick(stale.p, nv, s_);
        const bool incremental = ((prune && !full) || (exact_incr && numItrs > 1)) && !persistent;
        A.dein = incremental ? dein.p : nullptr;
        if (incremental && !prune) launch_fill_f64(dsum.p, 1, 0.0, s_);
        bool done_iter = false;
        if (use_wide) {
            launch_fill_i32((int32_t *)wbar.p, (i64)wbar.n, 0, s_);
            Profiler::get().begin(K_WIDE, s_);
            use_wide = launch_wide_iteration(W, wblocks, s_);
            Profiler::get().end(K_WIDE, 24.0 * g.nnz + 32.0 * nv, s_);
            if (use_wide) {
                launch_sum_partials(wpart.p, wblocks, res.p, s_);
                launch_sum_partials(wpart.p + wblocks, wblocks, res.p + 1, s_);
                double h[2];
                unsigned abort_flag = 0;
                Readback rb(s_);
                rb.add(h, res.p, sizeof(h));
                rb.add(&abort_flag, wbar.p, sizeof(unsigned));
                rb.run();
                if (abort_flag & 2)
                    fail(GLV_E_UNSUPPORTED, "a partition of a giant vertex's communities outgrew the LDS table");
                if (abort_flag) fail(GLV_E_DEVICE, "grid barrier of the wide sweep timed out");
                if (numItrs == 1 || !valid_colors) {
                    // a full pass: the first iteration, or classes that may hold an edge
                    launch_widen_ids(wc0.p, nv, C, s_);
                    currMod = colored_modularity(g, P, C, part.p, res.p, &exx);
                } else {
                    // moves of one color step are pairwise non-adjacent, so the
                    // internal weight changes by the sum of the movers' 2(e_yt - e_ys)
                    exx += h[0];
                    const double c = P.constant;
                    currMod = exx * (double)c - h[1] * (double)c * (double)c;
                }
                done_iter = true;
            } else {
                // launch refused: this phase continues with per-color launches
                launch_widen_ids(wc0.p, nv, C, s_);
                labels32 = false;
            }
        }
        if (done_iter) {
        } else if (persistent) {
            launch_fill_i32((int32_t *)bar.p, kBarrierWords, 0, s_);
            Profiler::get().begin(K_PERSIST, s_);
            persistent = launch_colored_iteration(A, plist, pstart, ncol, bar.p, ppart.p, pblocks, s_,
                                                  mode ? stale.p : nullptr, pC1.p, pa1.p, pn1.p,
                                                  pl16.n ? pl16.p : nullptr);
            Profiler::get().end(K_PERSIST, 24.0 * g.nnz + 32.0 * nv, s_);
            if (persistent) {
                launch_sum_partials(ppart.p, pblocks, res.p, s_);
                launch_sum_partials(ppart.p + pblocks, pblocks, res.p + 1, s_);
                double h[2];
                unsigned abort_flag = 0;
                Readback rb(s_);
                rb.add(h, res.p, sizeof(h));
                rb.add(&abort_flag, bar.p, sizeof(unsigned));
                rb.run();
                if (abort_flag) fail(GLV_E_DEVICE, "grid barrier of the persistent sweep timed out");
                const double c = P.constant;
                currMod = h[0] * (double)c - h[1] * (double)c * (double)c;
                exx = h[0];
                done_iter = true;
            } else if (prune && !full) {
                // persistent launch refused: redo this pruned iteration as a full one
                full = true;
                numItrs--;
                las->pruned_sweeps--;
                continue;
            }
            // persistent launch refused: this phase continues with per-color launches
        }
        if (!done_iter) {
            const std::vector<i64> &ptr = *hptr;
            auto lo = [&](int b, int c) { return ptr[(size_t)b * ncol + c]; };
            auto hi = [&](int b, int c) { return ptr[(size_t)b * ncol + c + 1]; };
            for (int c = 0; c < ncol; c++) {
                if (!color_nonempty[c]) continue;  // stale colors can leave classes empty
                Slices S;
                S.lo[0] = S.hi[0] = 0;  // isolated vertices were labelled -1 at init
                i64 ncnt = 0;
                for (int b = 1; b < kNumBins; b++) {
                    S.lo[b] = lo(b, c);
                    S.hi[b] = hi(b, c);
                    ncnt += S.hi[b] - S.lo[b];
                }
                if (ncnt == 0) continue;
                for (int b = 1; b < kNumBins; b++) {
                    if (S.hi[b] <= S.lo[b]) continue;
                    A.verts = plist + S.lo[b];
                    A.count = S.hi[b] - S.lo[b];
                    if (b == 9) giant_args(P, c, A);
                    Profiler::get().begin(sweep_class(b), s_);
                    launch_sweep(b, false, A, s_);
                    Profiler::get().end(sweep_class(b),
                                        hptr == &P.ptr ? 24.0 * P.nnz_of(b, c) + 32.0 * A.count : 0.0, s_);
                }
                Profiler::get().begin(K_APPLY, s_);
                launch_apply_colored(plist, S, tgt.p, C, P.k.p, P.a.p, P.n.p, s_,
                                     mode ? stale.p : nullptr, A.dein, incremental ? dsum.p : nullptr, dpart.p);
                Profiler::get().end(K_APPLY, 24.0 * ncnt, s_);
            }
            if (incremental) {
                // moves of one color step are pairwise non-adjacent, so the
                // internal weight changes by the sum of the movers' 2(e_yt - e_ys)
                launch_sumsq(P.a.p, nv, part.p, res.p + 1, s_);
                double h[2];
                Readback rb(s_);
                rb.add(&h[0], dsum.p, sizeof(double));
                rb.add(&h[1], res.p + 1, sizeof(double));
                rb.run();
                exx += h[0];
                const double c = P.constant;
                currMod = exx * (double)c - h[1] * (double)c * (double)c;
            } else {
                currMod = colored_modularity(g, P, C, part.p, res.p, &exx);
            }
        }
        if (thi > tlo) launch_las_label(las->tverts, tlo, thi, stale.p, las->tlabels, s_);
        if (trace) trace->push_back(currMod);
        if (prune && full && numItrs > 1) rechecks++;
        if ((currMod - prevMod) < thresh) {  // :260
            if (!prune || full) break;
            full = true;  // pruned sweeps stopped gaining: every vertex gets one more look
        } else if (prune) {
            if (full && las->max_rechecks > 0 && rechecks >= las->max_rechecks) {
                prevMod = currMod;
                break;
            }
            full = false;
        }
        if (numItrs >= kMaxPhaseIters) break;
        prevMod = currMod;
    }
    if (labels32) launch_widen_ids(wc0.p, nv, C, s_);
    if (wmoves.n) moves_ += fetch(wmoves.p);
    if (mode) sync();
    if (las_off) {
        las->swept_vertices += (i64)numItrs * nonisolated;
        las->swept_entries += (i64)numItrs * g.nnz;
        las->full_sweeps += numItrs;
    }
    if (iters) *iters = numItrs;
    return prevMod;  // :284
}

// ------------------------------------------------------------- plain phase
// parallelLouvianMethod / parallelLouvianMethodNoMap,
// BasicCommunitiesDetection/parallelLouvainMethod.cpp:65-326 and
// parallelLouvainMethodNoMap.cpp:48-268.
double Engine::plain_phase(const DevGraph &g, i64 *C, double Lower, double thresh, bool nomap,
                           int *iters, std::vector<double> *trace) {
    const i64 nv = g.nv;
    PhaseSetup P;
    setup(g, nullptr, 1, P);
    launch_init_comm(P.k.p, nv, P.a.p, P.n.p, s_);
    DBuf<i64> b0(nv, s_), b1(nv, s_), b2(nv, s_);
    i64 *past = b0.p, *curr = b1.p, *tgt = b2.p;
    launch_iota(past, nv, s_);
    if (nomap) {
        launch_init_comm_ass_opt(g, P.k.p, P.a.p, P.n.p, P.constant, curr, s_);
        launch_update_ax_opt(nv, curr, P.k.p, P.a.p, P.n.p, s_);
    } else {
        launch_iota(curr, nv, s_);
    }
    i64 nparts = kRedBlocks;
    for (int b = 1; b < kNumBins; b++) nparts += sweep_blocks(b, P.bin_hi(b) - P.bin_lo(b), nv);
    DBuf<double> part(nparts, s_), res(2, s_);
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
    double prevMod = -1, currMod = -1;
    int numItrs = 0;
    while (true) {
        numItrs++;
        SweepArgs A;
        A.off = g.off;
        A.tails = g.tails;
        A.w = g.w;
        A.C = curr;
        A.k = P.k.p;
        A.a = P.a.p;
        A.n = P.n.p;
        A.constant = P.constant;
        A.tgt = tgt;
        A.gkeys = gkeys.p;
        A.gvals = gvals.p;
        A.gslots = gslots;
    A.nv = nv;
        i64 off = 0;
        for (int b = 1; b < kNumBins; b++) {
            const i64 lo = P.bin_lo(b), cnt = P.bin_hi(b) - lo;
            if (cnt <= 0) continue;
            A.verts = P.perm.p + lo;
            A.count = cnt;
            A.eii_part = part.p + off;
            if (b == 9) giant_args(P, 0, A);
            Profiler::get().begin(sweep_class(b), s_);
            launch_sweep(b, true, A, s_);
            Profiler::get().end(sweep_class(b), 24.0 * P.bin_nnz(b) + 32.0 * cnt, s_);
            off += sweep_blocks(b, cnt, nv);
        }
        launch_isolated_targets(P.perm.p + P.bin_lo(0), P.bin_hi(0) - P.bin_lo(0), tgt, s_);
        if (off == 0) launch_fill_f64(res.p, 1, 0.0, s_);
        else launch_sum_partials(part.p, off, res.p, s_);
        launch_sumsq(P.a.p, nv, part.p + off, res.p + 1, s_);
        double h[2];
        GLV_HIP(hipMemcpyAsync(h, res.p, sizeof(h), hipMemcpyDeviceToHost, s_));
        GLV_HIP(hipStreamSynchronize(s_));
        const double c = P.constant;
        currMod = (h[0] * (double)c) - (h[1] * (double)c * (double)c);
        if (trace) trace->push_back(currMod);
        if ((currMod - prevMod) < thresh) break;  // :266
        if (numItrs >= kMaxPhaseIters) break;
        prevMod = currMod;
        if (prevMod < Lower) prevMod = Lower;  // :272-273
        Profiler::get().begin(K_APPLY, s_);
        launch_apply_plain(nv, tgt, curr, P.k.p, P.a.p, P.n.p, s_);
        Profiler::get().end(K_APPLY, 16.0 * nv, s_);
        i64 *t = past;
        past = curr;
        curr = tgt;
        tgt = t;
    }
    GLV_HIP(hipMemcpyAsync(C, past, nv * sizeof(i64), hipMemcpyDeviceToDevice, s_));  // :304-307
    if (iters) *iters = numItrs;
    return prevMod;
}

i64 Engine::color_step(const DevGraph &g, i64 *C, const int32_t *colors, int ncol, int which) {
    const i64 nv = g.nv;
    PhaseSetup P;
    setup(g, colors, ncol, P);
    launch_comm_from_labels(C, P.k.p, nv, P.a.p, P.n.p, s_);
    DBuf<i64> tgt(nv, s_);
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
    A.C = C;
    A.k = P.k.p;
    A.a = P.a.p;
    A.n = P.n.p;
    A.constant = P.constant;
    A.tgt = tgt.p;
    A.eii_part = nullptr;
    A.gkeys = gkeys.p;
    A.gvals = gvals.p;
    A.gslots = gslots;
    A.nv = nv;
    Slices S;
    S.lo[0] = S.hi[0] = 0;
    for (int b = 1; b < kNumBins; b++) {
        S.lo[b] = P.lo(b, which);
        S.hi[b] = P.hi(b, which);
        if (S.hi[b] <= S.lo[b]) continue;
        A.verts = P.perm.p + S.lo[b];
        A.count = S.hi[b] - S.lo[b];
        if (b == 9) giant_args(P, which, A);
        launch_sweep(b, false, A, s_);
    }
    // count movers before committing: compare on the host after the commit
    DBuf<i64> before(nv, s_);
    GLV_HIP(hipMemcpyAsync(before.p, C, nv * sizeof(i64), hipMemcpyDeviceToDevice, s_));
    launch_apply_colored(P.perm.p, S, tgt.p, C, P.k.p, P.a.p, P.n.p, s_);
    // isolated vertices of this color become -1 (parallelLouvainWithColoring.cpp:189-191)
    launch_isolated_targets(P.perm.p + P.lo(0, which), P.hi(0, which) - P.lo(0, which), C, s_);
    std::vector<i64> hb(nv), ha(nv);
    GLV_HIP(hipMemcpyAsync(hb.data(), before.p, nv * sizeof(i64), hipMemcpyDeviceToHost, s_));
    GLV_HIP(hipMemcpyAsync(ha.data(), C, nv * sizeof(i64), hipMemcpyDeviceToHost, s_));
    GLV_HIP(hipStreamSynchronize(s_));
    i64 moves = 0;
    for (i64 i = 0; i < nv; i++) moves += hb[i] != ha[i];
    return moves;
}

// ----------------------------------------------------------------- renumber
// renumberClustersContiguously, Utility/buildNextPhase.cpp:48-79: new id =
// rank of the community's first index; -1 stays -1.
i64 Engine::renumber(i64 *C, i64 n) {
    if (n <= 0) return 0;
    DBuf<i64> first(n, s_), flags(n, s_), pos(n, s_);
    DBuf<int32_t> err(1, s_);
    launch_fill_i32(err.p, 1, 0, s_);
    Profiler::get().begin(K_RENUMBER, s_);
    launch_fill_i64(first.p, n, n, s_);
    launch_first_index(C, n, first.p, err.p, s_);
    launch_head_flags(C, n, first.p, flags.p, s_);
    scan_exclusive_i64(flags.p, pos.p, n, s_);
    launch_relabel(C, n, first.p, pos.p, s_);
    Profiler::get().end(K_RENUMBER, 16.0 * n, s_);
    i64 last[2];
    int32_t e = 0;
    Readback rb(s_);
    rb.add(&last[0], pos.p + n - 1, sizeof(i64));
    rb.add(&last[1], flags.p + n - 1, sizeof(i64));
    rb.add(&e, err.p, sizeof(e));
    rb.run();
    if (e) fail(GLV_E_INVALID_ARG, "cluster id >= number of vertices");
    return last[0] + last[1];
}

// ------------------------------------------------------------------ coarsen
// buildNextLevelGraphOpt, Utility/buildNextPhase.cpp:83-255, as sort plus
// segmented reduce.  Output rows are sorted by tail.
i64 Engine::vertex_following(const DevGraph &g, i64 *C) {
    DBuf<i64> cnt(1, s_);
    launch_fill_i64(cnt.p, 1, 0, s_);
    launch_vertex_following(g, C, cnt.p, s_);
    return fetch(cnt.p);
}

void Engine::coarsen(const DevGraph &g, const i64 *C, i64 k, DevGraph &out, bool vf) {
    if (vf && k == 0) {
        // every vertex isolated and labelled -1: buildNewGraphVF yields an empty graph
        out.nv = out.nnz = out.num_edges = 0;
        out.off_buf.alloc(1, s_);
        launch_fill_i64(out.off_buf.p, 1, 0, s_);
        out.adopt();
        return;
    }
    if (k <= 0) fail(GLV_E_INVALID_ARG, "coarsening needs at least one cluster");
    const int kb = bits_for(k);
    if (2 * kb > 64) fail(GLV_E_UNSUPPORTED, "more than 2^32 clusters");
    const i64 nv = g.nv;
    Profiler::get().begin(K_COARSEN, s_);
    const i64 nch = coarsen_chunks(g.nnz);
    DBuf<i64> cnt(std::max<i64>(nch, 1), s_), pos(std::max<i64>(nch, 1), s_);
    launch_coarsen_count(g, C, cnt.p, s_);
    scan_exclusive_i64(cnt.p, pos.p, nch, s_);
    i64 m = 0;
    if (nch > 0) {
        i64 h[2];
        Readback rb(s_);
        rb.add(&h[0], pos.p + nch - 1, sizeof(i64));
        rb.add(&h[1], cnt.p + nch - 1, sizeof(i64));
        rb.run();
        m = h[0] + h[1];
    }
    const i64 tot = m + (vf ? 0 : k);
    DBuf<u64> keys(tot, s_);
    DBuf<double> vals(tot, s_);
    launch_coarsen_emit(g, C, pos.p, k, kb, keys.p, vals.p, s_, vf);
    cnt.release();
    pos.release();
    sort_pairs_u64_f64(keys.p, vals.p, tot, 2 * kb, s_);
    DBuf<u64> ukeys(tot, s_);
    DBuf<double> usums(tot, s_);
    const i64 nu = reduce_by_key_u64_f64(keys.p, vals.p, tot, ukeys.p, usums.p, s_);
    keys.release();
    vals.release();
    DBuf<u64> k2(2 * nu, s_);
    DBuf<double> v2(2 * nu, s_);
    DBuf<i64> nself(1, s_);
    launch_fill_i64(nself.p, 1, 0, s_);
    launch_coarsen_mirror(ukeys.p, usums.p, nu, kb, k2.p, v2.p, s_, nself.p);
    ukeys.release();
    usums.release();
    sort_pairs_u64_f64(k2.p, v2.p, 2 * nu, 2 * kb, s_);
    // self-loop entries appear once (k of them with the seeds)
    const i64 total = 2 * nu - (vf ? fetch(nself.p) : k);
    out.nv = k;
    out.nnz = total;
    out.num_edges = nu;  // NE_out + k (buildNextPhase.cpp:237)
    out.off_buf.alloc(k + 1, s_);
    out.tails_buf.alloc(total, s_);
    out.w_buf.alloc(total, s_);
    out.adopt();
    launch_coarsen_rows(k2.p, v2.p, total, k, kb, out.off_buf.p, out.tails_buf.p, out.w_buf.p, s_);
    // B_coarsen = 24 nnz_in + 8 NV_in + 24 nnz_out + 8 (k+1)
    Profiler::get().end(K_COARSEN, 24.0 * g.nnz + 8.0 * nv + 24.0 * total + 8.0 * (k + 1), s_);
}

double Engine::modularity(const DevGraph &g, const i64 *C) {
    PhaseSetup P;
    setup(g, nullptr, 1, P);
    launch_comm_from_labels(C, P.k.p, g.nv, P.a.p, P.n.p, s_);
    DBuf<double> part(mod_part_count(P), s_), res(2, s_);
    return colored_modularity(g, P, C, part.p, res.p);
}

// ------------------------------------------------------------ multi-phase
// runMultiPhaseColoring, Coloring/runMultiPhaseColoring.cpp:51-238 (coloring 1, 2 or 3).
int Engine::color_for_run(const DevGraph &g, int32_t *colors, const RunParams &p) {
    if (p.coloring == 3) {
        // runMultiPhaseColoring.cpp:78-86: two hashes, numColors / 4 iterations
        const int nItrs = std::max(p.numColors / 4, 1);
        const int mc = color_multihash(g, colors, 2, nItrs, s_);
        if (mc + 1 > kMaxColors) fail(GLV_E_TOO_MANY_COLORS, "multi-hash coloring needs more than 4096 colors");
        return mc + 1;
    }
    const int n = color(g, colors) + 1;
    if (p.coloring == 2) balance_colors(g, colors, n, 0, s_);  // :72-73
    return n;
}

void Engine::run_colored(const DevGraph &g0, i64 *C_orig, const RunParams &p, RunResult &r) {
    if (p.coloring < 1 || p.coloring > 3) fail(GLV_E_INVALID_ARG, "coloring must be 1, 2 or 3");
    if (p.las && p.coloring != 1)
        fail(GLV_E_UNSUPPORTED, "the learned active set needs independent color classes (coloring == 1)");
    const double t_start = now();
    const i64 NV = g0.nv;
    r = RunResult();
    if (NV == 0) return;
    DBuf<int32_t> colors(NV, s_);
    DBuf<i64> Cbuf(NV, s_);
    double t0 = now();
    int nColors = color_for_run(g0, colors.p, p);
    bool fresh_colors = true;  // colors were computed for the graph being swept
    r.t_coloring += now() - t0;
    r.colors_first = nColors;

    const DevGraph *G = &g0;
    DevGraph levels[2];
    int cur = -1;  // which `levels` entry holds G (-1: the input)
    double prevMod = -1, currMod = -1;
    int phase = 1, totItr = 0;
    i64 numClusters = 0;
    bool nonColor = false;
    while (true) {
        prevMod = currMod;
        PhaseRecord rec;
        rec.nv = G->nv;
        rec.nnz = G->nnz;
        int tmpItr = 0;
        t0 = now();
        const i64 swept0 = p.las ? p.las->swept_entries : 0;
        if (!nonColor) {
            currMod = colored_phase(*G, Cbuf.p, colors.p, nColors, currMod, p.C_threshold, &tmpItr,
                                    &r.trace, p.las, phase - 1, fresh_colors && p.coloring != 3);
            rec.colored = 1;
            rec.ncolors = nColors;
            r.edges_visited_mod += (i64)tmpItr * G->nnz;
        } else {
            currMod = plain_phase(*G, Cbuf.p, currMod, p.threshold, p.replaceMap == 1, &tmpItr,
                                  &r.trace);
        }
        rec.t_sweep = now() - t0;
        r.t_clustering += rec.t_sweep;
        totItr += tmpItr;
        rec.iters = tmpItr;
        rec.mod = currMod;
        if (!nonColor && p.las && p.las->mode == 2)
            r.edges_visited += p.las->swept_entries - swept0;  // entries the pruned sweeps read
        else
            r.edges_visited += (i64)tmpItr * G->nnz;
        t0 = now();
        numClusters = renumber(Cbuf.p, G->nv);
        launch_compose(C_orig, NV, Cbuf.p, phase == 1, s_);
        rec.t_renumber = now() - t0;
        r.t_renumber += rec.t_renumber;
        rec.clusters = numClusters;
        bool stop = (phase > p.max_phases) || (totItr > p.max_total_iters);
        if (!stop && (currMod - prevMod) > p.threshold && numClusters > 0) {
            t0 = now();
            const int nxt = cur == 0 ? 1 : 0;
            coarsen(*G, Cbuf.p, numClusters, levels[nxt]);
            sync();
            if (cur >= 0) {
                levels[cur].off_buf.release();
                levels[cur].tails_buf.release();
                levels[cur].w_buf.release();
            }
            cur = nxt;
            G = &levels[cur];
            rec.t_coarsen = now() - t0;
            r.t_building += rec.t_coarsen;
            launch_fill_i64(Cbuf.p, G->nv, -1, s_);
            phase++;
            fresh_colors = false;
            if (G->nv > p.minGraphSize && !nonColor) {
                t0 = now();
                nColors = color_for_run(*G, colors.p, p);
                fresh_colors = true;
                rec.t_color = now() - t0;
                r.t_coloring += rec.t_color;
            }
            r.phase.push_back(rec);
            continue;
        }
        r.phase.push_back(rec);
        if (stop) break;
        if (!nonColor && numClusters > 0)
            nonColor = true;
        else
            break;
    }
    sync();
    r.phases = phase;
    r.total_iterations = totItr;
    r.final_clusters = numClusters;
    r.final_modularity = prevMod;
    r.last_modularity = currMod;
    r.moves = moves_;
    r.t_total = now() - t_start;
}

// runMultiPhaseBasic, BasicCommunitiesDetection/runMultiPhaseBasic.cpp:52-170.
void Engine::run_basic(const DevGraph &g0, i64 *C_orig, const RunParams &p, RunResult &r) {
    const double t_start = now();
    const i64 NV = g0.nv;
    r = RunResult();
    if (NV == 0) return;
    DBuf<i64> Cbuf(NV, s_);
    const DevGraph *G = &g0;
    DevGraph levels[2];
    int cur = -1;
    double prevMod = -1, currMod = -1;
    int phase = 1, totItr = 0;
    i64 numClusters = 0;
    while (true) {
        prevMod = currMod;
        PhaseRecord rec;
        rec.nv = G->nv;
        rec.nnz = G->nnz;
        int tmpItr = 0;
        double t0 = now();
        currMod = plain_phase(*G, Cbuf.p, currMod, p.threshold, p.replaceMap == 1, &tmpItr, &r.trace);
        rec.t_sweep = now() - t0;
        r.t_clustering += rec.t_sweep;
        totItr += tmpItr;
        rec.iters = tmpItr;
        rec.mod = currMod;
        r.edges_visited += (i64)tmpItr * G->nnz;
        t0 = now();
        numClusters = renumber(Cbuf.p, G->nv);
        launch_compose(C_orig, NV, Cbuf.p, phase == 1, s_);
        rec.t_renumber = now() - t0;
        r.t_renumber += rec.t_renumber;
        rec.clusters = numClusters;
        r.phase.push_back(rec);
        if ((phase > p.max_phases) || (totItr > p.max_total_iters)) break;
        if ((currMod - prevMod) > p.threshold && numClusters > 0) {
            t0 = now();
            const int nxt = cur == 0 ? 1 : 0;
            coarsen(*G, Cbuf.p, numClusters, levels[nxt]);
            sync();
            if (cur >= 0) {
                levels[cur].off_buf.release();
                levels[cur].tails_buf.release();
                levels[cur].w_buf.release();
            }
            cur = nxt;
            G = &levels[cur];
            r.t_building += now() - t0;
            r.phase.back().t_coarsen = now() - t0;
            launch_fill_i64(Cbuf.p, G->nv, -1, s_);
            phase++;
        } else {
            break;
        }
    }
    sync();
    r.phases = phase;
    r.total_iterations = totItr;
    r.final_clusters = numClusters;
    r.final_modularity = prevMod;
    r.last_modularity = currMod;
    r.t_total = now() - t_start;
}

}  // namespace glv
