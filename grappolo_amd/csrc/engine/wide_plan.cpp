// wide_plan.cpp -- the work items of the whole-GPU persistent sweep (see
// wide.hpp), built on the host once per phase from the (degree bin, color)
// grouping of the vertices.
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <new>
#include <thread>

#include "louvain.hpp"
#include "wide.hpp"

namespace glv {

// The items of one color class (giant indices local to the class; the
// giants' pair and count offsets relative to the class).
struct ClassPlan {
    std::vector<WideItem> wi, vi;
    std::vector<i64> wstart, vstart;  // per workgroup, nb + 1 each
    std::vector<WideGiant> giants;
    i64 pairs = 0, counts = 0, parts = 1;
    bool ok = true;
};

static void plan_class(const PhaseSetup &P, const std::vector<i64> &vd, i64 hlo, int c, int nb, ClassPlan &out) {
    std::vector<WideItem> cls;  // the workgroup items of the class
    std::vector<std::vector<WideItem>> per(nb);
    std::vector<WideGiant> &giants = out.giants;
    // giants of the class: accumulation chunks first
    i64 pairs = 0, counts = 0;
    for (int b = 8; b < kNumBins; b++)
        for (i64 q = P.lo(b, c); q < P.hi(b, c); q++) {
            const i64 v = vd[2 * (q - hlo)], d = vd[2 * (q - hlo) + 1];
            if (d <= kWHeavyMaxDeg) continue;
            WideGiant G;
            G.v = v;
            G.nchunks = (int32_t)((d + kWChunk - 1) / kWChunk);
            if (G.nchunks > kWMaxChunks) {
                out.ok = false;
                return;
            }
            G.nparts = (int32_t)std::min<i64>(kWMaxParts, std::max<i64>(1, (d + kWPartKeys - 1) / kWPartKeys));
            G.pairs = pairs;
            G.cnt = counts;
            pairs += (i64)G.nchunks * kWChunk;
            counts += (i64)G.nchunks * (G.nparts + 1);
            out.parts = std::max<i64>(out.parts, G.nparts);
            const i64 gi = (i64)giants.size();
            giants.push_back(G);
            for (int32_t j = 0; j < G.nchunks; j++) cls.push_back(WideItem{gi, kWGiantChunk, j});
        }
    out.pairs = pairs;
    out.counts = counts;
    // heavy vertices, the larger bin first
    for (int b = kNumBins - 1; b >= 7; b--)
        for (i64 q = P.lo(b, c); q < P.hi(b, c); q++) {
            const i64 v = vd[2 * (q - hlo)], d = vd[2 * (q - hlo) + 1];
            if (d <= kWHeavyMaxDeg) cls.push_back(WideItem{v, kWHeavy, (int32_t)d});
        }
    // the giants' table slices, scored once every chunk is in
    for (size_t gi = 0; gi < giants.size(); gi++)
        for (int32_t j = 0; j < giants[gi].nparts; j++) cls.push_back(WideItem{(i64)gi, kWGiantPart, j});
    // longest processing time first: each item, costliest first, to the
    // workgroup with the least work so far; then every workgroup's list in
    // kind order (chunks, heavy vertices, slices), so that a slice only
    // waits on chunks, which never wait
    std::vector<std::pair<double, int>> heap;
    for (int w = 0; w < nb; w++) heap.push_back({-0.0, -w});
    std::make_heap(heap.begin(), heap.end());  // max-heap of (-load, -w): least loaded, lowest index
    std::vector<size_t> order(cls.size());
    for (size_t i = 0; i < cls.size(); i++) order[i] = i;
    auto item_cost = [](const WideItem &it) {
        return it.kind == kWHeavy ? kWCostHeavy + 3.0 * it.y / 1024.0 : it.kind == kWGiantChunk ? kWCostChunk : kWCostPart;
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return item_cost(cls[a]) > item_cost(cls[b]); });
    for (size_t i : order) {
        std::pop_heap(heap.begin(), heap.end());
        auto &h = heap.back();
        per[-h.second].push_back(cls[i]);
        h.first -= item_cost(cls[i]);
        std::push_heap(heap.begin(), heap.end());
    }
    out.wstart.assign(nb + 1, 0);
    for (int w = 0; w < nb; w++) {
        out.wstart[w] = (i64)out.wi.size();
        std::stable_sort(per[w].begin(), per[w].end(), [](const WideItem &a, const WideItem &b) {
            const int ka = a.kind == kWGiantChunk ? 0 : a.kind == kWHeavy ? 1 : 2;
            const int kb = b.kind == kWGiantChunk ? 0 : b.kind == kWHeavy ? 1 : 2;
            return ka < kb;
        });
        out.wi.insert(out.wi.end(), per[w].begin(), per[w].end());
    }
    out.wstart[nb] = (i64)out.wi.size();
    // wavefront items (mid vertices, then light ones from the largest bin
    // down), each to the workgroup whose work (its workgroup items plus its
    // wavefront items spread over its wavefronts) is least so far
    for (auto &q : per) q.clear();
    auto give = [&](const WideItem &it, double cost) {
        std::pop_heap(heap.begin(), heap.end());
        auto &h = heap.back();
        per[-h.second].push_back(it);
        h.first -= cost / kWWavesPerGroup;
        std::push_heap(heap.begin(), heap.end());
    };
    for (i64 q = P.lo(6, c); q < P.hi(6, c); q++) give(WideItem{q, kWMid, 1}, kWCostMid);
    for (int b = 5; b >= 1; b--) {
        const int per_item = kWave / kBinGroup[b];
        for (i64 q = P.lo(b, c); q < P.hi(b, c); q += per_item)
            give(WideItem{q, (int32_t)(kWLight1 + b - 1), (int32_t)std::min<i64>(per_item, P.hi(b, c) - q)},
                 kWCostLight);
    }
    out.vstart.assign(nb + 1, 0);
    for (int w = 0; w < nb; w++) {
        out.vstart[w] = (i64)out.vi.size();
        out.vi.insert(out.vi.end(), per[w].begin(), per[w].end());
    }
    out.vstart[nb] = (i64)out.vi.size();
}

// host threads for the planning: the CPUs this process may run on, at most 16
static int host_threads() {
    cpu_set_t set;
    int n = 1;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    return std::max(1, std::min(n, 16));
}

bool build_wide_plan(const DevGraph &g, const PhaseSetup &P, int nb, WidePlan &L, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const int ncol = P.ncol;
    // vertex and degree of every vertex of bins 7..9 (heavy and giant vertices)
    const i64 hlo = P.bin_lo(7), hhi = P.bin_hi(kNumBins - 1);
    std::vector<i64> vd(2 * (hhi - hlo));
    if (hhi > hlo) {
        DBuf<i64> dvd(vd.size(), s);
        launch_wide_degrees(P.perm.p, g.off, hlo, hhi, dvd.p, s);
        Readback rb(s);
        rb.add(vd.data(), dvd.p, vd.size() * sizeof(i64));
        rb.run();
    }
    // the classes are planned independently, on host threads, then concatenated
    std::vector<ClassPlan> cp(ncol);
    const int nthreads = (int)std::min<i64>(ncol, host_threads());
    std::atomic<int> next{0};
    std::atomic<bool> no_memory{false};
    auto worker = [&]() {
        for (int c = next++; c < ncol; c = next++) {
            try {
                plan_class(P, vd, hlo, c, nb, cp[c]);
            } catch (const std::bad_alloc &) {
                no_memory = true;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; t++) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    if (no_memory) throw std::bad_alloc();
    std::vector<WideItem> wi, vi;
    std::vector<i64> wstart((size_t)ncol * nb + 1, 0), vstart((size_t)ncol * nb + 1, 0);
    std::vector<WideGiant> giants;
    i64 gpairs = 0, gcounts = 0, gparts = 1, nwi = 0, nvi = 0;
    for (int c = 0; c < ncol; c++) {
        if (!cp[c].ok) {
            trace("  wide plan: a vertex of more than %ld entries", (long)(kWMaxChunks * kWChunk));
            return false;
        }
        nwi += (i64)cp[c].wi.size();
        nvi += (i64)cp[c].vi.size();
    }
    wi.reserve(nwi);
    vi.reserve(nvi);
    for (int c = 0; c < ncol; c++) {
        const ClassPlan &C = cp[c];
        const i64 gbase = (i64)giants.size();
        giants.insert(giants.end(), C.giants.begin(), C.giants.end());
        gpairs = std::max(gpairs, C.pairs);
        gcounts = std::max(gcounts, C.counts);
        gparts = std::max(gparts, C.parts);
        for (int w = 0; w < nb; w++) {
            wstart[(size_t)c * nb + w] = (i64)wi.size();
            for (i64 i = C.wstart[w]; i < C.wstart[w + 1]; i++) {
                WideItem it = C.wi[i];
                if (it.kind != kWHeavy) it.x += gbase;  // giant items: the phase's giant index
                wi.push_back(it);
            }
        }
        for (int w = 0; w < nb; w++) {
            vstart[(size_t)c * nb + w] = (i64)vi.size();
            vi.insert(vi.end(), C.vi.begin() + C.vstart[w], C.vi.begin() + C.vstart[w + 1]);
        }
    }
    wstart[(size_t)ncol * nb] = (i64)wi.size();
    vstart[(size_t)ncol * nb] = (i64)vi.size();
    L.nwitems = (i64)wi.size();
    L.nvitems = (i64)vi.size();
    L.ngiants = (i64)giants.size();
    L.gparts = gparts;
    L.gpairs = gpairs;
    L.gcounts = gcounts;
    L.witems.alloc(std::max<size_t>(wi.size(), 1), s);
    L.vitems.alloc(std::max<size_t>(vi.size(), 1), s);
    L.wstart.alloc(wstart.size(), s);
    L.vstart.alloc(vstart.size(), s);
    L.giants.alloc(std::max<size_t>(giants.size(), 1), s);
    const i64 ng = std::max<i64>(L.ngiants, 1);
    L.gpk.alloc(std::max<i64>(gpairs, 1), s);
    L.gpv.alloc(std::max<i64>(gpairs, 1), s);
    L.gcnt.alloc(std::max<i64>(gcounts, 1), s);
    L.gstate.alloc(3 * ng, s);
    L.gsc.alloc(2 * ng, s);
    L.gout.alloc(ng, s);
    L.gcand.alloc(3 * ng * gparts, s);
    L.gcid.alloc(ng * gparts, s);
    if (!wi.empty()) GLV_HIP(hipMemcpyAsync(L.witems.p, wi.data(), wi.size() * sizeof(WideItem), hipMemcpyHostToDevice, s));
    if (!vi.empty()) GLV_HIP(hipMemcpyAsync(L.vitems.p, vi.data(), vi.size() * sizeof(WideItem), hipMemcpyHostToDevice, s));
    GLV_HIP(hipMemcpyAsync(L.wstart.p, wstart.data(), wstart.size() * sizeof(i64), hipMemcpyHostToDevice, s));
    GLV_HIP(hipMemcpyAsync(L.vstart.p, vstart.data(), vstart.size() * sizeof(i64), hipMemcpyHostToDevice, s));
    if (!giants.empty())
        GLV_HIP(hipMemcpyAsync(L.giants.p, giants.data(), giants.size() * sizeof(WideGiant), hipMemcpyHostToDevice, s));
    launch_wide_clear_giants(L.gstate.p, L.gsc.p, L.gout.p, ng, s);
    GLV_HIP(hipStreamSynchronize(s));  // the host vectors are pageable
    trace("  wide plan: %ld workgroup items, %ld wavefront items, %ld giants (%ld pairs, %ld partitions at most), "
          "%.1f ms on the host",
          (long)L.nwitems, (long)L.nvitems, (long)L.ngiants, (long)gpairs, (long)gparts,
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    return true;
}

}  // namespace glv
