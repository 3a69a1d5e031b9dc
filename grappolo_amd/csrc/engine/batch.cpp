// batch.cpp -- host orchestration of the batched plain Louvain (batch.hpp).
//
// Each graph g of the batch runs runMultiPhaseBasic
// (BasicCommunitiesDetection/runMultiPhaseBasic.cpp:52-170) with the Jacobi
// sweeps of parallelLouvianMethod (parallelLouvainMethod.cpp:65-326): its own
// 1/(2m), its own modularity and termination test, its own phase decisions.
// The union of the graphs still running is one CSR: every iteration sweeps
// the vertices of all graphs whose current phase is still iterating (list
// compacted by graph), and every phase ends with one renumbering and one
// coarsening of the whole union (graphs are block-diagonal, so renumbering by
// first appearance gives each graph a contiguous id range and the coarse
// union is the union of the coarse graphs).
#include "batch.hpp"

#include <algorithm>
#include <chrono>

#include "active_set.hpp"
#include "kernels.hpp"
#include "profiler.hpp"

namespace glv {

namespace {
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
constexpr int kMaxPhaseIters = 100000;
constexpr int kLevelUnroll = 4;  // iterations captured in one graph launch
}  // namespace

void run_basic_batch(Engine &E, const DevGraph &g0, const std::vector<i64> &vb0, i64 *C_orig,
                     const BatchParams &p, BatchResult &r) {
    const double t0 = now_s();
    hipStream_t s = E.stream();
    const i64 ng = (i64)vb0.size() - 1;
    const i64 n0 = g0.nv;
    r = BatchResult();
    r.final_modularity.assign(ng, 0.0);
    r.last_modularity.assign(ng, 0.0);
    r.phases.assign(ng, 0);
    r.total_iterations.assign(ng, 0);
    r.clusters.assign(ng, 0);
    if (ng <= 0) return;
    launch_fill_i64(C_orig, n0, -1, s);
    if (n0 == 0) return;

    // graph of every original vertex; entries per graph (edgeless graphs end at once)
    DBuf<i64> d_vb0(ng + 1, s);
    GLV_HIP(hipMemcpyAsync(d_vb0.p, vb0.data(), (ng + 1) * sizeof(i64), hipMemcpyHostToDevice, s));
    DBuf<int32_t> gid0(n0, s);
    launch_fill_gid(d_vb0.p, ng, gid0.p, s);
    std::vector<i64> eoff(ng + 1);
    {
        DBuf<i64> t(ng + 1, s);
        launch_gather_i64(g0.off, d_vb0.p, ng + 1, t.p, s);
        GLV_HIP(hipMemcpyAsync(eoff.data(), t.p, (ng + 1) * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
    }
    std::vector<uint8_t> alive(ng, 0);
    std::vector<double> pm_prev(ng, -1.0), pm_curr(ng, -1.0);
    std::vector<int32_t> tot(ng, 0);
    for (i64 g = 0; g < ng; g++) {
        if (eoff[g + 1] > eoff[g]) {
            alive[g] = 1;
        } else {
            // edgeless graph: every vertex -1, modularity 0 (as the single-graph entry points)
            r.phases[g] = vb0[g + 1] > vb0[g] ? 1 : 0;
        }
    }

    const DevGraph *G = &g0;
    DevGraph levels[2];
    int cur = -1;
    std::vector<i64> vb = vb0;
    DBuf<i64> d_vb(ng + 1, s);
    DBuf<uint8_t> d_alive(ng, s), d_active(ng, s), d_apply(ng, s), d_stop(ng, s);
    DBuf<double> d_gconst(ng, s), d_lower(ng, s), d_prev(ng, s), d_curr(ng, s), d_exx(ng, s), d_a2(ng, s);
    DBuf<int32_t> d_itrs(ng, s);
    DBuf<i64> d_lo(ng, s), d_hi(ng, s), d_base(ng, s);
    std::vector<uint8_t> h_active(ng);
    std::vector<double> h_prev(ng), h_curr(ng);
    std::vector<int32_t> h_itrs(ng);
    std::vector<i64> h_lo(ng), h_hi(ng);
    int phase = 1;
    while (true) {
        const i64 nv = G->nv;
        GLV_HIP(hipMemcpyAsync(d_vb.p, vb.data(), (ng + 1) * sizeof(i64), hipMemcpyHostToDevice, s));
        GLV_HIP(hipMemcpyAsync(d_alive.p, alive.data(), ng, hipMemcpyHostToDevice, s));
        GLV_HIP(hipMemcpyAsync(d_lower.p, pm_curr.data(), ng * sizeof(double), hipMemcpyHostToDevice, s));
        DBuf<int32_t> gid(nv, s);
        launch_fill_gid(d_vb.p, ng, gid.p, s);
        PhaseSetup P;
        E.setup(*G, nullptr, 1, P);
        launch_graph_const(d_vb.p, ng, P.k.p, d_gconst.p, s);
        launch_init_comm(P.k.p, nv, P.a.p, P.n.p, s);
        DBuf<i64> b0(nv, s), b1(nv, s), b2(nv, s);
        i64 *past = b0.p, *curr = b1.p, *tgt = b2.p;
        launch_iota(past, nv, s);
        if (p.basicOpt == 1) {
            launch_init_comm_ass_opt(*G, P.k.p, P.a.p, P.n.p, 0.0, curr, s, gid.p, d_gconst.p);
            launch_update_ax_opt(nv, curr, P.k.p, P.a.p, P.n.p, s);
        } else {
            launch_iota(curr, nv, s);
        }
        launch_fill_f64(d_prev.p, ng, -1.0, s);
        launch_fill_i32(d_itrs.p, ng, 0, s);
        GLV_HIP(hipMemcpyAsync(d_active.p, d_alive.p, ng, hipMemcpyDeviceToDevice, s));
        DBuf<double> eii(nv, s);
        launch_fill_f64(eii.p, nv, 0.0, s);
        i64 nparts = kRedBlocks;
        for (int b = 1; b < kNumBins; b++) nparts += sweep_blocks(b, P.bin_hi(b) - P.bin_lo(b), nv);
        DBuf<double> part(nparts, s);
        DBuf<i64> gkeys;
        DBuf<double> gvals;
        i64 gslots = 0;
        if (P.maxdeg_global > 0) {
            gslots = 1;
            while (gslots < 2 * P.maxdeg_global) gslots <<= 1;
            const i64 nb = std::min<i64>(kGlobalHashBlocks, P.bin_hi(8) - P.bin_lo(8));
            gkeys.alloc(nb * gslots, s);
            gvals.alloc(nb * gslots, s);
        }
        const i64 nkeys = kNumBins;
        DBuf<i64> flags(nv, s), fpos(nv, s), list(nv, s), start(nkeys + 1, s);
        std::vector<i64> hstart(nkeys + 1);
        DBuf<int32_t> d_done(1, s);
        DBuf<i64> d_live(1, s);
        int32_t done = 0;
        while (!done) {
            // sweep lists of the graphs still iterating
            launch_compact_graphs(P.perm.p, nv, gid.p, d_active.p, P.dptr.p, nkeys, flags.p, fpos.p, list.p,
                                  start.p, s);
            GLV_HIP(hipMemcpyAsync(hstart.data(), start.p, (nkeys + 1) * sizeof(i64), hipMemcpyDeviceToHost, s));
            GLV_HIP(hipStreamSynchronize(s));
            const i64 listed = hstart[nkeys];
            if (listed == 0) break;  // every graph's phase has ended
            // kLevelUnroll iterations (sweeps, per-graph modularity, decisions,
            // commits, the level's status) captured as one graph for these
            // lists and launched until the level ends; once no graph iterates,
            // the captured sweeps return at once and nothing is committed.  A
            // graph whose phase ends meanwhile is still swept (its targets are
            // not used) until the lists are compacted again, after the launch
            // in which its phase ended (capturing and instantiating costs far
            // less than the four sweeps of a launch).
            launch_fill_i32(d_done.p, 1, 0, s);
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            GLV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            SweepArgs A;
            A.off = G->off;
            A.tails = G->tails;
            A.w = G->w;
            A.C = curr;
            A.k = P.k.p;
            A.a = P.a.p;
            A.n = P.n.p;
            A.constant = 0.0;
            A.tgt = tgt;
            A.gkeys = gkeys.p;
            A.gvals = gvals.p;
            A.gslots = gslots;
            A.nv = nv;
            A.gid = gid.p;
            A.gconst = d_gconst.p;
            A.eii_v = eii.p;
            A.skip = d_done.p;
            for (int u = 0; u < kLevelUnroll; u++) {
                for (int b = 1; b < kNumBins; b++) {
                    const i64 lo = hstart[b], cnt = hstart[b + 1] - lo;
                    if (cnt <= 0) continue;
                    A.verts = list.p + lo;
                    A.count = cnt;
                    A.eii_part = part.p;
                    if (b == 9) giant_args(P, 0, A);
                    launch_sweep(b, true, A, s);
                }
                launch_isolated_targets(P.perm.p + P.bin_lo(0), P.bin_hi(0) - P.bin_lo(0), tgt, s);
                launch_graph_mod(d_vb.p, ng, d_active.p, eii.p, P.a.p, d_exx.p, d_a2.p, s, d_done.p);
                launch_batch_decide(ng, d_exx.p, d_a2.p, d_gconst.p, d_lower.p, p.threshold, kMaxPhaseIters,
                                    d_prev.p, d_curr.p, d_itrs.p, d_active.p, d_apply.p, s);
                launch_batch_apply(nv, gid.p, d_apply.p, tgt, curr, past, P.k.p, P.a.p, P.n.p, s, d_done.p);
                launch_level_status(d_vb.p, ng, d_active.p, d_done.p, d_live.p, s);
            }
            GLV_HIP(hipStreamEndCapture(s, &graph));
            GLV_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            while (true) {
                GLV_HIP(hipGraphLaunch(exec, s));
                r.launches++;
                i64 live = 0;
                GLV_HIP(hipMemcpyAsync(&done, d_done.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
                GLV_HIP(hipMemcpyAsync(&live, d_live.p, sizeof(i64), hipMemcpyDeviceToHost, s));
                GLV_HIP(hipStreamSynchronize(s));
                if (done || live < listed) break;
            }
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
        }
        r.levels++;
        // per graph: the phase's value (prevMod at exit) and its iterations
        GLV_HIP(hipMemcpyAsync(h_prev.data(), d_prev.p, ng * sizeof(double), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipMemcpyAsync(h_itrs.data(), d_itrs.p, ng * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        // graphs that stopped earlier leave the union: their labels become -1
        launch_mask_graphs(d_vb.p, ng, d_alive.p, past, s);
        const i64 k = E.renumber(past, nv);
        launch_graph_label_range(d_vb.p, ng, past, d_lo.p, d_hi.p, s);
        launch_batch_compose(n0, gid0.p, d_alive.p, C_orig, past, phase == 1, s);
        GLV_HIP(hipMemcpyAsync(h_lo.data(), d_lo.p, ng * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipMemcpyAsync(h_hi.data(), d_hi.p, ng * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        // phase decisions of runMultiPhaseBasic.cpp:120-160, per graph
        std::vector<uint8_t> stop(ng, 0);
        std::vector<i64> base(ng, 0), kg(ng, 0);
        bool any_alive = false;
        int level_itrs = 0;  // the level's iterations: those of its longest-running graph
        for (i64 g = 0; g < ng; g++)
            if (alive[g]) level_itrs = std::max(level_itrs, h_itrs[g]);
        r.iterations += level_itrs;
        for (i64 g = 0; g < ng; g++) {
            kg[g] = h_hi[g] >= 0 ? h_hi[g] - h_lo[g] + 1 : 0;
            base[g] = h_hi[g] >= 0 ? h_lo[g] : 0;
            if (!alive[g]) continue;
            pm_prev[g] = pm_curr[g];
            pm_curr[g] = h_prev[g];
            tot[g] += h_itrs[g];
            const bool cap = (phase > p.max_phases) || (tot[g] > p.max_total_iters);
            const bool go = !cap && (pm_curr[g] - pm_prev[g]) > p.threshold && kg[g] > 0;
            if (!go) {
                stop[g] = 1;
                r.final_modularity[g] = pm_prev[g];
                r.last_modularity[g] = pm_curr[g];
                r.phases[g] = phase;
                r.total_iterations[g] = tot[g];
                r.clusters[g] = kg[g];
            } else {
                any_alive = true;
            }
        }
        GLV_HIP(hipMemcpyAsync(d_stop.p, stop.data(), ng, hipMemcpyHostToDevice, s));
        GLV_HIP(hipMemcpyAsync(d_base.p, base.data(), ng * sizeof(i64), hipMemcpyHostToDevice, s));
        launch_batch_finalize(n0, gid0.p, d_stop.p, d_base.p, C_orig, s);
        GLV_HIP(hipStreamSynchronize(s));  // host vectors above are pageable
        for (i64 g = 0; g < ng; g++)
            if (stop[g]) alive[g] = 0;
        if (!any_alive || k == 0) break;
        // next level: the coarse union; graph g's vertices are its renumbered ids
        std::vector<i64> nvb(ng + 1, 0);
        for (i64 g = 0; g < ng; g++) nvb[g + 1] = nvb[g] + kg[g];
        const int nxt = cur == 0 ? 1 : 0;
        E.coarsen(*G, past, k, levels[nxt]);
        E.sync();
        if (cur >= 0) {
            levels[cur].off_buf.release();
            levels[cur].tails_buf.release();
            levels[cur].w_buf.release();
        }
        cur = nxt;
        G = &levels[cur];
        vb.swap(nvb);
        phase++;
    }
    E.sync();
    r.t_total = now_s() - t0;
}

}  // namespace glv
