// capi.cpp -- the C ABI (include/grappolo_hip.h) and Grappolo's C++ entry
// points (include/grappolo/*.h) on top of the engine.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "grappolo/basic_comm.h"
#include "grappolo/basic_util.h"
#include "grappolo/color_comm.h"
#include "grappolo/coloring.h"
#include "batch.hpp"
#include "coloring_modes.hpp"
#include "compare.hpp"
#include "grappolo/utilityClusteringFunctions.h"
#include "grappolo_hip.h"
#include "kernels.hpp"
#include "louvain.hpp"
#include "noalloc.hpp"
#include "mlp.hpp"
#include "profiler.hpp"

using glv::DBuf;
using glv::DevGraph;
using glv::Engine;
using glv::i64;

namespace {

thread_local std::string g_last_error;

struct DeviceCtx {
    hipStream_t stream = nullptr;
    bool ready = false;
    bool pool = false;
};
std::mutex g_mu;
DeviceCtx g_ctx[64];

int current_device() {
    int d = 0;
    GLV_HIP(hipGetDevice(&d));
    return d;
}

void shutdown_all();

// The engine's stream on the current device (the caller's when given).  On
// first use of a device its default memory pool is set to keep freed blocks
// (release threshold = max), whichever stream the calls come on, so that
// stream-ordered allocations are reused across synchronisations instead of
// going back to the driver.  The engine's own stream is a blocking stream, so
// a caller that passes no stream gets the order of the legacy default stream
// (work queued there before the call completes before the engine reads its
// inputs).  The first use registers the engine's shutdown to run at process
// exit, before the HIP runtime's own teardown (registered earlier).
hipStream_t engine_stream(void *user) {
    int d = current_device();
    std::lock_guard<std::mutex> lk(g_mu);
    static bool registered = false;
    if (!registered) {
        std::atexit(shutdown_all);
        registered = true;
    }
    DeviceCtx &c = g_ctx[d & 63];
    if (!c.pool) {
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, d) == hipSuccess) {
            uint64_t thr = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
        c.pool = true;
    }
    if (user) return (hipStream_t)user;
    if (!c.ready) {
        GLV_HIP(hipStreamCreate(&c.stream));
        c.ready = true;
    }
    return c.stream;
}

// Releases what the engine holds: cached device blocks, pinned staging
// buffers, profiler events, its streams.  Later calls recreate them.
void shutdown_all() {
    glv::blocks_shutdown();
    glv::Profiler::get().shutdown();
    std::lock_guard<std::mutex> lk(g_mu);
    int was = -1;
    (void)hipGetDevice(&was);
    for (int d = 0; d < 64; d++) {
        DeviceCtx &c = g_ctx[d];
        if (!c.ready) continue;
        (void)hipSetDevice(d);
        (void)hipStreamSynchronize(c.stream);
        (void)hipStreamDestroy(c.stream);
        c.stream = nullptr;
        c.ready = false;
    }
    if (was >= 0) (void)hipSetDevice(was);
}

template <class F>
int guarded(F fn) {
    try {
        fn();
        return GLV_OK;
    } catch (const glv::Error &e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return GLV_E_OOM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return GLV_E_DEVICE;
    }
}

void check_host_csr(i64 nv, const int64_t *ptrs, const glv_edge *edges) {
    if (nv < 0) glv::fail(GLV_E_INVALID_ARG, "negative vertex count");
    if (nv > 0 && !ptrs) glv::fail(GLV_E_INVALID_ARG, "null edgeListPtrs");
    if (nv > 0 && ptrs[0] != 0) glv::fail(GLV_E_INVALID_GRAPH, "edgeListPtrs[0] != 0");
    if (nv > 0 && ptrs[nv] > 0 && !edges) glv::fail(GLV_E_INVALID_ARG, "null edgeList");
    if (nv > 0 && ptrs[nv] < 0) glv::fail(GLV_E_INVALID_GRAPH, "negative entry count");
}

// Validate a device CSR (offsets monotone and in range, tails in [0, nv)).
void check_dev_csr(i64 nv, i64 nnz, const i64 *off, const i64 *tails, hipStream_t s) {
    if (nv < 0 || nnz < 0) glv::fail(GLV_E_INVALID_ARG, "negative size");
    if (nv == 0) return;
    if (!off || (nnz > 0 && !tails)) glv::fail(GLV_E_INVALID_ARG, "null device pointer");
    i64 ends[2];
    GLV_HIP(hipMemcpyAsync(&ends[0], off, sizeof(i64), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipMemcpyAsync(&ends[1], off + nv, sizeof(i64), hipMemcpyDeviceToHost, s));
    DBuf<int32_t> err(1, s);
    glv::launch_fill_i32(err.p, 1, 0, s);
    glv::launch_check_graph(off, tails, nv, nnz, err.p, s);
    int32_t e = 0;
    GLV_HIP(hipMemcpyAsync(&e, err.p, sizeof(e), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    if (ends[0] != 0 || ends[1] != nnz) glv::fail(GLV_E_INVALID_GRAPH, "offsets do not span [0, nnz]");
    if (e & 1) glv::fail(GLV_E_INVALID_GRAPH, "offsets not monotone");
    if (e & 2) glv::fail(GLV_E_INVALID_GRAPH, "tail id out of range");
}

// Caller-supplied labels: every label in [-1, limit), -1 only on vertices
// without entries.
void check_labels(const DevGraph &g, const i64 *C, i64 limit, hipStream_t s) {
    if (g.nv == 0) return;
    if (!C) glv::fail(GLV_E_INVALID_ARG, "null labels");
    DBuf<int32_t> err(1, s);
    glv::launch_fill_i32(err.p, 1, 0, s);
    glv::launch_check_labels(g.off, C, g.nv, limit, err.p, s);
    int32_t e = 0;
    GLV_HIP(hipMemcpyAsync(&e, err.p, sizeof(e), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    if (e & 1) glv::fail(GLV_E_INVALID_ARG, "cluster id out of range");
    if (e & 2) glv::fail(GLV_E_INVALID_ARG, "vertex with edges has cluster id -1");
}

DevGraph view(i64 nv, i64 nnz, const int64_t *off, const int64_t *tails, const double *w) {
    DevGraph g;
    g.nv = nv;
    g.nnz = nnz;
    g.num_edges = nnz / 2;
    g.off = off;
    g.tails = tails;
    g.w = w;
    return g;
}

// Upload the reference's host arrays; the AoS records are split on the device.
void upload(i64 nv, const int64_t *ptrs, const glv_edge *edges, DevGraph &g, hipStream_t s) {
    check_host_csr(nv, ptrs, edges);
    const i64 nnz = nv > 0 ? ptrs[nv] : 0;
    g.nv = nv;
    g.nnz = nnz;
    g.num_edges = nnz / 2;
    g.off_buf.alloc(nv + 1, s);
    g.tails_buf.alloc(nnz, s);
    g.w_buf.alloc(nnz, s);
    g.adopt();
    if (nv + 1 > 0 && ptrs)
        GLV_HIP(hipMemcpyAsync(g.off_buf.p, ptrs, (nv + 1) * sizeof(i64), hipMemcpyHostToDevice, s));
    if (nnz > 0) {
        DBuf<glv_edge> aos(nnz, s);
        GLV_HIP(hipMemcpyAsync(aos.p, edges, nnz * sizeof(glv_edge), hipMemcpyHostToDevice, s));
        glv::launch_split_edges(aos.p, nnz, g.tails_buf.p, g.w_buf.p, s);
    }
    check_dev_csr(nv, nnz, g.off, g.tails, s);
}

void fill_stats(const glv::RunResult &r, glv_stats *st) {
    if (!st) return;
    st->phases = r.phases;
    st->total_iterations = r.total_iterations;
    st->final_clusters = r.final_clusters;
    st->final_modularity = r.final_modularity;
    st->last_modularity = r.last_modularity;
    st->time_clustering = r.t_clustering;
    st->time_building = r.t_building;
    st->time_coloring = r.t_coloring;
    st->time_renumber = r.t_renumber;
    st->time_total = r.t_total;
    st->edges_visited = r.edges_visited;
    st->edges_visited_mod = r.edges_visited_mod;
    st->colors_first = r.colors_first;
    st->moves = r.moves;
    st->n_phase = (int32_t)r.phase.size();
    for (int i = 0; i < (int)r.phase.size() && i < st->phase_cap; i++) {
        const auto &p = r.phase[i];
        if (st->phase_nv) st->phase_nv[i] = p.nv;
        if (st->phase_nnz) st->phase_nnz[i] = p.nnz;
        if (st->phase_iters) st->phase_iters[i] = p.iters;
        if (st->phase_colors) st->phase_colors[i] = p.ncolors;
        if (st->phase_colored) st->phase_colored[i] = p.colored;
        if (st->phase_mod) st->phase_mod[i] = p.mod;
        if (st->phase_clusters) st->phase_clusters[i] = p.clusters;
        if (st->phase_seconds) st->phase_seconds[i] = p.t_sweep + p.t_renumber + p.t_coarsen + p.t_color;
    }
    st->n_trace = (int32_t)r.trace.size();
    for (int i = 0; i < (int)r.trace.size() && i < st->trace_cap; i++)
        if (st->trace_q) st->trace_q[i] = r.trace[i];
}

// No entry of the union may join two graphs of the batch.
void check_batch_blocks(const DevGraph &g, const std::vector<i64> &vb, hipStream_t s) {
    const i64 ng = (i64)vb.size() - 1;
    if (ng <= 0 || g.nnz == 0) return;
    DBuf<i64> d_vb(ng + 1, s);
    DBuf<int32_t> err(1, s);
    GLV_HIP(hipMemcpyAsync(d_vb.p, vb.data(), (ng + 1) * sizeof(i64), hipMemcpyHostToDevice, s));
    glv::launch_fill_i32(err.p, 1, 0, s);
    glv::launch_check_blocks(g.off, g.tails, d_vb.p, ng, err.p, s);
    int32_t e = 0;
    GLV_HIP(hipMemcpyAsync(&e, err.p, sizeof(e), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    if (e) glv::fail(GLV_E_INVALID_GRAPH, "an entry joins two graphs of the batch");
}

void fill_las(const glv::LasParams &L, glv_las_stats *st) {
    if (!st) return;
    st->swept_vertices = L.swept_vertices;
    st->swept_entries = L.swept_entries;
    st->full_sweeps = L.full_sweeps;
    st->pruned_sweeps = L.pruned_sweeps;
    st->samples = L.count;
}

// A graph without any entry has no defined modularity (1/(2m) is infinite);
// the reference loops on NaN there.  Every vertex is isolated: label -1.
bool edgeless(const DevGraph &g, int64_t *d_C_orig, hipStream_t s, glv::RunResult &r) {
    if (g.nnz > 0) return false;
    glv::launch_fill_i64(d_C_orig, g.nv, -1, s);
    GLV_HIP(hipStreamSynchronize(s));
    r = glv::RunResult();
    r.phases = g.nv > 0 ? 1 : 0;
    r.final_modularity = 0;
    r.last_modularity = 0;
    return true;
}

void run_pipeline(bool colored, const DevGraph &g, int64_t *d_C_orig, const glv::RunParams &p,
                  hipStream_t s, glv::RunResult &r) {
    if (edgeless(g, d_C_orig, s, r)) return;
    Engine E(s);
    if (colored)
        E.run_colored(g, d_C_orig, p, r);
    else
        E.run_basic(g, d_C_orig, p, r);
}

}  // namespace

namespace glv {
bool is_engine_stream(hipStream_t s) {
    if (s == nullptr) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    for (const DeviceCtx &c : g_ctx)
        if (c.ready && c.stream == s) return true;
    return false;
}
}  // namespace glv

// ============================================================== C ABI
extern "C" {

int glv_shutdown(void) {
    return guarded([&] { shutdown_all(); });
}

const char *glv_last_error(void) { return g_last_error.c_str(); }

const char *glv_version(void) { return "grappolo_amd 0.1 (gfx950)"; }

int glv_set_device(int device) {
    return guarded([&] { GLV_HIP(hipSetDevice(device)); });
}

int glv_get_device_count(int *count) {
    return guarded([&] {
        if (!count) glv::fail(GLV_E_INVALID_ARG, "null count");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

void glv_free(void *p) { free(p); }

int glv_runMultiPhaseColoring(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                              int64_t *C_orig, int coloring, int numColors, int replaceMap,
                              int64_t minGraphSize, double threshold, double C_threshold,
                              int numThreads, int threadsOpt, glv_stats *stats) {
    (void)numThreads;
    (void)threadsOpt;
    return guarded([&] {
        if (nv > 0 && !C_orig) glv::fail(GLV_E_INVALID_ARG, "null C_orig");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        DBuf<i64> dC(nv, s);
        glv::RunParams p;
        p.coloring = coloring;
        p.numColors = numColors;
        p.replaceMap = replaceMap;
        p.minGraphSize = minGraphSize;
        p.threshold = threshold;
        p.C_threshold = C_threshold;
        glv::RunResult r;
        run_pipeline(true, g, dC.p, p, s, r);
        if (nv > 0) GLV_HIP(hipMemcpyAsync(C_orig, dC.p, nv * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        fill_stats(r, stats);
    });
}

int glv_runMultiPhaseBasic(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                           int64_t *C_orig, int basicOpt, int64_t minGraphSize, double threshold,
                           double C_threshold, int numThreads, int threadsOpt, glv_stats *stats) {
    (void)minGraphSize;
    (void)C_threshold;
    (void)numThreads;
    (void)threadsOpt;
    return guarded([&] {
        if (nv > 0 && !C_orig) glv::fail(GLV_E_INVALID_ARG, "null C_orig");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        DBuf<i64> dC(nv, s);
        glv::RunParams p;
        p.replaceMap = basicOpt == 1;
        p.threshold = threshold;
        p.max_total_iters = 100000;  // runMultiPhaseBasic.cpp:116
        glv::RunResult r;
        run_pipeline(false, g, dC.p, p, s, r);
        if (nv > 0) GLV_HIP(hipMemcpyAsync(C_orig, dC.p, nv * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        fill_stats(r, stats);
    });
}

int glv_algoDistanceOneVertexColoringOpt(int64_t nv, const int64_t *edgeListPtrs,
                                         const glv_edge *edgeList, int32_t *vtxColor,
                                         int32_t *maxColor, double *totTime) {
    return guarded([&] {
        if (nv > 0 && (!vtxColor || !maxColor)) glv::fail(GLV_E_INVALID_ARG, "null output");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        DBuf<int32_t> dc(nv, s);
        Engine E(s);
        auto t0 = std::chrono::steady_clock::now();
        int mc = E.color(g, dc.p);
        auto t1 = std::chrono::steady_clock::now();
        if (nv > 0) GLV_HIP(hipMemcpyAsync(vtxColor, dc.p, nv * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        if (maxColor) *maxColor = mc;
        if (totTime) *totTime = std::chrono::duration<double>(t1 - t0).count();
    });
}

int glv_algoLouvainWithDistOneColoring(int64_t nv, const int64_t *edgeListPtrs,
                                       const glv_edge *edgeList, int64_t *C, const int32_t *color,
                                       int numColor, double Lower, double thresh,
                                       double *modularity, double *totTime, int32_t *numItr) {
    return guarded([&] {
        if (nv > 0 && (!C || !color)) glv::fail(GLV_E_INVALID_ARG, "null C or color");
        if (numColor < 1 && nv > 0) glv::fail(GLV_E_INVALID_ARG, "numColor < 1");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        if (g.nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        DBuf<i64> dC(nv, s);
        DBuf<int32_t> dcol(nv, s);
        GLV_HIP(hipMemcpyAsync(dcol.p, color, nv * sizeof(int32_t), hipMemcpyHostToDevice, s));
        Engine E(s);
        int it = 0;
        auto t0 = std::chrono::steady_clock::now();
        double q = E.colored_phase(g, dC.p, dcol.p, numColor, Lower, thresh, &it);
        auto t1 = std::chrono::steady_clock::now();
        GLV_HIP(hipMemcpyAsync(C, dC.p, nv * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        if (modularity) *modularity = q;
        if (numItr) *numItr = it;
        if (totTime) *totTime = std::chrono::duration<double>(t1 - t0).count();
    });
}

int glv_parallelLouvianMethod(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                              int64_t *C, double Lower, double thresh, int noMap, double *modularity,
                              double *totTime, int32_t *numItr) {
    return guarded([&] {
        if (nv > 0 && !C) glv::fail(GLV_E_INVALID_ARG, "null C");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        if (g.nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        DBuf<i64> dC(nv, s);
        Engine E(s);
        int it = 0;
        auto t0 = std::chrono::steady_clock::now();
        double q = E.plain_phase(g, dC.p, Lower, thresh, noMap != 0, &it);
        auto t1 = std::chrono::steady_clock::now();
        GLV_HIP(hipMemcpyAsync(C, dC.p, nv * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        if (modularity) *modularity = q;
        if (numItr) *numItr = it;
        if (totTime) *totTime = std::chrono::duration<double>(t1 - t0).count();
    });
}

int glv_renumberClustersContiguously(int64_t *C, int64_t size, int64_t *numClusters) {
    return guarded([&] {
        if (size < 0 || (size > 0 && !C)) glv::fail(GLV_E_INVALID_ARG, "bad C/size");
        hipStream_t s = engine_stream(nullptr);
        DBuf<i64> dC(size, s);
        if (size > 0) GLV_HIP(hipMemcpyAsync(dC.p, C, size * sizeof(i64), hipMemcpyHostToDevice, s));
        Engine E(s);
        i64 k = E.renumber(dC.p, size);
        if (size > 0) GLV_HIP(hipMemcpyAsync(C, dC.p, size * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        if (numClusters) *numClusters = k;
    });
}

static int coarsen_host(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList, const int64_t *C,
                        int64_t numUniqueClusters, int64_t **outEdgeListPtrs, glv_edge **outEdgeList,
                        int64_t *outNumEdges, bool vf) {
    return guarded([&] {
        if (!outEdgeListPtrs || !outEdgeList || (nv > 0 && !C))
            glv::fail(GLV_E_INVALID_ARG, "null argument");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g, out;
        upload(nv, edgeListPtrs, edgeList, g, s);
        DBuf<i64> dC(nv, s);
        if (nv > 0) GLV_HIP(hipMemcpyAsync(dC.p, C, nv * sizeof(i64), hipMemcpyHostToDevice, s));
        check_labels(g, dC.p, numUniqueClusters, s);
        Engine E(s);
        E.coarsen(g, dC.p, numUniqueClusters, out, vf);
        const i64 k = out.nv, tot = out.nnz;
        int64_t *hp = (int64_t *)malloc((k + 1) * sizeof(int64_t));
        glv_edge *he = (glv_edge *)malloc((tot > 0 ? tot : 1) * sizeof(glv_edge));
        std::vector<i64> tails(tot);
        std::vector<double> w(tot);
        if (!hp || !he) {
            free(hp);
            free(he);
            glv::fail(GLV_E_OOM, "host allocation failed");
        }
        GLV_HIP(hipMemcpyAsync(hp, out.off, (k + 1) * sizeof(i64), hipMemcpyDeviceToHost, s));
        if (tot > 0) {
            GLV_HIP(hipMemcpyAsync(tails.data(), out.tails, tot * sizeof(i64), hipMemcpyDeviceToHost, s));
            GLV_HIP(hipMemcpyAsync(w.data(), out.w, tot * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        GLV_HIP(hipStreamSynchronize(s));
        for (i64 i = 0; i < k; i++)
            for (i64 j = hp[i]; j < hp[i + 1]; j++) he[j] = glv_edge{i, tails[j], w[j]};
        *outEdgeListPtrs = hp;
        *outEdgeList = he;
        if (outNumEdges) *outNumEdges = out.num_edges;
    });
}

int glv_buildNextLevelGraphOpt(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                               const int64_t *C, int64_t numUniqueClusters,
                               int64_t **outEdgeListPtrs, glv_edge **outEdgeList,
                               int64_t *outNumEdges) {
    return coarsen_host(nv, edgeListPtrs, edgeList, C, numUniqueClusters, outEdgeListPtrs, outEdgeList,
                        outNumEdges, false);
}

int glv_buildNewGraphVF(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList, const int64_t *C,
                        int64_t numUniqueClusters, int64_t **outEdgeListPtrs, glv_edge **outEdgeList,
                        int64_t *outNumEdges) {
    return coarsen_host(nv, edgeListPtrs, edgeList, C, numUniqueClusters, outEdgeListPtrs, outEdgeList,
                        outNumEdges, true);
}

int glv_vertexFollowing(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList, int64_t *C,
                        int64_t *numNode) {
    return guarded([&] {
        if (nv > 0 && !C) glv::fail(GLV_E_INVALID_ARG, "null argument");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        DBuf<i64> dC(nv, s);
        Engine E(s);
        const i64 n = E.vertex_following(g, dC.p);
        if (nv > 0) GLV_HIP(hipMemcpyAsync(C, dC.p, nv * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        if (numNode) *numNode = n;
    });
}

int glv_modularity(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                   const int64_t *C, double *modularity) {
    return guarded([&] {
        if (!modularity || (nv > 0 && !C)) glv::fail(GLV_E_INVALID_ARG, "null argument");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        if (g.nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        DBuf<i64> dC(nv, s);
        GLV_HIP(hipMemcpyAsync(dC.p, C, nv * sizeof(i64), hipMemcpyHostToDevice, s));
        check_labels(g, dC.p, nv, s);
        Engine E(s);
        *modularity = E.modularity(g, dC.p);
    });
}

// ---------------------------------------------------------------- device forms
int glv_runMultiPhaseColoring_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets,
                                  const int64_t *d_tails, const double *d_weights, int64_t *d_C_orig,
                                  int coloring, int numColors, int replaceMap, int64_t minGraphSize,
                                  double threshold, double C_threshold, void *stream,
                                  glv_stats *stats) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        glv::RunParams p;
        p.coloring = coloring;
        p.numColors = numColors;
        p.replaceMap = replaceMap;
        p.minGraphSize = minGraphSize;
        p.threshold = threshold;
        p.C_threshold = C_threshold;
        glv::RunResult r;
        run_pipeline(true, g, d_C_orig, p, s, r);
        fill_stats(r, stats);
    });
}

int glv_runMultiPhaseBasic_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets,
                               const int64_t *d_tails, const double *d_weights, int64_t *d_C_orig,
                               int basicOpt, double threshold, void *stream, glv_stats *stats) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        glv::RunParams p;
        p.replaceMap = basicOpt == 1;
        p.threshold = threshold;
        p.max_total_iters = 100000;
        glv::RunResult r;
        run_pipeline(false, g, d_C_orig, p, s, r);
        fill_stats(r, stats);
    });
}

int glv_color_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                  int32_t *d_colors, int32_t *maxColor, int32_t *rounds, void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, nullptr);
        Engine E(s);
        int r = 0;
        int mc = E.color(g, d_colors, &r);
        if (maxColor) *maxColor = mc;
        if (rounds) *rounds = r;
    });
}

int glv_colored_phase_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                          const double *d_weights, int64_t *d_C, const int32_t *d_colors,
                          int numColor, double Lower, double thresh, double *modularity,
                          int32_t *numItr, void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        if (nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        if (numColor < 1) glv::fail(GLV_E_INVALID_ARG, "numColor < 1");
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        Engine E(s);
        int it = 0;
        double q = E.colored_phase(g, d_C, d_colors, numColor, Lower, thresh, &it);
        E.sync();
        if (modularity) *modularity = q;
        if (numItr) *numItr = it;
    });
}

int glv_plain_phase_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                        const double *d_weights, int64_t *d_C, double Lower, double thresh,
                        int noMap, double *modularity, int32_t *numItr, void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        if (nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        Engine E(s);
        int it = 0;
        double q = E.plain_phase(g, d_C, Lower, thresh, noMap != 0, &it);
        E.sync();
        if (modularity) *modularity = q;
        if (numItr) *numItr = it;
    });
}

int glv_color_step_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                       const double *d_weights, int64_t *d_C, const int32_t *d_colors, int numColor,
                       int which, int64_t *moves, void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        if (nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        if (numColor < 1 || which < 0 || which >= numColor) glv::fail(GLV_E_INVALID_ARG, "bad color");
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        check_labels(g, d_C, nv, s);
        Engine E(s);
        i64 m = E.color_step(g, d_C, d_colors, numColor, which);
        if (moves) *moves = m;
    });
}

int glv_modularity_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                       const double *d_weights, const int64_t *d_C, double *modularity,
                       void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        if (nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        check_labels(g, d_C, nv, s);
        Engine E(s);
        double q = E.modularity(g, d_C);
        if (modularity) *modularity = q;
    });
}

int glv_modularity_noalloc_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                               const double *d_weights, const int64_t *d_C, double *modularity,
                               void *stream) {
    return guarded([&] {
        if (nv < 0 || nnz < 0) glv::fail(GLV_E_INVALID_ARG, "negative size");
        if (nv > glv::noalloc_max_nv())
            glv::fail(GLV_E_INVALID_ARG, "graph has more vertices than the no-allocation scratch holds (" +
                                             std::to_string(glv::noalloc_max_nv()) + ")");
        if (nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        if (!d_offsets || !d_tails || !d_weights || (nv > 0 && !d_C))
            glv::fail(GLV_E_INVALID_ARG, "null device pointer");
        hipStream_t s = engine_stream(stream);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        const double q = glv::modularity_noalloc(g, d_C, s);
        if (modularity) *modularity = q;
    });
}

int glv_noalloc_max_vertices(int64_t *max_nv) {
    return guarded([&] {
        if (!max_nv) glv::fail(GLV_E_INVALID_ARG, "null argument");
        *max_nv = glv::noalloc_max_nv();
    });
}

int glv_renumber_dev(int64_t n, int64_t *d_C, int64_t *numClusters, void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        Engine E(s);
        i64 k = E.renumber(d_C, n);
        if (numClusters) *numClusters = k;
    });
}

int glv_coarsen_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                    const double *d_weights, const int64_t *d_C, int64_t k, int64_t **d_out_offsets,
                    int64_t **d_out_tails, double **d_out_weights, int64_t *out_nnz,
                    int64_t *out_num_edges, void *stream) {
    return guarded([&] {
        if (!d_out_offsets || !d_out_tails || !d_out_weights) glv::fail(GLV_E_INVALID_ARG, "null output");
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights), out;
        check_labels(g, d_C, k, s);
        Engine E(s);
        E.coarsen(g, d_C, k, out);
        E.sync();
        // hand the buffers to the caller (released with glv_dev_free)
        *d_out_offsets = out.off_buf.p;
        *d_out_tails = out.tails_buf.p;
        *d_out_weights = out.w_buf.p;
        out.off_buf.p = nullptr;
        out.tails_buf.p = nullptr;
        out.w_buf.p = nullptr;
        if (out_nnz) *out_nnz = out.nnz;
        if (out_num_edges) *out_num_edges = out.num_edges;
    });
}

int glv_vertexFollowing_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                            int64_t *d_C, int64_t *numNode, void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, nullptr);
        Engine E(s);
        const i64 n = E.vertex_following(g, d_C);
        if (numNode) *numNode = n;
    });
}

int glv_buildNewGraphVF_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                            const double *d_weights, const int64_t *d_C, int64_t k, int64_t **d_out_offsets,
                            int64_t **d_out_tails, double **d_out_weights, int64_t *out_nnz,
                            int64_t *out_num_edges, void *stream) {
    return guarded([&] {
        if (!d_out_offsets || !d_out_tails || !d_out_weights) glv::fail(GLV_E_INVALID_ARG, "null output");
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights), out;
        check_labels(g, d_C, k, s);
        Engine E(s);
        E.coarsen(g, d_C, k, out, true);
        E.sync();
        *d_out_offsets = out.off_buf.p;
        *d_out_tails = out.tails_buf.p;
        *d_out_weights = out.w_buf.p;
        out.off_buf.p = nullptr;
        out.tails_buf.p = nullptr;
        out.w_buf.p = nullptr;
        if (out_nnz) *out_nnz = out.nnz;
        if (out_num_edges) *out_num_edges = out.num_edges;
    });
}

int glv_vertexFollowingPrepass_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                                   const double *d_weights, int64_t **d_out_offsets, int64_t **d_out_tails,
                                   double **d_out_weights, int64_t *out_nv, int64_t *out_nnz,
                                   int64_t *out_num_edges, int64_t *numNode, void *stream) {
    return guarded([&] {
        if (!d_out_offsets || !d_out_tails || !d_out_weights || !out_nv)
            glv::fail(GLV_E_INVALID_ARG, "null output");
        *d_out_offsets = nullptr;
        *d_out_tails = nullptr;
        *d_out_weights = nullptr;
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights), out;
        Engine E(s);
        DBuf<i64> C(std::max<i64>(nv, 1), s);
        const i64 n = E.vertex_following(g, C.p);
        if (numNode) *numNode = n;
        *out_nv = nv;
        if (out_nnz) *out_nnz = nnz;
        if (out_num_edges) *out_num_edges = nnz / 2;
        if (n == 0) return;  // graph unchanged (driverForGraphClustering.cpp:118)
        const i64 k = E.renumber(C.p, nv);
        E.coarsen(g, C.p, k, out, true);
        E.sync();
        *d_out_offsets = out.off_buf.p;
        *d_out_tails = out.tails_buf.p;
        *d_out_weights = out.w_buf.p;
        out.off_buf.p = nullptr;
        out.tails_buf.p = nullptr;
        out.w_buf.p = nullptr;
        *out_nv = k;
        if (out_nnz) *out_nnz = out.nnz;
        if (out_num_edges) *out_num_edges = out.num_edges;
    });
}

int glv_dev_free(void *d_ptr) {
    return guarded([&] {
        if (!d_ptr) return;
        hipStream_t s = engine_stream(nullptr);
        GLV_HIP(hipFreeAsync(d_ptr, s));
        GLV_HIP(hipStreamSynchronize(s));
    });
}

int glv_profile_enable(int on) {
    return guarded([&] { glv::Profiler::get().enable(on != 0); });
}

int glv_profile_reset(void) {
    return guarded([&] { glv::Profiler::get().reset(); });
}

int glv_profile_read(int kernel_class, double *ms, double *bytes, int64_t *launches) {
    return guarded([&] {
        if (kernel_class < 0 || kernel_class >= glv::K_NUM) glv::fail(GLV_E_INVALID_ARG, "bad class");
        glv::Profiler::get().collect();
        glv::KStat st = glv::Profiler::get().stat(kernel_class);
        if (ms) *ms = st.ms;
        if (bytes) *bytes = st.bytes;
        if (launches) *launches = st.launches;
    });
}

int glv_count_conflicts_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets,
                            const int64_t *d_tails, const int32_t *d_colors, int64_t *conflicts,
                            void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, nullptr);
        Engine E(s);
        i64 c = E.count_conflicts(g, d_colors);
        if (conflicts) *conflicts = c;
    });
}

int glv_mlp_param_count(void) { return glv::kMlpParams; }

int glv_mlp_forward_dev(const float *d_params, const uint16_t *d_feats, int64_t n, float *d_logits,
                        void *stream) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!d_params || !d_feats || !d_logits)))
            glv::fail(GLV_E_INVALID_ARG, "null pointer or negative size");
        hipStream_t s = engine_stream(stream);
        glv::launch_mlp_forward(d_params, d_feats, n, d_logits, s);
        GLV_HIP(hipStreamSynchronize(s));
    });
}

int glv_mlp_grad_dev(const float *d_params, const uint16_t *d_feats, const uint8_t *d_labels, int64_t n,
                     float pos_weight, float *d_grad, void *stream) {
    return guarded([&] {
        if (n <= 0 || !d_params || !d_feats || !d_labels || !d_grad)
            glv::fail(GLV_E_INVALID_ARG, "null pointer or empty sample set");
        hipStream_t s = engine_stream(stream);
        glv::MlpBatch B;
        B.feats = d_feats;
        B.labels = d_labels;
        B.n = n;
        B.batch = n;
        B.pos_weight = pos_weight;
        const i64 nw = glv::mlp_grad_waves(n);
        glv::DBuf<float> gpart(nw * glv::kMlpStride, s);
        glv::launch_mlp_grad_partials(d_params, B, gpart.p, s);
        glv::launch_mlp_reduce(gpart.p, nw, d_grad, s);
        GLV_HIP(hipStreamSynchronize(s));
    });
}

int glv_mlp_train_dev(float *d_params, float *d_m, float *d_v, const uint16_t *d_feats,
                      const uint8_t *d_labels, int64_t n, int64_t steps, int64_t batch, float lr,
                      float pos_weight, uint64_t seed, int64_t step0, float *d_losses, void *stream) {
    return guarded([&] {
        if (n <= 0 || batch <= 0 || steps < 0 || step0 < 0 || !d_params || !d_m || !d_v || !d_feats ||
            !d_labels)
            glv::fail(GLV_E_INVALID_ARG, "null pointer or non-positive size");
        hipStream_t s = engine_stream(stream);
        glv::MlpBatch B;
        B.feats = d_feats;
        B.labels = d_labels;
        B.n = n;
        B.batch = batch;
        B.random = 1;
        B.seed = seed;
        B.pos_weight = pos_weight;
        const i64 nw = glv::mlp_grad_waves(batch);
        glv::DBuf<float> gpart(nw * glv::kMlpStride, s);
        const double b1 = 0.9, b2 = 0.999;
        for (int64_t t = 0; t < steps; t++) {
            B.step = step0 + t;
            const double k = (double)(step0 + t + 1);
            const float c1 = (float)(1.0 / (1.0 - std::pow(b1, k)));
            const float c2 = (float)(1.0 / (1.0 - std::pow(b2, k)));
            glv::launch_mlp_grad_partials(d_params, B, gpart.p, s);
            glv::launch_mlp_adam(d_params, d_m, d_v, gpart.p, nw, lr, (float)b1, (float)b2, 1e-8f, c1, c2,
                                 d_losses ? d_losses + t : nullptr, s);
        }
        GLV_HIP(hipStreamSynchronize(s));
    });
}

int glv_las_trace_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                      const double *d_weights, int64_t *d_C_orig, int64_t minGraphSize, double threshold,
                      double C_threshold, double rate, uint64_t seed, uint16_t *d_feats, int64_t *d_verts,
                      uint8_t *d_labels, int64_t cap, void *stream, glv_stats *stats,
                      glv_las_stats *las) {
    return guarded([&] {
        if (cap < 0 || (cap > 0 && (!d_feats || !d_verts || !d_labels)) || !(rate >= 0 && rate <= 1))
            glv::fail(GLV_E_INVALID_ARG, "bad trace buffers or rate");
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        glv::LasParams L;
        L.mode = 1;
        L.rate = rate;
        L.seed = seed;
        L.tfeats = d_feats;
        L.tverts = d_verts;
        L.tlabels = d_labels;
        L.cap = cap;
        glv::RunParams p;
        p.minGraphSize = minGraphSize;
        p.threshold = threshold;
        p.C_threshold = C_threshold;
        p.las = &L;
        glv::RunResult r;
        run_pipeline(true, g, d_C_orig, p, s, r);
        fill_stats(r, stats);
        fill_las(L, las);
    });
}

int glv_runMultiPhaseColoring_las_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets,
                                      const int64_t *d_tails, const double *d_weights, int64_t *d_C_orig,
                                      int64_t minGraphSize, double threshold, double C_threshold,
                                      const float *d_params, float tau, int max_rechecks,
                                      void *stream, glv_stats *stats, glv_las_stats *las) {
    return guarded([&] {
        if (!d_params) glv::fail(GLV_E_INVALID_ARG, "null MLP parameters");
        if (!(tau >= 0.0f && tau < 1.0f)) glv::fail(GLV_E_INVALID_ARG, "tau outside [0, 1)");
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        glv::LasParams L;
        L.mode = 2;
        L.P = d_params;
        if (max_rechecks < 0) glv::fail(GLV_E_INVALID_ARG, "max_rechecks < 0");
        L.logit_tau = tau > 0.0f ? std::log(tau / (1.0f - tau)) : -INFINITY;
        L.max_rechecks = max_rechecks;
        glv::RunParams p;
        p.minGraphSize = minGraphSize;
        p.threshold = threshold;
        p.C_threshold = C_threshold;
        p.las = &L;
        glv::RunResult r;
        run_pipeline(true, g, d_C_orig, p, s, r);
        fill_stats(r, stats);
        fill_las(L, las);
    });
}

int glv_runMultiPhaseBasic_batch_dev(int64_t ngraphs, const int64_t *vbase, int64_t nv, int64_t nnz,
                                     const int64_t *d_offsets, const int64_t *d_tails, const double *d_weights,
                                     int64_t *d_C_orig, int basicOpt, double threshold, void *stream,
                                     glv_batch_stats *stats) {
    return guarded([&] {
        if (ngraphs < 0 || (ngraphs > 0 && !vbase)) glv::fail(GLV_E_INVALID_ARG, "bad graph table");
        if (ngraphs > 0 && (vbase[0] != 0 || vbase[ngraphs] != nv))
            glv::fail(GLV_E_INVALID_ARG, "vbase must start at 0 and end at nv");
        for (int64_t g = 0; g < ngraphs; g++)
            if (vbase[g + 1] < vbase[g]) glv::fail(GLV_E_INVALID_ARG, "vbase not monotone");
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        std::vector<i64> vb(vbase, vbase + ngraphs + 1);
        if (ngraphs == 0) vb.assign(1, 0);
        check_batch_blocks(g, vb, s);
        glv::BatchParams p;
        p.basicOpt = basicOpt;
        p.threshold = threshold;
        glv::BatchResult r;
        Engine E(s);
        glv::run_basic_batch(E, g, vb, d_C_orig, p, r);
        if (stats) {
            for (int64_t i = 0; i < ngraphs; i++) {
                if (stats->final_modularity) stats->final_modularity[i] = r.final_modularity[i];
                if (stats->last_modularity) stats->last_modularity[i] = r.last_modularity[i];
                if (stats->phases) stats->phases[i] = r.phases[i];
                if (stats->total_iterations) stats->total_iterations[i] = r.total_iterations[i];
                if (stats->final_clusters) stats->final_clusters[i] = r.clusters[i];
            }
            stats->levels = r.levels;
            stats->iterations = r.iterations;
            stats->time_total = r.t_total;
            stats->launches = r.launches;
        }
    });
}

int glv_get_bins(int64_t *max_deg) {
    return guarded([&] {
        if (!max_deg) glv::fail(GLV_E_INVALID_ARG, "null array");
        for (int b = 0; b < glv::kNumBins; b++) max_deg[b] = glv::current_bins().max_deg[b];
    });
}

int glv_set_bins(const int64_t *max_deg) {
    return guarded([&] {
        if (!max_deg) glv::fail(GLV_E_INVALID_ARG, "null array");
        glv::BinConfig c;
        for (int b = 0; b < glv::kNumBins; b++) c.max_deg[b] = max_deg[b];
        if (!glv::valid_bins(c)) glv::fail(GLV_E_INVALID_ARG, "bin ranges outside the kernels' capacities");
        glv::current_bins() = c;
    });
}

int glv_autotune_bins_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                          const double *d_weights, int reps, int64_t *max_deg, double *ms_default,
                          double *ms_tuned, void *stream) {
    return guarded([&] {
        if (!max_deg || reps < 1) glv::fail(GLV_E_INVALID_ARG, "null array or reps < 1");
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        if (nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        Engine E(s);
        const glv::BinConfig c = glv::autotune_bins(E, g, reps, ms_default, ms_tuned);
        for (int b = 0; b < glv::kNumBins; b++) max_deg[b] = c.max_deg[b];
    });
}

int glv_get_tuning(glv_tuning *t) {
    return guarded([&] {
        if (!t) glv::fail(GLV_E_INVALID_ARG, "null tuning");
        const glv::Tuning &c = glv::current_tuning();
        *t = glv_tuning{c.jp_blocks, c.persist_blocks, c.wide_blocks, 0, c.core_min_deg};
    });
}

int glv_set_tuning(const glv_tuning *t) {
    return guarded([&] {
        if (!t) glv::fail(GLV_E_INVALID_ARG, "null tuning");
        glv::Tuning c;
        c.jp_blocks = t->jp_blocks;
        c.persist_blocks = t->persist_blocks;
        c.wide_blocks = t->wide_blocks;
        c.core_min_deg = t->core_min_deg;
        if (!glv::valid_tuning(c)) glv::fail(GLV_E_INVALID_ARG, "launch shapes outside the kernels' limits");
        glv::current_tuning() = c;
    });
}

int glv_autotune_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                     const double *d_weights, int coloring, int64_t minGraphSize, double threshold,
                     double C_threshold, int reps, int runs, int apply, glv_tuning *best,
                     glv_tune_trial *trials, int cap, int *ntrials, void *stream) {
    return guarded([&] {
        if (!best || reps < 1 || cap < 0 || (cap > 0 && !trials)) glv::fail(GLV_E_INVALID_ARG, "bad output arguments");
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        if (nnz == 0) glv::fail(GLV_E_INVALID_GRAPH, "graph has no edges");
        DevGraph g = view(nv, nnz, d_offsets, d_tails, d_weights);
        glv::RunParams p;
        p.coloring = coloring;
        p.minGraphSize = minGraphSize;
        p.threshold = threshold;
        p.C_threshold = C_threshold;
        Engine E(s);
        std::vector<glv::TuneTrial> log;
        const glv::Tuning t = glv::autotune_launch(E, g, p, reps, runs != 0, &log);
        GLV_HIP(hipStreamSynchronize(s));
        *best = glv_tuning{t.jp_blocks, t.persist_blocks, t.wide_blocks, 0, t.core_min_deg};
        const int n = (int)std::min<size_t>(log.size(), (size_t)cap);
        for (int i = 0; i < n; i++) {
            const glv::Tuning &c = log[i].t;
            trials[i] = glv_tune_trial{{c.jp_blocks, c.persist_blocks, c.wide_blocks, 0, c.core_min_deg},
                                       log[i].stage, 0, log[i].ms};
        }
        if (ntrials) *ntrials = (int)log.size();
        if (apply) glv::current_tuning() = t;
    });
}

int glv_algoColoringMultiHashMaxMin_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                                        int32_t *d_colors, int nHash, int nItrs, int32_t *maxColor, void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, nullptr);
        const int mc = glv::color_multihash(g, d_colors, nHash, nItrs, s);
        GLV_HIP(hipStreamSynchronize(s));
        if (maxColor) *maxColor = mc;
    });
}

int glv_algoColoringMultiHashMaxMin(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                                    int32_t *vtxColor, int nHash, int nItrs, int32_t *maxColor) {
    return guarded([&] {
        if (nv > 0 && !vtxColor) glv::fail(GLV_E_INVALID_ARG, "null colors");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        DBuf<int32_t> c(nv, s);
        const int mc = glv::color_multihash(g, c.p, nHash, nItrs, s);
        if (nv > 0) GLV_HIP(hipMemcpyAsync(vtxColor, c.p, nv * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        if (maxColor) *maxColor = mc;
    });
}

int glv_vBaseRedistribution_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                                int32_t *d_colors, int ncolors, int type, int32_t *rounds, void *stream) {
    return guarded([&] {
        hipStream_t s = engine_stream(stream);
        check_dev_csr(nv, nnz, d_offsets, d_tails, s);
        DevGraph g = view(nv, nnz, d_offsets, d_tails, nullptr);
        const int r = glv::balance_colors(g, d_colors, ncolors, type, s);
        GLV_HIP(hipStreamSynchronize(s));
        if (rounds) *rounds = r;
    });
}

int glv_vBaseRedistribution(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList, int32_t *vtxColor,
                            int ncolors, int type, int32_t *rounds) {
    return guarded([&] {
        if (nv > 0 && !vtxColor) glv::fail(GLV_E_INVALID_ARG, "null colors");
        hipStream_t s = engine_stream(nullptr);
        DevGraph g;
        upload(nv, edgeListPtrs, edgeList, g, s);
        DBuf<int32_t> c(nv, s);
        if (nv > 0) GLV_HIP(hipMemcpyAsync(c.p, vtxColor, nv * sizeof(int32_t), hipMemcpyHostToDevice, s));
        const int r = glv::balance_colors(g, c.p, ncolors, type, s);
        if (nv > 0) GLV_HIP(hipMemcpyAsync(vtxColor, c.p, nv * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        if (rounds) *rounds = r;
    });
}

static void fill_comparison(const glv::CompareResult &r, glv_cluster_comparison *out) {
    out->nC1 = r.nC1;
    out->nC2 = r.nC2;
    out->SameSame = r.SameSame;
    out->SameDiff = r.SameDiff;
    out->DiffSame = r.DiffSame;
    out->Agree = r.Agree;
    out->precision = r.precision;
    out->recall = r.recall;
    out->fScore = r.fScore;
    out->gini1 = r.gini1;
    out->gini2 = r.gini2;
}

int glv_computeCommunityComparisons_dev(int64_t *d_C1, int64_t N1, int64_t *d_C2, int64_t N2,
                                        glv_cluster_comparison *out, void *stream) {
    return guarded([&] {
        if (!out) glv::fail(GLV_E_INVALID_ARG, "null output");
        if (N1 != N2) glv::fail(GLV_E_INVALID_ARG, "the two labellings differ in length");
        hipStream_t s = engine_stream(stream);
        glv::CompareResult r;
        glv::compare_communities(d_C1, d_C2, N1, r, s);
        fill_comparison(r, out);
    });
}

int glv_computeCommunityComparisons(int64_t *C1, int64_t N1, int64_t *C2, int64_t N2,
                                    glv_cluster_comparison *out) {
    return guarded([&] {
        if (!out || (N1 > 0 && (!C1 || !C2))) glv::fail(GLV_E_INVALID_ARG, "null argument");
        if (N1 != N2) glv::fail(GLV_E_INVALID_ARG, "the two labellings differ in length");
        if (N1 <= 0) glv::fail(GLV_E_INVALID_ARG, "comparison needs at least one vertex");
        hipStream_t s = engine_stream(nullptr);
        DBuf<i64> a(N1, s), b(N1, s);
        GLV_HIP(hipMemcpyAsync(a.p, C1, N1 * sizeof(i64), hipMemcpyHostToDevice, s));
        GLV_HIP(hipMemcpyAsync(b.p, C2, N1 * sizeof(i64), hipMemcpyHostToDevice, s));
        glv::CompareResult r;
        glv::compare_communities(a.p, b.p, N1, r, s);
        GLV_HIP(hipMemcpyAsync(C1, a.p, N1 * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipMemcpyAsync(C2, b.p, N1 * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        fill_comparison(r, out);
    });
}

int glv_computeGiniCoefficient_dev(int64_t *d_colorSize, int numColors, double *gini, void *stream) {
    return guarded([&] {
        if (!gini || (numColors > 0 && !d_colorSize)) glv::fail(GLV_E_INVALID_ARG, "null argument");
        hipStream_t s = engine_stream(stream);
        *gini = glv::gini_coefficient(d_colorSize, std::max(numColors, 0), numColors, s);
    });
}

int glv_computeGiniCoefficient(int64_t *colorSize, int numColors, double *gini) {
    return guarded([&] {
        if (!gini || (numColors > 0 && !colorSize)) glv::fail(GLV_E_INVALID_ARG, "null argument");
        hipStream_t s = engine_stream(nullptr);
        const i64 n = std::max(numColors, 0);
        DBuf<i64> d(n, s);
        if (n) GLV_HIP(hipMemcpyAsync(d.p, colorSize, n * sizeof(i64), hipMemcpyHostToDevice, s));
        *gini = glv::gini_coefficient(d.p, n, numColors, s);
        if (n) GLV_HIP(hipMemcpyAsync(colorSize, d.p, n * sizeof(i64), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"

// ===================================================== C++ drop-in (reference API)
namespace {

[[noreturn]] void die(const char *fn, int rc) {
    fprintf(stderr, "%s failed (status %d): %s\n", fn, rc, glv_last_error());
    exit(1);
}

void free_graph(graph *G) {
    if (!G) return;
    free(G->edgeListPtrs);
    free(G->edgeList);
    free(G);
}

void print_summary(const glv_stats &st, int coloring, int numThreads) {
    printf("********************************************\n");
    printf("*********    Compact Summary   *************\n");
    printf("********************************************\n");
    printf("Number of threads              : %d\n", numThreads);
    printf("Total number of phases         : %d\n", st.phases);
    printf("Total number of iterations     : %d\n", st.total_iterations);
    printf("Final number of clusters       : %ld\n", (long)st.final_clusters);
    printf("Final modularity               : %lf\n", st.final_modularity);
    printf("Total time for clustering      : %lf\n", st.time_clustering);
    printf("Total time for building phases : %lf\n", st.time_building);
    if (coloring == 1) printf("Total time for coloring        : %lf\n", st.time_coloring);
    printf("********************************************\n");
    printf("TOTAL TIME                     : %lf\n",
           st.time_clustering + st.time_building + st.time_coloring);
    printf("********************************************\n");
}

}  // namespace

void runMultiPhaseColoring(graph *G, long *C_orig, int coloring, int numColors, int replaceMap,
                           long minGraphSize, double threshold, double C_threshold, int numThreads,
                           int threadsOpt) {
    glv_stats st;
    memset(&st, 0, sizeof(st));
    int rc = glv_runMultiPhaseColoring(G->numVertices, G->edgeListPtrs, (const glv_edge *)G->edgeList,
                                       (int64_t *)C_orig, coloring, numColors, replaceMap, minGraphSize,
                                       threshold, C_threshold, numThreads, threadsOpt, &st);
    if (rc) die("runMultiPhaseColoring", rc);
    print_summary(st, coloring, numThreads);
    free_graph(G);  // runMultiPhaseColoring.cpp:227-232: the routine owns G
}

void runMultiPhaseBasic(graph *G, long *C_orig, int basicOpt, long minGraphSize, double threshold,
                        double C_threshold, int numThreads, int threadsOpt) {
    glv_stats st;
    memset(&st, 0, sizeof(st));
    int rc = glv_runMultiPhaseBasic(G->numVertices, G->edgeListPtrs, (const glv_edge *)G->edgeList,
                                    (int64_t *)C_orig, basicOpt, minGraphSize, threshold, C_threshold,
                                    numThreads, threadsOpt, &st);
    if (rc) die("runMultiPhaseBasic", rc);
    print_summary(st, 0, numThreads);
    free_graph(G);  // runMultiPhaseBasic.cpp:164-169
}

double algoLouvainWithDistOneColoring(graph *G, long *C, int nThreads, int *color, int numColor,
                                      double Lower, double thresh, double *totTime, int *numItr) {
    (void)nThreads;
    double q = 0, t = 0;
    int32_t it = 0;
    int rc = glv_algoLouvainWithDistOneColoring(G->numVertices, G->edgeListPtrs,
                                                (const glv_edge *)G->edgeList, (int64_t *)C,
                                                (const int32_t *)color, numColor, Lower, thresh, &q,
                                                &t, &it);
    if (rc) die("algoLouvainWithDistOneColoring", rc);
    if (totTime) *totTime = t;
    if (numItr) *numItr = it;
    return q;
}

double algoLouvainWithDistOneColoringNoMap(graph *G, long *C, int nThreads, int *color, int numColor,
                                           double Lower, double thresh, double *totTime, int *numItr) {
    // the NoMap variant computes the same targets (parallelLouvainWithColoringNoMap.cpp)
    return algoLouvainWithDistOneColoring(G, C, nThreads, color, numColor, Lower, thresh, totTime,
                                          numItr);
}

int algoDistanceOneVertexColoringOpt(graph *G, int *vtxColor, int nThreads, double *totTime) {
    (void)nThreads;
    int32_t mc = -1;
    double t = 0;
    int rc = glv_algoDistanceOneVertexColoringOpt(G->numVertices, G->edgeListPtrs,
                                                  (const glv_edge *)G->edgeList, (int32_t *)vtxColor,
                                                  &mc, &t);
    if (rc) die("algoDistanceOneVertexColoringOpt", rc);
    if (totTime) *totTime = t;
    return mc;
}

double parallelLouvianMethod(graph *G, long *C, int nThreads, double Lower, double thresh,
                             double *totTime, int *numItr) {
    (void)nThreads;
    double q = 0, t = 0;
    int32_t it = 0;
    int rc = glv_parallelLouvianMethod(G->numVertices, G->edgeListPtrs, (const glv_edge *)G->edgeList,
                                       (int64_t *)C, Lower, thresh, 0, &q, &t, &it);
    if (rc) die("parallelLouvianMethod", rc);
    if (totTime) *totTime = t;
    if (numItr) *numItr = it;
    return q;
}

double parallelLouvianMethodNoMap(graph *G, long *C, int nThreads, double Lower, double thresh,
                                  double *totTime, int *numItr) {
    (void)nThreads;
    double q = 0, t = 0;
    int32_t it = 0;
    int rc = glv_parallelLouvianMethod(G->numVertices, G->edgeListPtrs, (const glv_edge *)G->edgeList,
                                       (int64_t *)C, Lower, thresh, 1, &q, &t, &it);
    if (rc) die("parallelLouvianMethodNoMap", rc);
    if (totTime) *totTime = t;
    if (numItr) *numItr = it;
    return q;
}

long renumberClustersContiguously(long *C, long size) {
    int64_t k = 0;
    int rc = glv_renumberClustersContiguously((int64_t *)C, size, &k);
    if (rc) die("renumberClustersContiguously", rc);
    return k;
}

int algoColoringMultiHashMaxMin(graph *G, int *vtxColor, int nThreads, double *totTime, int nHash, int nItrs) {
    (void)nThreads;
    int32_t mc = 0;
    auto t0 = std::chrono::steady_clock::now();
    int rc = glv_algoColoringMultiHashMaxMin(G->numVertices, G->edgeListPtrs, (const glv_edge *)G->edgeList,
                                             (int32_t *)vtxColor, nHash, nItrs, &mc);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) die("algoColoringMultiHashMaxMin", rc);
    if (totTime) *totTime = std::chrono::duration<double>(t1 - t0).count();
    return mc;
}

int vBaseRedistribution(graph *G, int *vtxColor, int ncolors, int type) {
    int32_t rounds = 0;
    int rc = glv_vBaseRedistribution(G->numVertices, G->edgeListPtrs, (const glv_edge *)G->edgeList,
                                     (int32_t *)vtxColor, ncolors, type, &rounds);
    if (rc) die("vBaseRedistribution", rc);
    return rounds;
}

void computeCommunityComparisons(std::vector<long> &C1, long N1, std::vector<long> &C2, long N2) {
    if ((long)C1.size() < N1 || (long)C2.size() < N2) die("computeCommunityComparisons", GLV_E_INVALID_ARG);
    glv_cluster_comparison r;
    int rc = glv_computeCommunityComparisons((int64_t *)C1.data(), N1, (int64_t *)C2.data(), N2, &r);
    if (rc) die("computeCommunityComparisons", rc);
    // the report of utilityClusterComparisonMetrics.cpp:247-264
    printf("*******************************************\n");
    printf("Cluster comparison statistics: \n");
    printf("*******************************************\n");
    printf("|C1| (truth)          : %ld\n", N1);
    printf("Num communities in C1 : %ld\n", (long)r.nC1);
    printf("|C2| (output)         : %ld\n", N2);
    printf("Num communities in C2 : %ld\n", (long)r.nC2);
    printf("-------------------------------------------\n");
    printf("Same-Same (True positive)  : %ld\n", (long)r.SameSame);
    printf("Same-Diff (False negative) : %ld\n", (long)r.SameDiff);
    printf("Diff-Same (False positive) : %ld\n", (long)r.DiffSame);
    printf("-------------------------------------------\n");
    printf("Precision             :  %lf (%3.2lf%%)\n", r.precision, r.precision * 100);
    printf("Recall                :  %lf (%3.2lf%%)\n", r.recall, r.recall * 100);
    printf("F-score               :  %lf\n", r.fScore);
    printf("-------------------------------------------\n");
    printf("Gini coefficient, C1  :  %lf \n", r.gini1);
    printf("Gini coefficient, C2  :  %lf \n", r.gini2);
    printf("*******************************************\n");
}

double computeGiniCoefficient(long *colorSize, int numColors) {
    double g = 0;
    int rc = glv_computeGiniCoefficient((int64_t *)colorSize, numColors, &g);
    if (rc) die("computeGiniCoefficient", rc);
    return g;
}

long vertexFollowing(graph *G, long *C) {
    int64_t n = 0;
    int rc = glv_vertexFollowing(G->numVertices, G->edgeListPtrs, (const glv_edge *)G->edgeList,
                                 (int64_t *)C, &n);
    if (rc) die("vertexFollowing", rc);
    return n;
}

double buildNewGraphVF(graph *Gin, graph *Gout, long *C, long numUniqueClusters) {
    int64_t *ptrs = nullptr, ne = 0;
    glv_edge *edges = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    int rc = glv_buildNewGraphVF(Gin->numVertices, Gin->edgeListPtrs, (const glv_edge *)Gin->edgeList,
                                 (const int64_t *)C, numUniqueClusters, &ptrs, &edges, &ne);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) die("buildNewGraphVF", rc);
    Gout->numVertices = numUniqueClusters;
    Gout->sVertices = numUniqueClusters;
    Gout->numEdges = ne;
    Gout->edgeListPtrs = (long *)ptrs;
    Gout->edgeList = (edge *)edges;
    return std::chrono::duration<double>(t1 - t0).count();
}

double buildNextLevelGraphOpt(graph *Gin, graph *Gout, long *C, long numUniqueClusters, int nThreads) {
    (void)nThreads;
    int64_t *ptrs = nullptr, ne = 0;
    glv_edge *edges = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    int rc = glv_buildNextLevelGraphOpt(Gin->numVertices, Gin->edgeListPtrs,
                                        (const glv_edge *)Gin->edgeList, (const int64_t *)C,
                                        numUniqueClusters, &ptrs, &edges, &ne);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) die("buildNextLevelGraphOpt", rc);
    Gout->numVertices = numUniqueClusters;
    Gout->sVertices = numUniqueClusters;
    Gout->numEdges = ne;
    Gout->edgeListPtrs = (long *)ptrs;
    Gout->edgeList = (edge *)edges;
    return std::chrono::duration<double>(t1 - t0).count();
}
