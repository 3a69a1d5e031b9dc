// ref_harness.cpp -- a thin C-ABI around the compiled reference (Grappolo's
// own sources, built by oracle/Makefile from the reference tree into
// oracle/_ref/).  TEST INFRASTRUCTURE ONLY: used to pin the oracle against
// the real reference in this container and as the CPU baseline in bench.py.
//
// The reference functions are called unchanged; this file only converts the
// structure-of-arrays CSR used by the tests into the reference's `graph`
// (DefineStructure/defs.h:85-92) and collects results.
#include <fcntl.h>
#include <stdint.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include <omp.h>

#include "defs.h"  // must precede the other reference headers
#include "basic_comm.h"
#include "basic_util.h"
#include "color_comm.h"
#include "utilityClusteringFunctions.h"
#include "coloring.h"
#include "input_output.h"

namespace {

graph *make_graph(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
                  const double *w) {
    graph *G = (graph *)malloc(sizeof(graph));
    int64_t nnz = off[nv];
    G->numVertices = nv;
    G->sVertices = nv;
    G->numEdges = num_edges;
    G->edgeListPtrs = (long *)malloc((nv + 1) * sizeof(long));
    G->edgeList = (edge *)malloc((nnz > 0 ? nnz : 1) * sizeof(edge));
    for (int64_t i = 0; i <= nv; i++) G->edgeListPtrs[i] = off[i];
    for (int64_t i = 0; i < nv; i++)
        for (int64_t j = off[i]; j < off[i + 1]; j++) {
            G->edgeList[j].head = i;
            G->edgeList[j].tail = tails[j];
            G->edgeList[j].weight = w ? w[j] : 1.0;
        }
    return G;
}

void free_graph(graph *G) {
    if (!G) return;
    free(G->edgeListPtrs);
    free(G->edgeList);
    free(G);
}

// Redirect the reference's stdout chatter into a file while `fn` runs.
struct Quiet {
    int saved = -1;
    explicit Quiet(const char *path) {
        fflush(stdout);
        saved = dup(1);
        int fd = open(path ? path : "/dev/null", O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd >= 0) {
            dup2(fd, 1);
            close(fd);
        }
    }
    ~Quiet() {
        std::cout.flush();
        fflush(stdout);
        if (saved >= 0) {
            dup2(saved, 1);
            close(saved);
        }
    }
};

}  // namespace

extern "C" {

int ref_color(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
              const double *w, int32_t *colors, int nthreads) {
    graph *G = make_graph(nv, num_edges, off, tails, w);
    for (int64_t i = 0; i < nv; i++) colors[i] = -1;
    double t = 0;
    int mc;
    {
        Quiet q(nullptr);
        mc = algoDistanceOneVertexColoringOpt(G, colors, nthreads, &t);
    }
    free_graph(G);
    return mc;
}

double ref_colored_phase(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
                         const double *w, int64_t *C, int32_t *colors, int ncolors, double Lower,
                         double thresh, int nomap, int nthreads, int32_t *numItr) {
    graph *G = make_graph(nv, num_edges, off, tails, w);
    double t = 0, mod;
    int it = 0;
    {
        Quiet q(nullptr);
        if (nomap)
            mod = algoLouvainWithDistOneColoringNoMap(G, (long *)C, nthreads, colors, ncolors, Lower,
                                                      thresh, &t, &it);
        else
            mod = algoLouvainWithDistOneColoring(G, (long *)C, nthreads, colors, ncolors, Lower,
                                                 thresh, &t, &it);
    }
    *numItr = it;
    free_graph(G);
    return mod;
}

double ref_plain_phase(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
                       const double *w, int64_t *C, double Lower, double thresh, int nomap,
                       int nthreads, int32_t *numItr) {
    graph *G = make_graph(nv, num_edges, off, tails, w);
    double t = 0, mod;
    int it = 0;
    {
        Quiet q(nullptr);
        if (nomap)
            mod = parallelLouvianMethodNoMap(G, (long *)C, nthreads, Lower, thresh, &t, &it);
        else
            mod = parallelLouvianMethod(G, (long *)C, nthreads, Lower, thresh, &t, &it);
    }
    *numItr = it;
    free_graph(G);
    return mod;
}

int64_t ref_renumber(int64_t *C, int64_t n) {
    Quiet q(nullptr);
    return renumberClustersContiguously((long *)C, n);
}

// Coarsen; the output arrays are malloc()ed and returned through the out
// pointers (free with ref_free).  Rows come back in the reference's order.
int64_t ref_coarsen(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
                    const double *w, int64_t *C, int64_t k, int nthreads, int64_t **out_off,
                    int64_t **out_tails, double **out_w, int64_t *out_num_edges) {
    graph *G = make_graph(nv, num_edges, off, tails, w);
    graph Gout;
    {
        Quiet q(nullptr);
        buildNextLevelGraphOpt(G, &Gout, (long *)C, k, nthreads);
    }
    int64_t tot = Gout.edgeListPtrs[Gout.numVertices];
    *out_off = (int64_t *)malloc((k + 1) * sizeof(int64_t));
    *out_tails = (int64_t *)malloc((tot > 0 ? tot : 1) * sizeof(int64_t));
    *out_w = (double *)malloc((tot > 0 ? tot : 1) * sizeof(double));
    for (int64_t i = 0; i <= k; i++) (*out_off)[i] = Gout.edgeListPtrs[i];
    for (int64_t j = 0; j < tot; j++) {
        (*out_tails)[j] = Gout.edgeList[j].tail;
        (*out_w)[j] = Gout.edgeList[j].weight;
    }
    *out_num_edges = Gout.numEdges;
    free(Gout.edgeListPtrs);
    free(Gout.edgeList);
    free_graph(G);
    return tot;
}

// vertexFollowing (vertexFollowing.cpp:46-89) on the reference's own code.
int64_t ref_vertex_following(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
                             int64_t *C) {
    graph *G = make_graph(nv, num_edges, off, tails, nullptr);
    long n;
    {
        Quiet q(nullptr);
        n = vertexFollowing(G, (long *)C);
    }
    free_graph(G);
    return n;
}

// buildNewGraphVF (vertexFollowing.cpp:93-258); outputs as ref_coarsen.
int64_t ref_build_vf(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
                     const double *w, int64_t *C, int64_t k, int64_t **out_off, int64_t **out_tails,
                     double **out_w, int64_t *out_num_edges) {
    graph *G = make_graph(nv, num_edges, off, tails, w);
    graph Gout;
    {
        Quiet q(nullptr);
        buildNewGraphVF(G, &Gout, (long *)C, k);
    }
    int64_t tot = Gout.edgeListPtrs[Gout.numVertices];
    *out_off = (int64_t *)malloc((k + 1) * sizeof(int64_t));
    *out_tails = (int64_t *)malloc((tot > 0 ? tot : 1) * sizeof(int64_t));
    *out_w = (double *)malloc((tot > 0 ? tot : 1) * sizeof(double));
    for (int64_t i = 0; i <= k; i++) (*out_off)[i] = Gout.edgeListPtrs[i];
    for (int64_t j = 0; j < tot; j++) {
        (*out_tails)[j] = Gout.edgeList[j].tail;
        (*out_w)[j] = Gout.edgeList[j].weight;
    }
    *out_num_edges = Gout.numEdges;
    free(Gout.edgeListPtrs);
    free(Gout.edgeList);
    free_graph(G);
    return tot;
}

// computeCommunityComparisons (utilityClusterComparisonMetrics.cpp:50-304)
// with `nthreads` OpenMP threads; its report goes to `log`.  The vectors carry
// one element past N, as driverForClusterComparison.cpp's reader leaves them.
// C1 and C2 receive the relabelled vectors.
void ref_compare(int64_t *C1, int64_t *C2, int64_t N, int nthreads, const char *log) {
    std::vector<long> a(C1, C1 + N), b(C2, C2 + N);
    a.push_back(0);
    b.push_back(0);
    omp_set_num_threads(nthreads);
    {
        Quiet q(log);
        computeCommunityComparisons(a, N, b, N);
    }
    for (int64_t i = 0; i < N; i++) {
        C1[i] = a[i];
        C2[i] = b[i];
    }
}

// computeGiniCoefficient (utilityClusterComparisonMetrics.cpp:309-330); sorts in place.
double ref_gini(int64_t *sizes, int n) {
    Quiet q(nullptr);
    return computeGiniCoefficient((long *)sizes, n);
}

// algoColoringMultiHashMaxMin (coloringMultiHashMaxMin.cpp:60-226) at one
// thread after srand(seed); vals receives the rand() values it drew,
// vals[h * nv + v] (generateRandomNumbers2, :52-58, hash by hash).
int ref_multihash(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails, int nHash,
                  int nItrs, unsigned seed, int32_t *colors, double *vals) {
    graph *G = make_graph(nv, num_edges, off, tails, nullptr);
    double t = 0;
    int mc;
    srand(seed);
    {
        Quiet q(nullptr);
        mc = algoColoringMultiHashMaxMin(G, colors, 1, &t, nHash, nItrs);
    }
    srand(seed);
    for (int h = 0; h < nHash; h++)
        for (int64_t v = 0; v < nv; v++) vals[h * nv + v] = (double)rand();
    free_graph(G);
    return mc;
}

void ref_free(void *p) { free(p); }

// The reference's own file readers: parse_EdgeListBinaryNew (-f 9) and
// loadMetisFileFormat (-f 5).  Returns the graph as malloc()ed SoA arrays.
static int64_t take_graph(graph *G, int64_t **off, int64_t **tails, double **w, int64_t *ne) {
    int64_t nv = G->numVertices, nnz = G->edgeListPtrs[nv];
    *off = (int64_t *)malloc((nv + 1) * sizeof(int64_t));
    *tails = (int64_t *)malloc((nnz > 0 ? nnz : 1) * sizeof(int64_t));
    *w = (double *)malloc((nnz > 0 ? nnz : 1) * sizeof(double));
    for (int64_t i = 0; i <= nv; i++) (*off)[i] = G->edgeListPtrs[i];
    for (int64_t j = 0; j < nnz; j++) {
        (*tails)[j] = G->edgeList[j].tail;
        (*w)[j] = G->edgeList[j].weight;
    }
    *ne = G->numEdges;
    free_graph(G);
    return nv;
}

int64_t ref_load_binary(const char *path, int64_t **off, int64_t **tails, double **w, int64_t *ne) {
    graph *G = (graph *)calloc(1, sizeof(graph));
    {
        Quiet q(nullptr);
        parse_EdgeListBinaryNew(G, (char *)path);
    }
    return take_graph(G, off, tails, w, ne);
}

int64_t ref_load_metis(const char *path, int64_t **off, int64_t **tails, double **w, int64_t *ne) {
    graph *G = (graph *)calloc(1, sizeof(graph));
    {
        Quiet q(nullptr);
        loadMetisFileFormat(G, path);
    }
    return take_graph(G, off, tails, w, ne);
}

// writeGraphBinaryFormatNew (-f 9 writer)
void ref_write_binary(const char *path, int64_t nv, int64_t num_edges, const int64_t *off,
                      const int64_t *tails, const double *w) {
    graph *G = make_graph(nv, num_edges, off, tails, w);
    {
        Quiet q(nullptr);
        writeGraphBinaryFormatNew(G, (char *)path, 1);
    }
    free_graph(G);
}

// The reference's own drivers, unchanged.  They consume the graph and print
// their summary; the summary goes to `log_path` so the caller can read the
// printed "Final modularity" line.  C_orig receives the partition.
void ref_run_colored(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
                     const double *w, int64_t *C_orig, int replaceMap, int64_t minGraphSize,
                     double threshold, double C_threshold, int nthreads, const char *log_path) {
    graph *G = make_graph(nv, num_edges, off, tails, w);
    for (int64_t i = 0; i < nv; i++) C_orig[i] = -1;
    Quiet q(log_path);
    runMultiPhaseColoring(G, (long *)C_orig, 1, 0, replaceMap, minGraphSize, threshold, C_threshold,
                          nthreads, 1);
}

void ref_run_basic(int64_t nv, int64_t num_edges, const int64_t *off, const int64_t *tails,
                   const double *w, int64_t *C_orig, int basicOpt, double threshold, int nthreads,
                   const char *log_path) {
    graph *G = make_graph(nv, num_edges, off, tails, w);
    for (int64_t i = 0; i < nv; i++) C_orig[i] = -1;
    Quiet q(log_path);
    runMultiPhaseBasic(G, (long *)C_orig, basicOpt, 100000, threshold, 0.01, nthreads, 1);
}

}  // extern "C"
