// ref_bench.cpp -- times the compiled reference's colored multi-phase Louvain
// on host cores.  TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg).
//
// usage: ref_bench <graph.bin> <threads> [minGraphSize threshold C_threshold [driver <log>]]
// The graph file is the reference's -f 9 binary CSR (InputsOutput/
// writeBinary.cpp:2-29 layout).  The phase loop below is the control flow of
// Coloring/runMultiPhaseColoring.cpp:104-207 (coloring == 1, map sweeps) and
// calls the reference's own routines for every step, so that the edges visited
// per phase can be counted; the reference driver itself exposes no counts.
// With `driver <log>` the reference's runMultiPhaseColoring itself runs and
// the values of its printed summary are reported.
#include <fcntl.h>
#include <omp.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "defs.h"  // must precede the other reference headers
#include "basic_comm.h"
#include "basic_util.h"
#include "color_comm.h"
#include "coloring.h"

static graph *load(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(2);
    }
    long NV, NE, weighted;
    if (fread(&NV, 8, 1, f) != 1 || fread(&NE, 8, 1, f) != 1 || fread(&weighted, 8, 1, f) != 1)
        exit(2);
    graph *G = (graph *)malloc(sizeof(graph));
    G->numVertices = G->sVertices = NV;
    G->numEdges = NE;
    G->edgeListPtrs = (long *)malloc((NV + 1) * sizeof(long));
    G->edgeList = (edge *)malloc(2 * NE * sizeof(edge));
    if (fread(G->edgeListPtrs, 8, NV + 1, f) != (size_t)(NV + 1)) exit(2);
    if (fread(G->edgeList, sizeof(edge), 2 * NE, f) != (size_t)(2 * NE)) exit(2);
    fclose(f);
    return G;
}

// Value printed after `label` in the reference's compact summary
// (Coloring/runMultiPhaseColoring.cpp:209-224), or NaN.
static double summary_value(const char *path, const char *label) {
    FILE *f = fopen(path, "r");
    if (!f) return 0.0 / 0.0;
    char line[512];
    double v = 0.0 / 0.0;
    const size_t n = strlen(label);
    while (fgets(line, sizeof line, f)) {
        if (strncmp(line, label, n) != 0) continue;
        const char *colon = strchr(line, ':');
        if (colon) v = atof(colon + 1);
    }
    fclose(f);
    return v;
}

// The reference's own driver entry point, unchanged: runMultiPhaseColoring
// (Coloring/runMultiPhaseColoring.cpp:51-238) with coloring 1 and map sweeps,
// as driverForGraphClustering.cpp:163 calls it; its printed summary is parsed.
static int run_driver(graph *G, int nT, long minGraphSize, double threshold, double C_threshold,
                      const char *log_path) {
    long NV = G->numVertices;
    long *C_orig = (long *)malloc(NV * sizeof(long));
    for (long i = 0; i < NV; i++) C_orig[i] = -1;  // driverForGraphClustering.cpp:180-183
    int saved = dup(1);
    fflush(stdout);
    int fd = open(log_path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    dup2(fd, 1);
    close(fd);
    double t0 = omp_get_wtime();
    runMultiPhaseColoring(G, C_orig, 1, 16, 0, minGraphSize, threshold, C_threshold, nT, 1);
    double t1 = omp_get_wtime();
    fflush(stdout);
    dup2(saved, 1);
    close(saved);
    long isolated = 0;
    for (long i = 0; i < NV; i++) isolated += C_orig[i] < 0;
    free(C_orig);
    printf("{\"threads\": %d, \"seconds\": %.6f, \"total_time\": %.6f, \"final_modularity\": %.6f, "
           "\"phases\": %.0f, \"iterations\": %.0f, \"clusters\": %.0f, \"unlabelled\": %ld}\n",
           nT, t1 - t0, summary_value(log_path, "TOTAL TIME"),
           summary_value(log_path, "Final modularity"), summary_value(log_path, "Total number of phases"),
           summary_value(log_path, "Total number of iterations"),
           summary_value(log_path, "Final number of clusters"), isolated);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr,
                "usage: %s graph.bin threads [minGraphSize threshold C_threshold [driver log]]\n"
                "  driver: call the reference's runMultiPhaseColoring and parse its summary\n",
                argv[0]);
        return 2;
    }
    graph *G = load(argv[1]);
    if (argc > 7 && strcmp(argv[6], "driver") == 0)
        return run_driver(G, atoi(argv[2]), atol(argv[3]), atof(argv[4]), atof(argv[5]), argv[7]);
    int nT = atoi(argv[2]);
    long minGraphSize = argc > 3 ? atol(argv[3]) : 1000;
    double threshold = argc > 4 ? atof(argv[4]) : 1e-9;
    double C_threshold = argc > 5 ? atof(argv[5]) : 1e-6;
    omp_set_num_threads(nT);

    long NV = G->numVertices;
    long *C_orig = (long *)malloc(NV * sizeof(long));
    int *colors = (int *)malloc(NV * sizeof(int));
    long *C = (long *)malloc(NV * sizeof(long));
    int saved = dup(1);
    fflush(stdout);
    int fd = open("/dev/null", O_WRONLY);
    dup2(fd, 1);

    double t0 = omp_get_wtime();
    double tmpTime = 0, prevMod = -1, currMod = -1;
    int tmpItr = 0, totItr = 0;
    long phase = 1, numClusters = 0;
    double ev = 0;
    for (long i = 0; i < NV; i++) colors[i] = -1;
    int nColors = algoDistanceOneVertexColoringOpt(G, colors, nT, &tmpTime) + 1;
    int firstColors = nColors;
    for (long i = 0; i < NV; i++) C[i] = -1;
    bool nonColor = false;
    while (1) {
        prevMod = currMod;
        long nnz = G->edgeListPtrs[G->numVertices];
        if (!nonColor)
            currMod = algoLouvainWithDistOneColoring(G, C, nT, colors, nColors, currMod, C_threshold,
                                                     &tmpTime, &tmpItr);
        else
            currMod = parallelLouvianMethod(G, C, nT, currMod, threshold, &tmpTime, &tmpItr);
        totItr += tmpItr;
        ev += (double)tmpItr * (double)nnz;
        numClusters = renumberClustersContiguously(C, G->numVertices);
        if (phase == 1) {
            for (long i = 0; i < NV; i++) C_orig[i] = C[i];
        } else {
            for (long i = 0; i < NV; i++)
                if (C_orig[i] >= 0) C_orig[i] = C[C_orig[i]];
        }
        if ((phase > 200) || (totItr > 10000)) break;
        if ((currMod - prevMod) > threshold) {
            graph *Gnew = (graph *)malloc(sizeof(graph));
            buildNextLevelGraphOpt(G, Gnew, C, numClusters, nT);
            free(G->edgeListPtrs);
            free(G->edgeList);
            free(G);
            G = Gnew;
            free(C);
            C = (long *)malloc(numClusters * sizeof(long));
            for (long i = 0; i < numClusters; i++) C[i] = -1;
            phase++;
            if ((G->numVertices > minGraphSize) && !nonColor) {
                for (long i = 0; i < G->numVertices; i++) colors[i] = -1;
                nColors = algoDistanceOneVertexColoringOpt(G, colors, nT, &tmpTime) + 1;
            }
        } else {
            if (!nonColor)
                nonColor = true;
            else
                break;
        }
    }
    double t1 = omp_get_wtime();
    fflush(stdout);
    dup2(saved, 1);
    printf("{\"threads\": %d, \"seconds\": %.6f, \"edges_visited\": %.0f, \"ev_per_s\": %.6e, "
           "\"final_modularity\": %.12f, \"phases\": %ld, \"iterations\": %d, \"colors\": %d, "
           "\"clusters\": %ld}\n",
           nT, t1 - t0, ev, ev / (t1 - t0), prevMod, phase, totItr, firstColors, numClusters);
    return 0;
}
