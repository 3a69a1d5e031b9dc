// dropin_caller.cpp -- a caller written against the drop-in C++ headers
// (include/grappolo/*.h) the way Grappolo's own clustering driver calls the
// library (driverForGraphClustering.cpp:135-203): it loads a -f 9 binary
// graph into a malloc()ed `graph` (the layout parse_EdgeListBinaryNew fills,
// InputsOutput/loadBinary.cpp:6-48), sets C_orig to -1, hands G over to
// runMultiPhaseColoring / runMultiPhaseBasic (which free it) and writes
// C_orig one id per line (the -o file, driverForGraphClustering.cpp:192-203).
// The tests link it to libgrappolo_hip.so and run it as a fresh process.
//
// usage: dropin_caller <graph.bin> <out.txt> <mode> [minGraphSize]
//   mode color  : runMultiPhaseColoring(G, C_orig, 1, 16, 0, mgs, 1e-9, 1e-6, 1, 1)
//   mode basic  : runMultiPhaseBasic(G, C_orig, 0, mgs, 1e-9, 1e-6, 1, 1)
//   mode stages : coloring, one colored phase, renumber and coarsening through
//                 the stage entry points; the caller frees the coarse graph's
//                 arrays with free(), as runMultiPhaseColoring.cpp:158-160 does
//   mode badcolor : runMultiPhaseColoring with coloring = 4 (the reference
//                 asserts 0 < coloring < 4, runMultiPhaseColoring.cpp:55)
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "grappolo/basic_comm.h"
#include "grappolo/basic_util.h"
#include "grappolo/color_comm.h"
#include "grappolo/coloring.h"

static graph *load(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(2);
    }
    long NV, NE, weighted;
    if (fread(&NV, 8, 1, f) != 1 || fread(&NE, 8, 1, f) != 1 || fread(&weighted, 8, 1, f) != 1) exit(2);
    graph *G = (graph *)malloc(sizeof(graph));
    G->numVertices = G->sVertices = NV;
    G->numEdges = NE;
    G->edgeListPtrs = (long *)malloc((NV + 1) * sizeof(long));
    G->edgeList = (edge *)malloc((2 * NE > 0 ? 2 * NE : 1) * sizeof(edge));
    if (fread(G->edgeListPtrs, 8, NV + 1, f) != (size_t)(NV + 1)) exit(2);
    if (fread(G->edgeList, sizeof(edge), 2 * NE, f) != (size_t)(2 * NE)) exit(2);
    fclose(f);
    return G;
}

static void write_ids(const char *path, const long *ids, long n) {
    FILE *out = fopen(path, "w");
    if (!out) exit(2);
    for (long i = 0; i < n; i++) fprintf(out, "%ld\n", ids[i]);
    fclose(out);
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s graph.bin out.txt color|basic|stages|badcolor [minGraphSize]\n", argv[0]);
        return 2;
    }
    const char *mode = argv[3];
    const long mgs = argc > 4 ? atol(argv[4]) : 1000;
    graph *G = load(argv[1]);
    const long NV = G->numVertices;
    long *C_orig = (long *)malloc(NV * sizeof(long));
    for (long i = 0; i < NV; i++) C_orig[i] = -1;  // driverForGraphClustering.cpp:180-183
    if (strcmp(mode, "color") == 0) {
        runMultiPhaseColoring(G, C_orig, 1, 16, 0, mgs, 1e-9, 1e-6, 1, 1);  // :185, G is consumed
        write_ids(argv[2], C_orig, NV);
    } else if (strcmp(mode, "basic") == 0) {
        runMultiPhaseBasic(G, C_orig, 0, mgs, 1e-9, 1e-6, 1, 1);  // :189, G is consumed
        write_ids(argv[2], C_orig, NV);
    } else if (strcmp(mode, "badcolor") == 0) {
        runMultiPhaseColoring(G, C_orig, 4, 16, 0, mgs, 1e-9, 1e-6, 1, 1);
        return 0;  // not reached: the library reports the error and exits with status 1
    } else if (strcmp(mode, "stages") == 0) {
        int *colors = (int *)malloc(NV * sizeof(int));
        double t = 0;
        int itr = 0;
        const int nColors = algoDistanceOneVertexColoringOpt(G, colors, 1, &t) + 1;
        long *C = (long *)malloc(NV * sizeof(long));
        const double q = algoLouvainWithDistOneColoring(G, C, 1, colors, nColors, -1, 1e-6, &t, &itr);
        const long k = renumberClustersContiguously(C, NV);
        graph *Gnew = (graph *)malloc(sizeof(graph));
        buildNextLevelGraphOpt(G, Gnew, C, k, 1);
        // one id per line: the phase's labels, then the coarse graph's offsets
        FILE *out = fopen(argv[2], "w");
        if (!out) return 2;
        fprintf(out, "%d %d %.17g %ld %ld %ld\n", nColors, itr, q, k, Gnew->numVertices, Gnew->numEdges);
        for (long i = 0; i < NV; i++) fprintf(out, "%ld\n", C[i]);
        for (long i = 0; i <= Gnew->numVertices; i++) fprintf(out, "%ld\n", Gnew->edgeListPtrs[i]);
        fclose(out);
        // the caller owns both graphs here (runMultiPhaseColoring.cpp:158-160)
        free(Gnew->edgeListPtrs);
        free(Gnew->edgeList);
        free(Gnew);
        free(G->edgeListPtrs);
        free(G->edgeList);
        free(G);
        free(C);
        free(colors);
    } else {
        fprintf(stderr, "unknown mode %s\n", mode);
        return 2;
    }
    free(C_orig);
    // the heap is still sound after the library freed G: allocate and free again
    void *p = malloc(1 << 20);
    memset(p, 1, 1 << 20);
    free(p);
    printf("dropin ok\n");
    return 0;
}
