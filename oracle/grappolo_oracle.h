/*
 * grappolo_oracle -- CPU restatement of Grappolo's colored multi-phase Louvain.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the HIP engine is
 * compared against (tests/, __graft_entry__.smoke(), and bench.py's
 * cpu_baseline leg).  Nothing in the product path links, loads or calls it.
 *
 * Graphs are symmetric CSR in structure-of-arrays form:
 *   off[nv+1] (int64), tails[off[nv]] (int64), w[off[nv]] (double).
 * Every function follows one reference routine; the file:line it restates is
 * given at the definition in grappolo_oracle.c.
 */
#ifndef GRAPPOLO_ORACLE_H
#define GRAPPOLO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t phases;
    int32_t total_iterations;
    int64_t final_clusters;
    double final_modularity; /* prevMod at exit: the value the reference prints */
    double last_modularity;  /* currMod at exit */
    int64_t edges_visited;   /* sum over phases and iterations of nnz(G_phase) */
    int32_t colors_first;
    int32_t n_phase;         /* entries written to the per-phase arrays below */
    /* optional per-phase records (caller-owned, capacity phase_cap) */
    int32_t phase_cap;
    int64_t *phase_nv;
    int64_t *phase_nnz;
    int32_t *phase_iters;
    int32_t *phase_colors;    /* nColors used by that phase (0 for a plain phase) */
    int32_t *phase_colored;   /* 1 colored sweep, 0 plain sweep */
    double *phase_mod;        /* value returned by the phase's sweep routine */
    int64_t *phase_clusters;  /* k after renumbering */
} orc_stats;

/* utilityClusteringFunctions.cpp:70-96 */
void orc_degrees(int64_t nv, const int64_t *off, const double *w, double *k);
double orc_constant(int64_t nv, const double *k);

/* coloring: returns largest color, or -2 when a color >= 4096 would be needed */
int orc_color_greedy(int64_t nv, const int64_t *off, const int64_t *tails, int32_t *colors);
int orc_color_jp(int64_t nv, const int64_t *off, const int64_t *tails, int32_t *colors,
                 int32_t *rounds_out);
uint64_t orc_jp_priority(int64_t v);
int64_t orc_color_conflicts(int64_t nv, const int64_t *off, const int64_t *tails,
                            const int32_t *colors);

/* one phase of the colored sweep; frozen=1 reads neighbour labels from the
 * state at the start of each color step (the engine's semantics), frozen=0
 * updates labels immediately in vertex order (the reference at one thread). */
double orc_louvain_colored(int64_t nv, const int64_t *off, const int64_t *tails, const double *w,
                           int64_t *C, const int32_t *color, int numColor, double Lower,
                           double thresh, int frozen, int32_t *numItr, double *trace,
                           int32_t trace_cap);

/* one phase of the plain (Jacobi) sweep; nomap=1 starts from initCommAssOpt */
double orc_louvain_plain(int64_t nv, const int64_t *off, const int64_t *tails, const double *w,
                         int64_t *C, double Lower, double thresh, int nomap, int32_t *numItr,
                         double *trace, int32_t trace_cap);

/* one color step from an arbitrary labelling C (community info derived from
 * C); frozen semantics.  Returns the number of moved vertices. */
int64_t orc_color_step(int64_t nv, const int64_t *off, const int64_t *tails, const double *w,
                       int64_t *C, const int32_t *color, int32_t which_color);

/* modularity of a labelling (community degrees derived from C), long double
 * accumulation.  Labels must be -1 or in [0, nv). */
double orc_modularity(int64_t nv, const int64_t *off, const int64_t *tails, const double *w,
                      const int64_t *C);

/* buildNextPhase.cpp:48-79 */
int64_t orc_renumber(int64_t *C, int64_t size);

/* buildNextPhase.cpp:83-255, canonical (row-sorted) output.
 * First call with out_off == NULL to get the entry count (returned);
 * out_off must hold k+1 entries.  *out_num_edges receives NE_out + k. */
int64_t orc_coarsen(int64_t nv, const int64_t *off, const int64_t *tails, const double *w,
                    const int64_t *C, int64_t k, int64_t *out_off, int64_t *out_tails,
                    double *out_w, int64_t *out_num_edges);

/* vertexFollowing.cpp:93-258 (buildNewGraphVF), same calling form as
 * orc_coarsen: no self-loop seeds, weights truncated to integers,
 * *out_num_edges = NE_out + NE_self. */
int64_t orc_build_vf(int64_t nv, const int64_t *off, const int64_t *tails, const double *w,
                     const int64_t *C, int64_t k, int64_t *out_off, int64_t *out_tails,
                     double *out_w, int64_t *out_num_edges);

/* vertexFollowing.cpp:46-89; returns numNode */
int64_t orc_vertex_following(int64_t nv, const int64_t *off, const int64_t *tails, int64_t *C);

/* Results of computeCommunityComparisons (utilityClusterComparisonMetrics.cpp:50-304). */
typedef struct {
    int64_t nC1, nC2;
    int64_t same_same, same_diff, diff_same, agree;
    double precision, recall, f_score, gini1, gini2;
} orc_comparison;

/* utilityClusterComparisonMetrics.cpp:50-304 at one thread.  C1 (truth) and
 * C2 are relabelled in place as the reference does.  Returns 0, or -1 where
 * the reference asserts (N <= 0, no positive community count) or indexes out
 * of range (a label below -1). */
int orc_compare(int64_t *C1, int64_t *C2, int64_t N, orc_comparison *out);

/* utilityClusterComparisonMetrics.cpp:309-330; sorts colorSize in place */
double orc_gini(int64_t *colorSize, int numColors);

/* coloringMultiHashMaxMin.cpp:60-226 with hashed random values; returns maxColor */
double orc_mh_value(int h, int64_t v);
int orc_color_multihash(int64_t nv, const int64_t *off, const int64_t *tails, int nHash, int nItrs,
                        int32_t *colors);
int orc_color_multihash_vals(int64_t nv, const int64_t *off, const int64_t *tails, int nHash, int nItrs,
                             const double *vals, int32_t *colors);
/* vBase.cpp:6-168 (type 0) as deterministic rounds; recolors in place,
 * returns the number of rounds or < 0 */
int orc_color_balance(int64_t nv, const int64_t *off, const int64_t *tails, int32_t *colors, int ncolors,
                      int type);

/* runMultiPhaseColoring.cpp:51-238.  coloring_mode 0: greedy (reference at one
 * thread), 1: Jones-Plassmann (engine), 2: Jones-Plassmann then balanced,
 * 3: multi-hash incomplete coloring with numColors.  Returns 0 or < 0. */
int orc_run_colored(int64_t nv, const int64_t *off, const int64_t *tails, const double *w,
                    int64_t *C_orig, int coloring_mode, int numColors, int frozen, int replaceMap,
                    int64_t minGraphSize, double threshold, double C_threshold, orc_stats *st);

/* runMultiPhaseBasic.cpp:52-170 */
int orc_run_basic(int64_t nv, const int64_t *off, const int64_t *tails, const double *w,
                  int64_t *C_orig, int basicOpt, double threshold, orc_stats *st);

#ifdef __cplusplus
}
#endif
#endif
