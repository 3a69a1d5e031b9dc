/*
 * grappolo_hip.h -- C ABI of libgrappolo_hip.so, the MI355X (gfx950) engine
 * for Grappolo's colored multi-phase Louvain.
 *
 * Two families of entry points:
 *   glv_<ReferenceName>        graph given as the reference's host arrays
 *                              (edgeListPtrs[NV+1] + edge{head,tail,weight}[nnz]);
 *                              the engine uploads, runs on the GPU, downloads.
 *   glv_<stage>_dev            graph already resident in device memory as
 *                              structure-of-arrays CSR (offsets, tails, weights;
 *                              int64/int64/double), results left in device memory.
 * Each entry point names the reference function it replaces (file:line of
 * the reference tree).  Plain pointers and sizes only.  Every function returns
 * GLV_OK or a GLV_E_* status (glv_status.h); glv_last_error() gives the
 * message of the last failure on the calling thread.  Calls are not re-entrant
 * on one device: concurrent calls from several host threads must use distinct
 * streams (the _dev forms) and are serialised by the caller.
 *
 * Graph contract (the reference's, DefineStructure/defs.h:85-92): symmetric
 * CSR, every undirected edge in both rows, self-loops once, non-negative
 * weights, tails in [0, NV).  Violations return GLV_E_INVALID_GRAPH.
 */
#ifndef GRAPPOLO_HIP_H
#define GRAPPOLO_HIP_H

#include <stdint.h>

#include "glv_status.h"

#ifdef __cplusplus
extern "C" {
#endif

/* edge record, layout of DefineStructure/defs.h:78-83 (24 bytes) */
typedef struct glv_edge {
    int64_t head;
    int64_t tail;
    double weight;
} glv_edge;

/* Outcome of a multi-phase run.  The optional arrays are caller-owned; set
 * their pointer and capacity, the engine fills up to the capacity. */
typedef struct glv_stats {
    int32_t phases;           /* "Total number of phases" (runMultiPhaseColoring.cpp:213) */
    int32_t total_iterations; /* "Total number of iterations" */
    int64_t final_clusters;   /* "Final number of clusters" */
    double final_modularity;  /* "Final modularity" (prevMod at exit, :216) */
    double last_modularity;   /* currMod at exit */
    double time_clustering;   /* seconds: sweeps incl. their modularity passes */
    double time_building;     /* coarsening */
    double time_coloring;
    double time_renumber;     /* renumbering and C_orig updates */
    double time_total;        /* wall time of the whole call on the device side */
    int64_t edges_visited;    /* sum over phases of iterations * nnz(phase graph) */
    int64_t edges_visited_mod;/* modularity passes of colored phases */
    int32_t colors_first;     /* colors of the first coloring */
    int32_t n_phase;          /* phases recorded (may exceed phase_cap) */
    int32_t phase_cap;
    int32_t trace_cap;
    int32_t n_trace;          /* iterations recorded in trace_q (may exceed trace_cap) */
    int32_t reserved;
    int64_t *phase_nv;
    int64_t *phase_nnz;
    int32_t *phase_iters;
    int32_t *phase_colors;
    int32_t *phase_colored;
    double *phase_mod;
    int64_t *phase_clusters;
    double *trace_q;          /* modularity after every iteration of every phase */
    double *phase_seconds;    /* wall time of each phase: sweeps, renumbering, coarsening, coloring */
    int64_t moves;            /* vertices moved by the whole-GPU colored sweeps (graphs above
                                 262,144 vertices; each move is 2 FP64 and 2 int64 atomics) */
} glv_stats;

/* ---- device and errors ------------------------------------------------ */
int glv_set_device(int device);
int glv_get_device_count(int *count);
const char *glv_last_error(void);
/* The engine's build: "gfx950", plus version text. */
const char *glv_version(void);
/* No counterpart in the reference (its routines free what they allocate
 * before returning).  Releases what the engine keeps between calls: cached
 * device blocks, pinned staging buffers, timing events and its own streams.
 * Runs by itself at process exit (registered with atexit on the first call,
 * so before the HIP runtime's teardown); a later call starts afresh.  Must not
 * overlap another call. */
int glv_shutdown(void);

/* ---- whole-pipeline entry points (host arrays) --------------------------- */

/* runMultiPhaseColoring, DefineStructure/color_comm.h:7 and
 * Coloring/runMultiPhaseColoring.cpp:51-238.  coloring 1: distance-1
 * (Jones-Plassmann); 2: distance-1 then balanced (glv_vBaseRedistribution,
 * type 0); 3: incomplete multi-hash with nHash 2 and nItrs = max(numColors/4,
 * 1) (glv_algoColoringMultiHashMaxMin); numColors is used by coloring 3 only.
 * C_orig[NV] receives the partition (-1 for vertices without edges).
 * The graph is not modified (the C++ drop-in frees it, as the reference does).
 * numThreads/threadsOpt are accepted and ignored (no host threads are used). */
int glv_runMultiPhaseColoring(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                              int64_t *C_orig, int coloring, int numColors, int replaceMap,
                              int64_t minGraphSize, double threshold, double C_threshold,
                              int numThreads, int threadsOpt, glv_stats *stats);

/* runMultiPhaseBasic, DefineStructure/basic_comm.h:5 and
 * BasicCommunitiesDetection/runMultiPhaseBasic.cpp:52-170.  basicOpt 1 selects
 * the NoMap sweep (initCommAssOpt start). */
int glv_runMultiPhaseBasic(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                           int64_t *C_orig, int basicOpt, int64_t minGraphSize, double threshold,
                           double C_threshold, int numThreads, int threadsOpt, glv_stats *stats);

/* ---- single stages (host arrays) ------------------------------------------ */

/* algoDistanceOneVertexColoringOpt, DefineStructure/coloring.h:8 and
 * Coloring/coloringDistanceOne.cpp:50-233.  *maxColor = largest color used.
 * Jones-Plassmann rounds: a vertex takes the smallest color free among its
 * coloured neighbours once every neighbour of higher priority is coloured;
 * priority = (min(degree, 2^23-1) << 40) | (splitmix64(v) >> 24), equal
 * keys to the larger id.  A valid coloring, not the reference's colors. */
int glv_algoDistanceOneVertexColoringOpt(int64_t nv, const int64_t *edgeListPtrs,
                                         const glv_edge *edgeList, int32_t *vtxColor,
                                         int32_t *maxColor, double *totTime);

/* algoLouvainWithDistOneColoring(NoMap), DefineStructure/color_comm.h:10,13 and
 * Coloring/parallelLouvainWithColoring.cpp:47-286.  C[NV] receives the labels;
 * *modularity the value the reference returns (prevMod). */
int glv_algoLouvainWithDistOneColoring(int64_t nv, const int64_t *edgeListPtrs,
                                       const glv_edge *edgeList, int64_t *C, const int32_t *color,
                                       int numColor, double Lower, double thresh,
                                       double *modularity, double *totTime, int32_t *numItr);

/* parallelLouvianMethod / parallelLouvianMethodNoMap, DefineStructure/
 * basic_comm.h:21,28 and BasicCommunitiesDetection/parallelLouvainMethod.cpp:65-326. */
int glv_parallelLouvianMethod(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                              int64_t *C, double Lower, double thresh, int noMap,
                              double *modularity, double *totTime, int32_t *numItr);

/* renumberClustersContiguously, DefineStructure/basic_util.h:5 and
 * Utility/buildNextPhase.cpp:48-79. */
int glv_renumberClustersContiguously(int64_t *C, int64_t size, int64_t *numClusters);

/* vertexFollowing, DefineStructure/basic_util.h:14 and
 * Utility/vertexFollowing.cpp:46-89: C[i] = -1 for vertices without entries,
 * the neighbour's id for a vertex with one entry whose neighbour has more than
 * one entry or a smaller id, else i; *numNode = vertices to remove. */
int glv_vertexFollowing(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList, int64_t *C,
                        int64_t *numNode);
/* buildNewGraphVF, DefineStructure/basic_util.h:15 and
 * Utility/vertexFollowing.cpp:93-258: like glv_buildNextLevelGraphOpt but
 * without the zero self-loop per cluster and with every weight truncated to
 * an integer before summing; *outNumEdges = NE_out + NE_self. */
int glv_buildNewGraphVF(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList, const int64_t *C,
                        int64_t numUniqueClusters, int64_t **outEdgeListPtrs, glv_edge **outEdgeList,
                        int64_t *outNumEdges);

/* buildNextLevelGraphOpt, DefineStructure/basic_util.h:6 and
 * Utility/buildNextPhase.cpp:83-255.  Outputs are malloc()ed (free with free()
 * or glv_free); rows are sorted by tail; *outNumEdges = NE_out + k. */
int glv_buildNextLevelGraphOpt(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                               const int64_t *C, int64_t numUniqueClusters,
                               int64_t **outEdgeListPtrs, glv_edge **outEdgeList,
                               int64_t *outNumEdges);

/* The in-sweep modularity of Coloring/parallelLouvainWithColoring.cpp:226-249
 * for a given labelling (community degrees summed from C; -1 labels ignored). */
int glv_modularity(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                   const int64_t *C, double *modularity);

void glv_free(void *p);

/* ---- device-resident forms (stream may be NULL: the engine's own stream) -- */

int glv_runMultiPhaseColoring_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets,
                                  const int64_t *d_tails, const double *d_weights, int64_t *d_C_orig,
                                  int coloring, int numColors, int replaceMap, int64_t minGraphSize,
                                  double threshold, double C_threshold, void *stream,
                                  glv_stats *stats);
int glv_runMultiPhaseBasic_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets,
                               const int64_t *d_tails, const double *d_weights, int64_t *d_C_orig,
                               int basicOpt, double threshold, void *stream, glv_stats *stats);
int glv_color_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                  int32_t *d_colors, int32_t *maxColor, int32_t *rounds, void *stream);
int glv_colored_phase_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                          const double *d_weights, int64_t *d_C, const int32_t *d_colors,
                          int numColor, double Lower, double thresh, double *modularity,
                          int32_t *numItr, void *stream);
int glv_plain_phase_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                        const double *d_weights, int64_t *d_C, double Lower, double thresh,
                        int noMap, double *modularity, int32_t *numItr, void *stream);
/* One color step of the colored sweep from an arbitrary labelling d_C (the
 * community degrees and sizes are summed from it); d_C is updated in place.
 * *moves = vertices whose label changed. */
int glv_color_step_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                       const double *d_weights, int64_t *d_C, const int32_t *d_colors,
                       int numColor, int which, int64_t *moves, void *stream);
int glv_modularity_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                       const double *d_weights, const int64_t *d_C, double *modularity,
                       void *stream);
/* No counterpart in the reference.  glv_modularity_dev for callers that
 * cannot allocate device memory (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc): same validation and result (to FP64
 * rounding of the sums; exact on integer weights), but no allocation of any
 * kind -- its scratch is a fixed device array of 64 MiB + 64 KiB per GPU, held
 * from the library's first use on that GPU until the process ends.  Graphs of
 * more than glv_noalloc_max_vertices() vertices (4,194,304) are rejected with
 * GLV_E_INVALID_ARG.  Calls that overlap on one GPU share that array and race:
 * the caller serialises them. */
int glv_modularity_noalloc_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                               const double *d_weights, const int64_t *d_C, double *modularity,
                               void *stream);
int glv_noalloc_max_vertices(int64_t *max_nv);
int glv_renumber_dev(int64_t n, int64_t *d_C, int64_t *numClusters, void *stream);
/* Coarsened graph in engine-allocated device memory; release with
 * glv_dev_free on each of the three pointers. */
int glv_coarsen_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                    const double *d_weights, const int64_t *d_C, int64_t k, int64_t **d_out_offsets,
                    int64_t **d_out_tails, double **d_out_weights, int64_t *out_nnz,
                    int64_t *out_num_edges, void *stream);
/* device forms of glv_vertexFollowing and glv_buildNewGraphVF (outputs as glv_coarsen_dev) */
int glv_vertexFollowing_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                            int64_t *d_C, int64_t *numNode, void *stream);
int glv_buildNewGraphVF_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                            const double *d_weights, const int64_t *d_C, int64_t k, int64_t **d_out_offsets,
                            int64_t **d_out_tails, double **d_out_weights, int64_t *out_nnz,
                            int64_t *out_num_edges, void *stream);
/* The vertex-following hook of driverForGraphClustering.cpp:112-132 on the
 * device: vertexFollowing, renumberClustersContiguously and buildNewGraphVF.
 * When *numNode is 0 the graph is unchanged: the outputs are NULL and
 * *out_nv = nv.  Otherwise the folded graph is returned like glv_coarsen_dev. */
int glv_vertexFollowingPrepass_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                                   const double *d_weights, int64_t **d_out_offsets, int64_t **d_out_tails,
                                   double **d_out_weights, int64_t *out_nv, int64_t *out_nnz,
                                   int64_t *out_num_edges, int64_t *numNode, void *stream);
int glv_dev_free(void *d_ptr);
int glv_count_conflicts_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets,
                            const int64_t *d_tails, const int32_t *d_colors, int64_t *conflicts,
                            void *stream);

/* ---- kernel timing (no counterpart in the reference) ----------------------
 * When enabled, every launch of the instrumented kernel classes is bracketed
 * by HIP events on the engine's stream; glv_profile_read returns the summed
 * device time (ms), the summed algorithmic bytes and the launch count of one
 * class: 0..7 sweep kernels of degree bins 1..8 (algorithmic bytes
 * 24*entries + 32*vertices per launch), 8 modularity pass, 9 commits,
 * 10 coloring rounds, 11 coarsening, 12 renumbering, 13 one whole colored
 * iteration of a small graph in a single cooperative launch, 14 the
 * sequential coloring tail, 15 sweep kernels of degree bin 9 (more than
 * 16384 entries, several workgroups per vertex). */
int glv_profile_enable(int on);
int glv_profile_reset(void);
int glv_profile_read(int kernel_class, double *ms, double *bytes, int64_t *launches);

/* ---- learned active set (no counterpart in the reference) ----------------
 * An 8 -> 64 -> 64 -> 1 MLP (ReLU, logistic output; BF16 weights and
 * activations on the matrix cores, FP32 accumulation) predicts which vertices
 * the next colored sweep will move.  Parameters: glv_mlp_param_count() FP32
 * values in device memory, laid out W1[8][64], b1[64], W2[64][64], b2[64],
 * w3[64], b3.  Samples: 8 BF16 features per row, labels 0/1 (uint8). */
int glv_mlp_param_count(void);
/* logits[i] of the n feature rows */
int glv_mlp_forward_dev(const float *d_params, const uint16_t *d_feats, int64_t n, float *d_logits,
                        void *stream);
/* gradient of the mean weighted logistic loss over all n rows (label-1 rows
 * weighted pos_weight) into d_grad[0..param_count), the loss into
 * d_grad[param_count]; deterministic */
int glv_mlp_grad_dev(const float *d_params, const uint16_t *d_feats, const uint8_t *d_labels,
                     int64_t n, float pos_weight, float *d_grad, void *stream);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8) from step number step0 on
 * batches of `batch` rows drawn with replacement (splitmix64 of seed);
 * d_m, d_v: Adam moments (zero before the first step); d_losses[steps]
 * (nullable) receives the loss of every step */
int glv_mlp_train_dev(float *d_params, float *d_m, float *d_v, const uint16_t *d_feats,
                      const uint8_t *d_labels, int64_t n, int64_t steps, int64_t batch, float lr,
                      float pos_weight, uint64_t seed, int64_t step0, float *d_losses, void *stream);

typedef struct glv_las_stats {
    int64_t swept_vertices;  /* vertex visits of every sweep */
    int64_t swept_entries;   /* CSR entries those visits read */
    int32_t full_sweeps;     /* iterations over every vertex */
    int32_t pruned_sweeps;   /* iterations over the active vertices only */
    int64_t samples;         /* trace rows recorded (may exceed the capacity) */
} glv_las_stats;

/* The unpruned colored run (same partition as glv_runMultiPhaseColoring_dev)
 * that also records sweep traces: from the second iteration of every colored
 * phase, a `rate` fraction of the vertices (splitmix64 of seed) gets its
 * features (as recorded by the previous sweep) written to d_feats[row][8] and
 * the vertex id to d_verts[row]; d_labels[row] = 1 when the iteration then
 * moves it.  At most cap rows. */
int glv_las_trace_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                      const double *d_weights, int64_t *d_C_orig, int64_t minGraphSize,
                      double threshold, double C_threshold, double rate, uint64_t seed,
                      uint16_t *d_feats, int64_t *d_verts, uint8_t *d_labels, int64_t cap,
                      void *stream, glv_stats *stats, glv_las_stats *las);
/* runMultiPhaseColoring with learned pruning: after the first iteration of a
 * colored phase only vertices whose predicted move probability is >= tau are
 * swept.  When a pruned iteration gains less than C_threshold, the next
 * iteration sweeps every vertex; the phase ends when such a full iteration
 * gains less than C_threshold, or after max_rechecks of them (0: no cap).
 * stats->edges_visited counts the entries actually read. */
int glv_runMultiPhaseColoring_las_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets,
                                      const int64_t *d_tails, const double *d_weights,
                                      int64_t *d_C_orig, int64_t minGraphSize, double threshold,
                                      double C_threshold, const float *d_params, float tau,
                                      int max_rechecks, void *stream, glv_stats *stats,
                                      glv_las_stats *las);

/* ---- batched graphs (no counterpart in the reference) ---------------------
 * runMultiPhaseBasic (runMultiPhaseBasic.cpp:52-170) on ngraphs independent
 * graphs at once.  The graphs are one device-resident union CSR: graph g owns
 * the vertices [vbase[g], vbase[g+1]) (vbase on the host, vbase[0] = 0,
 * vbase[ngraphs] = nv) and no entry joins two graphs.  Each graph follows its
 * own iterations, phases and termination exactly as a separate
 * glv_runMultiPhaseBasic_dev call would; all graphs of a phase are swept by
 * the same launches.  d_C_orig (nv) receives graph-local cluster ids.
 * Per-graph results go to the caller's host arrays (each nullable, ngraphs
 * entries). */
typedef struct glv_batch_stats {
    double *final_modularity;
    double *last_modularity;
    int32_t *phases;
    int32_t *total_iterations;
    int64_t *final_clusters;
    int32_t levels;      /* phases of the longest-running graph */
    int32_t iterations;  /* batched iterations over all levels (per level: its longest-running graph's) */
    double time_total;   /* seconds */
    int32_t launches;    /* graph launches of the levels' iteration loops (4 iterations each) */
    int32_t reserved;
} glv_batch_stats;
int glv_runMultiPhaseBasic_batch_dev(int64_t ngraphs, const int64_t *vbase, int64_t nv, int64_t nnz,
                                     const int64_t *d_offsets, const int64_t *d_tails,
                                     const double *d_weights, int64_t *d_C_orig, int basicOpt,
                                     double threshold, void *stream, glv_batch_stats *stats);

/* ---- sweep bin ranges (no counterpart in the reference) ------------------
 * max_deg[0..9]: upper degree of each bin of the sweep (bin 0 = no entries:
 * 0; defaults 4, 8, 16, 32, 64, 256, 2048, 16384, INT64_MAX).  The upper ends
 * of bins 1..6 may be lowered (never raised); bins 7..9 are fixed.
 * glv_autotune_bins_dev times one plain sweep of the given graph per
 * candidate and writes the fastest ranges (it does not apply them);
 * glv_set_bins applies ranges to every later call in the process. */
int glv_get_bins(int64_t *max_deg);
int glv_set_bins(const int64_t *max_deg);
int glv_autotune_bins_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                          const double *d_weights, int reps, int64_t *max_deg, double *ms_default,
                          double *ms_tuned, void *stream);

/* ---- launch shapes (no counterpart in the reference) ---------------------
 * The grids of the persistent coloring rounds, the persistent sweep (coarse
 * phases) and the whole-GPU sweep (large phases), and the degree at which
 * the coloring's core starts on graphs above 16384 vertices.  None changes a
 * partition on integer weights.  glv_set_tuning applies shapes to every
 * later call in the process; glv_autotune_dev times candidates on the given
 * graph -- the coloring for jp_blocks and core_min_deg, and with runs != 0
 * whole colored runs (parameters as glv_runMultiPhaseColoring_dev) for
 * persist_blocks and wide_blocks -- one knob at a time from the current
 * shapes, keeping a change that saves at least 2 %, and writes the fastest
 * to *best (made current when apply != 0).  trials (cap entries, may be
 * null) receive every measured candidate; *ntrials their count. */
typedef struct {
    int32_t jp_blocks;      /* 1..256, default 64 */
    int32_t persist_blocks; /* 1..256, default 64 (graphs above 15360 vertices: at most 64 used) */
    int32_t wide_blocks;    /* 0: one workgroup per CU (default), else at most that many */
    int32_t reserved;
    int64_t core_min_deg;   /* >= 1, default 1024 */
} glv_tuning;
typedef struct {
    glv_tuning t;
    int32_t stage;          /* 0: timed on the coloring, 1: on a whole colored run */
    int32_t reserved;
    double ms;              /* best of reps */
} glv_tune_trial;
int glv_get_tuning(glv_tuning *t);
int glv_set_tuning(const glv_tuning *t);
int glv_autotune_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                     const double *d_weights, int coloring, int64_t minGraphSize, double threshold,
                     double C_threshold, int reps, int runs, int apply, glv_tuning *best,
                     glv_tune_trial *trials, int cap, int *ntrials, void *stream);

/* ---- the other colorings of runMultiPhaseColoring ------------------------
 * algoColoringMultiHashMaxMin, DefineStructure/coloring.h:12 and
 * coloringMultiHashMaxMin.cpp:60-226 (coloring == 3): nItrs iterations of
 * nHash passes; in pass (itr, h) an uncoloured vertex whose hash value is a
 * strict maximum (minimum) among its uncoloured or just-coloured neighbours
 * takes colour 2*itr*nHash + 2h (+1).  Vertices left uncoloured hold
 * *maxColor = 2*nHash*nItrs (smaller when an iteration finds every vertex
 * coloured, :165-170).  The hash value of vertex v under hash h is the top
 * 53 bits of splitmix64(v ^ ((h+1) << 56)); the reference draws rand(). */
int glv_algoColoringMultiHashMaxMin(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList,
                                    int32_t *vtxColor, int nHash, int nItrs, int32_t *maxColor);
int glv_algoColoringMultiHashMaxMin_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                                        int32_t *d_colors, int nHash, int nItrs, int32_t *maxColor, void *stream);
/* vBaseRedistribution, DefineStructure/coloring.h:15 and vBase.cpp:6-168
 * (coloring == 2 applies it with type 0 after the distance-1 coloring).
 * avg = ceil(NV / ncolors); colours whose class starts above avg are
 * oversize.  In rounds, every vertex of an oversize base class whose current
 * class still holds more than avg wishes for an eligible colour (not
 * oversize, fewer than avg vertices, held by no neighbour): the first one
 * (type 0) or the least used, smaller on ties (type 1).  Adjacent vertices
 * with the same wish keep only the higher priority (splitmix64 of the id);
 * each wished colour admits up to avg - size wishers and each class lets go
 * of up to size - avg of them, in decreasing priority.  Rounds end when one
 * moves nothing.  The reference applies the same rules one vertex at a time
 * inside a parallel loop (its outcome depends on the schedule). */
int glv_vBaseRedistribution(int64_t nv, const int64_t *edgeListPtrs, const glv_edge *edgeList, int32_t *vtxColor,
                            int ncolors, int type, int32_t *rounds);
int glv_vBaseRedistribution_dev(int64_t nv, int64_t nnz, const int64_t *d_offsets, const int64_t *d_tails,
                                int32_t *d_colors, int ncolors, int type, int32_t *rounds, void *stream);

/* ---- cluster comparison (the -r metrics and driverForClusterComparison) --
 * computeCommunityComparisons, Utility/utilityClusteringFunctions.h:71 and
 * utilityClusterComparisonMetrics.cpp:50-304, C1 = truth.  N1 must equal N2.
 * Both label arrays are relabelled in place as the reference does: -1 in C1
 * becomes a new last community; -1 in C2 becomes N2 + 1.  A labelling is
 * zero-based when it holds a 0, else one-based.  The pair counts are those of
 * the reference at one thread (its nested parallel region multiplies the
 * printed counts by its thread count; precision, recall and F-score do not
 * change).  Labels below -1 are rejected. */
typedef struct glv_cluster_comparison {
    int64_t nC1, nC2;     /* "Num communities in C1/C2" */
    int64_t SameSame;     /* pairs together in both (true positives) */
    int64_t SameDiff;     /* together in C1 only (false negatives) */
    int64_t DiffSame;     /* together in C2 only (false positives) */
    int64_t Agree;        /* = SameSame, counted from C2's side */
    double precision, recall, fScore;
    double gini1, gini2;  /* computeGiniCoefficient of the community sizes */
} glv_cluster_comparison;
int glv_computeCommunityComparisons(int64_t *C1, int64_t N1, int64_t *C2, int64_t N2,
                                    glv_cluster_comparison *out);
int glv_computeCommunityComparisons_dev(int64_t *d_C1, int64_t N1, int64_t *d_C2, int64_t N2,
                                        glv_cluster_comparison *out, void *stream);
/* computeGiniCoefficient, basic_util.h:18 and
 * utilityClusterComparisonMetrics.cpp:309-330: sorts colorSize[0..numColors)
 * ascending in place and returns 2*sum((i+1)*s_i)/(numColors*sum(s_i)) -
 * (numColors+1)/numColors (exact while the sums stay below 2^53). */
int glv_computeGiniCoefficient(int64_t *colorSize, int numColors, double *gini);
int glv_computeGiniCoefficient_dev(int64_t *d_colorSize, int numColors, double *gini, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GRAPPOLO_HIP_H */
