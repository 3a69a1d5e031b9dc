// batch.hpp -- batched plain Louvain (runMultiPhaseBasic) over many
// independent graphs held as one block-diagonal union CSR.  Every graph
// follows its own phases, iterations and termination exactly as a separate
// run would; all graphs of a phase are swept by the same launches.
#pragma once

#include <vector>

#include "common.hpp"
#include "louvain.hpp"

namespace glv {

struct BatchParams {
    int basicOpt = 0;             // 1: NoMap sweeps (initCommAssOpt start)
    double threshold = 1e-9;
    int max_phases = 200;         // runMultiPhaseBasic.cpp phase cap
    int max_total_iters = 100000;
};

struct BatchResult {
    std::vector<double> final_modularity, last_modularity;
    std::vector<int32_t> phases, total_iterations;
    std::vector<i64> clusters;
    int levels = 0;       // union levels (phases of the longest-running graph)
    int iterations = 0;   // batched iterations over all levels (per level: its longest-running graph's)
    int launches = 0;     // graph launches of the iteration loops (kLevelUnroll iterations each)
    double t_total = 0;
};

// g0: the union graph; vb0[ngraphs+1]: first vertex of every graph (host).
// C_orig (device, g0.nv) receives graph-local labels.
void run_basic_batch(Engine &E, const DevGraph &g0, const std::vector<i64> &vb0, i64 *C_orig,
                     const BatchParams &p, BatchResult &r);

// --- kernels (batch.hip) ---------------------------------------------------
void launch_fill_gid(const i64 *vb, i64 ng, int32_t *gid, hipStream_t s);
void launch_graph_const(const i64 *vb, i64 ng, const double *k, double *gconst, hipStream_t s);
// skip: nothing to do when *skip is set (after every graph's phase has ended)
void launch_graph_mod(const i64 *vb, i64 ng, const uint8_t *active, const double *eii, const double *a,
                      double *exx, double *a2, hipStream_t s, const int32_t *skip = nullptr);
void launch_batch_decide(i64 ng, const double *exx, const double *a2, const double *gconst, const double *lower,
                         double thresh, int maxitr, double *prev, double *curr, int32_t *itrs, uint8_t *active,
                         uint8_t *apply, hipStream_t s);
void launch_level_status(const i64 *vb, i64 ng, const uint8_t *active, int32_t *done, i64 *live, hipStream_t s);
void launch_batch_apply(i64 nv, const int32_t *gid, const uint8_t *apply, const i64 *tgt, i64 *curr, i64 *past,
                        const double *k, double *a, i64 *n, hipStream_t s, const int32_t *skip = nullptr);
void launch_compact_graphs(const i64 *list, i64 n, const int32_t *gid, const uint8_t *active, const i64 *start,
                           i64 nkeys, i64 *flags, i64 *pos, i64 *out, i64 *out_start, hipStream_t s);
void launch_graph_label_range(const i64 *vb, i64 ng, const i64 *C, i64 *lo, i64 *hi, hipStream_t s);
void launch_batch_compose(i64 n0, const int32_t *gid0, const uint8_t *alive, i64 *C_orig, const i64 *C,
                          bool first, hipStream_t s);
void launch_batch_finalize(i64 n0, const int32_t *gid0, const uint8_t *stop, const i64 *base, i64 *C_orig,
                           hipStream_t s);
void launch_gather_i64(const i64 *src, const i64 *idx, i64 n, i64 *dst, hipStream_t s);
void launch_mask_graphs(const i64 *vb, i64 ng, const uint8_t *keep, i64 *C, hipStream_t s);
void launch_check_blocks(const i64 *off, const i64 *tails, const i64 *vb, i64 ng, int32_t *err, hipStream_t s);

}  // namespace glv
