// louvain.hpp -- host orchestration of the engine: one phase of the colored
// sweep, one phase of the plain sweep, renumbering, coarsening, coloring and
// the multi-phase drivers.  All device work is enqueued on one stream.
#pragma once

#include <vector>

#include "common.hpp"

namespace glv {

struct PhaseRecord {
    i64 nv = 0, nnz = 0;
    int iters = 0;
    int ncolors = 0;  // colors used by the phase's sweep (0: plain sweep)
    int colored = 0;
    double mod = 0;   // value returned by the phase's sweep
    i64 clusters = 0;
    double t_sweep = 0, t_color = 0, t_coarsen = 0, t_renumber = 0;
};

struct RunResult {
    int phases = 0;
    int total_iterations = 0;
    i64 final_clusters = 0;
    double final_modularity = -1;  // prevMod at exit (what the reference prints)
    double last_modularity = -1;   // currMod at exit
    double t_clustering = 0, t_building = 0, t_coloring = 0, t_renumber = 0, t_total = 0;
    i64 edges_visited = 0;      // sum over phases of iterations * nnz(phase graph)
    i64 edges_visited_mod = 0;  // separate modularity passes (colored phases)
    int colors_first = 0;
    i64 moves = 0;              // vertices moved by the colored sweeps of large graphs
    std::vector<PhaseRecord> phase;
    std::vector<double> trace;   // modularity of every iteration, all phases in order
};

// Learned active set (no counterpart in the reference).  mode 1 records
// sweep traces of the unpruned method; mode 2 sweeps, after the first
// iteration of a colored phase, only the vertices the MLP marks active, and
// ends a phase only when an iteration over every vertex gains less than the
// threshold.
struct LasParams {
    int mode = 0;               // 0 off, 1 trace capture, 2 prune
    const float *P = nullptr;   // device MLP parameters (kMlpParams FP32)
    float logit_tau = 0.0f;     // prune: active when the logit is >= this
    int max_rechecks = 0;       // prune: a phase ends after this many full re-checks (0: no cap)
    // trace capture: features (BF16 [cap][8]), vertex ids and next-sweep move labels
    double rate = 0.0;
    u64 seed = 0;
    uint16_t *tfeats = nullptr;
    i64 *tverts = nullptr;
    uint8_t *tlabels = nullptr;
    i64 cap = 0;
    i64 count = 0;              // samples recorded (may exceed cap: the rest were dropped)
    // prune statistics
    i64 swept_vertices = 0, swept_entries = 0;
    int full_sweeps = 0, pruned_sweeps = 0;
};

struct RunParams {
    int coloring = 1;    // 1: distance-1, 2: distance-1 then balanced, 3: incomplete multi-hash
    int numColors = 16;  // coloring 3: nItrs = max(numColors / 4, 1) iterations of 2 hashes
    int replaceMap = 0;  // 1: NoMap plain sweeps (initCommAssOpt start)
    i64 minGraphSize = 1000;
    double threshold = 1e-9;
    double C_threshold = 1e-6;
    int max_phases = 200;       // runMultiPhaseColoring.cpp:147 (phase > 200)
    int max_total_iters = 10000;  // :147 (totItr > 10000); 100000 for runMultiPhaseBasic
    LasParams *las = nullptr;     // learned active set of the colored phases (null: off)
};

// Per-phase device state shared by the sweeps: vertex degrees, community
// degrees/sizes and the vertices grouped by (degree bin, color).
struct PhaseSetup {
    int ncol = 1;
    DBuf<double> k, a;
    DBuf<i64> n;
    DBuf<i64> perm;
    DBuf<i64> dptr;        // device copy of ptr
    std::vector<i64> ptr;  // (kNumBins*ncol + 1): slice starts in perm, key = bin*ncol + color
    std::vector<i64> pnnz; // entries of the vertices of each slice
    double constant = 0;   // 1 / sum of degrees
    double total_weight = 0;
    i64 maxdeg_global = 0;  // largest degree in the HBM-hash bin
    // giant vertices (bin 9): the HBM table offsets of color c's class are
    // gt_tab[gt_seg[c] ..] (class-relative); the tables hold the largest class
    DBuf<i64> gt_tab, gt_keys, gt_ri;
    DBuf<double> gt_vals, gt_sc, gt_rg;
    DBuf<float> gt_rt;
    std::vector<i64> gt_seg;
    i64 lo(int bin, int color) const { return ptr[(size_t)bin * ncol + color]; }
    i64 hi(int bin, int color) const { return ptr[(size_t)bin * ncol + color + 1]; }
    i64 bin_lo(int bin) const { return ptr[(size_t)bin * ncol]; }
    i64 bin_hi(int bin) const { return ptr[(size_t)(bin + 1) * ncol]; }
    i64 nnz_of(int bin, int color) const { return pnnz[(size_t)bin * ncol + color]; }
    i64 bin_nnz(int bin) const {
        i64 t = 0;
        for (int c = 0; c < ncol; c++) t += nnz_of(bin, c);
        return t;
    }
};

struct SweepArgs;
// the bin-9 (giant vertex) fields of A for color class c of P
void giant_args(const PhaseSetup &P, int c, SweepArgs &A);

// The bin ranges every engine uses (process-wide; the autotuner's result).
BinConfig &current_bins();

// Launch shapes every engine uses (process-wide; the autotuner's result).
// None of them changes a result: they choose grids and which vertices the
// coloring's core pass takes.
struct Tuning {
    int jp_blocks = 64;                // grid of the persistent coloring rounds (<= kJpPersistMaxBlocks)
    i64 core_min_deg = kJpCoreMinDeg;  // a coloring core of a large graph starts at this degree or above
    int persist_blocks = 64;           // grid of the persistent sweep at most (<= kPersistMaxBlocks)
    int wide_blocks = 0;               // grid of the whole-GPU sweep (0: one workgroup per CU)
};
Tuning &current_tuning();
bool valid_tuning(const Tuning &t);

class Engine {
public:
    explicit Engine(hipStream_t s) : s_(s), bins_(current_bins()), tune_(current_tuning()) {}
    hipStream_t stream() const { return s_; }
    void set_bins(const BinConfig &b) { bins_ = b; }
    void set_tuning(const Tuning &t) { tune_ = t; }
    const Tuning &tuning() const { return tune_; }
    // Times one plain sweep of g over every vertex (best of `reps`) with the
    // given bins, in ms.
    double time_sweep(const DevGraph &g, const BinConfig &b, int reps);

    void setup(const DevGraph &g, const int32_t *colors, int ncol, PhaseSetup &P);

    // algoDistanceOneVertexColoringOpt: returns the largest color
    int color(const DevGraph &g, int32_t *colors, int *rounds = nullptr);
    // the coloring step of runMultiPhaseColoring.cpp:68-87 for p.coloring;
    // returns nColors (largest color + 1)
    int color_for_run(const DevGraph &g, int32_t *colors, const RunParams &p);
    i64 count_conflicts(const DevGraph &g, const int32_t *colors);

    // algoLouvainWithDistOneColoring: C (device, nv) receives the labels.
    // valid_colors: the classes are independent sets of g (a fresh distance-1
    // coloring), so a color step's moves change the internal weight by exactly
    // the movers' 2 (e_vt - e_vs) and iterations after the first need no full
    // modularity pass
    double colored_phase(const DevGraph &g, i64 *C, const int32_t *colors, int ncol, double Lower,
                         double thresh, int *iters, std::vector<double> *trace = nullptr,
                         LasParams *las = nullptr, int phase_index = 0, bool valid_colors = false);
    // parallelLouvianMethod(NoMap): C receives pastCommAss
    double plain_phase(const DevGraph &g, i64 *C, double Lower, double thresh, bool nomap,
                       int *iters, std::vector<double> *trace = nullptr);
    // one color step from an arbitrary labelling (community info derived from C)
    i64 color_step(const DevGraph &g, i64 *C, const int32_t *colors, int ncol, int which);

    i64 renumber(i64 *C, i64 n);
    // vf: buildNewGraphVF (Utility/vertexFollowing.cpp:93-258) instead of
    // buildNextLevelGraphOpt: no zero self-loops, weights truncated to integers
    void coarsen(const DevGraph &g, const i64 *C, i64 k, DevGraph &out, bool vf = false);
    // vertexFollowing (Utility/vertexFollowing.cpp:46-89); returns the vertices to remove
    i64 vertex_following(const DevGraph &g, i64 *C);
    // modularity of a labelling with community degrees summed from C
    double modularity(const DevGraph &g, const i64 *C);

    void run_colored(const DevGraph &g, i64 *C_orig, const RunParams &p, RunResult &r);
    void run_basic(const DevGraph &g, i64 *C_orig, const RunParams &p, RunResult &r);

    void sync();

    // vertices moved by the whole-GPU sweeps since the engine was made
    i64 moves() const { return moves_; }

private:
    hipStream_t s_;
    BinConfig bins_;
    Tuning tune_;
    i64 moves_ = 0;
    // small pinned-free scratch for scalars copied back to the host
    template <class T>
    T fetch(const T *dptr);
    double colored_modularity(const DevGraph &g, const PhaseSetup &P, const i64 *C,
                              double *part, double *res, double *exx = nullptr);
};

// Sweep bin ranges that time fastest on g (autotune.cpp); the default
// ranges' and the result's sweep times in ms.
BinConfig autotune_bins(Engine &E, const DevGraph &g, int reps, double *ms_default, double *ms_tuned);

// One measured candidate of the launch-shape tuner.
struct TuneTrial {
    Tuning t;
    int stage;  // 0: coloring, 1: whole colored run
    double ms;
};
// Launch shapes that time fastest on g (autotune.cpp): the coloring's grid
// and core degree by timing the coloring of g, then (runs > 0) the sweep
// grids by timing whole colored runs of g with parameters p; candidates one
// knob at a time, each from the best so far, a change kept when it saves at
// least 2 %.  The trials go to log.
Tuning autotune_launch(Engine &E, const DevGraph &g, const RunParams &p, int reps, bool runs,
                       std::vector<TuneTrial> *log);

}  // namespace glv
