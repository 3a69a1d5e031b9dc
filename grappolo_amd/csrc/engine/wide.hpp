// wide.hpp -- whole-GPU persistent colored iteration for large graphs.
//
// One launch per colored iteration (all color steps of the iteration, the
// per-color loop of Coloring/parallelLouvainWithColoring.cpp:169-218 with the
// community updates of :213-217) on a grid of one 1024-thread workgroup per
// CU, with one grid barrier per color step.  The work of a color step is a
// list of items fixed when the phase starts:
//   workgroup items: a heavy vertex (257..kWHeavyMaxDeg entries, LDS table of
//     the whole workgroup), a chunk of kWChunk entries of a giant vertex
//     (summed per community in LDS, then written out as (community, weight)
//     pairs grouped by a hash partition of the community ids) and a
//     partition of a giant's communities (its pairs from every chunk summed in
//     LDS and scored; the last partition scored decides the move);
//   wavefront items: one mid vertex (65..256 entries, a wavefront's LDS table)
//     or 64/G light vertices of degree bin b (G lanes each, the in-order sums
//     of buildLocalMapCounter, Utility/utilityClusteringFunctions.cpp:172-194).
// Labels and the adjacency's neighbour ids are held as 32-bit copies for the
// phase (halving the bytes of the random label gathers and of the neighbour
// stream), so the sweep serves graphs of fewer than 2^31 vertices.
// Items are dealt to workgroups when the phase starts (longest first, each to
// the least loaded workgroup; a workgroup's wavefronts take its wavefront
// items in turn), so a run is reproducible.  Labels, community degrees and sizes are
// double-buffered as in the small-graph persistent sweep: step k reads the
// state after step k-1 and writes into the other buffer the moves of steps
// k-1 and k.
#pragma once

#include <vector>

#include "kernels.hpp"

namespace glv {

constexpr int kWBlock = 1024;            // threads per workgroup
constexpr int kWHeavySlots = 12288;      // LDS table of a workgroup: u32 keys + f64 values
constexpr i64 kWHeavyMaxDeg = 8192;      // larger vertices are giants (several workgroups)
constexpr i64 kWChunk = 4096;            // entries of a giant per accumulation item
constexpr i64 kWPartKeys = 4096;         // communities per partition of a giant (expected)
constexpr int kWMaxParts = 1024;         // partitions of a giant at most
constexpr int kWMaxChunks = 1024;        // chunks of a giant at most (4,194,304 entries; a phase
                                         // with a larger vertex takes the per-color launches)
constexpr int kWMidSlots = 512;          // table of a wavefront (mid vertices: <= 256 entries)
// Phases whose color classes hold more vertices than this on average take the
// per-color launches instead (a class that large fills the GPU by itself).
constexpr i64 kWideMaxPerColor = 100000;

// Relative costs of workgroup items (microseconds, measured order of
// magnitude): a heavy vertex kWCostHeavy plus three per 1024 entries.
constexpr double kWCostHeavy = 10.0, kWCostChunk = 10.0, kWCostPart = 10.0;
// wavefront items: a mid vertex, a batch of light vertices (a workgroup runs
// kWWavesPerGroup of them at a time)
constexpr double kWCostMid = 8.0, kWCostLight = 6.0;
constexpr int kWWavesPerGroup = kWBlock / kWave;

// Item kinds.  Workgroup items: x = vertex (heavy, y = its entries) or giant
// index (y = chunk or slice).  Wavefront items: x = first position in perm,
// y = vertices.
enum WideKind : int32_t {
    kWHeavy = 0,
    kWGiantChunk = 1,
    kWGiantPart = 2,
    kWMid = 3,
    kWLight1 = 4,  // .. kWLight1 + 4: light vertices of degree bins 1..5
};
struct WideItem {
    i64 x;
    int32_t kind;
    int32_t y;
};
// A giant vertex of the phase: chunk j writes its pairs at pairs + j *
// kWChunk (partition p's from cnt[j * (nparts + 1) + p] on, relative) and the
// partition bounds at cnt + j * (nparts + 1) (offsets into the pair arrays and
// the count array of its color class).
struct WideGiant {
    i64 v;
    i64 pairs;
    i64 cnt;
    int32_t nchunks;
    int32_t nparts;
};

// Degree and size of a community side by side (one 16-byte load serves both,
// as the reference's Comm record, DefineStructure/defs.h:66-70, holds them).
struct CommInfo {
    double a;
    i64 n;
};

struct WideArgs {
    SweepArgs A;                 // graph, k, constant, tgt
    int32_t *C0, *C1;            // labels, two 32-bit buffers (C0: the phase's labels)
    const int32_t *tails;        // the adjacency's neighbour ids, 32-bit
    double *a;                   // community degrees and sizes of the phase, read at the
    i64 *n;                      //   start of the iteration and written at its end
    CommInfo *m0, *m1;           // community degrees and sizes, two buffers
    const i64 *perm;             // vertices grouped by (degree bin, color)
    const i64 *cperm;            // vertices grouped by color
    const i64 *cstart;           // ncol + 1
    const WideItem *witems;      // workgroup items of color c for workgroup w at wstart[c*nb+w] .. wstart[c*nb+w+1]
    const i64 *wstart;
    const WideItem *vitems;      // wavefront items of color c for workgroup w at vstart[c*nb+w] .. (its
                                 //   wavefronts take them in turn)
    const i64 *vstart;
    const WideGiant *giants;
    i64 ngiants;
    uint32_t *gpk;               // giant pairs: community ids (one class at a time)
    double *gpv;                 //   and summed weights
    int32_t *gcnt;               // partition bounds of every chunk
    int32_t *gstate;             // per giant: chunks written, pairs ready, partitions scored
    double *gsc;                 // per giant: self-loop weight, weight to its own community
    double *gcand;               // per giant and partition: gain, weight to the candidate, its size
    i64 *gcid;                   // per giant and partition: candidate (-1: none)
    i64 gparts;                  // partitions per giant in gcand/gcid (the largest nparts)
    double *gout;                // per giant: 2(e_t - e_s) of its move this iteration, 0 if none
    int ncol;
    unsigned *bar;               // 16 + 16 * grid words, zero at launch
    double *part;                // per workgroup: [0, nb) movers' 2 (e_t - e_s), [nb, 2nb) sum of a^2
    i64 *moves;                  // vertices moved (added once per workgroup)
};

struct PhaseSetup;

// The items and giant tables of one phase.
struct WidePlan {
    DBuf<WideItem> witems, vitems;
    DBuf<i64> wstart, vstart;
    DBuf<WideGiant> giants;
    DBuf<uint32_t> gpk;
    DBuf<double> gpv, gsc, gcand, gout;
    DBuf<int32_t> gcnt, gstate;
    DBuf<i64> gcid;
    i64 ngiants = 0, gparts = 1, gpairs = 0, gcounts = 0;
    i64 nwitems = 0, nvitems = 0;
};
// Builds the items of every color class of P (vertices of g grouped by
// degree bin and color) for a grid of nb workgroups; false (nothing built)
// when a vertex has more entries than kWMaxChunks chunks hold.
bool build_wide_plan(const DevGraph &g, const PhaseSetup &P, int nb, WidePlan &L, hipStream_t s);

// false when the environment sets GLV_WIDE=0 (large graphs then take the
// per-color launches; for comparisons)
bool wide_enabled();
// Resident workgroups of the wide sweep on the current device (0: unavailable).
int wide_blocks();
// One colored iteration; false when the launch is not possible.
bool launch_wide_iteration(const WideArgs &W, int nblocks, hipStream_t s);
// (vertex, degree) of perm[lo .. hi) into vd[2 * (hi - lo)]
void launch_wide_degrees(const i64 *perm, const i64 *off, i64 lo, i64 hi, i64 *vd, hipStream_t s);
void launch_wide_clear_giants(int32_t *gstate, double *gsc, double *gout, i64 ng, hipStream_t s);
// 64-bit ids to 32-bit (all below 2^31) and back
void launch_narrow_ids(const i64 *src, i64 n, int32_t *dst, hipStream_t s);
void launch_widen_ids(const int32_t *src, i64 n, i64 *dst, hipStream_t s);

}  // namespace glv
