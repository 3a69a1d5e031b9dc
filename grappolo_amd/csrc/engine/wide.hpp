// wide.hpp -- whole-GPU persistent colored iteration for large graphs.
//
// One launch per colored iteration (all color steps of the iteration, the
// per-color loop of Coloring/parallelLouvainWithColoring.cpp:169-218 with the
// community updates of :213-217) on a grid of one 1024-thread workgroup per
// CU, with one grid barrier per color step.  The work of a color step is a
// list of items fixed when the phase starts:
//   workgroup items: a heavy vertex (257..kWHeavyMaxDeg entries, LDS table of
//     the whole workgroup), a chunk of kWChunk entries of a giant vertex
//     (summed in LDS, then added into the giant's HBM table) and a slice of a
//     giant's table to score (the last slice scored decides the move);
//   wavefront items: one mid vertex (65..256 entries, a wavefront's LDS table)
//     or 64/G light vertices of degree bin b (G lanes each, the in-order sums
//     of buildLocalMapCounter, Utility/utilityClusteringFunctions.cpp:172-194).
// Items are dealt to workgroups and wavefronts in a fixed (snake) order, so a
// run is reproducible.  Labels, community degrees and sizes are
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
constexpr i64 kWChunk = 8192;            // entries of a giant per accumulation item
constexpr i64 kWScoreSlice = 8192;       // slots of a giant's table per scoring item
constexpr int kWMidSlots = 512;          // table of a wavefront (mid vertices: <= 256 entries)
// Phases whose color classes hold more vertices than this on average take the
// per-color launches instead (a class that large fills the GPU by itself).
constexpr i64 kWideMaxPerColor = 100000;

// Relative costs of workgroup items (microseconds, measured order of
// magnitude): a heavy vertex kWCostHeavy plus one per 1024 entries.
constexpr double kWCostHeavy = 10.0, kWCostChunk = 16.0, kWCostSlice = 8.0;

// Item kinds.  Workgroup items: x = vertex (heavy, y = its entries) or giant
// index (y = chunk or slice).  Wavefront items: x = first position in perm,
// y = vertices.
enum WideKind : int32_t {
    kWHeavy = 0,
    kWGiantChunk = 1,
    kWGiantScore = 2,
    kWMid = 3,
    kWLight1 = 4,  // .. kWLight1 + 4: light vertices of degree bins 1..5
};
struct WideItem {
    i64 x;
    int32_t kind;
    int32_t y;
};
// A giant vertex of the phase: its HBM table (u32 keys, kEmptyKey when free;
// f64 values, 0 when free) of `slots` slots at table offset `tab`.
struct WideGiant {
    i64 v;
    i64 tab;
    i64 slots;
    int32_t nchunks;
    int32_t nslices;
};

// Degree and size of a community side by side (one 16-byte load serves both,
// as the reference's Comm record, DefineStructure/defs.h:66-70, holds them).
struct CommInfo {
    double a;
    i64 n;
};

struct WideArgs {
    SweepArgs A;                 // graph, k, constant, tgt
    i64 *C0, *C1;                // labels, two buffers (C0: the phase's labels)
    double *a;                   // community degrees and sizes of the phase, read at the
    i64 *n;                      //   start of the iteration and written at its end
    CommInfo *m0, *m1;           // community degrees and sizes, two buffers
    const i64 *perm;             // vertices grouped by (degree bin, color)
    const i64 *cperm;            // vertices grouped by color
    const i64 *cstart;           // ncol + 1
    const WideItem *witems;      // workgroup items of color c for workgroup w at wstart[c*nb+w] .. wstart[c*nb+w+1]
    const i64 *wstart;
    const WideItem *vitems;      // wavefront items, color c at vstart[c] .. vstart[c+1]
    const i64 *vstart;
    const WideGiant *giants;
    i64 ngiants;
    uint32_t *gkeys;             // giant tables (one class at a time)
    double *gvals;
    int32_t *gstate;             // per giant: chunks added, table ready, slices scored
    double *gsc;                 // per giant: self-loop weight, weight to its own community
    double *gcand;               // per giant and slice: gain, weight to the candidate, its size
    i64 *gcid;                   // per giant and slice: candidate (-1: none)
    i64 gslices;                 // slices per giant in gcand/gcid (the largest nslices)
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
    DBuf<uint32_t> gkeys;
    DBuf<double> gvals, gsc, gcand, gout;
    DBuf<int32_t> gstate;
    DBuf<i64> gcid;
    i64 ngiants = 0, gslices = 1, gtable = 0;
    i64 nwitems = 0, nvitems = 0;
};
// Builds the items of every color class of P (vertices of g grouped by
// degree bin and color) for a grid of nb workgroups; the giant tables start
// empty.
void build_wide_plan(const DevGraph &g, const PhaseSetup &P, int nb, WidePlan &L, hipStream_t s);

// false when the environment sets GLV_WIDE=0 (large graphs then take the
// per-color launches; for comparisons)
bool wide_enabled();
// Resident workgroups of the wide sweep on the current device (0: unavailable).
int wide_blocks();
// One colored iteration; false when the launch is not possible.
bool launch_wide_iteration(const WideArgs &W, int nblocks, hipStream_t s);
// (vertex, degree) of perm[lo .. hi) into vd[2 * (hi - lo)]
void launch_wide_degrees(const i64 *perm, const i64 *off, i64 lo, i64 hi, i64 *vd, hipStream_t s);
void launch_wide_clear_giants(uint32_t *gkeys, double *gvals, i64 slots, int32_t *gstate, double *gsc,
                              double *gout, i64 ng, hipStream_t s);

}  // namespace glv
