// kernels.hpp -- host-side launchers of the HIP kernels in kernels.hip.
// Every launcher enqueues on `s` and returns immediately.
#pragma once

#include "common.hpp"

namespace glv {

// Arguments shared by the sweep kernels.  `C`, `a`, `n` are the frozen state
// of the current color step (or Jacobi iteration); `tgt[v]` receives the
// chosen community of every vertex in `verts[0..count)`.
struct SweepArgs {
    const i64 *off;
    const i64 *tails;
    const double *w;
    const i64 *C;
    const double *k;
    const double *a;
    const i64 *n;
    double constant;
    const i64 *verts;
    i64 count;
    i64 *tgt;
    double *eii_part;  // plain sweeps: per-block sums of e_{i,C(i)}, else null
    i64 *gkeys;        // global hash (bin 8): per-block regions of gslots slots
    double *gvals;
    i64 gslots;
    i64 nv;            // vertices of the graph (community ids are < nv)
    // learned active set (both null on the plain path):
    float4 *feat = nullptr;  // per swept vertex: own-community fraction, top-two gains
    double *dein = nullptr;  // per swept vertex: 2 (e_{v,t} - e_{v,s}) of its move, 0 if none
    // persistent sweep above the dense-accumulator limit: nv zeros per workgroup
    // (kept zero between vertices) for the vertices too large for its LDS table
    double *dacc = nullptr;
    // giant vertices (bin 9) of the slice: HBM hash table of vertex q at
    // gt_keys/gt_vals + gt_tab[q], gt_tab[q+1] - gt_tab[q] slots (kEmpty / 0
    // between uses); gt_sc[2q..2q+1]: weight to its own community and
    // self-loop weight (0 between uses); gt_rg/gt_ri/gt_rt: the best move of
    // each of its kGiantBlocks workgroups
    // batched graphs: vertex v belongs to graph gid[v] whose 1/(2m) is
    // gconst[gid[v]] (null: one graph, `constant`); eii_v[v] receives the
    // weight from v to its own community (plain sweeps, per-graph modularity)
    const int32_t *gid = nullptr;
    const double *gconst = nullptr;
    double *eii_v = nullptr;
    // batched graphs: the sweep returns at once when *skip is set (a captured
    // iteration after every graph's phase has ended)
    const int32_t *skip = nullptr;
    const i64 *gt_tab = nullptr;
    i64 *gt_keys = nullptr;
    double *gt_vals = nullptr;
    double *gt_sc = nullptr;
    double *gt_rg = nullptr;
    i64 *gt_ri = nullptr;
    float *gt_rt = nullptr;
};

// Up to one slice per degree bin of a color class (perm[lo[b] .. hi[b])).
struct Slices {
    i64 lo[kNumBins];
    i64 hi[kNumBins];
};

// --- elementwise / setup -------------------------------------------------
void launch_fill_i64(i64 *p, i64 n, i64 v, hipStream_t s);
void launch_fill_f64(double *p, i64 n, double v, hipStream_t s);
void launch_fill_i32(int32_t *p, i64 n, int32_t v, hipStream_t s);
void launch_iota(i64 *p, i64 n, hipStream_t s);
void launch_init_labels(const i64 *off, i64 nv, i64 *C, hipStream_t s);  // C[i] = deg? i : -1
void launch_bin_keys(const i64 *off, i64 nv, const int32_t *colors, int ncol, uint32_t *keys,
                     i64 *vid, int32_t *err, hipStream_t s, const BinConfig &bins);
// slice starts (nkeys+1) and entries per slice (nkeys) of keys sorted with perm
void launch_slices(const uint32_t *sorted_keys, const i64 *perm, const i64 *off, i64 n, i64 nkeys,
                   i64 *start, i64 *pnnz, hipStream_t s);
void launch_degree(const DevGraph &g, const i64 *big, i64 nbig, double *k, hipStream_t s);
void launch_init_comm(const double *k, i64 nv, double *a, i64 *n, hipStream_t s);
void launch_comm_from_labels(const i64 *C, const double *k, i64 nv, double *a, i64 *n,
                             hipStream_t s);

// --- reductions ----------------------------------------------------------
constexpr int kRedBlocks = 1024;
// out[0] = sum(x[0..n)), two-level: kRedBlocks block partials into part, then one block
void launch_sum(const double *x, i64 n, double *part, double *out, hipStream_t s);
void launch_sumsq(const double *x, i64 n, double *part, double *out, hipStream_t s);
void launch_sum_partials(const double *part, i64 np, double *out, hipStream_t s);

// --- sweeps --------------------------------------------------------------
// number of block partials the sweep of `bin` writes for `count` vertices
i64 sweep_blocks(int bin, i64 count, i64 nv);
void launch_sweep(int bin, bool plain, const SweepArgs &A, hipStream_t s);
void launch_max_degree(const i64 *off, const i64 *verts, i64 count, i64 *out, hipStream_t s);
constexpr int kGlobalHashBlocks = 256;
// One colored iteration of a graph with at most kDenseMaxNV vertices in one
// persistent launch (all colors, commits, modularity partials: part[0..nb)
// internal weight, part[nb..2nb) sum of a^2).  bar: 3 zeroed words.  Returns
// false when the persistent launch is not possible (caller falls back).
int persistent_blocks();
// vertices grouped by color (stable: within a color in perm's order); cstart: ncol+1
void launch_color_major(const i64 *perm, const int32_t *colors, i64 n, int ncol, i64 *cperm,
                        i64 *cstart, hipStream_t s);
// C1/a1/n1: second buffers of the labels, community degrees and sizes (nv each);
// L16 (nullable; 2 nv): scratch of 16-bit label copies, given for graphs of
// at most kDenseMaxNV vertices on an even grid, which take the variant whose
// halves alternate
bool launch_colored_iteration(const SweepArgs &A, const i64 *perm, const i64 *start, int ncol,
                              unsigned *bar, double *part, int nblocks, hipStream_t s, uint8_t *stale,
                              i64 *C1, double *a1, i64 *n1, uint16_t *L16);
void launch_isolated_targets(const i64 *verts, i64 count, i64 *tgt, hipStream_t s);
// stale (nullable): reset to 0 for every mover; dsum (nullable): += sum of
// dein over the movers, in a fixed order (dpart: kApplyMaxBlocks scratch)
constexpr i64 kApplyMaxBlocks = 8192;
void launch_apply_colored(const i64 *perm, const Slices &S, const i64 *tgt, i64 *C, const double *k,
                          double *a, i64 *n, hipStream_t s, uint8_t *stale = nullptr,
                          const double *dein = nullptr, double *dsum = nullptr, double *dpart = nullptr);
void launch_apply_plain(i64 nv, const i64 *tgt, const i64 *curr, const double *k, double *a, i64 *n,
                        hipStream_t s);
void launch_init_comm_ass_opt(const DevGraph &g, const double *k, const double *a, const i64 *n,
                              double constant, i64 *curr, hipStream_t s, const int32_t *gid = nullptr,
                              const double *gconst = nullptr);
void launch_update_ax_opt(i64 nv, const i64 *curr, const double *k, double *a, i64 *n,
                          hipStream_t s);

// slots of the HBM hash table of a giant vertex of degree d
inline i64 giant_table_slots(i64 d) {
    i64 s = 64;
    while (s < 2 * d) s <<= 1;
    return s;
}
// per giant vertex q of verts[0..count): tab[q] = table slots (for a scan)
void launch_giant_table_sizes(const i64 *off, const i64 *verts, i64 count, i64 *tab, hipStream_t s);

// --- modularity ----------------------------------------------------------
i64 mod_blocks(int bin, i64 count);
void launch_mod_pass(int bin, const DevGraph &g, const i64 *C, const i64 *verts, i64 count,
                     double *part, hipStream_t s);

// --- coloring ------------------------------------------------------------
// Jones-Plassmann with degree-first priorities, computed as a topological
// sweep of the priority DAG (a vertex is coloured once its higher-priority
// neighbours are, with the smallest colour they do not use).  The round of a
// vertex in the classic formulation is its depth in that DAG (1 + the largest
// depth of its higher neighbours); the rounds end once at most kJpTail
// vertices are deeper than the last round, and those are coloured first-fit
// one at a time in increasing id.  The core -- the vertices of the largest
// degrees, a prefix of the priority order -- is coloured first by one
// workgroup from LDS (k_jp_core), the rest by frontier rounds that take the
// core's colours as given; depths are tracked, so the tail is the same set
// whatever order computed the colours.
//
// Frontier lists of the rounds by degree: small (<= 256, a wavefront each),
// large (<= kGiantDeg, a workgroup each), giant (kGiantBlocks workgroups
// each); counts[0..2] their lengths.
constexpr u64 kJpNoCore = u64(1) << 24;  // above every priority's degree field
struct JpLists {
    i64 *small, *large, *giant;
    i64 *counts;
    const u64 *prio = nullptr;  // priority of every vertex (launch_jp_prio)
    // one byte per adjacency entry j (row v, neighbour u = tails[j]), written
    // by the count: kJpRelease when u has the lower priority (v releases u),
    // kJpLarge / kJpHuge / kJpGiant when u has more than 256 / kJpMidDeg /
    // kGiantDeg entries, kJpSelf for u == v, kJpCore when u is a core vertex;
    // no release bit: u's colour counts for v
    uint8_t *code = nullptr;
    int32_t *dep = nullptr;      // depth of every coloured vertex
    u64 dcore = kJpNoCore;       // core: vertices of at least this degree (kJpNoCore: none)
    int32_t *core_cnt = nullptr; // per core vertex q: [2q] higher, [2q+1] lower core neighbours
};
constexpr uint8_t kJpRelease = 1, kJpLarge = 2, kJpGiant = 4, kJpSelf = 8, kJpHuge = 16, kJpCore = 32;
constexpr i64 kJpMidDeg = 1024;
// core vertices at most (one workgroup's LDS: state, counts and queue)
constexpr i64 kJpCoreMax = 16384;
// Jones-Plassmann priorities: (min(degree, 2^23 - 1) << 40) | (splitmix64(v) >> 24),
// equal keys broken by the larger id
void launch_jp_prio(const DevGraph &g, u64 *prio, hipStream_t s);
// count[v] = uncoloured non-core higher-priority neighbours of a non-core v;
// count-0 vertices form the first frontier.  A core vertex v holds its core
// index in count[v] on entry; its core neighbour counts go to L.core_cnt.
// heavy, giant: nv scratch each, heavy_count, giant_count: one zeroed word each.
void launch_jp_count(const DevGraph &g, int32_t *cnt, const JpLists &L, i64 *heavy, i64 *heavy_count,
                     i64 *giant, i64 *giant_count, hipStream_t s);
// color one frontier (n[0..2] vertices in cur's lists) and build the next
// one into next (whose counts are zero); gmark: n[2] * (kMaxColors/32 + 1) words
void launch_jp_frontier(const DevGraph &g, int32_t *colors, int32_t *cnt, const JpLists &cur,
                        const i64 n[3], const JpLists &next, uint32_t *gmark, int32_t *err,
                        hipStream_t s);
void launch_jp_finish(int32_t *state, i64 nv, int32_t *maxc, hipStream_t s);
// first-fit, one vertex at a time in increasing id, of the n uncoloured vertices
// (the n vertices are listed, then the colours of their coloured neighbours
// and their earlier tail neighbours are gathered in parallel, then one
// workgroup picks the colours in order).  flags, pos, list: nv each; tidx: nv;
// base: n * kMaxColors/32 words; adj: n * kJpTail/64 words; n <= kJpTail.
void launch_color_tail(const DevGraph &g, int32_t *state, i64 *flags, i64 *pos, i64 *list, i64 n,
                       int32_t *tidx, uint32_t *base, u64 *adj, int32_t *err, hipStream_t s);
constexpr i64 kJpTail = 1024;   // the tail: at most this many vertices
// Jones-Plassmann rounds in one persistent launch (same result as the
// host-driven rounds), continuing from the frontier in sA (small, count
// ring[0]) and lA (large, ring[1]) after colored0 vertices; false when the
// launch is refused.  ring: 10 words (zero except the frontier counts),
// ring[6] receives the rounds; bar: kJpBarrierWords zeroed words; err: 1 too
// many colors.  sA..lB: nv each.
int jp_persistent_blocks();
bool launch_jp_persistent(const DevGraph &g, int32_t *colors, int32_t *cnt, i64 *sA, i64 *sB, i64 *lA,
                          i64 *lB, i64 *ring, unsigned *bar, int32_t *err, i64 colored0, const u64 *prio,
                          uint8_t *code, int32_t *dep, i64 max_front, int nblocks, hipStream_t s);
// grid of the persistent rounds at most (bar: kJpBarrierWords words)
constexpr i64 kJpPersistMaxBlocks = 256;
constexpr int kJpBarrierWords = 16 + 16 * (int)kJpPersistMaxBlocks;
// frontiers below this many vertices finish in one persistent launch
constexpr i64 kJpPersistFrontier = 16384;
// a persistent launch hands back to full-GPU rounds once a frontier exceeds this
constexpr i64 kJpPersistExit = 32768;
// The core (jp_core.hip).  hist: kJpDegBuckets zeroed words, the vertex
// count per degree bucket (degrees below 8 exact, then 8 buckets per octave;
// jp_bucket_lo(b) is the smallest degree of bucket b).
constexpr int kJpDegBuckets = 512;
i64 jp_bucket_lo(int b);
void launch_jp_degree_hist(const DevGraph &g, unsigned long long *hist, hipStream_t s);
// hlist[0 .. K) = the vertices of degree >= dcore in increasing id, count[v] =
// their index (flags, pos: nv scratch)
void launch_jp_core_list(const DevGraph &g, u64 dcore, i64 *flags, i64 *pos, i64 *hlist, int32_t *cnt,
                         hipStream_t s);
// sizes[q] = core neighbours of core vertex q (q < K), sizes[K] = 0
void launch_jp_core_sizes(const int32_t *core_cnt, i64 K, i64 *sizes, hipStream_t s);
// ladj[loff[q] ..): the core indices of core vertex q's higher core
// neighbours, then of its lower ones (codes of the count)
void launch_jp_core_fill(const DevGraph &g, const uint8_t *code, const int32_t *cnt, const i64 *hlist, i64 K,
                         const i64 *loff, const int32_t *core_cnt, uint16_t *ladj, hipStream_t s);
// colours and depths of the K core vertices, one workgroup; stop >= 0 (the
// core is the whole graph): the rounds end after the first round that leaves
// at most stop vertices uncoloured, which get depth rounds + 1 (deeper than
// every coloured vertex) for the tail
void launch_jp_core(const i64 *hlist, i64 K, const i64 *loff, const int32_t *core_cnt, const uint16_t *ladj,
                    int32_t *colors, int32_t *dep, int32_t *err, int stop, hipStream_t s);
// largest depth (maxd: one word, zero on entry)
void launch_jp_depth_max(const int32_t *dep, i64 nv, int32_t *maxd, hipStream_t s);
// hist[i] = vertices of depth top - i, i < kJpDepthWindow (zeroed words)
constexpr int kJpDepthWindow = 4096;
void launch_jp_depth_hist(const int32_t *dep, i64 nv, int32_t top, unsigned *hist, hipStream_t s);
// vertices deeper than rounds become uncoloured (the tail)
void launch_jp_depth_cut(const int32_t *dep, i64 nv, int32_t rounds, int32_t *colors, hipStream_t s);
void launch_color_conflicts(const DevGraph &g, const int32_t *colors, i64 *count, hipStream_t s);

// --- renumber / compose --------------------------------------------------
void launch_first_index(const i64 *C, i64 n, i64 *first, int32_t *err, hipStream_t s);
void launch_head_flags(const i64 *C, i64 n, const i64 *first, i64 *flags, hipStream_t s);
void launch_relabel(i64 *C, i64 n, const i64 *first, const i64 *pos, hipStream_t s);
void launch_compose(i64 *C_orig, i64 n0, const i64 *C, bool first_phase, hipStream_t s);

// --- coarsen -------------------------------------------------------------
// The pairs (C[v], C[tail]) with C[tail] <= C[v] (C[v] >= 0) in entry order,
// counted and written per chunk of consecutive entries: ccount[coarsen_chunks(nnz)]
// the pairs of every chunk, cbase their exclusive prefix.
i64 coarsen_chunks(i64 nnz);
void launch_coarsen_count(const DevGraph &g, const i64 *C, i64 *ccount, hipStream_t s);
// vf: no seeded self-loops and integer-truncated weights (buildNewGraphVF)
void launch_coarsen_emit(const DevGraph &g, const i64 *C, const i64 *cbase, i64 k, int kb,
                         u64 *keys, double *vals, hipStream_t s, bool vf = false);
// nself (nullable, zeroed): += number of self-loop keys
void launch_coarsen_mirror(const u64 *ukeys, const double *uvals, i64 nu, int kb, u64 *keys,
                           double *vals, hipStream_t s, i64 *nself = nullptr);
// vertexFollowing: C[i] and *count (zeroed) += vertices to remove
void launch_vertex_following(const DevGraph &g, i64 *C, i64 *count, hipStream_t s);
void launch_coarsen_rows(const u64 *keys, const double *vals, i64 total, i64 k, int kb, i64 *off,
                         i64 *tails, double *w, hipStream_t s);

// --- AoS <-> SoA ---------------------------------------------------------
void launch_split_edges(const void *aos, i64 nnz, i64 *tails, double *w, hipStream_t s);
void launch_check_labels(const i64 *off, const i64 *C, i64 nv, i64 limit, int32_t *err,
                         hipStream_t s);
void launch_check_graph(const i64 *off, const i64 *tails, i64 nv, i64 nnz, int32_t *err,
                        hipStream_t s);

// --- scans / sorts (rocPRIM) ---------------------------------------------
// exclusive prefix sum of in[0..n) into out[0..n); returns nothing, total = out[n-1]+in[n-1]
void scan_exclusive_i64(const i64 *in, i64 *out, i64 n, hipStream_t s);
void sort_pairs_u32_i64(uint32_t *keys, i64 *vals, i64 n, int end_bit, hipStream_t s);
void sort_pairs_u64_f64(u64 *keys, double *vals, i64 n, int end_bit, hipStream_t s);
// returns number of unique keys (synchronises)
i64 reduce_by_key_u64_f64(const u64 *keys, const double *vals, i64 n, u64 *ukeys, double *usums,
                          hipStream_t s);

}  // namespace glv
