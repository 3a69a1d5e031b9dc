// active_set.hpp -- launchers of the learned active-set kernels: per-vertex
// inference of the MLP on the engine's state, sweep-trace capture for
// training, and compaction of the sweep lists to the active vertices.
#pragma once

#include "common.hpp"
#include "mlp_device.hpp"

namespace glv {

// active[v] = 1 when v has a community (C[v] >= 0) and either has no recorded
// features yet or the MLP logit of its features is >= logit_tau.
// counts[0] += active vertices, counts[1] += their CSR entries.
void launch_las_infer(const LasView &S, const float *P, float logit_tau, uint8_t *active, i64 *counts,
                      hipStream_t s);
// stale[v] = min(stale[v] + 1, 255) for every vertex
void launch_las_tick(uint8_t *stale, i64 nv, hipStream_t s);
// Appends the features of the sampled vertices (splitmix64(seed ^ v) below
// rate * 2^64, recorded features, a community) to the trace: feats[pos][8]
// BF16 and verts[pos]; *count advances by the sampled number, positions at
// or beyond cap are not written.
void launch_las_capture(const LasView &S, u64 seed, double rate, uint16_t *feats, i64 *verts,
                        i64 *count, i64 cap, hipStream_t s);
// labels[j] = (stale[verts[j]] == 0) for j in [lo, hi): the vertex moved in the sweep just committed
void launch_las_label(const i64 *verts, i64 lo, i64 hi, const uint8_t *stale, uint8_t *labels,
                      hipStream_t s);
// MLP input rows of the vertices verts[0..n) (for tests and offline use)
void launch_las_features(const LasView &S, const i64 *verts, i64 n, uint16_t *feats, hipStream_t s);

// Keeps the active entries of list[0..n) in order: out[0..total) and, for a
// list grouped into slices start[0..nkeys] (start[nkeys] = n), the slice
// starts of the kept entries in out_start[0..nkeys].  flags, pos: n scratch.
void launch_compact_active(const i64 *list, i64 n, const uint8_t *active, const i64 *start, i64 nkeys,
                           i64 *flags, i64 *pos, i64 *out, i64 *out_start, hipStream_t s);
// the same from flags[0..n) already set (0/1); pos: n scratch
void compact_flagged(const i64 *list, i64 n, i64 *flags, i64 *pos, const i64 *start, i64 nkeys, i64 *out,
                     i64 *out_start, hipStream_t s);

}  // namespace glv
