// wide.hip -- the whole-GPU persistent colored iteration (see wide.hpp).
//
// Per vertex, the decision is the reference's: the weights to the
// neighbouring communities (buildLocalMapCounter,
// Utility/utilityClusteringFunctions.cpp:172-194), the gain of each and the
// strictly positive maximum with ties to the smaller id and the singleton swap
// rule (max(), :226-259), against the state frozen at the start of the color
// step (the colored sweep, Coloring/parallelLouvainWithColoring.cpp:169-210),
// then the community degree and size updates of the move (:213-217).
#include "wide.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "sweep_device.hpp"

namespace glv {
namespace {

constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;
constexpr int kWWaves = kWBlock / kWave;
constexpr int kWUnroll = 4;  // loads in flight per thread in the gathers
constexpr size_t kWHeavyBytes = (size_t)kWHeavySlots * 12;
constexpr size_t kWWaveBytes = (size_t)kWMidSlots * 12;
static_assert(kWWaves * kWWaveBytes <= kWHeavyBytes, "wavefront tables exceed the LDS table");
static_assert(3 * 64 * 8 <= kWWaveBytes, "light-vertex staging exceeds a wavefront's LDS");

// slot of a community id in a table of T slots (multiplicative hash, then the
// high bits of the product with T)
__device__ inline uint32_t slot_of(uint32_t key, uint32_t T) {
    return (uint32_t)(((u64)(key * 0x9E3779B1u) * T) >> 32);
}

__device__ inline void lds_insert(uint32_t *keys, double *vals, uint32_t T, uint32_t key, double w) {
    uint32_t h = slot_of(key, T);
    while (true) {
        uint32_t k = keys[h];
        if (k == kEmptyKey) {
            k = atomicCAS(&keys[h], kEmptyKey, key);
            if (k == kEmptyKey) k = key;
        }
        if (k == key) {
            atomicAdd(&vals[h], w);
            return;
        }
        h = h + 1 == T ? 0 : h + 1;
    }
}

__device__ inline double lds_lookup(const uint32_t *keys, const double *vals, uint32_t T, uint32_t key) {
    uint32_t h = slot_of(key, T);
    while (true) {
        const uint32_t k = keys[h];
        if (k == key) return vals[h];
        if (k == kEmptyKey) return 0.0;
        h = h + 1 == T ? 0 : h + 1;
    }
}

// giant tables in HBM: power-of-two sizes, Fibonacci hashing
__device__ inline u64 gslot(uint32_t key, int lg) { return ((u64)key * kFib) >> (64 - lg); }

__device__ inline void hbm_insert(uint32_t *keys, double *vals, int lg, u64 mask, uint32_t key, double w) {
    u64 h = gslot(key, lg);
    while (true) {
        uint32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == kEmptyKey) {
            k = atomicCAS(&keys[h], kEmptyKey, key);
            if (k == kEmptyKey) k = key;
        }
        if (k == key) {
            atomicAdd(&vals[h], w);
            return;
        }
        h = (h + 1) & mask;
    }
}

__device__ inline double hbm_lookup(const uint32_t *keys, const double *vals, int lg, u64 mask, uint32_t key) {
    u64 h = gslot(key, lg);
    while (true) {
        const uint32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == key) return ld_agent(&vals[h]);
        if (k == kEmptyKey) return 0.0;
        h = (h + 1) & mask;
    }
}

// LDS writes of a wavefront are visible to its other lanes after this
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Reduction scratch of a workgroup (one slot per wavefront).
struct Red {
    double d[kWWaves];
    double e[kWWaves];
    i64 i[kWWaves];
    i64 n[kWWaves];
    int c[kWWaves];
};

// Workgroup sum in a fixed order (every thread gets it).
__device__ inline double wg_sum(double v, Red &R) {
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) R.d[w] = v;
    __syncthreads();
    double r = 0;
    for (int q = 0; q < kWWaves; q++) r += R.d[q];
    return r;
}

// The best move of a vertex so far: target community, its gain, the
// vertex's weight towards it and its size (for the singleton swap rule).
struct Cand {
    bool ok = false;
    double gain = 0;
    i64 id = INT64_MAX;
    double e = 0;
    i64 n = 0;
};

__device__ inline void offer(Cand &c, double g, i64 y, double e, i64 n) {
    if (better(g > 0, g, y, c.ok, c.gain, c.id)) {
        c.ok = true;
        c.gain = g;
        c.id = y;
        c.e = e;
        c.n = n;
    }
}

// best over a group of W lanes (every lane gets it)
template <int W>
__device__ inline void group_best(Cand &c) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        const double og = __shfl_xor(c.gain, o, W);
        const long long oid = __shfl_xor((long long)c.id, o, W);
        const double oe = __shfl_xor(c.e, o, W);
        const long long on = __shfl_xor((long long)c.n, o, W);
        const int ok = __shfl_xor((int)c.ok, o, W);
        if (better(ok != 0, og, oid, c.ok, c.gain, c.id)) {
            c.ok = true;
            c.gain = og;
            c.id = oid;
            c.e = oe;
            c.n = on;
        }
    }
}

// best over the workgroup (every thread gets it)
__device__ inline void wg_best(Cand &c, Red &R) {
    group_best<kWave>(c);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) {
        R.c[w] = c.ok;
        R.d[w] = c.gain;
        R.i[w] = c.id;
        R.e[w] = c.e;
        R.n[w] = c.n;
    }
    __syncthreads();
    Cand b;
    for (int q = 0; q < kWWaves; q++)
        if (better(R.c[q] != 0, R.d[q], R.i[q], b.ok, b.gain, b.id)) {
            b.ok = true;
            b.gain = R.d[q];
            b.id = R.i[q];
            b.e = R.e[q];
            b.n = R.n[q];
        }
    c = b;
}

// The state of one color step: read the state after the previous step,
// write the moves into the other buffer.
struct Step {
    const i64 *C;
    const CommInfo *m;
    i64 *Cq;
    CommInfo *mq;
};

__device__ inline void move_comm(i64 t, i64 s, double kv, CommInfo *m) {
    atomicAdd(&m[t].a, kv);
    atomicAdd(&m[s].a, -kv);
    atomicAdd((unsigned long long *)&m[t].n, 1ULL);
    atomicAdd((unsigned long long *)&m[s].n, ~0ULL);  // -1
}

// Move decision of vertex v (one thread): the best candidate, or stay; the
// swap rule of utilityClusteringFunctions.cpp:245-256 (two singletons move
// only towards the smaller id); the move goes into the buffer being written.
__device__ inline void decide(const WideArgs &W, const Step &S, i64 v, i64 s, double kv, i64 ns, const Cand &c,
                              double eix, double &dein, i64 &moved) {
    i64 target = s;
    if (c.ok) {
        target = c.id;
        if (ns == 1 && c.n == 1 && target > s) target = s;
    }
    W.A.tgt[v] = target;
    if (target != s) {
        move_comm(target, s, kv, S.mq);
        S.Cq[v] = target;
        dein += 2.0 * (c.e - eix);
        moved++;
    }
}

// Empties the first T slots of an LDS table (the caller synchronises).
__device__ inline void clear_table(uint32_t *keys, double *vals, uint32_t T, int first, int step) {
    for (uint32_t i = first; i < T; i += step) {
        keys[i] = kEmptyKey;
        vals[i] = 0.0;
    }
}

// Sums entries [beg, end) of vertex v per neighbouring community into the
// workgroup's LDS table of T slots (empty); returns this thread's part of the
// self-loop weight.  kWUnroll passes of loads are in flight at a time.
__device__ inline double wg_gather(const WideArgs &W, const Step &S, i64 v, i64 beg, i64 end, uint32_t *keys,
                                   double *vals, uint32_t T) {
    double sl = 0;
    for (i64 j0 = beg + threadIdx.x; j0 < end; j0 += (i64)kWUnroll * kWBlock) {
        i64 tt[kWUnroll], cl[kWUnroll];
        double wt[kWUnroll];
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const i64 j = j0 + (i64)u * kWBlock;
            tt[u] = j < end ? W.A.tails[j] : -1;
            wt[u] = j < end ? W.A.w[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) cl[u] = tt[u] >= 0 ? S.C[tt[u]] : 0;
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            if (tt[u] < 0) continue;
            if (tt[u] == v) sl += wt[u];
            lds_insert(keys, vals, T, (uint32_t)cl[u], wt[u]);
        }
    }
    __syncthreads();
    return sl;
}

// Scores slots [first, T) step `step` of an LDS table for vertex v: the
// degree and size of every candidate community are loaded together.
__device__ inline void score_lds(const Step &S, const uint32_t *keys, const double *vals, uint32_t T, int first,
                                 int step, uint32_t s, double kv, double eix, double ax, double cst, Cand &c) {
    for (int i0 = first; i0 < (int)T; i0 += kWUnroll * step) {
        uint32_t yk[kWUnroll];
        double ay[kWUnroll];
        i64 ny[kWUnroll];
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const int i = i0 + u * step;
            yk[u] = i < (int)T ? keys[i] : kEmptyKey;
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const bool y = yk[u] != kEmptyKey && yk[u] != s;
            const CommInfo m = y ? S.m[yk[u]] : CommInfo{0.0, 0};
            ay[u] = m.a;
            ny[u] = m.n;
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            if (yk[u] == kEmptyKey || yk[u] == s) continue;
            const double e = vals[i0 + u * step];
            offer(c, gain_of(e, eix, kv, ay[u], ax, cst), (i64)yk[u], e, ny[u]);
        }
    }
}

// A vertex of kWHeavy (the whole workgroup, LDS table).  Its row, label and
// degree come first; its community's degree and size are in flight during
// the gather.
__device__ void heavy_vertex(const WideArgs &W, const Step &S, i64 v, uint32_t *keys, double *vals, Red &R,
                             double &dein, i64 &moved) {
    const i64 beg = W.A.off[v], end = W.A.off[v + 1];
    const i64 s = S.C[v];
    const double kv = W.A.k[v];
    const uint32_t T = (uint32_t)std::min<i64>(kWHeavySlots, (2 * (end - beg) + kWBlock - 1) / kWBlock * kWBlock);
    clear_table(keys, vals, T, threadIdx.x, kWBlock);
    const CommInfo ms = S.m[s];
    const double as = ms.a;
    const i64 ns = ms.n;
    __syncthreads();
    double sl = wg_gather(W, S, v, beg, end, keys, vals, T);
    sl = wg_sum(sl, R);
    const double e_s = lds_lookup(keys, vals, T, (uint32_t)s);
    const double eix = e_s - sl, ax = as - kv;
    Cand c;
    score_lds(S, keys, vals, T, threadIdx.x, kWBlock, (uint32_t)s, kv, eix, ax, W.A.constant, c);
    wg_best(c, R);
    if (threadIdx.x == 0) decide(W, S, v, s, kv, ns, c, eix, dein, moved);
    __syncthreads();  // the table is cleared by the next item
}

// Chunk j of giant g: its entries summed per community in LDS, then added to
// the giant's HBM table; the workgroup adding the last chunk looks up the
// weight to the giant's own community and marks the table ready.
__device__ void giant_chunk(const WideArgs &W, const Step &S, i64 g, int j, uint32_t *keys, double *vals,
                            Red &R) {
    const WideGiant G = W.giants[g];
    const i64 v = G.v;
    const i64 beg0 = W.A.off[v], end0 = W.A.off[v + 1];
    const i64 beg = beg0 + (i64)j * kWChunk, end = std::min(beg + kWChunk, end0);
    const uint32_t T = (uint32_t)std::min<i64>(kWHeavySlots, (2 * (end - beg) + kWBlock - 1) / kWBlock * kWBlock);
    clear_table(keys, vals, T, threadIdx.x, kWBlock);
    __syncthreads();
    double sl = wg_gather(W, S, v, beg, end, keys, vals, T);
    sl = wg_sum(sl, R);
    int lg = 0;
    while (((i64)1 << lg) < G.slots) lg++;
    const u64 mask = (u64)G.slots - 1;
    uint32_t *gk = W.gkeys + G.tab;
    double *gv = W.gvals + G.tab;
    // every thread first claims the home slots of its kWUnroll keys at once
    // (most land there at load factor <= 1/2), then probes on for the rest
    for (uint32_t i0 = threadIdx.x; i0 < T; i0 += kWUnroll * kWBlock) {
        uint32_t k[kWUnroll], old[kWUnroll];
        double x[kWUnroll];
        u64 h[kWUnroll];
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const uint32_t i = i0 + u * kWBlock;
            k[u] = i < T ? keys[i] : kEmptyKey;
            x[u] = i < T ? vals[i] : 0.0;
            h[u] = gslot(k[u], lg);
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) old[u] = k[u] != kEmptyKey ? atomicCAS(&gk[h[u]], kEmptyKey, k[u]) : kEmptyKey;
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            if (k[u] == kEmptyKey) continue;
            if (old[u] == kEmptyKey || old[u] == k[u])
                atomicAdd(&gv[h[u]], x[u]);
            else
                hbm_insert(gk, gv, lg, mask, k[u], x[u]);
        }
    }
    if (threadIdx.x == 0 && sl != 0.0) atomicAdd(&W.gsc[2 * g], sl);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const int done = atomicAdd(&W.gstate[3 * g], 1) + 1;
        if (done == G.nchunks) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const double e_s = hbm_lookup(gk, gv, lg, mask, (uint32_t)S.C[v]);
            __hip_atomic_store(&W.gsc[2 * g + 1], e_s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&W.gstate[3 * g + 1], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
}

// Slice j of giant g's table: waits for the table, scores and clears its
// slots; the workgroup scoring the last slice merges the slices' candidates in
// slice order and decides the giant's move.  Returns false when the wait gave
// up (abort raised).
__device__ bool giant_score(const WideArgs &W, const Step &S, i64 g, int j, Red &R, i64 &moved) {
    __shared__ int s_ok;
    const WideGiant G = W.giants[g];
    const i64 v = G.v;
    const i64 s = S.C[v];
    const double kv = W.A.k[v];
    const CommInfo ms = S.m[s];
    const double as = ms.a;
    if (threadIdx.x == 0) {
        int ok = 1;
        unsigned spins = 0;
        while (__hip_atomic_load(&W.gstate[3 * g + 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            if (__hip_atomic_load(&W.bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > kSpinLimit) {
                atomicOr(&W.bar[0], 1u);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        s_ok = ok;
    }
    __syncthreads();
    if (!s_ok) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const double sl = ld_agent(&W.gsc[2 * g]), e_s = ld_agent(&W.gsc[2 * g + 1]);
    const double eix = e_s - sl, ax = as - kv;
    uint32_t *gk = W.gkeys + G.tab;
    double *gv = W.gvals + G.tab;
    const i64 lo = (i64)j * kWScoreSlice, hi = std::min(lo + kWScoreSlice, G.slots);
    Cand c;
    for (i64 i0 = lo + threadIdx.x; i0 < hi; i0 += (i64)kWUnroll * kWBlock) {
        uint32_t yk[kWUnroll];
        double e[kWUnroll], ay[kWUnroll];
        i64 ny[kWUnroll];
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const i64 i = i0 + (i64)u * kWBlock;
            yk[u] = i < hi ? __hip_atomic_load(&gk[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kEmptyKey;
            e[u] = yk[u] != kEmptyKey ? ld_agent(&gv[i]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const bool y = yk[u] != kEmptyKey && yk[u] != (uint32_t)s;
            const CommInfo m = y ? S.m[yk[u]] : CommInfo{0.0, 0};
            ay[u] = m.a;
            ny[u] = m.n;
            if (yk[u] != kEmptyKey) {  // empty between uses
                const i64 i = i0 + (i64)u * kWBlock;
                __hip_atomic_store(&gk[i], kEmptyKey, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&gv[i], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            if (yk[u] == kEmptyKey || yk[u] == (uint32_t)s) continue;
            offer(c, gain_of(e[u], eix, kv, ay[u], ax, W.A.constant), (i64)yk[u], e[u], ny[u]);
        }
    }
    wg_best(c, R);
    if (threadIdx.x == 0) {
        const i64 at = g * W.gslices + j;
        W.gcid[at] = c.ok ? c.id : -1;
        W.gcand[3 * at] = c.gain;
        W.gcand[3 * at + 1] = c.e;
        W.gcand[3 * at + 2] = (double)c.n;
        __threadfence();
        const int done = atomicAdd(&W.gstate[3 * g + 2], 1) + 1;
        if (done == G.nslices) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            Cand b;
            for (int q = 0; q < G.nslices; q++) {
                const i64 bq = g * W.gslices + q;
                const i64 iq = ld_agent(&W.gcid[bq]);
                if (iq < 0) continue;
                const double gq = ld_agent(&W.gcand[3 * bq]);
                if (better(true, gq, iq, b.ok, b.gain, b.id)) {
                    b.ok = true;
                    b.gain = gq;
                    b.id = iq;
                    b.e = ld_agent(&W.gcand[3 * bq + 1]);
                    b.n = (i64)ld_agent(&W.gcand[3 * bq + 2]);
                }
            }
            double dg = 0;
            decide(W, S, v, s, kv, ms.n, b, eix, dg, moved);
            W.gout[g] = dg;
            // ready for the next iteration
            __hip_atomic_store(&W.gstate[3 * g], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&W.gstate[3 * g + 1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&W.gstate[3 * g + 2], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&W.gsc[2 * g], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&W.gsc[2 * g + 1], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    return true;
}

// A mid vertex (65..256 entries): one wavefront, its own LDS table.
__device__ void mid_vertex(const WideArgs &W, const Step &S, i64 v, uint32_t *keys, double *vals, int lane,
                           double &dein, i64 &moved) {
    const i64 beg = W.A.off[v], end = W.A.off[v + 1];
    const i64 s = S.C[v];
    const double kv = W.A.k[v];
    uint32_t T = 64;
    while ((i64)T < 2 * (end - beg)) T <<= 1;
    clear_table(keys, vals, T, lane, kWave);
    const CommInfo ms = S.m[s];
    const double as = ms.a;
    const i64 ns = ms.n;
    wave_lds_sync();
    double sl = 0;
    for (i64 j0 = beg + lane; j0 < end; j0 += (i64)kWUnroll * kWave) {
        i64 tt[kWUnroll], cl[kWUnroll];
        double wt[kWUnroll];
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const i64 j = j0 + (i64)u * kWave;
            tt[u] = j < end ? W.A.tails[j] : -1;
            wt[u] = j < end ? W.A.w[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) cl[u] = tt[u] >= 0 ? S.C[tt[u]] : 0;
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            if (tt[u] < 0) continue;
            if (tt[u] == v) sl += wt[u];
            lds_insert(keys, vals, T, (uint32_t)cl[u], wt[u]);
        }
    }
    wave_lds_sync();
    sl = wave_sum(sl);
    const double e_s = lds_lookup(keys, vals, T, (uint32_t)s);
    const double eix = e_s - sl, ax = as - kv;
    Cand c;
    score_lds(S, keys, vals, T, lane, kWave, (uint32_t)s, kv, eix, ax, W.A.constant, c);
    group_best<kWave>(c);
    if (lane == 0) decide(W, S, v, s, kv, ns, c, eix, dein, moved);
    wave_lds_sync();  // every lane is done with the table before the next item clears it
}

// 64 / G light vertices perm[pos .. pos + count): G lanes each, one entry per
// lane; every lane walks its group's entries in adjacency order, so the
// weight to a community is the left-to-right sum buildLocalMapCounter forms.
// Community ids travel as 32 bits here (every id is below 2^31).
template <int G>
__device__ void light_vertices(const WideArgs &W, const Step &S, i64 pos, int count, unsigned char *wbase, int lane,
                               double &dein, i64 &moved) {
    int32_t *s_c = reinterpret_cast<int32_t *>(wbase);
    double *s_w = reinterpret_cast<double *>(wbase + 64 * 8);
    const int grp = lane / G, gl = lane % G;
    const bool valid = grp < count;
    i64 v = 0, beg = 0;
    int deg = 0;
    int32_t s = -3;
    double kv = 0;
    if (valid) {
        v = W.perm[pos + grp];
        beg = W.A.off[v];
        deg = (int)(W.A.off[v + 1] - beg);
        s = (int32_t)S.C[v];
        kv = W.A.k[v];
    }
    int32_t c = -2;
    double ww = 0;
    bool self = false;
    double as = 0;
    i64 ns = 0;
    if (gl < deg) {
        const i64 t = W.A.tails[beg + gl];
        ww = W.A.w[beg + gl];
        if (gl == 0) {
            const CommInfo ms = S.m[s];
            as = ms.a;
            ns = ms.n;
        }
        c = (int32_t)S.C[t];
        self = t == v;
    }
    // the self-loop's slot carries -3 - c (never a label) so one array holds both
    s_c[lane] = self ? -3 - c : c;
    s_w[lane] = ww;
    wave_lds_sync();
    double e_mine = 0, e_s = 0, sl = 0;
    bool first = true;
    const int gb = grp * G;
    for (int m = 0; m < deg; m++) {
        int32_t cm = s_c[gb + m];
        const double wm = s_w[gb + m];
        if (cm <= -3) {
            cm = -3 - cm;
            sl += wm;
        }
        if (cm == c) {
            e_mine += wm;
            if (m < gl) first = false;
        }
        if (cm == s) e_s += wm;
    }
    as = __shfl(as, gb, kWave);
    Cand b;
    if (gl < deg && first && c != s) {
        const CommInfo mc = S.m[c];
        // max() starts from maxGain = 0 and needs a strict increase
        offer(b, gain_of(e_mine, e_s - sl, kv, mc.a, as - kv, W.A.constant), c, e_mine, mc.n);
    }
    group_best<G>(b);
    if (valid && gl == 0) decide(W, S, v, s, kv, ns, b, e_s - sl, dein, moved);
    wave_lds_sync();  // the staging arrays are rewritten by the next item
}

// position of item round r for worker w of nw in the snake order
__device__ inline i64 snake(i64 r, i64 w, i64 nw) { return r * nw + ((r & 1) ? nw - 1 - w : w); }

__global__ __launch_bounds__(kWBlock) void k_wide_iteration(WideArgs W) {
    __shared__ __align__(16) unsigned char smem[kWHeavyBytes];
    __shared__ Red R;
    uint32_t *hkeys = reinterpret_cast<uint32_t *>(smem);
    double *hvals = reinterpret_cast<double *>(smem + (size_t)kWHeavySlots * 4);
    const unsigned nb = gridDim.x;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    unsigned char *wbase = smem + (size_t)wave * kWWaveBytes;
    uint32_t *wkeys = reinterpret_cast<uint32_t *>(wbase);
    double *wvals = reinterpret_cast<double *>(wbase + (size_t)kWMidSlots * 4);
    const i64 nv = W.A.nv;
    const i64 gtid = (i64)blockIdx.x * kWBlock + threadIdx.x, gstride = (i64)nb * kWBlock;
    // the second buffer starts equal to the first
    for (i64 i = gtid; i < nv; i += gstride) {
        W.C1[i] = W.C0[i];
        const CommInfo m{W.a[i], W.n[i]};
        W.m0[i] = m;
        W.m1[i] = m;
    }
    unsigned seq = 0;
    bool alive = grid_barrier(W.bar, nb, ++seq);
    double dein = 0;  // this thread's movers' 2 (e_t - e_s)
    i64 moved = 0;
    int par = 0;
    i64 plo = 0, phi = 0;  // the previous step's class in cperm
    // wavefront items in the snake order over the wavefronts
    const i64 nw = (i64)nb * kWWaves;
    const i64 gw = (i64)blockIdx.x * kWWaves + wave;
    for (int c = 0; c < W.ncol && alive; c++) {
        const i64 lo = W.cstart[c], hi = W.cstart[c + 1];
        if (hi == lo) continue;
        Step S;
        S.C = par ? W.C1 : W.C0;
        S.m = par ? W.m1 : W.m0;
        S.Cq = par ? W.C0 : W.C1;
        S.mq = par ? W.m0 : W.m1;
        // workgroup items
        const i64 wlo = W.wstart[(i64)c * nb + blockIdx.x], whi = W.wstart[(i64)c * nb + blockIdx.x + 1];
        for (i64 i = wlo; i < whi; i++) {
            const WideItem it = W.witems[i];
            if (it.kind == kWHeavy) {
                heavy_vertex(W, S, it.x, hkeys, hvals, R, dein, moved);
            } else if (it.kind == kWGiantChunk) {
                giant_chunk(W, S, it.x, it.y, hkeys, hvals, R);
            } else if (!giant_score(W, S, it.x, it.y, R, moved)) {
                alive = false;
                break;
            }
        }
        if (!alive) break;
        // wavefront items
        const i64 vlo = W.vstart[c], vn = W.vstart[c + 1] - vlo;
        for (i64 r = 0; r * nw < vn; r++) {
            const i64 i = snake(r, gw, nw);
            if (i >= vn) continue;
            const WideItem it = W.vitems[vlo + i];
            switch (it.kind) {
                case kWMid: mid_vertex(W, S, W.perm[it.x], wkeys, wvals, lane, dein, moved); break;
                case kWLight1: light_vertices<4>(W, S, it.x, it.y, wbase, lane, dein, moved); break;
                case kWLight1 + 1: light_vertices<8>(W, S, it.x, it.y, wbase, lane, dein, moved); break;
                case kWLight1 + 2: light_vertices<16>(W, S, it.x, it.y, wbase, lane, dein, moved); break;
                case kWLight1 + 3: light_vertices<32>(W, S, it.x, it.y, wbase, lane, dein, moved); break;
                default: light_vertices<64>(W, S, it.x, it.y, wbase, lane, dein, moved); break;
            }
        }
        // the previous step's moves, into the buffer this step writes
        for (i64 q = plo + gtid; q < phi; q += gstride) {
            const i64 v = W.cperm[q];
            const i64 t = W.A.tgt[v], s = S.Cq[v];
            if (s >= 0 && t != s) {  // vertices without entries (label -1) never move
                move_comm(t, s, W.A.k[v], S.mq);
                S.Cq[v] = t;
            }
        }
        alive = grid_barrier(W.bar, nb, ++seq);
        par ^= 1;
        plo = lo;
        phi = hi;
    }
    const CommInfo *mf = par ? W.m1 : W.m0;
    // modularity partials: the movers' change of the internal weight (giants
    // in giant order) and the sum of a^2
    double exx = 0, a2 = 0;
    if (alive) {
        exx = dein;
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (i64 g = 0; g < W.ngiants; g++) exx += W.gout[g];
        for (i64 i = gtid; i < nv; i += gstride) {
            const CommInfo m = mf[i];
            a2 += m.a * m.a;
            W.a[i] = m.a;
            W.n[i] = m.n;
        }
    }
    exx = wg_sum(exx, R);
    a2 = wg_sum(a2, R);
    // moves of the iteration (one atomic per workgroup)
    __shared__ unsigned long long s_moved;
    if (threadIdx.x == 0) s_moved = 0;
    __syncthreads();
    if (moved) atomicAdd(&s_moved, (unsigned long long)moved);
    __syncthreads();
    if (threadIdx.x == 0) {
        W.part[blockIdx.x] = exx;
        W.part[nb + blockIdx.x] = a2;
        if (s_moved) atomicAdd((unsigned long long *)W.moves, s_moved);
    }
    if (alive && par)  // the final labels sit in the second buffer
        for (i64 i = gtid; i < nv; i += gstride) W.C0[i] = W.C1[i];
}

__global__ void k_wide_degrees(const i64 *perm, const i64 *off, i64 lo, i64 hi, i64 *vd) {
    for (i64 q = lo + blockIdx.x * (i64)blockDim.x + threadIdx.x; q < hi; q += (i64)gridDim.x * blockDim.x) {
        const i64 v = perm[q];
        vd[2 * (q - lo)] = v;
        vd[2 * (q - lo) + 1] = off[v + 1] - off[v];
    }
}

__global__ void k_wide_clear_giants(uint32_t *gkeys, double *gvals, i64 slots, int32_t *gstate, double *gsc,
                                    double *gout, i64 ng) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < slots; i += stride) {
        gkeys[i] = kEmptyKey;
        gvals[i] = 0.0;
    }
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < ng; i += stride) {
        gstate[3 * i] = gstate[3 * i + 1] = gstate[3 * i + 2] = 0;
        gsc[2 * i] = gsc[2 * i + 1] = 0.0;
        gout[i] = 0.0;
    }
}

}  // namespace

#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

bool wide_enabled() {
    static const bool on = [] {
        const char *e = std::getenv("GLV_WIDE");
        return !(e && std::strcmp(e, "0") == 0);
    }();
    return on;
}

int wide_blocks() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0, cus = 0, per = 0;
        GLV_HIP(hipGetDevice(&dev));
        GLV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        GLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_wide_iteration, kWBlock, 0));
        cached = per > 0 ? cus * std::min(per, 1) : 0;
    }
    return cached;
}

bool launch_wide_iteration(const WideArgs &W, int nblocks, hipStream_t s) {
    if (nblocks <= 0 || nblocks > wide_blocks()) return false;
    WideArgs w = W;
    void *args[] = {&w};
    hipError_t e = hipLaunchKernel((const void *)k_wide_iteration, dim3(nblocks), dim3(kWBlock), args, 0, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trace("wide colored iteration (%d workgroups) refused: %s", nblocks, hipGetErrorString(e));
        return false;
    }
    return true;
}

void launch_wide_degrees(const i64 *perm, const i64 *off, i64 lo, i64 hi, i64 *vd, hipStream_t s) {
    if (hi <= lo) return;
    k_wide_degrees<<<grid_for(hi - lo, kBlock, 4096), kBlock, 0, s>>>(perm, off, lo, hi, vd);
    GLV_LAUNCH_CHECK();
}

void launch_wide_clear_giants(uint32_t *gkeys, double *gvals, i64 slots, int32_t *gstate, double *gsc,
                              double *gout, i64 ng, hipStream_t s) {
    const i64 n = std::max(slots, ng);
    if (n <= 0) return;
    k_wide_clear_giants<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(gkeys, gvals, slots, gstate, gsc, gout, ng);
    GLV_LAUNCH_CHECK();
}

}  // namespace glv
