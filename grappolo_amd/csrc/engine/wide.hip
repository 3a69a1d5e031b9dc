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

// partition of a community id among P (a hash independent of slot_of's, so
// the ids of one partition still spread over a table's slots)
__device__ inline int part_of(uint32_t key, int P) {
    uint32_t h = key * 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return (int)(((u64)h * (u64)P) >> 32);
}

// lds_insert that gives up (false) once every slot was probed
__device__ inline bool lds_insert_bounded(uint32_t *keys, double *vals, uint32_t T, uint32_t key, double w) {
    uint32_t h = slot_of(key, T);
    for (uint32_t probes = 0; probes < T; probes++) {
        uint32_t k = keys[h];
        if (k == kEmptyKey) {
            k = atomicCAS(&keys[h], kEmptyKey, key);
            if (k == kEmptyKey) k = key;
        }
        if (k == key) {
            atomicAdd(&vals[h], w);
            return true;
        }
        h = h + 1 == T ? 0 : h + 1;
    }
    return false;
}

// LDS writes of a wavefront are visible to its other lanes after this
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Reduction scratch of a workgroup (one slot per wavefront).
struct Red {
    double sl;  // self-loop weight of the vertex being gathered (one entry)
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

// Exclusive prefix sum over the workgroup (one value per thread); *total
// gets the sum.  Uses R.c.
__device__ inline int wg_exscan(int x, Red &R, int *total) {
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    int incl = x;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    __syncthreads();
    if (lane == kWave - 1) R.c[w] = incl;
    __syncthreads();
    int before = 0, all = 0;
    for (int q = 0; q < kWWaves; q++) {
        const int t = R.c[q];
        if (q < w) before += t;
        all += t;
    }
    *total = all;
    return before + incl - x;
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

// One pairing step of group_best over DPP control CTRL (or lane xor o when
// CTRL is 0): the other lane's candidate wins on a larger key (argmax_key)
// or an equal key and a smaller id; e and n follow the winner.
template <int CTRL>
__device__ inline void best_step(u64 &key, i64 &id, double &e, i64 &n, int o) {
    u64 ok, oe;
    i64 oid, on;
    if (CTRL) {
        ok = dpp64<CTRL>(key);
        oid = (i64)dpp64<CTRL>((u64)id);
        oe = dpp64<CTRL>((u64)__double_as_longlong(e));
        on = (i64)dpp64<CTRL>((u64)n);
    } else {
        ok = (u64)__shfl_xor((long long)key, o);
        oid = __shfl_xor((long long)id, o);
        oe = (u64)__shfl_xor(__double_as_longlong(e), o);
        on = __shfl_xor((long long)n, o);
    }
    const bool t = ok > key || (ok == key && oid < id);
    key = t ? ok : key;
    id = t ? oid : id;
    e = t ? __longlong_as_double((long long)oe) : e;
    n = t ? on : n;
}

// The best candidate over aligned groups of W lanes (W a power of two up to
// 64): every lane gets it, or keeps its own when the group has none.  The
// pairing is that of group_argmax_key (DPP within a row, bpermute across).
template <int W>
__device__ inline void group_best_key(u64 &key, i64 &id, double &e, i64 &n) {
    if (W >= 2) best_step<0xB1>(key, id, e, n, 0);   // quad_perm [1,0,3,2]
    if (W >= 4) best_step<0x4E>(key, id, e, n, 0);   // quad_perm [2,3,0,1]
    if (W >= 8) best_step<0x141>(key, id, e, n, 0);  // row_half_mirror
    if (W >= 16) best_step<0x140>(key, id, e, n, 0); // row_mirror
    if (W >= 32) best_step<0>(key, id, e, n, 16);
    if (W >= 64) best_step<0>(key, id, e, n, 32);
}
template <int W>
__device__ inline void group_best(Cand &c) {
    u64 key = argmax_key(c.ok, c.gain);
    i64 id = c.id, n = c.n;
    double e = c.e;
    group_best_key<W>(key, id, e, n);
    if (key != 0) {
        c.ok = true;
        c.gain = key_gain(key);
        c.id = id;
        c.e = e;
        c.n = n;
    }
}

// best over the workgroup, in thread 0 only (one barrier; R's slots are
// rewritten only after the caller's next barrier): the wavefronts' bests,
// then the first wavefront reduces the kWWaves of them
__device__ inline void wg_best0(Cand &c, Red &R) {
    group_best<kWave>(c);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) {
        R.c[w] = c.ok;
        R.d[w] = c.gain;
        R.i[w] = c.id;
        R.e[w] = c.e;
        R.n[w] = c.n;
    }
    __syncthreads();
    if (w != 0) return;
    const int q = l % kWWaves;
    u64 key = argmax_key(R.c[q] != 0, R.d[q]);
    i64 id = R.i[q], n = R.n[q];
    double e = R.e[q];
    group_best_key<kWWaves>(key, id, e, n);
    if (l != 0) return;
    Cand b;
    if (key != 0) {
        b.ok = true;
        b.gain = key_gain(key);
        b.id = id;
        b.e = e;
        b.n = n;
    }
    c = b;
}

// The state of one color step: read the state after the previous step,
// write the moves into the other buffer.
struct Step {
    const int32_t *C;
    const CommInfo *m;
    int32_t *Cq;
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
        S.Cq[v] = (int32_t)target;
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
// workgroup's LDS table of T slots (empty) and its self-loop weight into
// R.sl (zero); kWUnroll passes of loads are in flight at a time.
__device__ inline void wg_gather(const WideArgs &W, const Step &S, i64 v, i64 beg, i64 end, uint32_t *keys,
                                 double *vals, uint32_t T, Red &R) {
    for (i64 j0 = beg + threadIdx.x; j0 < end; j0 += (i64)kWUnroll * kWBlock) {
        i64 tt[kWUnroll], cl[kWUnroll];
        double wt[kWUnroll];
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const i64 j = j0 + (i64)u * kWBlock;
            tt[u] = j < end ? W.tails[j] : -1;
            wt[u] = j < end ? W.A.w[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) cl[u] = tt[u] >= 0 ? S.C[tt[u]] : 0;
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            if (tt[u] < 0) continue;
            if (tt[u] == v) atomicAdd(&R.sl, wt[u]);
            lds_insert(keys, vals, T, (uint32_t)cl[u], wt[u]);
        }
    }
    __syncthreads();
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
    if (threadIdx.x == 0) R.sl = 0.0;
    const CommInfo ms = S.m[s];
    const double as = ms.a;
    const i64 ns = ms.n;
    __syncthreads();
    wg_gather(W, S, v, beg, end, keys, vals, T, R);
    const double sl = R.sl;
    const double e_s = lds_lookup(keys, vals, T, (uint32_t)s);
    const double eix = e_s - sl, ax = as - kv;
    Cand c;
    score_lds(S, keys, vals, T, threadIdx.x, kWBlock, (uint32_t)s, kv, eix, ax, W.A.constant, c);
    wg_best0(c, R);
    if (threadIdx.x == 0) decide(W, S, v, s, kv, ns, c, eix, dein, moved);
    __syncthreads();  // the table is cleared by the next item
}

// Chunk j of giant g: its entries summed per community in LDS, then written
// out as (community, weight) pairs grouped by partition (bounds in gcnt), and
// its part of the self-loop weight and of the weight to the giant's own
// community added to gsc; the workgroup writing the last chunk marks the
// pairs ready.  pc: kWMaxParts + 1 ints of LDS.
__device__ void giant_chunk(const WideArgs &W, const Step &S, i64 g, int j, uint32_t *keys, double *vals, int *pc,
                            Red &R) {
    const WideGiant G = W.giants[g];
    const i64 v = G.v;
    const i64 s = S.C[v];
    const int P = G.nparts;
    const i64 beg0 = W.A.off[v], end0 = W.A.off[v + 1];
    const i64 beg = beg0 + (i64)j * kWChunk, end = std::min(beg + kWChunk, end0);
    const uint32_t T = (uint32_t)std::min<i64>(kWHeavySlots, (2 * (end - beg) + kWBlock - 1) / kWBlock * kWBlock);
    clear_table(keys, vals, T, threadIdx.x, kWBlock);
    for (int q = threadIdx.x; q <= P; q += kWBlock) pc[q] = 0;
    if (threadIdx.x == 0) R.sl = 0.0;
    __syncthreads();
    wg_gather(W, S, v, beg, end, keys, vals, T, R);
    const double sl = R.sl;
    const double es = lds_lookup(keys, vals, T, (uint32_t)s);
    for (uint32_t i = threadIdx.x; i < T; i += kWBlock) {
        const uint32_t k = keys[i];
        if (k != kEmptyKey) atomicAdd(&pc[part_of(k, P)], 1);
    }
    __syncthreads();
    int total = 0;
    const int ex = wg_exscan(threadIdx.x < (unsigned)P ? pc[threadIdx.x] : 0, R, &total);
    __syncthreads();  // every count was read
    int32_t *cnt = W.gcnt + G.cnt + (i64)j * (P + 1);
    if (threadIdx.x < (unsigned)P) {
        pc[threadIdx.x] = ex;
        cnt[threadIdx.x] = ex;
    }
    if (threadIdx.x == 0) cnt[P] = total;
    __syncthreads();
    uint32_t *pk = W.gpk + G.pairs + (i64)j * kWChunk;
    double *pv = W.gpv + G.pairs + (i64)j * kWChunk;
    for (uint32_t i = threadIdx.x; i < T; i += kWBlock) {
        const uint32_t k = keys[i];
        if (k == kEmptyKey) continue;
        const int at = atomicAdd(&pc[part_of(k, P)], 1);
        pk[at] = k;
        pv[at] = vals[i];
    }
    // every wavefront's pairs are written before thread 0 releases them (one
    // agent-scope fence for the workgroup)
    __syncthreads();
    if (threadIdx.x == 0) {
        if (sl != 0.0) atomicAdd(&W.gsc[2 * g], sl);
        if (es != 0.0) atomicAdd(&W.gsc[2 * g + 1], es);
        __threadfence();
        const int done = atomicAdd(&W.gstate[3 * g], 1) + 1;
        if (done == G.nchunks) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(&W.gstate[3 * g + 1], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
}

// Partition p of giant g: waits for the pairs, sums the partition's pairs of
// every chunk in the LDS table and scores them; the workgroup scoring the
// last partition merges the partitions' candidates in partition order and
// decides the giant's move.  pc: 2 * kWMaxChunks + 1 ints of LDS.  Returns
// false when the wait gave up or the partition outgrew the table (abort
// raised: bit 1, or bit 2 for the table).
__device__ bool giant_part(const WideArgs &W, const Step &S, i64 g, int p, uint32_t *keys, double *vals, int *pc,
                           Red &R, i64 &moved) {
    __shared__ int s_ok;
    const WideGiant G = W.giants[g];
    const i64 v = G.v;
    const i64 s = S.C[v];
    const double kv = W.A.k[v];
    const CommInfo ms = S.m[s];
    const int P = G.nparts, nc = G.nchunks;
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
    __syncthreads();  // thread 0's acquire covers the workgroup
    if (!s_ok) return false;
    const double sl = ld_agent(&W.gsc[2 * g]), e_s = ld_agent(&W.gsc[2 * g + 1]);
    const double eix = e_s - sl, ax = ms.a - kv;
    // this partition's segment of every chunk
    int lo = 0, len = 0;
    if (threadIdx.x < (unsigned)nc) {
        const int32_t *cnt = W.gcnt + G.cnt + (i64)threadIdx.x * (P + 1);
        lo = cnt[p];
        len = cnt[p + 1] - lo;
    }
    int total = 0;
    const int ex = wg_exscan(len, R, &total);
    int *pre = pc, *seg = pc + kWMaxChunks + 1;
    if (threadIdx.x < (unsigned)nc) {
        pre[threadIdx.x] = ex;
        seg[threadIdx.x] = lo;
    }
    if (threadIdx.x == 0) {
        pre[nc] = total;
        s_ok = 1;
    }
    const uint32_t T = (uint32_t)std::min<i64>(kWHeavySlots, ((i64)2 * total + kWBlock - 1) / kWBlock * kWBlock);
    clear_table(keys, vals, std::max<uint32_t>(T, kWBlock), threadIdx.x, kWBlock);
    __syncthreads();
    const uint32_t *pk = W.gpk + G.pairs;
    const double *pv = W.gpv + G.pairs;
    for (int i0 = threadIdx.x; i0 < total; i0 += kWUnroll * kWBlock) {
        uint32_t k[kWUnroll];
        double x[kWUnroll];
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
            const int i = i0 + u * kWBlock;
            k[u] = kEmptyKey;
            x[u] = 0.0;
            if (i < total) {
                int a = 0, b = nc;  // the chunk c with pre[c] <= i < pre[c + 1]
                while (b - a > 1) {
                    const int m = (a + b) / 2;
                    if (pre[m] <= i) a = m; else b = m;
                }
                const i64 at = (i64)a * kWChunk + seg[a] + (i - pre[a]);
                k[u] = pk[at];
                x[u] = pv[at];
            }
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++)
            if (k[u] != kEmptyKey && !lds_insert_bounded(keys, vals, std::max<uint32_t>(T, kWBlock), k[u], x[u]))
                s_ok = 0;
    }
    __syncthreads();
    if (!s_ok) {
        if (threadIdx.x == 0) atomicOr(&W.bar[0], 3u);
        return false;
    }
    Cand c;
    score_lds(S, keys, vals, std::max<uint32_t>(T, kWBlock), threadIdx.x, kWBlock, (uint32_t)s, kv, eix, ax,
              W.A.constant, c);
    wg_best0(c, R);
    if (threadIdx.x == 0) {
        const i64 at = g * W.gparts + p;
        W.gcid[at] = c.ok ? c.id : -1;
        W.gcand[3 * at] = c.gain;
        W.gcand[3 * at + 1] = c.e;
        W.gcand[3 * at + 2] = (double)c.n;
        __threadfence();
        const int done = atomicAdd(&W.gstate[3 * g + 2], 1) + 1;
        if (done == P) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            Cand b;
            for (int q = 0; q < P; q++) {
                const i64 bq = g * W.gparts + q;
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
            tt[u] = j < end ? W.tails[j] : -1;
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
        const i64 t = W.tails[beg + gl];
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

__global__ __launch_bounds__(kWBlock) void k_wide_iteration(WideArgs W) {
    __shared__ __align__(16) unsigned char smem[kWHeavyBytes];
    __shared__ Red R;
    __shared__ int pc[2 * kWMaxChunks + 2];  // giant items: partition counts or chunk segments
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
                giant_chunk(W, S, it.x, it.y, hkeys, hvals, pc, R);
            } else if (!giant_part(W, S, it.x, it.y, hkeys, hvals, pc, R, moved)) {
                alive = false;
                break;
            }
        }
        if (!alive) break;
        // wavefront items
        const i64 vlo = W.vstart[(i64)c * nb + blockIdx.x], vhi = W.vstart[(i64)c * nb + blockIdx.x + 1];
        for (i64 i = vlo + wave; i < vhi; i += kWWaves) {
            const WideItem it = W.vitems[i];
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
                S.Cq[v] = (int32_t)t;
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

__global__ void k_wide_clear_giants(int32_t *gstate, double *gsc, double *gout, i64 ng) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < ng; i += stride) {
        gstate[3 * i] = gstate[3 * i + 1] = gstate[3 * i + 2] = 0;
        gsc[2 * i] = gsc[2 * i + 1] = 0.0;
        gout[i] = 0.0;
    }
}

}  // namespace

#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

__global__ void k_narrow_ids(const i64 *src, i64 n, int32_t *dst) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        dst[i] = (int32_t)src[i];
}
__global__ void k_widen_ids(const int32_t *src, i64 n, i64 *dst) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        dst[i] = (i64)src[i];
}
void launch_narrow_ids(const i64 *src, i64 n, int32_t *dst, hipStream_t s) {
    if (n <= 0) return;
    k_narrow_ids<<<grid_for(n, kBlock, 16384), kBlock, 0, s>>>(src, n, dst);
    GLV_LAUNCH_CHECK();
}
void launch_widen_ids(const int32_t *src, i64 n, i64 *dst, hipStream_t s) {
    if (n <= 0) return;
    k_widen_ids<<<grid_for(n, kBlock, 16384), kBlock, 0, s>>>(src, n, dst);
    GLV_LAUNCH_CHECK();
}

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

void launch_wide_clear_giants(int32_t *gstate, double *gsc, double *gout, i64 ng, hipStream_t s) {
    if (ng <= 0) return;
    k_wide_clear_giants<<<grid_for(ng, kBlock, 8192), kBlock, 0, s>>>(gstate, gsc, gout, ng);
    GLV_LAUNCH_CHECK();
}

}  // namespace glv
