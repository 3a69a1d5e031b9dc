// kernels.hip -- the HIP kernels of the engine, written for gfx950 (CDNA4,
// 64-lane wavefronts, 160 KB LDS per CU, FP64 atomics in LDS and HBM).
//
// Reference correspondence (file:line of the CPU routine each one replaces):
//   k_degree*           sumVertexDegree       Utility/utilityClusteringFunctions.cpp:70-87
//   k_sweep_*           buildLocalMapCounter + max (:172-194, :226-259) inside the per-color
//                       loop of Coloring/parallelLouvainWithColoring.cpp:169-210 and the Jacobi
//                       loop of BasicCommunitiesDetection/parallelLouvainMethod.cpp:199-240
//   k_apply_*           cInfo += cUpdate      parallelLouvainWithColoring.cpp:213-217
//   k_mod_*             in-sweep modularity   parallelLouvainWithColoring.cpp:226-249
//   k_jp_*              distance-1 coloring   Coloring/coloringDistanceOne.cpp:50-233 (JP rounds)
//   k_first/flags/relab renumberClustersContiguously Utility/buildNextPhase.cpp:48-79
//   k_coarsen_*         buildNextLevelGraphOpt Utility/buildNextPhase.cpp:83-255
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "kernels.hpp"
#include "mlp_device.hpp"
#include "sweep_device.hpp"

namespace glv {

namespace {

__device__ inline u64 jp_priority(i64 v) {
    // splitmix64 finaliser: a bijection, so priorities of distinct vertices differ
    u64 z = (u64)v + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Jones-Plassmann order: prio[] (larger degree first, then a hash of the id),
// equal keys broken by the larger id, so the order is total.  The count's
// byte for entry (v, u) compares u's priority pu with v's pv; the degree
// class of u, and whether u is a core vertex (degree dcore or more), come
// from the degree field of its priority.
__device__ inline uint8_t jp_code(u64 pu, i64 u, u64 pv, i64 v, u64 dcore) {
    if (u == v) return kJpSelf;
    const u64 du = pu >> 40;
    uint8_t c = (pu < pv || (pu == pv && u < v)) ? kJpRelease : 0;
    if (du > (u64)kBinMaxDeg[6]) c |= kJpLarge;
    if (du > (u64)kJpMidDeg) c |= kJpHuge;
    if (du > (u64)kGiantDeg) c |= kJpGiant;
    if (du >= dcore) c |= kJpCore;
    return c;
}

// Block-wide sum, result valid in every thread.  Deterministic order.
__device__ double block_sum(double v) {
    __shared__ double sh[kBlock / kWave];
    __shared__ double res;
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int q = 0; q < kBlock / kWave; q++) r += sh[q];
        res = r;
    }
    __syncthreads();
    double r = res;
    __syncthreads();
    return r;
}

// Block-wide argmax, result valid in every thread (a block without a
// candidate leaves each thread's gain and id as they were).
__device__ void block_argmax(bool &cand, double &gain, i64 &id) {
    __shared__ u64 sk[kBlock / kWave];
    __shared__ i64 si[kBlock / kWave];
    u64 key = argmax_key(cand, gain);
    i64 rid = id, none = 0;
    group_argmax_key<kWave>(key, rid, none);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) {
        sk[w] = key;
        si[w] = rid;
    }
    __syncthreads();
    // every wavefront reduces the wavefronts' results (lane l takes wavefront l % kBlock/kWave)
    key = sk[l % (kBlock / kWave)];
    rid = si[l % (kBlock / kWave)];
    group_argmax_key<kBlock / kWave>(key, rid, none);
    cand = key != 0;
    if (cand) {
        gain = key_gain(key);
        id = rid;
    }
    __syncthreads();
}

// Block-wide top-two of per-thread (b1, b2), result valid in every thread.
template <int BS>
__device__ void block_top2(float &b1, float &b2) {
    __shared__ float s1[BS / kWave], s2[BS / kWave];
    top2_group<kWave>(b1, b2);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) {
        s1[w] = b1;
        s2[w] = b2;
    }
    __syncthreads();
    float x1 = s1[0], x2 = s2[0];
    for (int q = 1; q < BS / kWave; q++) top2_merge(x1, x2, s1[q], s2[q]);
    b1 = x1;
    b2 = x2;
    __syncthreads();
}

// Exclusive prefix sum over a workgroup of BS threads; *total gets the sum.
template <int BS>
__device__ inline int block_exscan(int x, int *wsum, int *total) {
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    int incl = x;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    __syncthreads();
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int before = 0, all = 0;
    for (int q = 0; q < BS / kWave; q++) {
        const int t = wsum[q];
        if (q < w) before += t;
        all += t;
    }
    *total = all;
    return before + incl - x;
}

// Lists the set bits of touched[0..nwords) (community ids) into list[], in
// increasing id; returns the count.  One word per thread per pass.
template <int BS>
__device__ inline int compact_touched(const uint32_t *touched, int nwords, uint16_t *list, int *wsum) {
    int n = 0;
    for (int base = 0; base < nwords; base += BS) {
        const int wi = base + threadIdx.x;
        uint32_t m = wi < nwords ? touched[wi] : 0u;
        int total = 0;
        int at = n + block_exscan<BS>(__popc(m), wsum, &total);
        while (m) {
            const int b = __ffs((int)m) - 1;
            m &= m - 1;
            list[at++] = (uint16_t)(wi * 32 + b);
        }
        n += total;
    }
    __syncthreads();
    return n;
}

// ---------------------------------------------------------------- elementwise
__global__ void k_fill_i64(i64 *p, i64 n, i64 v) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_f64(double *p, i64 n, double v) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_i32(int32_t *p, i64 n, int32_t v) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_iota(i64 *p, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        p[i] = i;
}
// initCommAss (utilityClusteringFunctions.cpp:98-104); vertices without
// entries are labelled -1 at once, which is what their first color step
// writes (parallelLouvainWithColoring.cpp:189-191,208).
__global__ void k_init_labels(const i64 *off, i64 nv, i64 *C) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nv; i += (i64)gridDim.x * blockDim.x)
        C[i] = (off[i + 1] != off[i]) ? i : -1;
}

__global__ void k_bin_keys(const i64 *off, i64 nv, const int32_t *colors, int ncol, uint32_t *keys,
                           i64 *vid, int32_t *err, BinConfig bins) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nv; i += (i64)gridDim.x * blockDim.x) {
        const int b = degree_bin_of(off[i + 1] - off[i], bins);
        const int c = colors ? colors[i] : 0;
        vid[i] = i;
        if (c < 0 || c >= ncol) {
            atomicOr(err, 1);
            keys[i] = 0;
            continue;
        }
        keys[i] = (uint32_t)(b * ncol + c);
    }
}
// start[k] = first position of key k in the sorted keys (n when absent past the end)
__global__ void k_slice_starts(const uint32_t *keys, i64 n, i64 nkeys, i64 *start) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const i64 k = keys[i];
        const i64 prev = i == 0 ? -1 : (i64)keys[i - 1];
        for (i64 kk = prev + 1; kk <= k; kk++) start[kk] = i;
        if (i == n - 1)
            for (i64 kk = k + 1; kk <= nkeys; kk++) start[kk] = n;
    }
}
__global__ void k_perm_degree(const i64 *perm, const i64 *off, i64 n, i64 *degp) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const i64 v = perm[i];
        degp[i] = off[v + 1] - off[v];
    }
}
__global__ void k_slice_nnz(const i64 *start, const i64 *dscan, const i64 *degp, i64 n, i64 nkeys,
                            i64 *pnnz) {
    const i64 total = dscan[n - 1] + degp[n - 1];
    for (i64 k = blockIdx.x * (i64)blockDim.x + threadIdx.x; k < nkeys; k += (i64)gridDim.x * blockDim.x) {
        const i64 a = start[k], b = start[k + 1];
        pnnz[k] = (b < n ? dscan[b] : total) - (a < n ? dscan[a] : total);
    }
}

// sumVertexDegree: one thread per row sums in row order (the reference's
// order) for rows up to 2048 entries; longer rows use a workgroup each.
// k[v] = its weights summed in adjacency order (vertices of at most 2048
// entries; the rest go to k_degree_big), one thread per row: for graphs of
// high mean degree.
__global__ void k_degree_rows_seq(const i64 *off, const double *w, i64 nv, double *k) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nv; i += (i64)gridDim.x * blockDim.x) {
        const i64 b = off[i], e = off[i + 1];
        if (e - b > kBinMaxDeg[7]) continue;
        double s = 0;
        for (i64 j = b; j < e; j++) s += w[j];
        k[i] = s;
    }
}
// The same for graphs of low mean degree: a workgroup takes kBlock
// consecutive rows, streams their entries through LDS in coalesced tiles and
// each thread adds its row's part of every tile in order.
constexpr int kDegTile = 2048;
constexpr i64 kDegTiledMean = 64;  // mean degree up to which the tiled form is used
__global__ __launch_bounds__(kBlock) void k_degree_rows(const i64 *off, const double *w, i64 nv, double *k) {
    __shared__ double tile[kDegTile];
    for (i64 r0 = (i64)blockIdx.x * kBlock; r0 < nv; r0 += (i64)gridDim.x * kBlock) {
        const i64 r = r0 + threadIdx.x;
        const i64 rb = r < nv ? off[r] : 0, re = r < nv ? off[r + 1] : 0;
        const bool mine = r < nv && re - rb <= kBinMaxDeg[7];
        const i64 e0 = off[r0], e1 = off[r0 + kBlock < nv ? r0 + kBlock : nv];
        double acc = 0;
        for (i64 t0 = e0; t0 < e1; t0 += kDegTile) {
            const i64 t1 = t0 + kDegTile < e1 ? t0 + kDegTile : e1;
            __syncthreads();  // the previous tile is read
            for (i64 j = t0 + threadIdx.x; j < t1; j += kBlock) tile[j - t0] = w[j];
            __syncthreads();
            if (mine) {
                const i64 lo = rb > t0 ? rb : t0, hi = re < t1 ? re : t1;
                for (i64 j = lo; j < hi; j++) acc += tile[j - t0];
            }
        }
        if (mine) k[r] = acc;
    }
}
__global__ void k_degree_big(const i64 *off, const double *w, const i64 *verts, i64 count, double *k) {
    for (i64 q = blockIdx.x; q < count; q += gridDim.x) {
        i64 v = verts[q];
        double s = 0;
        for (i64 j = off[v] + threadIdx.x; j < off[v + 1]; j += blockDim.x) s += w[j];
        s = block_sum(s);
        if (threadIdx.x == 0) k[v] = s;
    }
}
__global__ void k_init_comm(const double *k, i64 nv, double *a, i64 *n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nv; i += (i64)gridDim.x * blockDim.x) {
        a[i] = k[i];
        n[i] = 1;
    }
}
__global__ void k_comm_from_labels(const i64 *C, const double *k, i64 nv, double *a, i64 *n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nv; i += (i64)gridDim.x * blockDim.x) {
        i64 c = C[i];
        if (c >= 0) {
            atomicAdd(&a[c], k[i]);
            atomicAdd((unsigned long long *)&n[c], 1ULL);
        }
    }
}

// ----------------------------------------------------------------- reductions
template <bool SQ>
__global__ __launch_bounds__(kBlock) void k_red(const double *x, i64 n, double *part) {
    double s = 0;
    for (i64 i = blockIdx.x * (i64)kBlock + threadIdx.x; i < n; i += (i64)gridDim.x * kBlock) {
        double v = x[i];
        s += SQ ? v * v : v;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *part, i64 np, double *out) {
    double s = 0;
    for (i64 i = threadIdx.x; i < np; i += kBlock) s += part[i];
    s = block_sum(s);
    if (threadIdx.x == 0) out[0] = s;
}

// --------------------------------------------------------------------- sweeps
// Sub-wave groups of G lanes, one vertex (degree <= G) per group.  Each lane
// loads one entry; aggregation walks the group's entries in adjacency order
// from LDS, so every per-community weight is the same left-to-right sum that
// buildLocalMapCounter forms.  The lane holding the first occurrence of a
// community scores it; a butterfly over the group picks the best.
template <int G, bool PLAIN>
__global__ __launch_bounds__(kBlock) void k_sweep_group(SweepArgs A) {
    if (A.skip && *A.skip) return;
    constexpr int VPB = kBlock / G;
    __shared__ i64 s_c[kBlock];
    __shared__ double s_w[kBlock];
    __shared__ int s_self[kBlock];
    __shared__ double s_part[VPB];
    const int tid = threadIdx.x, lane = tid % G, grp = tid / G;
    const i64 idx = (i64)blockIdx.x * VPB + grp;
    const bool valid = idx < A.count;
    i64 v = 0, beg = 0;
    int deg = 0;
    if (valid) {
        v = A.verts[idx];
        beg = A.off[v];
        deg = (int)(A.off[v + 1] - beg);
    }
    i64 c = -2;
    double ww = 0;
    int self = 0;
    if (lane < deg) {
        i64 t = A.tails[beg + lane];
        ww = A.w[beg + lane];
        c = A.C[t];
        self = (t == v);
    }
    s_c[tid] = c;
    s_w[tid] = ww;
    s_self[tid] = self;
    __syncthreads();
    const i64 s = valid ? A.C[v] : -3;
    double e_mine = 0, e_s = 0, sl = 0;
    bool first = true;
    const int gb = grp * G;
    for (int m = 0; m < deg; m++) {
        i64 cm = s_c[gb + m];
        double wm = s_w[gb + m];
        if (cm == c) {
            e_mine += wm;
            if (m < lane) first = false;
        }
        if (cm == s) e_s += wm;
        if (s_self[gb + m]) sl += wm;
    }
    bool cand = lane < deg && first && c != s;
    double gain = 0;
    float t1 = -INFINITY, t2 = -INFINITY;
    const double cst = valid ? vconst(A, v) : A.constant;
    if (cand) {
        const double kv = A.k[v];
        gain = gain_of(e_mine, e_s - sl, kv, A.a[c], A.a[s] - kv, cst);
        if (A.feat) t1 = (float)(gain / (2.0 * kv));
        cand = gain > 0;  // max() starts from maxGain = 0 and needs a strict increase
    }
    const i64 my_c = c;
    const bool my_first = lane < deg && first;
    group_argmax<G>(cand, gain, c);
    // weight to the winner, from the lane that summed it (exactly one lane)
    double e_t = 0;
    if (A.dein) {
        e_t = cand && my_first && my_c == c ? e_mine : 0.0;
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) e_t += __shfl_xor(e_t, o, G);
    }
    if (A.feat) top2_group<G>(t1, t2);
    if (valid && lane == 0) {
        i64 target = s;
        if (cand) {
            target = c;
            if (A.n[target] == 1 && A.n[s] == 1 && target > s) target = s;  // swap protection
        }
        A.tgt[v] = target;
        if (A.feat || A.dein) {
            const double kv = A.k[v];
            if (A.feat) A.feat[v] = las_record(e_s, kv, t1, t2);
            if (A.dein)
                A.dein[v] = target != s ? 2.0 * (e_t - (e_s - sl)) : 0.0;
        }
        if (A.eii_v) A.eii_v[v] = e_s;
    }
    if (PLAIN) {
        if (lane == 0) s_part[grp] = valid ? e_s : 0.0;
        __syncthreads();
        if (tid == 0) {
            double t = 0;
            for (int q = 0; q < VPB; q++) t += s_part[q];
            A.eii_part[blockIdx.x] = t;
        }
    }
}

// Open-addressing hash of (community -> weight), linear probing, Fibonacci
// hashing of the id.  `keys` holds kEmpty in free slots.
__device__ inline void hash_insert(i64 *keys, double *vals, u64 mask, int shift, i64 c, double w) {
    u64 h = ((u64)c * kFib) >> shift;
    while (true) {
        i64 k = ld_relaxed(&keys[h]);
        if (k == kEmpty) {
            k = (i64)atomicCAS((unsigned long long *)&keys[h], (unsigned long long)kEmpty,
                               (unsigned long long)c);
            if (k == kEmpty) k = c;
        }
        if (k == c) {
            atomicAdd(&vals[h], w);
            return;
        }
        h = (h + 1) & mask;
    }
}
__device__ inline double hash_lookup(const i64 *keys, const double *vals, u64 mask, int shift, i64 c) {
    u64 h = ((u64)c * kFib) >> shift;
    while (true) {
        i64 k = ld_agent(&keys[h]);
        if (k == c) return ld_agent(&vals[h]);
        if (k == kEmpty) return 0.0;
        h = (h + 1) & mask;
    }
}

// Score every occupied slot of a table and reduce over the calling group.
__device__ inline void score_slots(const i64 *keys, const double *vals, i64 slots, int first, int step,
                                   const SweepArgs &A, i64 s, double kv, double eix, double ax, double cst,
                                   bool &cand, double &gain, i64 &id, float &t1, float &t2) {
    const double inv2k = A.feat ? 1.0 / (2.0 * kv) : 0.0;
    for (i64 h = first; h < slots; h += step) {
        i64 y = ld_agent(&keys[h]);
        if (y == kEmpty || y == s) continue;
        double g = gain_of(ld_agent(&vals[h]), eix, kv, A.a[y], ax, cst);
        if (A.feat) top2_push(t1, t2, (float)(g * inv2k));
        bool c = g > 0;
        if (better(c, g, y, cand, gain, id)) {
            cand = true;
            gain = g;
            id = y;
        }
    }
}

// One wavefront per vertex (degree <= 256), a private 512-slot LDS table.
template <bool PLAIN>
__global__ __launch_bounds__(kBlock) void k_sweep_wavehash(SweepArgs A) {
    if (A.skip && *A.skip) return;
    constexpr int WPB = kBlock / kWave;
    constexpr int SLOTS = kWaveHashSlots;
    constexpr int SHIFT = 64 - 9;
    __shared__ i64 s_keys[WPB][SLOTS];
    __shared__ double s_vals[WPB][SLOTS];
    __shared__ double s_part[WPB];
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const i64 idx = (i64)blockIdx.x * WPB + wid;
    const bool valid = idx < A.count;
    i64 *keys = s_keys[wid];
    double *vals = s_vals[wid];
    for (int h = lane; h < SLOTS; h += kWave) {
        keys[h] = kEmpty;
        vals[h] = 0;
    }
    __syncthreads();
    i64 v = valid ? A.verts[idx] : 0;
    i64 beg = valid ? A.off[v] : 0, end = valid ? A.off[v + 1] : 0;
    double sl = 0;
    for (i64 j = beg + lane; j < end; j += kWave) {
        i64 t = A.tails[j];
        double ww = A.w[j];
        if (t == v) sl += ww;
        hash_insert(keys, vals, SLOTS - 1, SHIFT, A.C[t], ww);
    }
    __syncthreads();
    sl = wave_sum(sl);
    const i64 s = valid ? A.C[v] : -3;
    const double e_s = valid ? hash_lookup(keys, vals, SLOTS - 1, SHIFT, s) : 0.0;
    bool cand = false;
    double gain = 0;
    i64 id = INT64_MAX;
    float t1 = -INFINITY, t2 = -INFINITY;
    const double cst = valid ? vconst(A, v) : A.constant;
    if (valid) {
        const double kv = A.k[v];
        score_slots(keys, vals, SLOTS, lane, kWave, A, s, kv, e_s - sl, A.a[s] - kv, cst, cand, gain, id,
                    t1, t2);
    }
    group_argmax<kWave>(cand, gain, id);
    if (A.feat) top2_group<kWave>(t1, t2);
    if (valid && lane == 0) {
        i64 target = s;
        if (cand) {
            target = id;
            if (A.n[target] == 1 && A.n[s] == 1 && target > s) target = s;
        }
        A.tgt[v] = target;
        if (A.feat || A.dein) {
            const double kv = A.k[v];
            if (A.feat) A.feat[v] = las_record(e_s, kv, t1, t2);
            if (A.dein)
                A.dein[v] = target != s
                                ? 2.0 * (hash_lookup(keys, vals, SLOTS - 1, SHIFT, target) - (e_s - sl))
                                : 0.0;
        }
        if (A.eii_v) A.eii_v[v] = e_s;
    }
    if (PLAIN) {
        if (lane == 0) s_part[wid] = valid ? e_s : 0.0;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0;
            for (int q = 0; q < WPB; q++) t += s_part[q];
            A.eii_part[blockIdx.x] = t;
        }
    }
}

// One workgroup per vertex.  GLOBAL=false: degree <= 2048, a 4096-slot LDS
// table.  GLOBAL=true: any degree, a table of the smallest power of two >= 2x
// the degree in the block's private HBM region; blocks stride over vertices.
template <bool PLAIN, bool GLOBAL>
__global__ __launch_bounds__(kBlock) void k_sweep_blockhash(SweepArgs A) {
    if (A.skip && *A.skip) return;
    constexpr int LSLOTS = GLOBAL ? 1 : kBlockHashSlots;
    __shared__ i64 l_keys[LSLOTS];
    __shared__ double l_vals[LSLOTS];
    double part = 0;
    for (i64 q = blockIdx.x; q < A.count; q += gridDim.x) {
        const i64 v = A.verts[q];
        const i64 beg = A.off[v], end = A.off[v + 1];
        i64 *keys;
        double *vals;
        i64 slots;
        if (GLOBAL) {
            keys = A.gkeys + (i64)blockIdx.x * A.gslots;
            vals = A.gvals + (i64)blockIdx.x * A.gslots;
            slots = 1;
            while (slots < 2 * (end - beg)) slots <<= 1;
        } else {
            keys = l_keys;
            vals = l_vals;
            slots = 64;
            while (slots < 2 * (end - beg)) slots <<= 1;
        }
        int lg = 0;
        while (((i64)1 << lg) < slots) lg++;
        const int shift = 64 - lg;
        const u64 mask = (u64)slots - 1;
        for (i64 h = threadIdx.x; h < slots; h += kBlock) {
            keys[h] = kEmpty;
            vals[h] = 0;
        }
        if (GLOBAL) __threadfence_block();
        __syncthreads();
        double sl = 0;
        for (i64 j = beg + threadIdx.x; j < end; j += kBlock) {
            i64 t = A.tails[j];
            double ww = A.w[j];
            if (t == v) sl += ww;
            hash_insert(keys, vals, mask, shift, A.C[t], ww);
        }
        if (GLOBAL) __threadfence_block();
        __syncthreads();
        sl = block_sum(sl);
        const i64 s = A.C[v];
        const double e_s = hash_lookup(keys, vals, mask, shift, s);
        const double kv = A.k[v];
        bool cand = false;
        double gain = 0;
        i64 id = INT64_MAX;
        float t1 = -INFINITY, t2 = -INFINITY;
        const double cst = vconst(A, v);
        score_slots(keys, vals, slots, threadIdx.x, kBlock, A, s, kv, e_s - sl, A.a[s] - kv, cst, cand,
                    gain, id, t1, t2);
        block_argmax(cand, gain, id);
        if (A.feat) block_top2<kBlock>(t1, t2);
        if (threadIdx.x == 0) {
            i64 target = s;
            if (cand) {
                target = id;
                if (A.n[target] == 1 && A.n[s] == 1 && target > s) target = s;
            }
            A.tgt[v] = target;
            if (A.feat) A.feat[v] = las_record(e_s, kv, t1, t2);
            if (A.dein)
                A.dein[v] = target != s ? 2.0 * (hash_lookup(keys, vals, mask, shift, target) - (e_s - sl)) : 0.0;
            if (A.eii_v) A.eii_v[v] = e_s;
        }
        part += e_s;
        __syncthreads();
    }
    if (PLAIN && threadIdx.x == 0) A.eii_part[blockIdx.x] = part;
}

// One workgroup per vertex with a dense LDS accumulator indexed by community
// id (graphs of at most kDenseMaxNV vertices: the coarse phases).  A bitmap
// records which communities were touched, so clearing and scoring cost the
// vertex's own entries plus nv/32 bitmap words, not nv slots.
template <bool PLAIN>
__global__ __launch_bounds__(kBlock) void k_sweep_dense(SweepArgs A) {
    if (A.skip && *A.skip) return;
    __shared__ double acc[kDenseMaxNV];
    __shared__ uint32_t touched[kDenseMaxNV / 32];
    __shared__ uint16_t list[kDenseListMax];
    __shared__ int wsum[kBlock / kWave];
    const int nv = (int)A.nv;
    const int nwords = (nv + 31) / 32;
    // acc[] and touched[] are all zero between vertices
    for (int i = threadIdx.x; i < nv; i += kBlock) acc[i] = 0.0;
    for (int i = threadIdx.x; i < nwords; i += kBlock) touched[i] = 0;
    double part = 0;
    __syncthreads();
    for (i64 q = blockIdx.x; q < A.count; q += gridDim.x) {
        const i64 v = A.verts[q];
        const i64 beg = A.off[v], end = A.off[v + 1];
        double sl = 0;
        for (i64 j = beg + threadIdx.x; j < end; j += kBlock) {
            const i64 t = A.tails[j];
            const double ww = A.w[j];
            if (t == v) sl += ww;
            const i64 c = A.C[t];
            atomicOr(&touched[c >> 5], 1u << (c & 31));
            atomicAdd(&acc[c], ww);
        }
        __syncthreads();
        sl = block_sum(sl);
        const int ntouched = compact_touched<kBlock>(touched, nwords, list, wsum);
        const i64 s = A.C[v];
        const double e_s = acc[s];
        const double kv = A.k[v];
        const double eix = e_s - sl, ax = A.a[s] - kv;
        const double cst = vconst(A, v);
        bool cand = false;
        double gain = 0;
        i64 id = INT64_MAX;
        float t1 = -INFINITY, t2 = -INFINITY;
        const double inv2k = A.feat ? 1.0 / (2.0 * kv) : 0.0;
        for (int k = threadIdx.x; k < ntouched; k += kBlock) {
            const i64 y = list[k];
            if (y == s) continue;
            const double g = gain_of(acc[y], eix, kv, A.a[y], ax, cst);
            if (A.feat) top2_push(t1, t2, (float)(g * inv2k));
            if (better(g > 0, g, y, cand, gain, id)) {
                cand = true;
                gain = g;
                id = y;
            }
        }
        block_argmax(cand, gain, id);  // ends with a barrier: every read of acc is done
        if (A.feat) block_top2<kBlock>(t1, t2);
        if (threadIdx.x == 0) {
            i64 target = s;
            if (cand) {
                target = id;
                if (A.n[target] == 1 && A.n[s] == 1 && target > s) target = s;
            }
            A.tgt[v] = target;
            if (A.feat) A.feat[v] = las_record(e_s, kv, t1, t2);
            if (A.dein) A.dein[v] = target != s ? 2.0 * (acc[target] - eix) : 0.0;
            if (A.eii_v) A.eii_v[v] = e_s;
        }
        part += e_s;
        __syncthreads();  // thread 0's read of acc[target] is done
        for (int k = threadIdx.x; k < ntouched; k += kBlock) acc[list[k]] = 0.0;
        for (int i = threadIdx.x; i < nwords; i += kBlock) touched[i] = 0;
        __syncthreads();
    }
    if (PLAIN && threadIdx.x == 0) A.eii_part[blockIdx.x] = part;
}

// Giant vertices (degree > kGiantDeg), kGiantBlocks workgroups each
// (blockIdx.y = vertex of the slice).  Accumulate: every workgroup sums
// chunks of kGiantChunk entries per community in an LDS table, then adds the
// chunk's sums into the vertex's HBM table; score: the workgroups split the
// vertex's table, score and clear their slots; decide: one wavefront per
// vertex merges the workgroups' candidates.
constexpr int kGiantChunk = 2048;
constexpr int kGiantLds = 2 * kGiantChunk;

__global__ __launch_bounds__(kBlock) void k_giant_accum(SweepArgs A) {
    if (A.skip && *A.skip) return;
    __shared__ i64 lk[kGiantLds];
    __shared__ double lv[kGiantLds];
    const i64 q = blockIdx.y;
    const i64 v = A.verts[q];
    const i64 beg = A.off[v], end = A.off[v + 1];
    const i64 s = A.C[v];
    const i64 tb = A.gt_tab[q], slots = A.gt_tab[q + 1] - tb;
    i64 *keys = A.gt_keys + tb;
    double *vals = A.gt_vals + tb;
    int lg = 0;
    while (((i64)1 << lg) < slots) lg++;
    const int gshift = 64 - lg;
    const u64 gmask = (u64)slots - 1;
    constexpr int LSHIFT = 64 - 12;  // log2(kGiantLds) = 12
    double sl = 0, es = 0;
    for (i64 c0 = beg + (i64)blockIdx.x * kGiantChunk; c0 < end; c0 += (i64)gridDim.x * kGiantChunk) {
        for (int h = threadIdx.x; h < kGiantLds; h += kBlock) {
            lk[h] = kEmpty;
            lv[h] = 0.0;
        }
        __syncthreads();
        const i64 c1 = c0 + kGiantChunk < end ? c0 + kGiantChunk : end;
        for (i64 j = c0 + threadIdx.x; j < c1; j += kBlock) {
            const i64 t = A.tails[j];
            const double ww = A.w[j];
            const i64 cc = A.C[t];
            if (t == v) sl += ww;
            if (cc == s) es += ww;
            hash_insert(lk, lv, kGiantLds - 1, LSHIFT, cc, ww);
        }
        __syncthreads();
        for (int h = threadIdx.x; h < kGiantLds; h += kBlock) {
            const i64 k = lk[h];
            if (k != kEmpty) hash_insert(keys, vals, gmask, gshift, k, lv[h]);
        }
        __syncthreads();
    }
    sl = block_sum(sl);
    es = block_sum(es);
    if (threadIdx.x == 0) {
        if (es != 0.0) atomicAdd(&A.gt_sc[2 * q], es);
        if (sl != 0.0) atomicAdd(&A.gt_sc[2 * q + 1], sl);
    }
}

__global__ __launch_bounds__(kBlock) void k_giant_score(SweepArgs A) {
    if (A.skip && *A.skip) return;
    const i64 q = blockIdx.y;
    const i64 v = A.verts[q];
    const i64 s = A.C[v];
    const i64 tb = A.gt_tab[q], slots = A.gt_tab[q + 1] - tb;
    i64 *keys = A.gt_keys + tb;
    double *vals = A.gt_vals + tb;
    const double kv = A.k[v];
    const double e_s = ld_agent(&A.gt_sc[2 * q]), sl = ld_agent(&A.gt_sc[2 * q + 1]);
    const double eix = e_s - sl, ax = A.a[s] - kv;
    const double cst = vconst(A, v);
    const double inv2k = A.feat ? 1.0 / (2.0 * kv) : 0.0;
    const i64 per = slots / gridDim.x;
    const i64 h0 = (i64)blockIdx.x * per;
    bool cand = false;
    double gain = 0;
    i64 id = INT64_MAX;
    float t1 = -INFINITY, t2 = -INFINITY;
    for (i64 h = h0 + threadIdx.x; h < h0 + per; h += kBlock) {
        const i64 y = ld_agent(&keys[h]);
        if (y == kEmpty) continue;
        const double ey = ld_agent(&vals[h]);
        keys[h] = kEmpty;
        vals[h] = 0.0;
        if (y == s) continue;
        const double g = gain_of(ey, eix, kv, A.a[y], ax, cst);
        if (A.feat) top2_push(t1, t2, (float)(g * inv2k));
        if (better(g > 0, g, y, cand, gain, id)) {
            cand = true;
            gain = g;
            id = y;
        }
    }
    block_argmax(cand, gain, id);
    if (A.feat) block_top2<kBlock>(t1, t2);
    if (threadIdx.x == 0) {
        const i64 r = q * kGiantBlocks + blockIdx.x;
        A.gt_rg[r] = cand ? gain : -1.0;
        A.gt_ri[r] = id;
        A.gt_rt[2 * r] = t1;
        A.gt_rt[2 * r + 1] = t2;
    }
}

template <bool PLAIN>
__global__ __launch_bounds__(kWave) void k_giant_decide(SweepArgs A) {
    if (A.skip && *A.skip) return;
    const i64 q = blockIdx.x;
    const int lane = threadIdx.x;
    const i64 v = A.verts[q];
    const i64 s = A.C[v];
    const i64 r = q * kGiantBlocks + lane;
    bool cand = false;
    double gain = 0;
    i64 id = INT64_MAX;
    float t1 = -INFINITY, t2 = -INFINITY;
    if (lane < kGiantBlocks) {
        const double g = A.gt_rg[r];
        cand = g > 0;
        gain = cand ? g : 0.0;
        id = A.gt_ri[r];
        t1 = A.gt_rt[2 * r];
        t2 = A.gt_rt[2 * r + 1];
    }
    group_argmax<kWave>(cand, gain, id);
    if (A.feat) top2_group<kWave>(t1, t2);
    if (lane == 0) {
        const double e_s = A.gt_sc[2 * q];
        i64 target = s;
        if (cand) {
            target = id;
            if (A.n[target] == 1 && A.n[s] == 1 && target > s) target = s;
        }
        A.tgt[v] = target;
        const double kv = A.k[v];
        if (A.feat) A.feat[v] = las_record(e_s, kv, t1, t2);
        if (A.dein) A.dein[v] = target != s ? gain + 2 * kv * (A.a[target] - (A.a[s] - kv)) * vconst(A, v) : 0.0;
        if (PLAIN) A.eii_part[q] = e_s;
        if (A.eii_v) A.eii_v[v] = e_s;
        A.gt_sc[2 * q] = 0.0;
        A.gt_sc[2 * q + 1] = 0.0;
    }
}

__global__ void k_giant_table_sizes(const i64 *off, const i64 *verts, i64 count, i64 *tab) {
    for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q < count; q += (i64)gridDim.x * blockDim.x) {
        const i64 v = verts[q];
        const i64 d = off[v + 1] - off[v];
        i64 sl = 64;
        while (sl < 2 * d) sl <<= 1;
        tab[q] = sl;
    }
}

// ---------------------------------------------------------- persistent sweep
// A whole colored iteration of a small graph (at most kDenseMaxNV vertices) in
// one persistent launch: for every color, the workgroups share out the
// vertices of the class (dense LDS accumulator per workgroup), meet at a grid
// barrier, commit their own vertices' moves, and meet again.  The modularity
// partials of the iteration follow the last color.  Grid barriers spin on a
// generation counter with a bounded wait; a workgroup that gives up raises
// `abort`, which releases every other wait, so every wave always exits.
constexpr int kPBlock = 1024;

__device__ inline double block_sum_p(double v, double *scratch) {
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) scratch[w] = v;
    __syncthreads();
    double r = 0;
    for (int q = 0; q < kPBlock / kWave; q++) r += scratch[q];
    return r;
}

// Block-wide argmax of the persistent sweeps (kPBlock threads; sk and si
// hold the wavefronts' keys and ids), result valid in every thread.
__device__ inline void block_argmax_p(bool &cand, double &gain, i64 &id, u64 *sk, i64 *si) {
    u64 key = argmax_key(cand, gain);
    i64 rid = id, none = 0;
    group_argmax_key<kWave>(key, rid, none);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) {
        sk[w] = key;
        si[w] = rid;
    }
    __syncthreads();
    // every wavefront reduces the wavefronts' results (lane l takes wavefront l % kPBlock/kWave)
    key = sk[l % (kPBlock / kWave)];
    rid = si[l % (kPBlock / kWave)];
    group_argmax_key<kPBlock / kWave>(key, rid, none);
    cand = key != 0;
    if (cand) {
        gain = key_gain(key);
        id = rid;
    }
}

// block_argmax_p that also carries `nid` (the size of the candidate community)
__device__ inline void block_argmax_pn(bool &cand, double &gain, i64 &id, i64 &nid, u64 *sk, i64 *si, i64 *sn) {
    u64 key = argmax_key(cand, gain);
    i64 rid = id, rn = nid;
    group_argmax_key<kWave, true>(key, rid, rn);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) {
        sk[w] = key;
        si[w] = rid;
        sn[w] = rn;
    }
    __syncthreads();
    // every wavefront reduces the wavefronts' results (lane l takes wavefront l % kPBlock/kWave)
    key = sk[l % (kPBlock / kWave)];
    rid = si[l % (kPBlock / kWave)];
    rn = sn[l % (kPBlock / kWave)];
    group_argmax_key<kPBlock / kWave, true>(key, rid, rn);
    cand = key != 0;
    if (cand) {
        gain = key_gain(key);
        id = rid;
        nid = rn;
    }
}

// LDS of the persistent sweep: either the dense accumulator of graphs with at
// most kDenseMaxNV vertices, or (HASH) an open-addressing table of
// kPHashSlots (community -> weight) slots for vertices of degree up to
// kPHashMaxDeg plus a touched-community bitmap of the whole graph for larger
// vertices, which accumulate into the workgroup's dense HBM row (A.dacc).
constexpr int kPHashSlots = 4096;
constexpr int kGatherUnroll = 4;
constexpr int kSlabMaxNV = 2 * (int)kDenseMaxNV;
constexpr i64 kPHashMaxDeg = kPHashSlots / 2;
constexpr int kPBitmapWords = (int)(kPersistMaxNV / 32);
constexpr size_t kPDenseBytes = kDenseMaxNV * 8 + kDenseMaxNV / 32 * 4 + kDenseListMax * 2;
constexpr size_t kPHashBytes = (size_t)kPHashSlots * 16 + (size_t)kPBitmapWords * 4;
constexpr size_t kPSmemBytes = kPDenseBytes > kPHashBytes ? kPDenseBytes : kPHashBytes;

// cperm: vertices grouped by color (cstart[c] .. cstart[c+1]).
//
// One grid barrier per color step: the labels, community degrees and sizes
// are double-buffered ((C, a, n) and (C1, a1, n1), equal at launch).  Step k
// reads buffer k % 2, which holds the state after step k-1, and brings the
// other buffer -- two steps old -- up to date by applying there the moves of
// step k-1 and of step k, so after the step's barrier it holds the state
// after step k.  Decisions of a step therefore never see a move of their own
// step (frozen state), as with a separate commit phase.  On exit the final
// state is copied to (C, a, n).
template <bool HASH>
__global__ __launch_bounds__(kPBlock) void k_colored_iteration(SweepArgs A, i64 *C, double *a,
                                                               i64 *n, const i64 *cperm,
                                                               const i64 *cstart, int ncol,
                                                               unsigned *bar, double *part,
                                                               uint8_t *stale, i64 *C1, double *a1,
                                                               i64 *n1) {
    __shared__ __align__(16) unsigned char smem[kPSmemBytes];
    __shared__ double red[kPBlock / kWave];
    __shared__ u64 sk[kPBlock / kWave];
    __shared__ i64 si[kPBlock / kWave];
    __shared__ int s_cnt;     // communities listed for the current vertex (dense view)
    __shared__ double s_sl;   // its self-loop weight
    __shared__ i64 s_nid;     // dense view: size of the chosen community
    // dense view
    double *acc = reinterpret_cast<double *>(smem);
    uint32_t *touched = reinterpret_cast<uint32_t *>(smem + kDenseMaxNV * 8);
    uint16_t *list = reinterpret_cast<uint16_t *>(smem + kDenseMaxNV * 8 + kDenseMaxNV / 32 * 4);
    // hash view
    i64 *hkeys = reinterpret_cast<i64 *>(smem);
    double *hvals = reinterpret_cast<double *>(smem + (size_t)kPHashSlots * 8);
    uint32_t *hbits = reinterpret_cast<uint32_t *>(smem + (size_t)kPHashSlots * 16);
    double *dacc = HASH ? A.dacc + (i64)blockIdx.x * A.nv : nullptr;
    const int nv = (int)A.nv;
    const int nwords = (nv + 31) / 32;
    const unsigned nb = gridDim.x;
    unsigned seq = 0;
    if (threadIdx.x == 0) {
        s_cnt = 0;
        s_sl = 0.0;
    }
    if (HASH) {
        for (int i = threadIdx.x; i < kPHashSlots; i += kPBlock) {
            hkeys[i] = kEmpty;
            hvals[i] = 0.0;
        }
        for (int i = threadIdx.x; i < nwords; i += kPBlock) hbits[i] = 0;
    } else {
        for (int i = threadIdx.x; i < nv; i += kPBlock) acc[i] = 0.0;
        for (int i = threadIdx.x; i < nwords; i += kPBlock) touched[i] = 0;
    }
    for (i64 i = (i64)blockIdx.x * kPBlock + threadIdx.x; i < nv; i += (i64)nb * kPBlock) {
        C1[i] = C[i];
        a1[i] = a[i];
        n1[i] = n[i];
    }
    bool alive = grid_barrier(bar, nb, ++seq);
    int par = 0;            // buffer holding the state after the previous step
    i64 plo = 0, ptot = 0;  // the previous step's class
    // Read-ahead of the workgroup's first vertex of the next step, issued along
    // with the loads of the current step so that a step starts from the labels
    // of its neighbours: id, entries, label (a vertex's label changes only in
    // its own step, so either buffer holds it) and degree, and per thread the
    // first kPBlock entries.
    int ra_c = -1;
    i64 ra_v = -1, ra_beg = 0, ra_end = 0, ra_s = -1, ra_t = 0;
    double ra_k = 0, ra_w = 0;
    // thread 0: the move made for its first vertex of the previous step
    i64 mv_v = -1, mv_t = -1, mv_s = -1;
    double mv_k = 0;
    for (int c = 0; c < ncol && alive; c++) {
        const i64 lo = cstart[c], hi = cstart[c + 1], total = hi - lo;
        if (total == 0) continue;
        const i64 hi2 = c + 1 < ncol ? cstart[c + 2] : hi;
        A.C = par ? C1 : C;
        A.a = par ? a1 : a;
        A.n = par ? n1 : n;
        i64 *Cn = par ? C : C1;
        double *an = par ? a : a1;
        i64 *nn = par ? n : n1;
        // the previous step's moves of this workgroup's vertices, into the other buffer
        if (threadIdx.x == 0) {
            if (mv_v >= 0 && mv_t != mv_s) {
                move_vertex(mv_t, mv_s, mv_k, an, nn);
                Cn[mv_v] = mv_t;
            }
            for (i64 q = blockIdx.x + nb; q < ptot; q += nb) {
                const i64 v = cperm[plo + q];
                const i64 t = A.tgt[v], s = Cn[v];
                if (t != s) {
                    move_vertex(t, s, A.k[v], an, nn);
                    Cn[v] = t;
                }
            }
        }
        mv_v = -1;
        // read-ahead, first load: the next step's vertex of this workgroup
        const i64 nx_v = (i64)blockIdx.x < hi2 - hi ? cperm[hi + blockIdx.x] : -1;
        i64 nx_beg = 0, nx_end = 0, nx_s = -1;
        double nx_k = 0;
        bool nx_done = false;
        for (i64 q = blockIdx.x; q < total; q += nb) {
            const bool first = q == (i64)blockIdx.x;
            const bool ra = first && ra_c == c;
            const i64 v = ra ? ra_v : cperm[lo + q];
            const i64 beg = ra ? ra_beg : A.off[v], end = ra ? ra_end : A.off[v + 1];
            const i64 s = ra ? ra_s : A.C[v];
            if (s < 0) {  // a vertex without entries (label -1) stays where it is
                if (threadIdx.x == 0) A.tgt[v] = s;
                continue;
            }
            const double as = A.a[s];
            i64 ns = 0;
            if (threadIdx.x == 0) ns = A.n[s];
            double sl = 0, e_s = 0;
            // HASH: LDS table of this vertex; when it is large, the dense view
            // one slab of kDenseMaxNV communities at a time (graphs of at most
            // kSlabMaxNV vertices, the slab of its own community first), else
            // the dense HBM row
            const bool big = HASH && end - beg > kPHashMaxDeg;
            const bool slab = big && nv <= kSlabMaxNV;
            const bool dense = !HASH || slab;
            const int nsl = slab ? (nv + (int)kDenseMaxNV - 1) / (int)kDenseMaxNV : 1;
            const int s_slab = slab ? (int)(s / kDenseMaxNV) : 0;
            i64 slots = 64;
            int shift = 58;
            if (HASH && !big) {
                while (slots < 2 * (end - beg)) slots <<= 1;
                int lg = 0;
                while (((i64)1 << lg) < slots) lg++;
                shift = 64 - lg;
            }
            const u64 mask = (u64)slots - 1;
            if (slab) {  // the dense view replaces the (empty) hash view
                for (int i = threadIdx.x; i < (int)kDenseMaxNV; i += kPBlock) acc[i] = 0.0;
                for (int i = threadIdx.x; i < (int)(kDenseMaxNV / 32); i += kPBlock) touched[i] = 0;
                __syncthreads();
            }
            const double kv = ra ? ra_k : A.k[v];
            const double ax = as - kv;
            const double inv2k = A.feat ? 1.0 / (2.0 * kv) : 0.0;
            bool cand = false;
            double gain = 0;
            i64 id = INT64_MAX;
            i64 nid = 0;  // dense view: size of the community `id`
            float t1 = -INFINITY, t2 = -INFINITY;
            int ntouched = 0;
            double eix = 0;
            for (int pi = 0; pi < nsl; pi++) {
                const i64 base = slab ? (i64)((s_slab + pi) % nsl) * kDenseMaxNV : 0;
                // kGatherUnroll passes of the workgroup at a time: their entry loads,
                // then their label loads, then the accumulation, so that a heavy
                // vertex waits for two loads per kGatherUnroll passes
                for (i64 j0 = beg + threadIdx.x; j0 < end; j0 += kGatherUnroll * kPBlock) {
                    i64 tt[kGatherUnroll], cl[kGatherUnroll];
                    double wt[kGatherUnroll];
#pragma unroll
                    for (int u = 0; u < kGatherUnroll; u++) {
                        const i64 j = j0 + (i64)u * kPBlock;
                        const bool pre = ra && pi == 0 && u == 0 && j0 == beg + threadIdx.x;
                        tt[u] = j < end ? (pre ? ra_t : A.tails[j]) : -1;
                        wt[u] = j < end ? (pre ? ra_w : A.w[j]) : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < kGatherUnroll; u++) cl[u] = tt[u] >= 0 ? A.C[tt[u]] : 0;
#pragma unroll
                    for (int u = 0; u < kGatherUnroll; u++) {
                        const i64 cc = cl[u] - base;
                        const bool valid = tt[u] >= 0;
                        const double ww = wt[u];
                        if (valid && pi == 0 && tt[u] == v) atomicAdd(&s_sl, ww);
                        if (!dense) {
                            if (!valid) continue;
                            if (big) {
                                atomicOr(&hbits[cc >> 5], 1u << (cc & 31));
                                atomicAdd(&dacc[cc], ww);
                            } else {
                                hash_insert(hkeys, hvals, mask, shift, cc, ww);
                            }
                        } else {
                            // the first lane to touch a community of the slab lists
                            // it (one LDS counter update per wavefront)
                            const bool in = valid && (!slab || (cc >= 0 && cc < kDenseMaxNV));
                            bool fresh = false;
                            if (in) {
                                const uint32_t bit = 1u << (cc & 31);
                                fresh = !(atomicOr(&touched[cc >> 5], bit) & bit);
                            }
                            const u64 fm = __ballot(fresh);
                            if (fm) {
                                const int lane = threadIdx.x % kWave;
                                const int leader = __ffsll((long long)fm) - 1;
                                int at = 0;
                                if (lane == leader) at = atomicAdd(&s_cnt, __popcll(fm));
                                at = __shfl(at, leader);
                                const u64 lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
                                if (fresh) list[at + __popcll(fm & lt)] = (uint16_t)cc;
                            }
                            if (in) atomicAdd(&acc[cc], ww);
                        }
                    }
                }
                if (HASH) __threadfence_block();
                __syncthreads();
                if (pi == 0) {
                    if (first && nx_v >= 0) {  // read-ahead, second load
                        nx_beg = A.off[nx_v];
                        nx_end = A.off[nx_v + 1];
                        nx_s = C[nx_v];
                        nx_k = A.k[nx_v];
                    }
                    sl = s_sl;
                }
                if (HASH && !dense) {
                    if (big) {
                        e_s = ld_agent(&dacc[s]);
                        eix = e_s - sl;
                        for (int wi = threadIdx.x; wi < nwords; wi += kPBlock) {
                            uint32_t m = hbits[wi];
                            while (m) {
                                const int b = __ffs((int)m) - 1;
                                m &= m - 1;
                                const i64 y = (i64)wi * 32 + b;
                                if (y == s) continue;
                                const double g = gain_of(ld_agent(&dacc[y]), eix, kv, A.a[y], ax, A.constant);
                                if (A.feat) top2_push(t1, t2, (float)(g * inv2k));
                                if (better(g > 0, g, y, cand, gain, id)) {
                                    cand = true;
                                    gain = g;
                                    id = y;
                                }
                            }
                        }
                    } else {
                        e_s = hash_lookup(hkeys, hvals, mask, shift, s);
                        score_slots(hkeys, hvals, slots, threadIdx.x, kPBlock, A, s, kv, e_s - sl, ax,
                                    A.constant, cand, gain, id, t1, t2);
                    }
                } else {
                    ntouched = s_cnt;
                    if (pi == 0) {
                        e_s = acc[s - base];
                        eix = e_s - sl;
                    }
                    for (int k0 = threadIdx.x; k0 < ntouched; k0 += kGatherUnroll * kPBlock) {
                        i64 yy[kGatherUnroll], nyy[kGatherUnroll];
                        double ay[kGatherUnroll];
#pragma unroll
                        for (int u = 0; u < kGatherUnroll; u++) {
                            const int k = k0 + u * kPBlock;
                            yy[u] = k < ntouched ? (i64)list[k] : -1;
                        }
#pragma unroll
                        for (int u = 0; u < kGatherUnroll; u++) {
                            ay[u] = yy[u] >= 0 ? A.a[base + yy[u]] : 0.0;
                            nyy[u] = yy[u] >= 0 ? A.n[base + yy[u]] : 0;
                        }
#pragma unroll
                        for (int u = 0; u < kGatherUnroll; u++) {
                            if (yy[u] < 0) continue;
                            const i64 y = base + yy[u];
                            if (y == s) continue;
                            const double g = gain_of(acc[yy[u]], eix, kv, ay[u], ax, A.constant);
                            if (A.feat) top2_push(t1, t2, (float)(g * inv2k));
                            if (better(g > 0, g, y, cand, gain, id)) {
                                cand = true;
                                gain = g;
                                id = y;
                                nid = nyy[u];
                            }
                        }
                    }
                    if (pi + 1 < nsl) {  // the next slab starts from a clear view
                        __syncthreads();
                        for (int k = threadIdx.x; k < ntouched; k += kPBlock) {
                            acc[list[k]] = 0.0;
                            touched[list[k] >> 5] = 0;
                        }
                        if (threadIdx.x == 0) s_cnt = 0;
                        __syncthreads();
                    }
                }
            }
            const bool my_cand = cand;
            const i64 my_id = id;
            block_argmax_p(cand, gain, id, sk, si);
            if (A.feat) block_top2<kPBlock>(t1, t2);
            if (!HASH) {  // the size of the winner, from the lane that scored it
                if (cand && my_cand && my_id == id) s_nid = nid;
                __syncthreads();
            }
            if (first && nx_v >= 0) {  // read-ahead, last load
                const i64 j = nx_beg + threadIdx.x;
                ra_t = j < nx_end ? A.tails[j] : 0;
                ra_w = j < nx_end ? A.w[j] : 0.0;
                ra_c = c + 1;
                ra_v = nx_v;
                ra_beg = nx_beg;
                ra_end = nx_end;
                ra_s = nx_s;
                ra_k = nx_k;
                nx_done = true;
            }
            if (threadIdx.x == 0) {
                i64 target = s;
                if (cand) {
                    target = id;
                    if (ns == 1 && target > s && (HASH ? A.n[target] : s_nid) == 1) target = s;
                }
                A.tgt[v] = target;
                if (A.feat) A.feat[v] = las_record(e_s, kv, t1, t2);
                if (target != s) {  // into the other buffer, read from the next step on
                    move_vertex(target, s, kv, an, nn);
                    Cn[v] = target;
                    if (stale) stale[v] = 0;
                }
                if (first) {
                    mv_v = v;
                    mv_t = target;
                    mv_s = s;
                    mv_k = kv;
                }
            }
            __syncthreads();  // every read of the accumulator is done
            if (HASH && big && !slab) {
                for (int wi = threadIdx.x; wi < nwords; wi += kPBlock) {
                    uint32_t m = hbits[wi];
                    while (m) {
                        const int b = __ffs((int)m) - 1;
                        m &= m - 1;
                        dacc[(i64)wi * 32 + b] = 0.0;
                    }
                    hbits[wi] = 0;
                }
                __threadfence_block();
            } else if (slab) {  // back to the empty hash view
                for (int h = threadIdx.x; h < kPHashSlots; h += kPBlock) {
                    hkeys[h] = kEmpty;
                    hvals[h] = 0.0;
                }
                for (int wi = threadIdx.x; wi < nwords; wi += kPBlock) hbits[wi] = 0;
            } else if (HASH) {
                for (i64 h = threadIdx.x; h < slots; h += kPBlock) {
                    hkeys[h] = kEmpty;
                    hvals[h] = 0.0;
                }
            } else {
                for (int k = threadIdx.x; k < ntouched; k += kPBlock) {
                    acc[list[k]] = 0.0;
                    touched[list[k] >> 5] = 0;
                }
            }
            if (threadIdx.x == 0) {
                s_cnt = 0;
                s_sl = 0.0;
            }
            __syncthreads();
        }
        if (!nx_done) {  // no vertex of this step took the read-ahead along
            ra_c = -1;
            if (nx_v >= 0) {
                ra_beg = A.off[nx_v];
                ra_end = A.off[nx_v + 1];
                ra_s = C[nx_v];
                ra_k = A.k[nx_v];
                const i64 j = ra_beg + threadIdx.x;
                ra_t = j < ra_end ? A.tails[j] : 0;
                ra_w = j < ra_end ? A.w[j] : 0.0;
                ra_v = nx_v;
                ra_c = c + 1;
            }
        }
        alive = grid_barrier(bar, nb, ++seq);
        par ^= 1;
        plo = lo;
        ptot = total;
    }
    A.C = par ? C1 : C;
    A.a = par ? a1 : a;
    A.n = par ? n1 : n;
    // modularity partials: internal weight over this workgroup's vertices, sum of a_c^2
    double exx = 0, a2 = 0;
    if (alive) {
        for (i64 v = blockIdx.x; v < nv; v += nb) {
            const i64 s = A.C[v];
            if (s < 0) continue;
            for (i64 j = A.off[v] + threadIdx.x; j < A.off[v + 1]; j += kPBlock)
                if (A.C[A.tails[j]] == s) exx += A.w[j];
        }
        for (i64 i = (i64)blockIdx.x * kPBlock + threadIdx.x; i < nv; i += (i64)nb * kPBlock)
            a2 += A.a[i] * A.a[i];
    }
    exx = block_sum_p(exx, red);
    a2 = block_sum_p(a2, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = exx;
        part[nb + blockIdx.x] = a2;
    }
    if (alive && par) {  // the final state sits in the second buffer
        for (i64 i = (i64)blockIdx.x * kPBlock + threadIdx.x; i < nv; i += (i64)nb * kPBlock) {
            C[i] = C1[i];
            a[i] = a1[i];
            n[i] = n1[i];
        }
    }
}

// Dense persistent sweep (graphs of at most kDenseMaxNV vertices) with the
// color steps dealt alternately to the two halves of the grid (workgroups of
// even and of odd index).  While one half decides step k, the other gathers
// the neighbour labels of its vertex of step k+1 from the state before step
// k, setting aside the entries whose neighbour belongs to step k's class
// (the only labels step k can change); after the barrier it reads those few
// labels again and adds them, so its accumulator holds exactly the entries a
// gather after the barrier would have added (the sum order of LDS atomics is
// unordered either way).  State, buffers and the modularity partials as in
// k_colored_iteration.  Gathers read the labels from 16-bit copies (L16:
// 2 nv, the copies of C and of C1; ids are below kDenseMaxNV), 30 KB that
// stay in a CU's L1 instead of a 128-byte line of L2 per entry, and the
// gathering half tests membership in step k's class with an LDS bitmap.
constexpr int kSpecFix = 256;  // set-aside entries per vertex; more: gather after the barrier
constexpr int kSpecUnroll = 4;  // entries in flight per thread when gathering
__global__ __launch_bounds__(kPBlock) void k_colored_iteration_spec(SweepArgs A, i64 *C, double *a, i64 *n,
                                                                    const i64 *cperm, const i64 *cstart,
                                                                    int ncol, unsigned *bar, double *part,
                                                                    uint8_t *stale, i64 *C1, double *a1,
                                                                    i64 *n1, uint16_t *L16) {
    __shared__ __align__(16) unsigned char smem[kPDenseBytes];
    __shared__ double red[kPBlock / kWave];
    __shared__ u64 sk[kPBlock / kWave];
    __shared__ i64 si[kPBlock / kWave];
    __shared__ i64 sn[kPBlock / kWave];
    __shared__ int s_cnt;                // communities listed for the gathered vertex
    __shared__ double s_sl;              // its self-loop weight
    __shared__ int s_fn;                 // set-aside entries (above kSpecFix: too many)
    __shared__ uint16_t s_ft[kSpecFix];  // their neighbours
    __shared__ uint32_t cbits[kDenseMaxNV / 32];  // gathering half: the vertices of step k's class
    __shared__ double s_fw[kSpecFix];    // and weights
    double *acc = reinterpret_cast<double *>(smem);
    uint32_t *touched = reinterpret_cast<uint32_t *>(smem + kDenseMaxNV * 8);
    uint16_t *list = reinterpret_cast<uint16_t *>(smem + kDenseMaxNV * 8 + kDenseMaxNV / 32 * 4);
    const int nv = (int)A.nv;
    const int nwords = (nv + 31) / 32;
    const unsigned nb = gridDim.x;
    const unsigned half = blockIdx.x & 1u, nbh = nb / 2, wid = blockIdx.x >> 1;
    const int lane = threadIdx.x % kWave;
    const u64 lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    unsigned seq = 0;
    if (threadIdx.x == 0) {
        s_cnt = 0;
        s_sl = 0.0;
        s_fn = 0;
    }
    for (int i = threadIdx.x; i < nv; i += kPBlock) acc[i] = 0.0;
    for (int i = threadIdx.x; i < nwords; i += kPBlock) {
        touched[i] = 0;
        cbits[i] = 0;
    }
    for (i64 i = (i64)blockIdx.x * kPBlock + threadIdx.x; i < nv; i += (i64)nb * kPBlock) {
        C1[i] = C[i];
        a1[i] = a[i];
        n1[i] = n[i];
        L16[i] = L16[nv + i] = (uint16_t)C[i];
    }
    // community cc gets weight ww; the first lane to touch it lists it (one LDS
    // counter update per wavefront); called by whole wavefronts
    auto insert = [&](bool valid, i64 cc, double ww) {
        bool fresh = false;
        if (valid) {
            const uint32_t bit = 1u << (cc & 31);
            fresh = !(atomicOr(&touched[cc >> 5], bit) & bit);
        }
        const u64 fm = __ballot(fresh);
        if (fm) {
            const int leader = __ffsll((long long)fm) - 1;
            int at = 0;
            if (lane == leader) at = atomicAdd(&s_cnt, __popcll(fm));
            at = __shfl(at, leader);
            if (fresh) list[at + __popcll(fm & lt)] = (uint16_t)cc;
        }
        if (valid) atomicAdd(&acc[cc], ww);
    };
    // empties the dense view after a vertex (ntouched listed communities)
    auto clear = [&](int ntouched) {
        for (int k = threadIdx.x; k < ntouched; k += kPBlock) {
            acc[list[k]] = 0.0;
            touched[list[k] >> 5] = 0;
        }
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_sl = 0.0;
            s_fn = 0;
        }
    };
    bool alive = grid_barrier(bar, nb, ++seq);
    int par = 0;            // buffer holding the state after the previous step; also the deciding half
    i64 plo = 0, ptot = 0;  // the previous step's class
    // the vertex gathered ahead by this workgroup: its class, id, entries, label, degree
    int sp_c = -1;
    i64 sp_v = -1, sp_beg = 0, sp_end = 0, sp_s = -1;
    double sp_k = 0;
    // thread 0: the move made for this workgroup's first vertex of the previous step
    i64 mv_v = -1, mv_t = -1, mv_s = -1;
    double mv_k = 0;
    for (int c = 0; c < ncol && alive; c++) {
        const i64 lo = cstart[c], hi = cstart[c + 1], total = hi - lo;
        if (total == 0) continue;
        A.C = par ? C1 : C;
        A.a = par ? a1 : a;
        A.n = par ? n1 : n;
        i64 *Cn = par ? C : C1;
        double *an = par ? a : a1;
        i64 *nn = par ? n : n1;
        const uint16_t *Lr = L16 + (par ? nv : 0);  // the labels of A.C
        uint16_t *Lw = L16 + (par ? 0 : nv);        // and of Cn
        const bool deciding = half == (unsigned)par;
        // decides vertex q of class c (into the other buffer); the first
        // vertex of a deciding workgroup (q == wid) may be gathered ahead
        auto decide = [&](i64 q, bool first) {
            const bool sp = first && sp_c == c;
            const i64 v = sp ? sp_v : cperm[lo + q];
            const i64 beg = sp ? sp_beg : A.off[v], end = sp ? sp_end : A.off[v + 1];
            const i64 s = sp ? sp_s : A.C[v];
            if (s < 0) {  // a vertex without entries (label -1) stays where it is
                if (threadIdx.x == 0) A.tgt[v] = s;
                return;
            }
            const double as = A.a[s];
            i64 ns = 0;
            if (threadIdx.x == 0) ns = A.n[s];
            if (sp) {
                // the set-aside entries with the labels after the previous step
                const int nf = s_fn;
                for (int i0 = 0; i0 < nf; i0 += kPBlock) {
                    const int i = i0 + threadIdx.x;
                    const bool valid = i < nf;
                    const i64 t = valid ? (i64)s_ft[i] : 0;
                    const double ww = valid ? s_fw[i] : 0.0;
                    const i64 cc = valid ? (i64)Lr[t] : 0;
                    insert(valid, cc, ww);
                }
            } else {
                for (i64 j0 = beg + threadIdx.x; j0 < end; j0 += kGatherUnroll * kPBlock) {
                    i64 tt[kGatherUnroll], cl[kGatherUnroll];
                    double wt[kGatherUnroll];
#pragma unroll
                    for (int u = 0; u < kGatherUnroll; u++) {
                        const i64 j = j0 + (i64)u * kPBlock;
                        tt[u] = j < end ? A.tails[j] : -1;
                        wt[u] = j < end ? A.w[j] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < kGatherUnroll; u++) cl[u] = tt[u] >= 0 ? (i64)Lr[tt[u]] : 0;
#pragma unroll
                    for (int u = 0; u < kGatherUnroll; u++) {
                        const bool valid = tt[u] >= 0;
                        if (valid && tt[u] == v) atomicAdd(&s_sl, wt[u]);
                        insert(valid, cl[u], wt[u]);
                    }
                }
            }
            __syncthreads();
            const double sl = s_sl;
            const double kv = sp ? sp_k : A.k[v];
            const double ax = as - kv;
            const double inv2k = A.feat ? 1.0 / (2.0 * kv) : 0.0;
            const int ntouched = s_cnt;
            const double e_s = acc[s];
            const double eix = e_s - sl;
            bool cand = false;
            double gain = 0;
            i64 id = INT64_MAX;
            i64 nid = 0;
            float t1 = -INFINITY, t2 = -INFINITY;
            for (int k0 = threadIdx.x; k0 < ntouched; k0 += kGatherUnroll * kPBlock) {
                i64 yy[kGatherUnroll], nyy[kGatherUnroll];
                double ay[kGatherUnroll];
#pragma unroll
                for (int u = 0; u < kGatherUnroll; u++) {
                    const int k = k0 + u * kPBlock;
                    yy[u] = k < ntouched ? (i64)list[k] : -1;
                }
#pragma unroll
                for (int u = 0; u < kGatherUnroll; u++) {
                    ay[u] = yy[u] >= 0 ? A.a[yy[u]] : 0.0;
                    nyy[u] = yy[u] >= 0 ? A.n[yy[u]] : 0;
                }
#pragma unroll
                for (int u = 0; u < kGatherUnroll; u++) {
                    const i64 y = yy[u];
                    if (y < 0 || y == s) continue;
                    const double g = gain_of(acc[y], eix, kv, ay[u], ax, A.constant);
                    if (A.feat) top2_push(t1, t2, (float)(g * inv2k));
                    if (better(g > 0, g, y, cand, gain, id)) {
                        cand = true;
                        gain = g;
                        id = y;
                        nid = nyy[u];
                    }
                }
            }
            block_argmax_pn(cand, gain, id, nid, sk, si, sn);
            if (A.feat) block_top2<kPBlock>(t1, t2);
            if (threadIdx.x == 0) {
                i64 target = s;
                if (cand) {
                    target = id;
                    if (ns == 1 && target > s && nid == 1) target = s;
                }
                A.tgt[v] = target;
                if (A.feat) A.feat[v] = las_record(e_s, kv, t1, t2);
                if (target != s) {  // into the other buffer, read from the next step on
                    move_vertex(target, s, kv, an, nn);
                    Cn[v] = target;
                    Lw[v] = (uint16_t)target;
                    if (stale) stale[v] = 0;
                }
                if (first) {
                    mv_v = v;
                    mv_t = target;
                    mv_s = s;
                    mv_k = kv;
                }
            }
            __syncthreads();  // every read of the accumulator is done
            clear(ntouched);
            __syncthreads();
        };
        if (!deciding) {
            // this workgroup decided the previous step: its moves, into the other buffer
            if (threadIdx.x == 0 && mv_v >= 0 && mv_t != mv_s) {
                move_vertex(mv_t, mv_s, mv_k, an, nn);
                Cn[mv_v] = mv_t;
                Lw[mv_v] = (uint16_t)mv_t;
            }
            // and the previous class's other vertices (q >= nbh, whoever decided them), one per thread
            for (i64 q = (i64)wid + nbh * (1 + (i64)threadIdx.x); q < ptot; q += nbh * (i64)kPBlock) {
                const i64 v = cperm[plo + q];
                const i64 t = A.tgt[v], s = Cn[v];
                if (t != s) {
                    move_vertex(t, s, A.k[v], an, nn);
                    Cn[v] = t;
                    Lw[v] = (uint16_t)t;
                }
            }
            mv_v = -1;
        }
        // this step's vertices: the deciding half takes q = wid (its first,
        // maybe gathered ahead) and q = nbh + wid + k nb, the other half (in
        // classes larger than a half) q = nb + wid + k nb
        {
            bool first = deciding;
            for (i64 q = deciding ? (i64)wid : (i64)nb + wid; q < total;) {
                decide(q, first);
                q = first ? (i64)nbh + wid : q + nb;
                first = false;
            }
        }
        if (!deciding) {
            // gather ahead: this workgroup's first vertex of the next non-empty class
            int c2 = c + 1;
            i64 lo2 = hi, hi2 = hi;
            while (c2 < ncol) {
                hi2 = cstart[c2 + 1];
                if (hi2 > lo2) break;
                lo2 = hi2;
                c2++;
            }
            sp_c = -1;
            if (c2 < ncol && (i64)wid < hi2 - lo2) {
                const i64 v = cperm[lo2 + wid];
                const i64 beg = A.off[v], end = A.off[v + 1];
                const i64 s = A.C[v];  // changes only in v's own step
                if (s >= 0) {
                    // this step's class, whose labels change, into the LDS bitmap
                    for (i64 q = threadIdx.x; q < total; q += kPBlock) {
                        const i64 u = cperm[lo + q];
                        atomicOr(&cbits[u >> 5], 1u << (u & 31));
                    }
                    __syncthreads();
                    for (i64 j0 = beg + threadIdx.x; j0 < end; j0 += kSpecUnroll * kPBlock) {
                        int32_t tt[kSpecUnroll], cl[kSpecUnroll];
                        double wt[kSpecUnroll];
#pragma unroll
                        for (int u = 0; u < kSpecUnroll; u++) {
                            const i64 j = j0 + (i64)u * kPBlock;
                            tt[u] = j < end ? (int32_t)A.tails[j] : -1;
                            wt[u] = j < end ? A.w[j] : 0.0;
                        }
#pragma unroll
                        for (int u = 0; u < kSpecUnroll; u++) cl[u] = tt[u] >= 0 ? (int32_t)Lr[tt[u]] : 0;
#pragma unroll
                        for (int u = 0; u < kSpecUnroll; u++) {
                            const bool valid = tt[u] >= 0;
                            if (valid && tt[u] == v) atomicAdd(&s_sl, wt[u]);
                            const bool aside = valid && tt[u] != v && ((cbits[tt[u] >> 5] >> (tt[u] & 31)) & 1u);
                            if (aside) {
                                const int at = atomicAdd(&s_fn, 1);
                                if (at < kSpecFix) {
                                    s_ft[at] = (uint16_t)tt[u];
                                    s_fw[at] = wt[u];
                                }
                            }
                            insert(valid && !aside, cl[u], wt[u]);
                        }
                    }
                    __syncthreads();
                    // every wavefront reads the counts before thread 0 may reset them in clear()
                    const int fn = s_fn, cn = s_cnt;
                    for (int i = threadIdx.x; i < nwords; i += kPBlock) cbits[i] = 0;
                    __syncthreads();
                    if (fn > kSpecFix) {  // too many set aside: gather again after the barrier
                        clear(cn);
                    } else {
                        sp_c = c2;
                        sp_v = v;
                        sp_beg = beg;
                        sp_end = end;
                        sp_s = s;
                        sp_k = A.k[v];
                    }
                    __syncthreads();
                }
            }
        } else {
            sp_c = -1;
        }
        alive = grid_barrier(bar, nb, ++seq);
        par ^= 1;
        plo = lo;
        ptot = total;
    }
    A.C = par ? C1 : C;
    A.a = par ? a1 : a;
    A.n = par ? n1 : n;
    // modularity partials: internal weight over this workgroup's vertices, sum of a_c^2
    double exx = 0, a2 = 0;
    if (alive) {
        for (i64 v = blockIdx.x; v < nv; v += nb) {
            const i64 s = A.C[v];
            if (s < 0) continue;
            for (i64 j = A.off[v] + threadIdx.x; j < A.off[v + 1]; j += kPBlock)
                if (A.C[A.tails[j]] == s) exx += A.w[j];
        }
        for (i64 i = (i64)blockIdx.x * kPBlock + threadIdx.x; i < nv; i += (i64)nb * kPBlock)
            a2 += A.a[i] * A.a[i];
    }
    exx = block_sum_p(exx, red);
    a2 = block_sum_p(a2, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = exx;
        part[nb + blockIdx.x] = a2;
    }
    if (alive && par) {  // the final state sits in the second buffer
        for (i64 i = (i64)blockIdx.x * kPBlock + threadIdx.x; i < nv; i += (i64)nb * kPBlock) {
            C[i] = C1[i];
            a[i] = a1[i];
            n[i] = n1[i];
        }
    }
}

__global__ void k_max_degree(const i64 *off, const i64 *verts, i64 count, i64 *out) {
    i64 m = 0;
    for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q < count; q += (i64)gridDim.x * blockDim.x) {
        const i64 v = verts[q];
        const i64 d = off[v + 1] - off[v];
        m = d > m ? d : m;
    }
    if (m) atomicMax((unsigned long long *)out, (unsigned long long)m);
}

__global__ void k_isolated_targets(const i64 *verts, i64 count, i64 *tgt) {
    for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q < count; q += (i64)gridDim.x * blockDim.x)
        tgt[verts[q]] = -1;
}


// Commit of one color step: community updates of the movers, then labels.
__global__ void k_apply_colored(const i64 *perm, Slices S, const i64 *tgt, i64 *C, const double *k,
                                double *a, i64 *n, uint8_t *stale, const double *dein, double *dpart) {
    i64 total = 0;
    for (int b = 0; b < kNumBins; b++) total += S.hi[b] - S.lo[b];
    double d = 0;
    for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q < total; q += (i64)gridDim.x * blockDim.x) {
        i64 r = q;
        int b = 0;
        while (r >= S.hi[b] - S.lo[b]) {
            r -= S.hi[b] - S.lo[b];
            b++;
        }
        const i64 v = perm[S.lo[b] + r];
        const i64 t = tgt[v], s = C[v];
        if (t != s) {
            move_vertex(t, s, k[v], a, n);
            C[v] = t;
            if (stale) stale[v] = 0;
            if (dein) d += dein[v];
        }
    }
    if (dpart) {
        d = block_sum(d);
        if (threadIdx.x == 0) dpart[blockIdx.x] = d;
    }
}
// *dsum += the n partials, in a fixed order (one workgroup)
__global__ __launch_bounds__(kBlock) void k_add_partials(const double *part, i64 n, double *dsum) {
    double x = 0;
    for (i64 i = threadIdx.x; i < n; i += kBlock) x += part[i];
    x = block_sum(x);
    if (threadIdx.x == 0) *dsum += x;
}
// cInfo += cUpdate of a Jacobi iteration (parallelLouvainMethod.cpp:223-236, :275-279).
__global__ void k_apply_plain(i64 nv, const i64 *tgt, const i64 *curr, const double *k, double *a,
                              i64 *n) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        const i64 t = tgt[v], s = curr[v];
        if (t != s && t != -1) move_vertex(t, s, k[v], a, n);
    }
}

// initCommAssOpt (utilityClusteringFunctions.cpp:108-169): one thread per
// vertex, quirks kept (integer-truncated self-loop in selfLoop, last
// self-loop weight as the own-cluster counter, duplicates not merged).
__global__ void k_init_comm_ass_opt(const i64 *off, const i64 *tails, const double *w, i64 nv,
                                    const double *k, const double *a, const i64 *n, double constant0,
                                    const int32_t *gid, const double *gconst, i64 *curr) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        const double constant = gid ? gconst[gid[v]] : constant0;
        double selfLoop = 0, cself = 0;
        for (i64 j = off[v]; j < off[v + 1]; j++)
            if (tails[j] == v) {
                selfLoop += (double)(long)w[j];
                cself = w[j];
            }
        i64 maxIndex = v;
        double maxGain = 0;
        const double eix = cself - selfLoop, ax = a[v] - k[v];
        for (i64 j = off[v]; j < off[v + 1]; j++) {
            i64 cid = tails[j];
            if (cid == v) continue;
            double g = gain_of(w[j], eix, k[v], a[cid], ax, constant);
            if ((g > maxGain) || ((g == maxGain) && (g != 0) && (cid < maxIndex))) {
                maxGain = g;
                maxIndex = cid;
            }
        }
        if (n[maxIndex] == 1 && n[v] == 1 && maxIndex > v) maxIndex = v;
        curr[v] = maxIndex;
    }
}
// updateAxForOpt (utilityClusteringFunctions.cpp:46-67).
__global__ void k_update_ax_opt(i64 nv, const i64 *curr, const double *k, double *a, i64 *n) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        if (curr[v] != v) move_vertex(curr[v], v, k[v], a, n);
}

// ----------------------------------------------------------------- modularity
// sum over vertices of the weight to their own community, G lanes per vertex.
template <int G>
__global__ __launch_bounds__(kBlock) void k_mod_pass(const i64 *off, const i64 *tails, const double *w,
                                                  const i64 *C, const i64 *verts, i64 count,
                                                  double *part) {
    constexpr int VPB = kBlock / G;
    constexpr int U = 4;  // vertices of a group in flight together
    const int lane = threadIdx.x % G, grp = threadIdx.x / G;
    const i64 stride = (i64)gridDim.x * VPB;
    double acc = 0;
    for (i64 q0 = (i64)blockIdx.x * VPB + grp; q0 < count; q0 += U * stride) {
        i64 vv[U], ss[U], bb[U], ee[U], tt[U];
        double ww[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const i64 q = q0 + u * stride;
            vv[u] = q < count ? verts[q] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            ss[u] = vv[u] >= 0 ? C[vv[u]] : -1;
            bb[u] = vv[u] >= 0 ? off[vv[u]] : 0;
            ee[u] = vv[u] >= 0 ? off[vv[u] + 1] : 0;
        }
        // the first pass of every vertex, loads first
#pragma unroll
        for (int u = 0; u < U; u++) {
            const i64 j = bb[u] + lane;
            const bool in = ss[u] >= 0 && j < ee[u];
            tt[u] = in ? tails[j] : -1;
            ww[u] = in ? w[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const i64 ct = tt[u] >= 0 ? C[tt[u]] : -2;
            if (ct == ss[u]) acc += ww[u];
        }
        // vertices of more than G entries
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (ss[u] < 0) continue;
            for (i64 j = bb[u] + lane + G; j < ee[u]; j += G)
                if (C[tails[j]] == ss[u]) acc += w[j];
        }
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Internal weight of giant vertices: kGiantBlocks workgroups per vertex.
__global__ __launch_bounds__(kBlock) void k_mod_giant(const i64 *off, const i64 *tails, const double *w,
                                                   const i64 *C, const i64 *verts, double *part) {
    const i64 q = blockIdx.y;
    const i64 v = verts[q];
    const i64 s = C[v];
    double acc = 0;
    if (s >= 0)
        for (i64 j = off[v] + (i64)blockIdx.x * kBlock + threadIdx.x; j < off[v + 1];
             j += (i64)gridDim.x * kBlock)
            if (C[tails[j]] == s) acc += w[j];
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[q * gridDim.x + blockIdx.x] = acc;
}

// ------------------------------------------------------------------- coloring

// Appends `v` to list[] when `pred`, one atomic per wavefront.
__device__ inline void wave_append(bool pred, i64 v, i64 *list, i64 *count) {
    const u64 m = __ballot(pred);
    if (!m) return;
    const int lane = threadIdx.x % kWave;
    const int leader = __ffsll((long long)m) - 1;
    i64 base = 0;
    if (lane == leader) base = (i64)atomicAdd((unsigned long long *)count, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    const u64 lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    if (pred) list[base + __popcll(m & lt)] = v;
}

// Workgroup-local staging of list appends.  Device-scope atomics on one
// counter serialise (about 12 ns each on MI355X, whatever the caller), so a
// round that appends from every wavefront is bound by the counter; staged
// appends take an LDS atomic per wavefront and stage_flush moves the staged
// entries with one device atomic per workgroup.  Entries that do not fit go
// straight to the list.  The order of a list is not significant to its users.
template <int CAP>
struct Stage {
    i64 buf[CAP];
    int n;
    i64 base;
};
template <int CAP>
__device__ inline void stage_init(Stage<CAP> &S) {
    if (threadIdx.x == 0) S.n = 0;
}
template <int CAP>
__device__ inline void stage_append(bool pred, i64 v, Stage<CAP> &S, i64 *list, i64 *count) {
    const u64 m = __ballot(pred);
    if (!m) return;
    const int lane = threadIdx.x % kWave;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&S.n, __popcll(m));
    base = __shfl(base, leader);
    const u64 lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    const int at = base + __popcll(m & lt);
    const bool fits = at < CAP;
    if (pred && fits) S.buf[at] = v;
    wave_append(pred && !fits, v, list, count);
}
// Called by every thread of the workgroup once its appends are done; leaves
// the stage empty.
template <int CAP>
__device__ inline void stage_flush(Stage<CAP> &S, i64 *list, i64 *count) {
    __syncthreads();
    const int n = S.n < CAP ? S.n : CAP;
    if (threadIdx.x == 0 && n > 0) S.base = (i64)atomicAdd((unsigned long long *)count, (unsigned long long)n);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) list[S.base + i] = S.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) S.n = 0;
    __syncthreads();
}

// Jones-Plassmann as a topological sweep of the priority DAG: count[v] is the
// number of uncoloured neighbours with a higher priority; a vertex enters the
// frontier of round r when its count reaches zero, which is exactly when round
// r of the classic formulation would find it a local maximum.  Each frontier
// vertex takes the smallest color absent from its (all coloured) higher
// neighbours and releases its lower neighbours, so total work is O(nnz).
__device__ inline void frontier_push(bool pred, i64 u, const i64 *off, const JpLists &L) {
    const i64 d = pred ? off[u + 1] - off[u] : 0;
    const bool big = pred && d > kBinMaxDeg[6];
    const bool giant = big && d > kGiantDeg;
    wave_append(pred && !big, u, L.small, &L.counts[0]);
    wave_append(big && !giant, u, L.large, &L.counts[1]);
    wave_append(giant, u, L.giant, &L.counts[2]);
}
// The same with the small list staged in S (the other two are rare) and the
// degree class of u from its entry's code.
template <int CAP>
__device__ inline void frontier_push(bool pred, i64 u, uint8_t code, const JpLists &L, Stage<CAP> &S) {
    const bool big = pred && (code & kJpLarge);
    const bool giant = pred && (code & kJpGiant);
    stage_append(pred && !big, u, S, L.small, &L.counts[0]);
    wave_append(big && !giant, u, L.large, &L.counts[1]);
    wave_append(giant, u, L.giant, &L.counts[2]);
}
constexpr int kStageCap = 2048;
// vertices of at most this many entries are counted by one thread (a serial
// loop of dependent loads); larger ones by a wavefront
constexpr i64 kJpCountThreadDeg = 32;

// One wavefront colours up to eight frontier entries front[lo .. lo+n): a
// vertex of at most kJpLightDeg entries takes a group of kJpGroup lanes (its
// first-fit colour is at most its entry count, so a 128-colour mask per lane,
// OR-ed over the group, decides it); heavier ones (up to 256 entries) take
// the whole wave one after another, with the wave's LDS bitmap mk.  One pass
// over a vertex's entries marks the colours and depths of its higher-priority
// neighbours (all coloured) and releases the lower-priority ones; frontier
// vertices are pairwise non-adjacent, so every colour read is final.
// push(ready, u) appends a released vertex (called by the whole wave).
constexpr int kJpGroup = 8;
constexpr int kJpUnroll = 4;
constexpr i64 kJpLightDeg = 64;
template <typename Push>
__device__ inline void jp_wave8(const i64 *off, const i64 *tails, int32_t *colors, int32_t *dep, int32_t *cnt,
                                const uint8_t *code, const i64 *front, i64 lo, i64 n, uint32_t *mk, int32_t *err,
                                Push push) {
    const int lane = threadIdx.x % kWave, g = lane / kJpGroup, gl = lane % kJpGroup;
    const bool in = g < n;
    const i64 v = in ? front[lo + g] : -1;
    const i64 b = in ? off[v] : 0, deg = in ? off[v + 1] - b : 0;
    const bool light = in && deg <= kJpLightDeg;
    u64 m0 = 0, m1 = 0;
    int dm = 0;
    i64 itmax = light ? (deg + kJpGroup - 1) / kJpGroup : 0;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const i64 x = __shfl_xor(itmax, o);
        itmax = x > itmax ? x : itmax;
    }
    // kJpUnroll iterations at a time: entry and code loads, then the count
    // updates and colour loads, then the appends
    for (i64 it0 = 0; it0 < itmax; it0 += kJpUnroll) {
        i64 uu[kJpUnroll];
        uint8_t cd[kJpUnroll];
        bool rd[kJpUnroll];
#pragma unroll
        for (int k = 0; k < kJpUnroll; k++) {
            const i64 j = (it0 + k) * kJpGroup + gl;
            const bool e = light && j < deg;
            uu[k] = e ? tails[b + j] : -1;
            cd[k] = e ? code[b + j] : kJpSelf;
        }
#pragma unroll
        for (int k = 0; k < kJpUnroll; k++) {
            rd[k] = false;
            const i64 u = uu[k];
            if (cd[k] & kJpSelf) continue;
            if (cd[k] & kJpRelease) {
                rd[k] = atomicSub(&cnt[u], 1) == 1;
            } else {
                const int c = colors[u];
                const int du = dep[u];
                dm = du > dm ? du : dm;
                if (c >= 0 && c < 64) m0 |= 1ull << c;
                else if (c >= 64 && c < 128) m1 |= 1ull << (c - 64);
            }
        }
#pragma unroll
        for (int k = 0; k < kJpUnroll; k++) push(rd[k], uu[k], cd[k]);
    }
#pragma unroll
    for (int o = 1; o < kJpGroup; o <<= 1) {
        m0 |= __shfl_xor(m0, o);
        m1 |= __shfl_xor(m1, o);
        const int x = __shfl_xor(dm, o);
        dm = x > dm ? x : dm;
    }
    if (light && gl == 0) {
        colors[v] = ~m0 ? __ffsll((long long)~m0) - 1 : 64 + __ffsll((long long)~m1) - 1;
        dep[v] = dm + 1;
    }
    u64 hm = __ballot(in && !light && gl == 0);
    while (hm) {
        const int src = __ffsll((long long)hm) - 1;
        hm &= hm - 1;
        const i64 hv = __shfl(v, src), hb = __shfl(b, src), hd = __shfl(deg, src);
        for (int i = lane; i < kMaxColors / 32; i += kWave) mk[i] = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        int hdm = 0;
        for (i64 j0 = 0; j0 < hd; j0 += kJpUnroll * kWave) {
            i64 uu[kJpUnroll];
            uint8_t cd[kJpUnroll];
            bool rd[kJpUnroll];
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) {
                const i64 j = j0 + k * kWave + lane;
                uu[k] = j < hd ? tails[hb + j] : -1;
                cd[k] = j < hd ? code[hb + j] : kJpSelf;
            }
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) {
                rd[k] = false;
                const i64 u = uu[k];
                if (cd[k] & kJpSelf) continue;
                if (cd[k] & kJpRelease) {
                    rd[k] = atomicSub(&cnt[u], 1) == 1;
                } else {
                    const int c = colors[u];
                    const int du = dep[u];
                    hdm = du > hdm ? du : hdm;
                    if (c >= 0) atomicOr(&mk[c >> 5], 1u << (c & 31));
                }
            }
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) push(rd[k], uu[k], cd[k]);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        int cand = kMaxColors;
        for (int i = lane; i < kMaxColors / 32; i += kWave) {
            const uint32_t m = mk[i];
            if (m != 0xFFFFFFFFu) {
                const int c = i * 32 + __ffs((int)~m) - 1;
                cand = c < cand ? c : cand;
            }
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const int other = __shfl_xor(cand, o);
            cand = other < cand ? other : cand;
            const int x = __shfl_xor(hdm, o);
            hdm = x > hdm ? x : hdm;
        }
        if (lane == 0) {
            int color = cand;
            if (color >= kMaxColors) {
                atomicOr(err, 1);
                color = 0;
            }
            colors[hv] = color;
            dep[hv] = hdm + 1;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// count[v] of vertices up to kJpCountThreadDeg entries, one thread each;
// heavier ones are listed in heavy[] (heavy_count) for k_jp_count_heavy.  A core vertex v
// (count[v] = its core index q on entry) gets its higher and lower core
// neighbours counted into core_cnt[2q], [2q+1] instead, and never enters a
// frontier: the core pass colours it.  A non-core vertex's higher core
// neighbours are coloured before the rounds and do not count.
__global__ __launch_bounds__(kBlock) void k_jp_count(const i64 *off, const i64 *tails, i64 nv, int32_t *cnt,
                                                     JpLists L, i64 *heavy, i64 *heavy_count) {
    __shared__ Stage<kStageCap> S;
    stage_init(S);
    __syncthreads();
    const i64 stride = (i64)gridDim.x * blockDim.x;
    const i64 first = blockIdx.x * (i64)blockDim.x + threadIdx.x;
    const i64 trips = (nv + stride - 1) / stride;
    for (i64 t = 0; t < trips; t++) {
        if (t > 0) {  // room for this trip's appends (a workgroup-uniform decision)
            __syncthreads();
            const int staged = S.n;
            __syncthreads();
            if (staged > kStageCap - kBlock) stage_flush(S, L.small, &L.counts[0]);
        }
        const i64 v = first + t * stride;
        int c = 0;
        bool heavy_v = false, core_v = false;
        if (v < nv) {
            const i64 b = off[v], e = off[v + 1];
            const u64 pv = L.prio[v];
            core_v = (pv >> 40) >= L.dcore;
            if (e - b > kJpCountThreadDeg) {
                heavy_v = true;
            } else {
                int hi = 0, lo = 0;
                for (i64 j = b; j < e; j++) {
                    const i64 u = tails[j];
                    const uint8_t cd = jp_code(L.prio[u], u, pv, v, L.dcore);
                    L.code[j] = cd;
                    if (cd & kJpSelf) continue;
                    if (cd & kJpCore) {
                        if (cd & kJpRelease) lo++;
                        else hi++;
                    } else if (!(cd & kJpRelease)) {
                        c++;
                    }
                }
                if (core_v) {
                    const int q = cnt[v];
                    L.core_cnt[2 * q] = hi;
                    L.core_cnt[2 * q + 1] = lo;
                } else {
                    cnt[v] = c;
                }
            }
        }
        wave_append(heavy_v, v, heavy, heavy_count);
        stage_append(v < nv && !heavy_v && !core_v && c == 0, v, S, L.small, &L.counts[0]);  // a small-list vertex
    }
    stage_flush(S, L.small, &L.counts[0]);
}
// count[v] of the heavy vertices up to kJpCountWaveDeg entries, one
// wavefront each (larger ones: k_jp_count_giant)
constexpr i64 kJpCountWaveDeg = 4096;
__global__ __launch_bounds__(kBlock) void k_jp_count_heavy(const i64 *off, const i64 *tails, int32_t *cnt,
                                                           const i64 *heavy, const i64 *heavy_count,
                                                           JpLists L, i64 *giant, i64 *giant_count) {
    constexpr int WPB = kBlock / kWave;
    const i64 nheavy = *heavy_count;
    const int lane = threadIdx.x % kWave;
    const i64 stride = (i64)gridDim.x * WPB;
    for (i64 q = (i64)blockIdx.x * WPB + threadIdx.x / kWave; q < nheavy; q += stride) {
        const i64 v = heavy[q];
        if (off[v + 1] - off[v] > kJpCountWaveDeg) {  // wave-uniform
            wave_append(lane == 0, v, giant, giant_count);
            continue;
        }
        const u64 pv = L.prio[v];
        const bool core_v = (pv >> 40) >= L.dcore;
        int c = 0, hi = 0, lo = 0;
        for (i64 j = off[v] + lane; j < off[v + 1]; j += kWave) {
            const i64 u = tails[j];
            const uint8_t cd = jp_code(L.prio[u], u, pv, v, L.dcore);
            L.code[j] = cd;
            if (cd & kJpSelf) continue;
            if (cd & kJpCore) {
                if (cd & kJpRelease) lo++;
                else hi++;
            } else if (!(cd & kJpRelease)) {
                c++;
            }
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            c += __shfl_xor(c, o);
            hi += __shfl_xor(hi, o);
            lo += __shfl_xor(lo, o);
        }
        if (core_v) {
            if (lane == 0) {
                const int cq = cnt[v];
                L.core_cnt[2 * cq] = hi;
                L.core_cnt[2 * cq + 1] = lo;
            }
        } else {
            if (lane == 0) cnt[v] = c;
            frontier_push(lane == 0 && c == 0, v, off, L);
        }
    }
}

// count[v] of the heavy vertices above kJpCountWaveDeg entries (listed by
// k_jp_count_heavy), one workgroup each (kJpUnroll passes of loads in flight)
__global__ __launch_bounds__(kBlock) void k_jp_count_giant(const i64 *off, const i64 *tails, int32_t *cnt,
                                                           const i64 *heavy, const i64 *heavy_count,
                                                           JpLists L) {
    __shared__ int s_c[3][kBlock / kWave];
    const i64 nheavy = *heavy_count;
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    for (i64 q = blockIdx.x; q < nheavy; q += gridDim.x) {
        const i64 v = heavy[q];
        const i64 b = off[v], deg = off[v + 1] - b;
        const u64 pv = L.prio[v];
        const bool core_v = (pv >> 40) >= L.dcore;
        int c = 0, hi = 0, lo = 0;
        for (i64 j0 = 0; j0 < deg; j0 += kJpUnroll * kBlock) {
            i64 uu[kJpUnroll];
            u64 pu[kJpUnroll];
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) {
                const i64 j = j0 + k * kBlock + threadIdx.x;
                uu[k] = j < deg ? tails[b + j] : -1;
            }
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) pu[k] = uu[k] >= 0 ? L.prio[uu[k]] : 0;
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) {
                if (uu[k] < 0) continue;
                const uint8_t cd = jp_code(pu[k], uu[k], pv, v, L.dcore);
                L.code[b + j0 + k * kBlock + threadIdx.x] = cd;
                if (cd & kJpSelf) continue;
                if (cd & kJpCore) {
                    if (cd & kJpRelease) lo++;
                    else hi++;
                } else if (!(cd & kJpRelease)) {
                    c++;
                }
            }
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            c += __shfl_xor(c, o);
            hi += __shfl_xor(hi, o);
            lo += __shfl_xor(lo, o);
        }
        if (lane == 0) {
            s_c[0][w] = c;
            s_c[1][w] = hi;
            s_c[2][w] = lo;
        }
        __syncthreads();
        if (w == 0) {
            c = hi = lo = 0;
            for (int i = 0; i < kBlock / kWave; i++) {
                c += s_c[0][i];
                hi += s_c[1][i];
                lo += s_c[2][i];
            }
            if (core_v) {
                if (lane == 0) {
                    const int cq = cnt[v];
                    L.core_cnt[2 * cq] = hi;
                    L.core_cnt[2 * cq + 1] = lo;
                }
            } else {
                if (lane == 0) cnt[v] = c;
                frontier_push(lane == 0 && c == 0, v, off, L);
            }
        }
        __syncthreads();
    }
}

// Frontier vertices of degree <= 256, one wavefront each (a private LDS color
// bitmap per wavefront); whole workgroups iterate together for the barriers.
__global__ __launch_bounds__(kBlock) void k_jp_frontier_wave(const i64 *off, const i64 *tails,
                                                             int32_t *colors, int32_t *cnt,
                                                             const i64 *front, i64 nfront, JpLists L,
                                                             int32_t *err) {
    constexpr int WPB = kBlock / kWave;
    constexpr i64 E = kWave / kJpGroup;  // entries per wavefront
    __shared__ uint32_t mark[WPB][kMaxColors / 32];
    __shared__ Stage<kStageCap> S;
    stage_init(S);
    __syncthreads();
    const int w = threadIdx.x / kWave;
    const i64 nw = (i64)gridDim.x * WPB, wid = (i64)blockIdx.x * WPB + w;
    auto push = [&](bool ready, i64 u, uint8_t cd) { frontier_push(ready, u, cd, L, S); };
    for (i64 lo = wid * E; lo < nfront; lo += nw * E) {
        const i64 n = nfront - lo < E ? nfront - lo : E;
        jp_wave8(off, tails, colors, L.dep, cnt, L.code, front, lo, n, mark[w], err, push);
    }
    stage_flush(S, L.small, &L.counts[0]);
}

// Frontier vertices of degree > 256, one workgroup each.
__global__ __launch_bounds__(kBlock) void k_jp_frontier_block(const i64 *off, const i64 *tails,
                                                              int32_t *colors, int32_t *cnt,
                                                              const i64 *front, i64 nfront, JpLists L,
                                                              int32_t *err) {
    __shared__ uint32_t s_mark[kMaxColors / 32];
    __shared__ int s_color, s_dep;
    __shared__ Stage<kStageCap> S;
    stage_init(S);
    for (i64 q = blockIdx.x; q < nfront; q += gridDim.x) {
        const i64 v = front[q];
        const i64 b = off[v], e = off[v + 1];
        if (threadIdx.x == 0) {
            s_color = kMaxColors;
            s_dep = 0;
        }
        for (int i = threadIdx.x; i < kMaxColors / 32; i += kBlock) s_mark[i] = 0;
        __syncthreads();
        // one pass: colours and depths of the higher-priority neighbours,
        // release of the lower ones; whole wavefronts iterate together for
        // the appends
        const i64 deg = e - b;
        int dm = 0;
        for (i64 j0 = 0; j0 < deg; j0 += kJpUnroll * kBlock) {
            i64 uu[kJpUnroll];
            uint8_t cd[kJpUnroll];
            bool rd[kJpUnroll];
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) {
                const i64 j = j0 + k * kBlock + threadIdx.x;
                uu[k] = j < deg ? tails[b + j] : -1;
                cd[k] = j < deg ? L.code[b + j] : kJpSelf;
            }
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) {
                rd[k] = false;
                const i64 u = uu[k];
                if (cd[k] & kJpSelf) continue;
                if (cd[k] & kJpRelease) {
                    rd[k] = atomicSub(&cnt[u], 1) == 1;
                } else {
                    const int c = colors[u];
                    const int du = L.dep[u];
                    dm = du > dm ? du : dm;
                    if (c >= 0) atomicOr(&s_mark[c >> 5], 1u << (c & 31));
                }
            }
#pragma unroll
            for (int k = 0; k < kJpUnroll; k++) frontier_push(rd[k], uu[k], cd[k], L, S);
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const int x = __shfl_xor(dm, o);
            dm = x > dm ? x : dm;
        }
        if (threadIdx.x % kWave == 0 && dm > 0) atomicMax(&s_dep, dm);
        __syncthreads();
        for (int i = threadIdx.x; i < kMaxColors / 32; i += kBlock) {
            const uint32_t m = s_mark[i];
            if (m != 0xFFFFFFFFu) atomicMin(&s_color, i * 32 + __ffs((int)~m) - 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int color = s_color;
            if (color >= kMaxColors) {
                atomicOr(err, 1);
                color = 0;
            }
            colors[v] = color;
            L.dep[v] = s_dep + 1;
        }
        __syncthreads();
        if (S.n > kStageCap / 2) stage_flush(S, L.small, &L.counts[0]);  // S.n is settled here
    }
    stage_flush(S, L.small, &L.counts[0]);
}

// Giant frontier vertices, kGiantBlocks workgroups each (blockIdx.y): mark
// the colors of the (coloured, higher) neighbours in the vertex's HBM bitmap
// gmark[y][0 .. kMaxColors/32) and their largest depth in gmark[y][kMaxColors/32]
// (zero on entry) ...
constexpr int kJpGiantMarkWords = kMaxColors / 32 + 1;
__global__ __launch_bounds__(kBlock) void k_jp_giant_mark(const i64 *off, const i64 *tails,
                                                          const int32_t *colors, const int32_t *dep,
                                                          const i64 *front, uint32_t *gmark) {
    const i64 v = front[blockIdx.y];
    uint32_t *mk = gmark + (i64)blockIdx.y * kJpGiantMarkWords;
    int dm = 0;
    for (i64 j = off[v] + (i64)blockIdx.x * kBlock + threadIdx.x; j < off[v + 1]; j += (i64)gridDim.x * kBlock) {
        const i64 u = tails[j];
        if (u == v) continue;
        const int c = colors[u];
        if (c >= 0) {
            atomicOr(&mk[c >> 5], 1u << (c & 31));
            const int du = dep[u];
            dm = du > dm ? du : dm;
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const int x = __shfl_xor(dm, o);
        dm = x > dm ? x : dm;
    }
    if (threadIdx.x % kWave == 0 && dm > 0) atomicMax(&mk[kMaxColors / 32], (uint32_t)dm);
}
// ... then take the first free color and release the lower neighbours.
__global__ __launch_bounds__(kBlock) void k_jp_giant_release(const i64 *off, const i64 *tails,
                                                             int32_t *colors, int32_t *cnt,
                                                             const i64 *front, const uint32_t *gmark,
                                                             JpLists L, int32_t *err) {
    __shared__ int s_color;
    __shared__ Stage<kStageCap> S;
    stage_init(S);
    __syncthreads();
    const i64 v = front[blockIdx.y];
    if (blockIdx.x == 0) {
        const uint32_t *mk = gmark + (i64)blockIdx.y * kJpGiantMarkWords;
        if (threadIdx.x == 0) s_color = kMaxColors;
        __syncthreads();
        for (int i = threadIdx.x; i < kMaxColors / 32; i += kBlock) {
            const uint32_t m = mk[i];
            if (m != 0xFFFFFFFFu) atomicMin(&s_color, i * 32 + __ffs((int)~m) - 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int color = s_color;
            if (color >= kMaxColors) {
                atomicOr(err, 1);
                color = 0;
            }
            colors[v] = color;
            L.dep[v] = (int32_t)mk[kMaxColors / 32] + 1;
        }
    }
    const i64 b = off[v], deg = off[v + 1] - b;
    const i64 step = (i64)gridDim.x * kBlock;
    for (i64 j0 = (i64)blockIdx.x * kBlock; j0 < deg; j0 += step) {
        const i64 j = j0 + threadIdx.x;
        bool ready = false;
        i64 u = 0;
        uint8_t cd = kJpSelf;
        if (j < deg) {
            u = tails[b + j];
            cd = L.code[b + j];
            if (cd & kJpRelease) ready = atomicSub(&cnt[u], 1) == 1;
        }
        frontier_push(ready, u, cd, L, S);
    }
    stage_flush(S, L.small, &L.counts[0]);
}

// Vertices still uncoloured, in increasing id (flags + scan + scatter).
__global__ void k_flag_uncolored(const int32_t *state, i64 nv, i64 *flags) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        flags[v] = state[v] < 0 ? 1 : 0;
}
__global__ void k_scatter_flagged(const i64 *flags, const i64 *pos, i64 nv, i64 *list) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        if (flags[v]) list[pos[v]] = v;
}

constexpr int kTailBlock = 1024;

// The tail in three steps.  tidx[v] = position of v in the tail list.
__global__ void k_tail_index(const i64 *list, i64 n, i64 nv, int32_t *tidx) {
    for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q < n; q += (i64)gridDim.x * blockDim.x)
        if (list[q] >= 0 && list[q] < nv) tidx[list[q]] = (int32_t)q;
}
// One workgroup per tail vertex q: the colours of its coloured neighbours
// into base[q] (fixed from here on), its earlier tail neighbours (position
// below q) into the bit row adj[q].
__global__ __launch_bounds__(kBlock) void k_tail_prep(const i64 *off, const i64 *tails, i64 nv,
                                                      const int32_t *colors, const i64 *list, const int32_t *tidx,
                                                      uint32_t *base, u64 *adj) {
    constexpr int W = (int)(kJpTail / 64);
    const i64 q = blockIdx.x;
    const i64 v = list[q];
    if (v < 0 || v >= nv) return;  // the list holds exactly the uncoloured vertices
    for (i64 j = off[v] + threadIdx.x; j < off[v + 1]; j += kBlock) {
        const i64 u = tails[j];
        if (u == v) continue;
        const int32_t cu = colors[u];
        if (cu >= 0) {
            atomicOr(&base[q * (kMaxColors / 32) + (cu >> 5)], 1u << (cu & 31));
        } else {
            const int32_t t = tidx[u];
            if (t >= 0 && t < q) atomicOr((unsigned long long *)&adj[q * W + (t >> 6)], 1ull << (t & 63));
        }
    }
}
// Sequential first-fit over the n <= kJpTail tail vertices in list order, in
// one workgroup with the bit rows in LDS: thread t holds the colour of tail
// vertex t once it is chosen; vertex i's forbidden colours are base[i] plus
// the colours of the set bits of row i.
constexpr int kTailBatch = 8;  // base rows fetched together
__global__ __launch_bounds__(kTailBlock) void k_tail_seq(const i64 *list, i64 n, i64 nv, const uint32_t *base,
                                                         const u64 *adj, int32_t *colors, int32_t *err) {
    constexpr int W = (int)(kJpTail / 64);
    constexpr int FW = kMaxColors / 32;
    __shared__ u64 sadj[kJpTail * W];
    __shared__ uint32_t forb[kTailBatch][FW];
    __shared__ int s_cw[FW / kWave];
    const int tid = threadIdx.x, lane = tid % kWave, w = tid / kWave;
    for (i64 x = tid; x < n * W; x += kTailBlock) sadj[x] = adj[x];
    int mine = 0;  // colour of tail vertex tid
    for (int i0 = 0; i0 < n; i0 += kTailBatch) {
        __syncthreads();  // the previous batch is done with forb
        for (int x = tid; x < kTailBatch * FW; x += kTailBlock) {
            const int r = i0 + x / FW;
            forb[x / FW][x % FW] = r < n ? base[(i64)r * FW + x % FW] : 0u;
        }
        __syncthreads();
        const int iend = (int)(i0 + kTailBatch < n ? i0 + kTailBatch : n);
        for (int i = i0; i < iend; i++) {
            const int r = i - i0;
            if (tid < i && ((sadj[i * W + (tid >> 6)] >> (tid & 63)) & 1ull))
                atomicOr(&forb[r][mine >> 5], 1u << (mine & 31));
            __syncthreads();
            int cw = kMaxColors;
            if (tid < FW) {
                const uint32_t m = forb[r][tid];
                if (m != 0xFFFFFFFFu) cw = tid * 32 + __ffs((int)~m) - 1;
            }
#pragma unroll
            for (int o = kWave / 2; o > 0; o >>= 1) {
                const int x = __shfl_xor(cw, o);
                cw = x < cw ? x : cw;
            }
            if (lane == 0 && w < FW / kWave) s_cw[w] = cw;
            __syncthreads();
            int c = kMaxColors;
#pragma unroll
            for (int q2 = 0; q2 < FW / kWave; q2++) c = s_cw[q2] < c ? s_cw[q2] : c;
            if (c >= kMaxColors) {
                if (tid == 0) atomicOr(err, 1);
                c = 0;
            }
            if (tid == i) mine = c;
        }
    }
    if (tid < n && list[tid] >= 0 && list[tid] < nv) colors[list[tid]] = mine;
}

// Jones-Plassmann rounds in one persistent launch: the same topological
// sweep as the host-driven rounds (count of higher-priority uncoloured
// neighbours, frontier = count zero, first-fit against the coloured higher
// neighbours), with a grid barrier between rounds instead of a launch and a
// host round trip.  Frontier vertices of degree <= 256 take a wavefront each,
// heavier ones a workgroup each.  Continues from the frontier in sA/lA
// (counts ring[0], ring[1]) after colored0 vertices were coloured, until
// every vertex is.  A frontier of more than max_front vertices ends the
// launch before its round (the host's full-GPU rounds take it; after r
// rounds the frontier is in sA/lA for even r, sB/lB for odd r).  ring: 10
// words, [2q], [2q+1] the small and large frontier counts of rounds r = q
// (mod 3), [6] receives the rounds run here, [7] receives 1 when the launch
// ended on a large frontier, [8] the coloured vertices at the end.
constexpr int kJpPBlock = 1024;
constexpr int kJpLargeGroups = 4;  // large frontier vertices in flight per workgroup when they are many
__global__ __launch_bounds__(kJpPBlock) void k_jp_persistent(const i64 *off, const i64 *tails, i64 nv,
                                                             int32_t *colors, int32_t *cnt, i64 *sA, i64 *sB,
                                                             i64 *lA, i64 *lB, i64 *ring, unsigned *bar,
                                                             int32_t *err, i64 colored0, uint8_t *code,
                                                             int32_t *dep, i64 max_front) {
    constexpr int WPB = kJpPBlock / kWave;
    __shared__ uint32_t mark[WPB][kMaxColors / 32];
    __shared__ int s_color[kJpLargeGroups], s_dep[kJpLargeGroups];
    __shared__ Stage<kStageCap> S;  // appends to the next small frontier
    stage_init(S);
    __syncthreads();
    const unsigned nb = gridDim.x;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    uint32_t *mk = mark[w];
    unsigned seq = 0;
    bool alive = true;
    i64 colored = colored0;
    int r = 0;
    bool exited = false;
    i64 *fs = sA, *fl = lA, *ns_ = sB, *nl_ = lB;
    while (alive) {
        const int q = r % 3, nq = (r + 1) % 3, zq = (r + 2) % 3;
        const i64 nsmall = __hip_atomic_load(&ring[2 * q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const i64 nlarge = __hip_atomic_load(&ring[2 * q + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (nsmall + nlarge == 0) break;
        if (nsmall + nlarge > max_front) {  // a frontier that needs the whole GPU: back to the host
            exited = true;
            break;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            __atomic_store_n(&ring[2 * zq], (i64)0, __ATOMIC_RELAXED);
            __atomic_store_n(&ring[2 * zq + 1], (i64)0, __ATOMIC_RELAXED);
        }
        i64 *cs = &ring[2 * nq], *cl = &ring[2 * nq + 1];
        // small frontier vertices: eight per wavefront
        {
            constexpr i64 E = kWave / kJpGroup;
            const i64 nw = (i64)nb * WPB, wid = (i64)blockIdx.x * WPB + w;
            auto push = [&](bool ready, i64 u, uint8_t cd) {
                const bool big = ready && (cd & kJpLarge);
                stage_append(ready && !big, u, S, ns_, cs);
                wave_append(big, u, nl_, cl);
            };
            for (i64 lo = wid * E; lo < nsmall; lo += nw * E) {
                const i64 n = nsmall - lo < E ? nsmall - lo : E;
                jp_wave8(off, tails, colors, dep, cnt, code, fs, lo, n, mk, err, push);
            }
        }
        __syncthreads();  // the marks are the large loop's from here
        // large frontier vertices: a workgroup each, or a quarter of one each
        // when there are more of them than workgroups (4 in flight per CU)
        const int NG = nlarge > (i64)nb ? kJpLargeGroups : 1;
        const int GT = kJpPBlock / NG, grp = threadIdx.x / GT, gtid = threadIdx.x % GT;
        uint32_t *gm = mark[grp * (WPB / NG)];
        for (i64 base = (i64)blockIdx.x * NG; base < nlarge; base += (i64)nb * NG) {
            const i64 qq = base + grp;
            const bool act = qq < nlarge;
            const i64 v = act ? fl[qq] : 0;
            const i64 b = act ? off[v] : 0, deg = act ? off[v + 1] - b : 0;
            for (int i = gtid; i < kMaxColors / 32; i += GT) gm[i] = 0;
            if (gtid == 0) {
                s_color[grp] = kMaxColors;
                s_dep[grp] = 0;
            }
            __syncthreads();
            // one pass: colours and depths of the higher-priority neighbours
            // (all coloured), release of the lower-priority ones (next round's
            // frontier)
            int dm = 0;
            for (i64 j0 = 0; j0 < deg; j0 += kJpUnroll * GT) {
                i64 uu[kJpUnroll];
                uint8_t cd[kJpUnroll];
                bool rd[kJpUnroll];
#pragma unroll
                for (int k = 0; k < kJpUnroll; k++) {
                    const i64 j = j0 + k * GT + gtid;
                    uu[k] = j < deg ? tails[b + j] : -1;
                    cd[k] = j < deg ? code[b + j] : kJpSelf;
                }
#pragma unroll
                for (int k = 0; k < kJpUnroll; k++) {
                    rd[k] = false;
                    const i64 u = uu[k];
                    if (cd[k] & kJpSelf) continue;
                    if (cd[k] & kJpRelease) {
                        rd[k] = atomicSub(&cnt[u], 1) == 1;
                    } else {
                        const int c = colors[u];
                        const int du = dep[u];
                        dm = du > dm ? du : dm;
                        if (c >= 0) atomicOr(&gm[c >> 5], 1u << (c & 31));
                    }
                }
#pragma unroll
                for (int k = 0; k < kJpUnroll; k++) {
                    const bool bg = rd[k] && (cd[k] & kJpLarge);
                    stage_append(rd[k] && !bg, uu[k], S, ns_, cs);
                    wave_append(bg, uu[k], nl_, cl);
                }
            }
            for (int o = kWave / 2; o > 0; o >>= 1) {
                const int x = __shfl_xor(dm, o);
                dm = x > dm ? x : dm;
            }
            if (lane == 0 && dm > 0) atomicMax(&s_dep[grp], dm);
            __syncthreads();
            if (gtid < kMaxColors / 32) {
                const uint32_t m = gm[gtid];
                if (m != 0xFFFFFFFFu) atomicMin(&s_color[grp], gtid * 32 + __ffs((int)~m) - 1);
            }
            __syncthreads();
            if (act && gtid == 0) {
                int color = s_color[grp];
                if (color >= kMaxColors) {
                    atomicOr(err, 1);
                    color = 0;
                }
                colors[v] = color;
                dep[v] = s_dep[grp] + 1;
            }
            __syncthreads();
            if (S.n > kStageCap / 2) stage_flush(S, ns_, cs);  // S.n is settled here
        }
        stage_flush(S, ns_, cs);
        alive = grid_barrier(bar, nb, ++seq);
        colored += nsmall + nlarge;
        r++;
        if (nv - colored <= 0) break;
        i64 *t = fs;
        fs = ns_;
        ns_ = t;
        t = fl;
        fl = nl_;
        nl_ = t;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ring[6] = r;
        ring[7] = exited ? 1 : 0;
        ring[8] = colored;
    }
}

__global__ void k_jp_prio(const i64 *off, i64 nv, u64 *prio) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        const u64 d = (u64)min(off[v + 1] - off[v], (i64(1) << 23) - 1);
        prio[v] = (d << 40) | (jp_priority(v) >> 24);
    }
}

__global__ void k_jp_finish(const int32_t *colors, i64 nv, int32_t *maxc) {
    int m = -1;
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        m = colors[v] > m ? colors[v] : m;
    atomicMax(maxc, m);
}

__global__ void k_color_conflicts(const i64 *off, const i64 *tails, i64 nv, const int32_t *colors,
                                  i64 *count) {
    i64 local = 0;
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        for (i64 j = off[v]; j < off[v + 1]; j++) {
            i64 u = tails[j];
            if (u != v && colors[u] == colors[v]) local++;
        }
    if (local) atomicAdd((unsigned long long *)count, (unsigned long long)local);
}

// ------------------------------------------------------------------- renumber
__global__ void k_first_index(const i64 *C, i64 n, i64 *first, int32_t *err) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        i64 c = C[i];
        if (c >= n) {
            atomicOr(err, 1);
            continue;
        }
        if (c >= 0) atomicMin((unsigned long long *)&first[c], (unsigned long long)i);
    }
}
__global__ void k_head_flags(const i64 *C, i64 n, const i64 *first, i64 *flags) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        i64 c = C[i];
        flags[i] = (c >= 0 && c < n && first[c] == i) ? 1 : 0;
    }
}
__global__ void k_relabel(i64 *C, i64 n, const i64 *first, const i64 *pos) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        i64 c = C[i];
        if (c >= 0 && c < n) C[i] = pos[first[c]];
    }
}
// C_orig update, runMultiPhaseColoring.cpp:132-144.
__global__ void k_compose(i64 *C_orig, i64 n0, const i64 *C, bool first_phase) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n0; i += (i64)gridDim.x * blockDim.x) {
        if (first_phase)
            C_orig[i] = C[i];
        else if (C_orig[i] >= 0)
            C_orig[i] = C[C_orig[i]];
    }
}

// -------------------------------------------------------------------- coarsen
// One wavefront per row: count / emit the entries i->j with C(i) >= C(j)
// (buildNextPhase.cpp:157), keeping CSR order.
// The contraction's entries: entry j of row v (a label cv >= 0) whose
// neighbour's label ct is <= cv gives the pair (cv, ct) with weight w[j].
// Entries are processed in chunks of kCoarsenChunk consecutive entries per
// workgroup (so rows of any degree cost the same per entry): the rows a
// chunk spans are found by binary search, their offsets and labels staged in
// LDS (up to kCoarsenRows of them; more rows -- long runs of rows without
// entries -- are searched in HBM), each entry finds its row in LDS.  One
// pass counts each chunk's pairs; a second writes them, in entry order, at
// the chunk's exclusive prefix.
constexpr i64 kCoarsenChunk = 8192;
constexpr int kCoarsenRows = 1024;
struct ChunkRows {
    i64 r0, nr;    // first row, rows spanned
    bool staged;   // offsets and labels in LDS
};
// rows of entries [e0, e1) (nonempty): r0 = the row of e0, r1 = that of e1 - 1
__device__ inline i64 row_of(const i64 *off, i64 lo, i64 hi, i64 j) {
    // the largest r in [lo, hi] with off[r] <= j
    while (lo < hi) {
        const i64 mid = (lo + hi + 1) >> 1;
        if (off[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__device__ inline ChunkRows stage_rows(const i64 *off, const i64 *C, i64 nv, i64 e0, i64 e1, i64 *s_off,
                                       i64 *s_c, i64 *s_r) {
    if (threadIdx.x == 0) {
        const i64 r0 = row_of(off, 0, nv - 1, e0);
        s_r[0] = r0;
        s_r[1] = row_of(off, r0, nv - 1, e1 - 1) - r0 + 1;
    }
    __syncthreads();
    ChunkRows R{s_r[0], s_r[1], s_r[1] <= kCoarsenRows};
    if (R.staged) {
        for (i64 i = threadIdx.x; i <= R.nr; i += blockDim.x) s_off[i] = off[R.r0 + i];
        for (i64 i = threadIdx.x; i < R.nr; i += blockDim.x) s_c[i] = C[R.r0 + i];
    }
    __syncthreads();
    return R;
}
// the label of entry j's row
__device__ inline i64 row_label(const ChunkRows &R, const i64 *s_off, const i64 *s_c, const i64 *off,
                                const i64 *C, i64 j) {
    if (R.staged) {
        int lo = 0, hi = (int)R.nr - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= j) lo = mid;
            else hi = mid - 1;
        }
        return s_c[lo];
    }
    return C[row_of(off, R.r0, R.r0 + R.nr - 1, j)];
}
__global__ __launch_bounds__(kBlock) void k_coarsen_count(const i64 *off, const i64 *tails, const i64 *C, i64 nv,
                                                       i64 nnz, i64 *ccount) {
    __shared__ i64 s_off[kCoarsenRows + 1], s_c[kCoarsenRows], s_r[2];
    __shared__ i64 s_red[kBlock / kWave];
    const i64 nchunks = (nnz + kCoarsenChunk - 1) / kCoarsenChunk;
    for (i64 ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const i64 e0 = ch * kCoarsenChunk, e1 = e0 + kCoarsenChunk < nnz ? e0 + kCoarsenChunk : nnz;
        const ChunkRows R = stage_rows(off, C, nv, e0, e1, s_off, s_c, s_r);
        i64 c = 0;
        for (i64 j = e0 + threadIdx.x; j < e1; j += kBlock) {
            const i64 cv = row_label(R, s_off, s_c, off, C, j);
            if (cv >= 0) c += C[tails[j]] <= cv;
        }
        for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_xor((long long)c, o);
        if (threadIdx.x % kWave == 0) s_red[threadIdx.x / kWave] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            i64 t = 0;
            for (int q = 0; q < kBlock / kWave; q++) t += s_red[q];
            ccount[ch] = t;
        }
        __syncthreads();
    }
}
// nseeds: k (buildNextLevelGraphOpt: one zero self-loop per cluster) or 0
// (buildNewGraphVF, vertexFollowing.cpp:127-131, which also sums weights
// truncated to integers, :147-151: trunc).  cbase: the chunks' exclusive
// prefix of pair counts.
__global__ __launch_bounds__(kBlock) void k_coarsen_emit(const i64 *off, const i64 *tails, const double *w,
                                                      const i64 *C, i64 nv, i64 nnz, const i64 *cbase, int kb,
                                                      u64 *keys, double *vals, i64 nseeds, bool trunc) {
    __shared__ i64 s_off[kCoarsenRows + 1], s_c[kCoarsenRows], s_r[2];
    __shared__ int s_wc[kBlock / kWave];
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
    const u64 lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    // seeds: one zero self-loop per cluster (buildNextPhase.cpp:122)
    for (i64 c = blockIdx.x * (i64)kBlock + threadIdx.x; c < nseeds; c += (i64)gridDim.x * kBlock) {
        keys[c] = ((u64)c << kb) | (u64)c;
        vals[c] = 0.0;
    }
    const i64 nchunks = (nnz + kCoarsenChunk - 1) / kCoarsenChunk;
    for (i64 ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const i64 e0 = ch * kCoarsenChunk, e1 = e0 + kCoarsenChunk < nnz ? e0 + kCoarsenChunk : nnz;
        const ChunkRows R = stage_rows(off, C, nv, e0, e1, s_off, s_c, s_r);
        i64 base = nseeds + cbase[ch];
        for (i64 j0 = e0; j0 < e1; j0 += kBlock) {
            const i64 j = j0 + threadIdx.x;
            i64 cv = -1, ct = 0;
            bool p = false;
            if (j < e1) {
                cv = row_label(R, s_off, s_c, off, C, j);
                if (cv >= 0) {
                    ct = C[tails[j]];
                    p = ct <= cv;
                }
            }
            const u64 m = __ballot(p);
            if (lane == 0) s_wc[wv] = __popcll(m);
            __syncthreads();
            i64 before = 0, all = 0;
            for (int q = 0; q < kBlock / kWave; q++) {
                before += q < wv ? s_wc[q] : 0;
                all += s_wc[q];
            }
            if (p) {
                const i64 at = base + before + __popcll(m & lt);
                keys[at] = ((u64)cv << kb) | (u64)ct;
                vals[at] = trunc ? (double)(long)w[j] : w[j];
            }
            base += all;
            __syncthreads();  // s_wc is rewritten by the next pass
        }
    }
}
// Each unique (a,b) becomes the entry of row a, and of row b when a != b;
// the second slot of a self-loop gets an all-ones key that sorts last.
__global__ void k_coarsen_mirror(const u64 *ukeys, const double *uvals, i64 nu, int kb, u64 *keys,
                                 double *vals, unsigned long long *nself) {
    const u64 mask = ((u64)1 << kb) - 1;
    for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q < nu; q += (i64)gridDim.x * blockDim.x) {
        const u64 key = ukeys[q];
        const u64 a = key >> kb, b = key & mask;
        keys[2 * q] = key;
        vals[2 * q] = uvals[q];
        if (a != b) {
            keys[2 * q + 1] = (b << kb) | a;
            vals[2 * q + 1] = uvals[q];
        } else {
            keys[2 * q + 1] = ~0ULL;
            vals[2 * q + 1] = 0.0;
            if (nself) atomicAdd(nself, 1ULL);
        }
    }
}
__global__ void k_coarsen_rows(const u64 *keys, const double *vals, i64 total, i64 k, int kb, i64 *off,
                               i64 *tails, double *w) {
    const u64 mask = ((u64)1 << kb) - 1;
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < total; i += (i64)gridDim.x * blockDim.x) {
        const u64 key = keys[i];
        const i64 r = (i64)(key >> kb);
        tails[i] = (i64)(key & mask);
        w[i] = vals[i];
        // row starts, empty rows included (the contraction without seeds can
        // leave a cluster without entries): every off[] slot has one writer
        const i64 rp = i == 0 ? -1 : (i64)(keys[i - 1] >> kb);
        for (i64 rr = rp + 1; rr <= r; ++rr) off[rr] = i;
        if (i == total - 1)
            for (i64 rr = r + 1; rr <= k; ++rr) off[rr] = total;
    }
}

// vertexFollowing (Utility/vertexFollowing.cpp:46-89): -1 for vertices
// without entries; a vertex with one entry follows its neighbour when that
// neighbour has more than one entry or a smaller id; *count += both kinds.
__global__ void k_vertex_following(const i64 *off, const i64 *tails, i64 nv, i64 *C,
                                   unsigned long long *count) {
    unsigned long long local = 0;
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nv; i += (i64)gridDim.x * blockDim.x) {
        const i64 a = off[i], b = off[i + 1];
        i64 c = i;
        if (a == b) {
            c = -1;
            local++;
        } else if (b - a == 1) {
            const i64 t = tails[a];
            if (off[t + 1] - off[t] > 1 || i > t) {
                c = t;
                local++;
            }
        }
        C[i] = c;
    }
    if (local) atomicAdd(count, local);
}

// ----------------------------------------------------------------- AoS/checks
// Labels supplied by a caller: bit 1 = label outside [-1, limit), bit 2 = a
// vertex with entries labelled -1 (only vertices without entries may be).
__global__ void k_check_labels(const i64 *off, const i64 *C, i64 nv, i64 limit, int32_t *err) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nv; i += (i64)gridDim.x * blockDim.x) {
        const i64 c = C[i];
        if (c < -1 || c >= limit) atomicOr(err, 1);
        else if (c == -1 && off[i + 1] > off[i]) atomicOr(err, 2);
    }
}
__global__ void k_split_edges(const i64 *aos, i64 nnz, i64 *tails, double *w) {
    for (i64 j = blockIdx.x * (i64)blockDim.x + threadIdx.x; j < nnz; j += (i64)gridDim.x * blockDim.x) {
        tails[j] = aos[3 * j + 1];
        w[j] = __longlong_as_double(aos[3 * j + 2]);
    }
}
__global__ void k_check_graph(const i64 *off, const i64 *tails, i64 nv, i64 nnz, int32_t *err) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < nv; i += (i64)gridDim.x * blockDim.x)
        if (off[i] > off[i + 1] || off[i] < 0 || off[i + 1] > nnz) atomicOr(err, 1);
    for (i64 j = blockIdx.x * (i64)blockDim.x + threadIdx.x; j < nnz; j += (i64)gridDim.x * blockDim.x)
        if (tails[j] < 0 || tails[j] >= nv) atomicOr(err, 2);
}

}  // namespace

// =================================================================== launchers
#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

void launch_fill_i64(i64 *p, i64 n, i64 v, hipStream_t s) {
    if (n <= 0) return;
    k_fill_i64<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(p, n, v);
    GLV_LAUNCH_CHECK();
}
void launch_fill_f64(double *p, i64 n, double v, hipStream_t s) {
    if (n <= 0) return;
    k_fill_f64<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(p, n, v);
    GLV_LAUNCH_CHECK();
}
void launch_fill_i32(int32_t *p, i64 n, int32_t v, hipStream_t s) {
    if (n <= 0) return;
    k_fill_i32<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(p, n, v);
    GLV_LAUNCH_CHECK();
}
void launch_iota(i64 *p, i64 n, hipStream_t s) {
    if (n <= 0) return;
    k_iota<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(p, n);
    GLV_LAUNCH_CHECK();
}
void launch_init_labels(const i64 *off, i64 nv, i64 *C, hipStream_t s) {
    if (nv <= 0) return;
    k_init_labels<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(off, nv, C);
    GLV_LAUNCH_CHECK();
}
void launch_bin_keys(const i64 *off, i64 nv, const int32_t *colors, int ncol, uint32_t *keys, i64 *vid,
                     int32_t *err, hipStream_t s, const BinConfig &bins) {
    if (nv <= 0) return;
    k_bin_keys<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(off, nv, colors, ncol, keys, vid, err, bins);
    GLV_LAUNCH_CHECK();
}
__global__ void k_color_keys(const i64 *perm, const int32_t *colors, i64 n, uint32_t *keys, i64 *vals) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const i64 v = perm[i];
        keys[i] = (uint32_t)colors[v];
        vals[i] = v;
    }
}
void launch_color_major(const i64 *perm, const int32_t *colors, i64 n, int ncol, i64 *cperm,
                        i64 *cstart, hipStream_t s) {
    if (n <= 0) return;
    DBuf<uint32_t> keys(n, s);
    k_color_keys<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(perm, colors, n, keys.p, cperm);
    GLV_LAUNCH_CHECK();
    int bits = 1;
    while ((1 << bits) <= ncol) bits++;
    sort_pairs_u32_i64(keys.p, cperm, n, bits, s);
    k_slice_starts<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(keys.p, n, ncol, cstart);
    GLV_LAUNCH_CHECK();
}
void launch_slices(const uint32_t *sorted_keys, const i64 *perm, const i64 *off, i64 n, i64 nkeys,
                   i64 *start, i64 *pnnz, hipStream_t s) {
    if (n <= 0) return;
    DBuf<i64> degp(n, s), dscan(n, s);
    k_slice_starts<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(sorted_keys, n, nkeys, start);
    GLV_LAUNCH_CHECK();
    k_perm_degree<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(perm, off, n, degp.p);
    GLV_LAUNCH_CHECK();
    scan_exclusive_i64(degp.p, dscan.p, n, s);
    k_slice_nnz<<<grid_for(nkeys, kBlock, 8192), kBlock, 0, s>>>(start, dscan.p, degp.p, n, nkeys, pnnz);
    GLV_LAUNCH_CHECK();
}
void launch_degree(const DevGraph &g, const i64 *big, i64 nbig, double *k, hipStream_t s) {
    if (g.nv <= 0) return;
    if (g.nnz <= kDegTiledMean * g.nv)
        k_degree_rows<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(g.off, g.w, g.nv, k);
    else
        k_degree_rows_seq<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(g.off, g.w, g.nv, k);
    GLV_LAUNCH_CHECK();
    if (nbig > 0) {
        k_degree_big<<<(unsigned)std::min<i64>(nbig, 4096), kBlock, 0, s>>>(g.off, g.w, big, nbig, k);
        GLV_LAUNCH_CHECK();
    }
}
void launch_init_comm(const double *k, i64 nv, double *a, i64 *n, hipStream_t s) {
    if (nv <= 0) return;
    k_init_comm<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(k, nv, a, n);
    GLV_LAUNCH_CHECK();
}
void launch_comm_from_labels(const i64 *C, const double *k, i64 nv, double *a, i64 *n, hipStream_t s) {
    launch_fill_f64(a, nv, 0.0, s);
    launch_fill_i64(n, nv, 0, s);
    if (nv <= 0) return;
    k_comm_from_labels<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(C, k, nv, a, n);
    GLV_LAUNCH_CHECK();
}

void launch_sum(const double *x, i64 n, double *part, double *out, hipStream_t s) {
    k_red<false><<<kRedBlocks, kBlock, 0, s>>>(x, n, part);
    GLV_LAUNCH_CHECK();
    launch_sum_partials(part, kRedBlocks, out, s);
}
void launch_sumsq(const double *x, i64 n, double *part, double *out, hipStream_t s) {
    k_red<true><<<kRedBlocks, kBlock, 0, s>>>(x, n, part);
    GLV_LAUNCH_CHECK();
    launch_sum_partials(part, kRedBlocks, out, s);
}
void launch_sum_partials(const double *part, i64 np, double *out, hipStream_t s) {
    k_sum_partials<<<1, kBlock, 0, s>>>(part, np, out);
    GLV_LAUNCH_CHECK();
}

static bool use_dense(int bin, i64 nv) { return bin >= 7 && nv <= kDenseMaxNV; }

i64 sweep_blocks(int bin, i64 count, i64 nv) {
    if (count <= 0) return 0;
    if (use_dense(bin, nv)) return std::min<i64>(count, kDenseBlocks);
    switch (bin) {
        case 1: return (count + kBlock / 4 - 1) / (kBlock / 4);
        case 2: return (count + kBlock / 8 - 1) / (kBlock / 8);
        case 3: return (count + kBlock / 16 - 1) / (kBlock / 16);
        case 4: return (count + kBlock / 32 - 1) / (kBlock / 32);
        case 5: return (count + kBlock / 64 - 1) / (kBlock / 64);
        case 6: return (count + kBlock / kWave - 1) / (kBlock / kWave);
        case 7: return count;
        case 8: return std::min<i64>(count, kGlobalHashBlocks);
        case 9: return count;  // one partial per giant vertex (written by its decision)
        default: return 0;
    }
}

template <bool PLAIN>
static void sweep_dispatch(int bin, const SweepArgs &A, hipStream_t s) {
    const unsigned g = (unsigned)sweep_blocks(bin, A.count, A.nv);
    if (g == 0) return;
    if (use_dense(bin, A.nv)) {
        k_sweep_dense<PLAIN><<<g, kBlock, 0, s>>>(A);
        GLV_LAUNCH_CHECK();
        return;
    }
    switch (bin) {
        case 1: k_sweep_group<4, PLAIN><<<g, kBlock, 0, s>>>(A); break;
        case 2: k_sweep_group<8, PLAIN><<<g, kBlock, 0, s>>>(A); break;
        case 3: k_sweep_group<16, PLAIN><<<g, kBlock, 0, s>>>(A); break;
        case 4: k_sweep_group<32, PLAIN><<<g, kBlock, 0, s>>>(A); break;
        case 5: k_sweep_group<64, PLAIN><<<g, kBlock, 0, s>>>(A); break;
        case 6: k_sweep_wavehash<PLAIN><<<g, kBlock, 0, s>>>(A); break;
        case 7: k_sweep_blockhash<PLAIN, false><<<g, kBlock, 0, s>>>(A); break;
        case 8: k_sweep_blockhash<PLAIN, true><<<g, kBlock, 0, s>>>(A); break;
        case 9: {
            const dim3 grid(kGiantBlocks, (unsigned)A.count);
            k_giant_accum<<<grid, kBlock, 0, s>>>(A);
            GLV_LAUNCH_CHECK();
            k_giant_score<<<grid, kBlock, 0, s>>>(A);
            GLV_LAUNCH_CHECK();
            k_giant_decide<PLAIN><<<(unsigned)A.count, kWave, 0, s>>>(A);
            break;
        }
        default: return;
    }
    GLV_LAUNCH_CHECK();
}
void launch_sweep(int bin, bool plain, const SweepArgs &A, hipStream_t s) {
    if (plain)
        sweep_dispatch<true>(bin, A, s);
    else
        sweep_dispatch<false>(bin, A, s);
}
void launch_giant_table_sizes(const i64 *off, const i64 *verts, i64 count, i64 *tab, hipStream_t s) {
    if (count <= 0) return;
    k_giant_table_sizes<<<grid_for(count, kBlock, 1024), kBlock, 0, s>>>(off, verts, count, tab);
    GLV_LAUNCH_CHECK();
}
void launch_max_degree(const i64 *off, const i64 *verts, i64 count, i64 *out, hipStream_t s) {
    if (count <= 0) return;
    k_max_degree<<<grid_for(count, kBlock, 4096), kBlock, 0, s>>>(off, verts, count, out);
    GLV_LAUNCH_CHECK();
}
// Launch of a kernel that synchronises its workgroups with its own flag
// barriers (grid_barrier): an ordinary launch in stream order whose grid never
// exceeds what the device holds resident at once (`resident`, from the
// kernel's occupancy), so every workgroup runs concurrently and each barrier
// completes; the barriers' bounded waits turn anything else into an error flag
// instead of a hang.
static bool launch_resident(const void *fn, int nblocks, int block, void **args, int resident, hipStream_t s,
                            const char *what) {
    if (nblocks <= 0 || nblocks > resident) {
        trace("%s: %d workgroups exceed the %d resident ones", what, nblocks, resident);
        return false;
    }
    hipError_t e = hipLaunchKernel(fn, dim3(nblocks), dim3(block), args, 0, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trace("%s (%d workgroups) refused: %s", what, nblocks, hipGetErrorString(e));
        return false;
    }
    return true;
}

int persistent_blocks() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0, cus = 0, per = 0, per_h = 0;
        GLV_HIP(hipGetDevice(&dev));
        GLV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        GLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_colored_iteration<false>, kPBlock, 0));
        GLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_h, k_colored_iteration<true>, kPBlock, 0));
        per = std::min(per, per_h);
        GLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_h, k_colored_iteration_spec, kPBlock, 0));
        per = std::min(per, per_h);
        cached = per > 0 ? cus * per : 0;
    }
    return cached;
}
bool launch_colored_iteration(const SweepArgs &A, const i64 *perm, const i64 *start, int ncol,
                              unsigned *bar, double *part, int nblocks, hipStream_t s, uint8_t *stale,
                              i64 *C1, double *a1, i64 *n1, uint16_t *L16) {
    const bool hash = A.nv > kDenseMaxNV;
    if (nblocks <= 0 || nblocks > (hash ? kPersistMaxBlocksHash : kPersistMaxBlocks) || A.nv > kPersistMaxNV)
        return false;
    // degrees above the LDS table's reach need the workgroup's dense HBM row
    if (hash && A.dacc == nullptr) return false;
    SweepArgs aa = A;
    i64 *C = const_cast<i64 *>(A.C);
    double *ca = const_cast<double *>(A.a);
    i64 *cn = const_cast<i64 *>(A.n);
    const i64 *p = perm;
    const i64 *st = start;
    int nc = ncol;
    unsigned *b = bar;
    double *pt = part;
    uint8_t *sf = stale;
    i64 *c1 = C1;
    double *ca1 = a1;
    i64 *cn1 = n1;
    uint16_t *l16 = L16;
    void *args[] = {&aa, &C, &ca, &cn, &p, &st, &nc, &b, &pt, &sf, &c1, &ca1, &cn1, &l16};
    const bool spec = !hash && L16 && nblocks >= 2 && nblocks % 2 == 0;
    const void *fn = spec ? (const void *)k_colored_iteration_spec
                          : hash ? (const void *)k_colored_iteration<true> : (const void *)k_colored_iteration<false>;
    return launch_resident(fn, nblocks, kPBlock, args, persistent_blocks(), s, "persistent colored iteration");
}
void launch_isolated_targets(const i64 *verts, i64 count, i64 *tgt, hipStream_t s) {
    if (count <= 0) return;
    k_isolated_targets<<<grid_for(count, kBlock, 8192), kBlock, 0, s>>>(verts, count, tgt);
    GLV_LAUNCH_CHECK();
}
void launch_apply_colored(const i64 *perm, const Slices &S, const i64 *tgt, i64 *C, const double *k,
                          double *a, i64 *n, hipStream_t s, uint8_t *stale, const double *dein,
                          double *dsum, double *dpart) {
    i64 total = 0;
    for (int b = 0; b < kNumBins; b++) total += S.hi[b] - S.lo[b];
    if (total <= 0) return;
    const unsigned blocks = grid_for(total, kBlock, kApplyMaxBlocks);
    k_apply_colored<<<blocks, kBlock, 0, s>>>(perm, S, tgt, C, k, a, n, stale, dein, dsum ? dpart : nullptr);
    GLV_LAUNCH_CHECK();
    if (dsum) {
        k_add_partials<<<1, kBlock, 0, s>>>(dpart, blocks, dsum);
        GLV_LAUNCH_CHECK();
    }
}
void launch_apply_plain(i64 nv, const i64 *tgt, const i64 *curr, const double *k, double *a, i64 *n,
                        hipStream_t s) {
    if (nv <= 0) return;
    k_apply_plain<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(nv, tgt, curr, k, a, n);
    GLV_LAUNCH_CHECK();
}
void launch_init_comm_ass_opt(const DevGraph &g, const double *k, const double *a, const i64 *n,
                              double constant, i64 *curr, hipStream_t s, const int32_t *gid,
                              const double *gconst) {
    if (g.nv <= 0) return;
    k_init_comm_ass_opt<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(g.off, g.tails, g.w, g.nv, k, a,
                                                                        n, constant, gid, gconst, curr);
    GLV_LAUNCH_CHECK();
}
void launch_update_ax_opt(i64 nv, const i64 *curr, const double *k, double *a, i64 *n, hipStream_t s) {
    if (nv <= 0) return;
    k_update_ax_opt<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(nv, curr, k, a, n);
    GLV_LAUNCH_CHECK();
}

i64 mod_blocks(int bin, i64 count) {
    if (count <= 0 || bin == 0) return 0;
    if (bin == 9) return count * kGiantBlocks;
    const int G = bin <= 5 ? kBinGroup[bin] : (bin == 6 ? kWave : kBlock);
    const i64 vpb = kBlock / G;
    return std::min<i64>((count + vpb - 1) / vpb, 8192);
}
void launch_mod_pass(int bin, const DevGraph &g, const i64 *C, const i64 *verts, i64 count, double *part,
                     hipStream_t s) {
    const unsigned gr = (unsigned)mod_blocks(bin, count);
    if (gr == 0) return;
    if (bin == 9) {
        k_mod_giant<<<dim3(kGiantBlocks, (unsigned)count), kBlock, 0, s>>>(g.off, g.tails, g.w, C, verts, part);
        GLV_LAUNCH_CHECK();
        return;
    }
    switch (bin) {
        case 1: k_mod_pass<4><<<gr, kBlock, 0, s>>>(g.off, g.tails, g.w, C, verts, count, part); break;
        case 2: k_mod_pass<8><<<gr, kBlock, 0, s>>>(g.off, g.tails, g.w, C, verts, count, part); break;
        case 3: k_mod_pass<16><<<gr, kBlock, 0, s>>>(g.off, g.tails, g.w, C, verts, count, part); break;
        case 4: k_mod_pass<32><<<gr, kBlock, 0, s>>>(g.off, g.tails, g.w, C, verts, count, part); break;
        case 5:
        case 6: k_mod_pass<64><<<gr, kBlock, 0, s>>>(g.off, g.tails, g.w, C, verts, count, part); break;
        default: k_mod_pass<256><<<gr, kBlock, 0, s>>>(g.off, g.tails, g.w, C, verts, count, part); break;
    }
    GLV_LAUNCH_CHECK();
}

void launch_jp_count(const DevGraph &g, int32_t *cnt, const JpLists &L, i64 *heavy, i64 *heavy_count,
                     i64 *giant, i64 *giant_count, hipStream_t s) {
    if (g.nv <= 0) return;
    k_jp_count<<<grid_for(g.nv, kBlock, 2048), kBlock, 0, s>>>(g.off, g.tails, g.nv, cnt, L, heavy,
                                                               heavy_count);
    GLV_LAUNCH_CHECK();
    k_jp_count_heavy<<<4096, kBlock, 0, s>>>(g.off, g.tails, cnt, heavy, heavy_count, L, giant, giant_count);
    GLV_LAUNCH_CHECK();
    k_jp_count_giant<<<1024, kBlock, 0, s>>>(g.off, g.tails, cnt, giant, giant_count, L);
    GLV_LAUNCH_CHECK();
}
void launch_jp_frontier(const DevGraph &g, int32_t *colors, int32_t *cnt, const JpLists &cur,
                        const i64 n[3], const JpLists &next, uint32_t *gmark, int32_t *err, hipStream_t s) {
    if (n[0] > 0) {
        // entries per workgroup: four passes of its wavefronts, so that the
        // staged appends cost one device atomic per 128 frontier entries
        const i64 per = 4 * (kBlock / kWave) * (kWave / kJpGroup);
        const unsigned blocks = (unsigned)std::max<i64>(1, std::min<i64>((n[0] + per - 1) / per, 16384));
        k_jp_frontier_wave<<<blocks, kBlock, 0, s>>>(g.off, g.tails, colors, cnt, cur.small, n[0], next, err);
        GLV_LAUNCH_CHECK();
    }
    if (n[1] > 0) {
        k_jp_frontier_block<<<(unsigned)std::min<i64>(n[1], 8192), kBlock, 0, s>>>(
            g.off, g.tails, colors, cnt, cur.large, n[1], next, err);
        GLV_LAUNCH_CHECK();
    }
    if (n[2] > 0) {
        launch_fill_i32((int32_t *)gmark, n[2] * kJpGiantMarkWords, 0, s);
        const dim3 grid(kGiantBlocks, (unsigned)n[2]);
        k_jp_giant_mark<<<grid, kBlock, 0, s>>>(g.off, g.tails, colors, cur.dep, cur.giant, gmark);
        GLV_LAUNCH_CHECK();
        k_jp_giant_release<<<grid, kBlock, 0, s>>>(g.off, g.tails, colors, cnt, cur.giant, gmark, next, err);
        GLV_LAUNCH_CHECK();
    }
}
void launch_color_tail(const DevGraph &g, int32_t *state, i64 *flags, i64 *pos, i64 *list, i64 n,
                       int32_t *tidx, uint32_t *base, u64 *adj, int32_t *err, hipStream_t s) {
    if (n <= 0) return;
    if (n > kJpTail) {  // callers pass at most kJpTail vertices
        launch_fill_i32(err, 1, 2, s);
        return;
    }
    k_flag_uncolored<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(state, g.nv, flags);
    GLV_LAUNCH_CHECK();
    scan_exclusive_i64(flags, pos, g.nv, s);
    k_scatter_flagged<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(flags, pos, g.nv, list);
    GLV_LAUNCH_CHECK();
    launch_fill_i32(tidx, g.nv, -1, s);
    launch_fill_i32((int32_t *)base, n * (kMaxColors / 32), 0, s);
    launch_fill_i64((i64 *)adj, n * (kJpTail / 64), 0, s);
    k_tail_index<<<grid_for(n, kBlock, 64), kBlock, 0, s>>>(list, n, g.nv, tidx);
    GLV_LAUNCH_CHECK();
    k_tail_prep<<<(unsigned)n, kBlock, 0, s>>>(g.off, g.tails, g.nv, state, list, tidx, base, adj);
    GLV_LAUNCH_CHECK();
    k_tail_seq<<<1, kTailBlock, 0, s>>>(list, n, g.nv, base, adj, state, err);
    GLV_LAUNCH_CHECK();
}
int jp_persistent_blocks() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0, cus = 0, per = 0;
        GLV_HIP(hipGetDevice(&dev));
        GLV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        GLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_jp_persistent, kJpPBlock, 0));
        cached = per > 0 ? cus * per : 0;
    }
    return cached;
}
bool launch_jp_persistent(const DevGraph &g, int32_t *colors, int32_t *cnt, i64 *sA, i64 *sB, i64 *lA,
                          i64 *lB, i64 *ring, unsigned *bar, int32_t *err, i64 colored0, const u64 *prio,
                          uint8_t *code, int32_t *dep, i64 max_front, int nblocks, hipStream_t s) {
    (void)prio;
    if (nblocks <= 0 || nblocks > kJpPersistMaxBlocks || g.nv <= 0) return false;
    const i64 *off = g.off, *tails = g.tails;
    i64 nv = g.nv, c0 = colored0, mf = max_front;
    uint8_t *cd = code;
    int32_t *dp = dep;
    void *args[] = {&off, &tails, &nv, &colors, &cnt, &sA, &sB, &lA, &lB, &ring, &bar, &err, &c0, &cd, &dp, &mf};
    return launch_resident((const void *)k_jp_persistent, nblocks, kJpPBlock, args, jp_persistent_blocks(), s,
                           "persistent JP rounds");
}
void launch_jp_prio(const DevGraph &g, u64 *prio, hipStream_t s) {
    if (g.nv <= 0) return;
    k_jp_prio<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(g.off, g.nv, prio);
    GLV_LAUNCH_CHECK();
}
void launch_jp_finish(int32_t *state, i64 nv, int32_t *maxc, hipStream_t s) {
    if (nv <= 0) return;
    k_jp_finish<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(state, nv, maxc);
    GLV_LAUNCH_CHECK();
}
void launch_color_conflicts(const DevGraph &g, const int32_t *colors, i64 *count, hipStream_t s) {
    if (g.nv <= 0) return;
    k_color_conflicts<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(g.off, g.tails, g.nv, colors, count);
    GLV_LAUNCH_CHECK();
}

void launch_first_index(const i64 *C, i64 n, i64 *first, int32_t *err, hipStream_t s) {
    if (n <= 0) return;
    k_first_index<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(C, n, first, err);
    GLV_LAUNCH_CHECK();
}
void launch_head_flags(const i64 *C, i64 n, const i64 *first, i64 *flags, hipStream_t s) {
    if (n <= 0) return;
    k_head_flags<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(C, n, first, flags);
    GLV_LAUNCH_CHECK();
}
void launch_relabel(i64 *C, i64 n, const i64 *first, const i64 *pos, hipStream_t s) {
    if (n <= 0) return;
    k_relabel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(C, n, first, pos);
    GLV_LAUNCH_CHECK();
}
void launch_compose(i64 *C_orig, i64 n0, const i64 *C, bool first_phase, hipStream_t s) {
    if (n0 <= 0) return;
    k_compose<<<grid_for(n0, kBlock, 8192), kBlock, 0, s>>>(C_orig, n0, C, first_phase);
    GLV_LAUNCH_CHECK();
}

i64 coarsen_chunks(i64 nnz) { return (nnz + kCoarsenChunk - 1) / kCoarsenChunk; }
void launch_coarsen_count(const DevGraph &g, const i64 *C, i64 *ccount, hipStream_t s) {
    const i64 nch = coarsen_chunks(g.nnz);
    if (g.nv <= 0 || nch <= 0) return;
    k_coarsen_count<<<(unsigned)std::min<i64>(nch, 65536), kBlock, 0, s>>>(g.off, g.tails, C, g.nv, g.nnz,
                                                                            ccount);
    GLV_LAUNCH_CHECK();
}
void launch_coarsen_emit(const DevGraph &g, const i64 *C, const i64 *cbase, i64 k, int kb, u64 *keys,
                         double *vals, hipStream_t s, bool vf) {
    const i64 nch = coarsen_chunks(g.nnz);
    i64 blocks = std::max<i64>(nch, vf ? 0 : (k + kBlock - 1) / kBlock);
    blocks = std::max<i64>(1, std::min<i64>(blocks, 65536));
    k_coarsen_emit<<<(unsigned)blocks, kBlock, 0, s>>>(g.off, g.tails, g.w, C, g.nv, g.nnz, cbase, kb, keys,
                                                       vals, vf ? 0 : k, vf);
    GLV_LAUNCH_CHECK();
}
void launch_coarsen_mirror(const u64 *ukeys, const double *uvals, i64 nu, int kb, u64 *keys, double *vals,
                           hipStream_t s, i64 *nself) {
    if (nu <= 0) return;
    k_coarsen_mirror<<<grid_for(nu, kBlock, 8192), kBlock, 0, s>>>(ukeys, uvals, nu, kb, keys, vals,
                                                                   (unsigned long long *)nself);
    GLV_LAUNCH_CHECK();
}
void launch_coarsen_rows(const u64 *keys, const double *vals, i64 total, i64 k, int kb, i64 *off,
                         i64 *tails, double *w, hipStream_t s) {
    if (total <= 0) {
        launch_fill_i64(off, k + 1, 0, s);
        return;
    }
    k_coarsen_rows<<<grid_for(total, kBlock, 8192), kBlock, 0, s>>>(keys, vals, total, k, kb, off, tails,
                                                                    w);
    GLV_LAUNCH_CHECK();
}

void launch_vertex_following(const DevGraph &g, i64 *C, i64 *count, hipStream_t s) {
    if (g.nv <= 0) return;
    k_vertex_following<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(g.off, g.tails, g.nv, C,
                                                                       (unsigned long long *)count);
    GLV_LAUNCH_CHECK();
}
void launch_split_edges(const void *aos, i64 nnz, i64 *tails, double *w, hipStream_t s) {
    if (nnz <= 0) return;
    k_split_edges<<<grid_for(nnz, kBlock, 8192), kBlock, 0, s>>>((const i64 *)aos, nnz, tails, w);
    GLV_LAUNCH_CHECK();
}
void launch_check_labels(const i64 *off, const i64 *C, i64 nv, i64 limit, int32_t *err, hipStream_t s) {
    if (nv <= 0) return;
    k_check_labels<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(off, C, nv, limit, err);
    GLV_LAUNCH_CHECK();
}
void launch_check_graph(const i64 *off, const i64 *tails, i64 nv, i64 nnz, int32_t *err, hipStream_t s) {
    i64 n = std::max(nv, nnz);
    if (n <= 0) return;
    k_check_graph<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(off, tails, nv, nnz, err);
    GLV_LAUNCH_CHECK();
}

// ------------------------------------------------------------------- rocPRIM
void scan_exclusive_i64(const i64 *in, i64 *out, i64 n, hipStream_t s) {
    if (n <= 0) return;
    size_t bytes = 0;
    GLV_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (i64)0, (size_t)n, rocprim::plus<i64>(), s));
    DBuf<char> tmp(bytes, s);
    GLV_HIP(rocprim::exclusive_scan(tmp.p, bytes, in, out, (i64)0, (size_t)n, rocprim::plus<i64>(), s));
}
void sort_pairs_u32_i64(uint32_t *keys, i64 *vals, i64 n, int end_bit, hipStream_t s) {
    if (n <= 1) return;
    DBuf<uint32_t> k2(n, s);
    DBuf<i64> v2(n, s);
    size_t bytes = 0;
    GLV_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, k2.p, vals, v2.p, (size_t)n, 0, end_bit, s));
    DBuf<char> tmp(bytes, s);
    GLV_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, keys, k2.p, vals, v2.p, (size_t)n, 0, end_bit, s));
    GLV_HIP(hipMemcpyAsync(keys, k2.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    GLV_HIP(hipMemcpyAsync(vals, v2.p, n * sizeof(i64), hipMemcpyDeviceToDevice, s));
}
void sort_pairs_u64_f64(u64 *keys, double *vals, i64 n, int end_bit, hipStream_t s) {
    if (n <= 1) return;
    DBuf<u64> k2(n, s);
    DBuf<double> v2(n, s);
    size_t bytes = 0;
    GLV_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, k2.p, vals, v2.p, (size_t)n, 0, end_bit, s));
    DBuf<char> tmp(bytes, s);
    GLV_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, keys, k2.p, vals, v2.p, (size_t)n, 0, end_bit, s));
    GLV_HIP(hipMemcpyAsync(keys, k2.p, n * sizeof(u64), hipMemcpyDeviceToDevice, s));
    GLV_HIP(hipMemcpyAsync(vals, v2.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
}
i64 reduce_by_key_u64_f64(const u64 *keys, const double *vals, i64 n, u64 *ukeys, double *usums,
                          hipStream_t s) {
    if (n <= 0) return 0;
    DBuf<size_t> cnt(1, s);
    size_t bytes = 0;
    GLV_HIP(rocprim::reduce_by_key(nullptr, bytes, keys, vals, (size_t)n, ukeys, usums, cnt.p,
                                   rocprim::plus<double>(), rocprim::equal_to<u64>(), s));
    DBuf<char> tmp(bytes, s);
    GLV_HIP(rocprim::reduce_by_key(tmp.p, bytes, keys, vals, (size_t)n, ukeys, usums, cnt.p,
                                   rocprim::plus<double>(), rocprim::equal_to<u64>(), s));
    size_t h = 0;
    GLV_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(size_t), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    return (i64)h;
}

}  // namespace glv
