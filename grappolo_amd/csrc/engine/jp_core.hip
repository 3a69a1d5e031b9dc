// jp_core.hip -- the core of the Jones-Plassmann coloring and the depth cut.
//
// The engine's coloring is Jones-Plassmann with degree-first priorities plus
// a sequential tail (the restated algorithm: colour every vertex whose
// priority beats all uncoloured neighbours with the smallest colour its
// coloured neighbours do not use, round after round, until at most kJpTail
// vertices remain; those take first-fit colours one at a time in increasing
// id).  A vertex coloured in round r takes the smallest colour absent from
// its higher-priority neighbours, all coloured in earlier rounds, and r is
// its depth in the priority DAG (1 + the largest depth of its higher
// neighbours).  So the colours of the rounds and the tail set follow from the
// DAG alone, whatever order computes them, as long as a vertex is coloured
// after its higher neighbours.
//
// On skewed graphs the DAG is thousands of rounds deep through the hubs, and
// a round of the whole GPU costs a grid barrier plus a hub's full adjacency.
// The core -- the vertices of degree dcore or more, a prefix of the priority
// order since the degree leads the priority -- is coloured here by one
// workgroup: its vertices' core neighbours are listed (16-bit core indices,
// higher ones first), a vertex's state (uncoloured: its count of uncoloured
// higher core neighbours; coloured: depth and colour) lives in LDS, and the
// rounds are separated by workgroup barriers only.  The frontier rounds then
// colour the rest, taking the core's colours and depths as given.  Finally
// the depths decide the tail: the smallest round r with at most kJpTail
// vertices deeper than r, those vertices uncoloured and handed to the
// sequential first-fit.
#include "kernels.hpp"

namespace glv {
namespace {

#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

constexpr int kCoreBlock = 1024;
constexpr int kCoreWaves = kCoreBlock / kWave;
constexpr int kColorWords = kMaxColors / 32;
constexpr uint32_t kColoured = 0x80000000u;  // state word: coloured (depth << 12 | colour)

static_assert(kMaxColors <= 4096, "colours take 12 bits of the core state");
static_assert(kJpCoreMax < (1 << 16), "core indices are 16-bit");

// degree bucket: degrees below 8 exactly, then 8 buckets per octave
__device__ inline int deg_bucket(i64 d) {
    if (d < 8) return (int)d;
    const int e = 63 - __clzll((long long)d);
    return 8 + (e - 3) * 8 + (int)((d >> (e - 3)) & 7);
}

__global__ __launch_bounds__(kBlock) void k_jp_degree_hist(const i64 *off, i64 nv, unsigned long long *hist) {
    __shared__ unsigned s_h[kJpDegBuckets];
    for (int i = threadIdx.x; i < kJpDegBuckets; i += kBlock) s_h[i] = 0;
    __syncthreads();
    for (i64 v = blockIdx.x * (i64)kBlock + threadIdx.x; v < nv; v += (i64)gridDim.x * kBlock)
        atomicAdd(&s_h[deg_bucket(off[v + 1] - off[v])], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kJpDegBuckets; i += kBlock)
        if (s_h[i]) atomicAdd(&hist[i], (unsigned long long)s_h[i]);
}

__global__ void k_jp_core_flags(const i64 *off, i64 nv, u64 dcore, i64 *flags) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        flags[v] = (u64)(off[v + 1] - off[v]) >= dcore ? 1 : 0;
}
__global__ void k_jp_core_scatter(const i64 *flags, const i64 *pos, i64 nv, i64 *hlist, int32_t *cnt) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        if (flags[v]) {
            hlist[pos[v]] = v;
            cnt[v] = (int32_t)pos[v];
        }
}
__global__ void k_jp_core_sizes(const int32_t *core_cnt, i64 K, i64 *sizes) {
    for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q <= K; q += (i64)gridDim.x * blockDim.x)
        sizes[q] = q < K ? (i64)core_cnt[2 * q] + core_cnt[2 * q + 1] : 0;
}

// One workgroup per core vertex q: its core neighbours' indices (count[u] of
// a core vertex u is its index) into ladj, the higher ones from loff[q], the
// lower ones from loff[q] + core_cnt[2q]; the order inside each part does
// not matter (colours, depths and counts only take sets and sums of them).
constexpr int kFillBlock = 1024;
constexpr int kFillUnroll = 4;
__global__ __launch_bounds__(kFillBlock) void k_jp_core_fill(const i64 *off, const i64 *tails, const uint8_t *code,
                                                             const int32_t *cnt, const i64 *hlist, i64 K,
                                                             const i64 *loff, const int32_t *core_cnt,
                                                             uint16_t *ladj) {
    __shared__ int s_hi, s_lo;
    const int lane = threadIdx.x % kWave;
    const u64 below = (1ull << lane) - 1;
    for (i64 q = blockIdx.x; q < K; q += gridDim.x) {
        const i64 v = hlist[q];
        const i64 b = off[v], deg = off[v + 1] - b;
        const i64 hb = loff[q], lb = loff[q] + core_cnt[2 * q];
        if (threadIdx.x == 0) {
            s_hi = 0;
            s_lo = 0;
        }
        __syncthreads();
        for (i64 j0 = 0; j0 < deg; j0 += kFillUnroll * kFillBlock) {
            uint8_t cd[kFillUnroll];
#pragma unroll
            for (int k = 0; k < kFillUnroll; k++) {
                const i64 j = j0 + k * kFillBlock + threadIdx.x;
                cd[k] = j < deg ? code[b + j] : kJpSelf;
            }
            int x[kFillUnroll];
#pragma unroll
            for (int k = 0; k < kFillUnroll; k++) {
                const bool core = (cd[k] & (kJpCore | kJpSelf)) == kJpCore;
                x[k] = core ? cnt[tails[b + j0 + k * kFillBlock + threadIdx.x]] : 0;
            }
#pragma unroll
            for (int k = 0; k < kFillUnroll; k++) {
                const bool core = (cd[k] & (kJpCore | kJpSelf)) == kJpCore;
                const bool lo = core && (cd[k] & kJpRelease), hi = core && !(cd[k] & kJpRelease);
                const u64 bh = __ballot(hi), bl = __ballot(lo);
                int ph = 0, pl = 0;
                if (lane == 0) {
                    if (bh) ph = atomicAdd(&s_hi, __popcll(bh));
                    if (bl) pl = atomicAdd(&s_lo, __popcll(bl));
                }
                ph = __shfl(ph, 0);
                pl = __shfl(pl, 0);
                if (hi) ladj[hb + ph + __popcll(bh & below)] = (uint16_t)x[k];
                if (lo) ladj[lb + pl + __popcll(bl & below)] = (uint16_t)x[k];
            }
        }
        __syncthreads();
    }
}

// The core pass, one workgroup, in rounds: the vertices queued in round t
// (their higher neighbours all coloured) are pairwise non-adjacent, so they
// are coloured together and their releases queue round t + 1.  st[q]: while
// q is uncoloured, its count of uncoloured higher core neighbours; then
// kColoured | depth << 12 | colour.  A round of at least kCoreWaves vertices
// gives each wavefront its share of them, one at a time; a smaller round
// gives each vertex a group of G wavefronts (G = kCoreWaves / n rounded down
// to a power of two), so a round along a chain of hubs spreads the hub's
// list over the whole workgroup: one load of it, the colour choice and three
// workgroup barriers.  With stop >= 0 (the core is the whole graph) the
// rounds end after the first one that leaves at most stop vertices
// uncoloured -- the tail's, first-fit afterwards -- which get colour -1 and
// depth rounds + 1, so the depth cut finds the same round.
constexpr int kCoreRegs = 4;  // list entries per thread held across the two passes
__global__ __launch_bounds__(kCoreBlock) void k_jp_core(const i64 *hlist, int K, const i64 *loff,
                                                        const int32_t *core_cnt, const uint16_t *ladj,
                                                        int32_t *colors, int32_t *dep, int32_t *err, int stop) {
    __shared__ uint32_t st[kJpCoreMax];
    __shared__ uint16_t queue[kJpCoreMax];
    __shared__ uint32_t mk[kCoreWaves][kColorWords];
    __shared__ int s_dm[kCoreWaves];
    __shared__ int s_tail;
    const int tid = threadIdx.x, lane = tid % kWave, w = tid / kWave;
    const u64 below = (1ull << lane) - 1;
    for (int q = tid; q < K; q += kCoreBlock) st[q] = (uint32_t)core_cnt[2 * q];
    if (tid == 0) s_tail = 0;
    __syncthreads();
    // appends the ready lanes' vertices to the next round (whole wavefront)
    auto enqueue = [&](bool ready, int x) {
        const u64 bal = __ballot(ready);
        if (!bal) return;
        int base = 0;
        if (lane == 0) base = atomicAdd(&s_tail, __popcll(bal));
        base = __shfl(base, 0);
        if (ready) queue[base + __popcll(bal & below)] = (uint16_t)x;
    };
    // first free colour of bitmap m (two words per lane) and the largest
    // depth dm, over the wavefront; valid in every lane
    auto first_free = [&](const uint32_t *m, int &dm) {
        int cand = kMaxColors;
        const uint32_t m0 = m[lane], m1 = m[lane + kWave];
        if (m0 != 0xFFFFFFFFu) cand = lane * 32 + __ffs((int)~m0) - 1;
        else if (m1 != 0xFFFFFFFFu) cand = (lane + kWave) * 32 + __ffs((int)~m1) - 1;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const int oc = __shfl_xor(cand, o);
            cand = oc < cand ? oc : cand;
            const int od = __shfl_xor(dm, o);
            dm = od > dm ? od : dm;
        }
        return cand;
    };
    auto state_of = [&](int color, int dm) {
        if (color >= kMaxColors) {
            atomicOr(err, 1);
            color = 0;
        }
        return kColoured | ((uint32_t)(dm + 1) << 12) | (uint32_t)color;
    };
    // colours and depth of the higher neighbours x (entry j of a list whose
    // first nhi entries are the higher ones) into bitmap m and dm
    auto mark = [&](int x, int j, int nhi, uint32_t *m, int &dm) {
        if (x < 0 || j >= nhi) return;
        const uint32_t sx = st[x];
        const int c = (int)(sx & 0xFFFu), d = (int)((sx >> 12) & 0x7FFFFu);
        atomicOr(&m[c >> 5], 1u << (c & 31));
        dm = d > dm ? d : dm;
    };
    for (int q0 = 0; q0 < K; q0 += kCoreBlock) {
        const int q = q0 + tid;
        enqueue(q < K && core_cnt[2 * q] == 0, q);
    }
    int head = 0, rounds = 0;
    bool stopped = false;
    while (true) {
        __syncthreads();  // the previous round is complete
        const int tail = s_tail;
        __syncthreads();  // and every thread knows where it ends before this round appends
        const int n = tail - head;
        if (n <= 0) break;
        // after round 1 or later, at most `stop` vertices left: they are the tail's
        if (stop >= 0 && rounds > 0 && K - head <= stop) {
            stopped = true;
            break;
        }
        rounds++;
        if (n >= kCoreWaves) {
            // a wavefront per vertex
            uint32_t *m = mk[w];
            for (int i = head + w; i < tail; i += kCoreWaves) {
                const int r = queue[i];
                const i64 b = loff[r];
                const int len = (int)(loff[r + 1] - b), nhi = core_cnt[2 * r];
                m[lane] = 0;
                m[lane + kWave] = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                int dm = 0;
                for (int j0 = 0; j0 < nhi; j0 += kCoreRegs * kWave) {
                    int x[kCoreRegs];
#pragma unroll
                    for (int k = 0; k < kCoreRegs; k++) {
                        const int j = j0 + k * kWave + lane;
                        x[k] = j < nhi ? (int)ladj[b + j] : -1;
                    }
#pragma unroll
                    for (int k = 0; k < kCoreRegs; k++) mark(x[k], j0 + k * kWave + lane, nhi, m, dm);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                const int color = first_free(m, dm);
                if (lane == 0) st[r] = state_of(color, dm);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                for (int j0 = nhi; j0 < len; j0 += kCoreRegs * kWave) {
                    int x[kCoreRegs];
#pragma unroll
                    for (int k = 0; k < kCoreRegs; k++) {
                        const int j = j0 + k * kWave + lane;
                        x[k] = j < len ? (int)ladj[b + j] : -1;
                    }
#pragma unroll
                    for (int k = 0; k < kCoreRegs; k++) enqueue(x[k] >= 0 && atomicSub(&st[x[k]], 1u) == 1u, x[k]);
                }
            }
        } else {
            // a group of G wavefronts per vertex, one vertex per group
            const int G = n == 1 ? 16 : n == 2 ? 8 : n <= 4 ? 4 : n <= 8 ? 2 : 1;
            const int g = w / G, gw = w % G, gt = gw * kWave + lane, gstride = G * kWave;
            const bool act = g < n;
            const int r = act ? queue[head + g] : 0;
            const i64 b = act ? loff[r] : 0;
            const int len = act ? (int)(loff[r + 1] - b) : 0, nhi = act ? core_cnt[2 * r] : 0;
            uint32_t *m = mk[g * G];
            if (act && gw == 0) {
                m[lane] = 0;
                m[lane + kWave] = 0;
                if (lane == 0) s_dm[g] = 0;
            }
            // the first kCoreRegs entries per thread of the whole list, for both passes
            int x[kCoreRegs];
#pragma unroll
            for (int k = 0; k < kCoreRegs; k++) {
                const int j = k * gstride + gt;
                x[k] = j < len ? (int)ladj[b + j] : -1;
            }
            __syncthreads();
            int dm = 0;
#pragma unroll
            for (int k = 0; k < kCoreRegs; k++) mark(x[k], k * gstride + gt, nhi, m, dm);
            for (int j0 = kCoreRegs * gstride; j0 < nhi; j0 += kCoreRegs * gstride) {
                int y[kCoreRegs];
#pragma unroll
                for (int k = 0; k < kCoreRegs; k++) {
                    const int j = j0 + k * gstride + gt;
                    y[k] = j < nhi ? (int)ladj[b + j] : -1;
                }
#pragma unroll
                for (int k = 0; k < kCoreRegs; k++) mark(y[k], j0 + k * gstride + gt, nhi, m, dm);
            }
#pragma unroll
            for (int o = kWave / 2; o > 0; o >>= 1) {
                const int od = __shfl_xor(dm, o);
                dm = od > dm ? od : dm;
            }
            if (act && lane == 0 && dm > 0) atomicMax(&s_dm[g], dm);
            __syncthreads();
            if (act && gw == 0) {
                int dmax = s_dm[g];
                const int color = first_free(m, dmax);
                if (lane == 0) st[r] = state_of(color, dmax);
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kCoreRegs; k++) {
                const int j = k * gstride + gt;
                enqueue(x[k] >= 0 && j >= nhi && atomicSub(&st[x[k]], 1u) == 1u, x[k]);
            }
            for (int j0 = kCoreRegs * gstride > nhi ? kCoreRegs * gstride : nhi; j0 < len;
                 j0 += kCoreRegs * gstride) {
                int y[kCoreRegs];
#pragma unroll
                for (int k = 0; k < kCoreRegs; k++) {
                    const int j = j0 + k * gstride + gt;
                    y[k] = j < len ? (int)ladj[b + j] : -1;
                }
#pragma unroll
                for (int k = 0; k < kCoreRegs; k++) enqueue(y[k] >= 0 && atomicSub(&st[y[k]], 1u) == 1u, y[k]);
            }
        }
        head = tail;
    }
    for (int q = tid; q < K; q += kCoreBlock) {
        const uint32_t sq = st[q];
        const i64 v = hlist[q];
        if (sq & kColoured) {
            colors[v] = (int32_t)(sq & 0xFFFu);
            dep[v] = (int32_t)((sq >> 12) & 0x7FFFFu);
        } else if (stopped) {
            colors[v] = -1;  // left for the tail
            dep[v] = rounds + 1;
        } else {
            atomicOr(err, 4);  // never queued: the counts do not match the lists
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_jp_depth_max(const int32_t *dep, i64 nv, int32_t *maxd) {
    int m = 0;
    for (i64 v = blockIdx.x * (i64)kBlock + threadIdx.x; v < nv; v += (i64)gridDim.x * kBlock)
        m = dep[v] > m ? dep[v] : m;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const int x = __shfl_xor(m, o);
        m = x > m ? x : m;
    }
    if (threadIdx.x % kWave == 0 && m > 0) atomicMax(maxd, m);
}
__global__ __launch_bounds__(kBlock) void k_jp_depth_hist(const int32_t *dep, i64 nv, int32_t top, unsigned *hist) {
    __shared__ unsigned s_h[kJpDepthWindow];
    for (int i = threadIdx.x; i < kJpDepthWindow; i += kBlock) s_h[i] = 0;
    __syncthreads();
    for (i64 v = blockIdx.x * (i64)kBlock + threadIdx.x; v < nv; v += (i64)gridDim.x * kBlock) {
        const int i = top - dep[v];
        if (i >= 0 && i < kJpDepthWindow) atomicAdd(&s_h[i], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kJpDepthWindow; i += kBlock)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}
__global__ void k_jp_depth_cut(const int32_t *dep, i64 nv, int32_t rounds, int32_t *colors) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        if (dep[v] > rounds) colors[v] = -1;
}

}  // namespace

i64 jp_bucket_lo(int b) {
    if (b < 8) return b;
    const int e = 3 + (b - 8) / 8, m = (b - 8) % 8;
    return (i64)(8 + m) << (e - 3);
}

void launch_jp_degree_hist(const DevGraph &g, unsigned long long *hist, hipStream_t s) {
    if (g.nv <= 0) return;
    k_jp_degree_hist<<<grid_for(g.nv, kBlock, 2048), kBlock, 0, s>>>(g.off, g.nv, hist);
    GLV_LAUNCH_CHECK();
}
void launch_jp_core_list(const DevGraph &g, u64 dcore, i64 *flags, i64 *pos, i64 *hlist, int32_t *cnt,
                         hipStream_t s) {
    if (g.nv <= 0) return;
    k_jp_core_flags<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(g.off, g.nv, dcore, flags);
    GLV_LAUNCH_CHECK();
    scan_exclusive_i64(flags, pos, g.nv, s);
    k_jp_core_scatter<<<grid_for(g.nv, kBlock, 8192), kBlock, 0, s>>>(flags, pos, g.nv, hlist, cnt);
    GLV_LAUNCH_CHECK();
}
void launch_jp_core_sizes(const int32_t *core_cnt, i64 K, i64 *sizes, hipStream_t s) {
    k_jp_core_sizes<<<grid_for(K + 1, kBlock, 1024), kBlock, 0, s>>>(core_cnt, K, sizes);
    GLV_LAUNCH_CHECK();
}
void launch_jp_core_fill(const DevGraph &g, const uint8_t *code, const int32_t *cnt, const i64 *hlist, i64 K,
                         const i64 *loff, const int32_t *core_cnt, uint16_t *ladj, hipStream_t s) {
    if (K <= 0) return;
    const unsigned blocks = (unsigned)std::min<i64>(K, 4096);
    k_jp_core_fill<<<blocks, kFillBlock, 0, s>>>(g.off, g.tails, code, cnt, hlist, K, loff, core_cnt, ladj);
    GLV_LAUNCH_CHECK();
}
void launch_jp_core(const i64 *hlist, i64 K, const i64 *loff, const int32_t *core_cnt, const uint16_t *ladj,
                    int32_t *colors, int32_t *dep, int32_t *err, int stop, hipStream_t s) {
    if (K <= 0) return;
    if (K > kJpCoreMax) {  // callers pass at most kJpCoreMax vertices
        launch_fill_i32(err, 1, 4, s);
        return;
    }
    k_jp_core<<<1, kCoreBlock, 0, s>>>(hlist, (int)K, loff, core_cnt, ladj, colors, dep, err, stop);
    GLV_LAUNCH_CHECK();
}
void launch_jp_depth_max(const int32_t *dep, i64 nv, int32_t *maxd, hipStream_t s) {
    if (nv <= 0) return;
    k_jp_depth_max<<<grid_for(nv, kBlock, 2048), kBlock, 0, s>>>(dep, nv, maxd);
    GLV_LAUNCH_CHECK();
}
void launch_jp_depth_hist(const int32_t *dep, i64 nv, int32_t top, unsigned *hist, hipStream_t s) {
    if (nv <= 0) return;
    k_jp_depth_hist<<<grid_for(nv, kBlock, 1024), kBlock, 0, s>>>(dep, nv, top, hist);
    GLV_LAUNCH_CHECK();
}
void launch_jp_depth_cut(const int32_t *dep, i64 nv, int32_t rounds, int32_t *colors, hipStream_t s) {
    if (nv <= 0) return;
    k_jp_depth_cut<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(dep, nv, rounds, colors);
    GLV_LAUNCH_CHECK();
}

}  // namespace glv
