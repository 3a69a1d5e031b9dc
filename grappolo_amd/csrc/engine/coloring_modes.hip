// coloring_modes.hip -- the other colorings of runMultiPhaseColoring
// (Coloring/runMultiPhaseColoring.cpp:68-87, :183-197):
//   coloring = 3: incomplete multi-hash max/min coloring
//                 (Coloring/coloringMultiHashMaxMin.cpp:60-226);
//   coloring = 2: distance-1 coloring rebalanced so no class holds far more
//                 than NV / ncolors vertices (Coloring/vBase.cpp:6-168).
// Both are HBM-bound neighbour scans; one wavefront serves one vertex whose
// degree exceeds kMhThreadDeg, one lane each of the others.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "batch.hpp"
#include "coloring_modes.hpp"
#include "kernels.hpp"

#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

namespace glv {

namespace {

constexpr int kCB = 256;
constexpr i64 kMhThreadDeg = 32;  // lane-per-vertex up to this degree

__device__ inline u64 mix64(u64 z) {
    // splitmix64 finaliser: a bijection of 64-bit words
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// random value of vertex v under hash h: the top 53 bits of splitmix64 of
// v ^ ((h + 1) << 56), exactly representable as a double
__device__ inline double mh_value(int h, i64 v) {
    return (double)(mix64((u64)(v ^ ((i64)(h + 1) << 56))) >> 11);
}

__device__ inline int block_count_add(int v, unsigned long long *acc) {
    __shared__ int sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    if (v) atomicAdd(&sh, v);
    __syncthreads();
    if (threadIdx.x == 0 && sh) atomicAdd(acc, (unsigned long long)sh);
    return 0;
}

// ---------------------------------------------------------------- multi-hash
// One (itr, hash) pass over the vertices of degree <= kMhThreadDeg, a lane
// each.  A pass reads colors < cur only to skip them; it writes cur or cur+1,
// so a neighbour coloured during the pass is compared either way and the
// order of the lanes does not matter.
__global__ __launch_bounds__(kCB) void k_mh_pass_light(const i64 *off, const i64 *tails, i64 nv, int32_t *colors,
                                                       int maxColor, int cur, int h, unsigned long long *count) {
    int got = 0;
    for (i64 v = blockIdx.x * (i64)kCB + threadIdx.x; v < nv; v += (i64)gridDim.x * kCB) {
        const i64 a = off[v], b = off[v + 1];
        if (b - a > kMhThreadDeg || colors[v] != maxColor) continue;
        const double rv = mh_value(h, v);
        bool isMax = true, isMin = true;
        for (i64 j = a; j < b; j++) {
            const i64 u = tails[j];
            if (u == v || colors[u] < cur) continue;
            const double ru = mh_value(h, u);
            isMax = isMax && !(rv <= ru);
            isMin = isMin && !(rv >= ru);
        }
        if (isMax) {
            colors[v] = cur;
            got++;
        } else if (isMin) {
            colors[v] = cur + 1;
            got++;
        }
    }
    block_count_add(got, count);
}

// the same for the heavy vertices (list), one wavefront each
__global__ __launch_bounds__(kCB) void k_mh_pass_heavy(const i64 *off, const i64 *tails, const i64 *list, i64 n,
                                                       int32_t *colors, int maxColor, int cur, int h,
                                                       unsigned long long *count) {
    const int lane = threadIdx.x % kWave;
    const i64 wid = (blockIdx.x * (i64)kCB + threadIdx.x) / kWave, nw = (i64)gridDim.x * kCB / kWave;
    int got = 0;
    for (i64 q = wid; q < n; q += nw) {
        const i64 v = list[q];
        if (colors[v] != maxColor) continue;  // wave-uniform
        const double rv = mh_value(h, v);
        bool isMax = true, isMin = true;
        for (i64 j = off[v] + lane; j < off[v + 1]; j += kWave) {
            const i64 u = tails[j];
            if (u == v || colors[u] < cur) continue;
            const double ru = mh_value(h, u);
            isMax = isMax && !(rv <= ru);
            isMin = isMin && !(rv >= ru);
        }
        isMax = __all(isMax);
        isMin = __all(isMin);
        if (lane == 0) {
            if (isMax) {
                colors[v] = cur;
                got++;
            } else if (isMin) {
                colors[v] = cur + 1;
                got++;
            }
        }
    }
    block_count_add(got, count);
}

__global__ void k_fill_i32v(int32_t *p, i64 n, int32_t v) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_flag_heavy(const i64 *off, i64 nv, i64 thresh, i64 *flags) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x)
        flags[v] = off[v + 1] - off[v] > thresh ? 1 : 0;
}

__global__ void k_scatter_flag(const i64 *flags, const i64 *pos, i64 n, i64 *out) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < n; v += (i64)gridDim.x * blockDim.x)
        if (flags[v]) out[pos[v]] = v;
}

// vertices with flags[v] != 0, in increasing id; returns the count
i64 list_flagged(const i64 *flags, i64 n, DBuf<i64> &list, hipStream_t s) {
    DBuf<i64> pos(std::max<i64>(n, 1), s);
    scan_exclusive_i64(flags, pos.p, n, s);
    i64 h[2] = {0, 0};
    GLV_HIP(hipMemcpyAsync(&h[0], pos.p + n - 1, sizeof(i64), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipMemcpyAsync(&h[1], flags + n - 1, sizeof(i64), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    const i64 m = h[0] + h[1];
    list.alloc(std::max<i64>(m, 1), s);
    k_scatter_flag<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(flags, pos.p, n, list.p);
    GLV_LAUNCH_CHECK();
    return m;
}

// ------------------------------------------------------------------ balanced
// eligible colours of a round: bit c of elig = (!over[c] && freq[c] < avg)
__global__ void k_bal_eligible(const i64 *freq, const uint8_t *over, int ncol, i64 avg, unsigned long long *elig,
                               int *n_elig) {
    const int w = threadIdx.x;  // one 64-bit word per lane, ncol <= 64 * blockDim.x
    unsigned long long bits = 0;
    for (int b = 0; b < 64; b++) {
        const int c = w * 64 + b;
        if (c < ncol && !over[c] && freq[c] < avg) bits |= 1ull << b;
    }
    elig[w] = bits;
    if (bits) atomicAdd(n_elig, __popcll(bits));
}

// wish of every candidate (base class oversize, current class above avg):
// among the eligible colours no neighbour holds, the first (type 0) or the
// one of fewest vertices, the smaller colour on ties (type 1); -1 if none.
// One wavefront per vertex with its neighbours' colours marked in an LDS
// bitmap.
template <int NW>
__global__ __launch_bounds__(kCB) void k_bal_wish(const i64 *off, const i64 *tails, i64 nv, const int32_t *colors,
                                                  const int32_t *base, const uint8_t *over, const i64 *freq, i64 avg,
                                                  const unsigned long long *elig, int type, int32_t *wish) {
    __shared__ unsigned long long mark[kCB / kWave][NW];
    const int lane = threadIdx.x % kWave, wl = threadIdx.x / kWave;
    const i64 wid = (blockIdx.x * (i64)kCB + threadIdx.x) / kWave, nw = (i64)gridDim.x * kCB / kWave;
    for (i64 v = wid; v < nv; v += nw) {
        const bool cand = over[base[v]] && freq[colors[v]] > avg;
        if (!cand) {
            if (lane == 0) wish[v] = -1;
            continue;
        }
        for (int q = lane; q < NW; q += kWave) mark[wl][q] = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        for (i64 j = off[v] + lane; j < off[v + 1]; j += kWave) {
            const i64 u = tails[j];
            if (u == v) continue;
            const int c = colors[u];
            atomicOr(&mark[wl][c >> 6], 1ull << (c & 63));
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        // key = (size << 12 | colour) for type 1, colour for type 0; minimum wins
        i64 best = INT64_MAX;
        for (int q = lane; q < NW; q += kWave) {
            unsigned long long free_bits = elig[q] & ~mark[wl][q];
            if (type == 0) {
                if (free_bits) best = std::min<i64>(best, q * 64 + __ffsll((long long)free_bits) - 1);
            } else {
                while (free_bits) {
                    const int b = __ffsll((long long)free_bits) - 1;
                    free_bits &= free_bits - 1;
                    const int c = q * 64 + b;
                    best = std::min<i64>(best, (freq[c] << 12) | c);
                }
            }
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) best = std::min<i64>(best, __shfl_xor(best, o));
        if (lane == 0) wish[v] = best == INT64_MAX ? -1 : (int32_t)(best & 4095);
        __builtin_amdgcn_wave_barrier();
    }
}

// a wish survives unless an adjacent vertex with the same wish has a higher
// priority; survivors are flagged
__global__ __launch_bounds__(kCB) void k_bal_block(const i64 *off, const i64 *tails, i64 nv, const int32_t *wish,
                                                   i64 *keep) {
    const int lane = threadIdx.x % kWave;
    const i64 wid = (blockIdx.x * (i64)kCB + threadIdx.x) / kWave, nw = (i64)gridDim.x * kCB / kWave;
    for (i64 v = wid; v < nv; v += nw) {
        const int c = wish[v];
        if (c < 0) {
            if (lane == 0) keep[v] = 0;
            continue;
        }
        const u64 pv = mix64((u64)v);
        bool ok = true;
        for (i64 j = off[v] + lane; j < off[v + 1]; j += kWave) {
            const i64 u = tails[j];
            if (u != v && wish[u] == c && mix64((u64)u) > pv) ok = false;
        }
        ok = __all(ok);
        if (lane == 0) keep[v] = ok ? 1 : 0;
    }
}

// sort keys of the listed vertices: ~priority (so ascending = decreasing
// priority) and the segment (wish, or the class being left)
__global__ void k_bal_keys(const i64 *list, i64 n, const int32_t *seg_of, u64 *pkey, uint32_t *skey, i64 *idx) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const i64 v = list[i];
        pkey[i] = ~mix64((u64)v);
        skey[i] = (uint32_t)seg_of[v];
        idx[i] = v;
    }
}

__global__ void k_gather_seg(const i64 *vs, i64 n, const int32_t *seg_of, uint32_t *skey) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        skey[i] = (uint32_t)seg_of[vs[i]];
}

__global__ void k_seg_start(const uint32_t *skey, i64 n, i64 *start) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        if (i == 0 || skey[i] != skey[i - 1]) start[skey[i]] = i;
}

// mode 0: admitted when rank < avg - freq[seg] (capacity of the wished colour)
// mode 1: kept when rank < freq[seg] - avg (departures of the class left)
__global__ void k_seg_admit(const uint32_t *skey, const i64 *vs, i64 n, const i64 *start, const i64 *freq, i64 avg,
                            int mode, i64 *flag) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const uint32_t c = skey[i];
        const i64 rank = i - start[c];
        const i64 lim = mode == 0 ? avg - freq[c] : freq[c] - avg;
        flag[i] = rank < lim ? 1 : 0;
    }
}

__global__ void k_bal_apply(const i64 *vs, const i64 *flag, i64 n, const int32_t *wish, int32_t *colors,
                            unsigned long long *freq, unsigned long long *moved) {
    int got = 0;
    for (i64 i = blockIdx.x * (i64)kCB + threadIdx.x; i < n; i += (i64)gridDim.x * kCB) {
        if (!flag[i]) continue;
        const i64 v = vs[i];
        const int from = colors[v], to = wish[v];
        atomicAdd(freq + from, (unsigned long long)-1ll);
        atomicAdd(freq + to, 1ull);
        colors[v] = to;
        got++;
    }
    block_count_add(got, moved);
}

__global__ void k_color_sizes(const int32_t *colors, i64 nv, int ncol, unsigned long long *freq, int *err) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        const int c = colors[v];
        if (c < 0 || c >= ncol) {
            atomicOr(err, 1);
            continue;
        }
        atomicAdd(freq + c, 1ull);
    }
}

__global__ void k_over_flags(const i64 *freq, int ncol, i64 avg, uint8_t *over) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x)
        over[c] = freq[c] > avg ? 1 : 0;
}

template <typename T>
T fetch_one(const T *d, hipStream_t s) {
    T h;
    GLV_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    return h;
}

// stable ordering of the listed items by (segment ascending, priority
// descending): sort by priority, then stably by segment
void order_by_segment(u64 *pkey, uint32_t *skey, i64 *vs, i64 n, int seg_bits, const int32_t *seg_of,
                      hipStream_t s) {
    if (n <= 1) return;
    {
        DBuf<u64> k2(n, s);
        DBuf<i64> v2(n, s);
        size_t bytes = 0;
        GLV_HIP(rocprim::radix_sort_pairs(nullptr, bytes, pkey, k2.p, vs, v2.p, (size_t)n, 0, 64, s));
        DBuf<char> tmp(bytes, s);
        GLV_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, pkey, k2.p, vs, v2.p, (size_t)n, 0, 64, s));
        GLV_HIP(hipMemcpyAsync(vs, v2.p, n * sizeof(i64), hipMemcpyDeviceToDevice, s));
    }
    k_gather_seg<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(vs, n, seg_of, skey);
    GLV_LAUNCH_CHECK();
    sort_pairs_u32_i64(skey, vs, n, seg_bits, s);  // radix sort: stable
}

int bits_for_int(int v) {
    int b = 1;
    while ((1 << b) <= v) b++;
    return b;
}

}  // namespace

int color_multihash(const DevGraph &g, int32_t *colors, int nHash, int nItrs, hipStream_t s) {
    if (nHash <= 0 || nItrs <= 0) fail(GLV_E_INVALID_ARG, "multi-hash coloring needs nHash, nItrs > 0");
    const i64 nv = g.nv;
    int maxColor = 2 * nHash * nItrs;
    if (nv <= 0) return maxColor;
    k_fill_i32v<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(colors, nv, maxColor);
    GLV_LAUNCH_CHECK();
    DBuf<i64> flags(nv, s), heavy;
    k_flag_heavy<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(g.off, nv, kMhThreadDeg, flags.p);
    GLV_LAUNCH_CHECK();
    const i64 nheavy = list_flagged(flags.p, nv, heavy, s);
    flags.release();
    DBuf<unsigned long long> cnt(1, s);
    i64 totalColored = 0;
    for (int itr = 0; itr < nItrs; itr++) {
        launch_fill_i64((i64 *)cnt.p, 1, 0, s);
        for (int h = 0; h < nHash; h++) {
            const int cur = 2 * itr * nHash + 2 * h;
            k_mh_pass_light<<<grid_for(nv, kCB, 4096), kCB, 0, s>>>(g.off, g.tails, nv, colors, maxColor, cur, h,
                                                                    cnt.p);
            GLV_LAUNCH_CHECK();
            if (nheavy > 0) {
                const unsigned blocks = (unsigned)std::min<i64>((nheavy + kCB / kWave - 1) / (kCB / kWave), 4096);
                k_mh_pass_heavy<<<blocks, kCB, 0, s>>>(g.off, g.tails, heavy.p, nheavy, colors, maxColor, cur, h,
                                                       cnt.p);
                GLV_LAUNCH_CHECK();
            }
        }
        const i64 iterFreq = (i64)fetch_one(cnt.p, s);
        totalColored += iterFreq;
        // coloringMultiHashMaxMin.cpp:165-173
        if (iterFreq == 0 && totalColored == nv) {
            maxColor = (2 * (itr - 1) * nHash) + 2 * nHash + 1;
            break;
        }
    }
    return maxColor;
}

int balance_colors(const DevGraph &g, int32_t *colors, int ncolors, int type, hipStream_t s) {
    if (ncolors <= 0 || ncolors > kMaxColors) fail(GLV_E_INVALID_ARG, "balanced coloring: bad color count");
    if (type != 0 && type != 1) fail(GLV_E_INVALID_ARG, "balanced coloring: type is 0 (first fit) or 1 (least used)");
    const i64 nv = g.nv;
    if (nv <= 0) return 0;
    const i64 avg = (i64)std::ceil((double)nv / (double)ncolors);
    constexpr int nwords = kMaxColors / 64;  // the eligibility bitmap always spans kMaxColors
    DBuf<i64> freq(ncolors, s);
    DBuf<uint8_t> over(ncolors, s);
    DBuf<int32_t> base(nv, s), wish(nv, s), err(1, s);
    DBuf<unsigned long long> elig(nwords, s), moved(1, s);
    DBuf<int> n_elig(1, s);
    launch_fill_i64(freq.p, ncolors, 0, s);
    launch_fill_i32(err.p, 1, 0, s);
    k_color_sizes<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(colors, nv, ncolors, (unsigned long long *)freq.p,
                                                                err.p);
    GLV_LAUNCH_CHECK();
    if (fetch_one(err.p, s)) fail(GLV_E_INVALID_ARG, "balanced coloring: a color outside [0, ncolors)");
    k_over_flags<<<grid_for(ncolors, kBlock, 64), kBlock, 0, s>>>(freq.p, ncolors, avg, over.p);
    GLV_LAUNCH_CHECK();
    GLV_HIP(hipMemcpyAsync(base.p, colors, nv * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    DBuf<i64> keep(nv, s);
    const int seg_bits = bits_for_int(ncolors);
    int rounds = 0;
    for (;;) {
        launch_fill_i32((int32_t *)n_elig.p, 1, 0, s);
        k_bal_eligible<<<1, kWave, 0, s>>>(freq.p, over.p, ncolors, avg, elig.p, n_elig.p);
        GLV_LAUNCH_CHECK();
        if (fetch_one(n_elig.p, s) == 0) break;  // no colour can take a vertex
        const unsigned wblocks = grid_for(nv * kWave, kCB, 8192);
        k_bal_wish<kMaxColors / 64><<<wblocks, kCB, 0, s>>>(g.off, g.tails, nv, colors, base.p, over.p, freq.p, avg,
                                                            elig.p, type, wish.p);
        GLV_LAUNCH_CHECK();
        k_bal_block<<<wblocks, kCB, 0, s>>>(g.off, g.tails, nv, wish.p, keep.p);
        GLV_LAUNCH_CHECK();
        DBuf<i64> vs;
        const i64 n = list_flagged(keep.p, nv, vs, s);
        if (n == 0) break;
        DBuf<u64> pkey(n, s);
        DBuf<uint32_t> skey(n, s);
        DBuf<i64> idx(n, s), start(ncolors, s), flag(n, s);
        // capacity of the wished colours
        k_bal_keys<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(vs.p, n, wish.p, pkey.p, skey.p, idx.p);
        GLV_LAUNCH_CHECK();
        order_by_segment(pkey.p, skey.p, idx.p, n, seg_bits, wish.p, s);
        k_seg_start<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(skey.p, n, start.p);
        GLV_LAUNCH_CHECK();
        k_seg_admit<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(skey.p, idx.p, n, start.p, freq.p, avg, 0, flag.p);
        GLV_LAUNCH_CHECK();
        DBuf<i64> adm;
        i64 m;
        {
            // admitted vertices, as flags over positions -> list of vertices
            DBuf<i64> pos_list;
            m = list_flagged(flag.p, n, pos_list, s);
            if (m == 0) break;
            adm.alloc(m, s);
            launch_gather_i64(idx.p, pos_list.p, m, adm.p, s);
        }
        // departures of the classes left
        DBuf<u64> pkey2(m, s);
        DBuf<uint32_t> skey2(m, s);
        DBuf<i64> idx2(m, s), flag2(m, s);
        k_bal_keys<<<grid_for(m, kBlock, 8192), kBlock, 0, s>>>(adm.p, m, colors, pkey2.p, skey2.p, idx2.p);
        GLV_LAUNCH_CHECK();
        order_by_segment(pkey2.p, skey2.p, idx2.p, m, seg_bits, colors, s);
        k_seg_start<<<grid_for(m, kBlock, 8192), kBlock, 0, s>>>(skey2.p, m, start.p);
        GLV_LAUNCH_CHECK();
        k_seg_admit<<<grid_for(m, kBlock, 8192), kBlock, 0, s>>>(skey2.p, idx2.p, m, start.p, freq.p, avg, 1,
                                                                 flag2.p);
        GLV_LAUNCH_CHECK();
        launch_fill_i64((i64 *)moved.p, 1, 0, s);
        k_bal_apply<<<grid_for(m, kCB, 4096), kCB, 0, s>>>(idx2.p, flag2.p, m, wish.p, colors,
                                                            (unsigned long long *)freq.p, moved.p);
        GLV_LAUNCH_CHECK();
        if (fetch_one(moved.p, s) == 0) break;
        rounds++;
    }
    return rounds;
}

}  // namespace glv
