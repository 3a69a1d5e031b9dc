// active_set.hip -- learned active-set kernels: MLP inference over the
// vertices of a phase, sweep-trace capture and labelling, list compaction.
#include <math.h>

#include <algorithm>

#include "active_set.hpp"
#include "kernels.hpp"

namespace glv {

namespace {

constexpr int kLasBlock = 256;
constexpr int kLasWpb = kLasBlock / kWave;

__device__ inline u64 las_mix(u64 z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// The network's input per vertex, for k_las_infer_x: the eight BF16 features
// (X[2v]: features 0-3, X[2v+1]: 4-7; zero when v has no sweep record) and
// flags (bit 0: v has a label, bit 1: a record, bit 2: giant degree) and
// its entry count (capped at 2^31 - 1).  One thread per vertex, so the
// dependent loads of the features (the label, then its community's size)
// overlap across many wavefronts.
__global__ void k_las_xfeat(LasView S, uint2 *X, uint8_t *flags, int32_t *deg) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < S.nv; v += (i64)gridDim.x * blockDim.x) {
        const i64 d = S.off[v + 1] - S.off[v];
        deg[v] = (int32_t)(d < INT32_MAX ? d : INT32_MAX);
        const bool has = S.C[v] >= 0;
        const bool rec = has && S.feat[v].w != 0.0f;
        const bool giant = has && d > kGiantDeg;
        uint2 x0 = make_uint2(0u, 0u), x1 = make_uint2(0u, 0u);
        if (rec) {
            float f[8];
            las_features4(S, v, 0, f);
            las_features4(S, v, 1, f + 4);
            x0 = make_uint2((uint32_t)bf16_bits(f[0]) | ((uint32_t)bf16_bits(f[1]) << 16),
                            (uint32_t)bf16_bits(f[2]) | ((uint32_t)bf16_bits(f[3]) << 16));
            x1 = make_uint2((uint32_t)bf16_bits(f[4]) | ((uint32_t)bf16_bits(f[5]) << 16),
                            (uint32_t)bf16_bits(f[6]) | ((uint32_t)bf16_bits(f[7]) << 16));
        }
        X[2 * v] = x0;
        X[2 * v + 1] = x1;
        flags[v] = (uint8_t)((has ? 1 : 0) | (rec ? 2 : 0) | (giant ? 4 : 0));
    }
}

// Inference over the prepared inputs: a wavefront per tile of 16 vertices,
// the next tile's inputs loaded while the current one goes through the three
// MFMA layers.  A vertex is active when it has a label and either no record,
// a giant degree or a logit at or above logit_tau.
__global__ __launch_bounds__(kLasBlock) void k_las_infer_x(const uint2 *X, const uint8_t *flags,
                                                           const int32_t *deg, i64 nv, const float *P,
                                                           float logit_tau, uint8_t *active, i64 *counts) {
    __shared__ i64 s_cnt[2];
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    MlpW W;
    mlp_load(P, W);
    const int lane = threadIdx.x % kWave, r = lane & 15, g = lane >> 4;
    const i64 wave = (i64)blockIdx.x * kLasWpb + threadIdx.x / kWave;
    const i64 step = (i64)gridDim.x * kLasWpb * 16;
    auto load = [&](i64 base, mlp_s4 &x, uint8_t &f, int32_t &d) {
        const i64 v = base + r;
        x = mlp_s4{0, 0, 0, 0};
        f = 0;
        d = 0;
        if (v < nv) {
            if (g < 2) {
                const uint2 u = X[2 * v + g];
                x = mlp_s4{(short)(u.x & 0xFFFFu), (short)(u.x >> 16), (short)(u.y & 0xFFFFu), (short)(u.y >> 16)};
            }
            f = flags[v];
            d = deg[v];
        }
    };
    i64 na = 0, ne = 0;
    i64 base = wave * 16;
    mlp_s4 x;
    uint8_t f;
    int32_t d;
    load(base, x, f, d);
    for (; base < nv; base += step) {
        mlp_s4 xn;
        uint8_t fn;
        int32_t dn;
        load(base + step, xn, fn, dn);
        MlpAct A;
        mlp_forward_tile(W, x, A);
        const i64 v = base + r;
        if (v < nv && lane < 16) {
            // giant vertices always sweep: their HBM tables are laid out for the full class
            const bool has = f & 1, rec = f & 2, giant = f & 4;
            const bool act = has && (!rec || giant || A.z >= logit_tau);
            active[v] = act ? 1 : 0;
            if (act) {
                na++;
                ne += d;
            }
        }
        x = xn;
        f = fn;
        d = dn;
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        na += __shfl_xor(na, o);
        ne += __shfl_xor(ne, o);
    }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&s_cnt[0], (unsigned long long)na);
        atomicAdd((unsigned long long *)&s_cnt[1], (unsigned long long)ne);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd((unsigned long long *)&counts[0], (unsigned long long)s_cnt[0]);
        atomicAdd((unsigned long long *)&counts[1], (unsigned long long)s_cnt[1]);
    }
}

__global__ void k_las_tick(uint8_t *stale, i64 nv) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        const uint8_t t = stale[v];
        stale[v] = t < 255 ? (uint8_t)(t + 1) : t;
    }
}

__device__ inline void write_row(const LasView &S, i64 v, uint16_t *row) {
    float f[8];
    las_features4(S, v, 0, f);
    las_features4(S, v, 1, f + 4);
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = (uint32_t)bf16_bits(f[2 * q]) | ((uint32_t)bf16_bits(f[2 * q + 1]) << 16);
    *reinterpret_cast<uint4 *>(row) = make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ void k_las_capture(LasView S, u64 seed, u64 cut, uint16_t *feats, i64 *verts, i64 *count,
                              i64 cap) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v - threadIdx.x < S.nv;
         v += (i64)gridDim.x * blockDim.x) {
        bool take = false;
        if (v < S.nv) take = las_mix(seed + (u64)v) < cut && S.C[v] >= 0 && S.feat[v].w != 0.0f;
        const u64 m = __ballot(take);
        if (!m) continue;
        const int lane = threadIdx.x % kWave;
        const int leader = __ffsll((long long)m) - 1;
        i64 base = 0;
        if (lane == leader) base = (i64)atomicAdd((unsigned long long *)count, (unsigned long long)__popcll(m));
        base = __shfl(base, leader);
        if (take) {
            const i64 pos = base + __popcll(m & ((1ULL << lane) - 1ULL));
            if (pos < cap) {
                write_row(S, v, feats + pos * kMlpIn);
                verts[pos] = v;
            }
        }
    }
}

__global__ void k_las_label(const i64 *verts, i64 lo, i64 hi, const uint8_t *stale, uint8_t *labels) {
    for (i64 j = lo + blockIdx.x * (i64)blockDim.x + threadIdx.x; j < hi; j += (i64)gridDim.x * blockDim.x)
        labels[j] = stale[verts[j]] == 0 ? 1 : 0;
}

__global__ void k_las_features(LasView S, const i64 *verts, i64 n, uint16_t *feats) {
    for (i64 j = blockIdx.x * (i64)blockDim.x + threadIdx.x; j < n; j += (i64)gridDim.x * blockDim.x)
        write_row(S, verts[j], feats + j * kMlpIn);
}

__global__ void k_flag_active(const i64 *list, i64 n, const uint8_t *active, i64 *flags) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        flags[i] = active[list[i]] ? 1 : 0;
}

__global__ void k_compact_list(const i64 *list, i64 n, const i64 *flags, const i64 *pos, i64 *out) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        if (flags[i]) out[pos[i]] = list[i];
}

__global__ void k_compact_starts(const i64 *start, i64 nkeys, const i64 *flags, const i64 *pos, i64 n,
                                 i64 *out_start) {
    const i64 total = pos[n - 1] + flags[n - 1];
    for (i64 k = blockIdx.x * (i64)blockDim.x + threadIdx.x; k <= nkeys; k += (i64)gridDim.x * blockDim.x) {
        const i64 at = start[k];
        out_start[k] = at >= n ? total : pos[at];
    }
}

}  // namespace

#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

void launch_las_infer(const LasView &S, const float *P, float logit_tau, uint8_t *active, i64 *counts,
                      hipStream_t s) {
    if (S.nv <= 0) return;
    // inputs first (one thread per vertex), then the network over them
    DBuf<uint2> X(2 * S.nv, s);
    DBuf<uint8_t> flags(S.nv, s);
    DBuf<int32_t> deg(S.nv, s);
    k_las_xfeat<<<grid_for(S.nv, kBlock, 8192), kBlock, 0, s>>>(S, X.p, flags.p, deg.p);
    GLV_LAUNCH_CHECK();
    const i64 tiles = (S.nv + 15) / 16;
    const unsigned blocks = (unsigned)std::min<i64>((tiles + kLasWpb - 1) / kLasWpb, 4096);
    k_las_infer_x<<<blocks, kLasBlock, 0, s>>>(X.p, flags.p, deg.p, S.nv, P, logit_tau, active, counts);
    GLV_LAUNCH_CHECK();
}

void launch_las_tick(uint8_t *stale, i64 nv, hipStream_t s) {
    if (nv <= 0) return;
    k_las_tick<<<grid_for(nv, kBlock, 8192), kBlock, 0, s>>>(stale, nv);
    GLV_LAUNCH_CHECK();
}

void launch_las_capture(const LasView &S, u64 seed, double rate, uint16_t *feats, i64 *verts, i64 *count,
                        i64 cap, hipStream_t s) {
    if (S.nv <= 0 || rate <= 0) return;
    const u64 cut = rate >= 1.0 ? ~0ULL : (u64)(rate * 18446744073709551616.0);
    k_las_capture<<<grid_for(S.nv, kBlock, 8192), kBlock, 0, s>>>(S, seed, cut, feats, verts, count, cap);
    GLV_LAUNCH_CHECK();
}

void launch_las_label(const i64 *verts, i64 lo, i64 hi, const uint8_t *stale, uint8_t *labels,
                      hipStream_t s) {
    if (hi <= lo) return;
    k_las_label<<<grid_for(hi - lo, kBlock, 8192), kBlock, 0, s>>>(verts, lo, hi, stale, labels);
    GLV_LAUNCH_CHECK();
}

void launch_las_features(const LasView &S, const i64 *verts, i64 n, uint16_t *feats, hipStream_t s) {
    if (n <= 0) return;
    k_las_features<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(S, verts, n, feats);
    GLV_LAUNCH_CHECK();
}

void launch_compact_active(const i64 *list, i64 n, const uint8_t *active, const i64 *start, i64 nkeys,
                           i64 *flags, i64 *pos, i64 *out, i64 *out_start, hipStream_t s) {
    if (n <= 0) return;
    k_flag_active<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(list, n, active, flags);
    GLV_LAUNCH_CHECK();
    compact_flagged(list, n, flags, pos, start, nkeys, out, out_start, s);
}

void compact_flagged(const i64 *list, i64 n, i64 *flags, i64 *pos, const i64 *start, i64 nkeys, i64 *out,
                     i64 *out_start, hipStream_t s) {
    if (n <= 0) return;
    scan_exclusive_i64(flags, pos, n, s);
    k_compact_list<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(list, n, flags, pos, out);
    GLV_LAUNCH_CHECK();
    k_compact_starts<<<grid_for(nkeys + 1, kBlock, 8192), kBlock, 0, s>>>(start, nkeys, flags, pos, n,
                                                                          out_start);
    GLV_LAUNCH_CHECK();
}

}  // namespace glv
