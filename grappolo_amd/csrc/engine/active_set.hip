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

// B operand of layer 1 for vertex v (lanes of group 0: features 0-3, group 1: 4-7)
__device__ inline mlp_s4 vertex_x(const LasView &S, i64 v, bool ok) {
    const int g = (threadIdx.x % kWave) >> 4;
    mlp_s4 x = {0, 0, 0, 0};
    if (ok && g < 2) {
        float f[4];
        las_features4(S, v, g, f);
#pragma unroll
        for (int i = 0; i < 4; i++) x[i] = (short)bf16_bits(f[i]);
    }
    return x;
}

__global__ __launch_bounds__(kLasBlock) void k_las_infer(LasView S, const float *P, float logit_tau,
                                                         uint8_t *active, i64 *counts) {
    __shared__ i64 s_cnt[2];
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    MlpW W;
    mlp_load(P, W);
    const int lane = threadIdx.x % kWave, r = lane & 15;
    const i64 wave = (i64)blockIdx.x * kLasWpb + threadIdx.x / kWave;
    const i64 nwaves = (i64)gridDim.x * kLasWpb;
    i64 na = 0, ne = 0;
    for (i64 base = wave * 16; base < S.nv; base += nwaves * 16) {
        const i64 v = base + r;
        const bool inb = v < S.nv;
        const bool has = inb && S.C[v] >= 0;
        const bool rec = has && S.feat[v].w != 0.0f;
        MlpAct A;
        mlp_forward_tile(W, vertex_x(S, v, rec), A);
        if (inb && lane < 16) {
            // giant vertices always sweep: their HBM tables are laid out for the full class
            const bool giant = has && S.off[v + 1] - S.off[v] > kGiantDeg;
            const bool act = has && (!rec || giant || A.z >= logit_tau);
            active[v] = act ? 1 : 0;
            if (act) {
                na++;
                ne += S.off[v + 1] - S.off[v];
            }
        }
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
    const i64 tiles = (S.nv + 15) / 16;
    const unsigned blocks = (unsigned)std::min<i64>((tiles + kLasWpb - 1) / kLasWpb, 4096);
    k_las_infer<<<blocks, kLasBlock, 0, s>>>(S, P, logit_tau, active, counts);
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
