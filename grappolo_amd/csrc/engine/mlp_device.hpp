// mlp_device.hpp -- device-side pieces of the active-set MLP shared by the
// training, inference and trace kernels: BF16 conversion, the matrix-core
// forward pass of one tile of 16 samples, and the vertex features.
//
// Matrix-core layout (v_mfma_f32_16x16x16_bf16, one 64-lane wavefront,
// g = lane / 16, r = lane % 16):
//   A operand (16 x 16, M x K): lane supplies A[r][4g + i], i = 0..3
//   B operand (16 x 16, K x N): lane supplies B[4g + i][r]
//   D result  (16 x 16):        lane holds    D[4g + i][r]
// Samples run along N, so the D tile of one layer (hidden units along M) is
// already the B operand of the next: no data movement between layers.
#pragma once

#include <math.h>

#include "common.hpp"
#include "mlp.hpp"

namespace glv {

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (NaN kept quiet, infinities exact)
__device__ inline uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf16_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ inline float bf16_round(float f) { return bf16_float(bf16_bits(f)); }

__device__ inline mlp_f4 mfma16(mlp_s4 a, mlp_s4 b, mlp_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Weights of the forward pass, as each lane holds them.
struct MlpW {
    mlp_s4 w1[4];      // layer 1, hidden tile t: W1^T[16t + r][4g + i] (features >= 8 are 0)
    mlp_s4 w2t[4][4];  // layer 2, [u][t]: W2^T[16u + r][16t + 4g + i]
    mlp_f4 b1[4];      // b1[16t + 4g + i]
    mlp_f4 b2[4];
    mlp_f4 w3[4];      // BF16-rounded w3[16u + 4g + i]
    float b3;
};

__device__ inline void mlp_load(const float *P, MlpW &W) {
    const int lane = threadIdx.x % kWave, g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int f = 4 * g + i;
            W.w1[t][i] = f < kMlpIn ? (short)bf16_bits(P[kMlpW1 + f * kMlpHidden + 16 * t + r]) : (short)0;
            W.b1[t][i] = P[kMlpB1 + 16 * t + 4 * g + i];
            W.b2[t][i] = P[kMlpB2 + 16 * t + 4 * g + i];
            W.w3[t][i] = bf16_round(P[kMlpW3 + 16 * t + 4 * g + i]);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                W.w2t[u][t][i] = (short)bf16_bits(P[kMlpW2 + (16 * t + 4 * g + i) * kMlpHidden + 16 * u + r]);
    W.b3 = P[kMlpB3];
}

// Activations of one tile of 16 samples (sample r of the tile in lane r + 16g).
struct MlpAct {
    mlp_f4 z1[4];  // pre-activation of layer 1, [t][i] = unit 16t + 4g + i
    mlp_s4 h1[4];  // relu(z1) in BF16: the B operand of layer 2
    mlp_f4 z2[4];
    float z;       // logit of sample r (identical in the 4 lanes r, r+16, r+32, r+48)
};

// x: the B operand of layer 1 (lane supplies features 4g..4g+3 of sample r; zeros for g >= 2)
__device__ inline void mlp_forward_tile(const MlpW &W, mlp_s4 x, MlpAct &A) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
        A.z1[t] = mfma16(W.w1[t], x, W.b1[t]);
#pragma unroll
        for (int i = 0; i < 4; i++) A.h1[t][i] = (short)bf16_bits(fmaxf(A.z1[t][i], 0.0f));
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        mlp_f4 acc = W.b2[u];
#pragma unroll
        for (int t = 0; t < 4; t++) acc = mfma16(W.w2t[u][t], A.h1[t], acc);
        A.z2[u] = acc;
    }
    float p = 0.0f;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int i = 0; i < 4; i++) p += W.w3[u][i] * fmaxf(A.z2[u][i], 0.0f);
    p += __shfl_xor(p, 16);
    p += __shfl_xor(p, 32);
    A.z = p + W.b3;
}

// ------------------------------------------------------------- vertex features
// The sweep records, per swept vertex v of degree k_v in community s:
//   feat.x = e_{v,s} / k_v           (weight to its own community, self-loop included)
//   feat.y = slog(best gain / 2k_v)  over communities y != s adjacent to v (-1: none)
//   feat.z = slog(second best ...)
//   feat.w = 1 once recorded
// slog(g) = sign(g) log2(1 + 1024|g|) / 11 spreads the small gains that
// decide a move.  The MLP input of v is then
//   0: log2(1 + entries of v) / 10     4: min(sweeps since v last moved, 16) / 16
//   1: feat.x                          5: color of v / number of colors
//   2: feat.y                          6: log2(1 + size of v's community) / 16
//   3: feat.z                          7: min(phase index, 15) / 16
__host__ __device__ inline float las_slog(float g) {
    const float m = log2f(1.0f + 1024.0f * fabsf(g)) / 11.0f;
    return g < 0 ? -m : m;
}

struct LasView {
    const i64 *off = nullptr;
    const float4 *feat = nullptr;
    const uint8_t *stale = nullptr;
    const int32_t *colors = nullptr;
    int ncol = 1;
    const i64 *C = nullptr;
    const i64 *n = nullptr;
    int phase = 0;
    i64 nv = 0;
};

// features 4h..4h+3 (h = 0 or 1) of vertex v
__device__ inline void las_features4(const LasView &S, i64 v, int h, float out[4]) {
    if (h == 0) {
        const float4 f = S.feat[v];
        out[0] = log2f(1.0f + (float)(S.off[v + 1] - S.off[v])) / 10.0f;
        out[1] = f.x;
        out[2] = f.y;
        out[3] = f.z;
    } else {
        const int st = S.stale[v];
        out[0] = (float)(st < 16 ? st : 16) / 16.0f;
        out[1] = (S.colors && S.ncol > 1) ? (float)S.colors[v] / (float)S.ncol : 0.0f;
        const i64 c = S.C[v];
        out[2] = c >= 0 ? log2f(1.0f + (float)S.n[c]) / 16.0f : 0.0f;
        out[3] = (float)(S.phase < 15 ? S.phase : 15) / 16.0f;
    }
}

// Top-two of a set of float gains held across lanes.
__device__ inline void top2_push(float &b1, float &b2, float g) {
    if (g > b1) {
        b2 = b1;
        b1 = g;
    } else if (g > b2) {
        b2 = g;
    }
}
__device__ inline void top2_merge(float &b1, float &b2, float o1, float o2) {
    const float lo = fminf(b1, o1), hi2 = fmaxf(b2, o2);
    b1 = fmaxf(b1, o1);
    b2 = fmaxf(lo, hi2);
}
template <int W>
__device__ inline void top2_group(float &b1, float &b2) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        const float o1 = __shfl_xor(b1, o, W), o2 = __shfl_xor(b2, o, W);
        top2_merge(b1, b2, o1, o2);
    }
}

__device__ inline float4 las_record(double e_s, double kv, float b1, float b2) {
    const float own = kv > 0 ? (float)(e_s / kv) : 0.0f;
    const float f1 = b1 == -INFINITY ? -1.0f : las_slog(b1);
    const float f2 = b2 == -INFINITY ? -1.0f : las_slog(b2);
    return make_float4(own, f1, f2, 1.0f);
}

}  // namespace glv
