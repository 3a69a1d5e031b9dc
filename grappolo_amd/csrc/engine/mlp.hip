// mlp.hip -- training and inference kernels of the active-set MLP on the
// gfx950 matrix cores (see mlp.hpp for the network and the loss, and
// mlp_device.hpp for the operand layout).
//
// Backward pass of one tile of 16 samples (dz = c_s (sigmoid(z) - y) / B):
//   dZ2 = relu'(z2) * w3 dz                      FP32, then BF16 for the matrix cores
//   dH1 = W2 dZ2                                 MFMA, A = W2 tiles, B = dZ2 (as produced)
//   dZ1 = relu'(z1) * dH1                        FP32, then BF16
//   gW2 = H1 dZ2^T,  gW1 = X dZ1^T               MFMA with the samples as the K dimension:
//                                                 H1, dZ2, dZ1 are transposed through LDS
//   gb1 = sum dZ1, gb2 = sum dZ2, gw3 = sum dz h2, gb3 = sum dz   FP32 (unrounded)
// Every wavefront writes its partial gradient to its own row; rows are
// summed in a fixed order, so a gradient is deterministic.
#include <math.h>

#include "mlp_device.hpp"

namespace glv {

namespace {

constexpr int kMlpBlock = 256;  // 4 wavefronts
constexpr int kWpb = kMlpBlock / kWave;

__device__ inline u64 mix64(u64 z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// B operand of layer 1 for the tile starting at sample `base` (index map idx()).
__device__ inline mlp_s4 load_x(const uint16_t *feats, i64 sample, bool ok) {
    const int lane = threadIdx.x % kWave, g = lane >> 4;
    mlp_s4 x = {0, 0, 0, 0};
    if (ok && g < 2) {
        const uint2 raw = *reinterpret_cast<const uint2 *>(feats + sample * kMlpIn + 4 * g);
        x[0] = (short)(raw.x & 0xffffu);
        x[1] = (short)(raw.x >> 16);
        x[2] = (short)(raw.y & 0xffffu);
        x[3] = (short)(raw.y >> 16);
    }
    return x;
}

// Transpose of a D-layout 16x16 BF16 tile through a wavefront's private LDS
// tile: lane returns M[r][4g .. 4g+3] of the matrix M whose D layout was `x`.
__device__ inline mlp_s4 transpose_tile(mlp_s4 x, uint16_t *buf) {
    const int lane = threadIdx.x % kWave, g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; i++) buf[(4 * g + i) * 16 + r] = (uint16_t)x[i];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint2 raw = *reinterpret_cast<const uint2 *>(buf + r * 16 + 4 * g);
    __builtin_amdgcn_wave_barrier();
    mlp_s4 y;
    y[0] = (short)(raw.x & 0xffffu);
    y[1] = (short)(raw.x >> 16);
    y[2] = (short)(raw.y & 0xffffu);
    y[3] = (short)(raw.y >> 16);
    return y;
}

__global__ __launch_bounds__(kMlpBlock) void k_mlp_forward(const float *P, const uint16_t *feats, i64 n,
                                                           float *logits) {
    MlpW W;
    mlp_load(P, W);
    const int lane = threadIdx.x % kWave, r = lane & 15;
    const i64 wave = (i64)blockIdx.x * kWpb + threadIdx.x / kWave;
    const i64 nwaves = (i64)gridDim.x * kWpb;
    for (i64 base = wave * 16; base < n; base += nwaves * 16) {
        const i64 smp = base + r;
        const bool ok = smp < n;
        MlpAct A;
        mlp_forward_tile(W, load_x(feats, smp, ok), A);
        if (ok && lane < 16) logits[smp] = A.z;
    }
}

__global__ __launch_bounds__(kMlpBlock) void k_mlp_grad(const float *P, MlpBatch B, float *gpart) {
    __shared__ uint16_t s_tr[kWpb][256];
    uint16_t *buf = s_tr[threadIdx.x / kWave];
    const int lane = threadIdx.x % kWave, g = lane >> 4, r = lane & 15;
    const i64 wave = (i64)blockIdx.x * kWpb + threadIdx.x / kWave;
    const i64 nwaves = (i64)gridDim.x * kWpb;
    MlpW W;
    mlp_load(P, W);
    // W2 as the A operand of dH1 = W2 dZ2: [t][u] -> W2[16t + r][16u + 4g + i]
    mlp_s4 w2[4][4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                w2[t][u][i] = (short)bf16_bits(P[kMlpW2 + (16 * t + r) * kMlpHidden + 16 * u + 4 * g + i]);
    mlp_f4 gW2[4][4], gW1[4], gb1[4], gb2[4], gw3[4];
    const mlp_f4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        gW1[t] = zero;
        gb1[t] = zero;
        gb2[t] = zero;
        gw3[t] = zero;
#pragma unroll
        for (int u = 0; u < 4; u++) gW2[t][u] = zero;
    }
    float gb3 = 0.f, loss = 0.f;
    const float inv_b = 1.0f / (float)B.batch;
    for (i64 base = wave * 16; base < B.batch; base += nwaves * 16) {
        const i64 j = base + r;  // position of lane's sample in the batch
        const bool ok = j < B.batch;
        i64 smp = j;
        if (B.random) smp = (i64)(mix64(B.seed + (u64)(B.step * B.batch + j)) % (u64)B.n);
        MlpAct A;
        const mlp_s4 x = load_x(B.feats, smp, ok);
        mlp_forward_tile(W, x, A);
        float dz = 0.f;
        if (ok) {
            const float y = B.labels[smp] ? 1.0f : 0.0f;
            const float c = (y > 0.5f ? B.pos_weight : 1.0f) * inv_b;
            const float z = A.z;
            const float p = 1.0f / (1.0f + expf(-z));
            dz = c * (p - y);
            // softplus(z) - y z, stable for either sign of z
            const float sp = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
            if (g == 0) {
                loss += c * (sp - y * z);
                gb3 += dz;
            }
        }
        // layer 3 and layer 2 deltas
        mlp_s4 dz2b[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float zz = A.z2[u][i];
                const float d = zz > 0.f ? W.w3[u][i] * dz : 0.f;
                gw3[u][i] += dz * fmaxf(zz, 0.f);
                gb2[u][i] += d;
                dz2b[u][i] = (short)bf16_bits(d);
            }
        }
        // dH1 = W2 dZ2, then layer-1 deltas
        mlp_s4 dz1b[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            mlp_f4 d = zero;
#pragma unroll
            for (int u = 0; u < 4; u++) d = mfma16(w2[t][u], dz2b[u], d);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float v = A.z1[t][i] > 0.f ? d[i] : 0.f;
                gb1[t][i] += v;
                dz1b[t][i] = (short)bf16_bits(v);
            }
        }
        // weight gradients: samples along K
        mlp_s4 dz2T[4], dz1T[4];
#pragma unroll
        for (int u = 0; u < 4; u++) dz2T[u] = transpose_tile(dz2b[u], buf);
#pragma unroll
        for (int u = 0; u < 4; u++) dz1T[u] = transpose_tile(dz1b[u], buf);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const mlp_s4 h1T = transpose_tile(A.h1[t], buf);
#pragma unroll
            for (int u = 0; u < 4; u++) gW2[t][u] = mfma16(h1T, dz2T[u], gW2[t][u]);
        }
        // X as the A operand: lane supplies X[sample 4g + i][feature r]
        mlp_s4 xa = {0, 0, 0, 0};
        if (r < kMlpIn) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const i64 jj = base + 4 * g + i;
                if (jj < B.batch) {
                    i64 s2 = jj;
                    if (B.random) s2 = (i64)(mix64(B.seed + (u64)(B.step * B.batch + jj)) % (u64)B.n);
                    xa[i] = (short)B.feats[s2 * kMlpIn + r];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) gW1[u] = mfma16(xa, dz1T[u], gW1[u]);
    }
    // reduce the per-sample sums over the 16 lanes of each group
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float a = gb1[t][i], b = gb2[t][i], c = gw3[t][i];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                a += __shfl_xor(a, o);
                b += __shfl_xor(b, o);
                c += __shfl_xor(c, o);
            }
            gb1[t][i] = a;
            gb2[t][i] = b;
            gw3[t][i] = c;
        }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        gb3 += __shfl_xor(gb3, o);
        loss += __shfl_xor(loss, o);
    }
    float *row = gpart + wave * kMlpStride;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                row[kMlpW2 + (16 * t + 4 * g + i) * kMlpHidden + 16 * u + r] = gW2[t][u][i];
    if (g < 2) {
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int i = 0; i < 4; i++) row[kMlpW1 + (4 * g + i) * kMlpHidden + 16 * u + r] = gW1[u][i];
    }
    if (r == 0) {
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                row[kMlpB1 + 16 * t + 4 * g + i] = gb1[t][i];
                row[kMlpB2 + 16 * t + 4 * g + i] = gb2[t][i];
                row[kMlpW3 + 16 * t + 4 * g + i] = gw3[t][i];
            }
    }
    if (lane == 0) {
        row[kMlpB3] = gb3;
        row[kMlpParams] = loss;
    }
}

__global__ void k_mlp_reduce(const float *gpart, i64 nw, float *grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float s = 0.f;
    for (i64 w = 0; w < nw; w++) s += gpart[w * kMlpStride + p];
    grad[p] = s;
}

__global__ void k_mlp_adam(float *P, float *m, float *v, const float *gpart, i64 nw, float lr, float b1,
                           float b2, float eps, float c1, float c2, float *loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float g = 0.f;
    for (i64 w = 0; w < nw; w++) g += gpart[w * kMlpStride + p];
    if (p == kMlpParams) {
        if (loss) *loss = g;
        return;
    }
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

}  // namespace

#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

void launch_mlp_forward(const float *P, const uint16_t *feats, i64 n, float *logits, hipStream_t s) {
    if (n <= 0) return;
    const i64 tiles = (n + 15) / 16;
    const unsigned blocks = (unsigned)std::min<i64>((tiles + kWpb - 1) / kWpb, 8192);
    k_mlp_forward<<<blocks, kMlpBlock, 0, s>>>(P, feats, n, logits);
    GLV_LAUNCH_CHECK();
}

i64 mlp_grad_waves(i64 batch) {
    const i64 tiles = (batch + 15) / 16;
    i64 w = std::min<i64>(tiles, kMlpMaxWaves);
    return ((w + kWpb - 1) / kWpb) * kWpb;
}

void launch_mlp_grad_partials(const float *P, const MlpBatch &B, float *gpart, hipStream_t s) {
    if (B.batch <= 0) return;
    if (B.random && B.n <= 0) fail(GLV_E_INVALID_ARG, "empty sample set");
    const i64 nw = mlp_grad_waves(B.batch);
    k_mlp_grad<<<(unsigned)(nw / kWpb), kMlpBlock, 0, s>>>(P, B, gpart);
    GLV_LAUNCH_CHECK();
}

void launch_mlp_reduce(const float *gpart, i64 nw, float *grad, hipStream_t s) {
    k_mlp_reduce<<<(kMlpParams + 1 + 255) / 256, 256, 0, s>>>(gpart, nw, grad);
    GLV_LAUNCH_CHECK();
}

void launch_mlp_adam(float *P, float *m, float *v, const float *gpart, i64 nw, float lr, float beta1,
                     float beta2, float eps, float c1, float c2, float *loss, hipStream_t s) {
    k_mlp_adam<<<(kMlpParams + 1 + 255) / 256, 256, 0, s>>>(P, m, v, gpart, nw, lr, beta1, beta2, eps, c1,
                                                            c2, loss);
    GLV_LAUNCH_CHECK();
}

}  // namespace glv
