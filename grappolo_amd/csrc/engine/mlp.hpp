// mlp.hpp -- the learned active-set predictor: an 8 -> 64 -> 64 -> 1 MLP
// (ReLU, logistic output) that estimates, for every vertex, the probability
// that the next colored sweep moves it.  Weight matrices and activations
// enter the matrix cores as BF16 (v_mfma_f32_16x16x16_bf16), accumulation is
// FP32; biases and the 64 -> 1 output layer stay FP32 (with BF16-rounded
// output weights).  There is no counterpart in the reference.
//
// Parameter vector (FP32, kMlpParams entries):
//   W1[f][h] at kMlpW1 + f*64 + h   (8 x 64)     b1[h] at kMlpB1 + h
//   W2[i][o] at kMlpW2 + i*64 + o   (64 x 64)    b2[o] at kMlpB2 + o
//   w3[h]    at kMlpW3 + h          (64)         b3    at kMlpB3
//
// Training minimises the mean weighted logistic loss
//   L = (1/B) sum_s c_s (softplus(z_s) - y_s z_s),  c_s = pos_weight if y_s else 1
// with Adam.  Features are 8 BF16 values per sample (see las_features in
// mlp_device.hpp for their definition on the engine's state).
#pragma once

#include "common.hpp"

namespace glv {

constexpr int kMlpIn = 8;
constexpr int kMlpHidden = 64;
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpHidden;     // 512
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;               // 576
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;  // 4672
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;               // 4736
constexpr int kMlpB3 = kMlpW3 + kMlpHidden;               // 4800
constexpr int kMlpParams = kMlpB3 + 1;                    // 4801
// per-wavefront gradient partial rows: params, then the loss
constexpr int kMlpStride = 4864;
constexpr int kMlpMaxWaves = 512;

// A batch of training samples.  random=0: sample j of the batch is j (the
// batch is the first `batch` samples); random=1: it is
// splitmix64(seed + step*batch + j) mod n.
struct MlpBatch {
    const uint16_t *feats = nullptr;  // [n][8] BF16
    const uint8_t *labels = nullptr;  // [n] 0/1
    i64 n = 0;
    i64 batch = 0;
    int random = 0;
    u64 seed = 0;
    i64 step = 0;
    float pos_weight = 1.0f;
};

// logits[i] of feats[i], i < n
void launch_mlp_forward(const float *P, const uint16_t *feats, i64 n, float *logits, hipStream_t s);
// number of wavefronts (partial rows) a gradient of `batch` samples uses
i64 mlp_grad_waves(i64 batch);
// per-wavefront partial gradients of the batch into gpart[w][kMlpStride]
void launch_mlp_grad_partials(const float *P, const MlpBatch &B, float *gpart, hipStream_t s);
// grad[0..kMlpParams) and grad[kMlpParams] = loss, summed over nw partial rows in order
void launch_mlp_reduce(const float *gpart, i64 nw, float *grad, hipStream_t s);
// reduce the partials, then one Adam step on P with moments m, v; loss of the
// step into *loss when non-null.  c1, c2: bias corrections 1/(1-beta^t).
void launch_mlp_adam(float *P, float *m, float *v, const float *gpart, i64 nw, float lr, float beta1,
                     float beta2, float eps, float c1, float c2, float *loss, hipStream_t s);

}  // namespace glv
