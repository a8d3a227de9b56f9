// cnn.h - parameter layout and launch interface of the CU-split CNN
// (cnn.hip), internal to librmh.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rmh {
namespace cnn {

// Flat fp32 parameter vector, in this order:
//   W1[32][3][3][1], b1[32], W2[32][3][3][32], b2[32], W3[32][3][3][32], b3[32],
//   W4[64][33], b4[64], W5[6][64], b5[6]
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + 32 * 9;
constexpr int kOffW2 = kOffB1 + 32;
constexpr int kOffB2 = kOffW2 + 32 * 9 * 32;
constexpr int kOffW3 = kOffB2 + 32;
constexpr int kOffB3 = kOffW3 + 32 * 9 * 32;
constexpr int kOffW4 = kOffB3 + 32;
constexpr int kOffB4 = kOffW4 + 64 * 33;
constexpr int kOffW5 = kOffB4 + 64;
constexpr int kOffB5 = kOffW5 + 6 * 64;
constexpr int kParams = kOffB5 + 6;
constexpr int kMlpParams = kParams - kOffW4;

// Device buffers of one network instance, sized for `max_batch` blocks.
struct Buffers {
    float* params;   // fp32 master weights
    float* grads;    // fp32 gradients
    float* adam_m;
    float* adam_v;
    uint16_t* w1;    // bf16 [32][16] (k padded to 16)
    uint16_t* w2;    // bf16 [32][288]
    uint16_t* w3;
    uint16_t* w2t;   // flipped, channel-swapped copies for the input-gradient pass
    uint16_t* w3t;
    uint16_t* x;     // bf16 [n][32][32]
    uint16_t* a1;    // bf16 [n][32][32][32]
    uint16_t* a2;
    uint16_t* a3;
    uint16_t* dz;    // output gradients (dz3, then dz1)
    uint16_t* dz2;
    float* pooled;   // [n][32]
    float* dpool;    // [n][32]
    float* probs;    // [n][6]
    float* part;     // per-image partial gradients, [n][9216] at most
};

// Forward pass: probs[n][6] (softmax).
hipError_t forward(const Buffers& b, const uint8_t* blocks, const float* qp, int n, float* probs, hipStream_t s);
// Gradients of the mean cross-entropy into grads[kParams] (b.grads or any
// device buffer of kParams floats) and per-sample losses into loss[n]. The
// target is labels[n] (one-hot) or, when labels is null, targets[n][6].
hipError_t gradients(const Buffers& b, const uint8_t* blocks, const float* qp, const uint8_t* labels,
                     const float* targets, int n, float* grads, float* loss, hipStream_t s);
// bf16 weight copies from the fp32 master weights.
hipError_t pack_weights(const Buffers& b, hipStream_t s);
// One Adam step (t >= 1) from b.grads, then pack_weights.
hipError_t adam(const Buffers& b, int t, float lr, hipStream_t s);

}  // namespace cnn
}  // namespace rmh
