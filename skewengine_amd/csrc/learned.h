// learned.h -- parameter layout and launchers of the learned sparse-view FBP
// (kernels_learned.hip).  No reference counterpart.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fbp.h"

namespace ske {

// Flat fp32 parameter vector: CNN weights and biases, then the symmetric
// filter taps sym[0..ndet-1] (full filter h[m] = sym[|m|]).
constexpr int kLC = 16;                 // hidden channels
constexpr int kW1 = 0;                  // [16][9]       conv 1 -> 16
constexpr int kB1 = kW1 + kLC * 9;      // [16]
constexpr int kW2 = kB1 + kLC;          // [16][9][16]   conv 16 -> 16
constexpr int kB2 = kW2 + kLC * 9 * kLC;
constexpr int kW3 = kB2 + kLC;          // [16][9][16]   conv 16 -> 16
constexpr int kB3 = kW3 + kLC * 9 * kLC;
constexpr int kW4 = kB3 + kLC;          // [9][16]       conv 16 -> 1
constexpr int kB4 = kW4 + 9 * kLC;      // [1]
constexpr int kTaps = kB4 + 1;          // [ndet]
static_assert(kTaps == 4945, "parameter layout");

// Per-pixel tensors of a batch (all batch x n x n pixels).
struct LearnedBufs {
    const float* r0;     // FBP reconstruction, fp32
    uint16_t* a1;        // bf16 NHWC, 16 channels
    uint16_t* a2;
    uint16_t* a3;
    float* out;          // fp32
    float* dz1;          // fp32 NHWC gradients w.r.t. the pre-activations
    float* dz2;
    float* dz3;
    float* dr0;          // fp32 gradient w.r.t. r0
    const uint16_t* wb2; // bf16 copies of W2, W3 ([co][t][ci])
    const uint16_t* wb3;
    const uint16_t* w1b; // bf16 W1 as [co][16] (9 taps, 7 zeros)
};

hipError_t launch_pack_weights(const float* prm, uint16_t* wb2, uint16_t* wb3, uint16_t* w1b, hipStream_t s);
// Training forward: a1..a3 stored (n a multiple of 16).
hipError_t launch_cnn_forward(const LearnedBufs& B, const float* prm, int n, int batch, const float* target,
                              float* dout, double* loss, hipStream_t s);
// Inference forward, fused in LDS (no activations stored; n a multiple of 16).
hipError_t launch_cnn_infer(const LearnedBufs& B, const float* prm, int n, int batch, hipStream_t s);
hipError_t launch_cnn_backward(const LearnedBufs& B, const float* prm, const float* dout, float* grad, int n,
                               int batch, hipStream_t s);
hipError_t launch_filter_backward(const float* dr0, float* dq, const float* sino, float* dfull, float* dsym,
                                  const float2* cs, const FbpGeom& g, int batch, hipStream_t s);
hipError_t launch_taps_expand(const float* sym, float* full, int ndet, hipStream_t s);
hipError_t launch_adam(float* prm, const float* grad, float* m1, float* m2, int count, float lr, float lr_taps,
                       int taps_off, int step, hipStream_t s);
hipError_t launch_phantom(const float* ell, int nell, float* img, int n, int batch, hipStream_t s);
hipError_t launch_noise(float* sino, size_t count, float i0, float mu, uint64_t seed, hipStream_t s);

}  // namespace ske
