// learned.h -- parameter layout and launchers of the learned sparse-view FBP
// (kernels_learned.hip).  No reference counterpart.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fbp.h"

namespace ske {

// Flat fp32 parameter vector for C hidden channels (16 or 32): CNN weights
// and biases, then the symmetric filter taps sym[0..ndet-1] (full filter
// h[m] = sym[|m|]).
//   W1[C][9] b1[C] W2[C][9][C] b2[C] W3[C][9][C] b3[C] W4[9][C] b4 taps
struct LearnedLayout {
    int W1, B1, W2, B2, W3, B3, W4, B4, Taps;
};
constexpr LearnedLayout learned_layout(int C) {
    return LearnedLayout{0,
                         C * 9,
                         C * 9 + C,
                         C * 9 + C + 9 * C * C,
                         C * 9 + 2 * C + 9 * C * C,
                         C * 9 + 2 * C + 18 * C * C,
                         C * 9 + 3 * C + 18 * C * C,
                         C * 9 + 3 * C + 18 * C * C + 9 * C,
                         C * 9 + 3 * C + 18 * C * C + 9 * C + 1};
}
static_assert(learned_layout(16).Taps == 4945, "parameter layout");
static_assert(learned_layout(32).Taps == 19105, "parameter layout");
constexpr bool learned_channels_ok(int C) { return C == 16 || C == 32; }

// Per-pixel tensors of a batch (all batch x n x n pixels).
struct LearnedBufs {
    const float* r0;     // FBP reconstruction, fp32
    uint16_t* a1;        // bf16 NHWC, C channels
    uint16_t* a2;
    uint16_t* a3;
    float* out;          // fp32
    float* dz1;          // fp32 NHWC gradients w.r.t. the pre-activations
    float* dz2;
    float* dz3;
    float* dr0;          // fp32 gradient w.r.t. r0
    const uint16_t* wb2; // bf16 copies of W2, W3 ([co][t][ci])
    const uint16_t* wb3;
    const uint16_t* wt2; // bf16 W2, W3 transposed ([ci][t][co]): the backward's operands
    const uint16_t* wt3;
    const uint16_t* w1b; // bf16 W1 as [co][16] (9 taps, 7 zeros)
    int C;               // hidden channels (16 or 32)
};

// wb: bf16 W2, W3 ([co][t][ci]), W1 as [co][16], then W2, W3 transposed
// ([ci][t][co]); learned_wb_size(C) elements.
constexpr size_t learned_wb_size(int C) { return 4 * 9 * (size_t)C * C + 16 * (size_t)C; }
hipError_t launch_pack_weights(const float* prm, int C, uint16_t* wb, hipStream_t s);
// Training forward: a1..a3 stored (n a multiple of 16).
hipError_t launch_cnn_forward(const LearnedBufs& B, const float* prm, int n, int batch, const float* target,
                              float* dout, double* loss, hipStream_t s);
// Inference forward, fused in LDS (no activations stored; n a multiple of 16).
hipError_t launch_cnn_infer(const LearnedBufs& B, const float* prm, int n, int batch, hipStream_t s);
hipError_t launch_cnn_backward(const LearnedBufs& B, const float* prm, const float* dout, float* grad, int n,
                               int batch, hipStream_t s);
// dq = the back-projection's adjoint of dr0 (written, not added; drT: batch
// n x n floats of scratch for dr0 transposed), then the filter gradient into
// dfull and its fold into the symmetric taps dsym.
hipError_t launch_filter_backward(const float* dr0, float* drT, float* dq, const float* sino, float* dfull,
                                  float* dsym, const float2* cs, const FbpGeom& g, int batch, hipStream_t s);
hipError_t launch_taps_expand(const float* sym, float* full, int ndet, hipStream_t s);
hipError_t launch_adam(float* prm, const float* grad, float* m1, float* m2, int count, float lr, float lr_taps,
                       int taps_off, int step, hipStream_t s);
hipError_t launch_phantom(const float* ell, int nell, float* img, int n, int batch, hipStream_t s);
hipError_t launch_noise(float* sino, size_t count, float i0, float mu, uint64_t seed, hipStream_t s);

}  // namespace ske
