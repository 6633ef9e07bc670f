// kernels_learned.hip -- learned sparse-view FBP (no reference counterpart):
// a learnable symmetric filter in place of Ram-Lak, then a 4-layer residual
// CNN on the reconstruction:
//
//   r0 = BP(filter_h(p))                      (fp32, n x n)
//   a1 = relu(conv3x3(r0;  1->32) + b1)       (bf16, NHWC)
//   a2 = relu(conv3x3(a1; 32->32) + b2)
//   a3 = relu(conv3x3(a2; 32->32) + b3)
//   out = r0 + conv3x3(a3; 32->1) + b4        (fp32)
//
// Weights are kept in fp32 and rounded to bf16 where they are read
// (bf16 weights); activations a1..a3 are stored in bf16; every sum
// accumulates in fp32.  Convolutions use zero padding; tap t = 3*ky + kx
// reads input pixel (i + ky - 1, j + kx - 1).  Weight layouts: W1[co][t],
// W2/W3[co][t][ci], W4[t][ci].  Image sides are a multiple of 32.
//
// Training kernels: the exact gradients of the loss mean((out - target)^2)
// with respect to every weight, the filter taps (through the adjoint of the
// back-projector) -- and an Adam step.  Also the data generators: additive
// ellipse phantoms and the Poisson-noise transmission model of the sinogram.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "fbp.h"
#include "learned.h"

namespace ske {

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    bf16x2 v;
    v.x = (__bf16)lo;
    v.y = (__bf16)hi;
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float rbf(float f) { return bf2f(f2bf(f)); }  // value as bf16

constexpr int C = kLC;  // hidden channels
constexpr int T = 16;  // tile side

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------

// W2, W3 fp32 [co][t][ci] -> bf16 copies in the same layout, and W1 as
// bf16 [co][16] (taps 0..8, then zeros): the matrix-core operands.
__global__ void k_pack_weights(const float* __restrict__ prm, uint16_t* __restrict__ wb2, uint16_t* __restrict__ wb3,
                               uint16_t* __restrict__ w1b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 9 * C * C) {
        wb2[i] = f2bf(prm[kW2 + i]);
        wb3[i] = f2bf(prm[kW3 + i]);
    }
    if (i < 16 * C) {
        const int co = i / 16, t = i % 16;
        w1b[i] = t < 9 ? f2bf(prm[kW1 + co * 9 + t]) : (uint16_t)0;
    }
}

// The convolutions with 32 output channels run on the matrix cores as
// implicit GEMMs, D[co][px] = sum_k W[co][k] X[k][px] over 32-pixel groups,
// with v_mfma_f32_32x32x16_bf16 (fp32 accumulate, the bias as the initial
// accumulator).  Operand lane layout: A[i][k] at lane i + 32*(k/8), element
// k%8; B[k][j] at lane j + 32*(k/8); D[i][j] at lane j + 32*((i/4)%2),
// register 4*(i/8) + i%4.
//   conv 1 -> 32: k = tap (9 taps padded to 16), one MFMA per group;
//   conv 32 -> 32: k = (tap, 16 channels), 18 MFMAs per group, taps in
//   order, each with channels 0-15 then 16-31.
// Activation tiles sit in LDS as bf16 with an 80-byte pixel pitch (32
// channels + 8 pad: conflict-free 16-byte reads).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int kMR = 8;        // tile rows of the unfused kernels
constexpr int kMW = 32;       // tile columns
constexpr int kMP = 40;       // LDS pixel pitch in bf16 (32 channels + 8 pad)
constexpr int kMHalo = (kMR + 2) * (kMW + 2);

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// The 8 bf16 B-operand entries of conv 1 -> 32 for one pixel: taps 8*hi ..
// 8*hi+7 of the bf16 image tile `s0` (row pitch w0) around (y, x).
__device__ __forceinline__ bf16x8 conv_in_operand(const uint16_t* s0, int w0, int y, int x, int hi) {
    uint32_t e[8];
    if (hi == 0) {
#pragma unroll
        for (int t = 0; t < 8; t++) e[t] = s0[(y + t / 3) * w0 + x + t % 3];
    } else {
        e[0] = s0[(y + 2) * w0 + x + 2];
#pragma unroll
        for (int t = 1; t < 8; t++) e[t] = 0;
    }
    return as_bf16x8(make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16), e[6] | (e[7] << 16)));
}

// relu, bf16, and the group's 16 values per lane into an NHWC tile (8-byte
// stores of 4 consecutive channels); `keep` false writes zeros.
__device__ __forceinline__ void store_group(uint16_t* dst, const f32x16& acc, bool keep) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint2 v;
        v.x = pack_bf16(fmaxf(acc[4 * q], 0.0f), fmaxf(acc[4 * q + 1], 0.0f));
        v.y = pack_bf16(fmaxf(acc[4 * q + 2], 0.0f), fmaxf(acc[4 * q + 3], 0.0f));
        if (!keep) v = make_uint2(0, 0);
        *(uint2*)(dst + 8 * q) = v;
    }
}

// conv 1 -> 32 + bias + relu.  Input r0 fp32, rounded to bf16 as it is
// staged.  Workgroup: 4 waves over an 8 x 32 pixel tile, two rows each.
__global__ __launch_bounds__(256) void k_conv_in_fwd(const float* __restrict__ r0, const uint16_t* __restrict__ w1b,
                                                     const float* __restrict__ prm, uint16_t* __restrict__ a1, int n) {
    __shared__ uint16_t s0[kMHalo];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * kMR, j0 = blockIdx.x * kMW;
    const float* img = r0 + (size_t)b * n * n;
    for (int t = threadIdx.x; t < kMHalo; t += 256) {
        const int ii = i0 + t / (kMW + 2) - 1, jj = j0 + t % (kMW + 2) - 1;
        s0[t] = (ii >= 0 && ii < n && jj >= 0 && jj < n) ? f2bf(img[(size_t)ii * n + jj]) : (uint16_t)0;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const bf16x8 w = as_bf16x8(*(const uint4*)(w1b + l32 * 16 + 8 * hi));
    float bias[16];
#pragma unroll
    for (int r = 0; r < 16; r++) bias[r] = prm[kB1 + 8 * (r >> 2) + 4 * hi + (r & 3)];
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < kMR / 4; rr++) {
        const int row = wave * (kMR / 4) + rr;
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; r++) acc[r] = bias[r];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, conv_in_operand(s0, kMW + 2, row, l32, hi), acc, 0, 0, 0);
        store_group(a1 + (((size_t)b * n + i0 + row) * n + j0 + l32) * C + 4 * hi, acc, true);
    }
}

// conv 32 -> 32 + bias + relu, bf16 NHWC in/out.  Workgroup: 4 waves over
// an 8 x 32 pixel tile, each wave two rows; the (8+2) x (32+2) halo in LDS;
// each wave keeps the 18 weight operands of its 32 output channels in
// registers.
__global__ __launch_bounds__(256) void k_conv_mid_fwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ wb,
                                                      const float* __restrict__ prm, int boff,
                                                      uint16_t* __restrict__ y, int n) {
    __shared__ __align__(16) uint16_t xs[kMHalo * kMP];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * kMR, j0 = blockIdx.x * kMW;
    const uint16_t* xb = x + (size_t)b * n * n * C;
    for (int t = threadIdx.x; t < kMHalo * 4; t += 256) {
        const int pix = t >> 2, part = t & 3;
        const int ii = i0 + pix / (kMW + 2) - 1, jj = j0 + pix % (kMW + 2) - 1;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (ii >= 0 && ii < n && jj >= 0 && jj < n) v = *(const uint4*)(xb + ((size_t)ii * n + jj) * C + 8 * part);
        *(uint4*)(xs + pix * kMP + 8 * part) = v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    bf16x8 w[18];
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
            w[2 * t + h] = as_bf16x8(*(const uint4*)(wb + (l32 * 9 + t) * C + 16 * h + 8 * hi));
        }
    float bias[16];
#pragma unroll
    for (int r = 0; r < 16; r++) bias[r] = prm[boff + 8 * (r >> 2) + 4 * hi + (r & 3)];
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < kMR / 4; rr++) {
        const int row = wave * (kMR / 4) + rr;
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; r++) acc[r] = bias[r];
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const uint16_t* xp = xs + ((row + t / 3) * (kMW + 2) + l32 + t % 3) * kMP + 8 * hi;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[2 * t + h], as_bf16x8(*(const uint4*)(xp + 16 * h)),
                                                              acc, 0, 0, 0);
            }
        }
        store_group(y + (((size_t)b * n + i0 + row) * n + j0 + l32) * C + 4 * hi, acc, true);
    }
}

// conv 32 -> 1 + bias, residual add; optionally the loss and its gradient.
// The a3 halo stays bf16 in LDS with a 17-word pixel pitch, so the lanes
// (consecutive pixels) read distinct banks.  Channel pairs are summed with
// v_dot2c_f32_bf16 (weights rounded to bf16) in two partial sums, channels
// 0-15 and 16-31 (taps in order, pairs in order within a tap), then
// (s0 + s1) + b4.
constexpr int kOP = C / 2 + 1;  // LDS pixel pitch in 32-bit words
__device__ __forceinline__ float dot2_bf16(uint32_t w, uint32_t x, float acc) {
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w), __builtin_bit_cast(bf16x2, x), acc, false);
}

__global__ __launch_bounds__(256) void k_conv_out_fwd(const uint16_t* __restrict__ a3, const float* __restrict__ r0,
                                                      const float* __restrict__ prm, float* __restrict__ out,
                                                      const float* __restrict__ target, float* __restrict__ dout,
                                                      float inv_count, double* __restrict__ loss, int n) {
    __shared__ uint32_t w[9 * C / 2];
    __shared__ uint32_t xs[(T + 2) * (T + 2) * kOP];
    __shared__ float red[256];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * T, j0 = blockIdx.x * T;
    for (int t = threadIdx.x; t < 9 * C / 2; t += 256) w[t] = pack_bf16(prm[kW4 + 2 * t], prm[kW4 + 2 * t + 1]);
    const float b4 = prm[kB4];
    const uint16_t* xb = a3 + (size_t)b * n * n * C;
    for (int t = threadIdx.x; t < (T + 2) * (T + 2) * 4; t += 256) {
        const int pix = t >> 2, part = t & 3;
        const int ii = i0 + pix / (T + 2) - 1, jj = j0 + pix % (T + 2) - 1;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (ii >= 0 && ii < n && jj >= 0 && jj < n) v = *(const uint4*)(xb + ((size_t)ii * n + jj) * C + 8 * part);
        uint32_t* d = xs + pix * kOP + 4 * part;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
    __syncthreads();
    const int ty = threadIdx.x / T, tx = threadIdx.x % T;
    float s0 = 0.0f, s1 = 0.0f;  // channels 0-15 and 16-31
    for (int tap = 0; tap < 9; tap++) {
        const uint32_t* xp = xs + ((ty + tap / 3) * (T + 2) + tx + tap % 3) * kOP;
#pragma unroll
        for (int c2 = 0; c2 < C / 4; c2++) s0 = dot2_bf16(w[tap * (C / 2) + c2], xp[c2], s0);
#pragma unroll
        for (int c2 = C / 4; c2 < C / 2; c2++) s1 = dot2_bf16(w[tap * (C / 2) + c2], xp[c2], s1);
    }
    const size_t p = ((size_t)b * n + i0 + ty) * n + j0 + tx;
    const float o = r0[p] + ((s0 + s1) + b4);
    out[p] = o;
    if (target) {
        const float d = o - target[p];
        dout[p] = 2.0f * d * inv_count;
        red[threadIdx.x] = d * d;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicAdd(loss, (double)red[0]);
    }
}

// The whole CNN forward for inference, fused: one workgroup per 16 x 16
// output tile computes a1 on the 22 x 22 region around it (from a 24 x 24
// tile of r0), a2 on 20 x 20, a3 on 18 x 18, then out -- all in LDS, so HBM
// sees r0 once and out once.  Region pixels outside the image are zero,
// the zero padding of the next layer.  The arithmetic per value is the
// unfused kernels' (same MFMA sequence, same FMA order for conv 32 -> 1),
// so the output is bit-identical to theirs.  8 waves; each layer's pixels
// go in 32-pixel groups, group g to wave g mod 8.
constexpr int kFT = 16;          // output tile side
constexpr int kF0 = kFT + 8;     // r0 region side
constexpr int kF1 = kFT + 6;     // a1
constexpr int kF2 = kFT + 4;     // a2
constexpr int kF3 = kFT + 2;     // a3

// One conv 32 -> 32 layer from the LDS region `in` (side wi) to `out` (side
// wi - 2) whose top-left pixel is image pixel (oy, ox).  A wave runs its
// groups g and g + 8 together (two independent accumulators), so the loads
// of one group overlap the matrix-core work of the other; the bias (32
// floats in LDS) seeds both accumulators.
__device__ __forceinline__ void fused_acc_init(f32x16& acc, const float* bias, int hi) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float4 b = *(const float4*)(bias + 8 * q + 4 * hi);
        acc[4 * q] = b.x;
        acc[4 * q + 1] = b.y;
        acc[4 * q + 2] = b.z;
        acc[4 * q + 3] = b.w;
    }
}

__device__ __forceinline__ void load_mid_weights(bf16x8 (&w)[18], const uint16_t* __restrict__ wb, int l32, int hi) {
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int h = 0; h < 2; h++) w[2 * t + h] = as_bf16x8(*(const uint4*)(wb + (l32 * 9 + t) * C + 16 * h + 8 * hi));
}

__device__ __forceinline__ void fused_mid_layer(const uint16_t* in, uint16_t* out, int wi, int oy, int ox, int n,
                                                const bf16x8 (&w)[18], const float* bias, int wave, int l32,
                                                int hi) {
    const int wo = wi - 2, count = wo * wo;
    for (int g0 = wave; g0 * 32 < count; g0 += 16) {
        const int g1 = g0 + 8;
        const int p0 = g0 * 32 + l32, p1 = g1 * 32 + l32;
        const int q0 = min(p0, count - 1), q1 = min(p1, count - 1);
        const int y0 = q0 / wo, x0 = q0 % wo, y1 = q1 / wo, x1 = q1 % wo;
        const uint16_t* b0 = in + (y0 * wi + x0) * kMP + 8 * hi;
        f32x16 acc0;
        fused_acc_init(acc0, bias, hi);
        if (g1 * 32 < count) {  // two groups (wave-uniform branch)
            const uint16_t* b1 = in + (y1 * wi + x1) * kMP + 8 * hi;
            f32x16 acc1;
            fused_acc_init(acc1, bias, hi);
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const int o = ((t / 3) * wi + t % 3) * kMP;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const bf16x8 v0 = as_bf16x8(*(const uint4*)(b0 + o + 16 * h));
                    const bf16x8 v1 = as_bf16x8(*(const uint4*)(b1 + o + 16 * h));
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[2 * t + h], v0, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[2 * t + h], v1, acc1, 0, 0, 0);
                }
            }
            const int iy = oy + y1, ix = ox + x1;
            if (p1 < count) store_group(out + p1 * kMP + 4 * hi, acc1, iy >= 0 && iy < n && ix >= 0 && ix < n);
        } else {
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const int o = ((t / 3) * wi + t % 3) * kMP;
#pragma unroll
                for (int h = 0; h < 2; h++)
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[2 * t + h],
                                                                   as_bf16x8(*(const uint4*)(b0 + o + 16 * h)), acc0,
                                                                   0, 0, 0);
            }
        }
        const int iy = oy + y0, ix = ox + x0;
        if (p0 < count) store_group(out + p0 * kMP + 4 * hi, acc0, iy >= 0 && iy < n && ix >= 0 && ix < n);
    }
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_cnn_fused_fwd(const float* __restrict__ r0, const float* __restrict__ prm,
                                                       const uint16_t* __restrict__ w1b,
                                                       const uint16_t* __restrict__ wb2,
                                                       const uint16_t* __restrict__ wb3, float* __restrict__ out,
                                                       int n) {
    __shared__ __align__(16) uint16_t sa[kF1 * kF1 * kMP];  // a1, then a3
    __shared__ __align__(16) uint16_t sb[kF2 * kF2 * kMP];  // a2
    __shared__ uint16_t s0[kF0 * kF0];                     // r0 as bf16
    __shared__ uint32_t w4[9 * C / 2];                     // W4 as bf16 pairs
    __shared__ __align__(16) float sbias[2 * C];            // b2, b3
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * kFT, j0 = blockIdx.x * kFT;
    const float* img = r0 + (size_t)b * n * n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    // all global loads of the prologue in flight together: W2 operands, W1,
    // biases, W4 and the r0 tile
    bf16x8 w[18];
    load_mid_weights(w, wb2, l32, hi);
    const bf16x8 w1 = as_bf16x8(*(const uint4*)(w1b + l32 * 16 + 8 * hi));
    float bias1[16];
#pragma unroll
    for (int r = 0; r < 16; r++) bias1[r] = prm[kB1 + 8 * (r >> 2) + 4 * hi + (r & 3)];
    for (int t = threadIdx.x; t < kF0 * kF0; t += 512) {
        const int ii = i0 - 4 + t / kF0, jj = j0 - 4 + t % kF0;
        s0[t] = (ii >= 0 && ii < n && jj >= 0 && jj < n) ? f2bf(img[(size_t)ii * n + jj]) : (uint16_t)0;
    }
    for (int t = threadIdx.x; t < 9 * C / 2; t += 512) w4[t] = pack_bf16(prm[kW4 + 2 * t], prm[kW4 + 2 * t + 1]);
    if (threadIdx.x < 2 * C) sbias[threadIdx.x] = prm[(threadIdx.x < C ? kB2 : kB3 - C) + threadIdx.x];
    __syncthreads();
    {   // a1 on the 22 x 22 region at (i0-3, j0-3)
        const int count = kF1 * kF1;
        for (int g = wave; g * 32 < count; g += 8) {
            const int p = g * 32 + l32;
            const int pc = min(p, count - 1);
            const int y = pc / kF1, x = pc % kF1;
            f32x16 acc;
#pragma unroll
            for (int r = 0; r < 16; r++) acc[r] = bias1[r];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, conv_in_operand(s0, kF0, y, x, hi), acc, 0, 0, 0);
            const int iy = i0 - 3 + y, ix = j0 - 3 + x;
            if (p < count) store_group(sa + p * kMP + 4 * hi, acc, iy >= 0 && iy < n && ix >= 0 && ix < n);
        }
    }
    __syncthreads();
    fused_mid_layer(sa, sb, kF1, i0 - 2, j0 - 2, n, w, sbias, wave, l32, hi);
    load_mid_weights(w, wb3, l32, hi);  // in flight while the other waves finish
    __syncthreads();
    fused_mid_layer(sb, sa, kF2, i0 - 1, j0 - 1, n, w, sbias + C, wave, l32, hi);
    __syncthreads();
    {   // conv 32 -> 1 + bias + residual: two threads per pixel (channels
        // 0-15 and 16-31), combined as in k_conv_out_fwd
        const int px = threadIdx.x >> 1, part = threadIdx.x & 1;
        const int y = px / kFT, x = px % kFT;
        float acc = 0.0f;
        for (int tap = 0; tap < 9; tap++) {
            const uint16_t* xp = sa + ((y + tap / 3) * kF3 + x + tap % 3) * kMP + 16 * part;
#pragma unroll
            for (int c8 = 0; c8 < 2; c8++) {
                const uint4 v = *(const uint4*)(xp + 8 * c8);
                const uint32_t* wp = w4 + tap * (C / 2) + 8 * part + 4 * c8;
                acc = dot2_bf16(wp[0], v.x, acc);
                acc = dot2_bf16(wp[1], v.y, acc);
                acc = dot2_bf16(wp[2], v.z, acc);
                acc = dot2_bf16(wp[3], v.w, acc);
            }
        }
        const float other = __shfl_xor(acc, 1, 64);
        if (part == 0) {
            const size_t p = ((size_t)b * n + i0 + y) * n + j0 + x;
            out[p] = img[(size_t)(i0 + y) * n + j0 + x] + ((acc + other) + prm[kB4]);
        }
    }
}

// ---------------------------------------------------------------------------
// backward: activation gradients (fp32, NHWC), masked by the relu of the
// layer input they flow into
// ---------------------------------------------------------------------------

// dA3[q][ci] = sum_t W4[t][ci] * dout[q - off(t)], times (a3 > 0).
__global__ __launch_bounds__(256) void k_conv_out_bwd(const float* __restrict__ dout, const uint16_t* __restrict__ a3,
                                                      const float* __restrict__ prm, float* __restrict__ dz3, int n) {
    __shared__ float w[9 * C];
    __shared__ float ds[(T + 2) * (T + 2)];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * T, j0 = blockIdx.x * T;
    for (int t = threadIdx.x; t < 9 * C; t += 256) w[t] = rbf(prm[kW4 + t]);
    const float* db = dout + (size_t)b * n * n;
    for (int t = threadIdx.x; t < (T + 2) * (T + 2); t += 256) {
        const int ii = i0 + t / (T + 2) - 1, jj = j0 + t % (T + 2) - 1;
        ds[t] = (ii >= 0 && ii < n && jj >= 0 && jj < n) ? db[(size_t)ii * n + jj] : 0.0f;
    }
    __syncthreads();
    const int ty = threadIdx.x / T, tx = threadIdx.x % T;
    const size_t p = ((size_t)b * n + i0 + ty) * n + j0 + tx;
    float g[9];
#pragma unroll
    for (int t = 0; t < 9; t++) g[t] = ds[(ty + 2 - t / 3) * (T + 2) + tx + 2 - t % 3];  // dout[q - off(t)]
    for (int ci = 0; ci < C; ci++) {
        float s = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; t++) s = __builtin_fmaf(w[t * C + ci], g[t], s);
        dz3[p * C + ci] = bf2f(a3[p * C + ci]) > 0.0f ? s : 0.0f;
    }
}

// dX[q][ci] = sum_{t,co} W[co][t][ci] * dZ[q - off(t)][co], times (x > 0).
// 16 x 8 pixel tiles so that the fp32 dZ halo and the weights fit in LDS.
constexpr int TY = 8;
__global__ __launch_bounds__(128) void k_conv_mid_bwd(const float* __restrict__ dz, const uint16_t* __restrict__ x,
                                                      const float* __restrict__ prm, int woff,
                                                      float* __restrict__ dzp, int n) {
    extern __shared__ float sm[];
    float* w = sm;               // [t][co][ci]
    float* ds = sm + 9 * C * C;  // (TY+2)*(T+2) pixels at a pitch of C+1 floats (distinct banks per lane)
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * T;
    for (int t = threadIdx.x; t < 9 * C * C; t += 128) {
        const int co = t / (9 * C), r = t % (9 * C), tap = r / C, ci = r % C;
        w[(tap * C + co) * C + ci] = rbf(prm[woff + t]);
    }
    const float* db = dz + (size_t)b * n * n * C;
    for (int t = threadIdx.x; t < (TY + 2) * (T + 2) * C; t += 128) {
        const int pix = t / C, c = t % C;
        const int ii = i0 + pix / (T + 2) - 1, jj = j0 + pix % (T + 2) - 1;
        ds[pix * (C + 1) + c] = (ii >= 0 && ii < n && jj >= 0 && jj < n) ? db[((size_t)ii * n + jj) * C + c] : 0.0f;
    }
    __syncthreads();
    const int ty = threadIdx.x / T, tx = threadIdx.x % T;
    float acc[C];
#pragma unroll
    for (int ci = 0; ci < C; ci++) acc[ci] = 0.0f;
    for (int tap = 0; tap < 9; tap++) {
        const float* gp = ds + ((ty + 2 - tap / 3) * (T + 2) + tx + 2 - tap % 3) * (C + 1);
        for (int co = 0; co < C; co++) {
            const float gv = gp[co];
            const float* wr = w + (tap * C + co) * C;
#pragma unroll
            for (int ci = 0; ci < C; ci++) acc[ci] = __builtin_fmaf(wr[ci], gv, acc[ci]);
        }
    }
    const size_t p = ((size_t)b * n + i0 + ty) * n + j0 + tx;
#pragma unroll
    for (int ci = 0; ci < C; ci++) dzp[p * C + ci] = bf2f(x[p * C + ci]) > 0.0f ? acc[ci] : 0.0f;
}

// dr0[q] = dout[q] + sum_{t,co} W1[co][t] * dZ1[q - off(t)][co].
__global__ __launch_bounds__(128) void k_conv_in_bwd(const float* __restrict__ dz1, const float* __restrict__ dout,
                                                     const float* __restrict__ prm, float* __restrict__ dr0, int n) {
    __shared__ float w[C * 9];
    __shared__ float ds[(TY + 2) * (T + 2) * (C + 1)];  // pixel pitch C+1: distinct banks per lane
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * T;
    for (int t = threadIdx.x; t < C * 9; t += 128) w[t] = rbf(prm[kW1 + t]);
    const float* db = dz1 + (size_t)b * n * n * C;
    for (int t = threadIdx.x; t < (TY + 2) * (T + 2) * C; t += 128) {
        const int pix = t / C, c = t % C;
        const int ii = i0 + pix / (T + 2) - 1, jj = j0 + pix % (T + 2) - 1;
        ds[pix * (C + 1) + c] = (ii >= 0 && ii < n && jj >= 0 && jj < n) ? db[((size_t)ii * n + jj) * C + c] : 0.0f;
    }
    __syncthreads();
    const int ty = threadIdx.x / T, tx = threadIdx.x % T;
    float s = 0.0f;
    for (int tap = 0; tap < 9; tap++) {
        const float* gp = ds + ((ty + 2 - tap / 3) * (T + 2) + tx + 2 - tap % 3) * (C + 1);
#pragma unroll
        for (int co = 0; co < C; co++) s = __builtin_fmaf(w[co * 9 + tap], gp[co], s);
    }
    const size_t p = ((size_t)b * n + i0 + ty) * n + j0 + tx;
    dr0[p] = dout[p] + s;
}

// ---------------------------------------------------------------------------
// backward: weight gradients, tile partial sums added into grad[] with atomics
// ---------------------------------------------------------------------------

// dW[co][t][ci] += sum_p dZ[p][co] * X[p + off(t)][ci]; db[co] += sum_p dZ[p][co].
// X: bf16 NHWC (xc channels) or fp32 single channel (xc == 1, xf != null).
// dZ: fp32 NHWC with zc channels.  16 x 8 pixel tiles; a workgroup sums its
// tile over the slices b = blockIdx.z, blockIdx.z + gridDim.z, ... in
// registers and adds the partial sums into grad[] with atomics once.
constexpr int kWgPer = 9 * C * C / 256;  // weights per thread (largest layer)
__global__ __launch_bounds__(256) void k_wgrad(const float* __restrict__ dz, int zc, const uint16_t* __restrict__ xh,
                                               const float* __restrict__ xf, int xc, float* __restrict__ grad, int woff,
                                               int boff, int wlayout_tc, int n, int batch) {
    extern __shared__ float sm[];
    float* zs = sm;                        // TY*T*zc
    float* xs = sm + TY * T * zc;          // (TY+2)(T+2)*xc
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * T;
    const int nw = zc * 9 * xc;
    float acc[kWgPer];
#pragma unroll
    for (int r = 0; r < kWgPer; r++) acc[r] = 0.0f;
    float bacc = 0.0f;
    for (int b = blockIdx.z; b < batch; b += gridDim.z) {
        __syncthreads();
        const float* zb = dz + (size_t)b * n * n * zc;
        for (int t = threadIdx.x; t < TY * T * zc; t += 256) {
            const int pix = t / zc, c = t % zc;
            zs[t] = zb[((size_t)(i0 + pix / T) * n + j0 + pix % T) * zc + c];
        }
        for (int t = threadIdx.x; t < (TY + 2) * (T + 2) * xc; t += 256) {
            const int pix = t / xc, c = t % xc;
            const int ii = i0 + pix / (T + 2) - 1, jj = j0 + pix % (T + 2) - 1;
            float v = 0.0f;
            if (ii >= 0 && ii < n && jj >= 0 && jj < n) {
                const size_t q = ((size_t)b * n + ii) * n + jj;
                v = xf ? rbf(xf[q]) : bf2f(xh[q * xc + c]);
            }
            xs[t] = v;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kWgPer; r++) {
            const int wi = threadIdx.x + 256 * r;
            if (wi < nw) {
                // weight index -> (co, tap, ci) in the layer's layout
                int co, tap, ci;
                if (wlayout_tc) {  // [t][ci] (zc == 1)
                    co = 0;
                    tap = wi / xc;
                    ci = wi % xc;
                } else {           // [co][t][ci]
                    co = wi / (9 * xc);
                    const int rm = wi % (9 * xc);
                    tap = rm / xc;
                    ci = rm % xc;
                }
                const int oy = tap / 3, ox = tap % 3;
                float s = acc[r];
                for (int py = 0; py < TY; py++)
                    for (int px = 0; px < T; px++)
                        s = __builtin_fmaf(zs[(py * T + px) * zc + co],
                                           xs[((py + oy) * (T + 2) + px + ox) * xc + ci], s);
                acc[r] = s;
            }
        }
        if ((int)threadIdx.x < zc)
            for (int pix = 0; pix < TY * T; pix++) bacc += zs[pix * zc + threadIdx.x];
    }
#pragma unroll
    for (int r = 0; r < kWgPer; r++) {
        const int wi = threadIdx.x + 256 * r;
        if (wi < nw) atomicAdd(grad + woff + wi, acc[r]);
    }
    if ((int)threadIdx.x < zc) atomicAdd(grad + boff + threadIdx.x, bacc);
}

// ---------------------------------------------------------------------------
// filter and back-projector gradients
// ---------------------------------------------------------------------------

// Adjoint of k_fbp_backproject: dq[a][k] += scale * sum over pixels of the
// pixel's linear weight on bin k.  Per 16 x 16 tile and angle, the tile's
// contributions land in a 48-bin window accumulated in LDS first.
constexpr int kAdjWin = 48;
__global__ __launch_bounds__(256) void k_bp_adjoint(const float* __restrict__ dr, float* __restrict__ dq,
                                                    const float2* __restrict__ cs, FbpGeom g, float scale) {
    __shared__ float win[kAdjWin];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * T, j0 = blockIdx.x * T;
    const int ty = threadIdx.x / T, tx = threadIdx.x % T;
    const int i = i0 + ty, j = j0 + tx;
    const float inv = 1.0f / g.delta;
    const float x = ((float)j + 0.5f) * g.delta - 1.0f, y = 1.0f - ((float)i + 0.5f) * g.delta;
    const float v = dr[((size_t)b * g.n + i) * g.n + j] * scale;
    // tile centre
    const float xc = ((float)j0 + 0.5f * T) * g.delta - 1.0f, yc = 1.0f - ((float)i0 + 0.5f * T) * g.delta;
    float* dqb = dq + (size_t)b * g.nang * g.ndet;
    for (int a = 0; a < g.nang; a++) {
        const float2 t2 = cs[a];
        const int base = (int)floorf((xc * t2.x + yc * t2.y) * inv + g.det_c) - kAdjWin / 2;
        if (threadIdx.x < kAdjWin) win[threadIdx.x] = 0.0f;
        __syncthreads();
        const float kf = (x * t2.x + y * t2.y) * inv + g.det_c;
        const float fk = floorf(kf);
        const int k0 = (int)fk;
        const float w = kf - fk;
        // |k0 - base - 24| <= ceil(T*sqrt(2)/2) + 1 < 24: inside the window
        atomicAdd(&win[k0 - base], v - w * v);
        atomicAdd(&win[k0 + 1 - base], w * v);
        __syncthreads();
        if (threadIdx.x < kAdjWin) {
            const int k = base + (int)threadIdx.x;
            const float s = win[threadIdx.x];
            if (k >= 0 && k < g.ndet && s != 0.0f) atomicAdd(&dqb[(size_t)a * g.ndet + k], s);
        }
        __syncthreads();
    }
}

// dh[m + ndet - 1] += sum_k dq[row][k] * p[row][k - m] for every row.
__global__ __launch_bounds__(256) void k_filter_grad(const float* __restrict__ dq, const float* __restrict__ p,
                                                     float* __restrict__ dh, int ndet) {
    extern __shared__ float sm[];
    float* gs = sm;
    float* ps = sm + ndet;
    const size_t row = blockIdx.x;
    for (int t = threadIdx.x; t < ndet; t += 256) {
        gs[t] = dq[row * ndet + t];
        ps[t] = p[row * ndet + t];
    }
    __syncthreads();
    for (int mi = threadIdx.x; mi < 2 * ndet - 1; mi += 256) {
        const int m = mi - (ndet - 1);
        const int klo = m > 0 ? m : 0, khi = m > 0 ? ndet : ndet + m;
        float s = 0.0f;
        for (int k = klo; k < khi; k++) s = __builtin_fmaf(gs[k], ps[k - m], s);
        atomicAdd(dh + mi, s);
    }
}

// Symmetric taps: full[m + ndet - 1] = sym[|m|].
__global__ void k_taps_expand(const float* __restrict__ sym, float* __restrict__ full, int ndet) {
    const int mi = blockIdx.x * blockDim.x + threadIdx.x;
    if (mi < 2 * ndet - 1) full[mi] = sym[abs(mi - (ndet - 1))];
}

// d sym[k] = d full[ndet-1+k] + d full[ndet-1-k] (k > 0), d full[ndet-1] (k = 0).
__global__ void k_taps_fold(const float* __restrict__ dfull, float* __restrict__ dsym, int ndet) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ndet) dsym[k] += dfull[ndet - 1 + k] + (k ? dfull[ndet - 1 - k] : 0.0f);
}

// ---------------------------------------------------------------------------
// optimiser and data generation
// ---------------------------------------------------------------------------

__global__ void k_adam(float* __restrict__ prm, const float* __restrict__ grad, float* __restrict__ m1,
                       float* __restrict__ m2, int count, float lr, float lr_taps, int taps_off, float b1, float b2,
                       float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.0f - b1) * g;
    const float v = b2 * m2[i] + (1.0f - b2) * g * g;
    m1[i] = a;
    m2[i] = v;
    const float step = (i >= taps_off) ? lr_taps : lr;
    prm[i] -= step * (a / c1) / (sqrtf(v / c2) + eps);
}

// Additive ellipses (x0, y0, a, b, phi, rho) sampled at pixel centres.
__global__ __launch_bounds__(256) void k_phantom(const float* __restrict__ ell, int nell, float* __restrict__ img,
                                                 int n) {
    const int j = blockIdx.x * 16 + (threadIdx.x & 15);
    const int i = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int b = blockIdx.z;
    if (i >= n || j >= n) return;
    const float x = ((float)j + 0.5f) / n * 2.0f - 1.0f, y = -(((float)i + 0.5f) / n * 2.0f - 1.0f);
    float v = 0.0f;
    const float* e = ell + (size_t)b * nell * 6;
    for (int k = 0; k < nell; k++) {
        const float cp = cosf(e[6 * k + 4]), sp = sinf(e[6 * k + 4]);
        const float dx = x - e[6 * k], dy = y - e[6 * k + 1];
        const float u = (dx * cp + dy * sp) / e[6 * k + 2], w = (-dx * sp + dy * cp) / e[6 * k + 3];
        if (u * u + w * w <= 1.0f) v += e[6 * k + 5];
    }
    img[((size_t)b * n + i) * n + j] = v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Transmission noise: I ~ Poisson(I0 exp(-mu p)) (normal approximation with
// mean and variance lambda, rounded, for lambda >= 20; exact inversion below),
// then p = -ln(max(I,1)/I0)/mu.
__global__ void k_noise(float* __restrict__ sino, size_t count, float i0, float mu, uint64_t seed) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const float lam = i0 * expf(-mu * sino[idx]);
    const uint64_t r1 = mix64(seed ^ (idx * 0x9E3779B97F4A7C15ull)), r2 = mix64(r1 + 0x632BE59BD9B4E019ull);
    const double u1 = ((r1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((r2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    float cnt;
    if (lam >= 20.0f) {
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        cnt = rintf((float)(lam + sqrt((double)lam) * z));
    } else {  // inversion of the Poisson cdf
        double pk = exp(-(double)lam), cdf = pk;
        int k = 0;
        while (u1 > cdf && k < 200) {
            k++;
            pk *= lam / k;
            cdf += pk;
        }
        cnt = (float)k;
    }
    sino[idx] = -logf(fmaxf(cnt, 1.0f) / i0) / mu;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

#define SKE_GRID(n, b) dim3((n) / T, (n) / T, (b))

hipError_t launch_pack_weights(const float* prm, uint16_t* wb2, uint16_t* wb3, uint16_t* w1b, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_weights, dim3((9 * C * C + 255) / 256), dim3(256), 0, s, prm, wb2, wb3, w1b);
    return hipGetLastError();
}

// Inference: the fused forward (n a multiple of 16).
hipError_t launch_cnn_infer(const LearnedBufs& B, const float* prm, int n, int batch, hipStream_t s) {
    hipLaunchKernelGGL(k_cnn_fused_fwd, dim3(n / kFT, n / kFT, batch), dim3(512), 0, s, B.r0, prm, B.w1b, B.wb2, B.wb3,
                       B.out, n);
    return hipGetLastError();
}

// n must be a multiple of 32 (the matrix-core tiles are 32 pixels wide).
hipError_t launch_cnn_forward(const LearnedBufs& B, const float* prm, int n, int batch, const float* target,
                              float* dout, double* loss, hipStream_t s) {
    const dim3 gm(n / kMW, n / kMR, batch);
    hipLaunchKernelGGL(k_conv_in_fwd, gm, dim3(256), 0, s, B.r0, B.w1b, prm, B.a1, n);
    hipLaunchKernelGGL(k_conv_mid_fwd, gm, dim3(256), 0, s, B.a1, B.wb2, prm, kB2, B.a2, n);
    hipLaunchKernelGGL(k_conv_mid_fwd, gm, dim3(256), 0, s, B.a2, B.wb3, prm, kB3, B.a3, n);
    const float inv = 1.0f / ((float)n * n * batch);
    hipLaunchKernelGGL(k_conv_out_fwd, SKE_GRID(n, batch), dim3(256), 0, s, B.a3, B.r0, prm, B.out, target, dout, inv,
                       loss, n);
    return hipGetLastError();
}

hipError_t launch_cnn_backward(const LearnedBufs& B, const float* prm, const float* dout, float* grad, int n,
                               int batch, hipStream_t s) {
    // activation gradients
    hipLaunchKernelGGL(k_conv_out_bwd, SKE_GRID(n, batch), dim3(256), 0, s, dout, B.a3, prm, B.dz3, n);
    const size_t lmid = sizeof(float) * (9 * C * C + (TY + 2) * (T + 2) * (C + 1));
    dim3 g8(n / T, n / TY, batch);
    hipLaunchKernelGGL(k_conv_mid_bwd, g8, dim3(128), lmid, s, B.dz3, B.a2, prm, kW3, B.dz2, n);
    hipLaunchKernelGGL(k_conv_mid_bwd, g8, dim3(128), lmid, s, B.dz2, B.a1, prm, kW2, B.dz1, n);
    hipLaunchKernelGGL(k_conv_in_bwd, g8, dim3(128), 0, s, B.dz1, dout, prm, B.dr0, n);
    // weight gradients: about 2048 workgroups, each summing over a share of
    // the slices
    const int tiles = (n / T) * (n / TY);
    const dim3 gw(n / T, n / TY, (unsigned)std::max(1, std::min(batch, 2048 / tiles)));
    const size_t lw4 = sizeof(float) * (TY * T * 1 + (TY + 2) * (T + 2) * C);
    hipLaunchKernelGGL(k_wgrad, gw, dim3(256), lw4, s, dout, 1, B.a3, (const float*)nullptr, C, grad, kW4, kB4, 1, n,
                       batch);
    const size_t lwm = sizeof(float) * (TY * T * C + (TY + 2) * (T + 2) * C);
    hipLaunchKernelGGL(k_wgrad, gw, dim3(256), lwm, s, B.dz3, C, B.a2, (const float*)nullptr, C, grad, kW3, kB3, 0, n,
                       batch);
    hipLaunchKernelGGL(k_wgrad, gw, dim3(256), lwm, s, B.dz2, C, B.a1, (const float*)nullptr, C, grad, kW2, kB2, 0, n,
                       batch);
    const size_t lw1 = sizeof(float) * (TY * T * C + (TY + 2) * (T + 2) * 1);
    hipLaunchKernelGGL(k_wgrad, gw, dim3(256), lw1, s, B.dz1, C, (const uint16_t*)nullptr, B.r0, 1, grad, kW1, kB1, 0,
                       n, batch);
    return hipGetLastError();
}

hipError_t launch_filter_backward(const float* dr0, float* dq, const float* sino, float* dfull, float* dsym,
                                  const float2* cs, const FbpGeom& g, int batch, hipStream_t s) {
    const float scale = (float)(M_PI / g.nang);
    hipLaunchKernelGGL(k_bp_adjoint, dim3(g.n / T, g.n / T, batch), dim3(256), 0, s, dr0, dq, cs, g, scale);
    const size_t lds = sizeof(float) * 2 * (size_t)g.ndet;
    hipLaunchKernelGGL(k_filter_grad, dim3((unsigned)((size_t)batch * g.nang)), dim3(256), lds, s, dq, sino, dfull,
                       g.ndet);
    hipLaunchKernelGGL(k_taps_fold, dim3((g.ndet + 255) / 256), dim3(256), 0, s, dfull, dsym, g.ndet);
    return hipGetLastError();
}

hipError_t launch_taps_expand(const float* sym, float* full, int ndet, hipStream_t s) {
    hipLaunchKernelGGL(k_taps_expand, dim3((2 * ndet - 1 + 255) / 256), dim3(256), 0, s, sym, full, ndet);
    return hipGetLastError();
}

hipError_t launch_adam(float* prm, const float* grad, float* m1, float* m2, int count, float lr, float lr_taps,
                       int taps_off, int step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((count + 255) / 256), dim3(256), 0, s, prm, grad, m1, m2, count, lr, lr_taps,
                       taps_off, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t launch_phantom(const float* ell, int nell, float* img, int n, int batch, hipStream_t s) {
    hipLaunchKernelGGL(k_phantom, dim3((n + 15) / 16, (n + 15) / 16, batch), dim3(256), 0, s, ell, nell, img, n);
    return hipGetLastError();
}

hipError_t launch_noise(float* sino, size_t count, float i0, float mu, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(k_noise, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, sino, count, i0, mu, seed);
    return hipGetLastError();
}

}  // namespace ske
