// device_common.h -- device helpers shared by the gfx950 kernels.
//
// Every kernel of this library is compiled with -ffp-contract=off: float
// products and sums are rounded separately, exactly as the reference's
// generic x86-64 CPU build computes them.  Fused multiply-adds appear only
// where written explicitly with __builtin_fma/__builtin_fmaf.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ske_types.h"

namespace ske {

// Word 3 of a raw buffer resource on gfx9-family GPUs (gfx950 included):
// 32-bit data format, no swizzle.  With stride 0, a load whose byte offset
// is at or past num_records returns zero instead of touching memory.
constexpr int kBufferRsrcWord3 = 0x00020000;

// Value of skewed cell (r, j): the two-tap scatter of
// src/skewEngine.cuh:612-621 restated as a gather.  Source rows are visited
// in ascending order there, so the w tap of row r-t-1 lands before the
// (1-w) tap of row r-t; both start from the zero of resetSkew (:567).
__device__ __forceinline__ float skew_cell(const float* __restrict__ src, int t, float w,
                                           int r, int j, int dim_i, int dim_o) {
    const int i1 = r - t - 1;
    const int i0 = r - t;
    float s = 0.0f;
    if (i1 >= 0 && i1 < dim_o) s = s + w * src[(size_t)i1 * dim_i + j];
    if (i0 >= 0 && i0 < dim_o) s = s + (1.0f - w) * src[(size_t)i0 * dim_i + j];
    return s;
}

// skew_cell with both taps read unconditionally (row indices clamped into
// the image) and masked afterwards, so several cells' gathers can be in
// flight at once; the same value bit for bit.
__device__ __forceinline__ float skew_cell_sel(const float* __restrict__ src, int t, float w,
                                               int r, int j, int dim_i, int dim_o) {
    const int i1 = r - t - 1;
    const int i0 = r - t;
    const float x1 = src[(size_t)min(max(i1, 0), dim_o - 1) * dim_i + j];
    const float x0 = src[(size_t)min(max(i0, 0), dim_o - 1) * dim_i + j];
    const float s = (i1 >= 0 && i1 < dim_o) ? 0.0f + w * x1 : 0.0f;
    return (i0 >= 0 && i0 < dim_o) ? s + (1.0f - w) * x0 : s;
}

__device__ __forceinline__ int fl_first(uint32_t v) { return (int)(v & 0xffffu); }
__device__ __forceinline__ int fl_last(uint32_t v) { return (int)(v >> 16); }

// Stage skewed row r of sector sd into LDS: row[jj] = S(r, f+jj) for
// jj in [0, L) and zeros in the `pad`-float guard bands on both sides.  With
// Sbuf == nullptr the row is gathered from the oriented input; otherwise it
// is read from a materialised skewInput buffer (row pitch dim_i).
__device__ __forceinline__ void stage_row(float* row, const SecDesc& sd, int r, int f, int L,
                                          const float* __restrict__ in4,
                                          const int* __restrict__ target,
                                          const float* __restrict__ weight,
                                          const float* __restrict__ Sbuf, int pad = kPad) {
    const float* src = in4 + sd.img + (size_t)sd.type * sd.N;
    for (int jj = (int)threadIdx.x - pad; jj < L + pad; jj += (int)blockDim.x) {
        float v = 0.0f;
        if (jj >= 0 && jj < L) {
            const int j = f + jj;
            if (Sbuf)
                v = Sbuf[(size_t)r * sd.dim_i + j];
            else
                v = skew_cell(src, target[sd.tab + j], weight[sd.tab + j], r, j, sd.dim_i, sd.dim_o);
        }
        row[jj] = v;
    }
}

}  // namespace ske
