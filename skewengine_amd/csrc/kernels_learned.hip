// kernels_learned.hip -- learned sparse-view FBP (no reference counterpart):
// a learnable symmetric filter in place of Ram-Lak, then a 4-layer residual
// CNN on the reconstruction:
//
//   r0 = BP(filter_h(p))                      (fp32, n x n)
//   a1 = relu(conv3x3(r0; 1->C) + b1)         (bf16, NHWC)
//   a2 = relu(conv3x3(a1; C->C) + b2)
//   a3 = relu(conv3x3(a2; C->C) + b3)
//   out = r0 + conv3x3(a3; C->1) + b4         (fp32)
//
// with C = 16 or 32 hidden channels (every CNN kernel is a template on C).
// Weights are kept in fp32 and rounded to bf16 where they are read
// (bf16 weights); activations a1..a3 are stored in bf16; every sum
// accumulates in fp32.  Convolutions use zero padding; tap t = 3*ky + kx
// reads input pixel (i + ky - 1, j + kx - 1).  Weight layouts: W1[co][t],
// W2/W3[co][t][ci], W4[t][ci].  Image sides are a multiple of 16.
//
// Training kernels: the exact gradients of the loss mean((out - target)^2)
// with respect to every weight, the filter taps (through the adjoint of the
// back-projector) -- and an Adam step.  Also the data generators: additive
// ellipse phantoms and the Poisson-noise transmission model of the sinogram.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "fbp.h"
#include "learned.h"

namespace ske {

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    const f32x2 f = {lo, hi};  // one v_cvt_pk_bf16_f32 for the pair
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2));
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float rbf(float f) { return bf2f(f2bf(f)); }  // value as bf16

constexpr int T = 16;  // tile side

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------

// W2, W3 fp32 [co][t][ci] -> bf16 copies in the same layout (the forward's,
// each 8-channel chunk in the stored channel order 0, 1, 4, 5, 2, 3, 6, 7 of
// the activation tiles) and transposed ([ci][t][co], the backward's), and W1
// as bf16 [co][16] with k = 4*ky + kx (kx = 3 and k >= 12 zero): the
// matrix-core operands.
template <int C>
__global__ void k_pack_weights(const float* __restrict__ prm, uint16_t* __restrict__ wb) {
    constexpr LearnedLayout L = learned_layout(C);
    uint16_t* const wb2 = wb;
    uint16_t* const wb3 = wb + 9 * C * C;
    uint16_t* const w1b = wb + 2 * 9 * C * C;
    uint16_t* const wt2 = w1b + 16 * C;
    uint16_t* const wt3 = wt2 + 9 * C * C;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 9 * C * C) {
        const int co = i / (9 * C), t = (i / C) % 9, ci = i % C;
        const uint16_t v2 = f2bf(prm[L.W2 + i]), v3 = f2bf(prm[L.W3 + i]);
        const int pw = (ci >> 1) & 3;                                     // bf16 pair within the chunk
        const int st = i - ci + (ci & ~7) + 2 * ((pw == 1 || pw == 2) ? 3 - pw : pw) + (ci & 1);
        wb2[st] = v2;
        wb3[st] = v3;
        wt2[(ci * 9 + t) * C + co] = v2;
        wt3[(ci * 9 + t) * C + co] = v3;
    }
    if (i < 16 * C) {
        const int co = i / 16, k = i % 16, ky = k / 4, kx = k % 4;
        w1b[i] = (k < 12 && kx < 3) ? f2bf(prm[L.W1 + co * 9 + 3 * ky + kx]) : (uint16_t)0;
    }
}

// The whole CNN forward, fused: one workgroup per 16 x 16 output tile
// computes a1 on the 22 x 22 region around it (from a 24 x 24 tile of r0),
// a2 on 20 x 20, a3 on 18 x 18, then out -- all in LDS, so HBM sees r0 once
// and out once (and, when training, the centre 16 x 16 of a1, a2 and a3,
// which the backward kernels read).  Region pixels outside the image are
// zero, the zero padding of the next layer.  Every value is computed by the
// same instruction sequence whichever tile or group it falls in, so the
// training forward and the inference forward agree bit for bit.
//
// Matrix cores: v_mfma_f32_16x16x32_bf16 computes D[co][px] += sum_k
// W[co][k] X[k][px] for a group of 16 pixels and a block of 16 output
// channels (fp32 accumulate, the bias as the initial accumulator); C = 32
// has two channel blocks, on two waves that read the same pixel operands
// (FwdSplit).  Lane layout: A[i][k]
// at lane i + 16*(k/8), element k%8; B[k][j] at lane j + 16*(k/8), element
// k%8; D[i][j] at lane j + 16*(i/4), register i%4.
//   conv 1 -> C: v_mfma_f32_16x16x16_bf16 (4 k per lane) with k = 4*ky + kx
//   over kx = 0..3 (the kx = 3 weight is zero) for the three rows, then
//   zeros: one MFMA per group and channel block; the operand is read as
//   bf16 pairs (x, x+1) and (x+2, x+3) from two copies of the r0 tile
//   packed at even and at odd columns;
//   conv C -> C: K blocks of 32 = (32/C) taps x C channels, arranged so that
//   one operand read of an input row serves the three output rows above it
//   (Sweep): for C = 16 the tap pairs ((dy,0), (dy,1)) for dy = 0, 1, 2 and
//   ((0,2), (1,2)), ((1,2) zero, (2,2)); for C = 32 the nine taps.  A wave
//   sweeps down a band of rows, reading each input row once.
//   conv C -> 1: the K blocks of conv C -> C with W4 in every row of A
//   (every row of D is the output), then + b4.
// Activation tiles are bf16 in LDS, one plane per 16 channels: each region
// pixel (row pitch: the region side) holds 32 bytes
// per plane, two 16-byte halves of 8 channels, swapped in columns with bit
// 2 set (chunk_off).  LDS on gfx950, as measured with SQ_LDS_BANK_CONFLICT
// on probe patterns: a ds_read_b128 takes four cycles of 16 lanes on 64
// banks, the lanes of a cycle being the 4-lane blocks {0,3,5,6},
// {1,2,4,7}, {8,11,13,14}, {9,10,12,15}; ds_write_b64 four cycles of 16
// consecutive lanes on 32 banks; ds_write_b128 eight cycles of 8 (idle
// lanes' cycles included); ds_read_b32 two cycles of 32 lanes on 32 banks.
// With this layout:
//  - a layer's groups are the first 16 pixels of a row (then blocks of the
//    leftover columns), whose operand reads meet no bank twice (the lanes of
//    a read cycle lie in one row at one column shift);
//  - every lane stores 8 bytes of a chunk after one v_permlane16_swap
//    (store_block); a store cycle's 16 lanes meet each bank at most twice
//    (the minimum for 16 pixels' same 8 bytes at a 32-byte pitch);
//  - a lane's addresses depend on its column and tap only, so the row loop
//    of a layer adds immediate offsets and no address arithmetic.
// (A plain NHWC layout with 4-channel stores and groups that wrap rows
// spends 52% of its LDS cycles in bank conflicts, measured.)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int kFT = 16;          // output tile side
constexpr int kF0 = kFT + 8;     // r0 region side
constexpr int kF1 = kFT + 6;     // a1
constexpr int kF2 = kFT + 4;     // a2
constexpr int kF3 = kFT + 2;     // a3
constexpr int kFWaves = 4;
constexpr int kS0 = 12;          // bf16 pairs per row of the packed r0 tile
// The odd-column copy of the r0 tile starts 16 words (mod 32) after the
// even-column one, so the conv 1 -> C operand reads (4 bytes a lane, 32
// lanes a cycle on 32 banks) of neighbouring columns fall on different banks.
constexpr int kS0Odd = kF0 * kS0 + 16;
constexpr int kP3 = kF3;  // row pitch of the a3 region
constexpr int kPlA = kF1 * kF1 * 16;  // plane stride (elements) of the a1/a3 region
constexpr int kPlB = kF2 * kF2 * 16;  // of the a2 region

template <int C>
struct MidK {
    static_assert(C == 16 || C == 32, "16 or 32 hidden channels");
    static constexpr int kTapsPer = 32 / C;                          // taps per K block
    static constexpr int kChunks = C / 8;                            // 8-channel chunks per tap
    static constexpr int kBlocks = (9 + kTapsPer - 1) / kTapsPer;   // K blocks per layer (5 or 9)
    static constexpr int kHalves = C / 16;                           // 16-channel output blocks
};

// Waves of a forward workgroup: kFWaves row bands; for C = 32 each band's two
// 16-channel output blocks go to two waves, which halves the weights and
// accumulators a wave holds (twice as many waves per SIMD hide the operand
// reads' latency, for twice the operand reads).  Wave w takes band w / kSplit:
// a band's waves are consecutive, so they run on different SIMDs and the
// bands with a leftover group more spread over the SIMDs.
template <int C>
struct FwdSplit {
    static constexpr int kSplit = C == 32 ? 2 : 1;         // waves per band
    static constexpr int kWaves = kFWaves * kSplit;        // waves per workgroup
    static constexpr int kHW = MidK<C>::kHalves / kSplit;  // channel blocks per wave
};

// The K blocks as reads of a sweep down the rows.  Block b of output row y
// multiplies the operand of "kind" kind(b) read at input row y + dy(b); lane
// block kb of a kind-k read at row r takes the pixel (r + krow(k, kb),
// j + kcol(k, kb)) of its column j.  C = 16: kind 0 = the taps (dy, 0) and
// (dy, 1) of one row (pixel (r, j + kb/2)), kind 1 = the taps (dy, 2) of two
// rows (pixel (r + kb/2, j + 2)); blocks 0..2 = kind 0 at dy = 0, 1, 2,
// blocks 3, 4 = kind 1 at dy = 0, 1, where block 4's first half (tap (1, 2)
// again) has zero weights.  C = 32: kind dx = the tap (dy, dx) of one row,
// block 3 dy + dx.  A read of input row r thus serves output rows r, r - 1
// and r - 2, and every block's lane taps are those of an ordinary 3 x 3
// window: tap(b, kb) = 3 (dy(b) + krow) + kcol.
template <int C>
struct Sweep {
    static constexpr int kKinds = C == 16 ? 2 : 3;
    static constexpr int kind(int b) { return C == 16 ? (b < 3 ? 0 : 1) : b % 3; }
    static constexpr int dy(int b) { return C == 16 ? (b < 3 ? b : b - 3) : b / 3; }
    // the i-th block added to an output row by the sweep (by dy, then kind):
    // groups outside the sweep add theirs in the same order, so a pixel's
    // value does not depend on the group that computes it
    static constexpr int order(int i) { return C == 16 ? (i & 1 ? 3 + (i >> 1) : i >> 1) : i; }
    static __device__ __forceinline__ int krow(int k, int kb) { return (C == 16 && k == 1) ? (kb >> 1) : 0; }
    static __device__ __forceinline__ int kcol(int k, int kb) { return C == 16 ? (k == 0 ? (kb >> 1) : 2) : k; }
    // row and column of the lane's tap in block b; `zero`: the half without weights
    static __device__ __forceinline__ void tap(int b, int kb, int& ty, int& tx, bool& zero) {
        ty = dy(b) + krow(kind(b), kb);
        tx = kcol(kind(b), kb);
        zero = C == 16 && b == 4 && kb < 2;
    }
};

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// B operand of conv 1 -> C (v_mfma_f32_16x16x16_bf16, k = 4*ky + kx) for an
// a1-region pixel (y, x): lane block kb < 3 holds the bf16 pairs (x, x+1),
// (x+2, x+3) of tile row y + kb, and block 3 those of row y + 2 again
// (finite values that meet the zero weights of k >= 12).  `p` points at the
// lane's pair in row y + min(kb, 2) of the copy packed at the parity of x
// (the even-column copy at s0, the odd one at s0 + kS0Odd); `off` adds
// whole rows.  The pair of words is 4-byte aligned only: two 32-bit reads.
typedef short s16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ s16x4 in_operand(const uint32_t* p, int off) {
    return __builtin_bit_cast(s16x4, make_uint2(p[off], p[off + 1]));
}

// relu, bf16, and a lane's 4 channels of one pixel (8 bytes); `keep` false
// writes zeros.  The relu acts on the rounded pairs as int16 maxima with 0
// (a bf16 is negative exactly when its int16 pattern is), which equals
// rounding the relu of the fp32 values.
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
    const s16x2 m = __builtin_elementwise_max(__builtin_bit_cast(s16x2, v), (s16x2){0, 0});
    return __builtin_bit_cast(uint32_t, m);
}
// Offset (elements) of 8-channel chunk c (0 or 1) of the pixel at row y,
// column x of a planar region with row pitch `pitch`, within its plane: the
// two 16-byte halves are swapped in columns with bit 2 set.  Eight
// consecutive pixels' same chunk then covers 128 distinct bytes (a store
// cycle), and the four 4-pixel blocks of a read cycle (two of chunk 0, two
// of chunk 1, 8 columns apart) fall on distinct banks.  The swap depends on
// the column only, so a lane that walks down a column keeps its offsets.
__device__ __forceinline__ int chunk_off(int y, int x, int pitch, int c) {
    return (y * pitch + x) * 16 + 8 * ((c ^ (x >> 2)) & 1);
}

// One MFMA result block (16 channels 16h .. 16h+15 of 16 pixels; lane
// (j, kb) holds channels 16h + 4kb .. +3 of pixel j) to a planar LDS region:
// the lane's two packed bf16 pairs x, y are exchanged between rows 2r and
// 2r + 1 (v_permlane16_swap: the even row gets the odd row's x, the odd row
// the even row's y), and every lane stores 8 bytes at `dst` = byte 8 (kb & 1)
// of chunk kb / 2.  Within a chunk the channels are thus stored in the order
// 0, 1, 4, 5, 2, 3, 6, 7 (pair_swap restores the order); the weights of the
// next layer are read in the same order.
__device__ __forceinline__ void store_block(uint16_t* dst, const f32x4& acc, bool keep, bool store) {
    uint32_t x = relu_bf16x2(pack_bf16(acc[0], acc[1]));
    uint32_t y = relu_bf16x2(pack_bf16(acc[2], acc[3]));
    if (!keep) x = y = 0u;
    const auto sw = __builtin_amdgcn_permlane16_swap(x, y, false, false);
    if (store) *(uint2*)dst = make_uint2(sw[0], sw[1]);
}
// The stored channel order of a chunk and back: swap its middle 4-byte words.
__device__ __forceinline__ uint4 pair_swap(uint4 v) { return make_uint4(v.x, v.z, v.y, v.w); }

// A operands of a conv C -> C layer: lane (co = l%16, kb = l/16) of channel
// block h0 + h holds W[16(h0 + h) + co][tap][ch .. ch+7] for K block b, with
// the lane's tap of Sweep<C> and ch = 8*(kb % (C/8)) (zeros in block 4's
// first half for C = 16), in the stored channel order, as packed.
template <int C, int HW>
__device__ __forceinline__ void load_mid_weights(bf16x8 (&w)[HW][MidK<C>::kBlocks], const uint16_t* __restrict__ wb,
                                                 int co, int kb, int h0) {
    using K = MidK<C>;
#pragma unroll
    for (int h = 0; h < HW; h++)
#pragma unroll
        for (int b = 0; b < K::kBlocks; b++) {
            int ty, tx;
            bool zero;
            Sweep<C>::tap(b, kb, ty, tx, zero);
            uint4 v = make_uint4(0, 0, 0, 0);
            if (!zero) v = *(const uint4*)(wb + ((16 * (h0 + h) + co) * 9 + 3 * ty + tx) * C + 8 * (kb % K::kChunks));
            w[h][b] = as_bf16x8(v);
        }
}

// The leftover columns 16 .. WO-1 of a mid layer with output side WO (20 or
// 18), in groups of 16 pixels: 16 / (WO - 16) rows a group, taken in an
// order that keeps each read cycle's lanes on distinct banks at input
// pitches 22 and 20.
template <int WO>
struct MidLeft {
    static constexpr int kK = WO - 16;                 // leftover columns
    static constexpr int kRows = 16 / kK;              // rows per group
    static constexpr int kGroups = (WO + kRows - 1) / kRows;
};
// R output rows of a mid layer's first 16 columns, swept downwards: each
// input row r = 0 .. R+1 of the band is read once per kind (lane bases
// `kin` at band row 0, rows as immediate offsets) and multiplies into the
// output rows r, r - 1, r - 2 it feeds; output row y is complete, and
// stored, once row y + 2 has been read.  yimg: the image row of band row 0.
// HW channel blocks, stored at rout + h * pout.
template <int C, int HW, bool kEdge, int WI, int PO, int R>
__device__ __forceinline__ void mid_sweep(const uint16_t* const (&kin)[Sweep<C>::kKinds], uint16_t* rout, int pout,
                                          const bf16x8 (&w)[HW][MidK<C>::kBlocks], const f32x4 (&bias)[HW],
                                          int yimg, int n, bool xin) {
    using K = MidK<C>;
    using S = Sweep<C>;
    f32x4 acc[R][HW];
#pragma unroll
    for (int y = 0; y < R; y++)
#pragma unroll
        for (int h = 0; h < HW; h++) acc[y][h] = bias[h];
#pragma unroll
    for (int r = 0; r < R + 2; r++) {
#pragma unroll
        for (int k = 0; k < S::kKinds; k++) {
            bool need = false;
#pragma unroll
            for (int b = 0; b < K::kBlocks; b++)
                if (S::kind(b) == k && r - S::dy(b) >= 0 && r - S::dy(b) < R) need = true;
            if (!need) continue;
            const bf16x8 v = as_bf16x8(*(const uint4*)(kin[k] + r * WI * 16));
#pragma unroll
            for (int b = 0; b < K::kBlocks; b++) {
                const int y = r - S::dy(b);
                if (S::kind(b) != k || y < 0 || y >= R) continue;
#pragma unroll
                for (int h = 0; h < HW; h++)
                    acc[y][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[h][b], v, acc[y][h], 0, 0, 0);
            }
        }
        if (r >= 2) {
            const int y = r - 2;
            const bool keep = xin && (!kEdge || (yimg + y >= 0 && yimg + y < n));
#pragma unroll
            for (int h = 0; h < HW; h++) store_block(rout + h * pout + y * PO * 16, acc[y][h], keep, true);
        }
    }
}

// One conv C -> C layer from the LDS region `in` (side and row pitch WI,
// plane stride pin) to `out` (side WI - 2, row pitch PO, plane stride pout)
// whose top-left pixel is image pixel (oy, ox).  Groups of 16 pixels: the
// first 16 columns of each row -- wave w sweeps a band of consecutive rows
// (side 20: five rows each; side 18: four, four, five, five) -- then the
// leftover columns, whose groups read every block's operand themselves.
// The wave computes HW channel blocks, stored from `out` (the region of its
// first block) at plane stride pout; `wave` is its row band.
template <int C, int HW, bool kEdge, int WI, int PO>
__device__ __forceinline__ void mid_layer(const uint16_t* in, int pin, uint16_t* out, int pout, int oy, int ox,
                                          int n, const bf16x8 (&w)[HW][MidK<C>::kBlocks], const f32x4 (&bias)[HW],
                                          int wave, int j16, int kb) {
    using K = MidK<C>;
    using S = Sweep<C>;
    constexpr int WO = WI - 2;
    static_assert(WO == 20 || WO == 18, "mid layer sides 20 and 18");
    const int ch = kb % K::kChunks;  // this lane's 8-channel chunk of every tap
    const uint16_t* const cin = in + (ch >> 1) * pin;
    {
        const bool five = WO == 20 || wave >= 2;  // wave-uniform
        const int y0 = WO == 20 ? 5 * wave : (wave < 2 ? 4 * wave : 8 + 5 * (wave - 2));
        const uint16_t* kin[S::kKinds];
#pragma unroll
        for (int k = 0; k < S::kKinds; k++) kin[k] = cin + chunk_off(y0 + S::krow(k, kb), j16 + S::kcol(k, kb), WI, ch);
        uint16_t* const rout = out + chunk_off(y0, j16, PO, kb >> 1) + 4 * (kb & 1);
        const bool xin = !kEdge || (ox + j16 >= 0 && ox + j16 < n);
        if (five)
            mid_sweep<C, HW, kEdge, WI, PO, 5>(kin, rout, pout, w, bias, oy + y0, n, xin);
        else
            mid_sweep<C, HW, kEdge, WI, PO, 4>(kin, rout, pout, w, bias, oy + y0, n, xin);
    }
    {   // leftover columns: lane j's pixel row is kRows * g + ro(j); only the
        // 18-wide layer's last group has rows past the side (read clamped, not stored)
        using M = MidLeft<WO>;
        const int r = j16 / M::kK;
        const int ro = M::kK == 4 ? (r ^ (r >> 1)) : ((r & 3) * 2 + (r >> 2));
        const int y0 = M::kRows * wave + ro, x = 16 + j16 % M::kK;
        const int yc = min(y0, WO - 1);
        const uint16_t* rin[K::kBlocks];
#pragma unroll
        for (int b = 0; b < K::kBlocks; b++) {
            int ty, tx;
            bool zero;
            S::tap(b, kb, ty, tx, zero);
            rin[b] = cin + chunk_off(yc + ty, x + tx, WI, ch);
        }
        uint16_t* const o = out + chunk_off(yc, x, PO, kb >> 1) + 4 * (kb & 1);
#pragma unroll
        for (int mm = 0; mm < (M::kGroups + kFWaves - 1) / kFWaves; mm++) {
            if (wave + kFWaves * mm >= M::kGroups) break;  // wave-uniform
            const int y = y0 + kFWaves * M::kRows * mm;
            constexpr int kStep = kFWaves * M::kRows;      // rows between a lane's groups
            f32x4 acc[HW];
#pragma unroll
            for (int h = 0; h < HW; h++) acc[h] = bias[h];
#pragma unroll
            for (int i = 0; i < K::kBlocks; i++) {
                const int b = S::order(i);
                const bf16x8 v = as_bf16x8(*(const uint4*)(rin[b] + mm * kStep * WI * 16));
#pragma unroll
                for (int h = 0; h < HW; h++)
                    acc[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[h][b], v, acc[h], 0, 0, 0);
            }
            const int iy = oy + y, ix = ox + x;
            const bool keep = !kEdge || (iy >= 0 && iy < n && ix >= 0 && ix < n);
#pragma unroll
            for (int h = 0; h < HW; h++) store_block(o + h * pout + mm * kStep * PO * 16, acc[h], keep, y < WO);
        }
    }
}

// conv 1 -> C on the 22 x 22 a1 region at image pixel (i0-3, j0-3), from
// the packed r0 tile s0, on v_mfma_f32_16x16x16_bf16 (the 12 products of a
// 3 x 3 window and 4 zeros per pixel).  Groups of 16 pixels: the first 16
// columns of each row (row band w: rows w, w + 4, ..., two at a time), then the
// columns 16..21 as blocks of 2 rows x 8 columns (lanes of columns 22 and
// 23 idle).  A lane's operand and store addresses are fixed; rows are
// immediate offsets.  The wave computes H channel blocks, stored from `sa`
// (the region of its first block).
template <int C, int H, bool kEdge>
__device__ __forceinline__ void conv_in(const uint32_t* s0, uint16_t* sa, const s16x4 (&w1)[H],
                                        const f32x4 (&bias1)[H], int i0, int j0, int n, int wave, int j16, int kb) {
    const int ry = min(kb, 2);  // tile row offset of this lane's operand
    auto base = [&](int x) { return ((x & 1) ? s0 + kS0Odd : s0) + (x >> 1); };
    {   // first 16 columns
        const uint32_t* const p = base(j16) + (wave + ry) * kS0;
        uint16_t* const o = sa + chunk_off(wave, j16, kF1, kb >> 1) + 4 * (kb & 1);
        const bool xin = !kEdge || (j0 - 3 + j16 >= 0 && j0 - 3 + j16 < n);
        // rows wave + 4r: r = 0..3, then 4 and (waves 0, 1) 5
        auto rows = [&](auto nr_tag, int r0) {
            constexpr int NR = decltype(nr_tag)::value;
            s16x4 v[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) v[r] = in_operand(p, 4 * (r0 + r) * kS0);
            f32x4 acc[NR][H];
#pragma unroll
            for (int r = 0; r < NR; r++)
#pragma unroll
                for (int h = 0; h < H; h++)
                    acc[r][h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[h], v[r], bias1[h], 0, 0, 0);
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int y = i0 - 3 + wave + 4 * (r0 + r);
                const bool keep = xin && (!kEdge || (y >= 0 && y < n));
#pragma unroll
                for (int h = 0; h < H; h++) store_block(o + h * kPlA + 4 * (r0 + r) * kF1 * 16, acc[r][h], keep, true);
            }
        };
        rows(std::integral_constant<int, 4>(), 0);
        if (wave + 20 < kF1)
            rows(std::integral_constant<int, 2>(), 4);
        else
            rows(std::integral_constant<int, 1>(), 4);
    }
    {   // columns 16..21: group g covers rows 2g, 2g + 1
        const int r = j16 >> 3, c = j16 & 7;
        const bool cin = c < kF1 - 16;
        const int x = cin ? 16 + c : kF1 - 1;
        const uint32_t* const p = base(x) + (2 * wave + r + ry) * kS0;
        uint16_t* const o = sa + chunk_off(2 * wave + r, x, kF1, kb >> 1) + 4 * (kb & 1);
        const bool xin = !kEdge || (j0 - 3 + x >= 0 && j0 - 3 + x < n);
#pragma unroll
        for (int mm = 0; mm < (kF1 / 2 + kFWaves - 1) / kFWaves; mm++) {
            if (wave + kFWaves * mm >= kF1 / 2) break;  // wave-uniform
            const int y = 2 * wave + r + 8 * mm;
            const s16x4 v = in_operand(p, 8 * mm * kS0);
            const bool keep = xin && (!kEdge || (i0 - 3 + y >= 0 && i0 - 3 + y < n));
            f32x4 acc[H];
#pragma unroll
            for (int h = 0; h < H; h++) acc[h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[h], v, bias1[h], 0, 0, 0);
#pragma unroll
            for (int h = 0; h < H; h++) store_block(o + h * kPlA + 8 * mm * kF1 * 16, acc[h], keep, cin);
        }
    }
}

// Copy the centre 16 x 16 pixels of a planar LDS region (side wi, plane
// stride pl, centre offset `off`) to the NHWC image `dst` (training: the
// backward kernels read it).
template <int C>
__device__ __forceinline__ void store_centre(const uint16_t* reg, int pl, int wi, int off,
                                             uint16_t* __restrict__ dst, int i0, int j0, int n) {
    constexpr int kV = C / 8;  // 16-byte vectors per pixel
    for (int t = threadIdx.x; t < kFT * kFT * kV; t += 64 * FwdSplit<C>::kWaves) {
        const int px = t / kV, h = t % kV;  // h: channels 8h .. 8h+7 (plane h / 2, half h % 2)
        const int y = px / kFT, x = px % kFT;
        const uint4 v = pair_swap(*(const uint4*)(reg + (h >> 1) * pl + chunk_off(y + off, x + off, wi, h & 1)));
        *(uint4*)(dst + ((size_t)(i0 + y) * n + j0 + x) * C + 8 * h) = v;
    }
}

// One output tile.  kEdge: some region pixel lies outside the image (zero
// padding to apply); interior tiles skip the bounds tests.
template <int C, bool kTrain, bool kEdge>
__device__ __forceinline__ void cnn_tile(const float* __restrict__ r0, const float* __restrict__ prm,
                                         const uint16_t* __restrict__ w1b, const uint16_t* __restrict__ wb2,
                                         const uint16_t* __restrict__ wb3, float* __restrict__ out, int n,
                                         uint16_t* __restrict__ a1g, uint16_t* __restrict__ a2g,
                                         uint16_t* __restrict__ a3g, const float* __restrict__ target,
                                         float* __restrict__ dout, float inv_count, double* __restrict__ loss,
                                         uint16_t* sa, uint16_t* sb, uint32_t* s0, uint32_t* w4,
                                         float* red) {
    using K = MidK<C>;
    using F = FwdSplit<C>;
    constexpr int H = F::kHW;  // channel blocks of this wave
    constexpr LearnedLayout L = learned_layout(C);
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * kFT, j0 = blockIdx.x * kFT;
    const float* img = r0 + (size_t)b * n * n;
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // scalar: wave-uniform tests branch
    const int wave = F::kSplit > 1 ? wid / F::kSplit : wid;     // row band
    const int h0 = F::kSplit > 1 ? (wid % F::kSplit) * H : 0;  // first channel block
    const int j16 = lane & 15, kb = lane >> 4;  // A row / B column, and k block
    // prologue loads in flight together: W2 operands, W1, biases, W4, r0 tile
    bf16x8 w[H][K::kBlocks];
    load_mid_weights<C, H>(w, wb2, j16, kb, h0);
    s16x4 w1[H];
    f32x4 bias1[H], bias2[H], bias3[H];
#pragma unroll
    for (int h = 0; h < H; h++) {
        const int co = 16 * (h0 + h);
        w1[h] = __builtin_bit_cast(s16x4, *(const uint2*)(w1b + (co + j16) * 16 + 4 * kb));  // W1[co][ky = kb][0..3]
#pragma unroll
        for (int r = 0; r < 4; r++) {
            bias1[h][r] = prm[L.B1 + co + 4 * kb + r];
            bias2[h][r] = prm[L.B2 + co + 4 * kb + r];
            bias3[h][r] = prm[L.B3 + co + 4 * kb + r];
        }
    }
    {   // the 24 x 24 r0 tile as 4-column quads: thread t < 144 loads quad
        // t % 6 of tile row t / 6 (columns j0 - 4 + 4q .. +3: a multiple of 4,
        // so a quad lies wholly inside or outside the image) and the next
        // column, and writes its two even- and two odd-column bf16 pairs
        constexpr int kQ = kF0 / 4;
        const int t = threadIdx.x;
        if (t < kF0 * kQ) {
            const int row = t / kQ, q = t % kQ;
            const int ii = i0 - 4 + row, jj = j0 - 4 + 4 * q;
            const bool rin = !kEdge || (ii >= 0 && ii < n);
            const float* src = img + (size_t)ii * n + jj;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float e = 0.0f;  // column 4q + 4 (the tile's pad column 24 for q = 5)
            if (rin && (!kEdge || (jj >= 0 && jj < n))) v = *(const float4*)src;
            if (rin && q + 1 < kQ && (!kEdge || (jj + 4 >= 0 && jj + 4 < n))) e = src[4];
            *(uint2*)(s0 + row * kS0 + 2 * q) = make_uint2(pack_bf16(v.x, v.y), pack_bf16(v.z, v.w));
            *(uint2*)(s0 + kS0Odd + row * kS0 + 2 * q) = make_uint2(pack_bf16(v.y, v.z), pack_bf16(v.w, e));
        }
    }
    for (int t = threadIdx.x; t < 9 * C / 2; t += 64 * F::kWaves) {  // in the stored channel order
        const int q = t & 3;
        w4[t - q + ((q == 1 || q == 2) ? 3 - q : q)] = pack_bf16(prm[L.W4 + 2 * t], prm[L.W4 + 2 * t + 1]);
    }
    __syncthreads();
    conv_in<C, H, kEdge>(s0, sa + h0 * kPlA, w1, bias1, i0, j0, n, wave, j16, kb);
    __syncthreads();
    const size_t img_off = (size_t)b * n * n * C;
    if (kTrain) store_centre<C>(sa, kPlA, kF1, 3, a1g + img_off, i0, j0, n);
    mid_layer<C, H, kEdge, kF1, kF2>(sa, kPlA, sb + h0 * kPlB, kPlB, i0 - 2, j0 - 2, n, w, bias2, wave, j16, kb);
    load_mid_weights<C, H>(w, wb3, j16, kb, h0);  // in flight while the other waves finish
    __syncthreads();
    if (kTrain) store_centre<C>(sb, kPlB, kF2, 2, a2g + img_off, i0, j0, n);
    mid_layer<C, H, kEdge, kF2, kP3>(sb, kPlB, sa + h0 * kPlA, kPlA, i0 - 1, j0 - 1, n, w, bias3, wave, j16, kb);
    __syncthreads();
    if (kTrain) store_centre<C>(sa, kPlA, kP3, 1, a3g + img_off, i0, j0, n);
    // conv C -> 1 + bias + residual on the matrix cores: the K blocks and
    // operand reads of the mid layers over the a3 region, with W4 in every
    // row of A, so every row of D is the output; row 0 (lanes 0..15,
    // register 0) is used.  Wave w of the workgroup's W sweeps output rows
    // R w .. R w + R - 1, R = 16 / W.
    float dsum = 0.0f;  // training: this thread's squared errors
    {
        using S = Sweep<C>;
        constexpr int R = kFT / F::kWaves;
        const int ch = kb % K::kChunks;
        const int y0 = R * wid;
        bf16x8 w4a[K::kBlocks];
#pragma unroll
        for (int q = 0; q < K::kBlocks; q++) {
            int ty, tx;
            bool zero;
            S::tap(q, kb, ty, tx, zero);
            const uint4 wv = *(const uint4*)(w4 + (3 * ty + tx) * (C / 2) + 4 * ch);  // W4[tap][8ch .. 8ch+7]
            w4a[q] = as_bf16x8(zero ? make_uint4(0, 0, 0, 0) : wv);
        }
        const uint16_t* kin[S::kKinds];
#pragma unroll
        for (int k = 0; k < S::kKinds; k++)
            kin[k] = sa + (ch >> 1) * kPlA + chunk_off(y0 + S::krow(k, kb), j16 + S::kcol(k, kb), kP3, ch & 1);
        const float b4 = prm[L.B4];
        f32x4 acc[R];
#pragma unroll
        for (int y = 0; y < R; y++) acc[y] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int r = 0; r < R + 2; r++) {
#pragma unroll
            for (int k = 0; k < S::kKinds; k++) {
                bool need = false;
#pragma unroll
                for (int q = 0; q < K::kBlocks; q++)
                    if (S::kind(q) == k && r - S::dy(q) >= 0 && r - S::dy(q) < R) need = true;
                if (!need) continue;
                const bf16x8 v = as_bf16x8(*(const uint4*)(kin[k] + r * kP3 * 16));
#pragma unroll
                for (int q = 0; q < K::kBlocks; q++) {
                    const int y = r - S::dy(q);
                    if (S::kind(q) != k || y < 0 || y >= R) continue;
                    acc[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w4a[q], v, acc[y], 0, 0, 0);
                }
            }
            if (r >= 2 && kb == 0) {
                const int y = y0 + r - 2;
                const size_t p = ((size_t)b * n + i0 + y) * n + j0 + j16;
                const float o = img[(size_t)(i0 + y) * n + j0 + j16] + (acc[r - 2][0] + b4);
                out[p] = o;
                if (kTrain && target) {
                    const float d = o - target[p];
                    dout[p] = 2.0f * d * inv_count;
                    dsum += d * d;
                }
            }
        }
    }
    if (kTrain && target) {  // uniform
        red[threadIdx.x] = dsum;
        __syncthreads();
        for (int st = 32 * F::kWaves; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicAdd(loss, (double)red[0]);
    }
}


// C = 16: at most 96 VGPRs, so the five workgroups per CU that LDS allows
// (five waves per SIMD) also fit the registers.  C = 32: eight waves, at
// most 128 VGPRs, two workgroups per CU (LDS): four waves per SIMD.
template <int C, bool kTrain>
__global__ __launch_bounds__(64 * FwdSplit<C>::kWaves, C == 16 ? 5 : 2) void k_cnn_fused_fwd(const float* __restrict__ r0,
                                                                const float* __restrict__ prm,
                                                                const uint16_t* __restrict__ w1b,
                                                                const uint16_t* __restrict__ wb2,
                                                                const uint16_t* __restrict__ wb3,
                                                                float* __restrict__ out, int n,
                                                                uint16_t* __restrict__ a1g, uint16_t* __restrict__ a2g,
                                                                uint16_t* __restrict__ a3g,
                                                                const float* __restrict__ target,
                                                                float* __restrict__ dout, float inv_count,
                                                                double* __restrict__ loss) {
    __shared__ __align__(16) uint16_t sa[(C / 16) * kPlA];  // a1, then a3 (planar)
    __shared__ __align__(16) uint16_t sb[(C / 16) * kPlB];  // a2 (planar)
    __shared__ __align__(16) uint32_t s0[kS0Odd + kF0 * kS0];             // r0 as bf16 pairs at even, then odd columns
    __shared__ __align__(16) uint32_t w4[9 * C / 2];      // W4 as bf16 pairs
    __shared__ float red[64 * FwdSplit<C>::kWaves];
    const int i0 = blockIdx.y * kFT, j0 = blockIdx.x * kFT;
    const bool interior = i0 >= 4 && i0 + kFT + 4 <= n && j0 >= 4 && j0 + kFT + 4 <= n;  // uniform
    if (interior)
        cnn_tile<C, kTrain, false>(r0, prm, w1b, wb2, wb3, out, n, a1g, a2g, a3g, target, dout, inv_count, loss, sa,
                                   sb, s0, w4, red);
    else
        cnn_tile<C, kTrain, true>(r0, prm, w1b, wb2, wb3, out, n, a1g, a2g, a3g, target, dout, inv_count, loss, sa,
                                  sb, s0, w4, red);
}

// ---------------------------------------------------------------------------
// backward: activation gradients (fp32, NHWC), masked by the relu of the
// layer input they flow into
// ---------------------------------------------------------------------------

// dA3[q][ci] = sum_t W4[t][ci] * dout[q - off(t)], times (a3 > 0).
template <int C>
__global__ __launch_bounds__(256) void k_conv_out_bwd(const float* __restrict__ dout, const uint16_t* __restrict__ a3,
                                                      const float* __restrict__ prm, float* __restrict__ dz3, int n) {
    __shared__ float w[9 * C];
    __shared__ float ds[(T + 2) * (T + 2)];
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * T, j0 = blockIdx.x * T;
    for (int t = threadIdx.x; t < 9 * C; t += 256) w[t] = rbf(prm[learned_layout(C).W4 + t]);
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

// 16 x 8 pixel tiles for the single-channel layers' backward kernels.
constexpr int TY = 8;
// dr0[q] = dout[q] + sum_{t,co} W1[co][t] * dZ1[q - off(t)][co].
template <int C>
__global__ __launch_bounds__(128) void k_conv_in_bwd(const float* __restrict__ dz1, const float* __restrict__ dout,
                                                     const float* __restrict__ prm, float* __restrict__ dr0, int n) {
    __shared__ float w[C * 9];
    __shared__ float ds[(TY + 2) * (T + 2) * (C + 1)];  // pixel pitch C+1: distinct banks per lane
    const int b = blockIdx.z;
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * T;
    for (int t = threadIdx.x; t < C * 9; t += 128) w[t] = rbf(prm[learned_layout(C).W1 + t]);
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
template <int C>
__global__ __launch_bounds__(256) void k_wgrad(const float* __restrict__ dz, int zc, const uint16_t* __restrict__ xh,
                                               const float* __restrict__ xf, int xc, float* __restrict__ grad, int woff,
                                               int boff, int wlayout_tc, int n, int batch) {
    extern __shared__ float sm[];
    float* zs = sm;                        // TY*T*zc
    float* xs = sm + TY * T * zc;          // (TY+2)(T+2)*xc
    const int i0 = blockIdx.y * TY, j0 = blockIdx.x * T;
    constexpr int kWgPer = 9 * C * C / 256;  // weights per thread (largest layer)
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
// backward of the C -> C layers on the matrix cores
// ---------------------------------------------------------------------------

// fp32 pair -> bf16 hi and lo parts (hi = bf16(v), lo = bf16(v - hi); v - hi
// is exact in fp32), packed as pairs: hi + lo equals v to 2^-17 of |v|, so an
// MFMA on each part keeps the product with a bf16 operand to that accuracy.
__device__ __forceinline__ void split_bf16(float a, float b, uint32_t& hi, uint32_t& lo) {
    hi = pack_bf16(a, b);
    lo = pack_bf16(a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xffff0000u));
}

// Activation gradient of a conv C -> C layer:
//   dzp[q][ci] = (x[q][ci] > 0) * sum_{t,co} W[co][t][ci] dz[q - off(t)][co],
// off(t) = (t/3 - 1, t%3 - 1).  One workgroup per 16 x 16 tile.  The fp32
// dz of the tile and a one-pixel halo goes to two planar bf16 LDS regions,
// its hi and lo parts (split_bf16), laid out like the forward's (chunk_off,
// channels in order); each K block runs one v_mfma_f32_16x16x32_bf16 on
// each part with fp32 accumulation.  A: the transposed weights wt[ci][t][co].
// K blocks as in the forward (tap pairs at C = 16, single taps at C = 32);
// tap t reads halo pixel (y + 2 - t/3, x + 2 - t%3).  Wave w takes rows w,
// w + 4, w + 8, w + 12, four accumulator chains.
template <int C>
__global__ __launch_bounds__(256) void k_conv_mid_bwd_mfma(const float* __restrict__ dz,
                                                           const uint16_t* __restrict__ x,
                                                           const uint16_t* __restrict__ wt,
                                                           float* __restrict__ dzp, int n) {
    using K = MidK<C>;
    constexpr int H = K::kHalves;
    constexpr int kW = kFT + 2;         // halo tile side
    constexpr int kPl = kW * kW * 16;   // plane stride (elements)
    __shared__ __align__(16) uint16_t sh[(C / 16) * kPl];
    __shared__ __align__(16) uint16_t sl[(C / 16) * kPl];
    const int b = blockIdx.z, i0 = blockIdx.y * kFT, j0 = blockIdx.x * kFT;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j16 = lane & 15, kb = lane >> 4;
    bf16x8 w[H][K::kBlocks];
#pragma unroll
    for (int h = 0; h < H; h++)
#pragma unroll
        for (int tp = 0; tp < K::kBlocks; tp++) {
            const int tap = K::kTapsPer * tp + kb / K::kChunks;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (tap < 9) v = *(const uint4*)(wt + ((16 * h + j16) * 9 + tap) * C + 8 * (kb % K::kChunks));
            w[h][tp] = as_bf16x8(v);
        }
    const float* db = dz + (size_t)b * n * n * C;
    for (int t = threadIdx.x; t < kW * kW * (C / 8); t += 64 * kFWaves) {
        const int px = t / (C / 8), c = t % (C / 8);  // halo pixel and 8-channel chunk
        const int y = px / kW, xx = px % kW;
        const int ii = i0 - 1 + y, jj = j0 - 1 + xx;
        float4 v0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v1 = v0;
        if (ii >= 0 && ii < n && jj >= 0 && jj < n) {
            const float4* src = (const float4*)(db + ((size_t)ii * n + jj) * C + 8 * c);
            v0 = src[0];
            v1 = src[1];
        }
        uint32_t hi[4], lo[4];
        split_bf16(v0.x, v0.y, hi[0], lo[0]);
        split_bf16(v0.z, v0.w, hi[1], lo[1]);
        split_bf16(v1.x, v1.y, hi[2], lo[2]);
        split_bf16(v1.z, v1.w, hi[3], lo[3]);
        const int off = (c >> 1) * kPl + chunk_off(y, xx, kW, c & 1);
        *(u32x4*)(sh + off) = (u32x4){hi[0], hi[1], hi[2], hi[3]};
        *(u32x4*)(sl + off) = (u32x4){lo[0], lo[1], lo[2], lo[3]};
    }
    __syncthreads();
    const int ch = kb % K::kChunks;
    const uint16_t* rh[K::kBlocks];
    const uint16_t* rl[K::kBlocks];
#pragma unroll
    for (int tp = 0; tp < K::kBlocks; tp++) {
        const int tap = min(K::kTapsPer * tp + kb / K::kChunks, 8);
        const int o = (ch >> 1) * kPl + chunk_off(wave + 2 - tap / 3, j16 + 2 - tap % 3, kW, ch & 1);
        rh[tp] = sh + o;
        rl[tp] = sl + o;
    }
    f32x4 acc[4][H];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int h = 0; h < H; h++) acc[r][h] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int tp = 0; tp < K::kBlocks; tp++) {
        bf16x8 vh[4], vl[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            vh[r] = as_bf16x8(*(const uint4*)(rh[tp] + 4 * r * kW * 16));
            vl[r] = as_bf16x8(*(const uint4*)(rl[tp] + 4 * r * kW * 16));
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int h = 0; h < H; h++) {
                acc[r][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[h][tp], vh[r], acc[r][h], 0, 0, 0);
                acc[r][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[h][tp], vl[r], acc[r][h], 0, 0, 0);
            }
    }
    // lane (j16, kb) holds channels 16h + 4kb .. +3 of pixel (wave + 4r, j16)
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const size_t q = ((size_t)b * n + i0 + wave + 4 * r) * n + j0 + j16;
#pragma unroll
        for (int h = 0; h < H; h++) {
            const uint2 xv = *(const uint2*)(x + q * C + 16 * h + 4 * kb);
            float4 o;
            o.x = __uint_as_float(xv.x << 16) > 0.0f ? acc[r][h][0] : 0.0f;
            o.y = __uint_as_float(xv.x & 0xffff0000u) > 0.0f ? acc[r][h][1] : 0.0f;
            o.z = __uint_as_float(xv.y << 16) > 0.0f ? acc[r][h][2] : 0.0f;
            o.w = __uint_as_float(xv.y & 0xffff0000u) > 0.0f ? acc[r][h][3] : 0.0f;
            *(float4*)(dzp + q * C + 16 * h + 4 * kb) = o;
        }
    }
}

// Weight gradient of a conv C -> C layer:
//   dW[co][t][ci] += sum_p dz[p][co] x[p + off(t)][ci],  db[co] += sum_p dz[p][co].
// One workgroup per 16 x 16 tile, summed over the slices b = blockIdx.z,
// blockIdx.z + gridDim.z, ... in registers, then added into grad[] with
// atomics once.  K = pixels: A[co][p] = dz, split into bf16 hi and lo parts
// (split_bf16) and transposed in LDS ([co][16 rows][16 columns], co stride
// 272 elements); B[p][ci] = x (bf16, exact), transposed with a one-pixel
// halo ([ci][18 rows][pitch 24]).  A K block is two tile rows: lane block kb
// holds row kb >> 1, columns 8 (kb & 1) .. +7.  Tap (ty, tx) shifts B by ty
// rows and tx columns; the column shift comes from one aligned 16-byte read,
// the next 4 bytes and v_alignbyte.  The co stride of A and the ci stride of
// B keep each 16-byte read cycle on distinct banks.  Wave w takes the taps
// w, w + 4 and (w = 0) 8.
template <int C>
__global__ __launch_bounds__(256) void k_wgrad_mfma(const float* __restrict__ dz, const uint16_t* __restrict__ x,
                                                    float* __restrict__ grad, int woff, int boff, int n, int batch) {
    constexpr int H = C / 16;
    constexpr int kAS = 272;         // A: co stride (elements)
    constexpr int kXP = 24;          // B: row pitch (elements)
    constexpr int kXS = 18 * kXP;    // B: ci stride
    __shared__ __align__(16) uint16_t ah[C * kAS];
    __shared__ __align__(16) uint16_t al[C * kAS];
    __shared__ __align__(16) uint16_t xs[C * kXS];
    __shared__ float bsum[C];
    const int i0 = blockIdx.y * kFT, j0 = blockIdx.x * kFT;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j16 = lane & 15, kb = lane >> 4;
    if ((int)threadIdx.x < C) bsum[threadIdx.x] = 0.0f;
    f32x4 acc[3][H][H];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int a = 0; a < H; a++)
#pragma unroll
            for (int c = 0; c < H; c++) acc[k][a][c] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    for (int b = blockIdx.z; b < batch; b += gridDim.z) {
        __syncthreads();
        // A: item (co, 8 pixels of a row)
        for (int t = threadIdx.x; t < C * 32; t += 64 * kFWaves) {
            const int co = t % C, pg = t / C;
            const int row = pg >> 1, c0 = 8 * (pg & 1);
            const float* src = dz + (((size_t)b * n + i0 + row) * n + j0 + c0) * C + co;
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = src[(size_t)e * C];
            uint32_t hi[4], lo[4];
#pragma unroll
            for (int e = 0; e < 4; e++) split_bf16(v[2 * e], v[2 * e + 1], hi[e], lo[e]);
            *(u32x4*)(ah + co * kAS + row * 16 + c0) = (u32x4){hi[0], hi[1], hi[2], hi[3]};
            *(u32x4*)(al + co * kAS + row * 16 + c0) = (u32x4){lo[0], lo[1], lo[2], lo[3]};
            atomicAdd(&bsum[co], ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7])));
        }
        // B: item (ci, halo row, 8 columns); columns 18..23 are zeros
        for (int t = threadIdx.x; t < C * 18 * 3; t += 64 * kFWaves) {
            const int ci = t % C, rg = t / C;
            const int row = rg / 3, g = rg % 3;
            const int ii = i0 - 1 + row;
            uint32_t v[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                uint32_t pr = 0;
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int col = 8 * g + 2 * e + u, jj = j0 - 1 + col;
                    uint32_t h = 0;
                    if (col < 18 && ii >= 0 && ii < n && jj >= 0 && jj < n)
                        h = x[(((size_t)b * n + ii) * n + jj) * C + ci];
                    pr |= h << (16 * u);
                }
                v[e] = pr;
            }
            *(u32x4*)(xs + ci * kXS + row * kXP + 8 * g) = (u32x4){v[0], v[1], v[2], v[3]};
        }
        __syncthreads();
#pragma unroll 2
        for (int kblk = 0; kblk < 8; kblk++) {
            const int prow = 2 * kblk + (kb >> 1), pc = 8 * (kb & 1);
            bf16x8 a_h[H], a_l[H];
#pragma unroll
            for (int a = 0; a < H; a++) {
                a_h[a] = as_bf16x8(*(const uint4*)(ah + (16 * a + j16) * kAS + prow * 16 + pc));
                a_l[a] = as_bf16x8(*(const uint4*)(al + (16 * a + j16) * kAS + prow * 16 + pc));
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int tap = wave + 4 * k;
                if (tap > 8) break;  // wave-uniform
                const int ty = tap / 3, tx = tap % 3;
#pragma unroll
                for (int c = 0; c < H; c++) {
                    const uint16_t* xr = xs + (16 * c + j16) * kXS + (prow + ty) * kXP + pc;
                    const uint4 d = *(const uint4*)xr;
                    const uint32_t d4 = *(const uint32_t*)(xr + 8);
                    uint4 s = d;  // columns pc + tx .. +7 of the row
                    if (tx == 1) {
                        s.x = __builtin_amdgcn_alignbyte(d.y, d.x, 2);
                        s.y = __builtin_amdgcn_alignbyte(d.z, d.y, 2);
                        s.z = __builtin_amdgcn_alignbyte(d.w, d.z, 2);
                        s.w = __builtin_amdgcn_alignbyte(d4, d.w, 2);
                    } else if (tx == 2) {
                        s = make_uint4(d.y, d.z, d.w, d4);
                    }
                    const bf16x8 bv = as_bf16x8(s);
#pragma unroll
                    for (int a = 0; a < H; a++) {
                        acc[k][a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h[a], bv, acc[k][a][c], 0, 0, 0);
                        acc[k][a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l[a], bv, acc[k][a][c], 0, 0, 0);
                    }
                }
            }
        }
    }
    // lane (j16, kb) holds D[co = 16a + 4kb + r][ci = 16c + j16]
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int tap = wave + 4 * k;
        if (tap > 8) break;
#pragma unroll
        for (int a = 0; a < H; a++)
#pragma unroll
            for (int c = 0; c < H; c++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    atomicAdd(grad + woff + ((16 * a + 4 * kb + r) * 9 + tap) * C + 16 * c + j16, acc[k][a][c][r]);
    }
    __syncthreads();
    if ((int)threadIdx.x < C) atomicAdd(grad + boff + threadIdx.x, bsum[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// filter and back-projector gradients
// ---------------------------------------------------------------------------

// Adjoint of k_fbp_backproject as a gather: dq[a][k] = sum over the pixels
// p with floor(kf_p) in {k - 1, k} of p's weight on bin k -- v - w v when
// floor(kf_p) = k, w v when it is k - 1 (w = kf_p - floor(kf_p), v = scale *
// dr[p]) -- with kf_p = (x c + y sn) / delta + det_c computed as the
// back-projection computes it.  One thread per (slice, angle, bin).  It walks
// the rows (the columns when |c| < |sn|) in ascending order; in each, kf is
// linear along the row with slope c (sn) per pixel, so the pixels with kf in
// [k - 1, k + 1) are a run of at most 2 / max(|c|, |sn|) + 1 <= 3.9 whose
// ends follow from the row's kf at pixel 0; the run, widened by one pixel
// each side, is tested pixel by pixel.  Fixed order, no atomics.
// drT: dr transposed per slice, read on column walks so that a wave's
// neighbouring bins read neighbouring addresses there too.
__global__ __launch_bounds__(256) void k_bp_adjoint_gather(const float* __restrict__ dr,
                                                           const float* __restrict__ drT, float* __restrict__ dq,
                                                           const float2* __restrict__ cs, FbpGeom g, float scale) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y, b = blockIdx.z;
    if (k >= g.ndet) return;
    const float2 t2 = cs[a];
    const float c = t2.x, sn = t2.y;
    const float inv = 1.0f / g.delta;
    const bool rows = fabsf(c) >= fabsf(sn);  // walk rows: kf varies along j with slope c
    const float* const im = (rows ? dr : drT) + (size_t)b * g.n * g.n;  // line u at im + u n
    const float slope = rows ? c : -sn;        // d kf / d (index along the walk), about
    const float rsl = 1.0f / slope;
    const int cnt = (int)(2.0f * fabsf(rsl)) + 3;  // candidates per line: the run, one more each side
    float acc = 0.0f;
    for (int u = 0; u < g.n; u++) {
        // kf at index 0 of line u (row u: j = 0; column u: i = 0), approximate
        const float fu = ((float)u + 0.5f) * g.delta;
        const float kf0 = rows ? ((0.5f * g.delta - 1.0f) * c + (1.0f - fu) * sn) * inv + g.det_c
                               : ((fu - 1.0f) * c + (1.0f - 0.5f * g.delta) * sn) * inv + g.det_c;
        // indices t with kf0 + t slope in [k - 1, k + 1)
        const float ta = ((float)k - 1.0f - kf0) * rsl, tb = ((float)k + 1.0f - kf0) * rsl;
        const int t0 = (int)floorf(fminf(ta, tb)) - 1;
        for (int q = 0; q < cnt; q++) {
            const int t = t0 + q;
            if (t < 0 || t >= g.n) continue;
            const int i = rows ? u : t, j = rows ? t : u;
            const float x = ((float)j + 0.5f) * g.delta - 1.0f, y = 1.0f - ((float)i + 0.5f) * g.delta;
            const float kf = (x * c + y * sn) * inv + g.det_c;
            const float fk = floorf(kf);
            const int k0 = (int)fk;
            if (k0 != k && k0 != k - 1) continue;
            const float w = kf - fk;
            const float v = im[(size_t)u * g.n + t] * scale;
            acc += (k0 == k) ? v - w * v : w * v;
        }
    }
    dq[((size_t)b * g.nang + a) * g.ndet + k] = acc;
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

template <int C>
hipError_t pack_weights_t(const float* prm, uint16_t* wb, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_weights<C>, dim3((9 * C * C + 255) / 256), dim3(256), 0, s, prm, wb);
    return hipGetLastError();
}

hipError_t launch_pack_weights(const float* prm, int C, uint16_t* wb, hipStream_t s) {
    return C == 32 ? pack_weights_t<32>(prm, wb, s) : pack_weights_t<16>(prm, wb, s);
}

template <int C, bool kTrain>
hipError_t fused_t(const LearnedBufs& B, const float* prm, int n, int batch, const float* target, float* dout,
                   double* loss, hipStream_t s) {
    const float inv = 1.0f / ((float)n * n * batch);
    hipLaunchKernelGGL((k_cnn_fused_fwd<C, kTrain>), dim3(n / kFT, n / kFT, batch),
                       dim3(64 * FwdSplit<C>::kWaves), 0, s, B.r0,
                       prm, B.w1b, B.wb2, B.wb3, B.out, n, kTrain ? B.a1 : nullptr, kTrain ? B.a2 : nullptr,
                       kTrain ? B.a3 : nullptr, target, dout, inv, loss);
    return hipGetLastError();
}

// Inference: the fused forward, no activations stored (n a multiple of 16).
hipError_t launch_cnn_infer(const LearnedBufs& B, const float* prm, int n, int batch, hipStream_t s) {
    return B.C == 32 ? fused_t<32, false>(B, prm, n, batch, nullptr, nullptr, nullptr, s)
                     : fused_t<16, false>(B, prm, n, batch, nullptr, nullptr, nullptr, s);
}

// Training forward: the same fused kernel, storing a1..a3 (centre tiles),
// out, and with a target the loss gradient and the summed squared error.
hipError_t launch_cnn_forward(const LearnedBufs& B, const float* prm, int n, int batch, const float* target,
                              float* dout, double* loss, hipStream_t s) {
    return B.C == 32 ? fused_t<32, true>(B, prm, n, batch, target, dout, loss, s)
                     : fused_t<16, true>(B, prm, n, batch, target, dout, loss, s);
}

template <int C>
hipError_t backward_t(const LearnedBufs& B, const float* prm, const float* dout, float* grad, int n, int batch,
                      hipStream_t s) {
    constexpr LearnedLayout L = learned_layout(C);
    // activation gradients
    hipLaunchKernelGGL(k_conv_out_bwd<C>, SKE_GRID(n, batch), dim3(256), 0, s, dout, B.a3, prm, B.dz3, n);
    const dim3 g16(n / kFT, n / kFT, batch);
    hipLaunchKernelGGL(k_conv_mid_bwd_mfma<C>, g16, dim3(64 * kFWaves), 0, s, B.dz3, B.a2, B.wt3, B.dz2, n);
    hipLaunchKernelGGL(k_conv_mid_bwd_mfma<C>, g16, dim3(64 * kFWaves), 0, s, B.dz2, B.a1, B.wt2, B.dz1, n);
    dim3 g8(n / T, n / TY, batch);
    hipLaunchKernelGGL(k_conv_in_bwd<C>, g8, dim3(128), 0, s, B.dz1, dout, prm, B.dr0, n);
    // weight gradients: about 2048 workgroups, each summing over a share of
    // the slices
    const int tiles = (n / T) * (n / TY);
    const dim3 gw(n / T, n / TY, (unsigned)std::max(1, std::min(batch, 2048 / tiles)));
    const size_t lw4 = sizeof(float) * (TY * T * 1 + (TY + 2) * (T + 2) * C);
    hipLaunchKernelGGL(k_wgrad<C>, gw, dim3(256), lw4, s, dout, 1, B.a3, (const float*)nullptr, C, grad, L.W4, L.B4, 1,
                       n, batch);
    const int tiles16 = (n / kFT) * (n / kFT);
    const dim3 gm(n / kFT, n / kFT, (unsigned)std::max(1, std::min(batch, 2048 / tiles16)));
    hipLaunchKernelGGL(k_wgrad_mfma<C>, gm, dim3(64 * kFWaves), 0, s, B.dz3, B.a2, grad, L.W3, L.B3, n, batch);
    hipLaunchKernelGGL(k_wgrad_mfma<C>, gm, dim3(64 * kFWaves), 0, s, B.dz2, B.a1, grad, L.W2, L.B2, n, batch);
    const size_t lw1 = sizeof(float) * (TY * T * C + (TY + 2) * (T + 2) * 1);
    hipLaunchKernelGGL(k_wgrad<C>, gw, dim3(256), lw1, s, B.dz1, C, (const uint16_t*)nullptr, B.r0, 1, grad, L.W1,
                       L.B1, 0, n, batch);
    return hipGetLastError();
}

hipError_t launch_cnn_backward(const LearnedBufs& B, const float* prm, const float* dout, float* grad, int n,
                               int batch, hipStream_t s) {
    return B.C == 32 ? backward_t<32>(B, prm, dout, grad, n, batch, s) : backward_t<16>(B, prm, dout, grad, n, batch, s);
}

// out[b][j][i] = in[b][i][j] for n x n slices, through 32 x 32 LDS tiles.
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ in, float* __restrict__ out, int n) {
    __shared__ float tile[32][33];
    const size_t off = (size_t)blockIdx.z * n * n;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8)
        if (y0 + k < n && x0 + tx < n) tile[k][tx] = in[off + (size_t)(y0 + k) * n + x0 + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += 8)
        if (x0 + k < n && y0 + tx < n) out[off + (size_t)(x0 + k) * n + y0 + tx] = tile[tx][k];
}

hipError_t launch_filter_backward(const float* dr0, float* drT, float* dq, const float* sino, float* dfull,
                                  float* dsym, const float2* cs, const FbpGeom& g, int batch, hipStream_t s) {
    const float scale = (float)(M_PI / g.nang);
    hipLaunchKernelGGL(k_transpose, dim3((g.n + 31) / 32, (g.n + 31) / 32, batch), dim3(256), 0, s, dr0, drT, g.n);
    hipLaunchKernelGGL(k_bp_adjoint_gather, dim3((g.ndet + 255) / 256, g.nang, batch), dim3(256), 0, s, dr0, drT, dq,
                       cs, g, scale);
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
