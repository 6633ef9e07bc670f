// kernels_sdem.hip -- total-viewshed operator (src/kernelSDEM.h:168-222) on
// gfx950.
//
// Layout: one workgroup (4 waves) per skewed row.  The row is staged once in
// LDS; each wave takes 64 consecutive observers and runs their right scans
// then their left scans with a wave-uniform step d = 1, 2, ...  Because d is
// uniform, 1/d and d*d come from a scalar table and every lane reads
// row[j +/- d] from consecutive LDS words (no bank conflicts).  Each lane
// keeps the reference's per-observer state (max angle, visible flag, opening
// distance) and accumulates cv in the reference's order, so cv and every
// visibility decision are bit-identical to the CPU path.
//
// Division.  The reference computes cAngle = (S[k]-h)/d with an IEEE divide.
// Variant 1 does exactly that.  Variant 0 uses q = fma(a, yh, a*yl) with
// yh = RN(1/d), yl = RN(1/d - yh): |q_exact - a/d| < 2^-47 |a/d| before the
// final rounding, while for d < 2^22 the quotient a/d of a float a is never
// closer than 2^-13 ulp to a rounding midpoint (a is a multiple of ulp(a/d)
// and a midpoint times d is an odd multiple of half an ulp with at least
// 25 significant bits, so it is never a float).  Hence fma(a, yh, a*yl)
// rounds to exactly RN(a/d): two VALU operations instead of the ~10 of the
// IEEE divide sequence.  The tests check both variants against each other
// and against the oracle.
#include <algorithm>

#include "device_common.h"
#include "launch.h"

namespace ske {

template <int kVariant>
__device__ __forceinline__ void sdem_step(float sk, float h, const float4 t, float& M, float& od2, bool& vis,
                                          float& cv) {
    const float a = sk - h;
    float q;
    if (kVariant == 0)
        q = __builtin_fmaf(a, t.x, a * t.y);
    else
        q = a / t.z;
    const bool above = q > M;
    M = above ? q : M;  // std::max(cAngle, max_angle)
    const bool opening = above && !vis;
    const bool closing = !above && vis;
    vis = above;
    od2 = opening ? t.w : od2;               // open_delta_d^2 (exact: integers < 2^24)
    cv = closing ? cv + (t.w - od2) : cv;    // cv += d*d - od*od
}

template <int kVariant>
__global__ __launch_bounds__(256) void k_op_sdem(const SecDesc* __restrict__ secs,
                                                 const RowWork* __restrict__ work,
                                                 const float* __restrict__ in4,
                                                 const int* __restrict__ target,
                                                 const float* __restrict__ weight,
                                                 const uint32_t* __restrict__ fl,
                                                 const float* __restrict__ Sbuf,
                                                 const float4* __restrict__ rtab, float povh,
                                                 float* __restrict__ arena) {
    extern __shared__ float lds[];
    float* row = lds + kPad;
    const RowWork w = work[blockIdx.x];
    const SecDesc sd = secs[w.sec];
    const int r = w.row;
    const uint32_t v = fl[sd.rows + r];
    const int f = fl_first(v);
    const int L = fl_last(v) - f;
    if (L <= 0) return;
    stage_row(row, sd, r, f, L, in4, target, weight, Sbuf);
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nchunk = (L + 63) >> 6;
    float* so = arena + sd.so + (size_t)r * sd.dim_i + f;
    for (int c = wave; c < nchunk; c += 4) {
        const int j0 = c << 6;
        const int jj = j0 + lane;
        const bool valid = jj < L;
        const float h = row[jj] + povh;
        float cv = 0.0f;
        {   // right scan: k = j+1 .. last-1
            float M = -2000.0f, od2 = 0.0f;
            bool vis = true;
            const int myR = valid ? (L - 1 - jj) : 0;
            const int dall = L - 1 - (j0 + 63);  // every lane active up to here
            const int dmax = L - 1 - j0;
            int d = 1;
            for (; d <= dall; ++d) sdem_step<kVariant>(row[jj + d], h, rtab[d], M, od2, vis, cv);
            for (; d <= dmax; ++d) {
                const float4 t = rtab[d];
                const float sk = row[jj + d];
                if (d <= myR) sdem_step<kVariant>(sk, h, t, M, od2, vis, cv);
            }
        }
        {   // left scan: k = j-1 .. first
            float M = -2000.0f, od2 = 0.0f;
            bool vis = true;
            const int myL = valid ? jj : 0;
            const int dall = (j0 + 63 < L) ? j0 : -1;
            const int dmax = (j0 + 63 < L) ? (j0 + 63) : (L - 1);
            int d = 1;
            for (; d <= dall; ++d) sdem_step<kVariant>(row[jj - d], h, rtab[d], M, od2, vis, cv);
            for (; d <= dmax; ++d) {
                const float4 t = rtab[d];
                const float sk = row[jj - d];
                if (d <= myL) sdem_step<kVariant>(sk, h, t, M, od2, vis, cv);
            }
        }
        if (valid) so[jj] = cv * sd.R;
    }
}

// ---------------------------------------------------------------------------
// Bit-block variant.  One scan direction of one observer sees visibility
// flags V_1..V_D (V_0 = 1: the observer's own cell).  The reference adds
// d_c^2 - d_o^2 for every visible run (d_o, d_c] that closes before the row
// end; since d_c^2 - d_o^2 = sum_{k=d_o}^{d_c-1} (2k+1), that total is
//     cv = sum_{k=0..D} V_k (2k+1)  -  sum over the trailing run of (2k+1)
//        = A - ((D+1)^2 - s^2),  s = 1 + last step with V = 0 (0 if none).
// Steps past a lane's own scan length are recorded as visible: they extend
// the trailing run and cancel exactly, so every lane runs the same
// wave-uniform D (a multiple of 32).  Per step a lane only computes the
// quotient, one compare, one max and appends the flag to a 32-bit word;
// every 32 steps popcounts of the word update A and s:
//     sum_{V=1} (2k+1) = (2 d0 + 65) popc(w) - 2 sum_p p w_p,
//     sum_p p w_p = sum_i 2^i popc(w & M_i),   k = d0 + 32 - p.
// All quantities are integers below 2^24 for rows of at most 4096 cells, so
// the float cv of the reference (whose partial sums are exact integers
// there) is reproduced bit for bit.
// ---------------------------------------------------------------------------
// Guard band on each side of a staged row: the one-chunk scans of a chunk
// run to its first lane's row end, up to 95 cells past the row (128); with
// lane compaction a block only runs while some lane's row continues, so
// every read stays within 32 cells of the row (and of the safe positions).
template <int kG>
constexpr int pad_bits() { return kG > 1 ? 32 : 128; }
constexpr int kChunksMax = 128;  // 32-cell chunks of a row of <= 4096 cells
// Chunk-maximum tables of a staged row, each indexed by chunk c in
// [-kTabLo, kChunksMax + kTabHi): level 0 cm[c] = max row[32c .. 32c+31],
// level 1..3 the maxima of 2, 4 and 8 consecutive chunks starting at c,
// then suffix and prefix maxima.  Chunks outside the row hold -inf (the
// prefix table repeats the row maximum past the end).
constexpr int kTabLo = 16, kTabHi = 16;
constexpr int kTabN = kTabLo + kChunksMax + kTabHi;
enum { kTabL0 = 0, kTabL1 = 1, kTabL2 = 2, kTabL3 = 3, kTabSuf = 4, kTabPre = 5, kTabs = 6 };

// LDS of the bit-block kernels: guard band, row, guard band, chunk tables.
inline size_t sdem_bits_lds(int maxL, int group = 1) {
    return sizeof(float) * ((size_t)maxL + 2 * (group > 1 ? pad_bits<4>() : pad_bits<1>()) + kTabs * kTabN);
}

__device__ __forceinline__ int lshl1_add(int a, int b) {
    int r;
    asm("v_lshl_add_u32 %0, %1, 1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

__device__ __forceinline__ void flush_block(uint32_t w, int d0, int& A, int& s) {
    const int pc = __popc(w);
    // sp in Horner form, one v_lshl_add per level (the compiler would
    // otherwise expand it into separate shifts and adds)
    int sp = __popc(w & 0xFFFF0000u);
    sp = lshl1_add(sp, __popc(w & 0xFF00FF00u));
    sp = lshl1_add(sp, __popc(w & 0xF0F0F0F0u));
    sp = lshl1_add(sp, __popc(w & 0xCCCCCCCCu));
    sp = lshl1_add(sp, __popc(w & 0xAAAAAAAAu));
    A += (int)__umul24(2 * d0 + 65, pc) - 2 * sp;  // operands < 2^13
    // s = d0 + 33 - ctz(~w) when the block has an invisible step: such a value
    // exceeds every earlier s (<= d0 + 1).  v_ffbl gives ~0u for ~w == 0, which
    // the saturating subtraction turns into 0, so the max keeps s.
    uint32_t f;
    asm("v_ffbl_b32 %0, %1" : "=v"(f) : "v"(~w));
    s = max(s, (int)__builtin_elementwise_sub_sat((uint32_t)(d0 + 33), f));
}

// The visibility flag is the sign of t = M - q: rounding preserves the sign
// of the exact difference and M == q gives +0, so sign(t) = (q > M).
// alignbit(w, t, 31) = (w << 1) | (t >> 31) appends it in one operation.
// The 32 reciprocal pairs of a block are loaded up front (wave-uniform:
// scalar loads, one wait), so the steps themselves wait only on LDS.
template <int kDir, bool kTail>
__device__ __forceinline__ uint32_t scan_block(const float* rowj, float h, const float4* __restrict__ rtab, int d0,
                                               int myLen, float& M) {
    const float2* __restrict__ ytab = reinterpret_cast<const float2*>(rtab + kMaxDim + 1);
    float2 y[32];
#pragma unroll
    for (int u = 0; u < 32; ++u) y[u] = ytab[d0 + 1 + u];
    uint32_t w = 0;
#pragma unroll
    for (int u = 1; u <= 32; ++u) {
        const int d = d0 + u;
        const float a = rowj[kDir * d] - h;
        const float q = __builtin_fmaf(a, y[u - 1].x, a * y[u - 1].y);
        float t = M - q;
        if (kTail) t = (d > myLen) ? -1.0f : t;  // past this lane's row end: record visible
        M = fmaxf(M, q);
        w = __builtin_amdgcn_alignbit(w, __float_as_uint(t), 31);
    }
    return w;
}

// 16 steps d0 + u0 + 1 .. d0 + u0 + 16 of a block (no tail handling),
// appended to w: the groups' slots run a block as two halves, so only 16
// reciprocal pairs are live in scalar registers at a time.
template <int kDir>
__device__ __forceinline__ uint32_t scan_half(const float* rowj, float h, const float2* __restrict__ ytab, int d0,
                                              int u0, float& M, uint32_t w) {
#pragma unroll
    for (int u = 1; u <= 16; ++u) {
        const int d = d0 + u0 + u;
        const float2 y = ytab[d];
        const float a = rowj[kDir * d] - h;
        const float q = __builtin_fmaf(a, y.x, a * y.y);
        const float t = M - q;
        M = fmaxf(M, q);
        w = __builtin_amdgcn_alignbit(w, __float_as_uint(t), 31);
    }
    return w;
}

// Early exit.  After the steps up to db, no later target of a lane can be
// visible once an upper bound of its quotient is <= M.  The remaining
// targets (distances e = db+1 .. myLen) are split into four segments that
// start at distances e, e+32, e+96 and e+224 (the last runs to the row end);
// with a_i = (largest height of segment i) - h, every later a_k is <= the
// a_i of its segment by monotone rounding, so every later quotient is
//     <= max( max_i a_i / e_i ,  (max_i a_i) / L ).
// (The second term covers a_i < 0, where a farther target has the larger
// quotient; L exceeds every distance.  When some a_i >= 0 it is dominated by
// the first term.)  Segment maxima come from the chunk tables over a
// superset of each segment's cells: extra cells only raise the bound.  The
// products use RN(1/e) and are rounded once, so the bound is raised by
// 2^-20 of its magnitude to stay above the rounded quotients.  When every
// lane of the wave is past its row end or bounded by its running maximum
// M, the scan stops: a lane with steps left has only invisible steps ahead,
// so its trailing visible run closes and counts (cv = A); a lane whose row
// ended keeps the usual correction of the padded steps.
// kClamp: clamp the nearest-target position into the row (callers whose
// lanes may lie far past their row end); without it p must lie in
// [-32, L + 31], which the chunk tables' padding covers.
template <int kDir, bool kClamp = true>
__device__ __forceinline__ float exit_bound(const float* __restrict__ tab, int jj, int L, float h, int db,
                                            const float2* __restrict__ ytab, float invL) {
    const int e = db + 1;
    int p = jj + kDir * e;  // nearest remaining target
    if (kClamp) p = p < 0 ? 0 : (p > L - 1 ? L - 1 : p);
    const float* t = tab + kTabLo + (p >> 5);
    float s1, s2, s3, s4;
    if (kDir > 0) {  // cells [p, p+32), [p+32, p+96), [p+96, p+224), [p+224, end)
        s1 = t[kTabL1 * kTabN];
        s2 = t[kTabL2 * kTabN + 1];
        s3 = t[kTabL3 * kTabN + 3];
        s4 = t[kTabSuf * kTabN + 7];
    } else {         // cells (p-32, p], (p-96, p-32], (p-224, p-96], [0, p-224]
        s1 = t[kTabL1 * kTabN - 1];
        s2 = t[kTabL2 * kTabN - 4];
        s3 = t[kTabL3 * kTabN - 10];
        s4 = t[kTabPre * kTabN - 7];
    }
    const float a1 = s1 - h, a2 = s2 - h, a3 = s3 - h, a4 = s4 - h;
    const float amax = fmaxf(fmaxf(a1, a2), fmaxf(a3, a4));
    float b = fmaxf(fmaxf(a1 * ytab[e].x, a2 * ytab[e + 32].x), fmaxf(a3 * ytab[e + 96].x, a4 * ytab[e + 224].x));
    b = fmaxf(b, amax * invL);
    return __builtin_fmaf(fabsf(b), 9.5367431640625e-7f, b);  // b + |b| 2^-20
}

template <int kDir>
__device__ __forceinline__ bool scan_done(const float* __restrict__ tab, int jj, int L, float h, float invL,
                                          const float2* __restrict__ ytab, int myLen, int db, float M) {
    const bool done = (myLen <= db) | (exit_bound<kDir>(tab, jj, L, h, db, ytab, invL) <= M);
    // every lane of the wave is active here; a ballot keeps the exit uniform
    return __builtin_amdgcn_readfirstlane((int)(__builtin_amdgcn_ballot_w64(!done) != 0)) == 0;
}

template <int kDir>
__device__ __forceinline__ int scan_dir(const float* rowj, float h, const float4* __restrict__ rtab, int myLen,
                                        int dfull, int D, const float* __restrict__ tab, int jj, int L,
                                        int& nblk) {
    const float2* __restrict__ ytab = reinterpret_cast<const float2*>(rtab + kMaxDim + 1);
    const float invL = ytab[L].x;
    float M = -2000.0f;
    int A = 1, s = 0;
    int d0 = 0;
    int e1 = dfull, e2 = D;  // loop ends, lowered to the current step when the scan stops
    for (; d0 + 32 <= e1; d0 += 32) {
        flush_block(scan_block<kDir, false>(rowj, h, rtab, d0, myLen, M), d0, A, s);
        if (scan_done<kDir>(tab, jj, L, h, invL, ytab, myLen, d0 + 32, M)) e1 = e2 = d0 + 32;
    }
    for (; d0 < e2; d0 += 32) {
        flush_block(scan_block<kDir, true>(rowj, h, rtab, d0, myLen, M), d0, A, s);
        if (scan_done<kDir>(tab, jj, L, h, invL, ytab, myLen, d0 + 32, M)) e2 = d0 + 32;
    }
    nblk += d0 >> 5;
    return myLen > d0 ? A : A - ((d0 + 1) * (d0 + 1) - s * s);
}

// Build the chunk tables of a staged row of L <= 4096 cells (see kTabN).
// Ends with a barrier.
template <int kWaves>
__device__ __forceinline__ void row_tables(const float* row, int L, float* tab) {
    const int nch = (L + 31) >> 5;
    float* t0 = tab + kTabLo;
    for (int i = threadIdx.x; i < kTabs * kTabN; i += 64 * kWaves) tab[i] = -INFINITY;
    __syncthreads();
    for (int t0i = 0; t0i < nch * 32; t0i += 64 * kWaves) {
        const int t = t0i + (int)threadIdx.x;
        float v = t < L ? row[t] : -INFINITY;
#pragma unroll
        for (int o = 16; o; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        if ((t & 31) == 0 && t < nch * 32) t0[kTabL0 * kTabN + (t >> 5)] = v;
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // one wave: suffix and prefix maxima, two chunks per lane
        const int l = threadIdx.x;
        const float x0 = t0[kTabL0 * kTabN + 2 * l], x1 = t0[kTabL0 * kTabN + 2 * l + 1];  // -inf past nch
        float vs = fmaxf(x0, x1), vp = vs;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const float us = __shfl_down(vs, o), up = __shfl_up(vp, o);  // own value when out of range
            vs = fmaxf(vs, us);
            vp = fmaxf(vp, up);
        }
        float ns = __shfl_down(vs, 1), np = __shfl_up(vp, 1);
        if (l == 63) ns = -INFINITY;
        if (l == 0) np = -INFINITY;
        const float s1 = fmaxf(x1, ns), p0 = fmaxf(x0, np), p1 = fmaxf(x1, p0);
        t0[kTabSuf * kTabN + 2 * l] = fmaxf(x0, s1);
        t0[kTabSuf * kTabN + 2 * l + 1] = s1;
        t0[kTabPre * kTabN + 2 * l] = p0;
        t0[kTabPre * kTabN + 2 * l + 1] = p1;
        if (l < kTabHi) t0[kTabPre * kTabN + kChunksMax + l] = __shfl(p1, 63);  // whole-row maximum
    }
    for (int lev = 1; lev <= 3; ++lev) {  // level lev from level lev-1: max of two runs 2^(lev-1) apart
        const int off = 1 << (lev - 1);
        for (int i = threadIdx.x; i < kTabN - off; i += 64 * kWaves)
            tab[lev * kTabN + i] = fmaxf(tab[(lev - 1) * kTabN + i], tab[(lev - 1) * kTabN + i + off]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Lane compaction.  With the early exit a wave still runs until its longest
// scan stops, while most of its lanes are done.  A wave therefore takes a
// group of kG chunks (kG x 64 scans of one direction, all starting at d = 1)
// held in kG register "slots" and steps all of them with the same uniform d.
// After each 32-step block, lanes whose scan is over retire (write their cv)
// and, once the live scans fit in fewer slots, they are packed into the
// lowest slots in (slot, lane) order: the p-th live scan moves to slot p/64,
// lane p%64.  Per-lane state moves with ds_bpermute; each destination lane
// finds its source from the slots' live masks (prefix counts, then the n-th
// set bit), so the packing needs no LDS.  Slots 0..k-1 hold at most 64k live
// scans, so destination slot k2 draws only from slots >= k2 and the packing
// can run in place, lowest slot first.
// A lane's steps past its row end are recorded as visible by or-ing the low
// bits of its block word; retired and empty lanes read from a position whose
// remaining steps stay inside the staged row and guard bands.  The
// right-scan pass stores cv as an integer in the output cell; the left-scan
// pass of the same group (same wave, later) adds its cv and stores the float
// (float)(cvR + cvL) * R, the same value as the one-chunk path.
// ---------------------------------------------------------------------------
// Position of the r-th (0-based) set bit of m; m must have more than r set bits.
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r) {
    int pos = 0;
    uint32_t w = (uint32_t)m;
    const int c0 = __popc(w);
    if (r >= c0) {
        r -= c0;
        pos = 32;
        w = (uint32_t)(m >> 32);
    }
    int c = __popc(w & 0xFFFFu);
    if (r >= c) { r -= c; pos += 16; w >>= 16; }
    c = __popc(w & 0xFFu);
    if (r >= c) { r -= c; pos += 8; w >>= 8; }
    c = __popc(w & 0xFu);
    if (r >= c) { r -= c; pos += 4; w >>= 4; }
    c = __popc(w & 0x3u);
    if (r >= c) { r -= c; pos += 2; w >>= 2; }
    if (r >= (int)(w & 1u)) pos += 1;
    return pos;
}

template <int kDir, int kG>
__device__ __forceinline__ void scan_group(const float* row, int L, float povh, int c0, int nck,
                                           const float* __restrict__ tab, const float4* __restrict__ rtab,
                                           float* __restrict__ so, float R, int& nblk) {
    const float2* __restrict__ ytab = reinterpret_cast<const float2*>(rtab + kMaxDim + 1);
    int* __restrict__ soi = reinterpret_cast<int*>(so);
    const float invL = ytab[L].x;
    const int lane = threadIdx.x & 63;
    const int safe = kDir > 0 ? 0 : L - 1;  // reads row[safe + kDir*d], d <= L + 31: inside the guard bands
    int jj[kG], A[kG], s[kG];
    float h[kG], M[kG];
    uint64_t act[kG];  // live scans of each slot (wave-uniform masks)
#pragma unroll
    for (int k = 0; k < kG; ++k) {
        const int j = 64 * (c0 + k) + lane;
        const int myLen = kDir > 0 ? L - 1 - j : j;
        const bool in = k < nck && j < L;
        if (in && myLen <= 0) {  // no targets in this direction: cv = 0
            if (kDir > 0)
                soi[j] = 0;
            else
                so[j] = (float)soi[j] * R;
        }
        const bool a = in && myLen > 0;
        act[k] = __builtin_amdgcn_ballot_w64(a);
        jj[k] = a ? j : safe;
        h[k] = row[jj[k]] + povh;
        M[k] = -2000.0f;
        A[k] = 1;
        s[k] = 0;
    }
    int ns = nck;
    for (int d0 = 0; ns > 0; d0 += 32) {
        bool ends = false;  // some lane's row ends in this block (uniform)
        uint32_t wk[kG];
#pragma unroll
        for (int k = 0; k < kG; ++k)
            if (k < ns) wk[k] = scan_half<kDir>(row + jj[k], h[k], ytab, d0, 0, M[k], 0u);
#pragma unroll
        for (int k = 0; k < kG; ++k)
            if (k < ns) wk[k] = scan_half<kDir>(row + jj[k], h[k], ytab, d0, 16, M[k], wk[k]);
#pragma unroll
        for (int k = 0; k < kG; ++k) {
            if (k < ns) {
                uint32_t w = wk[k];
                const int myLen = kDir > 0 ? L - 1 - jj[k] : jj[k];
                // steps past the row end (the last e of the block) are recorded as
                // visible; skipped when no lane's row ends in this block (a row
                // ending at the block's last step adds no steps: e = 0)
                if (__builtin_amdgcn_ballot_w64(myLen <= d0 + 32) != 0) {
                    const int e = min(max(d0 + 32 - myLen, 0), 31);
                    w |= (1u << e) - 1u;
                    ends = true;
                }
                flush_block(w, d0, A[k], s[k]);
            }
        }
        const int db = d0 + 32;
        nblk += ns;
        // The exit bound is tested after every other block (blocks 0, 2, 4,
        // ...), lanes past their row end retire after every block.  A scan
        // whose bound holds only runs further invisible steps, which leave A
        // as it is, so its cv does not change; the bound costs about a
        // seventh of a block's instructions.
        const bool tb = ((d0 >> 5) & 1) == 0;
        if (!tb && !ends) continue;  // no lane can retire: slots and masks stay as they are
        int T = 0;
#pragma unroll
        for (int k = 0; k < kG; ++k) {
            if (k < ns) {
                const bool a = (act[k] >> lane) & 1;
                const int myLen = kDir > 0 ? L - 1 - jj[k] : jj[k];
                bool done = myLen <= db;
                // active lanes and empty lanes (jj = safe) keep the bound's table reads in range
                if (tb) done = done | (exit_bound<kDir, false>(tab, jj[k], L, h[k], db, ytab, invL) <= M[k]);
                if (a && done) {
                    const int cv = myLen > db ? A[k] : A[k] - ((db + 1) * (db + 1) - s[k] * s[k]);
                    if (kDir > 0)
                        soi[jj[k]] = cv;
                    else
                        so[jj[k]] = (float)(cv + soi[jj[k]]) * R;
                }
                const bool live = a && !done;
                act[k] = __builtin_amdgcn_ballot_w64(live);
                jj[k] = live ? jj[k] : safe;
                T += (int)__popcll(act[k]);
            }
        }
        T = __builtin_amdgcn_readfirstlane(T);
        const int nn = (T + 63) >> 6;
        if (nn < ns) {
            // live scans before each slot (wave-uniform), from the masks as
            // they were before any slot is rewritten
            uint64_t am[kG];
            int pre[kG];
            int acc = 0;
#pragma unroll
            for (int k = 0; k < kG; ++k) {
                am[k] = k < ns ? act[k] : 0ull;
                pre[k] = acc;
                acc += (int)__popcll(am[k]);
            }
#pragma unroll
            for (int k2 = 0; k2 < kG; ++k2) {
                if (k2 < nn) {
                    // the p-th live scan lies in the last slot k with pre[k] <= p
                    // (k >= k2, since p >= 64 k2 >= pre[k2]) at its (p - pre[k])-th live lane
                    const int p = 64 * k2 + lane;
                    uint64_t m = am[k2];
                    int base = pre[k2], sk = k2;
#pragma unroll
                    for (int k = k2 + 1; k < kG; ++k) {
                        if (k < ns && p >= pre[k]) {
                            m = am[k];
                            base = pre[k];
                            sk = k;
                        }
                    }
                    if (p >= T) sk = -1;  // lane left empty
                    const int sa = nth_set_bit(m, p >= T ? 0 : p - base) << 2;  // ds_bpermute byte address
                    int nj = safe, nA = 1, nS = 0;
                    float nh = 0.0f, nM = -2000.0f;
#pragma unroll
                    for (int k = k2; k < kG; ++k) {
                        if (k < ns) {
                            const int tj = __builtin_amdgcn_ds_bpermute(sa, jj[k]);
                            const int tA = __builtin_amdgcn_ds_bpermute(sa, A[k]);
                            const int tS = __builtin_amdgcn_ds_bpermute(sa, s[k]);
                            const int th = __builtin_amdgcn_ds_bpermute(sa, __float_as_int(h[k]));
                            const int tM = __builtin_amdgcn_ds_bpermute(sa, __float_as_int(M[k]));
                            if (sk == k) {
                                nj = tj;
                                nA = tA;
                                nS = tS;
                                nh = __int_as_float(th);
                                nM = __int_as_float(tM);
                            }
                        }
                    }
                    jj[k2] = nj;
                    A[k2] = nA;
                    s[k2] = nS;
                    h[k2] = nh;
                    M[k2] = nM;
                    act[k2] = __builtin_amdgcn_ballot_w64(p < T);
                }
            }
            ns = nn;
        }
    }
}

// One skewed row: stage it, then every wave of the workgroup takes chunks
// of 64 observers (kG = 1), or groups of kG chunks with lane compaction,
// handed out by an LDS counter.  Ends with a barrier so the caller may
// restage LDS.
template <int kWaves, int kG>
__device__ __forceinline__ void sdem_bits_row(const RowWork w, const SecDesc* __restrict__ secs,
                                              const float* __restrict__ in4, const int* __restrict__ target,
                                              const float* __restrict__ weight, const uint32_t* __restrict__ fl,
                                              const float* __restrict__ Sbuf, const float4* __restrict__ rtab,
                                              float povh, float* __restrict__ arena, float* row,
                                              unsigned long long* __restrict__ steps, int* gnext) {
    const SecDesc sd = secs[w.sec];
    const int r = w.row;
    const uint32_t v = fl[sd.rows + r];
    const int f = fl_first(v);
    const int L = fl_last(v) - f;
    if (L <= 0) return;  // uniform; the work list holds only rows with L > 0
    stage_row(row, sd, r, f, L, in4, target, weight, Sbuf, pad_bits<kG>());
    if (kG > 1 && threadIdx.x == 0) *gnext = kWaves;
    __syncthreads();
    float* tab = row + L + pad_bits<kG>();  // past the right guard band
    row_tables<kWaves>(row, L, tab);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nchunk = (L + 63) >> 6;
    float* so = arena + sd.so + (size_t)r * sd.dim_i + f;
    int nblk = 0;  // 32-step blocks this wave ran (both directions)
    if (kG > 1) {
        const int ngroups = (nchunk + kG - 1) / kG;
        for (int g = wave; g < ngroups;) {
            const int c0 = g * kG;
            const int nck = min(kG, nchunk - c0);
            scan_group<+1, kG>(row, L, povh, c0, nck, tab, rtab, so, sd.R, nblk);
            scan_group<-1, kG>(row, L, povh, c0, nck, tab, rtab, so, sd.R, nblk);
            int ng = 0;
            if (lane == 0) ng = atomicAdd(gnext, 1);
            g = __builtin_amdgcn_readfirstlane(ng);
        }
    }
    for (int c = wave; kG == 1 && c < nchunk; c += kWaves) {
        const int j0 = c << 6;
        const int jj = j0 + lane;
        const bool valid = jj < L;
        const bool full = j0 + 63 < L;  // all 64 observers of the chunk exist
        const float h = row[jj] + povh;
        // right scan: k = j+1 .. last-1
        const int dmaxR = L - 1 - j0;
        const int cvR = scan_dir<+1>(row + jj, h, rtab, valid ? L - 1 - jj : 0, full ? L - 1 - (j0 + 63) : 0,
                                     (dmaxR + 31) & ~31, tab, jj, L, nblk);
        // left scan: k = j-1 .. first
        const int dmaxL = full ? j0 + 63 : L - 1;
        const int cvL = scan_dir<-1>(row + jj, h, rtab, valid ? jj : 0, full ? j0 : 0, (dmaxL + 31) & ~31, tab,
                                     jj, L, nblk);
        if (valid) so[jj] = (float)(cvR + cvL) * sd.R;
    }
    if (steps && lane == 0 && nblk) atomicAdd(steps, 2048ull * (unsigned long long)nblk);  // 32 steps x 64 lanes
    __syncthreads();
}

// One workgroup per row of the work list.
template <int kWaves, int kG>
__global__ __launch_bounds__(64 * kWaves) void k_op_sdem_bits(const SecDesc* __restrict__ secs,
                                                              const RowWork* __restrict__ work,
                                                              const float* __restrict__ in4,
                                                              const int* __restrict__ target,
                                                              const float* __restrict__ weight,
                                                              const uint32_t* __restrict__ fl,
                                                              const float* __restrict__ Sbuf,
                                                              const float4* __restrict__ rtab, float povh,
                                                              float* __restrict__ arena,
                                                              unsigned long long* __restrict__ steps) {
    extern __shared__ float lds[];
    __shared__ int gnext;
    sdem_bits_row<kWaves, kG>(work[blockIdx.x], secs, in4, target, weight, fl, Sbuf, rtab, povh, arena,
                              lds + pad_bits<kG>(), steps, &gnext);
}

// Persistent: a fixed grid pulls rows off the (cost-sorted) work list with
// an atomic counter until the list is exhausted; every workgroup exits once
// the counter passes nwork.  The counter is cleared by the launcher.
// kOcc > 1 asks the compiler for at least kOcc waves per SIMD (fewer VGPRs,
// less software pipelining of the LDS loads, more waves to hide latency).
template <int kWaves, int kOcc, int kG>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(kOcc, 8))) void k_op_sdem_bits_persistent(
    const SecDesc* __restrict__ secs, const RowWork* __restrict__ work, int nwork, int* __restrict__ counter,
    const float* __restrict__ in4, const int* __restrict__ target, const float* __restrict__ weight,
    const uint32_t* __restrict__ fl, const float* __restrict__ Sbuf, const float4* __restrict__ rtab, float povh,
    float* __restrict__ arena, unsigned long long* __restrict__ steps) {
    extern __shared__ float lds[];
    __shared__ int item, gnext;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(item);  // the same value in every lane
        __syncthreads();  // every lane has read `item` before thread 0 overwrites it
        if (it >= nwork) break;
        sdem_bits_row<kWaves, kG>(work[it], secs, in4, target, weight, fl, Sbuf, rtab, povh, arena,
                                  lds + pad_bits<kG>(), steps, &gnext);
    }
}

__global__ void k_zero_int(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_op_sdem_tuned(const DevPlan& p, const float* in4, const float* Sbuf, float* arena,
                                const float4* rtab, float povh, const SdemTuning& t, int* counter,
                                unsigned long long* steps, hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
    if (p.maxL > 4096) return launch_op_sdem(p, in4, Sbuf, arena, rtab, povh, 2, s);
    const int G = t.group > 1 ? 4 : 1;
    const size_t lds = sdem_bits_lds(p.maxL, G);
    if (t.persistent_wgs_per_cu > 0 && counter) {
        int dev = 0, ncu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, counter, 1);
        const int grid = std::min(p.nwork, ncu * t.persistent_wgs_per_cu);
#define SKE_PERS(W, O, G)                                                                                        \
    hipLaunchKernelGGL((k_op_sdem_bits_persistent<W, O, G>), dim3(grid), dim3(64 * W), lds, s, p.secs, p.work,    \
                       p.nwork, counter, in4, p.target, p.weight, p.fl, Sbuf, rtab, povh, arena, steps)
#define SKE_PERS_W(O, G)                                                                                         \
    if (t.waves == 2) SKE_PERS(2, O, G); else if (t.waves == 8) SKE_PERS(8, O, G); else SKE_PERS(4, O, G)
#define SKE_PERS_G(O)                                                                                            \
    if (G == 4) { SKE_PERS_W(O, 4); } else { SKE_PERS_W(O, 1); }
        if (t.occupancy >= 8) {
            SKE_PERS_G(8);
        } else if (t.occupancy >= 6) {
            SKE_PERS_G(6);
        } else {
            SKE_PERS_G(1);
        }
#undef SKE_PERS_G
#undef SKE_PERS_W
#undef SKE_PERS
        return hipGetLastError();
    }
#define SKE_GRID(W, G)                                                                                            \
    hipLaunchKernelGGL((k_op_sdem_bits<W, G>), dim3(p.nwork), dim3(64 * W), lds, s, p.secs, p.work, in4, p.target, \
                       p.weight, p.fl, Sbuf, rtab, povh, arena, steps)
    if (G == 4) {
        if (t.waves == 2) SKE_GRID(2, 4); else if (t.waves == 8) SKE_GRID(8, 4); else SKE_GRID(4, 4);
    } else {
        if (t.waves == 2) SKE_GRID(2, 1); else if (t.waves == 8) SKE_GRID(8, 1); else SKE_GRID(4, 1);
    }
#undef SKE_GRID
    return hipGetLastError();
}

hipError_t launch_op_sdem(const DevPlan& p, const float* in4, const float* Sbuf, float* arena,
                          const float4* rtab, float povh, int variant, hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
    if (variant == 0 && p.maxL <= 4096) {
        const size_t lds = sdem_bits_lds(p.maxL);
        hipLaunchKernelGGL((k_op_sdem_bits<4, 1>), dim3(p.nwork), dim3(256), lds, s, p.secs, p.work, in4, p.target,
                           p.weight, p.fl, Sbuf, rtab, povh, arena, nullptr);
        return hipGetLastError();
    }
    const size_t lds = sizeof(float) * (size_t)(p.maxL + 2 * kPad);
    if (variant == 1)
        hipLaunchKernelGGL(k_op_sdem<1>, dim3(p.nwork), dim3(256), lds, s, p.secs, p.work, in4, p.target,
                           p.weight, p.fl, Sbuf, rtab, povh, arena);
    else
        hipLaunchKernelGGL(k_op_sdem<0>, dim3(p.nwork), dim3(256), lds, s, p.secs, p.work, in4, p.target,
                           p.weight, p.fl, Sbuf, rtab, povh, arena);
    return hipGetLastError();
}

}  // namespace ske
