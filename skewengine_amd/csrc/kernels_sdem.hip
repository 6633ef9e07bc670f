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
constexpr int kPadBits = 128;
constexpr int kChunksMax = 128;  // 32-cell chunks of a row of <= 4096 cells

// LDS of the bit-block kernels: guard band, row, guard band, chunk maxima.
inline size_t sdem_bits_lds(int maxL) { return sizeof(float) * ((size_t)maxL + 2 * kPadBits + 2 * kChunksMax); }

__device__ __forceinline__ void flush_block(uint32_t w, int d0, int& A, int& s) {
    const int pc = __popc(w);
    const int sp = __popc(w & 0xAAAAAAAAu) + 2 * __popc(w & 0xCCCCCCCCu) + 4 * __popc(w & 0xF0F0F0F0u) +
                   8 * __popc(w & 0xFF00FF00u) + 16 * __popc(w & 0xFFFF0000u);
    A += (2 * d0 + 65) * pc - 2 * sp;
    if (w != 0xFFFFFFFFu) s = d0 + 33 - __builtin_ctz(~w);
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

// Early exit.  After the steps up to db, no later target k of this lane can
// be visible once an upper bound of its quotient is <= M: with a = Smax - h
// (Smax the largest height among the remaining targets, taken from the
// per-32-cell maxima `cmax`), every later a_k <= a by monotone rounding, so
// q_k <= RN(a / (db+1)) for a >= 0 and q_k <= RN(a / myLen) for a < 0.  The
// bound is evaluated with one-sided slack (2^-20 relative), which keeps it
// above those rounded quotients.  When every lane of the wave is past its row
// end or bounded, the scan stops: a lane with steps left has only invisible
// steps ahead, so its trailing visible run closes and counts (cv = A); a
// lane whose row ended keeps the usual correction of the padded steps.
template <int kDir>
__device__ __forceinline__ bool scan_done(const float* __restrict__ cmax, int jj, int L, float h, float rl,
                                          const float2* __restrict__ ytab, int myLen, int db, float M) {
    int pos = jj + kDir * (db + 1);
    pos = pos < 0 ? 0 : (pos > L - 1 ? L - 1 : pos);  // lanes with myLen <= db are done anyway
    const float a = cmax[pos >> 5] - h;
    const float bpos = (a * ytab[db + 1].x) * 1.00000095367431640625f;
    const float bneg = (a * rl) * 0.99999904632568359375f;
    const bool done = (myLen <= db) | ((a >= 0.0f ? bpos : bneg) <= M);
    // every lane of the wave is active here; a ballot keeps the exit uniform
    return __builtin_amdgcn_readfirstlane((int)(__builtin_amdgcn_ballot_w64(!done) != 0)) == 0;
}

template <int kDir>
__device__ __forceinline__ int scan_dir(const float* rowj, float h, const float4* __restrict__ rtab, int myLen,
                                        int dfull, int D, const float* __restrict__ cmax, int jj, int L,
                                        int& nblk) {
    const float2* __restrict__ ytab = reinterpret_cast<const float2*>(rtab + kMaxDim + 1);
    const float rl = 1.0f / (float)(myLen > 0 ? myLen : 1);
    float M = -2000.0f;
    int A = 1, s = 0;
    int d0 = 0;
    int e1 = dfull, e2 = D;  // loop ends, lowered to the current step when the scan stops
    for (; d0 + 32 <= e1; d0 += 32) {
        flush_block(scan_block<kDir, false>(rowj, h, rtab, d0, myLen, M), d0, A, s);
        if (scan_done<kDir>(cmax, jj, L, h, rl, ytab, myLen, d0 + 32, M)) e1 = e2 = d0 + 32;
    }
    for (; d0 < e2; d0 += 32) {
        flush_block(scan_block<kDir, true>(rowj, h, rtab, d0, myLen, M), d0, A, s);
        if (scan_done<kDir>(cmax, jj, L, h, rl, ytab, myLen, d0 + 32, M)) e2 = d0 + 32;
    }
    nblk += d0 >> 5;
    return myLen > d0 ? A : A - ((d0 + 1) * (d0 + 1) - s * s);
}

// Per-32-cell maxima of the staged row (cm[c] = max row[32c .. 32c+31]),
// then their suffix maxima (suf) and prefix maxima (pre), for the early exit
// of the right and left scans.  Rows of at most 4096 cells: <= 128 chunks.
template <int kWaves>
__device__ __forceinline__ void row_chunk_max(const float* row, int L, float* suf, float* pre) {
    const int nch = (L + 31) >> 5;
    for (int t0 = 0; t0 < nch * 32; t0 += 64 * kWaves) {
        const int t = t0 + (int)threadIdx.x;
        float v = t < L ? row[t] : -INFINITY;
#pragma unroll
        for (int o = 16; o; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        if ((t & 31) == 0 && t < nch * 32) suf[t >> 5] = v;
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // one wave: two chunks per lane
        const int l = threadIdx.x;
        const float x0 = 2 * l < nch ? suf[2 * l] : -INFINITY;
        const float x1 = 2 * l + 1 < nch ? suf[2 * l + 1] : -INFINITY;
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
        const float s1 = fmaxf(x1, ns), p0 = fmaxf(x0, np);
        if (2 * l < nch) {
            suf[2 * l] = fmaxf(x0, s1);
            pre[2 * l] = p0;
        }
        if (2 * l + 1 < nch) {
            suf[2 * l + 1] = s1;
            pre[2 * l + 1] = fmaxf(x1, p0);
        }
    }
}

// One skewed row: stage it, then every wave of the workgroup takes chunks
// of 64 observers.  Ends with a barrier so the caller may restage LDS.
template <int kWaves>
__device__ __forceinline__ void sdem_bits_row(const RowWork w, const SecDesc* __restrict__ secs,
                                              const float* __restrict__ in4, const int* __restrict__ target,
                                              const float* __restrict__ weight, const uint32_t* __restrict__ fl,
                                              const float* __restrict__ Sbuf, const float4* __restrict__ rtab,
                                              float povh, float* __restrict__ arena, float* row,
                                              unsigned long long* __restrict__ steps) {
    const SecDesc sd = secs[w.sec];
    const int r = w.row;
    const uint32_t v = fl[sd.rows + r];
    const int f = fl_first(v);
    const int L = fl_last(v) - f;
    if (L <= 0) return;  // uniform; the work list holds only rows with L > 0
    stage_row(row, sd, r, f, L, in4, target, weight, Sbuf, kPadBits);
    __syncthreads();
    float* suf = row + L + kPadBits;  // past the right guard band
    float* pre = suf + kChunksMax;
    row_chunk_max<kWaves>(row, L, suf, pre);
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nchunk = (L + 63) >> 6;
    float* so = arena + sd.so + (size_t)r * sd.dim_i + f;
    int nblk = 0;  // 32-step blocks this wave ran (both directions)
    for (int c = wave; c < nchunk; c += kWaves) {
        const int j0 = c << 6;
        const int jj = j0 + lane;
        const bool valid = jj < L;
        const bool full = j0 + 63 < L;  // all 64 observers of the chunk exist
        const float h = row[jj] + povh;
        // right scan: k = j+1 .. last-1
        const int dmaxR = L - 1 - j0;
        const int cvR = scan_dir<+1>(row + jj, h, rtab, valid ? L - 1 - jj : 0, full ? L - 1 - (j0 + 63) : 0,
                                     (dmaxR + 31) & ~31, suf, jj, L, nblk);
        // left scan: k = j-1 .. first
        const int dmaxL = full ? j0 + 63 : L - 1;
        const int cvL = scan_dir<-1>(row + jj, h, rtab, valid ? jj : 0, full ? j0 : 0, (dmaxL + 31) & ~31, pre,
                                     jj, L, nblk);
        if (valid) so[jj] = (float)(cvR + cvL) * sd.R;
    }
    if (steps && lane == 0 && nblk) atomicAdd(steps, 2048ull * (unsigned long long)nblk);  // 32 steps x 64 lanes
    __syncthreads();
}

// One workgroup per row of the work list.
template <int kWaves>
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
    sdem_bits_row<kWaves>(work[blockIdx.x], secs, in4, target, weight, fl, Sbuf, rtab, povh, arena,
                          lds + kPadBits, steps);
}

// Persistent: a fixed grid pulls rows off the (cost-sorted) work list with
// an atomic counter until the list is exhausted; every workgroup exits once
// the counter passes nwork.  The counter is cleared by the launcher.
// kOcc > 1 asks the compiler for at least kOcc waves per SIMD (fewer VGPRs,
// less software pipelining of the LDS loads, more waves to hide latency).
template <int kWaves, int kOcc>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(kOcc, 8))) void k_op_sdem_bits_persistent(
    const SecDesc* __restrict__ secs, const RowWork* __restrict__ work, int nwork, int* __restrict__ counter,
    const float* __restrict__ in4, const int* __restrict__ target, const float* __restrict__ weight,
    const uint32_t* __restrict__ fl, const float* __restrict__ Sbuf, const float4* __restrict__ rtab, float povh,
    float* __restrict__ arena, unsigned long long* __restrict__ steps) {
    extern __shared__ float lds[];
    __shared__ int item;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(item);  // the same value in every lane
        __syncthreads();  // every lane has read `item` before thread 0 overwrites it
        if (it >= nwork) break;
        sdem_bits_row<kWaves>(work[it], secs, in4, target, weight, fl, Sbuf, rtab, povh, arena, lds + kPadBits,
                              steps);
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
    const size_t lds = sdem_bits_lds(p.maxL);
    if (t.persistent_wgs_per_cu > 0 && counter) {
        int dev = 0, ncu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, counter, 1);
        const int grid = std::min(p.nwork, ncu * t.persistent_wgs_per_cu);
#define SKE_PERS(W, O)                                                                                           \
    hipLaunchKernelGGL((k_op_sdem_bits_persistent<W, O>), dim3(grid), dim3(64 * W), lds, s, p.secs, p.work,       \
                       p.nwork, counter, in4, p.target, p.weight, p.fl, Sbuf, rtab, povh, arena, steps)
        if (t.occupancy >= 8) {
            if (t.waves == 2) SKE_PERS(2, 8); else if (t.waves == 8) SKE_PERS(8, 8); else SKE_PERS(4, 8);
        } else {
            if (t.waves == 2) SKE_PERS(2, 1); else if (t.waves == 8) SKE_PERS(8, 1); else SKE_PERS(4, 1);
        }
#undef SKE_PERS
        return hipGetLastError();
    }
#define SKE_GRID(W)                                                                                             \
    hipLaunchKernelGGL(k_op_sdem_bits<W>, dim3(p.nwork), dim3(64 * W), lds, s, p.secs, p.work, in4, p.target, \
                       p.weight, p.fl, Sbuf, rtab, povh, arena, steps)
    if (t.waves == 2) SKE_GRID(2); else if (t.waves == 8) SKE_GRID(8); else SKE_GRID(4);
#undef SKE_GRID
    return hipGetLastError();
}

hipError_t launch_op_sdem(const DevPlan& p, const float* in4, const float* Sbuf, float* arena,
                          const float4* rtab, float povh, int variant, hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
    if (variant == 0 && p.maxL <= 4096) {
        const size_t lds = sdem_bits_lds(p.maxL);
        hipLaunchKernelGGL(k_op_sdem_bits<4>, dim3(p.nwork), dim3(256), lds, s, p.secs, p.work, in4, p.target,
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
