// kernels_radn.hip -- Radon operator (src/kernelRADN.h:212-258) on gfx950:
// the global transform (LRT=false) and the local one (LRT=true, pointRadon
// :114-182).
//
//   rowsum : one wave per skewed row, P_r = sum_j S(r,j) in double (the
//            reference sums in double too, kernelRADN.h:227-232).
//   filter : one workgroup per sector.  P is zero-padded to H2 = 2^ceil(log2 H)
//            in LDS and filtered with the reference's own transform: the
//            radix-2 decimation-in-frequency FFT of src/helper.h:59-106 with
//            its twiddles (repeated complex products, precomputed on the host
//            by the same recurrence), the one-sided ramp k/H for k <= H/2
//            (kernelRADN.h:247), and ifft = conj o fft o conj / H2
//            (helper.h:108-114).  Butterflies of one stage are independent,
//            so the parallel stage computes exactly the reference's values.
//   The smear back (kernelRADN.h:252-257) is folded into the deskew, which
//   reads q_r for every cell of row r in [first, last).
#include <limits.h>

#include "device_common.h"
#include "launch.h"

namespace ske {

__global__ __launch_bounds__(256) void k_radn_rowsum(const SecDesc* __restrict__ secs,
                                                     const RowWork* __restrict__ work, int nwork,
                                                     const float* __restrict__ in4,
                                                     const int* __restrict__ target,
                                                     const float* __restrict__ weight,
                                                     const uint32_t* __restrict__ fl,
                                                     const float* __restrict__ Sbuf, double* __restrict__ P) {
    const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= nwork) return;
    const int lane = threadIdx.x & 63;
    const RowWork w = work[item];
    const SecDesc sd = secs[w.sec];
    const uint32_t v = fl[sd.rows + w.row];
    const int f = fl_first(v), l = fl_last(v);
    const float* src = in4 + sd.img + (size_t)sd.type * sd.N;
    double acc = 0.0;
    int j = f + lane;
    if (!Sbuf) {
        // four cells per lane and pass: table reads, then gathers, in flight
        // together; the lane's sum still runs over j in ascending order
        for (; j + 192 < l; j += 256) {
            int t[4];
            float wt[4], x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                t[u] = target[sd.tab + j + 64 * u];
                wt[u] = weight[sd.tab + j + 64 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                x[u] = skew_cell_sel(src, t[u], wt[u], w.row, j + 64 * u, sd.dim_i, sd.dim_o);
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += (double)x[u];
        }
    }
    for (; j < l; j += 64) {
        float x;
        if (Sbuf)
            x = Sbuf[(size_t)w.row * sd.dim_i + j];
        else
            x = skew_cell(src, target[sd.tab + j], weight[sd.tab + j], w.row, j, sd.dim_i, sd.dim_o);
        acc += (double)x;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) P[sd.q + w.row] = acc;
}

// The fused row sums (no materialised skewInput): one wave per 64
// consecutive skewed rows of one sector, lane = row.  Each lane adds
// S(r, j) over its own [first, last) in ascending j in double -- the
// reference's own order (kernelRADN.h:227-232), so P is the reference's sum
// bit for bit -- with S(r, j) computed as skew_cell does.  The two taps of
// cell (r, j) lie in column j of the oriented copy, rows r - t_j - 1 and
// r - t_j; a column of one oriented copy is a row of another (the partner,
// read backwards for types 0 and 3):
//   type 0: in0[a][j] = in2[j][dimy-1-a]        type 1: in1[a][j] = in3[dimy-1-j][a]
//   type 2: in2[a][j] = in0[dimy-1-j][a]        type 3: in3[a][j] = in2[dimx-1-j][dimy-1-a]
// so at every j the 64 lanes' taps are 64 consecutive floats of one row and
// t_j, w_j are wave-uniform (scalar loads).  The byte offset of the lane's
// lower-addressed tap is A + B j + C t_j + D: three sector constants and a
// per-lane term.  The wave walks the union of its rows' ranges; cells
// outside a lane's own range are read but add zero (acc + 0 = acc: acc is
// never -0).
constexpr int kColsUnroll = 8;
// Columns [j0, j1) of the walk, adding to acc.  kMask: test the lane's own
// [f, l) (columns where some lane of the wave is outside its range).  The
// first tap's value is w * x1 without the reference's 0 + (it differs only
// in the sign of a zero, which no later sum can see: acc is never -0).
template <bool kRev, bool kMask>
__device__ __forceinline__ void rowsum_cols_span(const int* __restrict__ tt, const float* __restrict__ ww,
                                                 __amdgpu_buffer_rsrc_t src, int A, int B, int C, int D, int r,
                                                 int dim_o, int f, int l, int j0, int j1, double& acc) {
    auto cell = [&](int j, int t, float w, uint2 pr) -> float {
        // kRev: the tap of row r - t at the lower address
        const float x1 = __uint_as_float(kRev ? pr.y : pr.x), x0 = __uint_as_float(kRev ? pr.x : pr.y);
        const int i0 = r - t;
        const float s = ((unsigned)(i0 - 1) < (unsigned)dim_o) ? w * x1 : 0.0f;
        const float v = ((unsigned)i0 < (unsigned)dim_o) ? s + (1.0f - w) * x0 : s;
        return (!kMask || (unsigned)(j - f) < (unsigned)(l - f)) ? v : 0.0f;
    };
    auto load = [&](int j, int t) -> uint2 {
        const int o = A + B * j + C * t + D;
        return make_uint2(__builtin_amdgcn_raw_buffer_load_b32(src, o, 0, 0),
                          __builtin_amdgcn_raw_buffer_load_b32(src, o + 4, 0, 0));
    };
    int j = j0;
    for (; j + kColsUnroll <= j1; j += kColsUnroll) {
        int t[kColsUnroll];
        float w[kColsUnroll];
        uint2 pr[kColsUnroll];
#pragma unroll
        for (int u = 0; u < kColsUnroll; ++u) {
            t[u] = tt[j + u];
            w[u] = ww[j + u];
        }
#pragma unroll
        for (int u = 0; u < kColsUnroll; ++u) pr[u] = load(j + u, t[u]);
#pragma unroll
        for (int u = 0; u < kColsUnroll; ++u) acc += (double)cell(j + u, t[u], w[u], pr[u]);
    }
    for (; j < j1; ++j) {
        const int t = tt[j];
        acc += (double)cell(j, t, ww[j], load(j, t));
    }
}

// The walk over [lo, hi): columns [fmax, lmin), inside every lane's range,
// without the range test.
template <bool kRev>
__device__ __forceinline__ double rowsum_cols_walk(const int* __restrict__ tt, const float* __restrict__ ww,
                                                   __amdgpu_buffer_rsrc_t src, int A, int B, int C, int D, int r,
                                                   int dim_o, int f, int l, int lo, int hi, int fmax, int lmin) {
    double acc = 0.0;
    if (fmax < lmin) {
        rowsum_cols_span<kRev, true>(tt, ww, src, A, B, C, D, r, dim_o, f, l, lo, fmax, acc);
        rowsum_cols_span<kRev, false>(tt, ww, src, A, B, C, D, r, dim_o, f, l, fmax, lmin, acc);
        rowsum_cols_span<kRev, true>(tt, ww, src, A, B, C, D, r, dim_o, f, l, lmin, hi, acc);
    } else {
        rowsum_cols_span<kRev, true>(tt, ww, src, A, B, C, D, r, dim_o, f, l, lo, hi, acc);
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_radn_rowsum_cols(const SecDesc* __restrict__ secs,
                                                          const float* __restrict__ in4,
                                                          const int* __restrict__ target,
                                                          const float* __restrict__ weight,
                                                          const uint32_t* __restrict__ fl, double* __restrict__ P) {
    const SecDesc sd = secs[blockIdx.y];
    const int r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    if (r0 >= sd.H) return;  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int r = r0 + lane;
    const bool live = r < sd.H;
    int f = 0, l = 0;
    if (live) {
        const uint32_t v = fl[sd.rows + r];
        f = fl_first(v);
        l = fl_last(v);
    }
    if (!(live && l > f)) f = l = 0;  // no cell of this lane is added
    int lo = l > f ? f : INT_MAX, hi = l > f ? l : INT_MIN;
    // [fmax, lmin): inside every lane's range (empty when some lane has none)
    int fmax = l > f ? f : INT_MAX, lmin = l > f ? l : INT_MIN;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        lo = min(lo, __shfl_xor(lo, off, 64));
        hi = max(hi, __shfl_xor(hi, off, 64));
        fmax = max(fmax, __shfl_xor(fmax, off, 64));
        lmin = min(lmin, __shfl_xor(lmin, off, 64));
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    hi = __builtin_amdgcn_readfirstlane(hi);
    fmax = __builtin_amdgcn_readfirstlane(fmax);
    lmin = __builtin_amdgcn_readfirstlane(lmin);
    const int dim_o = sd.dim_o, dim_i = sd.dim_i, type = sd.type;
    const int partner = type == 0 ? 2 : type == 1 ? 3 : type == 2 ? 0 : 2;
    // the partner copy as a raw buffer of N floats: a tap outside it (a row
    // index outside the image, whose cell is masked anyway) reads zero
    const __amdgpu_buffer_rsrc_t src = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(in4 + sd.img + (size_t)partner * sd.N), 0, sd.N * 4, kBufferRsrcWord3);
    const int* const tt = target + sd.tab;
    const float* const ww = weight + sd.tab;
    double acc;
    if (type == 0)  // partner row j, position dim_o-1-a: tap r-t at 4 (j dim_o + dim_o-1 + t - r)
        acc = rowsum_cols_walk<true>(tt, ww, src, 4 * (dim_o - 1), 4 * dim_o, 4, -4 * r, r, dim_o, f, l, lo, hi, fmax,
                                     lmin);
    else if (type == 3)  // partner row dim_i-1-j, position dim_o-1-a
        acc = rowsum_cols_walk<true>(tt, ww, src, 4 * ((dim_i - 1) * dim_o + dim_o - 1), -4 * dim_o, 4, -4 * r, r,
                                     dim_o, f, l, lo, hi, fmax, lmin);
    else  // types 1, 2: partner row dim_i-1-j, position a: tap r-t-1 at 4 ((dim_i-1-j) dim_o + r - t - 1)
        acc = rowsum_cols_walk<false>(tt, ww, src, 4 * ((dim_i - 1) * dim_o - 1), -4 * dim_o, -4, 4 * r, r, dim_o,
                                      f, l, lo, hi, fmax, lmin);
    if (live) P[sd.q + r] = acc;
}

hipError_t launch_radn_rowsum(const DevPlan& p, const float* in4, const float* Sbuf, double* P, hipStream_t s) {
    if (!Sbuf) {  // writes every row r < H of every sector (rows without cells: 0)
        if (p.nsec == 0 || p.maxH <= 0) return hipSuccess;
        hipLaunchKernelGGL(k_radn_rowsum_cols, dim3((p.maxH + 255) / 256, p.nsec), dim3(256), 0, s, p.secs, in4,
                           p.target, p.weight, p.fl, P);
        return hipGetLastError();
    }
    if (p.nwork == 0) return hipSuccess;  // writes the rows of the work list only: P zeroed by the caller
    hipLaunchKernelGGL(k_radn_rowsum, dim3((p.nwork + 3) / 4), dim3(256), 0, s, p.secs, p.work, p.nwork, in4,
                       p.target, p.weight, p.fl, Sbuf, P);
    return hipGetLastError();
}

// One DIF pass over re/im in LDS (src/helper.h:65-82) followed by the
// bit-reversal permutation (:84-100).
__device__ void lds_fft(double* re, double* im, int N, int m, const double2* __restrict__ tw) {
    for (int n = N; n > 1; n >>= 1) {
        const int k = n >> 1;
        const double2* T = tw + (N - n);
        for (int beta = threadIdx.x; beta < (N >> 1); beta += blockDim.x) {
            const int l = beta & (k - 1);
            const int a = (beta >> (__ffs(k) - 1)) * n + l;  // (beta / k) * n + l
            const int b = a + k;
            const double2 t = T[l];
            const double tr = re[a] - re[b], ti = im[a] - im[b];
            re[a] = re[a] + re[b];
            im[a] = im[a] + im[b];
            re[b] = tr * t.x - ti * t.y;
            im[b] = tr * t.y + ti * t.x;
        }
        __syncthreads();
    }
    for (int a = threadIdx.x; a < N; a += blockDim.x) {
        const int b = (int)(__brev((unsigned)a) >> (32 - m));
        if (b > a) {
            double t = re[a]; re[a] = re[b]; re[b] = t;
            t = im[a]; im[a] = im[b]; im[b] = t;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_radn_filter(const SecDesc* __restrict__ secs, const double* __restrict__ P,
                                                     float* __restrict__ q, const double2* __restrict__ tw,
                                                     const int* __restrict__ tw_off) {
    extern __shared__ double buf[];
    const SecDesc sd = secs[blockIdx.x];
    const int H = sd.H;
    const int m = (H <= 1) ? 0 : 32 - __clz(H - 1);  // H2 = 2^ceil(log2 H)
    const int H2 = 1 << m;
    double* re = buf;
    double* im = buf + H2;
    for (int i = threadIdx.x; i < H2; i += blockDim.x) {
        re[i] = (i < H) ? P[sd.q + i] : 0.0;
        im[i] = 0.0;
    }
    __syncthreads();
    if (H2 > 1) lds_fft(re, im, H2, m, tw + tw_off[m]);
    // ramp: zero above H/2, gain (double)i/H below (kernelRADN.h:247);
    // then the conj of ifft's first step.
    for (int i = threadIdx.x; i < H2; i += blockDim.x) {
        if (i > H / 2) {
            re[i] = 0.0;
            im[i] = -0.0;
        } else {
            const double g = (double)i / H;
            re[i] = g * re[i];
            im[i] = -(g * im[i]);
        }
    }
    __syncthreads();
    if (H2 > 1) lds_fft(re, im, H2, m, tw + tw_off[m]);
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        const double x = re[i] / H2;
        q[sd.q + i] = (float)(0.0 + x);  // skewOutput (zero) += Re
    }
}

hipError_t launch_radn_filter(const DevPlan& p, const double* P, float* q, const double* tw, const int* tw_off,
                              int maxH2, hipStream_t s) {
    if (p.nsec == 0) return hipSuccess;
    const size_t lds = sizeof(double) * 2 * (size_t)maxH2;
    hipLaunchKernelGGL(k_radn_filter, dim3(p.nsec), dim3(256), lds, s, p.secs, P, q,
                       reinterpret_cast<const double2*>(tw), tw_off);
    return hipGetLastError();
}

// Smear the filtered projection back along each row (kernelRADN.h:252-257)
// into a materialised skewOutput (the per-sector engine API).
__global__ __launch_bounds__(256) void k_radn_smear(const SecDesc* __restrict__ secs, const RowWork* __restrict__ work,
                                                    int nwork, const uint32_t* __restrict__ fl,
                                                    const float* __restrict__ q, float* __restrict__ SO) {
    const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= nwork) return;
    const int lane = threadIdx.x & 63;
    const RowWork w = work[item];
    const SecDesc sd = secs[w.sec];
    const uint32_t v = fl[sd.rows + w.row];
    const float x = q[sd.q + w.row];
    for (int j = fl_first(v) + lane; j < fl_last(v); j += 64) SO[sd.so + (size_t)w.row * sd.dim_i + j] = x;
}

hipError_t launch_radn_smear(const DevPlan& p, const float* q, float* SO, hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
    hipLaunchKernelGGL(k_radn_smear, dim3((p.nwork + 3) / 4), dim3(256), 0, s, p.secs, p.work, p.nwork, p.fl, q, SO);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Local Radon transform (LRT=true).  For skewed cell (i, j) the reference
// (pointRadon, kernelRADN.h:114-182) forms 64 samples
//   x_kk = g_kk * sum_{l=j-32}^{j+32} S(k, l) [first_k <= l < last_k],
//   k = i - 32 + kk (zero for k outside [0, H)),
// g_kk = exp(-(kk-32)^2 / (2 sigma^2)), sigma = (64/5)/scale, then
//   X = FFT64(x)/8,  y_f = sin(f/24.6 * pi/2) Re X_f for f <= 25, 0 above,
//   Y = FFT64(y)/8,  out = |Y_{m-1}|^2 at the first m in [5, 32) with
//   |Y_m|^2 > |Y_{m-1}|^2 (0 if none; getParameter :96-108).
// Here in two passes:
//   lrt_rowsum: one workgroup per skewed row; the row is staged in LDS and
//     every column's 65-wide window sum is formed in double in the
//     reference's order (ascending l), for all dim_i columns -> W.
//   op_lrt: one wave per skewed row, lanes over its cells.  Each cell reads
//     its 64 window sums down column j of W (coalesced across the lanes),
//     and evaluates the 26 needed real parts of X and then |Y_m|^2 for
//     m = 4, 5, ... directly as cosine/sine sums (FP64), stopping at the
//     first rise.  The direct sums differ from the reference's radix-2 FFT
//     only by FP64 rounding.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(128) void k_lrt_rowsum(const SecDesc* __restrict__ secs, const RowWork* __restrict__ work,
                                                    const float* __restrict__ in4, const int* __restrict__ target,
                                                    const float* __restrict__ weight, const uint32_t* __restrict__ fl,
                                                    const float* __restrict__ Sbuf, double* __restrict__ W) {
    extern __shared__ float lds[];
    float* row = lds + kPad;
    const RowWork w = work[blockIdx.x];
    const SecDesc sd = secs[w.sec];
    const uint32_t v = fl[sd.rows + w.row];
    const int f = fl_first(v);
    const int L = fl_last(v) - f;
    if (L <= 0) return;
    stage_row(row, sd, w.row, f, L, in4, target, weight, Sbuf);
    __syncthreads();
    double* out = W + sd.so + (size_t)w.row * sd.dim_i;
    for (int j = threadIdx.x; j < sd.dim_i; j += blockDim.x) {
        const int lo = max(j - 32, f) - f, hi = min(j + 32, f + L - 1) - f;
        double acc = 0.0;
        for (int l = lo; l <= hi; ++l) acc += (double)row[l];
        out[j] = acc;
    }
}

// tab: the sine ramp sin(f/24.6 * pi/2) for f <= 25 at [160,186), the two
// DFT matrices at kLrtC and kLrtY (ske_types.h); gauss: 64 weights per folded
// angle (newAngle).
__global__ __launch_bounds__(64) void k_op_lrt(const SecDesc* __restrict__ secs, const RowWork* __restrict__ work,
                                               const uint32_t* __restrict__ fl, const double* __restrict__ W,
                                               const double* __restrict__ tab, const double* __restrict__ gauss,
                                               float* __restrict__ arena) {
    const RowWork w = work[blockIdx.x];
    const SecDesc sd = secs[w.sec];
    const int i = w.row;
    const uint32_t v = fl[sd.rows + i];
    const int f = fl_first(v), l = fl_last(v);
    const double* g = gauss + 64 * sd.newAngle;
    const double* Wb = W + sd.so;
    const int k0 = max(0, 32 - i), k1 = min(64, sd.H - i + 32);  // kk with 0 <= i-32+kk < H
    for (int j = f + (int)threadIdx.x; j < l; j += 64) {
        double re1[26];
#pragma unroll
        for (int q = 0; q < 26; ++q) re1[q] = 0.0;
        int kk = k0;
        // four window sums in flight per pass; re1 still accumulates in kk order
        for (; kk + 4 <= k1; kk += 4) {
            double xs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) xs[u] = Wb[(size_t)(i - 32 + kk + u) * sd.dim_i + j];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double x = xs[u] * g[kk + u];
                const double* C = tab + kLrtC + 32 * (kk + u);  // wave-uniform row: scalar loads
#pragma unroll
                for (int q = 0; q < 26; ++q) re1[q] = __builtin_fma(x, C[q], re1[q]);
            }
        }
        for (; kk < k1; ++kk) {
            const double x = Wb[(size_t)(i - 32 + kk) * sd.dim_i + j] * g[kk];
            const double* C = tab + kLrtC + 32 * kk;
#pragma unroll
            for (int q = 0; q < 26; ++q) re1[q] = __builtin_fma(x, C[q], re1[q]);
        }
#pragma unroll
        for (int q = 0; q < 26; ++q) re1[q] *= 0.125 * tab[160 + q];
        double prev = 0.0, val = 0.0;
        for (int m = 4; m < 32; ++m) {
            double yr = 0.0, yi = 0.0;
            const double* Y = tab + kLrtY + 64 * m;
#pragma unroll
            for (int q = 0; q < 26; ++q) {
                yr = __builtin_fma(re1[q], Y[2 * q], yr);
                yi = __builtin_fma(re1[q], Y[2 * q + 1], yi);
            }
            yr *= 0.125;
            yi *= 0.125;
            const double cur = yr * yr + yi * yi;
            if (m > 4 && cur > prev) {
                val = prev;
                break;
            }
            prev = cur;
        }
        arena[sd.so + (size_t)i * sd.dim_i + j] = (float)val;
    }
}

// The vectors the reference records at a sample point (pointRadon,
// kernelRADN.h:117-170, for the cells of samplePoints): x = the 64 windowed
// sums (from W) times the Gaussian, the Gaussian-weighted row through the
// cell (column j - 32 + kk of row i, inside the extents of row k = i - 32 + kk
// as the reference reads them; 0 for k outside [0, H)), y = the filtered
// real spectrum and Re FFT64(y)/8.  One 64-lane wave per request; direct
// sums in FP64 (cos/sin(2 pi t/64) from tab).
__global__ __launch_bounds__(64) void k_lrt_samples(const SecDesc* __restrict__ secs,
                                                    const SampleReq* __restrict__ req,
                                                    const uint32_t* __restrict__ fl, const double* __restrict__ W,
                                                    const float* __restrict__ in4, const int* __restrict__ target,
                                                    const float* __restrict__ weight, const double* __restrict__ tab,
                                                    const double* __restrict__ gauss, float* __restrict__ out) {
    __shared__ double xs[64], ys[64];
    const SampleReq r = req[blockIdx.x];
    const SecDesc sd = secs[r.sec];
    const int kk = threadIdx.x;
    const double g = gauss[64 * sd.newAngle + kk];
    const int k = r.i - 32 + kk;
    double win = 0.0, axis = 0.0;
    if (k >= 0 && k < sd.H) {
        win = W[sd.so + (size_t)k * sd.dim_i + r.j];
        const uint32_t v = fl[sd.rows + k];
        const int c = r.j - 32 + kk;
        if (c >= fl_first(v) && c < fl_last(v)) {
            const float* src = in4 + sd.img + (size_t)sd.type * sd.N;
            axis = g * (double)skew_cell(src, target[sd.tab + c], weight[sd.tab + c], r.i, c, sd.dim_i, sd.dim_o);
        }
    }
    const double x = win * g;
    xs[kk] = x;
    __syncthreads();
    double re = 0.0;
    for (int m = 0; m < 64; ++m) re = __builtin_fma(xs[m], tab[2 * ((kk * m) & 63)], re);
    const double y = kk <= 25 ? tab[160 + kk] * (re * 0.125) : 0.0;  // cutoff 0.4 * 64, sine ramp
    ys[kk] = y;
    __syncthreads();
    double yr = 0.0;
    for (int m = 0; m < 64; ++m) yr = __builtin_fma(ys[m], tab[2 * ((kk * m) & 63)], yr);
    float* o = out + (size_t)r.slot * 256;
    o[kk] = (float)x;
    o[64 + kk] = (float)axis;
    o[128 + kk] = (float)y;
    o[192 + kk] = (float)(yr * 0.125);
}

hipError_t launch_lrt_samples(const DevPlan& p, const SampleReq* req, int nreq, const float* in4, const double* W,
                              const double* tab, const double* gauss, float* out, hipStream_t s) {
    if (nreq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lrt_samples, dim3(nreq), dim3(64), 0, s, p.secs, req, p.fl, W, in4, p.target, p.weight, tab,
                       gauss, out);
    return hipGetLastError();
}

hipError_t launch_lrt(const DevPlan& p, const float* in4, const float* Sbuf, double* W, float* arena,
                      const double* tab, const double* gauss, hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
    const size_t lds = sizeof(float) * (size_t)(p.maxL + 2 * kPad);
    hipLaunchKernelGGL(k_lrt_rowsum, dim3(p.nwork), dim3(128), lds, s, p.secs, p.work, in4, p.target, p.weight, p.fl,
                       Sbuf, W);
    hipLaunchKernelGGL(k_op_lrt, dim3(p.nwork), dim3(64), 0, s, p.secs, p.work, p.fl, W, tab, gauss, arena);
    return hipGetLastError();
}

}  // namespace ske
