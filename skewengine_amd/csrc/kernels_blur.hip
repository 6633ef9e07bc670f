// kernels_blur.hip -- cepstral blur operator (src/kernelBLUR.h:267-302,
// lineCepstrum :153-228, select_subarray :69-85, getBlurParameter2 :88-110)
// on gfx950, FP64 throughout like the reference.
//
// Per skewed cell j of row r the reference takes the 64-sample window
// S(r, j-32 .. j+31) (zero outside [first,last)), X = DFT64/8, the log power
// of bins 0..31, C = DFT32/sqrt(32) of those logs, v_k = -min(0, Re C_k) for
// k < 16 and a peak rule on v, times 1e4.
//
// Layout: one wave per skewed row; the row is staged in LDS with 64 zeros of
// guard band on each side (which is exactly the window's zero fill).  Lane l
// owns a contiguous segment of cells.  It computes the 32 needed bins of the
// segment's first window directly and then slides the window one cell at a
// time: X_k(j+1) = e^{+2 pi i k/64} (X_k(j) + (x[j+32] - x[j-32])/8).  A
// segment is at most 16 cells, so rounding drift stays at the 1e-15 level.
// Re C_k is the cosine sum of the real log spectrum, which is what the
// reference's complex DFT of a real input yields in its real part.
#include "device_common.h"
#include "launch.h"

namespace ske {

// Natural log of a double: x = m 2^e with m in [1, 2); the top 8 fraction
// bits i of m pick c_i ~ 1/(1 + (i + 1/2)/256) from the table, so that
// u = m c_i - 1 (one fma) has |u| < 2^-9, and
//     log x = e ln2 + (-log c_i) + log1p(u),
//     log1p(u) = u - u^2/2 + u^3/3 - u^4/4 + u^5/5   (the rest < 2^-55).
// -log c_i and ln 2 (as a double-double) come from the host in extended
// precision.  x = 0 gives -inf like the libm log.  `lt` = the table in LDS.
__device__ __forceinline__ double fast_log(double x, const double* lt) {
    const double mant = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
    const int ex = __builtin_amdgcn_frexp_exp(x) - 1;
    const int i = (int)((uint32_t)(__double_as_longlong(mant) >> 44) & 0xFFu);
    const double c = lt[2 * i], t = lt[2 * i + 1];
    const double u = __builtin_fma(2.0 * mant, c, -1.0);
    const double q = __builtin_fma(u, __builtin_fma(u, __builtin_fma(u, 0.2, -0.25), 1.0 / 3.0), -0.5);
    const double p = __builtin_fma(u * u, q, u);
    const double e = (double)ex;
    const double r = __builtin_fma(e, lt[513], t + p);
    const double y = __builtin_fma(e, lt[512], r);
    return x == 0.0 ? -INFINITY : y;
}

// tab layout (double): [0,128): cos/sin(2 pi t/64) interleaved; [128,160): cos(2 pi t/32);
// from kLogTab the log table of fast_log; from kBlurD the window DFT by rows.
__global__ __launch_bounds__(64) void k_op_blur(const SecDesc* __restrict__ secs, const RowWork* __restrict__ work,
                                                const float* __restrict__ in4, const int* __restrict__ target,
                                                const float* __restrict__ weight, const uint32_t* __restrict__ fl,
                                                const float* __restrict__ Sbuf, const double* __restrict__ tab,
                                                float* __restrict__ arena, int p_maxL) {
    extern __shared__ float lds[];
    float* row = lds + kPad;
    double* lt = reinterpret_cast<double*>(lds) + ((size_t)(p_maxL + 2 * kPad) + 1) / 2;  // after the row
    for (int t = threadIdx.x; t < 514; t += 64) lt[t] = tab[kLogTab + t];
    const RowWork w = work[blockIdx.x];
    const SecDesc sd = secs[w.sec];
    const int r = w.row;
    const uint32_t v = fl[sd.rows + r];
    const int f = fl_first(v);
    const int L = fl_last(v) - f;
    if (L <= 0) return;
    stage_row(row, sd, r, f, L, in4, target, weight, Sbuf);
    __syncthreads();

    const int lane = threadIdx.x;
    const int seg = (L + 63) >> 6;
    const int c0 = lane * seg;
    const int c1 = min(c0 + seg, L);
    if (c0 >= c1) return;
    float* so = arena + sd.so + (size_t)r * sd.dim_i + f;

    double Xr[32], Xi[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) { Xr[k] = 0.0; Xi[k] = 0.0; }
    for (int m = 0; m < 64; ++m) {
        const double x = (double)row[c0 - 32 + m] * 0.125;
        const double* D = tab + kBlurD + 64 * m;  // wave-uniform row: contiguous scalar loads
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            Xr[k] = __builtin_fma(x, D[2 * k], Xr[k]);
            Xi[k] = __builtin_fma(-x, D[2 * k + 1], Xi[k]);
        }
    }
    const double inv_sqrt32 = 0.17677669529663687;  // 1/sqrt(32)
    for (int c = c0; c < c1; ++c) {
        if (c > c0) {
            const double delta = ((double)row[c + 31] - (double)row[c - 33]) * 0.125;
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const double cr = tab[2 * k], sr = tab[2 * k + 1];
                const double a = Xr[k] + delta, b = Xi[k];
                Xr[k] = __builtin_fma(a, cr, -(b * sr));
                Xi[k] = __builtin_fma(a, sr, b * cr);
            }
        }
        double Lg[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) Lg[k] = fast_log(Xr[k] * Xr[k] + Xi[k] * Xi[k], lt);
        double vv[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            double C = 0.0;
#pragma unroll
            for (int n = 0; n < 32; ++n) C = __builtin_fma(Lg[n], tab[128 + ((k * n) & 31)], C);
            C *= inv_sqrt32;
            vv[k] = (C < 0.0) ? -C : -0.0;  // -std::min(0.0, C)
        }
        // getBlurParameter2(v, 16): first i >= 5 with v[i] > v[i-1]; the
        // value is max(v[i-1], v[i+1..15]) (v[i] itself is skipped).
        int i0 = 16;
#pragma unroll
        for (int i = 15; i >= 5; --i)
            if (vv[i] > vv[i - 1]) i0 = i;
        double V = 0.0;
        if (i0 < 16) {
#pragma unroll
            for (int i = 5; i < 16; ++i)
                if (i == i0) V = vv[i - 1];
#pragma unroll
            for (int mm = 6; mm < 16; ++mm)
                if (mm > i0 && vv[mm] > V) V = vv[mm];
        }
        so[c] = (float)(V * 10000);
    }
}

hipError_t launch_op_blur(const DevPlan& p, const float* in4, const float* Sbuf, float* arena, const double* tab,
                          hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
    const size_t row_floats = ((size_t)(p.maxL + 2 * kPad) + 1) / 2 * 2;  // keep the log table 8-byte aligned
    const size_t lds = sizeof(float) * row_floats + sizeof(double) * 514;
    hipLaunchKernelGGL(k_op_blur, dim3(p.nwork), dim3(64), lds, s, p.secs, p.work, in4, p.target, p.weight, p.fl,
                       Sbuf, tab, arena, p.maxL);
    return hipGetLastError();
}

}  // namespace ske
