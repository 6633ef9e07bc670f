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

hipError_t launch_op_sdem(const DevPlan& p, const float* in4, const float* Sbuf, float* arena,
                          const float4* rtab, float povh, int variant, hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
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
