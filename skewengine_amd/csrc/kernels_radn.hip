// kernels_radn.hip -- global Radon operator (src/kernelRADN.h:221-258,
// LRT=false) on gfx950.
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
    for (int j = f + lane; j < l; j += 64) {
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

hipError_t launch_radn_rowsum(const DevPlan& p, const float* in4, const float* Sbuf, double* P, hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
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

}  // namespace ske
