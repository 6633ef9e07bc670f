// kernels_fbp.hip -- standard parallel-beam geometry (no reference code; the
// counterpart of the reference's skew-space Radon, src/kernelRADN.h:212-259,
// in the 725-bin geometry of the benchmark configurations):
//
//   image   n x n on [-1,1]^2, pixel pitch delta = 2/n, row 0 at y = +1;
//   angles  theta_a (host table of cos/sin);
//   bins    s_k = (k - (ndet-1)/2) * delta.
//
//   project     p(a,k) = delta * sum_t f(s_k cos - t sin, s_k sin + t cos),
//               t on a delta grid over the chord of the ray through the
//               image square, f bilinear (zero outside); lanes run along
//               whichever of bins or samples keeps the gathers on one image
//               row;
//   filter      q(a,k) = sum_k' p(a,k') h(k-k'), h = delta * Ram-Lak
//               (h0 = 1/(4 delta^2), odd m: -1/(m^2 pi^2 delta^2), even: 0) or
//               learned taps;
//   backproject f(x,y) = (pi/nang) sum_a q(a, s(x,y)), linear in s.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "fbp.h"

namespace ske {

// The projector reads a copy of each image with a zero border of kProjPad
// pixels, so the bilinear taps need no bounds tests: the samples of
// ray_range lie within 1 + 2 delta of the centre along each axis, i.e. their
// taps within 3 pixels of the image (checked exhaustively on the CPU for
// n = 2..1024, 2..2n+3 bins and arbitrary angles), and zeros there equal the
// zero-outside rule.
__global__ __launch_bounds__(256) void k_fbp_pad(const float* __restrict__ img, float* __restrict__ pad, int n,
                                                 int np, size_t total) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const size_t per = (size_t)np * np;
    const size_t b = e / per;
    const int r = (int)((e - b * per) / np) - kProjPad, c = (int)(e % np) - kProjPad;
    pad[e] = (r >= 0 && r < n && c >= 0 && c < n) ? img[b * n * n + (size_t)r * n + c] : 0.0f;
}

__device__ __forceinline__ float bilinear_padded(const float* __restrict__ pim, int np, float u, float v) {
    // u: column coordinate, v: row coordinate, pixel centres at integers
    const float fu = floorf(u), fv = floorf(v);
    // 24-bit multiply: row and np are below 2^14, the offset below 2^31
    const float* r = pim + (__mul24((int)fv + kProjPad, np) + ((int)fu + kProjPad));
    const float wu = u - fu, wv = v - fv;
    const float a = r[0], b = r[1], c = r[np], d = r[np + 1];
    const float top = a + wu * (b - a);
    const float bot = c + wu * (d - c);
    return top + wv * (bot - top);
}

// Sample range [i0, i1] of the ray of bin k at angle (c, sn): the samples
// t_i = (i - (nt-1)/2) delta inside the chord of the line x c + y sn = s
// through the square [-B, B]^2, B = 1 + delta (the bilinear image is zero
// farther than 1 + delta/2 from the centre along either axis, so samples
// outside add exact zeros).  x = s c - t sn, y = s sn + t c.  Empty: i1 < i0.
__device__ __forceinline__ void ray_range(const FbpGeom& g, float s, float c, float sn, int& i0, int& i1) {
    const float B = 1.0f + g.delta;
    float tlo = -g.tmax, thi = g.tmax;
    if (fabsf(sn) > 1e-7f) {
        const float t1 = (s * c - B) / sn, t2v = (s * c + B) / sn;
        tlo = fmaxf(tlo, fminf(t1, t2v));
        thi = fminf(thi, fmaxf(t1, t2v));
    } else if (fabsf(s * c) > B) {
        thi = tlo - 1.0f;
    }
    if (fabsf(c) > 1e-7f) {
        const float t1 = (-B - s * sn) / c, t2v = (B - s * sn) / c;
        tlo = fmaxf(tlo, fminf(t1, t2v));
        thi = fminf(thi, fmaxf(t1, t2v));
    } else if (fabsf(s * sn) > B) {
        thi = tlo - 1.0f;
    }
    i0 = 0;
    i1 = -1;
    if (thi >= tlo) {
        const float half = 0.5f * (float)(g.nt - 1);
        i0 = max((int)ceilf(tlo / g.delta + half) - 1, 0);
        i1 = min((int)floorf(thi / g.delta + half) + 1, g.nt - 1);
    }
}

__device__ __forceinline__ float ray_sample(const float* __restrict__ pim, int np, const FbpGeom& g, float s,
                                            float c, float sn, int i) {
    const float half = 0.5f * (float)(g.nt - 1);
    const float inv = 1.0f / g.delta;
    const float t = ((float)i - half) * g.delta;
    const float x = s * c - t * sn;
    const float y = s * sn + t * c;
    return bilinear_padded(pim, np, (x + 1.0f) * inv - 0.5f, (1.0f - y) * inv - 0.5f);
}

// Angles with |cos| >= |sin|: one thread per (bin, angle, slice) summing its
// ray's samples in order.  Neighbouring lanes (bins) sample points
// delta (cos, sin) apart: nearly the same image row, so the gathers
// coalesce.  `ang` lists the angle indices this launch covers.
__global__ __launch_bounds__(256) void k_fbp_project_bins(const float* __restrict__ pimg, float* __restrict__ sino,
                                                          const float2* __restrict__ cs, const int* __restrict__ ang,
                                                          FbpGeom g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = ang[blockIdx.y];
    const int b = blockIdx.z;
    if (k >= g.ndet) return;
    const float2 t2 = cs[a];
    const float s = ((float)k - g.det_c) * g.delta;
    int i0, i1;
    ray_range(g, s, t2.x, t2.y, i0, i1);
    const int np = g.n + 2 * kProjPad;
    const float* im = pimg + (size_t)b * np * np;
    // four samples' gathers in flight per pass, summed in sample order
    float acc = 0.0f;
    int i = i0;
    for (; i + 3 <= i1; i += 4) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ray_sample(im, np, g, s, t2.x, t2.y, i + u);
        acc += v[0];
        acc += v[1];
        acc += v[2];
        acc += v[3];
    }
    for (; i <= i1; ++i) acc += ray_sample(im, np, g, s, t2.x, t2.y, i);
    sino[((size_t)b * g.nang + a) * g.ndet + k] = acc * g.delta;
}

// Angles with |sin| > |cos|: one wave per ray, lanes over its samples
// (neighbouring samples are delta (-sin, cos) apart: again nearly one image
// row), then a wave reduction.
__global__ __launch_bounds__(256) void k_fbp_project_rays(const float* __restrict__ pimg, float* __restrict__ sino,
                                                          const float2* __restrict__ cs, const int* __restrict__ ang,
                                                          FbpGeom g) {
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int a = ang[blockIdx.y];
    const int b = blockIdx.z;
    if (k >= g.ndet) return;
    const float2 t2 = cs[a];
    const float s = ((float)k - g.det_c) * g.delta;
    int i0, i1;
    ray_range(g, s, t2.x, t2.y, i0, i1);
    const int np = g.n + 2 * kProjPad;
    const float* im = pimg + (size_t)b * np * np;
    float acc = 0.0f;
    int i = i0 + lane;
    for (; i + 64 <= i1; i += 128) {
        const float v0 = ray_sample(im, np, g, s, t2.x, t2.y, i);
        const float v1 = ray_sample(im, np, g, s, t2.x, t2.y, i + 64);
        acc += v0;
        acc += v1;
    }
    if (i <= i1) acc += ray_sample(im, np, g, s, t2.x, t2.y, i);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) sino[((size_t)b * g.nang + a) * g.ndet + k] = acc * g.delta;
}

// One workgroup per sinogram row: the row and the 2*ndet-1 filter taps in
// LDS.  A thread owns 4 consecutive outputs and slides a 4-tap register
// window along h, so each step costs one broadcast read of p, one read of h
// and 4 FMAs; every output still sums k' = 0, 1, ... in order.
__global__ __launch_bounds__(256) void k_fbp_filter(const float* __restrict__ sino, float* __restrict__ filt,
                                                    const float* __restrict__ taps, int ndet) {
    extern __shared__ float sm[];
    float* p = sm;
    float* h = sm + ndet;
    const size_t row = blockIdx.x;
    const float* src = sino + row * ndet;
    for (int i = threadIdx.x; i < ndet; i += blockDim.x) p[i] = src[i];
    for (int i = threadIdx.x; i < 2 * ndet - 1; i += blockDim.x) h[i] = taps[i];
    __syncthreads();
    for (int k = 4 * (int)threadIdx.x; k < ndet; k += 4 * (int)blockDim.x) {
        // outputs k..k+3 (those >= ndet are computed and dropped); h index
        // of output k+m at k' is k + m - k' + ndet - 1, in [0, 2 ndet - 1)
        // for k + m < ndet; taps past the end read as 0
        const int top = 2 * ndet - 1;
        float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f;
        int hb = k + ndet - 1;  // h index of output k at k' = 0
        float w0 = h[hb];
        float w1 = hb + 1 < top ? h[hb + 1] : 0.0f;
        float w2 = hb + 2 < top ? h[hb + 2] : 0.0f;
        float w3 = hb + 3 < top ? h[hb + 3] : 0.0f;
        for (int kp = 0; kp < ndet; ++kp) {
            const float x = p[kp];
            acc0 = __builtin_fmaf(x, w0, acc0);
            acc1 = __builtin_fmaf(x, w1, acc1);
            acc2 = __builtin_fmaf(x, w2, acc2);
            acc3 = __builtin_fmaf(x, w3, acc3);
            w3 = w2;
            w2 = w1;
            w1 = w0;
            --hb;
            w0 = hb >= 0 ? h[hb] : 0.0f;
        }
        float* o = filt + row * ndet;
        o[k] = acc0;
        if (k + 1 < ndet) o[k + 1] = acc1;
        if (k + 2 < ndet) o[k + 2] = acc2;
        if (k + 3 < ndet) o[k + 3] = acc3;
    }
}

// One thread per pixel, one workgroup per 16 x 16 pixel tile.  A tile
// projects onto at most 16*sqrt(2) + 2 < 25 consecutive bins of any angle,
// so per angle only a 32-bin window of the filtered row, starting at
// floor(projection of the tile centre) - 16, is staged in LDS (zeros outside
// the detector, which is what the bounds test of the linear interpolation
// gives); angles go in chunks of up to kBpChunk.
constexpr int kBpWin = 32;
constexpr int kBpChunk = 128;
__global__ __launch_bounds__(256) void k_fbp_backproject(const float* __restrict__ filt, float* __restrict__ img,
                                                         const float2* __restrict__ cs, FbpGeom g, float scale) {
    __shared__ float q[kBpChunk * kBpWin];
    __shared__ int base[kBpChunk];
    const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
    const int j = j0 + (threadIdx.x & 15);
    const int i = i0 + (threadIdx.x >> 4);
    const int b = blockIdx.z;
    const bool in = i < g.n && j < g.n;
    const float x = ((float)j + 0.5f) * g.delta - 1.0f;
    const float y = 1.0f - ((float)i + 0.5f) * g.delta;
    const float xc = ((float)j0 + 8.0f) * g.delta - 1.0f;
    const float yc = 1.0f - ((float)i0 + 8.0f) * g.delta;
    const float inv = 1.0f / g.delta;
    const float* fb = filt + (size_t)b * g.nang * g.ndet;
    float acc = 0.0f;
    for (int a0 = 0; a0 < g.nang; a0 += kBpChunk) {
        const int na = min(kBpChunk, g.nang - a0);
        __syncthreads();
        for (int t = threadIdx.x; t < na; t += blockDim.x) {
            const float2 t2 = cs[a0 + t];
            base[t] = (int)floorf((xc * t2.x + yc * t2.y) * inv + g.det_c) - kBpWin / 2;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < na * kBpWin; t += blockDim.x) {
            const int aa = t / kBpWin;
            const int k = base[aa] + (t % kBpWin);
            q[t] = (k >= 0 && k < g.ndet) ? fb[(size_t)(a0 + aa) * g.ndet + k] : 0.0f;
        }
        __syncthreads();
        if (in) {
            for (int aa = 0; aa < na; ++aa) {
                const float2 t2 = cs[a0 + aa];
                const float sp = x * t2.x + y * t2.y;
                const float kf = sp * inv + g.det_c;
                const float fk = floorf(kf);
                const int k0 = (int)fk;
                const float w = kf - fk;
                const int r = aa * kBpWin + k0 - base[aa];  // in [4, 28) (tile radius < 12 bins)
                const float v0 = q[r];
                const float v1 = q[r + 1];
                acc += v0 + w * (v1 - v0);
            }
        }
    }
    if (in) img[(size_t)b * g.n * g.n + (size_t)i * g.n + j] = acc * scale;
}

hipError_t launch_fbp_project(const float* img, float* padded, float* sino, const float2* cs, const int* ang_bins,
                              int n_bins, const int* ang_rays, int n_rays, const FbpGeom& g, int batch, hipStream_t s) {
    const int np = g.n + 2 * kProjPad;
    const size_t total = (size_t)batch * np * np;
    hipLaunchKernelGGL(k_fbp_pad, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, img, padded, g.n, np, total);
    if (n_bins > 0)
        hipLaunchKernelGGL(k_fbp_project_bins, dim3((g.ndet + 255) / 256, n_bins, batch), dim3(256), 0, s, padded,
                           sino, cs, ang_bins, g);
    if (n_rays > 0)
        hipLaunchKernelGGL(k_fbp_project_rays, dim3((g.ndet + 3) / 4, n_rays, batch), dim3(256), 0, s, padded, sino,
                           cs, ang_rays, g);
    return hipGetLastError();
}

hipError_t launch_fbp_filter(const float* sino, float* filt, const float* taps, const FbpGeom& g, int batch,
                             hipStream_t s) {
    const size_t lds = sizeof(float) * (size_t)(3 * g.ndet - 1);
    hipLaunchKernelGGL(k_fbp_filter, dim3((unsigned)((size_t)batch * g.nang)), dim3(256), lds, s, sino, filt, taps,
                       g.ndet);
    return hipGetLastError();
}

hipError_t launch_fbp_backproject(const float* filt, float* img, const float2* cs, const FbpGeom& g, int batch,
                                  float scale, hipStream_t s) {
    dim3 grid((g.n + 15) / 16, (g.n + 15) / 16, batch);
    hipLaunchKernelGGL(k_fbp_backproject, grid, dim3(256), 0, s, filt, img, cs, g, scale);
    return hipGetLastError();
}

}  // namespace ske
