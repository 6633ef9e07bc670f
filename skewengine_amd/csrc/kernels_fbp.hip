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
//               image square, f bilinear (zero outside); a wave's lanes
//               cover 8 bins x 8 samples, a compact block of the image;
//   filter      q(a,k) = sum_k' p(a,k') h(k-k'), h = delta * Ram-Lak
//               (h0 = 1/(4 delta^2), odd m: -1/(m^2 pi^2 delta^2), even: 0) or
//               learned taps, as a GEMM with a Toeplitz matrix on the FP32
//               matrix cores;
//   backproject f(x,y) = (pi/nang) sum_a q(a, s(x,y)), linear in s.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>

#include "fbp.h"

namespace ske {

// The projector reads a copy of each image with a zero border of kProjPad
// pixels, so the bilinear taps (ray_sample) need no bounds tests: the samples of
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

// Sample i of the ray of bin offset s at angle (c, sn) in pixel coordinates:
// with t = (i - half) delta, u = (s c - t sn + 1)/delta - 1/2 = U0 - i sn and
// v = (1 - s sn - t c)/delta - 1/2 = V0 - i c, so a sample costs two fma.
struct RayCoord {
    float u0, v0;
};
__device__ __forceinline__ RayCoord ray_coord(const FbpGeom& g, float s, float c, float sn) {
    const float half = 0.5f * (float)(g.nt - 1);
    const float inv = 1.0f / g.delta;
    return RayCoord{((s * c + 1.0f) * inv - 0.5f) + half * sn, ((1.0f - s * sn) * inv - 0.5f) + half * c};
}

__device__ __forceinline__ float ray_sample(const float* __restrict__ pim, int np, const RayCoord& rc, float c,
                                            float sn, int i) {
    const float fi = (float)i;
    const float u = __builtin_fmaf(-fi, sn, rc.u0), v = __builtin_fmaf(-fi, c, rc.v0);
    const float fu = floorf(u), fv = floorf(v);
    // 24-bit multiply: row and np are below 2^14, the offset below 2^31
    const float* r = pim + (__mul24((int)fv + kProjPad, np) + ((int)fu + kProjPad));
    const float wu = u - fu, wv = v - fv;
    const float a = r[0], b = r[1], cc = r[np], d = r[np + 1];
    const float top = __builtin_fmaf(wu, b - a, a);
    const float bot = __builtin_fmaf(wu, d - cc, cc);
    return __builtin_fmaf(wv, bot - top, top);
}

// One wave per 8 bins of one angle and slice: lane (kb, ib) = (lane % 8,
// lane / 8) takes the samples i = i_lo + ib, i_lo + ib + 8, ... of bin
// k0 + kb, where [i_lo, i_hi] covers the 8 bins' sample ranges; samples
// outside the lane's own range [i0, i1] are skipped (they would add exact
// zeros, and may lie beyond the zero border).  At every pass the 64 lanes'
// samples form an 8 x 8 block of (bin, sample) points, a rotated square of
// side 8 delta: their bilinear taps fall on a handful of image rows and
// cache lines at every angle.  Four passes' gathers are in flight together;
// each lane sums its samples in order, then the 8 lanes of a bin are added
// in a fixed butterfly order.
__global__ __launch_bounds__(256) void k_fbp_project(const float* __restrict__ pimg, float* __restrict__ sino,
                                                     const float2* __restrict__ cs, const int* __restrict__ ang,
                                                     FbpGeom g) {
    const int lane = threadIdx.x & 63;
    const int kb = lane & 7, ib = lane >> 3;
    const int kbase = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 8;
    const int a = ang[blockIdx.y];
    const int b = blockIdx.z;
    if (kbase >= g.ndet) return;  // uniform per wave
    const int k = min(kbase + kb, g.ndet - 1);
    const float2 t2 = cs[a];
    const float s = ((float)k - g.det_c) * g.delta;
    int i0, i1;
    ray_range(g, s, t2.x, t2.y, i0, i1);
    if (kbase + kb >= g.ndet) i1 = i0 - 1;  // padding lanes: no samples
    const RayCoord rc = ray_coord(g, s, t2.x, t2.y);
    // common range of the wave's 8 bins (every lane holds its own bin's range)
    int lo = i1 >= i0 ? i0 : INT_MAX, hi = i1 >= i0 ? i1 : INT_MIN;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        lo = min(lo, __shfl_xor(lo, off, 64));
        hi = max(hi, __shfl_xor(hi, off, 64));
    }
    const int np = g.n + 2 * kProjPad;
    const float* im = pimg + (size_t)b * np * np;
    float acc = 0.0f;
    if (lo <= hi) {
        int i = lo + ib;
        for (; i + 24 <= hi; i += 32) {
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int iu = i + 8 * u;
                v[u] = 0.0f;
                if (iu >= i0 && iu <= i1) v[u] = ray_sample(im, np, rc, t2.x, t2.y, iu);  // lanes masked: no load
            }
            acc += v[0];
            acc += v[1];
            acc += v[2];
            acc += v[3];
        }
        for (; i <= hi; i += 8)
            if (i >= i0 && i <= i1) acc += ray_sample(im, np, rc, t2.x, t2.y, i);
    }
    acc += __shfl_xor(acc, 8, 64);
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    if (ib == 0 && kbase + kb < g.ndet) sino[((size_t)b * g.nang + a) * g.ndet + kbase + kb] = acc * g.delta;
}

// The filter as a GEMM on the matrix cores: Q = P T over the sinogram rows
// P (rows x ndet), with the Toeplitz T[k'][k] = h[k - k' + ndet - 1], in
// FP32 (v_mfma_f32_16x16x4_f32: fp32 products, fp32 accumulation).  A
// workgroup owns 16 rows and kFiltTiles 16-column output tiles; each of its
// waves keeps kFiltNB tiles in flight, so one A fragment feeds kFiltNB
// MFMAs.  K runs in chunks of up to kFiltKC: the chunk's 16 rows (zero past
// ndet and past the last row) and the window of taps its B fragments read
// (zero outside [0, 2 ndet - 1)) are staged in LDS.  Lane layout: A[i][k]
// at lane i + 16k, B[k][j] at lane j + 16k, D[i][j] at lane j + 16(i/4),
// register i%4.
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kFiltNB = 4;                    // output tiles per wave pass
constexpr int kFiltTiles = 4 * kFiltNB;       // output tiles per workgroup (4 waves)
constexpr int kFiltKC = 768;                  // K chunk (a multiple of 4)
__global__ __launch_bounds__(256) void k_fbp_filter_mfma(const float* __restrict__ sino, float* __restrict__ filt,
                                                         const float* __restrict__ taps, int ndet, int rows,
                                                         int kpad) {
    __shared__ float P[16 * kFiltKC];                 // rows r0.., columns kc..kc+kl
    __shared__ float hw[kFiltTiles * 16 + kFiltKC];   // taps window of the chunk
    const int r0 = blockIdx.x * 16;
    const int c0 = blockIdx.y * kFiltTiles * 16;  // first output column of this workgroup
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kk = lane >> 4;
    f32x4 acc[kFiltNB];
#pragma unroll
    for (int u = 0; u < kFiltNB; u++) acc[u] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    for (int kc = 0; kc < kpad; kc += kFiltKC) {
        const int kl = min(kFiltKC, kpad - kc);
        __syncthreads();  // the previous chunk's reads are done
        for (int e = threadIdx.x; e < 16 * kl; e += 256) {
            const int r = e / kl, k = e - r * kl;
            P[r * kFiltKC + k] = (r0 + r < rows && kc + k < ndet) ? sino[(size_t)(r0 + r) * ndet + kc + k] : 0.0f;
        }
        // column n = c0 + m (m < 16 kFiltTiles) at k' = kc + q (q < kl) needs
        // h[n - k' + ndet - 1], stored at hw[m - q + kl - 1]
        const int hbase = c0 + ndet - kc - kl;
        for (int y = threadIdx.x; y < kFiltTiles * 16 + kl; y += 256) {
            const int idx = hbase + y;
            hw[y] = (idx >= 0 && idx < 2 * ndet - 1) ? taps[idx] : 0.0f;
        }
        __syncthreads();
        const float* pa = P + i * kFiltKC + kk;
        const float* pb = hw + 16 * wave * kFiltNB + i - kk + kl - 1;
        for (int q = 0; q < kl; q += 4) {
            const float a = pa[q];
#pragma unroll
            for (int u = 0; u < kFiltNB; u++)
                acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, pb[16 * u - q], acc[u], 0, 0, 0);
        }
    }
#pragma unroll
    for (int u = 0; u < kFiltNB; u++) {
        const int col = c0 + 16 * (wave * kFiltNB + u) + i;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = r0 + 4 * kk + r;
            if (row < rows && col < ndet) filt[(size_t)row * ndet + col] = acc[u][r];
        }
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
    __shared__ float2 csi[kBpChunk];  // (cos, sin) / delta of each angle of the chunk
    __shared__ int base[kBpChunk];    // window offset: aa * kBpWin - first bin of the window
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
            const float2 ci = make_float2(t2.x * inv, t2.y * inv);
            csi[t] = ci;
            // bin position of the tile centre, as the pixels compute theirs
            const float kc = __builtin_fmaf(xc, ci.x, __builtin_fmaf(yc, ci.y, g.det_c));
            base[t] = t * kBpWin - ((int)floorf(kc) - kBpWin / 2);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < na * kBpWin; t += blockDim.x) {
            const int aa = t / kBpWin;
            const int k = (t % kBpWin) - (base[aa] - aa * kBpWin);
            q[t] = (k >= 0 && k < g.ndet) ? fb[(size_t)(a0 + aa) * g.ndet + k] : 0.0f;
        }
        __syncthreads();
        if (in) {
            for (int aa = 0; aa < na; ++aa) {
                const float2 ci = csi[aa];
                const float kf = __builtin_fmaf(x, ci.x, __builtin_fmaf(y, ci.y, g.det_c));
                const float fk = floorf(kf);
                const float w = kf - fk;
                const int r = (int)fk + base[aa];  // aa * kBpWin + [4, 28): the tile radius is < 12 bins
                const float v0 = q[r];
                const float v1 = q[r + 1];
                acc += __builtin_fmaf(w, v1 - v0, v0);
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
    const unsigned groups = (unsigned)((g.ndet + 31) / 32);  // 4 waves x 8 bins per workgroup
    if (n_bins > 0)
        hipLaunchKernelGGL(k_fbp_project, dim3(groups, n_bins, batch), dim3(256), 0, s, padded, sino, cs, ang_bins, g);
    if (n_rays > 0)
        hipLaunchKernelGGL(k_fbp_project, dim3(groups, n_rays, batch), dim3(256), 0, s, padded, sino, cs, ang_rays, g);
    return hipGetLastError();
}

hipError_t launch_fbp_filter(const float* sino, float* filt, const float* taps, const FbpGeom& g, int batch,
                             hipStream_t s) {
    const int rows = batch * g.nang;
    const int kpad = (g.ndet + 3) & ~3;  // K in steps of 4
    const dim3 grid((unsigned)((rows + 15) / 16), (unsigned)((g.ndet + 16 * kFiltTiles - 1) / (16 * kFiltTiles)));
    hipLaunchKernelGGL(k_fbp_filter_mfma, grid, dim3(256), 0, s, sino, filt, taps, g.ndet, rows, kpad);
    return hipGetLastError();
}

hipError_t launch_fbp_backproject(const float* filt, float* img, const float2* cs, const FbpGeom& g, int batch,
                                  float scale, hipStream_t s) {
    dim3 grid((g.n + 15) / 16, (g.n + 15) / 16, batch);
    hipLaunchKernelGGL(k_fbp_backproject, grid, dim3(256), 0, s, filt, img, cs, g, scale);
    return hipGetLastError();
}

}  // namespace ske
