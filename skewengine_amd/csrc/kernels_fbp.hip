// kernels_fbp.hip -- standard parallel-beam geometry (no reference code; the
// counterpart of the reference's skew-space Radon, src/kernelRADN.h:212-259,
// in the 725-bin geometry of the benchmark configurations):
//
//   image   n x n on [-1,1]^2, pixel pitch delta = 2/n, row 0 at y = +1;
//   angles  theta_a (host table of cos/sin);
//   bins    s_k = (k - (ndet-1)/2) * delta.
//
//   project     p(a,k) = delta * sum_t f(s_k cos - t sin, s_k sin + t cos),
//               t on a delta grid of nt = ndet samples, f bilinear (zero
//               outside); the image is cut into tiles staged in LDS, each
//               tile's share of every ray summed there, the shares added in
//               tile order;
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

// Sample i of the ray of bin offset s at angle (c, sn) in pixel coordinates
// (u: column, v: row): with t = (i - half) delta,
// u = (s c - t sn + 1)/delta - 1/2 = U0 - i sn and
// v = (1 - s sn - t c)/delta - 1/2 = V0 - i c, so a sample costs two fma.
// Its bilinear base pixel is (floor v, floor u); it reads (v, u) .. (v+1,
// u+1), zero outside the image.  In exact arithmetic, with P = k - det_c,
// Q = half - i and m = n/2 - 1/2: u - m = P c + Q sn, v - m = -P sn + Q c.
struct RayCoord {
    float u0, v0;
};
// inv = 1 / delta (IEEE division)
__device__ __forceinline__ RayCoord ray_coord(const FbpGeom& g, float s, float c, float sn, float inv) {
    const float half = 0.5f * (float)(g.nt - 1);
    return RayCoord{((s * c + 1.0f) * inv - 0.5f) + half * sn, ((1.0f - s * sn) * inv - 0.5f) + half * c};
}

// The projector in LDS tiles.  A sample contributes only if one of its taps
// is an image pixel, i.e. its base pixel lies in [-1, n)^2; that square is
// cut into ntx x ntx tiles of side ts (ProjTiling).  A workgroup stages one
// tile of one slice -- base pixels and their right/lower neighbours, (ts+1)^2
// values, zero outside the image -- in LDS once, then for each angle of its
// chunk sums, per bin, the samples whose base pixel lies in the tile: every
// sample is counted in exactly one tile, with the value the direct sampling
// gives (same coordinates, same taps, same lerp).  The partial sums go to
// part[slice][angle][tile][k - klo]; k_fbp_project_sum adds each ray's
// partials in tile order.  Bins and samples a tile can hold are bounded with
// margins (tile_krange, sample_range) and every sample is tested exactly.
constexpr int kPtPitch = 136;  // LDS row pitch (floats) of a staged tile, >= kPtMax + 1
constexpr int kPtWaves = 8;
constexpr int kPtKB = 16;               // bins per work item (lanes of a pass along the bins)
constexpr int kPtIB = 64 / kPtKB;      // samples per pass along the ray

// Bins [klo, khi] whose rays can sample a base pixel of [x0, x1) x [y0, y1):
// P = (u - m) c - (v - m) sn over the corners, two bins of margin.
__device__ __forceinline__ void tile_krange(const FbpGeom& g, float c, float sn, int x0, int x1, int y0, int y1,
                                            int& klo, int& khi) {
    const float m = 0.5f * (float)g.n - 0.5f;
    const float ua = (float)x0 - m, ub = (float)x1 - m, va = (float)y0 - m, vb = (float)y1 - m;
    const float p0 = ua * c - va * sn, p1 = ub * c - va * sn, p2 = ua * c - vb * sn, p3 = ub * c - vb * sn;
    const float pmin = fminf(fminf(p0, p1), fminf(p2, p3)), pmax = fmaxf(fmaxf(p0, p1), fmaxf(p2, p3));
    klo = max((int)floorf(pmin + g.det_c) - 2, 0);
    khi = min((int)ceilf(pmax + g.det_c) + 2, g.ndet - 1);
}

// Samples [ilo, ihi] of bin offset P whose base pixel can lie in the tile:
// the line's chord through the tile's square widened by one pixel (Q from
// both slabs, rs = 1/sn and rc = 1/c rounded), two samples of margin; empty:
// ihi < ilo.
__device__ __forceinline__ void sample_range(const FbpGeom& g, float c, float sn, float rc, float rs, float P, int x0,
                                             int x1, int y0, int y1, int& ilo, int& ihi) {
    const float m = 0.5f * (float)g.n - 0.5f;
    const float X0 = (float)x0 - m - 1.0f, X1 = (float)x1 - m + 1.0f;
    const float Y0 = (float)y0 - m - 1.0f, Y1 = (float)y1 - m + 1.0f;
    float qlo = -1e30f, qhi = 1e30f;
    const float pc = P * c, ps = P * sn;
    if (fabsf(sn) > 1e-6f) {  // u - m = P c + Q sn in [X0, X1]
        const float q1 = (X0 - pc) * rs, q2 = (X1 - pc) * rs;
        qlo = fmaxf(qlo, fminf(q1, q2));
        qhi = fminf(qhi, fmaxf(q1, q2));
    } else if (pc < X0 || pc > X1) {
        qhi = qlo - 1.0f;
    }
    if (fabsf(c) > 1e-6f) {  // v - m = -P sn + Q c in [Y0, Y1]
        const float q1 = (Y0 + ps) * rc, q2 = (Y1 + ps) * rc;
        qlo = fmaxf(qlo, fminf(q1, q2));
        qhi = fminf(qhi, fmaxf(q1, q2));
    } else if (-ps < Y0 || -ps > Y1) {
        qhi = qlo - 1.0f;
    }
    ilo = 0;
    ihi = -1;
    if (qhi >= qlo) {
        const float half = 0.5f * (float)(g.nt - 1);
        ilo = max((int)floorf(half - qhi) - 2, 0);
        ihi = min((int)ceilf(half - qlo) + 2, g.nt - 1);
    }
}

// One workgroup per (tile, chunk of kPtAngles angles, slice).  Work items are
// (angle, block of kPtKB bins) of the chunk, taken by the waves in turn;
// lane (kb, ib) = (lane % kPtKB, lane / kPtKB) sums the samples ilo + ib,
// ilo + ib + kPtIB, ... of bin k0 + kb over the union of the block's sample
// ranges (a kPtKB x kPtIB block of (bin, sample) points a pass: a compact
// rotated patch of base pixels), four passes in flight; then the kPtIB lanes
// of a bin are added in a fixed butterfly order.
__global__ __launch_bounds__(64 * kPtWaves) void k_fbp_project_tiles(const float* __restrict__ img,
                                                                      float* __restrict__ part,
                                                                      const float2* __restrict__ cs, FbpGeom g,
                                                                      ProjTiling pt) {
    __shared__ float L[(kPtMax + 1) * kPtPitch];
    const int T = blockIdx.x, b = blockIdx.z;
    const int x0 = -1 + (T % pt.ntx) * pt.ts, y0 = -1 + (T / pt.ntx) * pt.ts;
    const int x1 = min(x0 + pt.ts, g.n), y1 = min(y0 + pt.ts, g.n);  // base pixels [x0, x1) x [y0, y1)
    const float* im = img + (size_t)b * g.n * g.n;
    const int side = pt.ts + 1;
    for (int e = threadIdx.x; e < side * side; e += blockDim.x) {
        const int r = e / side, q = e - r * side;
        const int y = y0 + r, x = x0 + q;
        L[r * kPtPitch + q] = (y >= 0 && y < g.n && x >= 0 && x < g.n) ? im[(size_t)y * g.n + x] : 0.0f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, kb = lane % kPtKB, ib = lane / kPtKB;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int a0 = blockIdx.y * kPtAngles, na = min(kPtAngles, g.nang - a0);
    const float inv = 1.0f / g.delta;
    // base pixels of the tile as floats, for clamping and the membership test
    const float fx0 = (float)x0, fx1 = (float)(x1 - 1), fy0 = (float)y0, fy1 = (float)(y1 - 1);
    const int lt = -(y0 * kPtPitch + x0);  // L[lt + y * kPtPitch + x]: image pixel (y, x) of the tile
    int item = wave;  // flat index over (angle, bin block) pairs of the chunk
    int base = 0;     // items of the angles before aa
    for (int aa = 0; aa < na; ++aa) {
        const int a = a0 + aa;
        const float2 t2 = cs[a];
        const float c = t2.x, sn = t2.y;
        const float rc = 1.0f / c, rs = 1.0f / sn;  // used only where |c|, |sn| > 1e-6
        int klo, khi;
        tile_krange(g, c, sn, x0, x1, y0, y1, klo, khi);
        const int nkb = khi >= klo ? (khi - klo + kPtKB) / kPtKB : 0;
        float* const out = part + (((size_t)b * g.nang + a) * pt.ntx * pt.ntx + T) * pt.kw;
        for (; item < base + nkb; item += kPtWaves) {
            const int k = klo + kPtKB * (item - base) + kb;
            const bool kin = k <= khi;
            const float P = (float)k - g.det_c;
            int i0, i1;
            sample_range(g, c, sn, rc, rs, P, x0, x1, y0, y1, i0, i1);
            if (!kin) i1 = i0 - 1;
            int lo = i1 >= i0 ? i0 : INT_MAX, hi = i1 >= i0 ? i1 : INT_MIN;
#pragma unroll
            for (int off = 1; off < kPtKB; off <<= 1) {
                lo = min(lo, __shfl_xor(lo, off, 64));
                hi = max(hi, __shfl_xor(hi, off, 64));
            }
            const RayCoord ro = ray_coord(g, P * g.delta, c, sn, inv);
            // branch-free, so that the reads of several samples are in flight
            // together: a sample outside the tile reads a clamped in-tile
            // address and is dropped
            auto sample = [&](float fi) -> float {
                const float u = __builtin_fmaf(-fi, sn, ro.u0), v = __builtin_fmaf(-fi, c, ro.v0);
                const float fu = floorf(u), fv = floorf(v);
                const float cu = __builtin_amdgcn_fmed3f(fu, fx0, fx1), cv = __builtin_amdgcn_fmed3f(fv, fy0, fy1);
                const bool in = kin && cu == fu && cv == fv;
                // cv * pitch + cu is an integer below 2^24: exact in float
                const float* r = L + (lt + (int)__builtin_fmaf(cv, (float)kPtPitch, cu));
                const float wu = u - fu, wv = v - fv;
                const float ta = r[0], tb = r[1], tc = r[kPtPitch], td = r[kPtPitch + 1];
                const float top = __builtin_fmaf(wu, tb - ta, ta);
                const float bot = __builtin_fmaf(wu, td - tc, tc);
                const float val = __builtin_fmaf(wv, bot - top, top);
                return in ? val : 0.0f;
            };
            float acc = 0.0f;
            if (lo <= hi) {
                int i = lo + ib;
                float fi = (float)i;  // i, exactly (below 2^24)
                for (; i + 3 * kPtIB <= hi; i += 4 * kPtIB, fi += (float)(4 * kPtIB)) {
                    float v[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) v[u] = sample(fi + (float)(kPtIB * u));
                    acc += v[0];
                    acc += v[1];
                    acc += v[2];
                    acc += v[3];
                }
                for (; i <= hi; i += kPtIB, fi += (float)kPtIB) acc += sample(fi);
            }
#pragma unroll
            for (int off = kPtKB; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
            if (ib == 0 && kin) out[k - klo] = acc;
        }
        base += nkb;
    }
}

// sino(a, k) = delta * the sum of the partials of the tiles its ray crosses,
// in tile order.
__global__ __launch_bounds__(256) void k_fbp_project_sum(const float* __restrict__ part, float* __restrict__ sino,
                                                         const float2* __restrict__ cs, FbpGeom g, ProjTiling pt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y, b = blockIdx.z;
    if (k >= g.ndet) return;
    const float2 t2 = cs[a];
    const int nt2 = pt.ntx * pt.ntx;
    const float* const p = part + ((size_t)b * g.nang + a) * nt2 * pt.kw;
    float acc = 0.0f;
    for (int T = 0; T < nt2; ++T) {
        const int x0 = -1 + (T % pt.ntx) * pt.ts, y0 = -1 + (T / pt.ntx) * pt.ts;
        int klo, khi;
        tile_krange(g, t2.x, t2.y, x0, min(x0 + pt.ts, g.n), y0, min(y0 + pt.ts, g.n), klo, khi);
        if (k >= klo && k <= khi) acc += p[(size_t)T * pt.kw + (k - klo)];
    }
    sino[((size_t)b * g.nang + a) * g.ndet + k] = acc * g.delta;
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

hipError_t launch_fbp_project(const float* img, float* part, float* sino, const float2* cs, const FbpGeom& g,
                              const ProjTiling& pt, int batch, hipStream_t s) {
    if (pt.ts < 1 || pt.ts > kPtMax || pt.ntx * pt.ts < g.n + 1) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(pt.ntx * pt.ntx), (unsigned)((g.nang + kPtAngles - 1) / kPtAngles), (unsigned)batch);
    hipLaunchKernelGGL(k_fbp_project_tiles, grid, dim3(64 * kPtWaves), 0, s, img, part, cs, g, pt);
    hipLaunchKernelGGL(k_fbp_project_sum, dim3((unsigned)((g.ndet + 255) / 256), (unsigned)g.nang, (unsigned)batch),
                       dim3(256), 0, s, part, sino, cs, g, pt);
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
