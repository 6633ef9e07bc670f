// kernels_core.hip -- orientation copies, skew, identity operator, deskew and
// reduce on gfx950 (wave64).  Float arithmetic follows the reference CPU
// path operation by operation (built with -ffp-contract=off), so the
// per-sector and accumulated results are bit-identical to a one-thread run
// of the reference.
#include "device_common.h"
#include "launch.h"

namespace ske {

// ---------------------------------------------------------------------------
// prepare (src/skewEngine.cuh:962-987).  32x32 tiles through LDS so that the
// two transposed copies are written coalesced.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_prepare(const float* __restrict__ in0, float* __restrict__ in1,
                                                 float* __restrict__ in2, float* __restrict__ in3,
                                                 int dimx, int dimy) {
    __shared__ float tile[32][33];
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int i = y0 + k, j = x0 + tx;
        if (i < dimy && j < dimx) {
            const float v = in0[(size_t)i * dimx + j];
            tile[k][tx] = v;
            in3[(size_t)i * dimx + (dimx - 1 - j)] = v;  // in3[dimx*i + cj]
        }
    }
    __syncthreads();
    // in1[dimy*cj + ci] and in2[dimy*j + ci]: output row indexed by x, column by y
    for (int k = ty; k < 32; k += 8) {
        const int j = x0 + k;       // source column -> output row
        const int i = y0 + tx;      // source row -> output column
        if (i < dimy && j < dimx) {
            const float v = tile[tx][k];
            const int ci = dimy - 1 - i, cj = dimx - 1 - j;
            in1[(size_t)dimy * cj + ci] = v;
            in2[(size_t)dimy * j + ci] = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_f32(float* __restrict__ p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_f64(double* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

static unsigned fill_blocks(size_t n) {
    const size_t b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? (b ? b : 1) : 65536);
}

hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3(fill_blocks(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(fill_blocks(n)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_prepare(const float* in0, float* in1, float* in2, float* in3, int dimx, int dimy,
                          hipStream_t s) {
    dim3 grid((dimx + 31) / 32, (dimy + 31) / 32);
    hipLaunchKernelGGL(k_prepare, grid, dim3(256), 0, s, in0, in1, in2, in3, dimx, dimy);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// skew, materialised: every cell of the H x dim_i skewed block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_skew_full(const SecDesc* __restrict__ secs,
                                                   const float* __restrict__ in4,
                                                   const int* __restrict__ target,
                                                   const float* __restrict__ weight,
                                                   float* __restrict__ S) {
    const SecDesc sd = secs[0];
    const int r = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= sd.dim_i) return;
    const float* src = in4 + sd.img + (size_t)sd.type * sd.N;
    S[(size_t)r * sd.dim_i + j] = skew_cell(src, target[sd.tab + j], weight[sd.tab + j], r, j, sd.dim_i, sd.dim_o);
}

hipError_t launch_skew_full(const DevPlan& p, const SecDesc& h, const float* in4, float* S, hipStream_t s) {
    dim3 grid((h.dim_i + 255) / 256, h.H);
    hipLaunchKernelGGL(k_skew_full, grid, dim3(256), 0, s, p.secs, in4, p.target, p.weight, S);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// identity operator (src/kernelUNIT.h:16-23): SO = S on [first, last).
// One wave per row work item, four rows per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_op_unit(const SecDesc* __restrict__ secs, const RowWork* __restrict__ work,
                                                 int nwork, const float* __restrict__ in4,
                                                 const int* __restrict__ target, const float* __restrict__ weight,
                                                 const uint32_t* __restrict__ fl, const float* __restrict__ Sbuf,
                                                 float* __restrict__ arena) {
    const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= nwork) return;
    const int lane = threadIdx.x & 63;
    const RowWork w = work[item];
    const SecDesc sd = secs[w.sec];
    const uint32_t v = fl[sd.rows + w.row];
    const int f = fl_first(v), l = fl_last(v);
    const float* src = in4 + sd.img + (size_t)sd.type * sd.N;
    float* so = arena + sd.so + (size_t)w.row * sd.dim_i;
    for (int j = f + lane; j < l; j += 64) {
        float x;
        if (Sbuf)
            x = Sbuf[(size_t)w.row * sd.dim_i + j];
        else
            x = skew_cell(src, target[sd.tab + j], weight[sd.tab + j], w.row, j, sd.dim_i, sd.dim_o);
        so[j] = x;
    }
}

hipError_t launch_op_unit(const DevPlan& p, const float* in4, const float* Sbuf, float* arena, hipStream_t s) {
    if (p.nwork == 0) return hipSuccess;
    hipLaunchKernelGGL(k_op_unit, dim3((p.nwork + 3) / 4), dim3(256), 0, s, p.secs, p.work, p.nwork, in4,
                       p.target, p.weight, p.fl, Sbuf, arena);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// deskew (src/skewEngine.cuh:755-793) of all sectors of a plan, one thread per
// column of a few oriented output rows, sectors in ascending order per
// orientation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float stdmaxf(float a, float b) { return (a < b) ? b : a; }

// skewOutput cell (r, j) of sector index si as the deskew gather sees it.
template <int kSource>
__device__ __forceinline__ float so_cell(const SecDesc* __restrict__ secs, int si, const SecDesc& sd,
                                         const uint32_t* __restrict__ fl, const float* __restrict__ src,
                                         int r, int j) {
    if (kSource == kDeskewDense) return src[(size_t)r * sd.dim_i + j];  // single reused buffer
    if (r < sd.H) {
        const uint32_t v = fl[sd.rows + r];
        if (j < fl_first(v) || j >= fl_last(v)) return 0.0f;
        if (kSource == kDeskewArena) return src[sd.so + (size_t)r * sd.dim_i + j];
        return src[sd.q + r];
    }
    // r == H happens only when newAngle == 0 (weight 0).  The reference reads
    // the row the previously processed sector left in its reused buffer.
    if (kSource == kDeskewArena && sd.stale >= 0) {
        const SecDesc pv = secs[sd.stale];
        if (r < pv.H && pv.dim_i == sd.dim_i) {
            const uint32_t v = fl[pv.rows + r];
            if (j >= fl_first(v) && j < fl_last(v)) return src[pv.so + (size_t)r * sd.dim_i + j];
        }
    }
    return 0.0f;
}

// Tap (r, j) of sector si with the common case r < H read without branches:
// the row index is clamped into the sector (in bounds of its H x dim_i slot)
// and the value masked by the row's first/last range afterwards.  Rows r >= H
// are patched by the caller through so_cell.
template <int kSource>
__device__ __forceinline__ float so_tap(const SecDesc& sd, const uint32_t* __restrict__ fl,
                                        const float* __restrict__ src, int r, int j) {
    if (kSource == kDeskewDense) return src[(size_t)r * sd.dim_i + j];
    const int rc = min(r, sd.H - 1);
    const uint32_t v = fl[sd.rows + rc];
    const float x = (kSource == kDeskewArena) ? src[sd.so + (size_t)rc * sd.dim_i + j] : src[sd.q + rc];
    return (j >= fl_first(v) && j < fl_last(v)) ? x : 0.0f;
}

template <int kMax>
__device__ __forceinline__ float deskew_acc(float O, float a, float b, float wj, int angle) {
    if (kMax) {
        O = stdmaxf(O, stdmaxf(a, b));
        return (float)(angle + 1000 * (int)(O / 1000));
    }
    return O + ((1.0f - wj) * a + wj * b);
}

// One output cell over sectors [sb, se), in ascending order.
template <int kSource, int kMax>
__device__ void deskew_cell(const SecDesc* __restrict__ secs, const int* __restrict__ target,
                            const float* __restrict__ weight, const uint32_t* __restrict__ fl,
                            const float* __restrict__ src, int sb, int se, int i, int j, float* o) {
    float O = *o;
    for (int si = sb; si < se; ++si) {
        const SecDesc& sd = secs[si];
        const int tj = target[sd.tab + j];
        const float wj = weight[sd.tab + j];
        const float a = so_cell<kSource>(secs, si, sd, fl, src, i + tj, j);
        const float b = so_cell<kSource>(secs, si, sd, fl, src, i + tj + 1, j);
        O = deskew_acc<kMax>(O, a, b, wj, sd.angle);
    }
    *o = O;
}

constexpr int kDeskewRows = 16;  // output rows per thread

// kRows consecutive output rows per thread: the kRows + 1 taps of a sector
// are shared by neighbouring rows.  kU sectors per pass: their table reads,
// then their gathers are issued together, and each row accumulates in
// ascending sector order afterwards.  Orientation t of one image: sectors
// [sb, se), accumulators acc (four dimx*dimy buffers).
template <int kSource, int kMax, int kRows, int kU>
__device__ __forceinline__ void deskew_body(const SecDesc* __restrict__ secs, const int* __restrict__ target,
                                            const float* __restrict__ weight, const uint32_t* __restrict__ fl,
                                            const float* __restrict__ src, int t, int sb, int se,
                                            float* __restrict__ acc, int dimx, int dimy) {
    if (sb >= se) return;
    const bool tr = (t == 1 || t == 2);
    const int dim_i = tr ? dimy : dimx, dim_o = tr ? dimx : dimy;
    const int i0 = blockIdx.y * kRows;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 >= dim_o || j >= dim_i) return;
    const size_t N = (size_t)dimx * dimy;
    float* o = acc + (size_t)t * N + (size_t)i0 * dim_i + j;
    if (i0 + kRows > dim_o) {  // last rows of the image
        for (int k = 0; k < dim_o - i0; ++k)
            deskew_cell<kSource, kMax>(secs, target, weight, fl, src, sb, se, i0 + k, j, o + (size_t)k * dim_i);
        return;
    }
    float O[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) O[k] = o[(size_t)k * dim_i];
    int si = sb;
    for (; si + kU <= se; si += kU) {
        int tj[kU];
        float wj[kU], v[kU][kRows + 1];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const SecDesc& sd = secs[si + u];
            tj[u] = target[sd.tab + j];
            wj[u] = weight[sd.tab + j];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
#pragma unroll
            for (int m = 0; m <= kRows; ++m) v[u][m] = so_tap<kSource>(secs[si + u], fl, src, i0 + tj[u] + m, j);
        }
        if (kSource != kDeskewDense) {
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const SecDesc& sd = secs[si + u];
                if (i0 + tj[u] + kRows >= sd.H) {
#pragma unroll
                    for (int m = 0; m <= kRows; ++m)
                        if (i0 + tj[u] + m >= sd.H) v[u][m] = so_cell<kSource>(secs, si + u, sd, fl, src, i0 + tj[u] + m, j);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int angle = secs[si + u].angle;
#pragma unroll
            for (int k = 0; k < kRows; ++k) O[k] = deskew_acc<kMax>(O[k], v[u][k], v[u][k + 1], wj[u], angle);
        }
    }
    for (; si < se; ++si) {
        const SecDesc& sd = secs[si];
        const int tj = target[sd.tab + j];
        const float wj = weight[sd.tab + j];
        float v[kRows + 1];
#pragma unroll
        for (int m = 0; m <= kRows; ++m) v[m] = so_cell<kSource>(secs, si, sd, fl, src, i0 + tj + m, j);
#pragma unroll
        for (int k = 0; k < kRows; ++k) O[k] = deskew_acc<kMax>(O[k], v[k], v[k + 1], wj, sd.angle);
    }
#pragma unroll
    for (int k = 0; k < kRows; ++k) o[(size_t)k * dim_i] = O[k];
}

__device__ __forceinline__ int range_begin(int t, int4 lo, int4 hi) {
    return (t == 0) ? lo.x : (t == 1) ? lo.z : (t == 2) ? hi.x : hi.z;
}
__device__ __forceinline__ int range_end(int t, int4 lo, int4 hi) {
    return (t == 0) ? lo.y : (t == 1) ? lo.w : (t == 2) ? hi.y : hi.w;
}

// One image: blockIdx.z = orientation.
template <int kSource, int kMax, int kRows, int kU>
__global__ __launch_bounds__(256) void k_deskew(const SecDesc* __restrict__ secs, const int* __restrict__ target,
                                                const float* __restrict__ weight, const uint32_t* __restrict__ fl,
                                                const float* __restrict__ src, int4 rlo, int4 rhi,
                                                float* __restrict__ acc, int dimx, int dimy) {
    const int t = blockIdx.z;
    deskew_body<kSource, kMax, kRows, kU>(secs, target, weight, fl, src, t, range_begin(t, rlo, rhi),
                                          range_end(t, rlo, rhi), acc, dimx, dimy);
}

// Every image of a chunk in one launch: blockIdx.z = 4 * part + orientation,
// so the images' cells run side by side instead of one launch after another.
template <int kSource, int kMax, int kRows, int kU>
__global__ __launch_bounds__(256) void k_deskew_parts(const SecDesc* __restrict__ secs,
                                                      const int* __restrict__ target,
                                                      const float* __restrict__ weight,
                                                      const uint32_t* __restrict__ fl, const float* __restrict__ src,
                                                      const DeskewPart* __restrict__ parts, float* __restrict__ acc) {
    const DeskewPart& P = parts[blockIdx.z >> 2];
    const int t = blockIdx.z & 3;
    deskew_body<kSource, kMax, kRows, kU>(secs, target, weight, fl, src, t, P.range[2 * t], P.range[2 * t + 1],
                                          acc + P.acc_off, P.dimx, P.dimy);
}

// The Radon deskew with the taps staged in LDS.  A Radon skewOutput row r
// holds q[r] on [first_r, last_r), so a sector's taps for a workgroup's 256
// columns j0 .. j0+255 and rows i0 .. i0+15 are q and first|last of the rows
// i0 + t_j .. i0 + t_j + 16: since t is nondecreasing with steps of 0 or 1,
// the rows i0 + t_{j0} .. i0 + t_{jlast} + 16, at most 272.  A pass stages
// them for kRadonU sectors (rows past H as empty rows), then every thread
// reads its 17 taps per sector from LDS: the values so_cell gives, added in
// ascending sector order, so the result is the k_deskew one bit for bit.
constexpr int kRadonU = 8, kRadonStage = 288;  // 18 KB of LDS: eight sectors per pass
__device__ __forceinline__ void deskew_radon_body(const SecDesc* __restrict__ secs, const int* __restrict__ target,
                                                  const float* __restrict__ weight,
                                                  const uint32_t* __restrict__ fl, const float* __restrict__ q,
                                                  int t, int sb, int se, float* __restrict__ acc, int dimx,
                                                  int dimy) {
    constexpr int kRows = kDeskewRows;
    __shared__ float sq[kRadonU * kRadonStage];
    __shared__ uint32_t sfl[kRadonU * kRadonStage];
    if (sb >= se) return;  // uniform
    const bool tr = (t == 1 || t == 2);
    const int dim_i = tr ? dimy : dimx, dim_o = tr ? dimx : dimy;
    const int i0 = blockIdx.y * kRows;
    const int j0 = blockIdx.x * blockDim.x;
    if (i0 >= dim_o || j0 >= dim_i) return;  // uniform
    const int j = min(j0 + (int)threadIdx.x, dim_i - 1);
    const bool col = j0 + (int)threadIdx.x < dim_i;
    const int jlast = min(j0 + (int)blockDim.x - 1, dim_i - 1);
    const size_t N = (size_t)dimx * dimy;
    float* o = acc + (size_t)t * N + (size_t)i0 * dim_i + j;
    const int nrow = min(kRows, dim_o - i0);
    float O[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) O[k] = (col && k < nrow) ? o[(size_t)k * dim_i] : 0.0f;
    for (int si = sb; si < se; si += kRadonU) {
        const int nu = min(kRadonU, se - si);
        __syncthreads();  // the previous pass's reads are done
        for (int e = threadIdx.x; e < nu * kRadonStage; e += blockDim.x) {
            const int u = e / kRadonStage, x = e - u * kRadonStage;
            const SecDesc& sd = secs[si + u];
            const int rlo = i0 + target[sd.tab + j0];
            const int len = min(target[sd.tab + jlast] - target[sd.tab + j0] + kRows + 1, kRadonStage);
            const int r = rlo + x;
            const bool in = x < len && r < sd.H;
            sq[e] = in ? q[sd.q + r] : 0.0f;
            sfl[e] = in ? fl[sd.rows + r] : 0u;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kRadonU; ++u) {
            if (u >= nu) break;  // uniform
            const SecDesc& sd = secs[si + u];
            const int tj = target[sd.tab + j];
            const float wj = weight[sd.tab + j];
            // stage index of row i0 + tj (tj - t_{j0} <= j - j0 <= 255; the clamp only
            // guards the stage's bounds)
            const int base = u * kRadonStage + min(tj - target[sd.tab + j0], kRadonStage - kRows - 1);
            float v[kRows + 1];
#pragma unroll
            for (int m = 0; m <= kRows; ++m) {
                const uint32_t f = sfl[base + m];
                v[m] = (j >= fl_first(f) && j < fl_last(f)) ? sq[base + m] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < kRows; ++k) O[k] = deskew_acc<0>(O[k], v[k], v[k + 1], wj, sd.angle);
        }
    }
    if (col) {
#pragma unroll
        for (int k = 0; k < kRows; ++k)
            if (k < nrow) o[(size_t)k * dim_i] = O[k];
    }
}

__global__ __launch_bounds__(256) void k_deskew_radon(const SecDesc* __restrict__ secs,
                                                      const int* __restrict__ target,
                                                      const float* __restrict__ weight,
                                                      const uint32_t* __restrict__ fl, const float* __restrict__ q,
                                                      int4 rlo, int4 rhi, float* __restrict__ acc, int dimx,
                                                      int dimy) {
    const int t = blockIdx.z;
    deskew_radon_body(secs, target, weight, fl, q, t, range_begin(t, rlo, rhi), range_end(t, rlo, rhi), acc, dimx,
                      dimy);
}

__global__ __launch_bounds__(256) void k_deskew_radon_parts(const SecDesc* __restrict__ secs,
                                                            const int* __restrict__ target,
                                                            const float* __restrict__ weight,
                                                            const uint32_t* __restrict__ fl,
                                                            const float* __restrict__ q,
                                                            const DeskewPart* __restrict__ parts,
                                                            float* __restrict__ acc) {
    const DeskewPart& P = parts[blockIdx.z >> 2];
    const int t = blockIdx.z & 3;
    deskew_radon_body(secs, target, weight, fl, q, t, P.range[2 * t], P.range[2 * t + 1], acc + P.acc_off, P.dimx,
                      P.dimy);
}

// 16 rows, one sector per pass: measured fastest on the batched viewshed
// (129 VGPRs; 8 rows 1.2x, 4 rows 1.9x, 1 row 4.1x slower, 24 or 32 rows
// 1.45x slower at 151 / 191 VGPRs; a second sector per pass only adds
// registers).  The Radon deskew reads a short vector q per sector (L1/L2
// resident), so its time is the chain of dependent table and tap reads over
// the sectors: four sectors per pass.
constexpr int kDeskewU = 1, kDeskewRadonU = 4;

hipError_t launch_deskew(const DevPlan& p, const int range[8], const float* src, int source, int is_max,
                         float* acc, int dimx, int dimy, long long img_acc, hipStream_t s) {
    const int dl = dimx > dimy ? dimx : dimy;
    dim3 grid((dl + 255) / 256, (dl + kDeskewRows - 1) / kDeskewRows, 4);
    int4 lo = make_int4(range[0], range[1], range[2], range[3]);
    int4 hi = make_int4(range[4], range[5], range[6], range[7]);
    float* a = acc + img_acc;
#define SKE_DESKEW(SRC, MX)                                                                                  \
    hipLaunchKernelGGL((k_deskew<SRC, MX, kDeskewRows, kDeskewU>), grid, dim3(256), 0, s, p.secs, p.target, \
                       p.weight, p.fl, src, lo, hi, a, dimx, dimy)
    if (source == kDeskewDense) {
        if (is_max) SKE_DESKEW(kDeskewDense, 1); else SKE_DESKEW(kDeskewDense, 0);
    } else if (source == kDeskewArena) {
        if (is_max) SKE_DESKEW(kDeskewArena, 1); else SKE_DESKEW(kDeskewArena, 0);
    } else if ((size_t)grid.x * grid.y * grid.z >= 1024) {
        hipLaunchKernelGGL(k_deskew_radon, grid, dim3(256), 0, s, p.secs, p.target, p.weight, p.fl, src, lo, hi, a,
                           dimx, dimy);
    } else {
        hipLaunchKernelGGL((k_deskew<kDeskewRadon, 0, kDeskewRows, kDeskewRadonU>), grid, dim3(256), 0, s, p.secs,
                           p.target, p.weight, p.fl, src, lo, hi, a, dimx, dimy);
    }
#undef SKE_DESKEW
    return hipGetLastError();
}

hipError_t launch_deskew_parts(const DevPlan& p, const DeskewPart* parts, int nparts, int max_dim, const float* src,
                               int source, int is_max, float* acc, hipStream_t s) {
    if (nparts <= 0) return hipSuccess;
    if (nparts > 16383) return hipErrorInvalidValue;
    dim3 grid((max_dim + 255) / 256, (max_dim + kDeskewRows - 1) / kDeskewRows, 4 * nparts);
#define SKE_DESKEW_P(SRC, MX, U)                                                                                  \
    hipLaunchKernelGGL((k_deskew_parts<SRC, MX, kDeskewRows, U>), grid, dim3(256), 0, s, p.secs, p.target, p.weight, \
                       p.fl, src, parts, acc)
    if (source == kDeskewArena) {
        if (is_max) SKE_DESKEW_P(kDeskewArena, 1, kDeskewU); else SKE_DESKEW_P(kDeskewArena, 0, kDeskewU);
    } else if (source == kDeskewRadon) {
        // staged taps pay off with four or more workgroups per CU (a batch of
        // images); a single small image runs its sector chains with four
        // sectors' gathers in flight instead (one 512^2 image: 50 us against
        // 87 us staged)
        if ((size_t)grid.x * grid.y * grid.z >= 1024)
            hipLaunchKernelGGL(k_deskew_radon_parts, grid, dim3(256), 0, s, p.secs, p.target, p.weight, p.fl, src,
                               parts, acc);
        else
            SKE_DESKEW_P(kDeskewRadon, 0, kDeskewRadonU);
    } else {
        return hipErrorInvalidValue;  // the dense source has one image per call
    }
#undef SKE_DESKEW_P
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// reduce (src/skewEngine.cuh:885-912).  out[i,j] (+)= o0[i,j], o1[cj,ci],
// o2[j,ci], o3[i,cj] in that order.  o1/o2 are read through an LDS tile so
// that their transposed accesses stay coalesced.
// ---------------------------------------------------------------------------
template <int kMax>
__global__ __launch_bounds__(256) void k_reduce(const float* __restrict__ acc, int4 present,
                                                float* __restrict__ out, int dimx, int dimy) {
    __shared__ float t1[32][33];
    __shared__ float t2[32][33];
    const size_t N = (size_t)dimx * dimy;
    const float* o0 = acc;
    const float* o1 = acc + N;
    const float* o2 = acc + 2 * N;
    const float* o3 = acc + 3 * N;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // o1 rows cj = dimx-1-j (j in [x0, x0+32)), columns ci = dimy-1-i;
    // o2 rows j, columns ci.  Load both tiles with ci along threadIdx.x.
    for (int k = ty; k < 32; k += 8) {
        const int j = x0 + k;
        const int i = y0 + 31 - tx;  // ci = dimy-1-i runs forward with tx
        if (j < dimx && i >= 0 && i < dimy) {
            const int ci = dimy - 1 - i, cj = dimx - 1 - j;
            if (present.y) t1[k][31 - tx] = o1[(size_t)cj * dimy + ci];
            if (present.z) t2[k][31 - tx] = o2[(size_t)j * dimy + ci];
        }
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int i = y0 + k, j = x0 + tx;
        if (i >= dimy || j >= dimx) continue;
        const size_t oi = (size_t)i * dimx + j;
        float v = out[oi];
        const int cj = dimx - 1 - j;
        if (kMax) {
            if (present.x) { const float u = o0[oi]; v = (v > u) ? v : u; }
            if (present.y) { const float u = t1[tx][k]; v = (v > u) ? v : u; }
            if (present.z) { const float u = t2[tx][k]; v = (v > u) ? v : u; }
            if (present.w) { const float u = o3[(size_t)i * dimx + cj]; v = (v > u) ? v : u; }
        } else {
            if (present.x) v += o0[oi];
            if (present.y) v += t1[tx][k];
            if (present.z) v += t2[tx][k];
            if (present.w) v += o3[(size_t)i * dimx + cj];
        }
        out[oi] = v;
    }
}

hipError_t launch_reduce(const float* acc, const int present[4], float* out, int dimx, int dimy, int is_max,
                         hipStream_t s) {
    dim3 grid((dimx + 31) / 32, (dimy + 31) / 32);
    int4 pr = make_int4(present[0], present[1], present[2], present[3]);
    if (is_max)
        hipLaunchKernelGGL(k_reduce<1>, grid, dim3(256), 0, s, acc, pr, out, dimx, dimy);
    else
        hipLaunchKernelGGL(k_reduce<0>, grid, dim3(256), 0, s, acc, pr, out, dimx, dimy);
    return hipGetLastError();
}

}  // namespace ske
