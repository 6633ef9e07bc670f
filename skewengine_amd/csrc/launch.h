// launch.h -- host-side launchers of the gfx950 kernels (kernels_*.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ske_types.h"

namespace ske {

// Device view of a planned set of sectors (one image or a batch).
struct DevPlan {
    const SecDesc* secs = nullptr;
    int nsec = 0;
    const RowWork* work = nullptr;  // rows with at least one valid cell
    int nwork = 0;
    const int* target = nullptr;
    const float* weight = nullptr;
    const uint32_t* fl = nullptr;  // first | last << 16 per row
    int maxL = 0;                   // longest valid row (LDS sizing)
    int maxH = 0;                   // tallest skewed block (rows per sector)
};

enum DeskewSource { kDeskewDense = 0, kDeskewArena = 1, kDeskewRadon = 2 };

// Launch shape of the viewshed operator, chosen by the autotuner.
struct SdemTuning {
    int waves = 4;                 // waves per workgroup (observer chunks of one row in parallel): 2, 4 or 8
    int persistent_wgs_per_cu = 0; // 0: one workgroup per row; >0: persistent grid of this many per CU
    int occupancy = 0;             // persistent grid: 8 = compile for 8 waves per SIMD (fewer VGPRs)
    int group = 4;                 // 1: one 64-observer chunk per wave; 4: groups of 4 chunks, lane compaction
};

// Graph-safe clears (kernels, not hipMemsetAsync).
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s);
hipError_t launch_fill_f64(double* p, size_t n, hipStream_t s);

// prepare(): the three mirrored/transposed copies (src/skewEngine.cuh:962-987).
hipError_t launch_prepare(const float* in0, float* in1, float* in2, float* in3, int dimx, int dimy,
                          hipStream_t s);

// Materialise the skewInput rows (H*dim_i) of one sector; cells without a
// tap are zero, as after resetSkew + skew (src/skewEngine.cuh:567,612-621).
hipError_t launch_skew_full(const DevPlan& p, const SecDesc& h, const float* in4, float* S,
                            hipStream_t s);

// Operators.  `Sbuf` == nullptr gathers rows from the oriented inputs
// (fused path); otherwise rows come from a materialised skewInput.
hipError_t launch_op_unit(const DevPlan& p, const float* in4, const float* Sbuf, float* arena,
                          hipStream_t s);
hipError_t launch_op_sdem(const DevPlan& p, const float* in4, const float* Sbuf, float* arena,
                          const float4* rtab, float povh, int variant, hipStream_t s);
// Bit-block viewshed with an explicit launch shape; `counter` (one int of
// device memory) feeds the persistent grid.
// `steps`, when not null, receives += the lane-steps the scans executed
// (32-step blocks x 64 lanes; the early exit skips the rest).
hipError_t launch_op_sdem_tuned(const DevPlan& p, const float* in4, const float* Sbuf, float* arena,
                                const float4* rtab, float povh, const SdemTuning& t, int* counter,
                                unsigned long long* steps, hipStream_t s);
// tab: cos/sin(2 pi t/64) interleaved (128 doubles) then cos(2 pi t/32) (32).
hipError_t launch_op_blur(const DevPlan& p, const float* in4, const float* Sbuf, float* arena,
                          const double* tab, hipStream_t s);
hipError_t launch_radn_rowsum(const DevPlan& p, const float* in4, const float* Sbuf, double* P,
                              hipStream_t s);
// Per-sector ramp filter of the row sums in LDS (FFT length H2 <= 2^15).
hipError_t launch_radn_filter(const DevPlan& p, const double* P, float* q, const double* tw,
                              const int* tw_off, int maxH2, hipStream_t s);

// Local Radon transform (LRT=true): window sums of every skewed row into W
// (doubles at the arena offsets of the rows, zero-filled beforehand by the
// caller), then the per-cell transform into the arena.  tab as for the
// blur operator plus the sine ramp at [160,186); gauss: 64 weights per
// folded angle 0..90.
hipError_t launch_lrt(const DevPlan& p, const float* in4, const float* Sbuf, double* W, float* arena,
                      const double* tab, const double* gauss, hipStream_t s);

// The four 64-float vectors of the local transform at `nreq` sample cells
// (after launch_lrt filled W for the same plan) into out[slot * 256 ...].
hipError_t launch_lrt_samples(const DevPlan& p, const SampleReq* req, int nreq, const float* in4, const double* W,
                              const double* tab, const double* gauss, float* out, hipStream_t s);

// q -> skewOutput rows on [first, last) (materialised per-sector path).
hipError_t launch_radn_smear(const DevPlan& p, const float* q, float* SO, hipStream_t s);

// Deskew sectors [first of each type ... ] of a plan into the four
// orientation accumulators acc[4*N] (src/skewEngine.cuh:755-793), in
// ascending sector order per orientation.  range[t] = {begin, end} indices
// into p.secs of the sectors of type t.
hipError_t launch_deskew(const DevPlan& p, const int range[8], const float* src, int source,
                         int is_max, float* acc, int dimx, int dimy, long long img_acc,
                         hipStream_t s);
// The same for every image of a chunk in one launch: parts[k] holds image k's
// sector ranges, accumulator offset and sides (max_dim: the largest side).
hipError_t launch_deskew_parts(const DevPlan& p, const DeskewPart* parts, int nparts, int max_dim, const float* src,
                               int source, int is_max, float* acc, hipStream_t s);

// reduce(): out (+)= o0, o1, o2, o3 with the inverse orientation maps
// (src/skewEngine.cuh:885-912).  present[t] == 0 skips buffer t.
hipError_t launch_reduce(const float* acc, const int present[4], float* out, int dimx, int dimy,
                         int is_max, hipStream_t s);

}  // namespace ske
