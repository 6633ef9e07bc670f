// fbp.h -- standard parallel-beam geometry shared by the FBP host object and
// its kernels (kernels_fbp.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace ske {

struct FbpGeom {
    int n;        // image side
    int nang;     // projection angles
    int ndet;     // detector bins
    int nt;       // ray samples per bin (= ndet)
    float delta;  // pixel pitch and bin spacing, 2/n
    float det_c;  // (ndet-1)/2
    float tmax;   // (nt-1)/2 * delta
    float pad;
};

// Tiles of the projector: the base pixels [-1, n) of each axis in ntx tiles
// of side ts <= kPtMax; kw partial sums per (slice, angle, tile), enough for
// every bin that can cross a tile (ts sqrt(2) plus margins).
constexpr int kPtMax = 129;
constexpr int kPtAngles = 12;  // angles per projector workgroup
struct ProjTiling {
    int ntx;
    int ts;
    int kw;
};
inline ProjTiling proj_tiling(int n) {
    ProjTiling t;
    t.ntx = (n + 1 + kPtMax - 1) / kPtMax;
    t.ts = (n + 1 + t.ntx - 1) / t.ntx;
    t.kw = (t.ts * 1415 + 999) / 1000 + 8;
    return t;
}
// img: batch x n x n; part: batch x nang x ntx^2 x kw floats of scratch.
hipError_t launch_fbp_project(const float* img, float* part, float* sino, const float2* cs, const FbpGeom& g,
                              const ProjTiling& pt, int batch, hipStream_t s);
hipError_t launch_fbp_filter(const float* sino, float* filt, const float* taps, const FbpGeom& g, int batch,
                             hipStream_t s);
hipError_t launch_fbp_backproject(const float* filt, float* img, const float2* cs, const FbpGeom& g, int batch,
                                  float scale, hipStream_t s);

}  // namespace ske
