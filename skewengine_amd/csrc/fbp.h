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

// Zero border of the projector's image copy (padded: batch x (n + 2 kProjPad)^2).
constexpr int kProjPad = 4;
// ang_bins: angles with |cos| >= |sin|; ang_rays: the others (the two lists
// together cover every angle once).
hipError_t launch_fbp_project(const float* img, float* padded, float* sino, const float2* cs, const int* ang_bins,
                              int n_bins, const int* ang_rays, int n_rays, const FbpGeom& g, int batch, hipStream_t s);
hipError_t launch_fbp_filter(const float* sino, float* filt, const float* taps, const FbpGeom& g, int batch,
                             hipStream_t s);
hipError_t launch_fbp_backproject(const float* filt, float* img, const float2* cs, const FbpGeom& g, int batch,
                                  float scale, hipStream_t s);

}  // namespace ske
