// fbp_object.h -- the ske_fbp object (standard-geometry projector / FBP),
// shared by its C ABI (fbp.cpp) and the learned reconstruction (learned.cpp).
#pragma once
#include <vector>

#include "fbp.h"
#include "host_util.h"
#include "skewengine.h"

struct ske_fbp {
    int device = 0;
    hipStream_t stream = nullptr;
    ske::FbpGeom g{};
    int max_batch = 0;
    ske::DevBuf<float2> cs;
    ske::DevBuf<float> taps;
    std::vector<float> taps_host;
    ske::DevBuf<float> buf[4];  // image, sinogram, filtered, reconstruction
    ske::ProjTiling pt{};
    ske::DevBuf<float> part;    // the projector's partial sums per tile
    ske::LaunchTimer timer;
    ~ske_fbp() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
    size_t elems(int which, int batch) const {
        const size_t per = (which == SKE_FBP_IMAGE || which == SKE_FBP_RECON) ? (size_t)g.n * g.n
                                                                              : (size_t)g.nang * g.ndet;
        return per * (size_t)batch;
    }
};


namespace ske {
// delta * Ram-Lak taps h[m + ndet - 1], |m| < ndet (fbp.cpp).
void ramlak_taps(int ndet, double delta, std::vector<float>& taps);
}  // namespace ske
