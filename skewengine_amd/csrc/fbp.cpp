// fbp.cpp -- standard-geometry Radon projector and filtered back-projection
// (C ABI ske_fbp_*).  No reference counterpart: the reference's Radon
// operator works in its own skewed geometry (src/kernelRADN.h:212-259); this
// object provides the 725-bin parallel-beam projector/FBP of the benchmark
// configurations and the substrate of the learned reconstruction.
#include <math.h>
#include <string.h>

#include <memory>
#include <vector>

#include "fbp_object.h"

using namespace ske;

// delta * Ram-Lak: h(0) = 1/(4 delta^2), h(m odd) = -1/(m^2 pi^2 delta^2),
// h(m even != 0) = 0, stored at index m + ndet - 1 for |m| < ndet.
void ske::ramlak_taps(int ndet, double delta, std::vector<float>& taps) {
    taps.assign(2 * (size_t)ndet - 1, 0.0f);
    for (int m = -(ndet - 1); m <= ndet - 1; m++) {
        double h = 0.0;
        if (m == 0)
            h = 1.0 / (4.0 * delta * delta);
        else if (m % 2)
            h = -1.0 / ((double)m * m * M_PI * M_PI * delta * delta);
        taps[m + ndet - 1] = (float)(delta * h);
    }
}

namespace {

int check_batch(ske_fbp* f, int batch) {
    if (!f) return fail(SKE_EINVAL, "fbp is NULL");
    if (batch < 1 || batch > f->max_batch) return fail(SKE_EINVAL, "batch outside [1, max_batch]");
    return SKE_OK;
}

int check_which(int which) {
    if (which < SKE_FBP_IMAGE || which > SKE_FBP_RECON) return fail(SKE_EINVAL, "unknown FBP buffer id");
    return SKE_OK;
}

}  // namespace

extern "C" {

int ske_fbp_create(int n, int n_angles, int n_det, const float* angles, int max_batch, int device_id, ske_fbp** out) {
    if (!out) return fail(SKE_EINVAL, "out is NULL");
    *out = nullptr;
    if (n < 2 || n > 8192 || n_angles < 1 || n_angles > 4096 || n_det < 2 || n_det > 5000 || max_batch < 1)
        return fail(SKE_EINVAL, "FBP geometry out of range (n <= 8192, n_angles <= 4096, n_det <= 5000)");
    int ndev = 0;
    SKE_HIP(hipGetDeviceCount(&ndev));
    if (device_id < 0 || device_id >= ndev) return fail(SKE_EINVAL, "device not present");
    SKE_HIP(hipSetDevice(device_id));
    std::unique_ptr<ske_fbp> f(new ske_fbp());
    f->device = device_id;
    f->max_batch = max_batch;
    f->g.n = n;
    f->g.nang = n_angles;
    f->g.ndet = n_det;
    f->g.nt = n_det;
    f->g.delta = (float)(2.0 / n);
    f->g.det_c = 0.5f * (float)(n_det - 1);
    f->g.tmax = 0.5f * (float)(n_det - 1) * f->g.delta;
    SKE_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
    std::vector<float2> cs(n_angles);
    for (int a = 0; a < n_angles; a++) {
        const double th = angles ? (double)angles[a] : M_PI * a / n_angles;
        cs[a] = make_float2((float)cos(th), (float)sin(th));
    }
    SKE_TRY(f->cs.upload(cs.data(), cs.size(), f->stream));
    ramlak_taps(n_det, 2.0 / n, f->taps_host);
    SKE_TRY(f->taps.upload(f->taps_host.data(), f->taps_host.size(), f->stream));
    for (int w = 0; w < 4; w++) SKE_TRY(f->buf[w].ensure(f->elems(w, max_batch)));
    f->pt = proj_tiling(n);
    SKE_TRY(f->part.ensure((size_t)max_batch * n_angles * f->pt.ntx * f->pt.ntx * f->pt.kw));
    SKE_TRY(f->timer.sync(f->stream));
    *out = f.release();
    return SKE_OK;
}

int ske_fbp_destroy(ske_fbp* f) {
    if (!f) return fail(SKE_EINVAL, "fbp is NULL");
    (void)hipSetDevice(f->device);
    delete f;
    return SKE_OK;
}

int ske_fbp_upload(ske_fbp* f, int which, int batch, const float* host) {
    SKE_TRY(check_batch(f, batch));
    SKE_TRY(check_which(which));
    if (!host) return fail(SKE_EINVAL, "host is NULL");
    SKE_HIP(hipMemcpyAsync(f->buf[which].p, host, f->elems(which, batch) * sizeof(float), hipMemcpyHostToDevice,
                           f->stream));
    return f->timer.sync(f->stream);
}

int ske_fbp_download(ske_fbp* f, int which, int batch, float* host) {
    SKE_TRY(check_batch(f, batch));
    SKE_TRY(check_which(which));
    if (!host) return fail(SKE_EINVAL, "host is NULL");
    SKE_HIP(hipMemcpyAsync(host, f->buf[which].p, f->elems(which, batch) * sizeof(float), hipMemcpyDeviceToHost,
                           f->stream));
    return f->timer.sync(f->stream);
}

int ske_fbp_buffer(ske_fbp* f, int which, float** d_ptr) {
    if (!f || !d_ptr) return fail(SKE_EINVAL, "fbp or d_ptr is NULL");
    SKE_TRY(check_which(which));
    *d_ptr = f->buf[which].p;
    return SKE_OK;
}

int ske_fbp_project(ske_fbp* f, int batch) {
    SKE_TRY(check_batch(f, batch));
    return f->timer.launch(f->stream, "fbp_project", [&] {
        return launch_fbp_project(f->buf[SKE_FBP_IMAGE].p, f->part.p, f->buf[SKE_FBP_SINO].p, f->cs.p, f->g, f->pt,
                                  batch, f->stream);
    });
}

int ske_fbp_filter(ske_fbp* f, int batch) {
    SKE_TRY(check_batch(f, batch));
    return f->timer.launch(f->stream, "fbp_filter", [&] {
        return launch_fbp_filter(f->buf[SKE_FBP_SINO].p, f->buf[SKE_FBP_FILTERED].p, f->taps.p, f->g, batch, f->stream);
    });
}

int ske_fbp_backproject(ske_fbp* f, int batch) {
    SKE_TRY(check_batch(f, batch));
    const float scale = (float)(M_PI / f->g.nang);
    return f->timer.launch(f->stream, "fbp_backproject", [&] {
        return launch_fbp_backproject(f->buf[SKE_FBP_FILTERED].p, f->buf[SKE_FBP_RECON].p, f->cs.p, f->g, batch, scale,
                                      f->stream);
    });
}

int ske_fbp_reconstruct(ske_fbp* f, int batch) {
    SKE_TRY(ske_fbp_filter(f, batch));
    return ske_fbp_backproject(f, batch);
}

int ske_fbp_set_taps(ske_fbp* f, const float* taps) {
    if (!f) return fail(SKE_EINVAL, "fbp is NULL");
    if (taps)
        f->taps_host.assign(taps, taps + 2 * (size_t)f->g.ndet - 1);
    else
        ramlak_taps(f->g.ndet, 2.0 / f->g.n, f->taps_host);
    SKE_TRY(f->taps.upload(f->taps_host.data(), f->taps_host.size(), f->stream));
    return f->timer.sync(f->stream);
}

int ske_fbp_get_taps(ske_fbp* f, float* taps) {
    if (!f || !taps) return fail(SKE_EINVAL, "fbp or taps is NULL");
    memcpy(taps, f->taps_host.data(), f->taps_host.size() * sizeof(float));
    return SKE_OK;
}

int ske_fbp_synchronize(ske_fbp* f) {
    if (!f) return fail(SKE_EINVAL, "fbp is NULL");
    return f->timer.sync(f->stream);
}

int ske_fbp_set_option(ske_fbp* f, int option, long long value) {
    if (!f) return fail(SKE_EINVAL, "fbp is NULL");
    if (option != SKE_OPT_PROFILE) return fail(SKE_EINVAL, "unknown FBP option");
    f->timer.enabled = value != 0;
    return SKE_OK;
}

int ske_fbp_profile_get(ske_fbp* f, const char* name, double* total_ms, int* count) {
    if (!f || !name) return fail(SKE_EINVAL, "fbp or name is NULL");
    return f->timer.get(f->stream, name, total_ms, count);
}

int ske_fbp_profile_reset(ske_fbp* f) {
    if (!f) return fail(SKE_EINVAL, "fbp is NULL");
    SKE_TRY(f->timer.sync(f->stream));
    f->timer.totals.clear();
    return SKE_OK;
}

}  // extern "C"
