// learned.cpp -- learned sparse-view FBP (C ABI ske_learned_*): a learnable
// symmetric filter in place of Ram-Lak, then a residual 4-layer CNN
// (kernels_learned.hip), trained on the device by its own forward/backward
// kernels and Adam.  No reference counterpart; it extends the reference's
// Radon/FBP operator (src/kernelRADN.h:212-259) in the standard geometry of
// ske_fbp.  The object borrows the geometry, stream, timer and buffers of
// an ske_fbp: SINO is its input, RECON its output, IMAGE the training target.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <memory>
#include <vector>

#include "fbp_object.h"
#include "launch.h"
#include "learned.h"

using namespace ske;

struct ske_learned {
    ske_fbp* f = nullptr;
    int C = 16;                    // hidden channels (16 or 32)
    LearnedLayout L = learned_layout(16);
    int nparam = 0;
    int step = 0;                  // Adam steps taken
    bool derived = false;          // full taps and bf16 weights match prm (cleared whenever prm changes)
    DevBuf<float> prm, grad, m1, m2, full, dq, dfull, dout;
    DevBuf<double> loss;
    DevBuf<float> r0, dz, dr0, drT;  // dz: three layers back to back; drT: dr0 transposed
    DevBuf<uint16_t> act, wb;      // act: a1, a2, a3 back to back; wb: bf16 weights (learned_wb_size)
    DevBuf<float> ell;
    size_t pix(int batch) const { return (size_t)f->g.n * f->g.n * (size_t)batch; }
    LearnedBufs bufs(int batch, bool train) const {
        LearnedBufs B{};
        const size_t a = pix(f->max_batch) * C;
        B.C = C;
        B.r0 = r0.p;
        B.a1 = act.p;
        B.a2 = act.p + a;
        B.a3 = act.p + 2 * a;
        B.out = f->buf[SKE_FBP_RECON].p;
        if (train) {
            B.dz1 = dz.p;
            B.dz2 = dz.p + a;
            B.dz3 = dz.p + 2 * a;
            B.dr0 = dr0.p;
        }
        B.wb2 = wb.p;
        B.wb3 = wb.p + 9 * C * C;
        B.w1b = wb.p + 2 * 9 * C * C;
        B.wt2 = wb.p + 2 * 9 * C * C + 16 * C;
        B.wt3 = wb.p + 3 * 9 * C * C + 16 * C;
        (void)batch;
        return B;
    }
};

namespace {

// splitmix64 stream: uniform doubles in [0, 1).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {  // Box-Muller
        const double u1 = uniform() + 1e-300, u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
};

void random_ellipses(uint64_t seed, int batch, int nell, float* out) {
    Rng r(seed);
    for (int i = 0; i < batch * nell; i++) {
        const double rad = sqrt(r.uniform()), t = 2.0 * M_PI * r.uniform();
        float* e = out + 6 * (size_t)i;
        e[0] = (float)(rad * cos(t));
        e[1] = (float)(rad * sin(t));
        e[2] = (float)(0.05 + 0.35 * r.uniform());
        e[3] = (float)(0.05 + 0.35 * r.uniform());
        e[4] = (float)(M_PI * r.uniform());
        e[5] = (float)(-0.5 + 1.5 * r.uniform());
    }
}

int check(ske_learned* l, int batch) {
    if (!l) return fail(SKE_EINVAL, "learned is NULL");
    if (batch < 1 || batch > l->f->max_batch) return fail(SKE_EINVAL, "batch outside [1, max_batch]");
    return SKE_OK;
}

int forward(ske_learned* l, int batch, bool train) {
    ske_fbp* f = l->f;
    hipStream_t s = f->stream;
    const FbpGeom& g = f->g;
    const LearnedBufs B = l->bufs(batch, train);
    if (!l->derived) {  // the parameters changed: expanded taps and bf16 matrix-core weights
        SKE_TRY(f->timer.launch(s, "learned_params", [&] {
            hipError_t e = launch_taps_expand(l->prm.p + l->L.Taps, l->full.p, g.ndet, s);
            if (e != hipSuccess) return e;
            const int C = l->C;
            return launch_pack_weights(l->prm.p, C, l->wb.p, s);
        }));
        l->derived = true;
    }
    SKE_TRY(f->timer.launch(s, "learned_filter", [&] {
        return launch_fbp_filter(f->buf[SKE_FBP_SINO].p, f->buf[SKE_FBP_FILTERED].p, l->full.p, g, batch, s);
    }));
    SKE_TRY(f->timer.launch(s, "learned_backproject", [&] {
        return launch_fbp_backproject(f->buf[SKE_FBP_FILTERED].p, l->r0.p, f->cs.p, g, batch, (float)(M_PI / g.nang),
                                      s);
    }));
    return f->timer.launch(s, "cnn_forward", [&] {
        if (!train) return launch_cnn_infer(B, l->prm.p, g.n, batch, s);
        hipError_t e = launch_fill_f64(l->loss.p, 1, s);
        if (e != hipSuccess) return e;
        return launch_cnn_forward(B, l->prm.p, g.n, batch, train ? f->buf[SKE_FBP_IMAGE].p : nullptr,
                                  train ? l->dout.p : nullptr, train ? l->loss.p : nullptr, s);
    });
}

int ensure_train(ske_learned* l) {
    const int mb = l->f->max_batch;
    const FbpGeom& g = l->f->g;
    SKE_TRY(l->dz.ensure(3 * l->pix(mb) * l->C));
    SKE_TRY(l->dr0.ensure(l->pix(mb)));
    SKE_TRY(l->drT.ensure(l->pix(mb)));
    SKE_TRY(l->dout.ensure(l->pix(mb)));
    SKE_TRY(l->dq.ensure((size_t)mb * g.nang * g.ndet));
    SKE_TRY(l->dfull.ensure(2 * (size_t)g.ndet - 1));
    SKE_TRY(l->grad.ensure(l->nparam));
    SKE_TRY(l->loss.ensure(1));
    return SKE_OK;
}

// Forward with the loss against IMAGE, then the gradient into l->grad.
int gradient(ske_learned* l, int batch) {
    SKE_TRY(ensure_train(l));
    ske_fbp* f = l->f;
    hipStream_t s = f->stream;
    const FbpGeom& g = f->g;
    SKE_TRY(forward(l, batch, true));
    const LearnedBufs B = l->bufs(batch, true);
    SKE_TRY(f->timer.launch(s, "cnn_backward", [&] {
        hipError_t e = launch_fill_f32(l->grad.p, (size_t)l->nparam, 0.0f, s);
        if (e != hipSuccess) return e;
        return launch_cnn_backward(B, l->prm.p, l->dout.p, l->grad.p, g.n, batch, s);
    }));
    return f->timer.launch(s, "filter_backward", [&] {
        hipError_t e = launch_fill_f32(l->dfull.p, 2 * (size_t)g.ndet - 1, 0.0f, s);
        if (e != hipSuccess) return e;
        return launch_filter_backward(l->dr0.p, l->drT.p, l->dq.p, f->buf[SKE_FBP_SINO].p, l->dfull.p,
                                      l->grad.p + l->L.Taps, f->cs.p, g, batch, s);
    });
}

int read_loss(ske_learned* l, int batch, double* loss) {
    if (!loss) return SKE_OK;
    double sum = 0.0;
    SKE_HIP(hipMemcpyAsync(&sum, l->loss.p, sizeof(double), hipMemcpyDeviceToHost, l->f->stream));
    SKE_TRY(l->f->timer.sync(l->f->stream));
    *loss = sum / ((double)l->pix(batch));
    return SKE_OK;
}

}  // namespace

extern "C" {

int ske_learned_create(ske_fbp* f, ske_learned** out) { return ske_learned_create_channels(f, 16, out); }

int ske_learned_create_channels(ske_fbp* f, int channels, ske_learned** out) {
    if (!out) return fail(SKE_EINVAL, "out is NULL");
    *out = nullptr;
    if (!f) return fail(SKE_EINVAL, "fbp is NULL");
    if (!learned_channels_ok(channels)) return fail(SKE_EINVAL, "hidden channels must be 16 or 32");
    if (f->g.n % 32) return fail(SKE_EUNSUPPORTED, "learned FBP needs an image side that is a multiple of 32");
    SKE_HIP(hipSetDevice(f->device));
    std::unique_ptr<ske_learned> l(new ske_learned());
    l->f = f;
    l->C = channels;
    l->L = learned_layout(channels);
    l->nparam = l->L.Taps + f->g.ndet;
    const size_t pix = l->pix(f->max_batch);
    SKE_TRY(l->prm.ensure(l->nparam));
    SKE_TRY(l->r0.ensure(pix));
    SKE_TRY(l->act.ensure(3 * pix * channels));
    SKE_TRY(l->wb.ensure(learned_wb_size(channels)));
    SKE_TRY(l->full.ensure(2 * (size_t)f->g.ndet - 1));
    SKE_TRY(ske_learned_init(l.get(), 1));
    *out = l.release();
    return SKE_OK;
}

int ske_learned_destroy(ske_learned* l) {
    if (!l) return fail(SKE_EINVAL, "learned is NULL");
    (void)hipSetDevice(l->f->device);
    (void)hipStreamSynchronize(l->f->stream);
    delete l;
    return SKE_OK;
}

int ske_learned_param_count(ske_learned* l, int* count) {
    if (!l || !count) return fail(SKE_EINVAL, "learned or count is NULL");
    *count = l->nparam;
    return SKE_OK;
}

int ske_learned_layout(ske_learned* l, int* channels, int* taps_offset) {
    if (!l) return fail(SKE_EINVAL, "learned is NULL");
    if (channels) *channels = l->C;
    if (taps_offset) *taps_offset = l->L.Taps;
    return SKE_OK;
}

int ske_learned_init(ske_learned* l, unsigned long long seed) {
    if (!l) return fail(SKE_EINVAL, "learned is NULL");
    std::vector<float> p(l->nparam, 0.0f);
    Rng r(seed);
    // He-normal hidden layers, zero biases, zero output layer: out = FBP.
    const int C = l->C;
    const LearnedLayout& L = l->L;
    for (int i = 0; i < C * 9; i++) p[L.W1 + i] = (float)(r.normal() * sqrt(2.0 / 9.0));
    for (int i = 0; i < 9 * C * C; i++) p[L.W2 + i] = (float)(r.normal() * sqrt(2.0 / (9.0 * C)));
    for (int i = 0; i < 9 * C * C; i++) p[L.W3 + i] = (float)(r.normal() * sqrt(2.0 / (9.0 * C)));
    std::vector<float> rl;
    ramlak_taps(l->f->g.ndet, 2.0 / l->f->g.n, rl);
    for (int k = 0; k < l->f->g.ndet; k++) p[L.Taps + k] = rl[l->f->g.ndet - 1 + k];
    return ske_learned_set_params(l, p.data());
}

int ske_learned_set_params(ske_learned* l, const float* params) {
    if (!l || !params) return fail(SKE_EINVAL, "learned or params is NULL");
    SKE_TRY(l->prm.upload(params, l->nparam, l->f->stream));
    l->derived = false;
    l->step = 0;
    if (l->m1.p) {
        SKE_HIP(launch_fill_f32(l->m1.p, l->nparam, 0.0f, l->f->stream));
        SKE_HIP(launch_fill_f32(l->m2.p, l->nparam, 0.0f, l->f->stream));
    }
    return l->f->timer.sync(l->f->stream);
}

int ske_learned_get_params(ske_learned* l, float* params) {
    if (!l || !params) return fail(SKE_EINVAL, "learned or params is NULL");
    SKE_HIP(hipMemcpyAsync(params, l->prm.p, l->nparam * sizeof(float), hipMemcpyDeviceToHost, l->f->stream));
    return l->f->timer.sync(l->f->stream);
}

int ske_learned_reconstruct(ske_learned* l, int batch) {
    SKE_TRY(check(l, batch));
    return forward(l, batch, false);
}

int ske_learned_gradient(ske_learned* l, int batch, double* loss, float* grad) {
    SKE_TRY(check(l, batch));
    SKE_TRY(gradient(l, batch));
    if (grad)
        SKE_HIP(hipMemcpyAsync(grad, l->grad.p, l->nparam * sizeof(float), hipMemcpyDeviceToHost, l->f->stream));
    SKE_TRY(read_loss(l, batch, loss));
    return l->f->timer.sync(l->f->stream);
}

int ske_learned_train_step(ske_learned* l, int batch, float lr, float lr_taps, double* loss) {
    SKE_TRY(check(l, batch));
    if (!(lr >= 0.0f) || !(lr_taps >= 0.0f)) return fail(SKE_EINVAL, "learning rates must be >= 0");
    if (!l->m1.p) {
        SKE_TRY(l->m1.ensure(l->nparam));
        SKE_TRY(l->m2.ensure(l->nparam));
        SKE_HIP(launch_fill_f32(l->m1.p, l->nparam, 0.0f, l->f->stream));
        SKE_HIP(launch_fill_f32(l->m2.p, l->nparam, 0.0f, l->f->stream));
    }
    SKE_TRY(gradient(l, batch));
    l->step++;
    SKE_TRY(l->f->timer.launch(l->f->stream, "adam", [&] {
        return launch_adam(l->prm.p, l->grad.p, l->m1.p, l->m2.p, l->nparam, lr, lr_taps, l->L.Taps, l->step,
                           l->f->stream);
    }));
    l->derived = false;
    return read_loss(l, batch, loss);
}

int ske_learned_download(ske_learned* l, int which, int batch, void* host) {
    SKE_TRY(check(l, batch));
    if (!host) return fail(SKE_EINVAL, "host is NULL");
    const size_t pix = l->pix(batch), a = l->pix(l->f->max_batch) * l->C;
    const void* src = nullptr;
    size_t bytes = 0;
    switch (which) {
        case SKE_LEARNED_R0: src = l->r0.p; bytes = pix * sizeof(float); break;
        case SKE_LEARNED_A1:
        case SKE_LEARNED_A2:
        case SKE_LEARNED_A3:
            src = l->act.p + (which - SKE_LEARNED_A1) * a;
            bytes = pix * l->C * sizeof(uint16_t);
            break;
        case SKE_LEARNED_DR0:
            if (!l->dr0.p) return fail(SKE_ESTATE, "no gradient computed yet");
            src = l->dr0.p;
            bytes = pix * sizeof(float);
            break;
        case SKE_LEARNED_FILTERED_GRAD:
            if (!l->dq.p) return fail(SKE_ESTATE, "no gradient computed yet");
            src = l->dq.p;
            bytes = (size_t)batch * l->f->g.nang * l->f->g.ndet * sizeof(float);
            break;
        default: return fail(SKE_EINVAL, "unknown learned buffer id");
    }
    SKE_HIP(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, l->f->stream));
    return l->f->timer.sync(l->f->stream);
}

int ske_random_ellipses(unsigned long long seed, int batch, int n_ellipses, float* out) {
    if (!out || batch < 0 || n_ellipses < 0) return fail(SKE_EINVAL, "bad ellipse request");
    random_ellipses(seed, batch, n_ellipses, out);
    return SKE_OK;
}

int ske_learned_render(ske_learned* l, int batch, int n_ellipses, const float* ellipses) {
    SKE_TRY(check(l, batch));
    if (!ellipses || n_ellipses < 1) return fail(SKE_EINVAL, "ellipses is NULL or empty");
    SKE_TRY(l->ell.upload(ellipses, (size_t)6 * batch * n_ellipses, l->f->stream));
    SKE_TRY(l->f->timer.launch(l->f->stream, "phantom", [&] {
        return launch_phantom(l->ell.p, n_ellipses, l->f->buf[SKE_FBP_IMAGE].p, l->f->g.n, batch, l->f->stream);
    }));
    return l->f->timer.sync(l->f->stream);  // the host array may go once this returns
}

int ske_learned_make_batch(ske_learned* l, int batch, int n_ellipses, unsigned long long seed, float i0, float mu) {
    SKE_TRY(check(l, batch));
    if (n_ellipses < 1) return fail(SKE_EINVAL, "n_ellipses < 1");
    if (i0 > 0.0f && !(mu > 0.0f)) return fail(SKE_EINVAL, "mu must be > 0 with noise");
    std::vector<float> e((size_t)6 * batch * n_ellipses);
    random_ellipses(seed, batch, n_ellipses, e.data());
    SKE_TRY(ske_learned_render(l, batch, n_ellipses, e.data()));
    SKE_TRY(ske_fbp_project(l->f, batch));
    if (i0 > 0.0f) {
        const size_t count = (size_t)batch * l->f->g.nang * l->f->g.ndet;
        SKE_TRY(l->f->timer.launch(l->f->stream, "noise", [&] {
            return launch_noise(l->f->buf[SKE_FBP_SINO].p, count, i0, mu, seed * 0x2545F4914F6CDD1Dull + 1,
                                l->f->stream);
        }));
    }
    return l->f->timer.sync(l->f->stream);
}

}  // extern "C"
