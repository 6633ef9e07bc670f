// engine.cpp -- the skewEngine object on one MI355X and the C ABI of
// include/skewengine.h.
//
// Two paths share the kernels:
//  * the per-sector path (ske_skew / ske_kernel / ske_deskew / ske_reduce)
//    mirrors the reference methods one to one, with materialised
//    skewInput/skewOutput buffers reused across sectors like the reference's;
//  * the all-sector path (ske_run) plans every sector once, runs the operator
//    over all rows of a chunk of sectors in one launch (rows gathered straight
//    from the oriented inputs, results in a per-sector arena), deskews the
//    chunk in ascending sector order and reduces on the device.
// Both give bit-identical results to a one-thread run of the reference CPU
// path for SDEM/identity, and match it to FP64 rounding for Radon/cepstrum.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "launch.h"
#include "planner.h"
#include "skewengine.h"

using namespace ske;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define SKE_HIP(x)                                                                               \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) return fail(SKE_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SKE_TRY(x)                \
    do {                          \
        int rc_ = (x);            \
        if (rc_ != SKE_OK) return rc_; \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Grow to hold at least `count` elements (contents not preserved).
    int ensure(size_t count) {
        if (count <= n && p) return SKE_OK;
        release();
        size_t c = std::max<size_t>(count, 1);
        hipError_t e = hipMalloc(&p, c * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(SKE_ENOMEM, std::string("hipMalloc(") + std::to_string(c * sizeof(T)) +
                                        " bytes): " + hipGetErrorString(e));
        }
        n = c;
        return SKE_OK;
    }
    int upload(const T* h, size_t count, hipStream_t s) {
        SKE_TRY(ensure(count));
        if (count) SKE_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
        return SKE_OK;
    }
};

struct DevPlanBuf {
    ImagePlan ip;
    DevBuf<SecDesc> secs;
    DevBuf<RowWork> work;
    DevBuf<int> target;
    DevBuf<float> weight;
    DevBuf<uint32_t> fl;
    int upload(hipStream_t s) {
        SKE_TRY(secs.upload(ip.secs.data(), ip.secs.size(), s));
        SKE_TRY(work.upload(ip.work.data(), ip.work.size(), s));
        SKE_TRY(target.upload(ip.target.data(), ip.target.size(), s));
        SKE_TRY(weight.upload(ip.weight.data(), ip.weight.size(), s));
        SKE_TRY(fl.upload(ip.fl.data(), ip.fl.size(), s));
        return SKE_OK;
    }
    DevPlan view() const {
        DevPlan v;
        v.secs = secs.p;
        v.nsec = (int)ip.secs.size();
        v.work = work.p;
        v.nwork = (int)ip.work.size();
        v.target = target.p;
        v.weight = weight.p;
        v.fl = fl.p;
        v.maxL = ip.maxL;
        return v;
    }
};

struct RunPlan {
    int s0 = 0, s1 = 0, order = 0;
    long long budget = 0;
    std::vector<std::unique_ptr<DevPlanBuf>> chunks;
    long long max_so = 0, max_q = 0;
    int maxH = 0;
    double cells = 0, steps = 0;
};

int next_pow2_log(int H) {
    int m = 0;
    while ((1 << m) < H) m++;
    return m;
}

bool valid_kernel(int k) { return k >= SKE_KERNEL_SDEM && k <= SKE_KERNEL_CEPSTRUM; }

constexpr int kMaxFftLog = 12;  // LDS FFT up to 4096 points (64 KB of complex doubles)

}  // namespace

struct ske_engine {
    int dimx = 0, dimy = 0, N = 0, dim_l = 0, device = 0;
    hipStream_t stream = nullptr;
    DevBuf<float> in4;   // four oriented copies
    DevBuf<float> acc;   // four orientation accumulators
    DevBuf<float> out;   // reduced output
    int present[4] = {0, 0, 0, 0};
    // constant tables
    DevBuf<float4> rtab;
    DevBuf<double> tw;
    DevBuf<int> tw_off;
    DevBuf<double> ctab;
    // options
    int sdem_variant = 0;
    long long arena_bytes = 24LL << 30;
    bool profile = false;
    std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int>> prof;
    // per-sector path
    int kernel_id = SKE_KERNEL_IDENTITYDEM;
    float povh = 0.15f;
    ske_device_operator custom = nullptr;
    void* custom_user = nullptr;
    bool have_sector = false;
    bool have_kernel_out = false;
    SectorPlan cur;
    DevPlanBuf one;
    DevBuf<float> S, SO;          // 2N + dim_l each, like skewAlloc (src/skewEngine.cuh:425-426)
    DevBuf<uint16_t> first1, last1;
    DevBuf<double> P1;
    DevBuf<float> q1;
    // all-sector path
    std::unique_ptr<RunPlan> plan;
    DevBuf<float> arena;
    DevBuf<double> P;
    DevBuf<float> q;

    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t ev = event_pool.back();
            event_pool.pop_back();
            return ev;
        }
        hipEvent_t ev = nullptr;
        (void)hipEventCreate(&ev);
        return ev;
    }

    // Launch through `fn` (returning hipError_t); with profiling on, bracket
    // it with HIP events on the engine stream.
    template <typename F>
    int launch(const char* name, F&& fn) {
        hipEvent_t a = nullptr, b = nullptr;
        if (profile) {
            a = take_event();
            b = take_event();
            SKE_HIP(hipEventRecord(a, stream));
        }
        hipError_t e = fn();
        if (e != hipSuccess) return fail(SKE_EHIP, std::string(name) + ": " + hipGetErrorString(e));
        if (profile) {
            SKE_HIP(hipEventRecord(b, stream));
            pending.emplace_back(name, a, b);
        }
        return SKE_OK;
    }

    int sync() {
        SKE_HIP(hipStreamSynchronize(stream));
        for (auto& t : pending) {
            float ms = 0.0f;
            SKE_HIP(hipEventElapsedTime(&ms, std::get<1>(t), std::get<2>(t)));
            auto& slot = prof[std::get<0>(t)];
            slot.first += ms;
            slot.second += 1;
            event_pool.push_back(std::get<1>(t));
            event_pool.push_back(std::get<2>(t));
        }
        pending.clear();
        return SKE_OK;
    }

    int init_tables() {
        std::vector<float> r;
        recip_table(kMaxDim + 1, r);
        SKE_TRY(rtab.upload(reinterpret_cast<const float4*>(r.data()), kMaxDim + 1, stream));
        std::vector<double> t;
        std::vector<int> off;
        fft_twiddles(kMaxFftLog, t, off);
        SKE_TRY(tw.upload(t.data(), t.size(), stream));
        SKE_TRY(tw_off.upload(off.data(), off.size(), stream));
        std::vector<double> c;
        cepstrum_table(c);
        SKE_TRY(ctab.upload(c.data(), c.size(), stream));
        SKE_HIP(hipStreamSynchronize(stream));  // host vectors go out of scope
        return SKE_OK;
    }

    ~ske_engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& t : pending) {
            (void)hipEventDestroy(std::get<1>(t));
            (void)hipEventDestroy(std::get<2>(t));
        }
        for (auto ev : event_pool) (void)hipEventDestroy(ev);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int check_dims(int dimx, int dimy) {
    if (dimx < 2 || dimy < 2) return fail(SKE_EINVAL, "grid must be at least 2x2");
    if (dimx > kMaxDim || dimy > kMaxDim)
        return fail(SKE_EUNSUPPORTED, "grid side above " + std::to_string(kMaxDim));
    // For dimx != dimy skewShape sets last[] = dim_l > dim_i on some rows
    // (src/skewEngine.cuh:581), so the reference's row scans run past the
    // row end into the next row.  Only square grids are accepted.
    if (dimx != dimy) return fail(SKE_EUNSUPPORTED, "only square grids are supported (dimx == dimy)");
    return SKE_OK;
}

int create_common(int dimx, int dimy, const float* input, bool on_device, int device_id, ske_engine** out) {
    if (!out) return fail(SKE_EINVAL, "out is NULL");
    *out = nullptr;
    if (!input) return fail(SKE_EINVAL, "input is NULL");
    SKE_TRY(check_dims(dimx, dimy));
    int ndev = 0;
    SKE_HIP(hipGetDeviceCount(&ndev));
    if (device_id < 0 || device_id >= ndev)
        return fail(SKE_EINVAL, "device " + std::to_string(device_id) + " not present");
    SKE_HIP(hipSetDevice(device_id));
    std::unique_ptr<ske_engine> e(new ske_engine());
    e->dimx = dimx;
    e->dimy = dimy;
    e->N = dimx * dimy;
    e->dim_l = std::max(dimx, dimy);
    e->device = device_id;
    SKE_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    const size_t N = (size_t)e->N;
    SKE_TRY(e->in4.ensure(4 * N));
    SKE_TRY(e->acc.ensure(4 * N));
    SKE_TRY(e->out.ensure(N));
    SKE_HIP(hipMemcpyAsync(e->in4.p, input, N * sizeof(float),
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, e->stream));
    SKE_TRY(e->launch("prepare", [&] {
        return launch_prepare(e->in4.p, e->in4.p + N, e->in4.p + 2 * N, e->in4.p + 3 * N, dimx, dimy, e->stream);
    }));
    SKE_HIP(hipMemsetAsync(e->acc.p, 0, 4 * N * sizeof(float), e->stream));
    SKE_TRY(e->init_tables());
    *out = e.release();
    return SKE_OK;
}

// Plan (or reuse) the chunks of sectors [s0,s1).  A chunk holds as many
// consecutive sectors as fit the arena budget.
int ensure_plan(ske_engine* e, int s0, int s1, int order) {
    if (e->plan && e->plan->s0 == s0 && e->plan->s1 == s1 && e->plan->order == order &&
        e->plan->budget == e->arena_bytes)
        return SKE_OK;
    std::unique_ptr<RunPlan> rp(new RunPlan());
    rp->s0 = s0;
    rp->s1 = s1;
    rp->order = order;
    rp->budget = e->arena_bytes;
    const long long budget_floats = std::max<long long>(e->arena_bytes / 4, 1);
    SectorPlan sp;
    int a = s0;
    while (a < s1) {
        long long acc_floats = 0;
        int b = a;
        while (b < s1) {
            plan_sector(e->dimx, e->dimy, b, sp);
            const long long need = (long long)sp.H * sp.dim_i;
            if (b > a && acc_floats + need > budget_floats) break;
            acc_floats += need;
            b++;
        }
        std::unique_ptr<DevPlanBuf> c(new DevPlanBuf());
        plan_image(e->dimx, e->dimy, a, b, 0, 0, 0, order, c->ip);
        SKE_TRY(c->upload(e->stream));
        rp->max_so = std::max(rp->max_so, c->ip.so_total);
        rp->max_q = std::max(rp->max_q, c->ip.q_total);
        rp->maxH = std::max(rp->maxH, c->ip.maxH);
        rp->cells += c->ip.cells;
        rp->steps += c->ip.steps;
        rp->chunks.push_back(std::move(c));
        a = b;
    }
    SKE_HIP(hipStreamSynchronize(e->stream));
    e->plan = std::move(rp);
    return SKE_OK;
}

int run_all(ske_engine* e, int kernel_id, int s0, int s1, float povh) {
    if (!valid_kernel(kernel_id)) return fail(SKE_EINVAL, "unknown operator id " + std::to_string(kernel_id));
    if (s0 < 0 || s1 > SKE_MAX_SECTOR || s0 >= s1) return fail(SKE_EINVAL, "sector range must be within [0,180)");
    SKE_HIP(hipSetDevice(e->device));
    const int order = (kernel_id == SKE_KERNEL_SDEM) ? kOrderCostDesc : kOrderNatural;
    SKE_TRY(ensure_plan(e, s0, s1, order));
    RunPlan& rp = *e->plan;
    const bool radon = kernel_id == SKE_KERNEL_RADON;
    const int is_max = kernel_id == SKE_KERNEL_CEPSTRUM ? 1 : 0;
    if (radon && next_pow2_log(rp.maxH) > kMaxFftLog)
        return fail(SKE_EUNSUPPORTED, "Radon projection longer than 4096 bins");
    const size_t N = (size_t)e->N;
    if (radon) {
        SKE_TRY(e->P.ensure((size_t)rp.max_q));
        SKE_TRY(e->q.ensure((size_t)rp.max_q));
    } else {
        SKE_TRY(e->arena.ensure((size_t)rp.max_so));
    }
    SKE_HIP(hipMemsetAsync(e->acc.p, 0, 4 * N * sizeof(float), e->stream));
    for (int t = 0; t < 4; t++) e->present[t] = 0;
    for (auto& cp : rp.chunks) {
        const DevPlan v = cp->view();
        const ImagePlan& ip = cp->ip;
        switch (kernel_id) {
        case SKE_KERNEL_SDEM:
            SKE_TRY(e->launch("op_sdem", [&] {
                return launch_op_sdem(v, e->in4.p, nullptr, e->arena.p, e->rtab.p, povh, e->sdem_variant, e->stream);
            }));
            break;
        case SKE_KERNEL_CEPSTRUM:
            SKE_TRY(e->launch("op_blur", [&] {
                return launch_op_blur(v, e->in4.p, nullptr, e->arena.p, e->ctab.p, e->stream);
            }));
            break;
        case SKE_KERNEL_RADON:
            SKE_HIP(hipMemsetAsync(e->P.p, 0, (size_t)ip.q_total * sizeof(double), e->stream));
            SKE_TRY(e->launch("radn_rowsum", [&] { return launch_radn_rowsum(v, e->in4.p, nullptr, e->P.p, e->stream); }));
            SKE_TRY(e->launch("radn_filter", [&] {
                return launch_radn_filter(v, e->P.p, e->q.p, e->tw.p, e->tw_off.p, 1 << next_pow2_log(ip.maxH),
                                          e->stream);
            }));
            break;
        default:
            SKE_TRY(e->launch("op_unit", [&] { return launch_op_unit(v, e->in4.p, nullptr, e->arena.p, e->stream); }));
            break;
        }
        const float* src = radon ? e->q.p : e->arena.p;
        const int source = radon ? kDeskewRadon : kDeskewArena;
        SKE_TRY(e->launch("deskew", [&] {
            return launch_deskew(v, ip.range, src, source, is_max, e->acc.p, e->dimx, e->dimy, 0, e->stream);
        }));
        for (int t = 0; t < 4; t++)
            if (ip.range[2 * t] < ip.range[2 * t + 1]) e->present[t] = 1;
    }
    SKE_HIP(hipMemsetAsync(e->out.p, 0, N * sizeof(float), e->stream));
    SKE_TRY(e->launch("reduce", [&] {
        return launch_reduce(e->acc.p, e->present, e->out.p, e->dimx, e->dimy, is_max, e->stream);
    }));
    return SKE_OK;
}

// ---- per-sector path ------------------------------------------------------

int sector_skew(ske_engine* e, int angle) {
    if (angle < 0 || angle >= SKE_MAX_SECTOR) return fail(SKE_EINVAL, "angle must be in [0,180)");
    SKE_HIP(hipSetDevice(e->device));
    plan_sector(e->dimx, e->dimy, angle, e->cur);
    const SectorPlan& sp = e->cur;
    ImagePlan& ip = e->one.ip;
    plan_image(e->dimx, e->dimy, angle, angle + 1, 0, 0, 0, kOrderNatural, ip);
    ip.secs[0].stale = -1;  // the reused buffer carries the stale row itself
    SKE_TRY(e->one.upload(e->stream));
    const size_t sk = 2 * (size_t)e->N + e->dim_l;
    if (!e->S.p) {
        SKE_TRY(e->S.ensure(sk));
        SKE_TRY(e->SO.ensure(sk));
        SKE_HIP(hipMemsetAsync(e->S.p, 0, sk * sizeof(float), e->stream));
        SKE_HIP(hipMemsetAsync(e->SO.p, 0, sk * sizeof(float), e->stream));
        SKE_TRY(e->first1.ensure(2 * (size_t)e->dim_l + 1));
        SKE_TRY(e->last1.ensure(2 * (size_t)e->dim_l + 1));
        SKE_TRY(e->P1.ensure(2 * (size_t)e->dim_l + 2));
        SKE_TRY(e->q1.ensure(2 * (size_t)e->dim_l + 2));
    }
    SKE_HIP(hipMemcpyAsync(e->first1.p, sp.first.data(), sp.first.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                           e->stream));
    SKE_HIP(hipMemcpyAsync(e->last1.p, sp.last.data(), sp.last.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                           e->stream));
    // resetSkew: zero skewHeight*dim_i of both buffers (src/skewEngine.cuh:567-568)
    const size_t cells = (size_t)sp.H * sp.dim_i;
    SKE_HIP(hipMemsetAsync(e->S.p, 0, cells * sizeof(float), e->stream));
    SKE_HIP(hipMemsetAsync(e->SO.p, 0, cells * sizeof(float), e->stream));
    const DevPlan v = e->one.view();
    SKE_TRY(e->launch("skew_full", [&] { return launch_skew_full(v, ip.secs[0], e->in4.p, e->S.p, e->stream); }));
    SKE_HIP(hipStreamSynchronize(e->stream));  // host plan vectors are reused on the next call
    e->have_sector = true;
    e->have_kernel_out = false;
    return SKE_OK;
}

int sector_kernel(ske_engine* e) {
    if (!e->have_sector) return fail(SKE_ESTATE, "kernel() before skew()");
    SKE_HIP(hipSetDevice(e->device));
    const DevPlan v = e->one.view();
    const SectorPlan& sp = e->cur;
    if (e->custom) {
        e->custom(e->SO.p, e->S.p, sp.dim_i, sp.H, e->last1.p, e->first1.p, e->povh, sp.newAngle,
                  (void*)e->stream, e->custom_user);
        SKE_HIP(hipGetLastError());
    } else {
        switch (e->kernel_id) {
        case SKE_KERNEL_SDEM:
            SKE_TRY(e->launch("op_sdem", [&] {
                return launch_op_sdem(v, e->in4.p, e->S.p, e->SO.p, e->rtab.p, e->povh, e->sdem_variant, e->stream);
            }));
            break;
        case SKE_KERNEL_CEPSTRUM:
            SKE_TRY(e->launch("op_blur", [&] { return launch_op_blur(v, e->in4.p, e->S.p, e->SO.p, e->ctab.p, e->stream); }));
            break;
        case SKE_KERNEL_RADON: {
            if (next_pow2_log(sp.H) > kMaxFftLog) return fail(SKE_EUNSUPPORTED, "Radon projection longer than 4096 bins");
            SKE_HIP(hipMemsetAsync(e->P1.p, 0, (size_t)sp.H * sizeof(double), e->stream));
            SKE_TRY(e->launch("radn_rowsum", [&] { return launch_radn_rowsum(v, e->in4.p, e->S.p, e->P1.p, e->stream); }));
            SKE_TRY(e->launch("radn_filter", [&] {
                return launch_radn_filter(v, e->P1.p, e->q1.p, e->tw.p, e->tw_off.p, 1 << next_pow2_log(sp.H), e->stream);
            }));
            SKE_TRY(e->launch("radn_smear", [&] { return launch_radn_smear(v, e->q1.p, e->SO.p, e->stream); }));
            break;
        }
        default:
            SKE_TRY(e->launch("op_unit", [&] { return launch_op_unit(v, e->in4.p, e->S.p, e->SO.p, e->stream); }));
            break;
        }
    }
    e->have_kernel_out = true;
    return SKE_OK;
}

int sector_deskew(ske_engine* e, int is_max) {
    if (!e->have_sector) return fail(SKE_ESTATE, "deskew() before skew()");
    SKE_HIP(hipSetDevice(e->device));
    const DevPlan v = e->one.view();
    const int t = e->cur.type;
    int range[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    range[2 * t] = 0;
    range[2 * t + 1] = 1;
    SKE_TRY(e->launch("deskew", [&] {
        return launch_deskew(v, range, e->SO.p, kDeskewDense, is_max ? 1 : 0, e->acc.p, e->dimx, e->dimy, 0, e->stream);
    }));
    e->present[t] = 1;
    return SKE_OK;
}

}  // namespace

// ---- C ABI ------------------------------------------------------------------

extern "C" {

const char* ske_last_error(void) { return g_err.c_str(); }

int ske_version(void) { return 10000; }

int ske_device_count(int* count) {
    if (!count) return fail(SKE_EINVAL, "count is NULL");
    SKE_HIP(hipGetDeviceCount(count));
    return SKE_OK;
}

int ske_create(int dimx, int dimy, const float* input, int device_id, ske_engine** out) {
    return create_common(dimx, dimy, input, false, device_id, out);
}

int ske_create_device(int dimx, int dimy, const float* d_input, int device_id, ske_engine** out) {
    return create_common(dimx, dimy, d_input, true, device_id, out);
}

int ske_destroy(ske_engine* e) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    (void)hipSetDevice(e->device);
    delete e;
    return SKE_OK;
}

int ske_set_kernel(ske_engine* e, int kernel_id, float povh) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    if (!valid_kernel(kernel_id)) return fail(SKE_EINVAL, "unknown operator id " + std::to_string(kernel_id));
    e->kernel_id = kernel_id;
    e->povh = povh;
    e->custom = nullptr;
    return SKE_OK;
}

int ske_set_device_operator(ske_engine* e, ske_device_operator op, void* user) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    e->custom = op;
    e->custom_user = user;
    return SKE_OK;
}

int ske_skew(ske_engine* e, int angle) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    return sector_skew(e, angle);
}

int ske_kernel(ske_engine* e) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    return sector_kernel(e);
}

int ske_deskew(ske_engine* e, int is_max) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    return sector_deskew(e, is_max);
}

int ske_reduce(ske_engine* e, float* output, int is_max) {
    if (!e || !output) return fail(SKE_EINVAL, "engine or output is NULL");
    SKE_HIP(hipSetDevice(e->device));
    const size_t N = (size_t)e->N;
    SKE_HIP(hipMemcpyAsync(e->out.p, output, N * sizeof(float), hipMemcpyHostToDevice, e->stream));
    SKE_TRY(e->launch("reduce", [&] {
        return launch_reduce(e->acc.p, e->present, e->out.p, e->dimx, e->dimy, is_max ? 1 : 0, e->stream);
    }));
    SKE_HIP(hipMemcpyAsync(output, e->out.p, N * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    return e->sync();
}

int ske_get_params(ske_engine* e, ske_sector_params* p) {
    if (!e || !p) return fail(SKE_EINVAL, "engine or params is NULL");
    if (!e->have_sector) return fail(SKE_ESTATE, "no sector skewed yet");
    const SectorPlan& sp = e->cur;
    p->a = sp.angle;
    p->newAngle = sp.newAngle;
    p->sectorType = sp.type;
    p->dim_i = sp.dim_i;
    p->dim_o = sp.dim_o;
    p->offset = sp.offset;
    p->skewHeight = sp.H;
    p->skewWidth = sp.skewWidth;
    p->scale = sp.scale;
    p->skewness = sp.skewness;
    return SKE_OK;
}

int ske_get_tables(ske_engine* e, int* target, float* weight, unsigned short* first, unsigned short* last) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    if (!e->have_sector) return fail(SKE_ESTATE, "no sector skewed yet");
    const SectorPlan& sp = e->cur;
    if (target) memcpy(target, sp.target.data(), sizeof(int) * sp.dim_i);
    if (weight) memcpy(weight, sp.weight.data(), sizeof(float) * sp.dim_i);
    if (first) memcpy(first, sp.first.data(), sizeof(uint16_t) * sp.H);
    if (last) memcpy(last, sp.last.data(), sizeof(uint16_t) * sp.H);
    return SKE_OK;
}

int ske_get_skew(ske_engine* e, float* skewInput, float* skewOutput) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    if (!e->have_sector) return fail(SKE_ESTATE, "no sector skewed yet");
    const size_t cells = (size_t)e->cur.H * e->cur.dim_i;
    if (skewInput) SKE_HIP(hipMemcpyAsync(skewInput, e->S.p, cells * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    if (skewOutput) SKE_HIP(hipMemcpyAsync(skewOutput, e->SO.p, cells * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    return e->sync();
}

int ske_device_pointers(ske_engine* e, float** d_skewInput, float** d_skewOutput, unsigned short** d_first,
                        unsigned short** d_last) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    if (!e->have_sector) return fail(SKE_ESTATE, "no sector skewed yet");
    if (d_skewInput) *d_skewInput = e->S.p;
    if (d_skewOutput) *d_skewOutput = e->SO.p;
    if (d_first) *d_first = e->first1.p;
    if (d_last) *d_last = e->last1.p;
    return SKE_OK;
}

int ske_run(ske_engine* e, int kernel_id, int sector_begin, int sector_end, float povh, float* output) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    SKE_TRY(run_all(e, kernel_id, sector_begin, sector_end, povh));
    if (output) {
        SKE_HIP(hipMemcpyAsync(output, e->out.p, (size_t)e->N * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        return e->sync();
    }
    return SKE_OK;
}

int ske_output_device(ske_engine* e, const float** d_output) {
    if (!e || !d_output) return fail(SKE_EINVAL, "engine or d_output is NULL");
    *d_output = e->out.p;
    return SKE_OK;
}

int ske_synchronize(ske_engine* e) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    return e->sync();
}

int ske_set_option(ske_engine* e, int option, long long value) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    switch (option) {
    case SKE_OPT_SDEM_VARIANT:
        if (value != 0 && value != 1) return fail(SKE_EINVAL, "SDEM variant must be 0 or 1");
        e->sdem_variant = (int)value;
        return SKE_OK;
    case SKE_OPT_ARENA_BYTES:
        if (value < (1LL << 20)) return fail(SKE_EINVAL, "arena budget below 1 MiB");
        e->arena_bytes = value;
        return SKE_OK;
    case SKE_OPT_PROFILE:
        e->profile = value != 0;
        return SKE_OK;
    default:
        return fail(SKE_EINVAL, "unknown option " + std::to_string(option));
    }
}

int ske_profile_get(ske_engine* e, const char* name, double* total_ms, int* count) {
    if (!e || !name) return fail(SKE_EINVAL, "engine or name is NULL");
    SKE_TRY(e->sync());
    auto it = e->prof.find(name);
    if (total_ms) *total_ms = it == e->prof.end() ? 0.0 : it->second.first;
    if (count) *count = it == e->prof.end() ? 0 : it->second.second;
    return SKE_OK;
}

int ske_profile_reset(ske_engine* e) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    SKE_TRY(e->sync());
    e->prof.clear();
    return SKE_OK;
}

int ske_work(int dimx, int dimy, int s0, int s1, double* cells, double* steps) {
    if (check_dims(dimx, dimy) != SKE_OK) return SKE_EINVAL;
    if (s0 < 0 || s1 > SKE_MAX_SECTOR || s0 > s1) return fail(SKE_EINVAL, "bad sector range");
    ImagePlan ip;
    plan_image(dimx, dimy, s0, s1, 0, 0, 0, kOrderNatural, ip);
    if (cells) *cells = ip.cells;
    if (steps) *steps = ip.steps;
    return SKE_OK;
}

int ske_execute(int kernel_id, int dimx, int dimy, const float* input, float* output, float povh, int sector_begin,
                int sector_end) {
    if (!output || !input) return fail(SKE_EINVAL, "input or output is NULL");
    if (!valid_kernel(kernel_id)) return fail(SKE_EINVAL, "unknown operator id " + std::to_string(kernel_id));
    if (sector_begin < 0 || sector_end > SKE_MAX_SECTOR || sector_begin >= sector_end)
        return fail(SKE_EINVAL, "sector range must be within [0,180)");
    SKE_TRY(check_dims(dimx, dimy));
    ske_engine* e = nullptr;
    SKE_TRY(ske_create(dimx, dimy, input, 0, &e));
    int rc = ske_run(e, kernel_id, sector_begin, sector_end, povh, output);
    std::string msg = g_err;
    ske_destroy(e);
    if (rc != SKE_OK) g_err = msg;
    return rc;
}

}  // extern "C"
