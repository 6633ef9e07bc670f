// engine.cpp -- the skewEngine object on one MI355X and the C ABI of
// include/skewengine.h.
//
// Core holds the device state of one or more square grids (their four
// oriented copies, orientation accumulators and outputs), the constant
// tables, the planned chunks of (grid, sector) work and the skewOutput
// arena.  Two front ends share it:
//  * ske_engine: one grid.  Its per-sector path (ske_skew / ske_kernel /
//    ske_deskew / ske_reduce) mirrors the reference methods one to one with
//    materialised skewInput/skewOutput buffers reused across sectors like the
//    reference's; ske_run goes through the all-sector path below.
//  * ske_batch: many grids of uneven sizes in one call.
// All-sector path: every (grid, sector) is planned once; a chunk of them that
// fits the arena budget gets one operator launch over all its rows (rows are
// gathered straight from the oriented inputs; the viewshed uses a persistent
// grid pulling cost-sorted rows off an atomic counter), then one deskew per
// grid in ascending sector order, and finally one reduce per grid.  The whole
// sequence can be captured once in a hipGraph and replayed.
// Results are bit-identical to a one-thread run of the reference CPU path for
// the viewshed and identity operators, and equal to it up to FP64 rounding
// for Radon and cepstrum.
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

#include "host_util.h"
#include "launch.h"
#include "planner.h"
#include "skewengine.h"

using namespace ske;

namespace {

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
        v.maxH = ip.maxH;
        return v;
    }
};

struct Image {
    int n = 0;               // side (square grid)
    long long in_off = 0;    // element offset of the four oriented copies
    long long acc_off = 0;   // element offset of the four accumulators
    long long out_off = 0;   // element offset of the output
};

// A chunk: consecutive (grid, sector) pairs whose skewOutput rows fit the arena.
struct Chunk {
    DevPlanBuf plan;
    struct Part {
        int img;
        int range[8];
    };
    std::vector<Part> parts;
    DevBuf<DeskewPart> dparts;  // the parts for the one-launch deskew
    int max_n = 0;              // largest grid side among the parts
    DevBuf<SampleReq> sreq;  // sample points of the local Radon transform in this chunk
    int nreq = 0;
    int sample_version = -1;  // Core::sample_version the requests were built for
};

// Sample-point requests of one planned chunk: for every sector of grid
// `img_off` (four-copy offset) and every point (x, y), the skewed cell whose
// source is (x, y) -- getXY (src/kernelRADN.h:28-41) inverted -- when it lies
// inside its row's [first, last), the cells pointRadon visits.
void sample_requests(const ImagePlan& ip, long long img_off, int n, const std::vector<int>& xy,
                     std::vector<SampleReq>& out) {
    out.clear();
    const int npts = (int)xy.size() / 2;
    for (int s = 0; s < (int)ip.secs.size(); s++) {
        const SecDesc& sd = ip.secs[s];
        if (sd.img != img_off) continue;
        const bool tr = sd.type == 1 || sd.type == 2, rx = sd.type == 1 || sd.type == 3,
                   ry = sd.type == 1 || sd.type == 2;
        for (int p = 0; p < npts; p++) {
            const int x = xy[2 * p], y = xy[2 * p + 1];
            if (x < 0 || x >= n || y < 0 || y >= n) continue;
            const int xx = rx ? n - 1 - x : x, yy = ry ? n - 1 - y : y;
            const int j = tr ? yy : xx;
            const int i = (tr ? xx : yy) + ip.target[sd.tab + j];
            if (i < 0 || i >= sd.H) continue;
            const uint32_t v = ip.fl[sd.rows + i];
            if (j < (int)(v & 0xffffu) || j >= (int)(v >> 16)) continue;
            out.push_back({s, i, j, sd.angle * npts + p});
        }
    }
}

struct RunPlan {
    int s0 = 0, s1 = 0, order = 0;
    long long budget = 0;
    int cell_bytes = 4;
    std::vector<std::unique_ptr<Chunk>> chunks;
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

int check_side(int n) {
    if (n < 2) return fail(SKE_EINVAL, "grid must be at least 2x2");
    if (n > kMaxDim) return fail(SKE_EUNSUPPORTED, "grid side above " + std::to_string(kMaxDim));
    return SKE_OK;
}

int check_dims(int dimx, int dimy) {
    if (dimx < 2 || dimy < 2) return fail(SKE_EINVAL, "grid must be at least 2x2");
    if (dimx > kMaxDim || dimy > kMaxDim) return fail(SKE_EUNSUPPORTED, "grid side above " + std::to_string(kMaxDim));
    // For dimx != dimy skewShape sets last[] = dim_l > dim_i on some rows
    // (src/skewEngine.cuh:581), so the reference's row scans run past the
    // row end into the next row.  Only square grids are accepted.
    if (dimx != dimy) return fail(SKE_EUNSUPPORTED, "only square grids are supported (dimx == dimy)");
    return SKE_OK;
}

struct Core {
    int device = 0;
    int ncu = 0;
    hipStream_t stream = nullptr;
    std::vector<Image> imgs;
    DevBuf<float> in4, acc, out;
    int present[4] = {0, 0, 0, 0};  // per run, same for every grid (same sector range)
    // constant tables
    DevBuf<float4> rtab;
    DevBuf<double> tw;
    DevBuf<int> tw_off;
    DevBuf<double> ctab;
    DevBuf<double> gtab;  // local Radon Gaussian windows
    DevBuf<int> counter;
    // options
    int sdem_variant = 0;
    bool radon_local = false;  // the reference's LRT flag (src/kernelRADN.h:212)
    SdemTuning tune;
    long long arena_bytes = 48LL << 30;
    bool profile = false;
    bool use_graph = false;
    bool waves_set = false;  // waves per workgroup chosen by the caller or the autotuner
    bool count_steps = false;  // viewshed: count executed lane-steps (counter[2..3] as a u64)
    // sample points of the local Radon transform (saveSampleData,
    // src/global.cpp:47-56): image cells (x, y) of grid 0, and the recorded
    // vectors, 180 sectors x points x 4 x 64 floats
    std::vector<int> sample_xy;
    int sample_version = 0;
    DevBuf<float> sout;
    std::map<std::string, double> counters;
    // profiling
    std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int>> prof;
    bool capturing = false;
    // planned runs
    std::unique_ptr<RunPlan> plan;
    DevBuf<float> arena;
    DevBuf<double> W;  // local Radon window sums, one double per arena cell
    DevBuf<double> P;
    DevBuf<float> q;
    // captured graph of the last all-sector run
    hipGraphExec_t gexec = nullptr;
    std::string gkey;

    ~Core() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        for (auto& t : pending) {
            (void)hipEventDestroy(std::get<1>(t));
            (void)hipEventDestroy(std::get<2>(t));
        }
        for (auto ev : event_pool) (void)hipEventDestroy(ev);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        gkey.clear();
    }

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

    // Launch through `fn` (returning hipError_t); with profiling on (and not
    // capturing a graph), bracket it with HIP events on the stream.
    template <typename F>
    int launch(const char* name, F&& fn) {
        const bool timed = profile && !capturing;
        hipEvent_t a = nullptr, b = nullptr;
        if (timed) {
            a = take_event();
            b = take_event();
            SKE_HIP(hipEventRecord(a, stream));
        }
        hipError_t e = fn();
        if (e != hipSuccess) return fail(SKE_EHIP, std::string(name) + ": " + hipGetErrorString(e));
        if (timed) {
            SKE_HIP(hipEventRecord(b, stream));
            pending.emplace_back(name, a, b);
        }
        return SKE_OK;
    }

    unsigned long long* step_counter() { return reinterpret_cast<unsigned long long*>(counter.p + 2); }

    int sync() {
        SKE_HIP(hipStreamSynchronize(stream));
        if (count_steps) {
            unsigned long long v = 0;
            SKE_HIP(hipMemcpy(&v, step_counter(), sizeof(v), hipMemcpyDeviceToHost));
            SKE_HIP(hipMemset(step_counter(), 0, sizeof(v)));
            counters["op_sdem_lane_steps"] += (double)v;
        }
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

    int init(int device_id) {
        int ndev = 0;
        SKE_HIP(hipGetDeviceCount(&ndev));
        if (device_id < 0 || device_id >= ndev) return fail(SKE_EINVAL, "device " + std::to_string(device_id) + " not present");
        device = device_id;
        SKE_HIP(hipSetDevice(device));
        SKE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        SKE_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        std::vector<float> r;
        recip_table(kMaxDim + 1, r);
        // float4 table of kMaxDim+1 entries, then the float2 pairs (read as
        // float4 storage: 2*(kMaxDim+33) floats, an even number of float4s)
        SKE_TRY(rtab.upload(reinterpret_cast<const float4*>(r.data()), r.size() / 4, stream));
        std::vector<double> t;
        std::vector<int> off;
        fft_twiddles(kMaxFftLog, t, off);
        SKE_TRY(tw.upload(t.data(), t.size(), stream));
        SKE_TRY(tw_off.upload(off.data(), off.size(), stream));
        std::vector<double> c;
        cepstrum_table(c);
        SKE_TRY(ctab.upload(c.data(), c.size(), stream));
        std::vector<double> gt;
        lrt_gauss_table(gt);
        SKE_TRY(gtab.upload(gt.data(), gt.size(), stream));
        SKE_TRY(counter.ensure(16));
        SKE_HIP(hipMemsetAsync(counter.p, 0, 16 * sizeof(int), stream));
        SKE_HIP(hipStreamSynchronize(stream));  // host vectors go out of scope
        return SKE_OK;
    }

    // Upload the grids and build their oriented copies (prepare).
    int add_images(int count, const int* sides, const float* const* inputs, bool on_device) {
        long long tot = 0;
        for (int i = 0; i < count; i++) {
            SKE_TRY(check_side(sides[i]));
            if (!inputs[i]) return fail(SKE_EINVAL, "input " + std::to_string(i) + " is NULL");
            Image im;
            im.n = sides[i];
            const long long N = (long long)im.n * im.n;
            im.in_off = 4 * tot;
            im.acc_off = 4 * tot;
            im.out_off = tot;
            tot += N;
            imgs.push_back(im);
        }
        SKE_TRY(in4.ensure(4 * (size_t)tot));
        SKE_TRY(acc.ensure(4 * (size_t)tot));
        SKE_TRY(out.ensure((size_t)tot));
        for (int i = 0; i < count; i++) {
            const Image& im = imgs[i];
            const size_t N = (size_t)im.n * im.n;
            float* base = in4.p + im.in_off;
            SKE_HIP(hipMemcpyAsync(base, inputs[i], N * sizeof(float),
                                   on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream));
            SKE_TRY(launch("prepare", [&] {
                return launch_prepare(base, base + N, base + 2 * N, base + 3 * N, im.n, im.n, stream);
            }));
        }
        SKE_TRY(launch("fill", [&] { return launch_fill_f32(acc.p, 4 * (size_t)tot, 0.0f, stream); }));
        SKE_TRY(launch("fill", [&] { return launch_fill_f32(out.p, (size_t)tot, 0.0f, stream); }));
        return sync();
    }

    // Plan (or reuse) the chunks of sectors [s0,s1) of every grid; `budget`
    // bytes of device memory per chunk at `cell_bytes` per skewed cell.
    int ensure_plan(int s0, int s1, int order, long long budget, int cell_bytes) {
        if (plan && plan->s0 == s0 && plan->s1 == s1 && plan->order == order && plan->budget == budget &&
            plan->cell_bytes == cell_bytes)
            return SKE_OK;
        drop_graph();
        std::unique_ptr<RunPlan> rp(new RunPlan());
        rp->s0 = s0;
        rp->s1 = s1;
        rp->order = order;
        rp->budget = budget;
        rp->cell_bytes = cell_bytes;
        const long long budget_floats = std::max<long long>(budget / cell_bytes, 1);
        std::unique_ptr<Chunk> cur(new Chunk());
        SectorPlan sp;
        auto flush = [&]() -> int {
            if (cur->parts.empty()) return SKE_OK;
            plan_finish(cur->plan.ip, order);
            SKE_TRY(cur->plan.upload(stream));
            std::vector<DeskewPart> dp(cur->parts.size());
            for (size_t k = 0; k < cur->parts.size(); k++) {
                const Image& im = imgs[cur->parts[k].img];
                for (int r = 0; r < 8; r++) dp[k].range[r] = cur->parts[k].range[r];
                dp[k].acc_off = im.acc_off;
                dp[k].dimx = dp[k].dimy = im.n;
                cur->max_n = std::max(cur->max_n, im.n);
            }
            SKE_TRY(cur->dparts.upload(dp.data(), dp.size(), stream));
            rp->max_so = std::max(rp->max_so, cur->plan.ip.so_total);
            rp->max_q = std::max(rp->max_q, cur->plan.ip.q_total);
            rp->maxH = std::max(rp->maxH, cur->plan.ip.maxH);
            rp->cells += cur->plan.ip.cells;
            rp->steps += cur->plan.ip.steps;
            rp->chunks.push_back(std::move(cur));
            cur.reset(new Chunk());
            return SKE_OK;
        };
        for (int gi = 0; gi < (int)imgs.size(); gi++) {
            const int n = imgs[gi].n;
            int a = s0;
            while (a < s1) {
                // extend this grid's run of sectors while it fits the chunk
                int b = a;
                long long have = cur->plan.ip.so_total;
                while (b < s1) {
                    plan_sector(n, n, b, sp);
                    const long long need = (long long)sp.H * sp.dim_i;
                    if (have + need > budget_floats && !(cur->parts.empty() && b == a)) break;
                    have += need;
                    b++;
                }
                if (b == a) {  // nothing fits next to what the chunk holds: start a new chunk
                    SKE_TRY(flush());
                    continue;
                }
                Chunk::Part part;
                part.img = gi;
                plan_append(n, n, a, b, imgs[gi].in_off, cur->plan.ip, part.range);
                cur->parts.push_back(part);
                a = b;
                if (a < s1) SKE_TRY(flush());
            }
        }
        SKE_TRY(flush());
        SKE_HIP(hipStreamSynchronize(stream));
        present[0] = present[1] = present[2] = present[3] = 0;
        for (int a = s0; a < s1 && !imgs.empty(); a++) {
            plan_sector(imgs[0].n, imgs[0].n, a, sp);
            present[sp.type] = 1;
        }
        plan = std::move(rp);
        return SKE_OK;
    }

    int prepare_run(int kernel_id, int s0, int s1) {
        if (!valid_kernel(kernel_id)) return fail(SKE_EINVAL, "unknown operator id " + std::to_string(kernel_id));
        if (s0 < 0 || s1 > SKE_MAX_SECTOR || s0 >= s1) return fail(SKE_EINVAL, "sector range must be within [0,180)");
        if (imgs.empty()) return fail(SKE_ESTATE, "no grids");
        SKE_HIP(hipSetDevice(device));
        const int order = (kernel_id == SKE_KERNEL_SDEM) ? kOrderCostDesc : kOrderNatural;
        const bool local = kernel_id == SKE_KERNEL_RADON && radon_local;
        SKE_TRY(ensure_plan(s0, s1, order, arena_bytes, local ? 12 : 4));
        if (local) {
            if (arena.n < (size_t)plan->max_so || W.n < (size_t)plan->max_so) drop_graph();
            SKE_TRY(arena.ensure((size_t)plan->max_so));
            SKE_TRY(W.ensure((size_t)plan->max_so));
            if (!sample_xy.empty()) {
                const size_t nout = (size_t)SKE_MAX_SECTOR * (sample_xy.size() / 2) * 256;
                if (sout.n < nout) drop_graph();
                SKE_TRY(sout.ensure(nout));
                for (auto& cp : plan->chunks) {
                    if (cp->sample_version == sample_version) continue;
                    std::vector<SampleReq> req;
                    sample_requests(cp->plan.ip, imgs[0].in_off, imgs[0].n, sample_xy, req);
                    if (!req.empty()) SKE_TRY(cp->sreq.upload(req.data(), req.size(), stream));
                    cp->nreq = (int)req.size();
                    cp->sample_version = sample_version;
                }
                SKE_HIP(hipStreamSynchronize(stream));  // the host request vectors are gone
            }
        } else if (kernel_id == SKE_KERNEL_RADON) {
            if (next_pow2_log(plan->maxH) > kMaxFftLog) return fail(SKE_EUNSUPPORTED, "Radon projection longer than 4096 bins");
            if (P.n < (size_t)plan->max_q || q.n < (size_t)plan->max_q) drop_graph();
            SKE_TRY(P.ensure((size_t)plan->max_q));
            SKE_TRY(q.ensure((size_t)plan->max_q));
        } else {
            if (arena.n < (size_t)plan->max_so) drop_graph();
            SKE_TRY(arena.ensure((size_t)plan->max_so));
        }
        return SKE_OK;
    }

    // Without a caller's or the autotuner's choice, a grid workgroup whose
    // rows hold fewer than four work units (chunk groups, or chunks) per scan
    // direction runs 2 waves instead of 4, so that no wave of it idles
    // (500 x 500 grid: 2.36 instead of 3.19 ms).
    SdemTuning effective_tuning(int maxL) const {
        SdemTuning t = tune;
        if (!waves_set && t.persistent_wgs_per_cu == 0) {
            const int chunks = (maxL + 63) / 64;
            const int units = t.group > 1 ? (chunks + t.group - 1) / t.group : chunks;
            t.waves = units < 4 ? 2 : 4;
        }
        return t;
    }

    // Enqueue the whole all-sector run (no allocation, no synchronisation:
    // safe to capture).
    int enqueue_run(int kernel_id, float povh) {
        const bool local = kernel_id == SKE_KERNEL_RADON && radon_local;
        const bool radon = kernel_id == SKE_KERNEL_RADON && !local;
        const int is_max = kernel_id == SKE_KERNEL_CEPSTRUM ? 1 : 0;
        const size_t acc_n = acc.n;
        SKE_TRY(launch("fill", [&] { return launch_fill_f32(acc.p, acc_n, 0.0f, stream); }));
        const bool samples = local && !sample_xy.empty();
        if (samples) {
            const size_t nout = (size_t)SKE_MAX_SECTOR * (sample_xy.size() / 2) * 256;
            SKE_TRY(launch("fill", [&] { return launch_fill_f32(sout.p, nout, 0.0f, stream); }));
        }
        for (auto& cp : plan->chunks) {
            const DevPlan v = cp->plan.view();
            const ImagePlan& ip = cp->plan.ip;
            switch (kernel_id) {
            case SKE_KERNEL_SDEM:
                if (sdem_variant == 0) {
                    const SdemTuning t = effective_tuning(v.maxL);
                    SKE_TRY(launch("op_sdem", [&] {
                        return launch_op_sdem_tuned(v, in4.p, nullptr, arena.p, rtab.p, povh, t, counter.p,
                                                    count_steps ? step_counter() : nullptr, stream);
                    }));
                } else {
                    SKE_TRY(launch("op_sdem", [&] {
                        return launch_op_sdem(v, in4.p, nullptr, arena.p, rtab.p, povh, sdem_variant, stream);
                    }));
                }
                break;
            case SKE_KERNEL_CEPSTRUM:
                SKE_TRY(launch("op_blur", [&] { return launch_op_blur(v, in4.p, nullptr, arena.p, ctab.p, stream); }));
                break;
            case SKE_KERNEL_RADON:
                if (local) {
                    SKE_TRY(launch("fill", [&] { return launch_fill_f64(W.p, (size_t)ip.so_total, stream); }));
                    SKE_TRY(launch("op_lrt", [&] {
                        return launch_lrt(v, in4.p, nullptr, W.p, arena.p, ctab.p, gtab.p, stream);
                    }));
                    if (samples && cp->nreq) {
                        SKE_TRY(launch("lrt_samples", [&] {
                            return launch_lrt_samples(v, cp->sreq.p, cp->nreq, in4.p, W.p, ctab.p, gtab.p, sout.p,
                                                      stream);
                        }));
                    }
                    break;
                }
                SKE_TRY(launch("radn_rowsum", [&] { return launch_radn_rowsum(v, in4.p, nullptr, P.p, stream); }));
                SKE_TRY(launch("radn_filter", [&] {
                    return launch_radn_filter(v, P.p, q.p, tw.p, tw_off.p, 1 << next_pow2_log(ip.maxH), stream);
                }));
                break;
            default:
                SKE_TRY(launch("op_unit", [&] { return launch_op_unit(v, in4.p, nullptr, arena.p, stream); }));
                break;
            }
            const float* src = radon ? q.p : arena.p;
            const int source = radon ? kDeskewRadon : kDeskewArena;
            if (radon) {  // latency-bound per image: every image of the chunk in one launch
                SKE_TRY(launch("deskew", [&] {
                    return launch_deskew_parts(v, cp->dparts.p, (int)cp->parts.size(), cp->max_n, src, source, is_max,
                                               acc.p, stream);
                }));
            } else {  // grids sized to each image (a chunk's images may differ in size)
                for (const Chunk::Part& part : cp->parts) {
                    const Image& im = imgs[part.img];
                    SKE_TRY(launch("deskew", [&] {
                        return launch_deskew(v, part.range, src, source, is_max, acc.p, im.n, im.n, im.acc_off,
                                             stream);
                    }));
                }
            }
        }
        const size_t out_n = out.n;
        SKE_TRY(launch("fill", [&] { return launch_fill_f32(out.p, out_n, 0.0f, stream); }));
        for (const Image& im : imgs) {
            SKE_TRY(launch("reduce", [&] {
                return launch_reduce(acc.p + im.acc_off, present, out.p + im.out_off, im.n, im.n, is_max, stream);
            }));
        }
        return SKE_OK;
    }

    int run(int kernel_id, int s0, int s1, float povh) {
        SKE_TRY(prepare_run(kernel_id, s0, s1));
        if (!use_graph) return enqueue_run(kernel_id, povh);
        char key[160];
        unsigned pbits = 0;
        memcpy(&pbits, &povh, sizeof pbits);
        // every setting that changes what enqueue_run records is part of the key
        // (waves_set: unset wave counts are chosen per chunk by effective_tuning)
        snprintf(key, sizeof key, "%d:%d:%d:%08x:%d:%d:%d:%d:%d:%d:%lld:%d:%d:%d", kernel_id, s0, s1, pbits,
                 sdem_variant, tune.waves, waves_set ? 1 : 0, tune.persistent_wgs_per_cu, tune.occupancy, tune.group,
                 arena_bytes, radon_local ? 1 : 0, count_steps ? 1 : 0, sample_version);
        if (!gexec || gkey != key) {
            drop_graph();
            SKE_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            capturing = true;
            int rc = enqueue_run(kernel_id, povh);
            capturing = false;
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(stream, &g);
            if (rc != SKE_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(SKE_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                gexec = nullptr;
                return fail(SKE_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
            gkey = key;
        }
        return launch("graph", [&] { return hipGraphLaunch(gexec, stream); });
    }

    int set_option(int option, long long value) {
        switch (option) {
        case SKE_OPT_SDEM_VARIANT:
            if (value < 0 || value > 2) return fail(SKE_EINVAL, "SDEM variant must be 0, 1 or 2");
            sdem_variant = (int)value;
            return SKE_OK;
        case SKE_OPT_ARENA_BYTES:
            if (value < (1LL << 20)) return fail(SKE_EINVAL, "arena budget below 1 MiB");
            arena_bytes = value;
            return SKE_OK;
        case SKE_OPT_PROFILE:
            profile = value != 0;
            return SKE_OK;
        case SKE_OPT_GRAPH:
            use_graph = value != 0;
            if (!use_graph) drop_graph();
            return SKE_OK;
        case SKE_OPT_SDEM_WAVES:
            if (value != 2 && value != 4 && value != 8) return fail(SKE_EINVAL, "waves per workgroup must be 2, 4 or 8");
            tune.waves = (int)value;
            waves_set = true;
            return SKE_OK;
        case SKE_OPT_SDEM_PERSISTENT:
            if (value < 0 || value > 64) return fail(SKE_EINVAL, "persistent workgroups per CU must be in [0,64]");
            tune.persistent_wgs_per_cu = (int)value;
            return SKE_OK;
        case SKE_OPT_SDEM_OCCUPANCY:
            if (value != 0 && value != 6 && value != 8) return fail(SKE_EINVAL, "occupancy hint must be 0, 6 or 8");
            tune.occupancy = (int)value;
            return SKE_OK;
        case SKE_OPT_RADON_LOCAL:
            radon_local = value != 0;
            drop_graph();
            return SKE_OK;
        case SKE_OPT_SDEM_GROUP:
            if (value != 1 && value != 4) return fail(SKE_EINVAL, "viewshed chunk group must be 1 or 4");
            tune.group = (int)value;
            drop_graph();
            return SKE_OK;
        case SKE_OPT_COUNT_STEPS:
            SKE_TRY(sync());  // earlier launches are counted (or not) as they were issued
            count_steps = value != 0;
            drop_graph();
            return SKE_OK;
        default:
            return fail(SKE_EINVAL, "unknown option " + std::to_string(option));
        }
    }

    int get_option(int option, long long* value) {
        if (!value) return fail(SKE_EINVAL, "value is NULL");
        switch (option) {
        case SKE_OPT_SDEM_VARIANT: *value = sdem_variant; break;
        case SKE_OPT_ARENA_BYTES: *value = arena_bytes; break;
        case SKE_OPT_PROFILE: *value = profile ? 1 : 0; break;
        case SKE_OPT_GRAPH: *value = use_graph ? 1 : 0; break;
        case SKE_OPT_SDEM_WAVES: *value = tune.waves; break;
        case SKE_OPT_SDEM_PERSISTENT: *value = tune.persistent_wgs_per_cu; break;
        case SKE_OPT_SDEM_OCCUPANCY: *value = tune.occupancy; break;
        case SKE_OPT_RADON_LOCAL: *value = radon_local ? 1 : 0; break;
        case SKE_OPT_COUNT_STEPS: *value = count_steps ? 1 : 0; break;
        case SKE_OPT_SDEM_GROUP: *value = tune.group; break;
        default: return fail(SKE_EINVAL, "unknown option " + std::to_string(option));
        }
        return SKE_OK;
    }

    int counter_get(const char* name, double* value) {
        SKE_TRY(sync());
        auto it = counters.find(name);
        if (value) *value = it == counters.end() ? 0.0 : it->second;
        return SKE_OK;
    }

    int profile_get(const char* name, double* total_ms, int* count) {
        SKE_TRY(sync());
        auto it = prof.find(name);
        if (total_ms) *total_ms = it == prof.end() ? 0.0 : it->second.first;
        if (count) *count = it == prof.end() ? 0 : it->second.second;
        return SKE_OK;
    }

    // Time candidate viewshed launch shapes on a sample (one grid per distinct
    // side, every 12th sector): a warm-up launch, then the best of three timed
    // ones per shape; keep the first shape in list order within 2% of the
    // fastest, so that shapes a box's noise cannot separate give one answer.
    int autotune_sdem(float povh, int* waves, int* wgs, int* occ, double* best_ms) {
        SKE_HIP(hipSetDevice(device));
        std::unique_ptr<DevPlanBuf> sample(new DevPlanBuf());
        std::vector<int> seen;
        for (const Image& im : imgs) {
            if (std::find(seen.begin(), seen.end(), im.n) != seen.end()) continue;
            seen.push_back(im.n);
            int range[8];
            for (int a = 0; a < SKE_MAX_SECTOR; a += 12) plan_append(im.n, im.n, a, a + 1, im.in_off, sample->ip, range);
        }
        plan_finish(sample->ip, kOrderCostDesc);
        SKE_TRY(sample->upload(stream));
        DevBuf<float> scratch;
        SKE_TRY(scratch.ensure((size_t)sample->ip.so_total));
        const DevPlan v = sample->view();
        // Listed with the shape that measures fastest on the 180-sector batches first.
        const SdemTuning cands[] = {{4, 0, 0, 4}, {4, 0, 0, 1}, {8, 0, 0, 4}, {2, 0, 0, 4}, {4, 8, 6, 4},
                                    {4, 8, 0, 4}, {4, 8, 8, 1}, {4, 16, 8, 1}, {8, 8, 8, 1}, {2, 16, 8, 1}};
        hipEvent_t a = take_event(), b = take_event();
        constexpr int kCands = (int)(sizeof(cands) / sizeof(cands[0]));
        double t[kCands];
        double best = 1e300;
        for (int ci = 0; ci < kCands; ci++) {
            const SdemTuning& c = cands[ci];
            SKE_HIP(launch_op_sdem_tuned(v, in4.p, nullptr, scratch.p, rtab.p, povh, c, counter.p, nullptr, stream));
            t[ci] = 1e300;
            for (int rep = 0; rep < 3; rep++) {
                SKE_HIP(hipEventRecord(a, stream));
                SKE_HIP(launch_op_sdem_tuned(v, in4.p, nullptr, scratch.p, rtab.p, povh, c, counter.p, nullptr, stream));
                SKE_HIP(hipEventRecord(b, stream));
                SKE_HIP(hipEventSynchronize(b));
                float ms = 0.0f;
                SKE_HIP(hipEventElapsedTime(&ms, a, b));
                t[ci] = std::min(t[ci], (double)ms);
            }
            best = std::min(best, t[ci]);
        }
        SdemTuning pick = tune;
        for (int ci = 0; ci < kCands; ci++)
            if (t[ci] <= 1.02 * best) {
                pick = cands[ci];
                best = t[ci];
                break;
            }
        event_pool.push_back(a);
        event_pool.push_back(b);
        SKE_HIP(hipStreamSynchronize(stream));
        tune = pick;
        waves_set = true;
        drop_graph();
        if (waves) *waves = pick.waves;
        if (wgs) *wgs = pick.persistent_wgs_per_cu;
        if (occ) *occ = pick.occupancy;
        if (best_ms) *best_ms = best;
        return SKE_OK;
    }
};

}  // namespace

struct ske_batch {
    Core core;
};

struct ske_engine {
    Core core;
    int dimx = 0, dimy = 0, N = 0, dim_l = 0;
    int present[4] = {0, 0, 0, 0};  // per-sector path
    int kernel_id = SKE_KERNEL_IDENTITYDEM;
    float povh = 0.15f;
    ske_device_operator custom = nullptr;
    void* custom_user = nullptr;
    bool have_sector = false;
    SectorPlan cur;
    DevPlanBuf one;
    DevBuf<float> S, SO;  // 2N + dim_l each, like skewAlloc (src/skewEngine.cuh:425-426)
    DevBuf<uint16_t> first1, last1;
    DevBuf<double> P1;
    DevBuf<float> q1;
    DevBuf<double> W1;  // local Radon window sums of the current sector
    DevBuf<SampleReq> sreq1;  // its sample-point requests
};

namespace {

int create_common(int dimx, int dimy, const float* input, bool on_device, int device_id, ske_engine** out) {
    if (!out) return fail(SKE_EINVAL, "out is NULL");
    *out = nullptr;
    if (!input) return fail(SKE_EINVAL, "input is NULL");
    SKE_TRY(check_dims(dimx, dimy));
    std::unique_ptr<ske_engine> e(new ske_engine());
    SKE_TRY(e->core.init(device_id));
    e->dimx = dimx;
    e->dimy = dimy;
    e->N = dimx * dimy;
    e->dim_l = std::max(dimx, dimy);
    const int side = dimx;
    SKE_TRY(e->core.add_images(1, &side, &input, on_device));
    *out = e.release();
    return SKE_OK;
}

// ---- per-sector path ------------------------------------------------------

int sector_skew(ske_engine* e, int angle) {
    if (angle < 0 || angle >= SKE_MAX_SECTOR) return fail(SKE_EINVAL, "angle must be in [0,180)");
    Core& c = e->core;
    SKE_HIP(hipSetDevice(c.device));
    plan_sector(e->dimx, e->dimy, angle, e->cur);
    const SectorPlan& sp = e->cur;
    ImagePlan& ip = e->one.ip;
    plan_image(e->dimx, e->dimy, angle, angle + 1, 0, 0, 0, kOrderNatural, ip);
    ip.secs[0].stale = -1;  // the reused buffer carries the stale row itself
    SKE_TRY(e->one.upload(c.stream));
    const size_t sk = 2 * (size_t)e->N + e->dim_l;
    if (!e->S.p) {
        SKE_TRY(e->S.ensure(sk));
        SKE_TRY(e->SO.ensure(sk));
        SKE_HIP(hipMemsetAsync(e->S.p, 0, sk * sizeof(float), c.stream));
        SKE_HIP(hipMemsetAsync(e->SO.p, 0, sk * sizeof(float), c.stream));
        SKE_TRY(e->first1.ensure(2 * (size_t)e->dim_l + 1));
        SKE_TRY(e->last1.ensure(2 * (size_t)e->dim_l + 1));
        SKE_TRY(e->P1.ensure(2 * (size_t)e->dim_l + 2));
        SKE_TRY(e->q1.ensure(2 * (size_t)e->dim_l + 2));
    }
    SKE_HIP(hipMemcpyAsync(e->first1.p, sp.first.data(), sp.first.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                           c.stream));
    SKE_HIP(hipMemcpyAsync(e->last1.p, sp.last.data(), sp.last.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                           c.stream));
    // resetSkew: zero skewHeight*dim_i of both buffers (src/skewEngine.cuh:567-568)
    const size_t cells = (size_t)sp.H * sp.dim_i;
    SKE_HIP(hipMemsetAsync(e->S.p, 0, cells * sizeof(float), c.stream));
    SKE_HIP(hipMemsetAsync(e->SO.p, 0, cells * sizeof(float), c.stream));
    const DevPlan v = e->one.view();
    SKE_TRY(c.launch("skew_full", [&] { return launch_skew_full(v, ip.secs[0], c.in4.p, e->S.p, c.stream); }));
    SKE_HIP(hipStreamSynchronize(c.stream));  // host plan vectors are reused on the next call
    e->have_sector = true;
    return SKE_OK;
}

int sector_kernel(ske_engine* e) {
    if (!e->have_sector) return fail(SKE_ESTATE, "kernel() before skew()");
    Core& c = e->core;
    SKE_HIP(hipSetDevice(c.device));
    const DevPlan v = e->one.view();
    const SectorPlan& sp = e->cur;
    if (e->custom) {
        e->custom(e->SO.p, e->S.p, sp.dim_i, sp.H, e->last1.p, e->first1.p, e->povh, sp.newAngle, (void*)c.stream,
                  e->custom_user);
        SKE_HIP(hipGetLastError());
        return SKE_OK;
    }
    switch (e->kernel_id) {
    case SKE_KERNEL_SDEM:
        SKE_TRY(c.launch("op_sdem", [&] {
            return launch_op_sdem(v, c.in4.p, e->S.p, e->SO.p, c.rtab.p, e->povh, c.sdem_variant, c.stream);
        }));
        break;
    case SKE_KERNEL_CEPSTRUM:
        SKE_TRY(c.launch("op_blur", [&] { return launch_op_blur(v, c.in4.p, e->S.p, e->SO.p, c.ctab.p, c.stream); }));
        break;
    case SKE_KERNEL_RADON: {
        if (c.radon_local) {
            SKE_TRY(e->W1.ensure(2 * (size_t)e->N + e->dim_l));
            SKE_TRY(c.launch("fill", [&] { return launch_fill_f64(e->W1.p, (size_t)sp.H * sp.dim_i, c.stream); }));
            SKE_TRY(c.launch("op_lrt", [&] {
                return launch_lrt(v, c.in4.p, e->S.p, e->W1.p, e->SO.p, c.ctab.p, c.gtab.p, c.stream);
            }));
            if (!c.sample_xy.empty()) {
                const size_t nout = (size_t)SKE_MAX_SECTOR * (c.sample_xy.size() / 2) * 256;
                if (c.sout.n < nout) {
                    SKE_TRY(c.sout.ensure(nout));
                    SKE_TRY(c.launch("fill", [&] { return launch_fill_f32(c.sout.p, nout, 0.0f, c.stream); }));
                }
                std::vector<SampleReq> req;
                sample_requests(e->one.ip, 0, e->dimx, c.sample_xy, req);
                if (!req.empty()) {
                    SKE_TRY(e->sreq1.upload(req.data(), req.size(), c.stream));
                    SKE_TRY(c.launch("lrt_samples", [&] {
                        return launch_lrt_samples(v, e->sreq1.p, (int)req.size(), c.in4.p, e->W1.p, c.ctab.p,
                                                  c.gtab.p, c.sout.p, c.stream);
                    }));
                    SKE_HIP(hipStreamSynchronize(c.stream));  // req goes out of scope
                }
            }
            break;
        }
        if (next_pow2_log(sp.H) > kMaxFftLog) return fail(SKE_EUNSUPPORTED, "Radon projection longer than 4096 bins");
        SKE_HIP(hipMemsetAsync(e->P1.p, 0, (size_t)sp.H * sizeof(double), c.stream));
        SKE_TRY(c.launch("radn_rowsum", [&] { return launch_radn_rowsum(v, c.in4.p, e->S.p, e->P1.p, c.stream); }));
        SKE_TRY(c.launch("radn_filter", [&] {
            return launch_radn_filter(v, e->P1.p, e->q1.p, c.tw.p, c.tw_off.p, 1 << next_pow2_log(sp.H), c.stream);
        }));
        SKE_TRY(c.launch("radn_smear", [&] { return launch_radn_smear(v, e->q1.p, e->SO.p, c.stream); }));
        break;
    }
    default:
        SKE_TRY(c.launch("op_unit", [&] { return launch_op_unit(v, c.in4.p, e->S.p, e->SO.p, c.stream); }));
        break;
    }
    return SKE_OK;
}

int sector_deskew(ske_engine* e, int is_max) {
    if (!e->have_sector) return fail(SKE_ESTATE, "deskew() before skew()");
    Core& c = e->core;
    SKE_HIP(hipSetDevice(c.device));
    const DevPlan v = e->one.view();
    const int t = e->cur.type;
    int range[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    range[2 * t] = 0;
    range[2 * t + 1] = 1;
    if (!e->present[t]) {  // zeroed on first use, like the reference's O buffers (src/skewEngine.cuh:759-765)
        SKE_TRY(c.launch("fill", [&] { return launch_fill_f32(c.acc.p + (size_t)t * e->N, (size_t)e->N, 0.0f, c.stream); }));
    }
    SKE_TRY(c.launch("deskew", [&] {
        return launch_deskew(v, range, e->SO.p, kDeskewDense, is_max ? 1 : 0, c.acc.p, e->dimx, e->dimy, 0, c.stream);
    }));
    e->present[t] = 1;
    return SKE_OK;
}

}  // namespace

// ---- C ABI ------------------------------------------------------------------

extern "C" {

const char* ske_last_error(void) { return last_error().c_str(); }

int ske_version(void) { return 10100; }

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
    (void)hipSetDevice(e->core.device);
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
    Core& c = e->core;
    SKE_HIP(hipSetDevice(c.device));
    const size_t N = (size_t)e->N;
    SKE_HIP(hipMemcpyAsync(c.out.p, output, N * sizeof(float), hipMemcpyHostToDevice, c.stream));
    SKE_TRY(c.launch("reduce", [&] {
        return launch_reduce(c.acc.p, e->present, c.out.p, e->dimx, e->dimy, is_max ? 1 : 0, c.stream);
    }));
    SKE_HIP(hipMemcpyAsync(output, c.out.p, N * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    return c.sync();
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
    Core& c = e->core;
    const size_t cells = (size_t)e->cur.H * e->cur.dim_i;
    if (skewInput) SKE_HIP(hipMemcpyAsync(skewInput, e->S.p, cells * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    if (skewOutput) SKE_HIP(hipMemcpyAsync(skewOutput, e->SO.p, cells * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    return c.sync();
}

int ske_set_skew_output(ske_engine* e, const float* skewOutput) {
    if (!e || !skewOutput) return fail(SKE_EINVAL, "engine or skewOutput is NULL");
    if (!e->have_sector) return fail(SKE_ESTATE, "no sector skewed yet");
    Core& c = e->core;
    const size_t cells = (size_t)e->cur.H * e->cur.dim_i;
    SKE_HIP(hipMemcpyAsync(e->SO.p, skewOutput, cells * sizeof(float), hipMemcpyHostToDevice, c.stream));
    return c.sync();
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
    Core& c = e->core;
    SKE_TRY(c.run(kernel_id, sector_begin, sector_end, povh));
    // the run reused the orientation accumulators: the per-sector path
    // starts over from zeroed ones
    for (int t = 0; t < 4; t++) e->present[t] = 0;
    if (output) {
        SKE_HIP(hipMemcpyAsync(output, c.out.p, (size_t)e->N * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        return c.sync();
    }
    return SKE_OK;
}

int ske_output_device(ske_engine* e, const float** d_output) {
    if (!e || !d_output) return fail(SKE_EINVAL, "engine or d_output is NULL");
    *d_output = e->core.out.p;
    return SKE_OK;
}

int ske_synchronize(ske_engine* e) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    return e->core.sync();
}

int ske_set_option(ske_engine* e, int option, long long value) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    return e->core.set_option(option, value);
}

int ske_get_option(ske_engine* e, int option, long long* value) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    return e->core.get_option(option, value);
}

int ske_profile_get(ske_engine* e, const char* name, double* total_ms, int* count) {
    if (!e || !name) return fail(SKE_EINVAL, "engine or name is NULL");
    return e->core.profile_get(name, total_ms, count);
}

int ske_profile_reset(ske_engine* e) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    SKE_TRY(e->core.sync());
    e->core.prof.clear();
    e->core.counters.clear();
    return SKE_OK;
}

int ske_counter_get(ske_engine* e, const char* name, double* value) {
    if (!e || !name) return fail(SKE_EINVAL, "engine or name is NULL");
    return e->core.counter_get(name, value);
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

int ske_set_sample_points(ske_engine* e, int count, const int* xy) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    if (count < 0 || count > kMaxSamplePoints)
        return fail(SKE_EINVAL, "sample point count must be in [0, " + std::to_string(kMaxSamplePoints) + "]");
    if (count > 0 && !xy) return fail(SKE_EINVAL, "xy is NULL");
    Core& c = e->core;
    SKE_TRY(c.sync());
    c.sample_xy.assign(xy, xy + 2 * count);
    c.sample_version++;
    c.drop_graph();
    if (count) {  // a fresh, zeroed record
        const size_t nout = (size_t)SKE_MAX_SECTOR * count * 256;
        SKE_HIP(hipSetDevice(c.device));
        c.sout.release();
        SKE_TRY(c.sout.ensure(nout));
        SKE_TRY(c.launch("fill", [&] { return launch_fill_f32(c.sout.p, nout, 0.0f, c.stream); }));
        SKE_TRY(c.sync());
    }
    return SKE_OK;
}

int ske_get_sample_data(ske_engine* e, float* out) {
    if (!e || !out) return fail(SKE_EINVAL, "engine or out is NULL");
    Core& c = e->core;
    if (c.sample_xy.empty()) return fail(SKE_ESTATE, "no sample points set");
    SKE_TRY(c.sync());
    const size_t nout = (size_t)SKE_MAX_SECTOR * (c.sample_xy.size() / 2) * 256;
    SKE_HIP(hipMemcpy(out, c.sout.p, nout * sizeof(float), hipMemcpyDeviceToHost));
    return SKE_OK;
}

int ske_plan_sector(int dimx, int dimy, int angle, ske_sector_params* p, int* target, float* weight,
                    unsigned short* first, unsigned short* last) {
    SKE_TRY(check_dims(dimx, dimy));
    if (angle < 0 || angle >= SKE_MAX_SECTOR) return fail(SKE_EINVAL, "angle must be in [0,180)");
    SectorPlan sp;
    plan_sector(dimx, dimy, angle, sp);
    if (p) {
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
    }
    if (target) memcpy(target, sp.target.data(), sizeof(int) * sp.dim_i);
    if (weight) memcpy(weight, sp.weight.data(), sizeof(float) * sp.dim_i);
    if (first) memcpy(first, sp.first.data(), sizeof(uint16_t) * sp.H);
    if (last) memcpy(last, sp.last.data(), sizeof(uint16_t) * sp.H);
    return SKE_OK;
}

int ske_reset_skew(ske_engine* e) {
    if (!e) return fail(SKE_EINVAL, "engine is NULL");
    if (!e->have_sector) return fail(SKE_ESTATE, "no sector skewed yet");
    Core& c = e->core;
    SKE_HIP(hipSetDevice(c.device));
    const size_t cells = (size_t)e->cur.H * e->cur.dim_i;
    SKE_HIP(hipMemsetAsync(e->S.p, 0, cells * sizeof(float), c.stream));
    SKE_HIP(hipMemsetAsync(e->SO.p, 0, cells * sizeof(float), c.stream));
    return c.sync();
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
    std::string msg = last_error();
    ske_destroy(e);
    if (rc != SKE_OK) last_error() = msg;
    return rc;
}

// ---- batches ----------------------------------------------------------------

static int batch_create(int count, const int* sides, const float* const* inputs, bool on_device, int device_id,
                        ske_batch** out) {
    if (!out) return fail(SKE_EINVAL, "out is NULL");
    *out = nullptr;
    if (count <= 0 || !sides || !inputs) return fail(SKE_EINVAL, "empty batch or NULL arrays");
    for (int i = 0; i < count; i++) SKE_TRY(check_side(sides[i]));
    std::unique_ptr<ske_batch> b(new ske_batch());
    SKE_TRY(b->core.init(device_id));
    SKE_TRY(b->core.add_images(count, sides, inputs, on_device));
    *out = b.release();
    return SKE_OK;
}

int ske_batch_create(int count, const int* sides, const float* const* inputs, int device_id, ske_batch** out) {
    return batch_create(count, sides, inputs, false, device_id, out);
}

int ske_batch_create_device(int count, const int* sides, const float* const* d_inputs, int device_id,
                            ske_batch** out) {
    return batch_create(count, sides, d_inputs, true, device_id, out);
}

int ske_batch_destroy(ske_batch* b) {
    if (!b) return fail(SKE_EINVAL, "batch is NULL");
    (void)hipSetDevice(b->core.device);
    delete b;
    return SKE_OK;
}

int ske_batch_run(ske_batch* b, int kernel_id, int sector_begin, int sector_end, float povh) {
    if (!b) return fail(SKE_EINVAL, "batch is NULL");
    return b->core.run(kernel_id, sector_begin, sector_end, povh);
}

int ske_batch_output(ske_batch* b, int index, float* output) {
    if (!b || !output) return fail(SKE_EINVAL, "batch or output is NULL");
    Core& c = b->core;
    if (index < 0 || index >= (int)c.imgs.size()) return fail(SKE_EINVAL, "grid index out of range");
    const Image& im = c.imgs[index];
    SKE_HIP(hipMemcpyAsync(output, c.out.p + im.out_off, (size_t)im.n * im.n * sizeof(float), hipMemcpyDeviceToHost,
                           c.stream));
    return c.sync();
}

int ske_batch_output_device(ske_batch* b, int index, const float** d_output) {
    if (!b || !d_output) return fail(SKE_EINVAL, "batch or d_output is NULL");
    Core& c = b->core;
    if (index < 0 || index >= (int)c.imgs.size()) return fail(SKE_EINVAL, "grid index out of range");
    *d_output = c.out.p + c.imgs[index].out_off;
    return SKE_OK;
}

int ske_batch_synchronize(ske_batch* b) {
    if (!b) return fail(SKE_EINVAL, "batch is NULL");
    return b->core.sync();
}

int ske_batch_set_option(ske_batch* b, int option, long long value) {
    if (!b) return fail(SKE_EINVAL, "batch is NULL");
    return b->core.set_option(option, value);
}

int ske_batch_get_option(ske_batch* b, int option, long long* value) {
    if (!b) return fail(SKE_EINVAL, "batch is NULL");
    return b->core.get_option(option, value);
}

int ske_batch_profile_get(ske_batch* b, const char* name, double* total_ms, int* count) {
    if (!b || !name) return fail(SKE_EINVAL, "batch or name is NULL");
    return b->core.profile_get(name, total_ms, count);
}

int ske_batch_profile_reset(ske_batch* b) {
    if (!b) return fail(SKE_EINVAL, "batch is NULL");
    SKE_TRY(b->core.sync());
    b->core.prof.clear();
    b->core.counters.clear();
    return SKE_OK;
}

int ske_batch_counter_get(ske_batch* b, const char* name, double* value) {
    if (!b || !name) return fail(SKE_EINVAL, "batch or name is NULL");
    return b->core.counter_get(name, value);
}

int ske_batch_autotune(ske_batch* b, int kernel_id, float povh, int* waves, int* wgs_per_cu, int* occupancy,
                       double* best_ms) {
    if (!b) return fail(SKE_EINVAL, "batch is NULL");
    if (kernel_id != SKE_KERNEL_SDEM) return fail(SKE_EUNSUPPORTED, "the autotuner covers the viewshed operator");
    return b->core.autotune_sdem(povh, waves, wgs_per_cu, occupancy, best_ms);
}

}  // extern "C"
