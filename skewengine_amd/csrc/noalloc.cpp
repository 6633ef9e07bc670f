// noalloc.cpp -- ske_execute_noalloc: the all-sector run of one square grid
// (the work of ske_execute) with no device allocation, for callers that
// cannot call hipMalloc (code under HIP graph capture elsewhere in their
// program, tight serving loops, harnesses that forbid allocation).
//
// Every device buffer the run needs beyond the caller's input and output --
// the four oriented copies, the four orientation accumulators, the constant
// tables, the planned sector tables, the row sums of the Radon operator and
// the skewOutput arena -- is carved out of `scratch`, one file-scope array of
// kNoAllocScratchBytes per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel or a copy: its device address comes
// from hipGetSymbolAddress, once per device, and only that address is used.
//
// Calls on one GPU share its array and keep the last call's contents; calls
// that overlap on one GPU race.  The constant tables are written once per
// device under a lock; everything else is rewritten by every call (the
// accumulators and the output are cleared with fill kernels first).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "host_util.h"
#include "launch.h"
#include "planner.h"
#include "skewengine.h"

using namespace ske;

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)SKE_NOALLOC_SCRATCH_MB << 20;
constexpr int kMaxDevices = 64;
constexpr int kFftLogMax = 12;  // the Radon operator's LDS FFT (4096 points)

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Per-device state: the address of `scratch` on that device and whether the
// constant tables have been written there.
struct DeviceScratch {
    unsigned char* ptr = nullptr;
    bool tables = false;
};
std::mutex g_lock;
DeviceScratch g_dev[kMaxDevices];

// Byte layout of the constant tables at the start of the array.
struct ConstLayout {
    size_t rtab = 0, tw = 0, tw_off = 0, ctab = 0, gtab = 0, counter = 0, end = 0;
    size_t rtab_n = 0, tw_n = 0, tw_off_n = 0, ctab_n = 0, gtab_n = 0;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Resolve the array's address on `device` (current) and write the constant
// tables once.  Called with g_lock held.
int setup_device(int device, hipStream_t stream, ConstLayout& cl) {
    std::vector<float> r;
    recip_table(kMaxDim + 1, r);
    std::vector<double> t;
    std::vector<int> off;
    fft_twiddles(kFftLogMax, t, off);
    std::vector<double> c;
    cepstrum_table(c);
    std::vector<double> g;
    lrt_gauss_table(g);
    cl.rtab_n = r.size();
    cl.tw_n = t.size();
    cl.tw_off_n = off.size();
    cl.ctab_n = c.size();
    cl.gtab_n = g.size();
    size_t at = 0;
    cl.rtab = at;
    at = align256(at + cl.rtab_n * sizeof(float));
    cl.tw = at;
    at = align256(at + cl.tw_n * sizeof(double));
    cl.tw_off = at;
    at = align256(at + cl.tw_off_n * sizeof(int));
    cl.ctab = at;
    at = align256(at + cl.ctab_n * sizeof(double));
    cl.gtab = at;
    at = align256(at + cl.gtab_n * sizeof(double));
    cl.counter = at;
    at = align256(at + 16 * sizeof(int));
    cl.end = at;

    DeviceScratch& ds = g_dev[device];
    if (!ds.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SKE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ds.ptr = static_cast<unsigned char*>(p);
    }
    if (!ds.tables) {
        unsigned char* base = ds.ptr;
        SKE_HIP(hipMemcpyAsync(base + cl.rtab, r.data(), cl.rtab_n * sizeof(float), hipMemcpyHostToDevice, stream));
        SKE_HIP(hipMemcpyAsync(base + cl.tw, t.data(), cl.tw_n * sizeof(double), hipMemcpyHostToDevice, stream));
        SKE_HIP(hipMemcpyAsync(base + cl.tw_off, off.data(), cl.tw_off_n * sizeof(int), hipMemcpyHostToDevice,
                               stream));
        SKE_HIP(hipMemcpyAsync(base + cl.ctab, c.data(), cl.ctab_n * sizeof(double), hipMemcpyHostToDevice, stream));
        SKE_HIP(hipMemcpyAsync(base + cl.gtab, g.data(), cl.gtab_n * sizeof(double), hipMemcpyHostToDevice, stream));
        SKE_HIP(hipStreamSynchronize(stream));  // the host vectors go out of scope
        ds.tables = true;
    }
    return SKE_OK;
}

// One chunk of consecutive sectors whose skewOutput rows fit the arena.
struct NaChunk {
    ImagePlan ip;
    int range[8];
    size_t secs = 0, work = 0, target = 0, weight = 0, fl = 0;  // byte offsets of its tables
};

int next_pow2_log(int H) {
    int m = 0;
    while ((1 << m) < H) m++;
    return m;
}

int execute_noalloc(int kernel_id, int n, const float* d_input, float* d_output, float povh, int s0, int s1,
                    int radon_local, hipStream_t stream) {
    if (!d_input || !d_output) return fail(SKE_EINVAL, "input or output is NULL");
    if (kernel_id < SKE_KERNEL_SDEM || kernel_id > SKE_KERNEL_CEPSTRUM)
        return fail(SKE_EINVAL, "unknown operator id " + std::to_string(kernel_id));
    if (s0 < 0 || s1 > SKE_MAX_SECTOR || s0 >= s1) return fail(SKE_EINVAL, "sector range must be within [0,180)");
    if (n < 2) return fail(SKE_EINVAL, "grid must be at least 2x2");
    if (n > SKE_NOALLOC_MAX_SIDE)
        return fail(SKE_EINVAL, "ske_execute_noalloc: grid side " + std::to_string(n) + " above the maximum " +
                                    std::to_string(SKE_NOALLOC_MAX_SIDE) + " of its " +
                                    std::to_string(SKE_NOALLOC_SCRATCH_MB) + " MiB scratch array");
    int device = 0;
    SKE_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) return fail(SKE_EUNSUPPORTED, "device index above 63");

    ConstLayout cl;
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_lock);
        SKE_TRY(setup_device(device, stream, cl));
        base = g_dev[device].ptr;
    }
    const bool local = kernel_id == SKE_KERNEL_RADON && radon_local;
    const bool radon = kernel_id == SKE_KERNEL_RADON && !local;
    const int is_max = kernel_id == SKE_KERNEL_CEPSTRUM ? 1 : 0;
    const int cell_bytes = local ? 12 : 4;  // skewOutput float (+ the local transform's window sum double)

    // Sizes of the whole sector range: tables of every sector, and the row
    // count (the Radon row sums of one chunk never exceed it).
    const size_t N = (size_t)n * n;
    ImagePlan all;
    plan_image(n, n, s0, s1, 0, 0, 0, kOrderNatural, all);
    if (radon && next_pow2_log(all.maxH) > kFftLogMax) return fail(SKE_EUNSUPPORTED, "Radon projection too long");
    const int nsec = s1 - s0;
    // every chunk's five tables start on a 256-byte boundary
    const size_t tables_bytes = all.secs.size() * sizeof(SecDesc) + all.work.size() * sizeof(RowWork) +
                                all.target.size() * sizeof(int) + all.weight.size() * sizeof(float) +
                                all.fl.size() * sizeof(uint32_t) + (size_t)nsec * 5 * 256;
    size_t at = cl.end;
    const size_t o_in4 = at;
    at = align256(at + 4 * N * sizeof(float));
    const size_t o_acc = at;
    at = align256(at + 4 * N * sizeof(float));
    const size_t o_tab = at;
    at = align256(at + tables_bytes);
    const size_t o_P = at;
    at = align256(at + (radon ? (size_t)all.q_total * sizeof(double) : 0));
    const size_t o_q = at;
    at = align256(at + (radon ? (size_t)all.q_total * sizeof(float) : 0));
    const size_t o_arena = at;
    if (at >= kNoAllocScratchBytes) return fail(SKE_EINVAL, "ske_execute_noalloc: grid too large for the scratch array");
    // (256 bytes kept back for the alignment of the window sums after the floats)
    const long long arena_cells = (long long)((kNoAllocScratchBytes - o_arena - 256) / cell_bytes);

    // Chunks: runs of consecutive sectors whose rows fit the arena.
    std::vector<NaChunk> chunks;
    SectorPlan sp;
    int present[4] = {0, 0, 0, 0};
    size_t tab_at = o_tab;
    for (int a = s0; a < s1;) {
        int b = a;
        long long have = 0;
        while (b < s1) {
            plan_sector(n, n, b, sp);
            present[sp.type] = 1;
            const long long need = (long long)sp.H * sp.dim_i;
            if (have + need > arena_cells) break;
            have += need;
            b++;
        }
        if (b == a) return fail(SKE_EINVAL, "ske_execute_noalloc: one sector does not fit the scratch array");
        chunks.emplace_back();
        NaChunk& ch = chunks.back();
        plan_append(n, n, a, b, 0, ch.ip, ch.range);
        plan_finish(ch.ip, kernel_id == SKE_KERNEL_SDEM ? kOrderCostDesc : kOrderNatural);
        ch.secs = tab_at;
        tab_at = align256(tab_at + ch.ip.secs.size() * sizeof(SecDesc));
        ch.work = tab_at;
        tab_at = align256(tab_at + ch.ip.work.size() * sizeof(RowWork));
        ch.target = tab_at;
        tab_at = align256(tab_at + ch.ip.target.size() * sizeof(int));
        ch.weight = tab_at;
        tab_at = align256(tab_at + ch.ip.weight.size() * sizeof(float));
        ch.fl = tab_at;
        tab_at = align256(tab_at + ch.ip.fl.size() * sizeof(uint32_t));
        a = b;
    }
    if (tab_at > o_P) return fail(SKE_EINVAL, "ske_execute_noalloc: sector tables overflow their region");

    float* in4 = reinterpret_cast<float*>(base + o_in4);
    float* acc = reinterpret_cast<float*>(base + o_acc);
    double* P = reinterpret_cast<double*>(base + o_P);
    float* q = reinterpret_cast<float*>(base + o_q);
    float* arena = reinterpret_cast<float*>(base + o_arena);
    // the local transform keeps one double of window sums per arena cell after the floats
    double* W = reinterpret_cast<double*>(base + align256(o_arena + (size_t)arena_cells * sizeof(float)));
    const float4* rtab = reinterpret_cast<const float4*>(base + cl.rtab);
    const double* tw = reinterpret_cast<const double*>(base + cl.tw);
    const int* tw_off = reinterpret_cast<const int*>(base + cl.tw_off);
    const double* ctab = reinterpret_cast<const double*>(base + cl.ctab);
    const double* gtab = reinterpret_cast<const double*>(base + cl.gtab);
    int* counter = reinterpret_cast<int*>(base + cl.counter);
    // every chunk holds at most arena_cells skewed cells (floats in the arena,
    // and for the local transform as many doubles in W, which ends inside the array)
    for (const NaChunk& ch : chunks)
        if (ch.ip.so_total > arena_cells) return fail(SKE_EINVAL, "ske_execute_noalloc: chunk overflows the arena");
    if (local && reinterpret_cast<unsigned char*>(W + arena_cells) > base + kNoAllocScratchBytes)
        return fail(SKE_EINVAL, "ske_execute_noalloc: window sums overflow the scratch array");

    // Oriented copies of the caller's grid (prepare) and cleared accumulators.
    SKE_HIP(hipMemcpyAsync(in4, d_input, N * sizeof(float), hipMemcpyDeviceToDevice, stream));
    SKE_HIP(launch_prepare(in4, in4 + N, in4 + 2 * N, in4 + 3 * N, n, n, stream));
    SKE_HIP(launch_fill_f32(acc, 4 * N, 0.0f, stream));
    SKE_HIP(launch_fill_f32(reinterpret_cast<float*>(counter), 16, 0.0f, stream));  // zero bits
    for (NaChunk& ch : chunks) {
        SKE_HIP(hipMemcpyAsync(base + ch.secs, ch.ip.secs.data(), ch.ip.secs.size() * sizeof(SecDesc),
                               hipMemcpyHostToDevice, stream));
        SKE_HIP(hipMemcpyAsync(base + ch.work, ch.ip.work.data(), ch.ip.work.size() * sizeof(RowWork),
                               hipMemcpyHostToDevice, stream));
        SKE_HIP(hipMemcpyAsync(base + ch.target, ch.ip.target.data(), ch.ip.target.size() * sizeof(int),
                               hipMemcpyHostToDevice, stream));
        SKE_HIP(hipMemcpyAsync(base + ch.weight, ch.ip.weight.data(), ch.ip.weight.size() * sizeof(float),
                               hipMemcpyHostToDevice, stream));
        SKE_HIP(hipMemcpyAsync(base + ch.fl, ch.ip.fl.data(), ch.ip.fl.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice, stream));
    }
    for (NaChunk& ch : chunks) {
        DevPlan v;
        v.secs = reinterpret_cast<const SecDesc*>(base + ch.secs);
        v.nsec = (int)ch.ip.secs.size();
        v.work = reinterpret_cast<const RowWork*>(base + ch.work);
        v.nwork = (int)ch.ip.work.size();
        v.target = reinterpret_cast<const int*>(base + ch.target);
        v.weight = reinterpret_cast<const float*>(base + ch.weight);
        v.fl = reinterpret_cast<const uint32_t*>(base + ch.fl);
        v.maxL = ch.ip.maxL;
        v.maxH = ch.ip.maxH;
        switch (kernel_id) {
        case SKE_KERNEL_SDEM: {
            // the engine's default launch shape (two waves for short rows)
            SdemTuning t;
            const int nchunks = (v.maxL + 63) / 64;
            const int units = (nchunks + t.group - 1) / t.group;
            t.waves = units < 4 ? 2 : 4;
            SKE_HIP(launch_op_sdem_tuned(v, in4, nullptr, arena, rtab, povh, t, counter, nullptr, stream));
            break;
        }
        case SKE_KERNEL_CEPSTRUM:
            SKE_HIP(launch_op_blur(v, in4, nullptr, arena, ctab, stream));
            break;
        case SKE_KERNEL_RADON:
            if (local) {
                SKE_HIP(launch_fill_f64(W, (size_t)ch.ip.so_total, stream));
                SKE_HIP(launch_lrt(v, in4, nullptr, W, arena, ctab, gtab, stream));
            } else {
                SKE_HIP(launch_radn_rowsum(v, in4, nullptr, P, stream));
                SKE_HIP(launch_radn_filter(v, P, q, tw, tw_off, 1 << next_pow2_log(ch.ip.maxH), stream));
            }
            break;
        default:
            SKE_HIP(launch_op_unit(v, in4, nullptr, arena, stream));
            break;
        }
        SKE_HIP(launch_deskew(v, ch.range, radon ? q : arena, radon ? kDeskewRadon : kDeskewArena, is_max, acc, n, n,
                              0, stream));
    }
    SKE_HIP(launch_fill_f32(d_output, N, 0.0f, stream));
    SKE_HIP(launch_reduce(acc, present, d_output, n, n, is_max, stream));
    SKE_HIP(hipStreamSynchronize(stream));  // the host plan vectors go out of scope
    return SKE_OK;
}

}  // namespace

extern "C" int ske_execute_noalloc(int kernel_id, int n, const float* d_input, float* d_output, float povh,
                                   int sector_begin, int sector_end, int radon_local, void* stream) {
    try {
        return execute_noalloc(kernel_id, n, d_input, d_output, povh, sector_begin, sector_end, radon_local,
                               static_cast<hipStream_t>(stream));
    } catch (...) {  // host planning memory
        return fail(SKE_ENOMEM, "ske_execute_noalloc: out of host memory");
    }
}
