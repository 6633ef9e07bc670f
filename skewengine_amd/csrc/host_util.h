// host_util.h -- host-side helpers shared by the engine and the FBP object:
// error state of the C ABI, device buffers, and HIP-event launch timing.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "skewengine.h"

namespace ske {

// Message of the last failure on the calling thread (ske_last_error).
inline std::string& last_error() {
    static thread_local std::string s;
    return s;
}

inline int fail(int code, const std::string& msg) {
    last_error() = msg;
    return code;
}

#define SKE_HIP(x)                                                                                        \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess) return ::ske::fail(SKE_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SKE_TRY(x)                     \
    do {                               \
        int rc_ = (x);                 \
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
        const size_t c = std::max<size_t>(count, 1);
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

// HIP-event timing of named launches on one stream.
struct LaunchTimer {
    bool enabled = false;
    bool capturing = false;
    std::vector<std::tuple<std::string, hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, int>> totals;

    ~LaunchTimer() {
        for (auto& t : pending) {
            (void)hipEventDestroy(std::get<1>(t));
            (void)hipEventDestroy(std::get<2>(t));
        }
        for (auto ev : pool) (void)hipEventDestroy(ev);
    }

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t ev = pool.back();
            pool.pop_back();
            return ev;
        }
        hipEvent_t ev = nullptr;
        (void)hipEventCreate(&ev);
        return ev;
    }

    template <typename F>
    int launch(hipStream_t stream, const char* name, F&& fn) {
        const bool timed = enabled && !capturing;
        hipEvent_t a = nullptr, b = nullptr;
        if (timed) {
            a = take();
            b = take();
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

    // Synchronise the stream and fold finished launches into the totals.
    int sync(hipStream_t stream) {
        SKE_HIP(hipStreamSynchronize(stream));
        for (auto& t : pending) {
            float ms = 0.0f;
            SKE_HIP(hipEventElapsedTime(&ms, std::get<1>(t), std::get<2>(t)));
            auto& slot = totals[std::get<0>(t)];
            slot.first += ms;
            slot.second += 1;
            pool.push_back(std::get<1>(t));
            pool.push_back(std::get<2>(t));
        }
        pending.clear();
        return SKE_OK;
    }

    int get(hipStream_t stream, const char* name, double* total_ms, int* count) {
        SKE_TRY(sync(stream));
        auto it = totals.find(name);
        if (total_ms) *total_ms = it == totals.end() ? 0.0 : it->second.first;
        if (count) *count = it == totals.end() ? 0 : it->second.second;
        return SKE_OK;
    }
};

}  // namespace ske
