// stokesdt_amd internal definitions: errors, device buffers, complex types.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/stokesdt_amd.h"

namespace sdt {

// ------------------------------------------------------------- errors ------
// Thread-local last error message, read back through sdt_last_error().
void set_error(const std::string& msg);
const char* last_error();

struct Error {
    int code;
    std::string msg;
};

#define SDT_HIP_CHECK(expr)                                                   \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            throw ::sdt::Error{-5, std::string(#expr) + " failed: " +        \
                                       hipGetErrorString(e_)};                \
        }                                                                     \
    } while (0)

#define SDT_REQUIRE(cond, code, msg)                                          \
    do {                                                                      \
        if (!(cond)) throw ::sdt::Error{(code), (msg)};                       \
    } while (0)

// status codes (SDT_OK, SDT_EINVAL, ...) come from the public header


// ------------------------------------------- capture and device-wide calls --
// Graph captures and the calls that can synchronise with every stream of the
// device (allocation and freeing, synchronous copies, stream and graph
// creation and destruction) exclude each other across threads. A capture is
// thread-local, but such a call from another thread can still wait on the
// capturing stream, which ends that capture in error. Recursive: a capture
// section may allocate through the same thread's calls.
inline std::recursive_mutex& capture_mutex() {
    static std::recursive_mutex m;
    return m;
}
typedef std::lock_guard<std::recursive_mutex> CaptureLock;

inline hipError_t sync_memcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    CaptureLock lk(capture_mutex());
    return hipMemcpy(dst, src, bytes, kind);
}
inline hipError_t stream_create(hipStream_t* s) {
    CaptureLock lk(capture_mutex());
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}
inline hipError_t stream_destroy(hipStream_t s) {
    CaptureLock lk(capture_mutex());
    return hipStreamDestroy(s);
}

// ------------------------------------------------------ device buffer ------
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() {}
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) {
            CaptureLock lk(capture_mutex());
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        n = 0;
    }
    // grow-only allocation; contents are not preserved
    void reserve(size_t count) {
        if (count <= n && ptr) return;
        CaptureLock lk(capture_mutex());
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&ptr, count * sizeof(T));
        if (e != hipSuccess) {
            ptr = nullptr;
            throw Error{SDT_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                        " bytes failed: " + hipGetErrorString(e)};
        }
        n = count;
    }
    T* get() const { return ptr; }
    size_t bytes() const { return n * sizeof(T); }
};

// page-locked host buffer (hipHostMalloc), grow-only; released only when no
// copy into or out of it can be pending
template <typename T>
struct PinnedBuf {
    T* ptr = nullptr;
    size_t n = 0;
    PinnedBuf() {}
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    void release() {
        if (ptr) {
            CaptureLock lk(capture_mutex());
            (void)hipHostFree(ptr);
        }
        ptr = nullptr;
        n = 0;
    }
    void reserve(size_t count) {
        if (count <= n && ptr) return;
        CaptureLock lk(capture_mutex());
        release();
        if (count == 0) count = 1;
        hipError_t e = hipHostMalloc((void**)&ptr, count * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            throw Error{SDT_ENOMEM, "hipHostMalloc of " + std::to_string(count * sizeof(T)) +
                                        " bytes failed: " + hipGetErrorString(e)};
        }
        n = count;
    }
    T* get() const { return ptr; }
};

template <typename T>
struct cplx {
    T x, y;
};

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): linear id i runs on XCD i % 8. This maps the linear id to a logical
// id so that every XCD receives one contiguous range of the logical order
// (spatial neighbours share that XCD's L2). A bijection on [0, n).
__device__ __forceinline__ unsigned xcd_logical_id(unsigned i, unsigned n) {
    constexpr unsigned kXcd = 8;
    const unsigned q = n / kXcd, r = n % kXcd;
    const unsigned x = i % kXcd, k = i / kXcd;
    return x * q + min(x, r) + k;
}

// FFT plan: radices whose product is K (each in {2,3,4,5}).
struct FFTPlan {
    int K = 0;
    int nst = 0;
    int radix[16] = {0};
};
// order: 0 radix 8 first, then 4, 2, 3, 5; 1 radix 4 first (no radix 8);
// 2 radices 3 and 5 first, then 8, 4, 2
bool make_fft_plan(int K, FFTPlan* plan, int order = 0);

}  // namespace sdt

namespace sdt {

// Per-stage device timing with HIP events recorded on the launching stream.
// Stages: 0 spread, 1 forward x/y FFT, 2 z FFT + influence, 3 inverse y/x
// FFT, 4 interp, 5 real space, 6 Krylov block products, 7 steric (pair list
// + force), 8 mobility update (sort + pair list).
struct StageTimer {
    static constexpr int kStages = 10;
    bool on = false;
    double sum_ms[kStages] = {0};
    long long count[kStages] = {0};
    long long cols[kStages] = {0};   // right-hand sides over the launches
    struct Rec { int stage; long long cols; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        SDT_HIP_CHECK(hipEventCreate(&e));
        return e;
    }
    hipEvent_t begin(hipStream_t s) {
        hipEvent_t e = take();
        SDT_HIP_CHECK(hipEventRecord(e, s));
        return e;
    }
    void end(int stage, hipEvent_t a, hipStream_t s, long long ncols) {
        hipEvent_t b = take();
        SDT_HIP_CHECK(hipEventRecord(b, s));
        pending.push_back(Rec{stage, ncols, a, b});
    }
    // call after the stream has been synchronised
    void collect() {
        for (const Rec& r : pending) {
            float ms = 0.f;
            SDT_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
            sum_ms[r.stage] += ms;
            count[r.stage] += 1;
            cols[r.stage] += r.cols;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    void reset() {
        for (int i = 0; i < kStages; i++) { sum_ms[i] = 0.0; count[i] = 0; cols[i] = 0; }
    }
    ~StageTimer() {
        for (const Rec& r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// RAII stage scope: no-op when the timer is null or off
struct StageScope {
    StageTimer* t;
    int stage;
    hipStream_t s;
    long long ncols;
    hipEvent_t a = nullptr;
    // ncols: right-hand sides the launch group processes (0 when not
    // meaningful)
    StageScope(StageTimer* t_, int stage_, hipStream_t s_, long long ncols_ = 0)
        : t(t_), stage(stage_), s(s_), ncols(ncols_) {
        if (t && t->on) a = t->begin(s);
    }
    ~StageScope() noexcept(false) {
        if (t && t->on && a) t->end(stage, a, s, ncols);
    }
};

}  // namespace sdt
