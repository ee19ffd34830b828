// stokesdt_amd internal definitions: errors, device buffers, complex types.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
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
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    // grow-only allocation; contents are not preserved
    void reserve(size_t count) {
        if (count <= n && ptr) return;
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
};

template <typename T>
struct cplx {
    T x, y;
};

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// FFT plan: radices whose product is K (each in {2,3,4,5}).
struct FFTPlan {
    int K = 0;
    int nst = 0;
    int radix[16] = {0};
};
bool make_fft_plan(int K, FFTPlan* plan);

}  // namespace sdt
