// Host-side helpers shared by the library's translation units: error
// reporting through the C ABI and owned device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <exception>
#include <string>

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

inline thread_local std::string g_err;  // message of the last failed C-ABI call (per thread)

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(const std::string &m, int code = 1) { throw Failure{code, m}; }

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            fail(std::string(#expr) + ": " + hipGetErrorString(e_), 2);                   \
    } while (0)

template <class F>
inline int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const Failure &e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 3;
    }
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 16;
        HIPCHK(hipMalloc(&p, n));
        bytes = n;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

