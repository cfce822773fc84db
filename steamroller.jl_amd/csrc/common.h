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
    size_t bytes = 0;  // size in use
    size_t cap = 0;    // size allocated
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), cap(o.cap) {
        o.p = nullptr;
        o.bytes = o.cap = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            cap = o.cap;
            o.p = nullptr;
            o.bytes = o.cap = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = cap = 0;
    }
    // n bytes; the allocation is kept when it holds n and wastes at most a
    // quarter of n (regrids change sizes by a few percent); contents are not
    // preserved either way
    void alloc(size_t n) {
        if (n == 0) n = 16;
        if (p && cap >= n && cap - n <= n / 4) {
            bytes = n;
            return;
        }
        release();
        HIPCHK(hipMalloc(&p, n));
        bytes = cap = n;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

