// Host-side helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/espec_hip.h"

#define ESPEC_HIP_TRY(expr)                                                        \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess) {                                                     \
            std::fprintf(stderr, "espec_hip: %s failed: %s (%s:%d)\n", #expr,     \
                         hipGetErrorString(_e), __FILE__, __LINE__);                \
            return _e == hipErrorOutOfMemory ? ESPEC_ERR_NOMEM : ESPEC_ERR_HIP;     \
        }                                                                           \
    } while (0)

namespace espec {

// Interleave split (re, im) host arrays into V2<R> layout.
template <typename R>
std::vector<R> interleave(const double *re, const double *im, size_t n)
{
    std::vector<R> out(2 * n);
    for (size_t i = 0; i < n; ++i) {
        out[2 * i] = (R)re[i];
        out[2 * i + 1] = (R)(im ? im[i] : 0.0);
    }
    return out;
}

// (cos, sin) of c * t[i], computed in FP64 and rounded once.
template <typename R>
std::vector<R> phase_table(const double *t, size_t n, double c)
{
    std::vector<R> out(2 * n);
    for (size_t i = 0; i < n; ++i) {
        double th = c * t[i];
        out[2 * i] = (R)std::cos(th);
        out[2 * i + 1] = (R)std::sin(th);
    }
    return out;
}

// RAII device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t b)
    {
        release();
        bytes = b;
        return hipMalloc(&p, b ? b : 16);
    }
    hipError_t upload(const void *h, size_t b)
    {
        hipError_t e = alloc(b);
        if (e != hipSuccess) return e;
        return hipMemcpy(p, h, b, hipMemcpyHostToDevice);
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

inline int ntot(int ndim, const int *np)
{
    int n = 1;
    for (int a = 0; a < ndim; ++a) n *= np[a];
    return n;
}

}  // namespace espec
