// Shared device helpers: complex arithmetic on double2/float2, FFT stage
// descriptors, line geometry and FP64 block reductions.
#pragma once
#include <hip/hip_runtime.h>

#include "tables.hpp"

namespace espec {

template <typename R> struct V2Of;
template <> struct V2Of<double> { using T = double2; };
template <> struct V2Of<float> { using T = float2; };
template <typename R> using V2 = typename V2Of<R>::T;

template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return {a.x + b.x, a.y + b.y}; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return {a.x - b.x, a.y - b.y}; }
// a * w
template <typename V> __device__ __forceinline__ V cmul(V a, V w)
{
    return {a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
}
// a * conj(w): with w = (cos t, sin t) this is a * exp(-i t), the phase
// convention of the reference (U,V) <- (c U + s V, c V - s U) (spofft.f:121-125).
template <typename V> __device__ __forceinline__ V cmulc(V a, V w)
{
    return {a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y};
}
template <typename V, typename R> __device__ __forceinline__ V cscale(V a, R s) { return {a.x * s, a.y * s}; }
// SIGN > 0: multiply by w, else by conj(w).
template <int SIGN, typename V> __device__ __forceinline__ V ctw(V a, V w)
{
    return SIGN > 0 ? cmul(a, w) : cmulc(a, w);
}

// Stage plan of one axis on the device (see FftPlan).
// All sizes are powers of two, so the geometry is kept as shift counts:
// element e of line l sits at base(l) + (e << les), with
// base(l) = ((l & (lmod-1)) << lsa) + ((l >> llmod) << lsb).
struct DevAxis {
    int n, ln, nst, es, les;        // length, log2 length, stages, element stride, log2 stride
    int nl, lnl, lmod, llmod, lsa, lsb;
    int tw0;                        // offset of this axis' twiddles in the table
    int rad[kMaxStages], lm[kMaxStages], lL[kMaxStages], off[kMaxStages];
};

__device__ __forceinline__ int line_base(const DevAxis &a, int l)
{
    return ((l & (a.lmod - 1)) << a.lsa) + ((l >> a.llmod) << a.lsb);
}

// Host-side: fill the shift form of a line geometry given as
// base(l) = (l % lmod) * la + (l / lmod) * lb (la, lb powers of two or 0).
inline void set_line_geometry(DevAxis &d, int es, int nl, int lmod, int la, int lb)
{
    d.es = es;
    d.les = ilog2(es);
    d.nl = nl;
    d.lnl = ilog2(nl);
    d.lmod = lmod;
    d.llmod = ilog2(lmod);
    d.lsa = la > 0 ? ilog2(la) : 0;
    d.lsb = lb > 0 ? ilog2(lb) : 0;
}

// Wave-level FP64 sum (64 lanes).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of three FP64 values; the result is valid in thread 0.
template <int NT>
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *red)
{
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    if (NT > 64) {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        __syncthreads();
        if (lane == 0) {
            red[3 * w] = a;
            red[3 * w + 1] = b;
            red[3 * w + 2] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < NT / 64; ++k) {
                a += red[3 * k];
                b += red[3 * k + 1];
                c += red[3 * k + 2];
            }
        }
        __syncthreads();
    }
}

}  // namespace espec
