// In-LDS FFT stages for one workgroup.  A stage of an axis processes all
// lines of that axis; each thread takes whole butterflies.  Radix 2 and 4,
// decimation in frequency forward (natural in, digit-reversed out) and its
// exact stage-by-stage inverse (digit-reversed in, natural out).
#pragma once
#include "common.hpp"

namespace espec {

template <int SIGN, typename V>
__device__ __forceinline__ void bfly4(V &a0, V &a1, V &a2, V &a3)
{
    V akp = cadd(a0, a2), akm = csub(a0, a2), ajp = cadd(a1, a3), ajm = csub(a1, a3);
    a0 = cadd(akp, ajp);
    a2 = csub(akp, ajp);
    V ijm = {-ajm.y, ajm.x};  // i * ajm
    if (SIGN > 0) {
        a1 = cadd(akm, ijm);
        a3 = csub(akm, ijm);
    } else {
        a1 = csub(akm, ijm);
        a3 = cadd(akm, ijm);
    }
}

template <typename V> __device__ __forceinline__ void bfly2(V &a0, V &a1)
{
    V t = csub(a0, a1);
    a0 = cadd(a0, a1);
    a1 = t;
}

// Geometry of butterfly g of stage s: LDS index of its first element and
// the element step between its r inputs; j is the twiddle offset.
struct Bfly {
    int base, step, j, e0;
};

// LDS position of element i.  With PAD, one slot is skipped after every 16
// elements: butterflies whose inputs are 16 or 64 elements apart (the late
// stages, where lanes step from block to block) then fall in different
// banks instead of 4-way conflicts on the same 256-byte bank row.
template <bool PAD> __device__ __forceinline__ int lx(int i) { return PAD ? i + (i >> 4) : i; }
// LDS slots to reserve for n padded elements.
constexpr int lds_slots(int n, bool pad) { return pad ? n + (n >> 4) + 1 : n; }

// Threads walk butterflies of one line first when the line is contiguous
// (es == 1) and across lines first otherwise, so that neighbouring lanes
// touch neighbouring LDS words in both cases.
__device__ __forceinline__ Bfly bfly_at(const DevAxis &ax, int s, int g, int lnb)
{
    const int lm = ax.lm[s];
    int l, bf;
    if (ax.es == 1) {
        l = g >> lnb;
        bf = g & ((1 << lnb) - 1);
    } else {
        l = g & (ax.nl - 1);
        bf = g >> ax.lnl;
    }
    Bfly b;
    b.j = bf & ((1 << lm) - 1);
    b.e0 = ((bf >> lm) << ax.lL[s]) + b.j;
    b.base = line_base(ax, l) + (b.e0 << ax.les);
    b.step = 1 << (lm + ax.les);
    return b;
}

// One forward DIF stage: butterfly, then twiddle W^{q j} on output q.
template <typename R, int NT, int SIGN, bool PAD = false>
__device__ void dif_stage(V2<R> *x, const DevAxis &ax, int s, const V2<R> *__restrict__ tw)
{
    using V = V2<R>;
    const int r = ax.rad[s], m = 1 << ax.lm[s];
    const int lnb = ax.ln - (r == 4 ? 2 : 1);
    const int total = ax.nl << lnb;
    const V *t = tw + ax.tw0 + ax.off[s];
    for (int g = threadIdx.x; g < total; g += NT) {
        Bfly b = bfly_at(ax, s, g, lnb);
        if (r == 4) {
            V a0 = x[lx<PAD>(b.base)], a1 = x[lx<PAD>(b.base + b.step)], a2 = x[lx<PAD>(b.base + 2 * b.step)],
              a3 = x[lx<PAD>(b.base + 3 * b.step)];
            bfly4<SIGN>(a0, a1, a2, a3);
            if (b.j) {
                a1 = ctw<SIGN>(a1, t[b.j]);
                a2 = ctw<SIGN>(a2, t[m + b.j]);
                a3 = ctw<SIGN>(a3, t[2 * m + b.j]);
            }
            x[lx<PAD>(b.base)] = a0;
            x[lx<PAD>(b.base + b.step)] = a1;
            x[lx<PAD>(b.base + 2 * b.step)] = a2;
            x[lx<PAD>(b.base + 3 * b.step)] = a3;
        } else {
            V a0 = x[lx<PAD>(b.base)], a1 = x[lx<PAD>(b.base + b.step)];
            bfly2(a0, a1);
            if (b.j) a1 = ctw<SIGN>(a1, t[b.j]);
            x[lx<PAD>(b.base)] = a0;
            x[lx<PAD>(b.base + b.step)] = a1;
        }
    }
    __syncthreads();
}

// Inverse of a DIF stage of the opposite sign: twiddle first, then the
// radix-r DFT with sign SIGN (unnormalised).
template <typename R, int NT, int SIGN, bool PAD = false>
__device__ void dit_stage(V2<R> *x, const DevAxis &ax, int s, const V2<R> *__restrict__ tw)
{
    using V = V2<R>;
    const int r = ax.rad[s], m = 1 << ax.lm[s];
    const int lnb = ax.ln - (r == 4 ? 2 : 1);
    const int total = ax.nl << lnb;
    const V *t = tw + ax.tw0 + ax.off[s];
    for (int g = threadIdx.x; g < total; g += NT) {
        Bfly b = bfly_at(ax, s, g, lnb);
        if (r == 4) {
            V a0 = x[lx<PAD>(b.base)], a1 = x[lx<PAD>(b.base + b.step)], a2 = x[lx<PAD>(b.base + 2 * b.step)],
              a3 = x[lx<PAD>(b.base + 3 * b.step)];
            if (b.j) {
                a1 = ctw<SIGN>(a1, t[b.j]);
                a2 = ctw<SIGN>(a2, t[m + b.j]);
                a3 = ctw<SIGN>(a3, t[2 * m + b.j]);
            }
            bfly4<SIGN>(a0, a1, a2, a3);
            x[lx<PAD>(b.base)] = a0;
            x[lx<PAD>(b.base + b.step)] = a1;
            x[lx<PAD>(b.base + 2 * b.step)] = a2;
            x[lx<PAD>(b.base + 3 * b.step)] = a3;
        } else {
            V a0 = x[lx<PAD>(b.base)], a1 = x[lx<PAD>(b.base + b.step)];
            if (b.j) a1 = ctw<SIGN>(a1, t[b.j]);
            bfly2(a0, a1);
            x[lx<PAD>(b.base)] = a0;
            x[lx<PAD>(b.base + b.step)] = a1;
        }
    }
    __syncthreads();
}

// Stage 0 of axis 0 fused with the elementwise work between two steps:
// optionally the inverse stage (SIGN -1) that completes the previous
// step, then op(index, value) on each natural-order element, then
// optionally the forward stage (SIGN +1) that starts the next step.
template <typename R, int NT, bool INV, bool FWD, bool PAD = false, class Op>
__device__ void fused_outer(V2<R> *x, const DevAxis &ax, const V2<R> *__restrict__ tw, Op op)
{
    using V = V2<R>;
    const int s = 0, r = ax.rad[0], m = 1 << ax.lm[0];
    const int lnb = ax.ln - (r == 4 ? 2 : 1);
    const int total = ax.nl << lnb;
    const V *t = tw + ax.tw0 + ax.off[0];
    for (int g = threadIdx.x; g < total; g += NT) {
        Bfly b = bfly_at(ax, s, g, lnb);
        if (r == 4) {
            int i0 = b.base, i1 = i0 + b.step, i2 = i1 + b.step, i3 = i2 + b.step;
            V a0 = x[lx<PAD>(i0)], a1 = x[lx<PAD>(i1)], a2 = x[lx<PAD>(i2)], a3 = x[lx<PAD>(i3)];
            V w1, w2, w3;
            if (b.j) {
                w1 = t[b.j];
                w2 = t[m + b.j];
                w3 = t[2 * m + b.j];
            }
            if (INV) {
                if (b.j) {
                    a1 = ctw<-1>(a1, w1);
                    a2 = ctw<-1>(a2, w2);
                    a3 = ctw<-1>(a3, w3);
                }
                bfly4<-1>(a0, a1, a2, a3);
            }
            op(i0, a0);
            op(i1, a1);
            op(i2, a2);
            op(i3, a3);
            if (FWD) {
                bfly4<1>(a0, a1, a2, a3);
                if (b.j) {
                    a1 = ctw<1>(a1, w1);
                    a2 = ctw<1>(a2, w2);
                    a3 = ctw<1>(a3, w3);
                }
            }
            x[lx<PAD>(i0)] = a0;
            x[lx<PAD>(i1)] = a1;
            x[lx<PAD>(i2)] = a2;
            x[lx<PAD>(i3)] = a3;
        } else {
            int i0 = b.base, i1 = i0 + b.step;
            V a0 = x[lx<PAD>(i0)], a1 = x[lx<PAD>(i1)];
            V w1;
            if (b.j) w1 = t[b.j];
            if (INV) {
                if (b.j) a1 = ctw<-1>(a1, w1);
                bfly2(a0, a1);
            }
            op(i0, a0);
            op(i1, a1);
            if (FWD) {
                bfly2(a0, a1);
                if (b.j) a1 = ctw<1>(a1, w1);
            }
            x[lx<PAD>(i0)] = a0;
            x[lx<PAD>(i1)] = a1;
        }
    }
    __syncthreads();
}

// Last stage (m = 1, no twiddles) of the last axis fused with the
// momentum-space multiply and the first inverse stage.
template <typename R, int NT, bool PAD = false, class Op>
__device__ void fused_inner(V2<R> *x, const DevAxis &ax, Op op)
{
    using V = V2<R>;
    const int s = ax.nst - 1, r = ax.rad[s];
    const int lnb = ax.ln - (r == 4 ? 2 : 1);
    const int total = ax.nl << lnb;
    for (int g = threadIdx.x; g < total; g += NT) {
        Bfly b = bfly_at(ax, s, g, lnb);
        if (r == 4) {
            int i0 = b.base, i1 = i0 + b.step, i2 = i1 + b.step, i3 = i2 + b.step;
            V a0 = x[lx<PAD>(i0)], a1 = x[lx<PAD>(i1)], a2 = x[lx<PAD>(i2)], a3 = x[lx<PAD>(i3)];
            bfly4<1>(a0, a1, a2, a3);
            op(i0, a0);
            op(i1, a1);
            op(i2, a2);
            op(i3, a3);
            bfly4<-1>(a0, a1, a2, a3);
            x[lx<PAD>(i0)] = a0;
            x[lx<PAD>(i1)] = a1;
            x[lx<PAD>(i2)] = a2;
            x[lx<PAD>(i3)] = a3;
        } else {
            int i0 = b.base, i1 = i0 + b.step;
            V a0 = x[lx<PAD>(i0)], a1 = x[lx<PAD>(i1)];
            bfly2(a0, a1);
            op(i0, a0);
            op(i1, a1);
            bfly2(a0, a1);
            x[lx<PAD>(i0)] = a0;
            x[lx<PAD>(i1)] = a1;
        }
    }
    __syncthreads();
}

// Gather along an axis: y[line + k es] = x[line + perm[k] es] (undoes the
// digit reversal of a full DIF pass on that axis).
template <typename R, int NT, bool PAD = false>
__device__ void permute_axis(const V2<R> *x, V2<R> *y, const DevAxis &ax, const int *__restrict__ perm)
{
    const int total = ax.nl << ax.ln;
    for (int g = threadIdx.x; g < total; g += NT) {
        int l, k;
        if (ax.es == 1) {
            l = g >> ax.ln;
            k = g & (ax.n - 1);
        } else {
            l = g & (ax.nl - 1);
            k = g >> ax.lnl;
        }
        int lb = line_base(ax, l);
        y[lx<PAD>(lb + (k << ax.les))] = x[lx<PAD>(lb + (perm[k] << ax.les))];
    }
    __syncthreads();
}

}  // namespace espec
