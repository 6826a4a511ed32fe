// Register-resident line FFTs ("pencils") of length N = 16 M, M = 4, 8 or
// 16: M lanes of one wavefront hold a line, 16 elements per lane, and a
// transform is two rounds of in-register DFTs with one LDS exchange between
// them (one wave-local round trip instead of one per radix-4 stage).
//
// Index split (s = +1 forward, -1 inverse; W = exp(2 pi i / N)):
//   n = M n1 + n2,  k = k1 + 16 k2,   n1, k1 in [0,16),  n2, k2 in [0,M)
//   X[k1 + 16 k2] = sum_n2 e^{s 2pi i n2 k2 / M} W^{s n2 k1}
//                         sum_n1 x[M n1 + n2] e^{s 2pi i n1 k1 / 16}
// Natural layout: lane j holds x[M r + j] in register r.
// Spectral layout: lane j holds, in register r = q M + k2 (q in [0,16/M)),
// X[k] with k = (16/M) j + q + 16 k2; it is stored at line position
// p = r M + j, so that every register is a coalesced row of M lanes in
// both layouts.  spectral_index() gives k for a position p.
#pragma once
#include "common.hpp"
#include "fft_lds.hpp"

namespace espec {
namespace pencil {

// Frequency index held at spectral-layout position p of a line of 16 M.
inline int spectral_index(int p, int m)
{
    const int r = p / m, j = p % m, q = r / m, k2 = r % m;
    return j * (16 / m) + q + 16 * k2;
}

// Orders this wavefront's LDS accesses around an exchange inside its own
// slots: a wavefront's LDS instructions execute in issue order, so the
// compiler only has to keep the writes before the reads (no workgroup
// barrier, which would also stall the other wavefronts of a workgroup).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Scalar type of a complex register type.
template <typename V> struct ScalarOf;
template <> struct ScalarOf<double2> { using T = double; };
template <> struct ScalarOf<float2> { using T = float; };

// Single-precision arithmetic with split constants.  A constant c is
// carried as hi + lo (hi = float(c), lo = float(c - hi)), exact to about
// 2^-48, so that a product with it has only data-dependent roundings.  A
// constant rounded once to float would scale every step's result by the
// same factor, and the norm would drift linearly with the number of steps.
// Complex constants travel as float4 (hi.re, hi.im, lo.re, lo.im).
__device__ __forceinline__ float2 cmul_split(float2 a, float4 w)
{
    float re = fmaf(-a.y, w.w, a.x * w.z);
    re = fmaf(-a.y, w.y, re);
    re = fmaf(a.x, w.x, re);
    float im = fmaf(a.y, w.z, a.x * w.w);
    im = fmaf(a.y, w.x, im);
    im = fmaf(a.x, w.y, im);
    return {re, im};
}
// a * conj(w)
__device__ __forceinline__ float2 cmulc_split(float2 a, float4 w)
{
    float re = fmaf(a.y, w.w, a.x * w.z);
    re = fmaf(a.y, w.y, re);
    re = fmaf(a.x, w.x, re);
    float im = fmaf(-a.x, w.w, a.y * w.z);
    im = fmaf(-a.x, w.y, im);
    im = fmaf(a.y, w.x, im);
    return {re, im};
}
// (cos, sin) in double -> split float4.
__device__ __forceinline__ float4 split4(double c, double s)
{
    const float ch = (float)c, sh = (float)s;
    return {ch, sh, (float)(c - (double)ch), (float)(s - (double)sh)};
}

// Product with a constant in the arithmetic's constant form: double2
// constants for double2 data, split float4 constants for float2 data.
__device__ __forceinline__ double2 cmul_w(double2 a, double2 w) { return cmul(a, w); }
__device__ __forceinline__ double2 cmulc_w(double2 a, double2 w) { return cmulc(a, w); }
__device__ __forceinline__ float2 cmul_w(float2 a, float4 w) { return cmul_split(a, w); }
__device__ __forceinline__ float2 cmulc_w(float2 a, float4 w) { return cmulc_split(a, w); }
template <int SIGN, typename V, typename W> __device__ __forceinline__ V ctw_w(V a, W w)
{
    return SIGN > 0 ? cmul_w(a, w) : cmulc_w(a, w);
}

// Constant (table) type for complex register type V.
template <typename V> struct ConstOf;
template <> struct ConstOf<double2> { using T = double2; };
template <> struct ConstOf<float2> { using T = float4; };

// a * exp(SIGN i pi e / 8) for compile-time e.
template <int SIGN, int E, typename V>
__device__ __forceinline__ V rot16(V a)
{
    using T = typename ScalarOf<V>::T;
    constexpr int e = ((SIGN > 0 ? E : -E) % 16 + 16) % 16;
    constexpr double c8 = 0.92387953251128675613, s8 = 0.38268343236508977173,
                     r2 = 0.70710678118654752440;
    constexpr double cs[16] = {1, c8, r2, s8, 0, -s8, -r2, -c8, -1, -c8, -r2, -s8, 0, s8, r2, c8};
    constexpr double sn[16] = {0, s8, r2, c8, 1, c8, r2, s8, 0, -s8, -r2, -c8, -1, -c8, -r2, -s8};
    if constexpr (e == 0) return a;
    else if constexpr (e == 4) return {-a.y, a.x};
    else if constexpr (e == 8) return {-a.x, -a.y};
    else if constexpr (e == 12) return {a.y, -a.x};
    else if constexpr (sizeof(T) == 4) {
        constexpr float ch = (float)cs[e], sh = (float)sn[e];
        constexpr float cl = (float)(cs[e] - (double)ch), sl = (float)(sn[e] - (double)sh);
        return cmul_split(a, float4{ch, sh, cl, sl});
    } else {
        constexpr T c = (T)cs[e], t = (T)sn[e];
        return {a.x * c - a.y * t, a.x * t + a.y * c};
    }
}

// In-register DFT of length L in {4, 8, 16} on v[O], v[O+1], ..., natural
// order in and out, sign SIGN, unnormalised.
template <int SIGN, int O, typename V>
__device__ __forceinline__ void dft4(V *v)
{
    bfly4<SIGN>(v[O], v[O + 1], v[O + 2], v[O + 3]);
}

template <int SIGN, int O, typename V>
__device__ __forceinline__ void dft8(V *v)
{
    // n = 2 n1 + n2: DFT4 over n1, twiddle e^{s 2pi i n2 k1 / 8}, DFT2 over n2.
    bfly4<SIGN>(v[O + 0], v[O + 2], v[O + 4], v[O + 6]);
    bfly4<SIGN>(v[O + 1], v[O + 3], v[O + 5], v[O + 7]);
    v[O + 3] = rot16<SIGN, 2>(v[O + 3]);
    v[O + 5] = rot16<SIGN, 4>(v[O + 5]);
    v[O + 7] = rot16<SIGN, 6>(v[O + 7]);
    bfly2(v[O + 0], v[O + 1]);
    bfly2(v[O + 2], v[O + 3]);
    bfly2(v[O + 4], v[O + 5]);
    bfly2(v[O + 6], v[O + 7]);
    // position 2 k1 + k2 holds X[k1 + 4 k2]
    V t[8];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 2; ++k2) t[k1 + 4 * k2] = v[O + 2 * k1 + k2];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[O + i] = t[i];
}

template <int SIGN, int O, typename V>
__device__ __forceinline__ void dft16(V *v)
{
    // n = 4 n1 + n2: DFT4 over n1, twiddle e^{s 2pi i n2 k1 / 16}, DFT4 over n2.
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) bfly4<SIGN>(v[O + n2], v[O + 4 + n2], v[O + 8 + n2], v[O + 12 + n2]);
    v[O + 5] = rot16<SIGN, 1>(v[O + 5]);
    v[O + 6] = rot16<SIGN, 2>(v[O + 6]);
    v[O + 7] = rot16<SIGN, 3>(v[O + 7]);
    v[O + 9] = rot16<SIGN, 2>(v[O + 9]);
    v[O + 10] = rot16<SIGN, 4>(v[O + 10]);
    v[O + 11] = rot16<SIGN, 6>(v[O + 11]);
    v[O + 13] = rot16<SIGN, 3>(v[O + 13]);
    v[O + 14] = rot16<SIGN, 6>(v[O + 14]);
    v[O + 15] = rot16<SIGN, 9>(v[O + 15]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) bfly4<SIGN>(v[O + 4 * k1], v[O + 4 * k1 + 1], v[O + 4 * k1 + 2], v[O + 4 * k1 + 3]);
    // position 4 k1 + k2 holds X[k1 + 4 k2]
    V t[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) t[k1 + 4 * k2] = v[O + 4 * k1 + k2];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[O + i] = t[i];
}

template <int L, int SIGN, int O, typename V>
__device__ __forceinline__ void dft(V *v)
{
    static_assert(L == 4 || L == 8 || L == 16, "pencil DFT length");
    if constexpr (L == 4) dft4<SIGN, O>(v);
    else if constexpr (L == 8) dft8<SIGN, O>(v);
    else dft16<SIGN, O>(v);
}

// v[O + e] *= W^{SIGN k e} for e in [0, CNT), W = exp(2 pi i / N), from two
// table entries (W^k, W^{4k}; 4k < N) and at most three products each.
template <int SIGN, int CNT, int O, typename V>
__device__ __forceinline__ void twiddle_powers(V *v, const V *__restrict__ tw, int k)
{
    const V w1 = tw[k];
    const V w2 = cmul(w1, w1), w3 = cmul(w2, w1);
    V w4 = {1, 0}, w8 = {1, 0}, w12 = {1, 0};
    if constexpr (CNT > 4) {
        w4 = tw[4 * k];
        w8 = cmul(w4, w4);
        w12 = cmul(w8, w4);
    }
#pragma unroll
    for (int e = 1; e < CNT; ++e) {
        const V lo = (e & 3) == 1 ? w1 : (e & 3) == 2 ? w2 : w3;
        const V hi = (e >> 2) == 1 ? w4 : (e >> 2) == 2 ? w8 : w12;
        const V w = (e & 3) == 0 ? hi : (e >> 2) == 0 ? lo : cmul(hi, lo);
        v[O + e] = ctw<SIGN>(v[O + e], w);
    }
}

// v[O + e] *= W^{SIGN k e} for e in [1, CNT) from a table of split
// constants: tt[e * M] = W^{k e} (the caller offsets tt by its k's row).
template <int SIGN, int CNT, int O, int M>
__device__ __forceinline__ void twiddle_table(float2 *v, const float4 *__restrict__ tt)
{
#pragma unroll
    for (int e = 1; e < CNT; ++e) v[O + e] = ctw_w<SIGN>(v[O + e], tt[e * M]);
}

// sin and cos of th for |th| < kSincosMax: Cody-Waite reduction by pi in
// three FMA steps (pi split as P1 + P2 + P3, each rounded to double) to
// th = k pi + r, |r| <= pi/2, then the Taylor series of sin to r^19 and of
// cos to r^20 (truncation below 3e-16) and the sign (-1)^k of both.  The
// half-period reduction leaves one sign flip where a quarter-period one
// needs a four-way select; the two extra terms cost less than the select
// and the 64-bit quadrant.  Cheaper in registers than a general sincos,
// whose large-argument reduction the callers never need.
constexpr double kSincosMax = 1.0e6;
__device__ __forceinline__ void sincos_reduced(double th, double *sn, double *cs)
{
    constexpr double P1 = 3.141592653589793, P2 = 1.2246467991473532e-16,
                     P3 = -2.9947698097183397e-33, INV_PI = 0.3183098861837907;
    const double k = rint(th * INV_PI);
    double r = fma(-k, P1, th);
    r = fma(-k, P2, r);
    r = fma(-k, P3, r);
    const double r2 = r * r;
    // 1/n! for the odd (sin) and even (cos) terms
    double ps = 1.0 / 121645100408832000.0;                  // 1/19!
    ps = fma(ps, r2, -1.0 / 355687428096000.0);              // 1/17!
    ps = fma(ps, r2, 1.0 / 1307674368000.0);                 // 1/15!
    ps = fma(ps, r2, -1.0 / 6227020800.0);                   // 1/13!
    ps = fma(ps, r2, 1.0 / 39916800.0);                      // 1/11!
    ps = fma(ps, r2, -1.0 / 362880.0);                       // 1/9!
    ps = fma(ps, r2, 1.0 / 5040.0);                          // 1/7!
    ps = fma(ps, r2, -1.0 / 120.0);                          // 1/5!
    ps = fma(ps, r2, 1.0 / 6.0);                             // 1/3!
    const double s = fma(-ps * r2, r, r);                    // r - r^3/3! + ...
    double pc = 1.0 / 2432902008176640000.0;                 // 1/20!
    pc = fma(pc, r2, -1.0 / 6402373705728000.0);             // 1/18!
    pc = fma(pc, r2, 1.0 / 20922789888000.0);                // 1/16!
    pc = fma(pc, r2, -1.0 / 87178291200.0);                  // 1/14!
    pc = fma(pc, r2, 1.0 / 479001600.0);                     // 1/12!
    pc = fma(pc, r2, -1.0 / 3628800.0);                      // 1/10!
    pc = fma(pc, r2, 1.0 / 40320.0);                         // 1/8!
    pc = fma(pc, r2, -1.0 / 720.0);                          // 1/6!
    pc = fma(pc, r2, 1.0 / 24.0);                            // 1/4!
    pc = fma(pc, r2, -0.5);                                  // 1/2!
    const double c = fma(pc, r2, 1.0);
    // |k| < 2^19 here, so the conversion to int is exact; bit 63 flips both signs
    const unsigned long long flip = (unsigned long long)((int)k & 1) << 63;
    *sn = __longlong_as_double(__double_as_longlong(s) ^ (long long)flip);
    *cs = __longlong_as_double(__double_as_longlong(c) ^ (long long)flip);
}

// LDS slot of exchange element (k1, n2) of a line: row k1 of M entries,
// one pad slot per 16 so that column reads (stride M) hit distinct banks.
template <int M> __device__ __forceinline__ int xslot(int k1, int n2)
{
    const int i = k1 * M + n2;
    return i + (i >> 4);
}
// Exchange slots per line.
template <int M> constexpr int xslots() { return 17 * M; }

// Exchange element i (= k1 M + n2) of a line.  A pointer is the line's own
// block of xslots() slots (xslot's padding); a Strided view is a line whose
// elements lie pitch slots apart in a larger LDS array (a column of an
// LDS-resident grid), exchanged in place.
template <typename V> __device__ __forceinline__ V &xat(V *xl, int i) { return xl[i + (i >> 4)]; }
template <typename V> struct Strided {
    V *p;
    int pitch;
};
template <typename V> __device__ __forceinline__ V &xat(const Strided<V> &xl, int i)
{
    return xl.p[i * xl.pitch];
}

// Natural -> spectral, sign SIGN.  xl: this line's exchange slots (a
// pointer or a Strided view, see xat), which only this wavefront touches
// during the transform.  The exchange may overlay the slots the line was
// read from: one wavefront's LDS accesses complete in issue order, and
// every write here depends on all sixteen inputs.
// tw: double2 data take exp(2 pi i m / N), m in [0, N), and form the
// powers they need; float2 data take the split table of every twiddle,
// entry k1 M + j = W^{j k1} for the forward, then at N + (q M + n2) M + j
// W^{(j Q + q) n2} for the inverse.
template <int M, int SIGN, typename V, typename X>
__device__ __forceinline__ void forward(V (&v)[16], int j, X xl,
                                        const typename ConstOf<V>::T *__restrict__ tw)
{
    constexpr int Q = 16 / M;
    dft16<SIGN, 0>(v);                         // v[k1] = sum_n1 x[M n1 + j] ...
    if constexpr (sizeof(v[0].x) == 4)
        twiddle_table<SIGN, 16, 0, M>(v, tw + j);  // * W^{s j k1}
    else
        twiddle_powers<SIGN, 16, 0>(v, tw, j);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) xat(xl, k1 * M + j) = v[k1];
    wave_sync();
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int n2 = 0; n2 < M; ++n2) v[q * M + n2] = xat(xl, (j * Q + q) * M + n2);
#pragma unroll
    for (int q = 0; q < Q; ++q) dft<M, SIGN, 0>(v + q * M);
}

// Spectral -> natural, sign SIGN (exact inverse of forward<M, -SIGN> up to
// the factor N).
template <int M, int SIGN, typename V, typename X>
__device__ __forceinline__ void inverse(V (&v)[16], int j, X xl,
                                        const typename ConstOf<V>::T *__restrict__ tw)
{
    constexpr int Q = 16 / M;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        dft<M, SIGN, 0>(v + q * M);
        if constexpr (sizeof(v[0].x) == 4)
            twiddle_table<SIGN, M, 0, M>(v + q * M, tw + 16 * M + q * M * M + j);
        else
            twiddle_powers<SIGN, M, 0>(v + q * M, tw, j * Q + q);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int n2 = 0; n2 < M; ++n2) xat(xl, (j * Q + q) * M + n2) = v[q * M + n2];
    wave_sync();
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[k1] = xat(xl, k1 * M + j);
    dft16<SIGN, 0>(v);
}

}  // namespace pencil

// Wide lines: a line of L = 64, 128 or 256 points on L / 4 lanes, four
// points per lane (lane j holds x[j + (L/4) r] in register r), so that a
// wavefront holds 256 / L lines.  A transform is radix-4 Stockham stages
// (Govindaraju et al., SC 2008, "High performance discrete Fourier
// transforms on graphics processors", radix-R Stockham iteration), with
// a final radix-2 stage for L = 128 and an LDS exchange between stages.
// Natural order in and out, so no permutation is needed either way.  A
// quarter of the pencil's per-lane work on four times the lanes: for
// single packets, whose passes are the latency of one wavefront's chain.
namespace wide {

template <int L> constexpr int slots() { return L + L / 16; }
__device__ __forceinline__ int slot(int i) { return i + (i >> 4); }

// X[k] = sum_n x[n] exp(SIGN 2 pi i n k / L), unnormalised; tw[m] =
// exp(2 pi i m / L).  Radix-4 stage s (ns = 4^s): lane j twiddles its
// inputs by exp(SIGN 2 pi i r (j mod ns) / (4 ns)), takes a 4-point DFT and
// writes output r to position (j / ns) 4 ns + (j mod ns) + r ns.  L = 128
// ends with a radix-2 stage (ns = 64) on the register pairs (0, 2) and
// (1, 3), butterflies j and j + 32.
template <int L, int SIGN>
__device__ __forceinline__ void stockham(double2 (&v)[4], int j, double2 *xl,
                                         const double2 *__restrict__ tw)
{
    static_assert(L == 64 || L == 128 || L == 256, "wide line length");
    constexpr int LW = L / 4, S4 = L == 256 ? 4 : 3;
#pragma unroll
    for (int s = 0; s < S4; ++s) {
        const int ns = 1 << (2 * s);
        const int k = j & (ns - 1);
        if (s > 0) {
            const int base = k * (LW >> (2 * s));
            v[1] = ctw<SIGN>(v[1], tw[base]);
            v[2] = ctw<SIGN>(v[2], tw[2 * base]);
            v[3] = ctw<SIGN>(v[3], tw[3 * base]);
        }
        bfly4<SIGN>(v[0], v[1], v[2], v[3]);
        if (s == S4 - 1 && L != 128) break;  // the last stage writes position j + LW r
        const int d = ((j >> (2 * s)) << (2 * s + 2)) + k;
        pencil::wave_sync();  // the previous stage's reads of xl are done
#pragma unroll
        for (int r = 0; r < 4; ++r) xl[slot(d + r * ns)] = v[r];
        pencil::wave_sync();
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = xl[slot(j + LW * r)];
    }
    if constexpr (L == 128) {
        const double2 a = ctw<SIGN>(v[2], tw[j]), b = ctw<SIGN>(v[3], tw[j + 32]);
        v[2] = csub(v[0], a);
        v[0] = cadd(v[0], a);
        v[3] = csub(v[1], b);
        v[1] = cadd(v[1], b);
    }
}

}  // namespace wide
}  // namespace espec
