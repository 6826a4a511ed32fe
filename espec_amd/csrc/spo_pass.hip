// Multi-pass split-operator step for grids larger than one workgroup's LDS.
//
// Per step (2-D shown; 3-D adds a z pass like the y pass):
//   row pass  : [inverse x stages of step k]  B e^{-iV dt/2}  (C(t) sample)
//               e^{-iV dt/2}  [forward x stages of step k+1]  e^{-iTx dt}
//               -> state stored with x digit-reversed, y natural
//   y pass    : forward y, e^{-iTy dt}, inverse y on tiles of contiguous
//               x positions, y natural again
// Since e^{-i(Tx+Ty)dt} = e^{-iTx dt} e^{-iTy dt}, the kinetic phase splits
// per axis and each pass keeps one axis of the transform.  The state makes
// one round trip through L2/MALL per pass; the V phases ride on the row
// pass and the kinetic phases on the butterflies.  Each kernel issues all
// of its global loads (state, V, psi0, twiddles, kinetic phases) in one
// prologue and then works from LDS only, so a pass costs one memory
// latency rather than one per FFT stage.  C(t) partial sums are written
// per row block and reduced after the run (deterministic order).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include <hip/hip_ext.h>

#include "api_common.hpp"
#include "fft_lds.hpp"
#include "pencil.hpp"
#include "spo_pass.hpp"

namespace espec {

constexpr int kMaxEpt = 16;  // elements per thread held in registers by the prologue

struct PassArgs {
    int n, npk, np0, xblocks, cnt, ept;
    DevAxis ax;            // LDS geometry of this pass (twiddles at offset 0)
    const double2 *tw;     // this axis' twiddles, ntw entries
    int ntw;
    const double2 *kt;     // this axis' kinetic phases (digit-reversed order)
    void *x;               // state, storage precision
    const void *psi0;      // initial state, storage precision
    const void *vpot;      // potential, storage precision
    const double2 *vph;    // exp(-i cst2 V) as (cos, sin), FP64 (phase-table variant)
    double *part;
    long long *dstep;
    double b, cst2;
    int lc, maxsamp;
    int S, L, ltile, last;
    int spread;            // only workgroups with blockIdx.x % spread == 0 work (XCD placement)
    int ktps;              // kinetic table stride between packets (0: shared)
    int xmap;              // strided pass: give each XCD a contiguous range of tiles
    const double2 *twr;    // pencil kernels: exp(2 pi i m / L), m in [0, L)
    const double2 *kreg;   // pencil kernels: kinetic phases in spectral-layout order
    const float4 *twr_s;   // single-precision pencils: split twiddle table (2 L per axis)
    const float4 *kreg_s;  // and split kinetic phases (kreg's layout)
    const double2 *knat;   // wide lines: kinetic phases in natural frequency order
    const double *mask;    // [npk][n] .SMOOTHW mask after the closing V half phase, or null
                           // (row pass on LDS stages only)
};

// Constants of the pencil kernels in the form their arithmetic takes:
// double2 for FP64 storage, split float4 (hi, lo) for FP32 storage.
template <typename S> using PenW = typename pencil::ConstOf<V2<S>>::T;
template <typename S> __device__ __forceinline__ const PenW<S> *pen_twr(const PassArgs &a);
template <> __device__ __forceinline__ const double2 *pen_twr<double>(const PassArgs &a) { return a.twr; }
template <> __device__ __forceinline__ const float4 *pen_twr<float>(const PassArgs &a) { return a.twr_s; }
template <typename S> __device__ __forceinline__ const PenW<S> *pen_kreg(const PassArgs &a);
template <> __device__ __forceinline__ const double2 *pen_kreg<double>(const PassArgs &a) { return a.kreg; }
template <> __device__ __forceinline__ const float4 *pen_kreg<float>(const PassArgs &a) { return a.kreg_s; }

template <typename S>
__device__ __forceinline__ double2 widen(V2<S> v)
{
    return {(double)v.x, (double)v.y};
}

// S: storage precision of psi, psi0 and V in HBM; LDS and arithmetic FP64.
// EPT: elements per thread (compile time so that the prologue's staging
// arrays stay in VGPRs); twiddle and kinetic tables have at most EPT*NT entries.
// PHT: read the FP64 V phase table (16 B/point) instead of V in storage
// precision plus an FP64 sincos per point; the table wins when the pass is
// latency-bound (one packet), V+sincos when it is HBM-bound (large batches).
//
// Work items (row blocks over all packets) are taken in a grid-stride loop:
// with fewer workgroups than items, each workgroup issues the global loads
// of its next item into registers before it transforms the current one
// in LDS, so HBM latency overlaps the FFT.
// Copies a table of cnt double2 entries to LDS (slot cnt takes the
// clamped lanes).  Clamped indices and unconditional stores keep the
// staging in VGPRs and every load used, so no load is sunk into a branch.
template <int NT, int EPT>
__device__ __forceinline__ void stage_table(double2 *dst, const double2 *src, int cnt)
{
    double tr[EPT], ti[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int j = threadIdx.x + k * NT;
        const double2 v = src[j < cnt ? j : 0];
        tr[k] = v.x;
        ti[k] = v.y;
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int j = threadIdx.x + k * NT;
        dst[j < cnt ? j : cnt] = {tr[k], ti[k]};
    }
}

// Register staging of one row block; loop-carried arrays are scalar
// (arrays of double2 carried across the loop are not promoted to VGPRs).
template <typename S, int NT, int EPT, bool PHT, bool INV>
__device__ __forceinline__ void rows_fetch(const PassArgs &a, int w, V2<S> (&st)[EPT],
                                           S (&vv)[EPT], double (&wc)[EPT], double (&ws)[EPT],
                                           V2<S> (&q)[EPT])
{
    const int pk = w / a.xblocks, blk = w - pk * a.xblocks;
    const long long po = (long long)pk * a.n + (long long)blk * a.cnt;
    const V2<S> *gx = reinterpret_cast<const V2<S> *>(a.x) + po;
    const S *gv = reinterpret_cast<const S *>(a.vpot) + po;
    const V2<S> *g0 = reinterpret_cast<const V2<S> *>(a.psi0) + po;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int i = threadIdx.x + k * NT;
        st[k] = gx[i];
        if (PHT) {
            const double2 t = a.vph[po + i];
            wc[k] = t.x;
            ws[k] = t.y;
        } else {
            vv[k] = gv[i];
        }
        if (INV) q[k] = g0[i];
    }
}

template <typename S, int NT, bool INV, bool FWD, int EPT, bool PHT>
__global__ __launch_bounds__(NT) void pass_rows(PassArgs a)
{
    using V = double2;
    using VS = V2<S>;
    extern __shared__ double smem_d[];
    const int cnt = a.cnt;
    V *x = reinterpret_cast<V *>(smem_d);
    V *ph = x + lds_slots(cnt, true);  // exp(+i cst2 V) as (cos, sin)
    V *p0 = ph + cnt;     // psi0
    V *twl = p0 + cnt;        // twiddles (+1 dummy slot)
    V *kl = twl + a.ntw + 1;  // kinetic phases (+1 dummy slot)
    double *red = reinterpret_cast<double *>(kl + a.np0 + 1);
    // Workgroups are dealt to the XCDs round robin by linear id; keeping
    // only every spread-th one active confines a small grid to 8/spread
    // XCDs so that its passes hand the state over inside one L2.
    if (a.spread > 1 && blockIdx.x % a.spread) return;
    const int nact = gridDim.x / a.spread, total = a.xblocks * a.npk;
    int w = blockIdx.x / a.spread;
    if (w >= total) return;
    const double b = a.b;
    const long long g = INV ? *a.dstep : 0;  // the step this pass completes
    const bool samp = INV && (g % a.lc) == 0;

    stage_table<NT, EPT>(twl, a.tw, a.ntw);  // twiddles once per workgroup
    VS st[EPT];
    S vv[EPT];
    double wc[EPT], ws[EPT];
    VS q[EPT];
    rows_fetch<S, NT, EPT, PHT, INV>(a, w, st, vv, wc, ws, q);
    int kpk = -1;  // packet whose kinetic phases are in LDS
    for (;;) {
        const int pk = w / a.xblocks, blk = w - pk * a.xblocks;
        const long long po = (long long)pk * a.n + (long long)blk * cnt;
        VS *gx = reinterpret_cast<VS *>(a.x) + po;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int i = threadIdx.x + k * NT;
            x[lx<true>(i)] = widen<S>(st[k]);
            if (PHT) {
                ph[i] = {wc[k], ws[k]};
            } else {
                double sn, cs;
                sincos(a.cst2 * (double)vv[k], &sn, &cs);
                ph[i] = {cs, sn};
            }
            if (INV) p0[i] = widen<S>(q[k]);
        }
        if (FWD && pk != kpk) {  // kinetic phases of a new packet
            stage_table<NT, EPT>(kl, a.kt + (long long)pk * a.ktps, a.np0);
            kpk = pk;
        }
        __syncthreads();
        const int wn = w + nact;
        if (wn < total) rows_fetch<S, NT, EPT, PHT, INV>(a, wn, st, vv, wc, ws, q);

        if (INV)
            for (int s = a.ax.nst - 1; s >= 1; --s) dit_stage<double, NT, -1, true>(x, a.ax, s, twl);
        double cr = 0, ci = 0, nr = 0;
        fused_outer<double, NT, INV, FWD, true>(x, a.ax, twl, [&](int i, V &v) {
            const V wph = ph[i];
            if (INV) {
                v = cmulc(v, wph);
                v.x *= b;
                v.y *= b;
                if (a.mask) {  // .SMOOTHW (spofft.f:171-174)
                    const double mk = a.mask[po + i];
                    v.x *= mk;
                    v.y *= mk;
                }
                if (samp) {
                    const V p = p0[i];
                    cr += p.x * v.x + p.y * v.y;
                    ci += p.x * v.y - p.y * v.x;
                    nr += v.x * v.x + v.y * v.y;
                }
            }
            if (FWD) v = cmulc(v, wph);
        });
        if (FWD) {
            for (int s = 1; s < a.ax.nst; ++s) dif_stage<double, NT, 1, true>(x, a.ax, s, twl);
            for (int i = threadIdx.x; i < cnt; i += NT) {
                const int li = lx<true>(i);
                x[li] = cmulc(x[li], kl[i & (a.np0 - 1)]);
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < cnt; i += NT) {
            const V t = x[lx<true>(i)];
            gx[i] = {(S)t.x, (S)t.y};
        }
        if (samp) {
            block_sum3<NT>(cr, ci, nr, red);
            const long long nc = g / a.lc;
            if (threadIdx.x == 0 && nc < a.maxsamp) {
                double *c = a.part + (((long long)nc * a.npk + pk) * a.xblocks + blk) * 3;
                c[0] = cr;
                c[1] = ci;
                c[2] = nr;
            }
        }
        if (wn >= total) break;
        w = wn;
        __syncthreads();  // LDS is rewritten by the next item
    }
}

// Forward FFT, kinetic phase and inverse FFT along an axis of stride S and
// length L, on tiles of 2^ltile contiguous elements (grid-stride over the
// tiles of all packets, next tile prefetched as in pass_rows).
template <typename S, int NT, int EPT>
__device__ __forceinline__ long long strided_base(const PassArgs &a, int w, int ntl, int *pk_out)
{
    const int pk = w / ntl;
    int t = w - pk * ntl;
    // Neighbouring tiles share cache lines (a tile is 2^ltile contiguous
    // elements); dealt round robin they land on different XCDs, whose L2s
    // then each fetch and write back the same lines.  With xmap, XCD
    // (t mod 8) takes the contiguous tile range [x n/8, (x+1) n/8).
    if (a.xmap) t = (t & 7) * (ntl >> 3) + (t >> 3);
    const int spt = a.S >> a.ltile;  // tiles across the lower axes
    *pk_out = pk;
    return (long long)pk * a.n + (long long)(t / spt) * a.S * a.L +
           (long long)(t % spt) * (1 << a.ltile);
}

template <typename S, int NT, int EPT>
__global__ __launch_bounds__(NT) void pass_strided(PassArgs a)
{
    using V = double2;
    using VS = V2<S>;
    extern __shared__ double smem_d[];
    const int cnt = a.cnt;
    V *x = reinterpret_cast<V *>(smem_d);
    V *twl = x + lds_slots(cnt, true);  // twiddles (+1 dummy slot)
    V *kl = twl + a.ntw + 1;  // kinetic phases (+1 dummy slot)
    const int tile = 1 << a.ltile;
    if (a.spread > 1 && blockIdx.x % a.spread) return;
    const int ntl = a.n / (tile * a.L);  // tiles per packet
    const int nact = gridDim.x / a.spread, total = ntl * a.npk;
    int w = blockIdx.x / a.spread;
    if (w >= total) return;
    stage_table<NT, EPT>(twl, a.tw, a.ntw);
    VS st[EPT];
    int pk;
    long long base = strided_base<S, NT, EPT>(a, w, ntl, &pk);
    auto fetch = [&](long long bs) {
        const VS *gx = reinterpret_cast<const VS *>(a.x) + bs;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int i = threadIdx.x + k * NT;
            st[k] = gx[(long long)(i >> a.ltile) * a.S + (i & (tile - 1))];
        }
    };
    fetch(base);
    int kpk = -1;
    for (;;) {
        VS *gx = reinterpret_cast<VS *>(a.x) + base;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int j = threadIdx.x + k * NT;
            x[lx<true>(j)] = widen<S>(st[k]);
        }
        if (pk != kpk) {  // kinetic phases of a new packet
            stage_table<NT, EPT>(kl, a.kt + (long long)pk * a.ktps, a.L);
            kpk = pk;
        }
        __syncthreads();
        const int wn = w + nact;
        int pkn = pk;
        long long basen = base;
        if (wn < total) {
            basen = strided_base<S, NT, EPT>(a, wn, ntl, &pkn);
            fetch(basen);
        }
        for (int s = 0; s < a.ax.nst - 1; ++s) dif_stage<double, NT, 1, true>(x, a.ax, s, twl);
        fused_inner<double, NT, true>(x, a.ax, [&](int i, V &v) { v = cmulc(v, kl[i >> a.ltile]); });
        for (int s = a.ax.nst - 2; s >= 0; --s) dit_stage<double, NT, -1, true>(x, a.ax, s, twl);
        for (int i = threadIdx.x; i < cnt; i += NT) {
            const V t = x[lx<true>(i)];
            gx[(long long)(i >> a.ltile) * a.S + (i & (tile - 1))] = {(S)t.x, (S)t.y};
        }
        if (a.last && w == 0 && threadIdx.x == 0) *a.dstep += 1;
        if (wn >= total) break;
        w = wn;
        pk = pkn;
        base = basen;
        __syncthreads();
    }
}

// ------------------------------------------------------------ pencils
// The same two passes on register pencils (pencil.hpp) for axis lengths
// L = 16 M, M in {4, 8, 16}.  A workgroup is one wavefront holding C = 64/M
// lines; the state between passes keeps x in the spectral layout of the
// pencil (a permutation of k that makes every register a coalesced row).
// Work items are taken in a grid-stride loop; with single-precision
// storage the next item's state and V are loaded while the current one is
// transformed.

// Loads of one item: 16 elements per lane at po + r * stride.
template <typename S, bool V>
__device__ __forceinline__ void pen_fetch(const PassArgs &a, long long po, long long stride,
                                          S (&sr)[16], S (&si)[16], S (&sv)[16])
{
    const V2<S> *gx = reinterpret_cast<const V2<S> *>(a.x) + po;
    const S *gv = reinterpret_cast<const S *>(a.vpot) + po;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const V2<S> t = gx[r * stride];
        sr[r] = t.x;
        si[r] = t.y;
        if (V) sv[r] = gv[r * stride];
    }
}

// Two waves per SIMD are requested: the exchange slots (17 KiB per wave)
// cap residency near that anyway.  (Prefetching the next item here costs
// spills: the row pass holds psi0 and V beside the state.)  The row pass
// computes in FP64 for either storage: its FP64 sincos of the potential
// phase dominates the arithmetic, and in single precision the phase would
// have to stay split to FP64 accuracy anyway, since a fixed per-point
// rounding of it accumulates coherently over the steps.
template <typename S, int M, bool INV, bool FWD, bool PHT>
__global__ __launch_bounds__(64, 2) void pass_rows_pen(PassArgs a)
{
    constexpr int C = 64 / M, L = 16 * M;
    __shared__ double2 xs[C * pencil::xslots<M>()];
    const int j = threadIdx.x % M, l = threadIdx.x / M;
    double2 *xl = xs + l * pencil::xslots<M>();
    const int total = a.xblocks * a.npk;
    int w = blockIdx.x;
    if (w >= total) return;
    const long long g = INV ? *a.dstep : 0;
    const bool samp = INV && (g % a.lc) == 0;
    const double b = a.b;
    auto origin = [&](int wi, int &pk, int &blk) {
        pk = wi / a.xblocks;
        blk = wi - pk * a.xblocks;
        return (long long)pk * a.n + (long long)(blk * C + l) * L + j;
    };
    S sr[16], si[16], sv[16];
    int pk, blk;
    long long po = origin(w, pk, blk);
    for (;;) {
        // Single precision: V and psi0 are loaded with the state and are in
        // flight during the inverse transform.  Double precision: after it,
        // since holding them across the transform spills.
        constexpr bool early = sizeof(S) == 4;
        pen_fetch<S, !PHT && early>(a, po, M, sr, si, sv);
        double2 v[16];
        S cv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            v[r] = {(double)sr[r], (double)si[r]};
            cv[r] = sv[r];
        }
        const int wn = w + gridDim.x;
        int pkn = pk, blkn = blk;
        const long long pon = wn < total ? origin(wn, pkn, blkn) : po;
        S q0r[16], q0i[16];
        auto load_psi0 = [&]() {
            const V2<S> *g0 = reinterpret_cast<const V2<S> *>(a.psi0) + po;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const V2<S> t = g0[r * M];
                q0r[r] = t.x;
                q0i[r] = t.y;
            }
        };
        if (samp && early) load_psi0();
        if (INV) pencil::inverse<M, -1>(v, j, xl, a.twr);
        if (samp && !early) load_psi0();
        if (!PHT && !early) {
            const S *gv = reinterpret_cast<const S *>(a.vpot) + po;
#pragma unroll
            for (int r = 0; r < 16; ++r) cv[r] = gv[r * M];
        }
        double cr = 0, ci = 0, nr = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            double2 ph;
            if (PHT) {
                ph = a.vph[po + r * M];
            } else {
                double sn, cs;
                pencil::sincos_reduced(a.cst2 * (double)cv[r], &sn, &cs);
                ph = {cs, sn};
            }
            if (INV) {
                v[r] = cmulc(v[r], ph);
                v[r].x *= b;
                v[r].y *= b;
                if (samp) {
                    const double px = q0r[r], py = q0i[r];
                    cr += px * v[r].x + py * v[r].y;
                    ci += px * v[r].y - py * v[r].x;
                    nr += v[r].x * v[r].x + v[r].y * v[r].y;
                }
            }
            if (FWD) v[r] = cmulc(v[r], ph);
        }
        if (FWD) {
            pencil::forward<M, 1>(v, j, xl, a.twr);
            const double2 *kt = a.kreg + (long long)pk * a.ktps + j;
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = cmulc(v[r], kt[r * M]);
        }
        V2<S> *gx = reinterpret_cast<V2<S> *>(a.x) + po;
#pragma unroll
        for (int r = 0; r < 16; ++r) gx[r * M] = {(S)v[r].x, (S)v[r].y};
        if (samp) {
            cr = wave_sum(cr);
            ci = wave_sum(ci);
            nr = wave_sum(nr);
            const long long nc = g / a.lc;
            if (threadIdx.x == 0 && nc < a.maxsamp) {
                double *c = a.part + (((long long)nc * a.npk + pk) * a.xblocks + blk) * 3;
                c[0] = cr;
                c[1] = ci;
                c[2] = nr;
            }
        }
        if (wn >= total) break;
        w = wn;
        pk = pkn;
        blk = blkn;
        po = pon;
        __syncthreads();  // exchange slots are rewritten by the next item
    }
}

// Row pass on pencils for complex64 storage with single-precision
// transforms: the inverse and forward line FFTs and the kinetic phase run
// in FP32 with split (hi + lo) constants, as the strided pass does, while
// the V phase, the 1/N scale and the C(t) and norm sums stay in FP64 (a
// fixed per-point rounding of the phase would accumulate coherently over
// the steps).  Half the registers and exchange slots of pass_rows_pen, so
// more wavefronts hide the memory latency.
template <int M, bool INV, bool FWD, bool PHT>
__global__ __launch_bounds__(64, 2) void pass_rows_pen_sp(PassArgs a)
{
    constexpr int C = 64 / M, L = 16 * M;
    __shared__ float2 xs[C * pencil::xslots<M>()];
    const int j = threadIdx.x % M, l = threadIdx.x / M;
    float2 *xl = xs + l * pencil::xslots<M>();
    const int total = a.xblocks * a.npk;
    const long long g = INV ? *a.dstep : 0;
    const bool samp = INV && (g % a.lc) == 0;
    const double b = a.b;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        const int pk = w / a.xblocks, blk = w - pk * a.xblocks;
        const long long po = (long long)pk * a.n + (long long)(blk * C + l) * L + j;
        const float2 *gx = reinterpret_cast<const float2 *>(a.x) + po;
        const float *gv = reinterpret_cast<const float *>(a.vpot) + po;
        const float2 *g0 = reinterpret_cast<const float2 *>(a.psi0) + po;
        float2 v[16], q0[16];
        float cv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = gx[r * M];
        if (INV) pencil::inverse<M, -1>(v, j, xl, a.twr_s);
        // V and psi0 after the transform: fewer registers live across it
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            if (!PHT) cv[r] = gv[r * M];
            if (samp) q0[r] = g0[r * M];
        }
        double cr = 0, ci = 0, nr = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            double2 ph;
            if (PHT) {
                ph = a.vph[po + r * M];
            } else {
                double sn, cs;
                pencil::sincos_reduced(a.cst2 * (double)cv[r], &sn, &cs);
                ph = {cs, sn};
            }
            double2 x = {(double)v[r].x, (double)v[r].y};
            if (INV) {
                x = cmulc(x, ph);
                x.x *= b;
                x.y *= b;
                if (samp) {
                    const double px = q0[r].x, py = q0[r].y;
                    cr += px * x.x + py * x.y;
                    ci += px * x.y - py * x.x;
                    nr += x.x * x.x + x.y * x.y;
                }
            }
            if (FWD) x = cmulc(x, ph);
            v[r] = {(float)x.x, (float)x.y};
            // one point's sincos at a time: interleaved, the sixteen
            // evaluations need more registers than the occupancy allows
            __builtin_amdgcn_sched_barrier(0);
        }
        // keep the forward transform's twiddle and kinetic loads below the
        // phase loop: hoisted above it they hold ~120 registers across it
        asm volatile("" ::: "memory");
        if (FWD) {
            pencil::forward<M, 1>(v, j, xl, a.twr_s);
            const float4 *kt = a.kreg_s + (long long)pk * a.ktps + j;
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = pencil::cmulc_w(v[r], kt[r * M]);
        }
        float2 *gy = reinterpret_cast<float2 *>(a.x) + po;
#pragma unroll
        for (int r = 0; r < 16; ++r) gy[r * M] = v[r];
        if (samp) {
            cr = wave_sum(cr);
            ci = wave_sum(ci);
            nr = wave_sum(nr);
            const long long nc = g / a.lc;
            if (threadIdx.x == 0 && nc < a.maxsamp) {
                double *c = a.part + (((long long)nc * a.npk + pk) * a.xblocks + blk) * 3;
                c[0] = cr;
                c[1] = ci;
                c[2] = nr;
            }
        }
        __syncthreads();  // exchange slots are rewritten by the next item
    }
}

// Strided pass on pencils: lane = c + C j, line c of the item's C
// consecutive columns, lane j within the line; arithmetic in precision S.
// PF: load the next item's state while transforming the current one.
template <typename S, int M, bool PF>
__global__ __launch_bounds__(64, 2) void pass_strided_pen(PassArgs a)
{
    using VS = V2<S>;
    constexpr int C = 64 / M;
    __shared__ VS xs[C * pencil::xslots<M>()];
    const int c = threadIdx.x % C, j = threadIdx.x / C;
    VS *xl = xs + c * pencil::xslots<M>();
    const PenW<S> *twr = pen_twr<S>(a);
    const int groups = a.n / (a.L * C);  // column groups per packet
    const int total = groups * a.npk;
    int w = blockIdx.x;
    if (w >= total) return;
    const long long stride = (long long)M * a.S;
    auto origin = [&](int wi, int &pk) {
        // With xmap, XCD (wi mod 8) takes a contiguous range of column
        // groups, so that the groups sharing a cache line meet in one L2.
        if (a.xmap) wi = (wi & 7) * (total >> 3) + (wi >> 3);
        pk = wi / groups;
        const int col = (wi - pk * groups) * C + c;
        return (long long)pk * a.n + (col % a.S) + (long long)(col / a.S) * a.S * a.L +
               (long long)j * a.S;
    };
    S sr[16], si[16], sv[16];
    int pk;
    long long po = origin(w, pk);
    if (PF) pen_fetch<S, false>(a, po, stride, sr, si, sv);
    for (;;) {
        if (!PF) pen_fetch<S, false>(a, po, stride, sr, si, sv);
        VS v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = {sr[r], si[r]};
        const int wn = w + gridDim.x;
        int pkn = pk;
        const long long pon = wn < total ? origin(wn, pkn) : po;
        if (PF && wn < total) pen_fetch<S, false>(a, pon, stride, sr, si, sv);
        pencil::forward<M, 1>(v, j, xl, twr);
        const PenW<S> *kt = pen_kreg<S>(a) + (long long)pk * a.ktps + j;
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = pencil::cmulc_w(v[r], kt[r * M]);
        pencil::inverse<M, -1>(v, j, xl, twr);
        VS *gx = reinterpret_cast<VS *>(a.x) + po;
#pragma unroll
        for (int r = 0; r < 16; ++r) gx[r * stride] = v[r];
        if (a.last && w == 0 && threadIdx.x == 0) *a.dstep += 1;
        if (wn >= total) break;
        w = wn;
        pk = pkn;
        po = pon;
        __syncthreads();
    }
}

// Strided pass on pencils through an LDS tile: a workgroup of four
// wavefronts takes W = 4 C = 256 / M consecutive columns (one 128-byte row
// segment in single precision), so every global load and store covers
// whole cache lines; each wavefront transforms C of the columns as
// register pencils (lane c + C j, register r = row M r + j) in precision
// S.  The staging tile and the exchange slots share one LDS array: each is
// idle while the other is in use.  One item per workgroup; requires
// S % W == 0.
template <typename S, int M>
__global__ __launch_bounds__(256, 2) void pass_strided_tile(PassArgs a)
{
    using VS = V2<S>;
    // Row pitch of the tile: rows shifted by C elements, so that the 16
    // lanes of a quarter-wavefront (C columns by 16 / C rows) read distinct
    // banks when they take their pencils.
    constexpr int C = 64 / M, L = 16 * M, W = 4 * C, P = C < 16 ? W + C : W;
    constexpr int XS = C * pencil::xslots<M>();  // exchange slots of one wavefront
    constexpr int NS = L * P > 4 * XS ? L * P : 4 * XS;
    __shared__ VS xs[NS];
    VS *tile = xs;
    const int wv = threadIdx.x / 64, c = threadIdx.x % C, j = (threadIdx.x % 64) / C;
    VS *xl = xs + wv * XS + c * pencil::xslots<M>();
    const int tiles = a.S / W;                     // column tiles per slab
    const int per_pk = a.n / (L * a.S) * tiles;    // items per packet
    const int total = per_pk * a.npk;
    int w = blockIdx.x;
    if (w >= total) return;
    if (a.xmap) w = (w & 7) * (total >> 3) + (w >> 3);
    const int pk = w / per_pk, rem = w - pk * per_pk;
    const int slab = rem / tiles, ct = rem - slab * tiles;
    VS *gx = reinterpret_cast<VS *>(a.x) + (long long)pk * a.n + (long long)slab * a.S * L +
             (long long)ct * W;
    const int tc = threadIdx.x % W, tr = threadIdx.x / W;  // M rows per sweep of 256 threads
    {
        VS t[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) t[i] = gx[(long long)(tr + M * i) * a.S + tc];
#pragma unroll
        for (int i = 0; i < 16; ++i) tile[(tr + M * i) * P + tc] = t[i];
    }
    __syncthreads();
    VS v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = tile[(M * r + j) * P + C * wv + c];
    __syncthreads();  // the exchange slots overlay the tile
    const PenW<S> *twr = pen_twr<S>(a);
    pencil::forward<M, 1>(v, j, xl, twr);
    const PenW<S> *kt = pen_kreg<S>(a) + (long long)pk * a.ktps + j;
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = pencil::cmulc_w(v[r], kt[r * M]);
    pencil::inverse<M, -1>(v, j, xl, twr);
    __syncthreads();  // every wavefront is done with its exchange slots
#pragma unroll
    for (int r = 0; r < 16; ++r) tile[(M * r + j) * P + C * wv + c] = v[r];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) gx[(long long)(tr + M * i) * a.S + tc] = tile[(tr + M * i) * P + tc];
    if (a.last && blockIdx.x == 0 && threadIdx.x == 0) *a.dstep += 1;
}

// Row pass on wide lines (x length L in {64, 128, 256}): 256 / L lines per
// 64-thread workgroup, lane j + (L/4) c holding x[j + (L/4) r] of line c;
// FP64 arithmetic for either storage; the state between passes keeps x in
// natural frequency order.  Same sequence as pass_rows_pen.
template <typename S, int L, bool INV, bool FWD, bool PHT>
__global__ __launch_bounds__(64) void pass_rows_wide(PassArgs a)
{
    constexpr int LW = L / 4, C = 64 / LW;
    __shared__ double2 xs[C * wide::slots<L>()];
    const int j = threadIdx.x % LW, c = threadIdx.x / LW;
    double2 *xl = xs + c * wide::slots<L>();
    const int blocks = a.n / (L * C);  // workgroups per packet
    const int w = blockIdx.x;
    if (w >= blocks * a.npk) return;
    const int pk = w / blocks, blk = w - pk * blocks;
    const long long po = (long long)pk * a.n + (long long)(blk * C + c) * L + j;
    const long long g = INV ? *a.dstep : 0;
    const bool samp = INV && (g % a.lc) == 0;
    const V2<S> *gx = reinterpret_cast<const V2<S> *>(a.x) + po;
    double2 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = widen<S>(gx[LW * r]);
    if (INV) wide::stockham<L, -1>(v, j, xl, a.twr);
    double cr = 0, ci = 0, nr = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        double2 ph;
        if (PHT) {
            ph = a.vph[po + LW * r];
        } else {
            double sn, cs;
            pencil::sincos_reduced(a.cst2 * (double)reinterpret_cast<const S *>(a.vpot)[po + LW * r],
                                   &sn, &cs);
            ph = {cs, sn};
        }
        if (INV) {
            v[r] = cmulc(v[r], ph);
            v[r].x *= a.b;
            v[r].y *= a.b;
            if (samp) {
                const double2 q = widen<S>(reinterpret_cast<const V2<S> *>(a.psi0)[po + LW * r]);
                cr += q.x * v[r].x + q.y * v[r].y;
                ci += q.x * v[r].y - q.y * v[r].x;
                nr += v[r].x * v[r].x + v[r].y * v[r].y;
            }
        }
        if (FWD) v[r] = cmulc(v[r], ph);
    }
    if (FWD) {
        wide::stockham<L, 1>(v, j, xl, a.twr);
        const double2 *kt = a.knat + (long long)pk * a.ktps + j;
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = cmulc(v[r], kt[LW * r]);
    }
    V2<S> *gy = reinterpret_cast<V2<S> *>(a.x) + po;
#pragma unroll
    for (int r = 0; r < 4; ++r) gy[LW * r] = {(S)v[r].x, (S)v[r].y};
    if (samp) {
        cr = wave_sum(cr);
        ci = wave_sum(ci);
        nr = wave_sum(nr);
        const long long nc = g / a.lc;
        if (threadIdx.x == 0 && nc < a.maxsamp) {
            double *cp = a.part + (((long long)nc * a.npk + pk) * a.xblocks + blk) * 3;
            cp[0] = cr;
            cp[1] = ci;
            cp[2] = nr;
        }
    }
}

// Strided pass on wide lines (axis length L): 256 / L consecutive columns
// per 64-thread workgroup, lane c + C j holding rows j + (L/4) r of column
// c; FP64 arithmetic.
template <typename S, int L>
__global__ __launch_bounds__(64) void pass_strided_wide(PassArgs a)
{
    constexpr int LW = L / 4, C = 64 / LW;
    __shared__ double2 xs[C * wide::slots<L>()];
    const int c = threadIdx.x % C, j = threadIdx.x / C;
    double2 *xl = xs + c * wide::slots<L>();
    const int groups = a.n / (L * C);  // column groups per packet
    const int total = groups * a.npk;
    int w = blockIdx.x;
    if (w >= total) return;
    if (a.xmap) w = (w & 7) * (total >> 3) + (w >> 3);
    const int pk = w / groups, col = (w - pk * groups) * C + c;
    const long long po = (long long)pk * a.n + (col % a.S) + (long long)(col / a.S) * a.S * L +
                         (long long)j * a.S;
    const long long st = (long long)LW * a.S;
    V2<S> *gx = reinterpret_cast<V2<S> *>(a.x) + po;
    double2 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = widen<S>(gx[r * st]);
    wide::stockham<L, 1>(v, j, xl, a.twr);
    const double2 *kt = a.knat + (long long)pk * a.ktps + j;
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = cmulc(v[r], kt[LW * r]);
    wide::stockham<L, -1>(v, j, xl, a.twr);
#pragma unroll
    for (int r = 0; r < 4; ++r) gx[r * st] = {(S)v[r].x, (S)v[r].y};
    if (a.last && blockIdx.x == 0 && threadIdx.x == 0) *a.dstep += 1;
}

// ------------------------------------------------ LDS-resident pencils
// Two-dimensional grids of L0 x L1 points (L = 16 M, axis 0 contiguous)
// whose state fits one workgroup's LDS in storage precision: one
// workgroup per packet keeps the grid in LDS for every step of a launch,
// so per step HBM carries only the FP64 V phase table and psi0, 24 B per
// point in complex64 instead of the 44 B of a row pass and a strided pass
// (the table, not V and a sincos: the kernel is bound by its arithmetic).
// Each step is the arithmetic of those two passes on the LDS copy: the row
// pass of pass_rows_pen_sp (complex64; pass_rows_pen for complex128) on
// every row, then the strided pass of pass_strided_pen on every column,
// with every pencil exchange done in place in the line's own slots.  The
// state is therefore bit for bit the multi-pass plan's with those kernels,
// and C(t) is too: the partial sums are formed per row block of C0 rows
// and added in block order, as reduce_parts adds them.
// LDS layout: row y at y RP, element x at x + x / 16 (xat's padding); the
// pitch RP = 17 L0 / 16 + 1 puts the rows that one column exchange reads
// together in different banks.
struct FusedArgs {
    int n, npk, nsteps, lc, maxsamp, ktps;
    long long step0;
    void *x;
    const void *psi0;
    const double2 *vph;  // exp(-i cst2 V) as (cos, sin), FP64
    double *cbuf;
    double b;
    const double2 *twr[2], *kreg[2];    // FP64 pencil tables of axes 0 and 1
    const float4 *twr_s[2], *kreg_s[2];  // their split forms (complex64 storage)
};

template <typename S> __device__ __forceinline__ const PenW<S> *fz_twr(const FusedArgs &a, int ax);
template <> __device__ __forceinline__ const double2 *fz_twr<double>(const FusedArgs &a, int ax) { return a.twr[ax]; }
template <> __device__ __forceinline__ const float4 *fz_twr<float>(const FusedArgs &a, int ax) { return a.twr_s[ax]; }
template <typename S> __device__ __forceinline__ const PenW<S> *fz_kreg(const FusedArgs &a, int ax);
template <> __device__ __forceinline__ const double2 *fz_kreg<double>(const FusedArgs &a, int ax) { return a.kreg[ax]; }
template <> __device__ __forceinline__ const float4 *fz_kreg<float>(const FusedArgs &a, int ax) { return a.kreg_s[ax]; }

constexpr int fz_pitch(int l0) { return l0 + l0 / 16 + 1; }

// NW wavefronts per workgroup; row blocks and column groups are dealt to
// them round robin.  The whole launch advances nsteps steps from step0
// (the state in HBM is in natural order before and after).
template <typename S, int M0, int M1, int NW>
__global__ __launch_bounds__(64 * NW, 2) void pass_fused2d(FusedArgs a)
{
    using VS = V2<S>;
    constexpr int L0 = 16 * M0, L1 = 16 * M1, C0 = 64 / M0, C1 = 64 / M1;
    constexpr int RP = fz_pitch(L0), NB = L1 / C0, NT = 64 * NW;
    extern __shared__ double smem_d[];
    VS *xs = reinterpret_cast<VS *>(smem_d);
    double *red = reinterpret_cast<double *>(xs + L1 * RP);  // [NB][3] row-block sums
    const int pk = blockIdx.x, wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    const long long po = (long long)pk * (L0 * L1);
    VS *gx = reinterpret_cast<VS *>(a.x) + po;
    for (int i = threadIdx.x; i < L0 * L1; i += NT) {
        const int y = i / L0, x = i % L0;
        xs[y * RP + x + (x >> 4)] = gx[i];
    }
    __syncthreads();
    const PenW<S> *tw0 = fz_twr<S>(a, 0), *tw1 = fz_twr<S>(a, 1);
    const PenW<S> *kt0 = fz_kreg<S>(a, 0) + (long long)pk * a.ktps;
    const PenW<S> *kt1 = fz_kreg<S>(a, 1) + (long long)pk * a.ktps;
    const double b = a.b;
    for (int st = 0;; ++st) {
        const bool inv = st > 0, fwd = st < a.nsteps;
        const long long g = a.step0 + st;
        const bool samp = inv && (g % a.lc) == 0;
        // Row phase: [inverse x of step g]  B e^{-iV dt/2}  (C(t) sample)
        // e^{-iV dt/2}  [forward x, e^{-iTx dt} of step g + 1].
        {
            const int j = lane % M0, l = lane / M0;
            for (int blk = wv; blk < NB; blk += NW) {
                const int y = blk * C0 + l;
                VS *xl = xs + y * RP;
                const long long pr = po + (long long)y * L0 + j;
                VS v[16], q0[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) v[r] = pencil::xat(xl, r * M0 + j);
                pencil::wave_sync();  // every lane's reads before the exchange writes
                if (inv) pencil::inverse<M0, -1>(v, j, xl, tw0);
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    if (samp) q0[r] = reinterpret_cast<const VS *>(a.psi0)[pr + r * M0];
                double cr = 0, ci = 0, nr = 0;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const double2 ph = a.vph[pr + r * M0];
                    double2 x = {(double)v[r].x, (double)v[r].y};
                    if (inv) {
                        x = cmulc(x, ph);
                        x.x *= b;
                        x.y *= b;
                        if (samp) {
                            const double px = q0[r].x, py = q0[r].y;
                            cr += px * x.x + py * x.y;
                            ci += px * x.y - py * x.x;
                            nr += x.x * x.x + x.y * x.y;
                        }
                    }
                    if (fwd) x = cmulc(x, ph);
                    v[r] = {(S)x.x, (S)x.y};
                    __builtin_amdgcn_sched_barrier(0);
                }
                asm volatile("" ::: "memory");
                if (fwd) {
                    pencil::forward<M0, 1>(v, j, xl, tw0);
#pragma unroll
                    for (int r = 0; r < 16; ++r) v[r] = pencil::cmulc_w(v[r], kt0[j + r * M0]);
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) pencil::xat(xl, r * M0 + j) = v[r];
                if (samp) {
                    cr = wave_sum(cr);
                    ci = wave_sum(ci);
                    nr = wave_sum(nr);
                    if (lane == 0) {
                        red[3 * blk] = cr;
                        red[3 * blk + 1] = ci;
                        red[3 * blk + 2] = nr;
                    }
                }
            }
        }
        __syncthreads();
        if (samp && threadIdx.x == 0) {
            double s0 = 0, s1 = 0, s2 = 0;
            for (int k = 0; k < NB; ++k) {
                s0 += red[3 * k];
                s1 += red[3 * k + 1];
                s2 += red[3 * k + 2];
            }
            const long long nc = g / a.lc;
            if (nc < a.maxsamp) {
                double *c = a.cbuf + ((long long)pk * a.maxsamp + nc) * 3;
                c[0] = s0;
                c[1] = s1;
                c[2] = s2;
            }
        }
        if (!fwd) break;
        // Column phase: forward y, e^{-iTy dt}, inverse y.
        {
            const int c = lane % C1, j = lane / C1;
            for (int cb = wv; cb < L0 / C1; cb += NW) {
                const int x = cb * C1 + c;
                const pencil::Strided<VS> col{xs + x + (x >> 4), RP};
                VS v[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) v[r] = pencil::xat(col, r * M1 + j);
                pencil::forward<M1, 1>(v, j, col, tw1);
#pragma unroll
                for (int r = 0; r < 16; ++r) v[r] = pencil::cmulc_w(v[r], kt1[j + r * M1]);
                pencil::wave_sync();  // the forward exchange's reads before the inverse's writes
                pencil::inverse<M1, -1>(v, j, col, tw1);
#pragma unroll
                for (int r = 0; r < 16; ++r) pencil::xat(col, r * M1 + j) = v[r];
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < L0 * L1; i += NT) {
        const int y = i / L0, x = i % L0;
        gx[i] = xs[y * RP + x + (x >> 4)];
    }
}

__global__ void set_step(long long *d, long long v) { *d = v; }
// One-dimensional grids have no strided pass to advance the step counter.
__global__ void add_step(long long *d) { *d += 1; }

// cbuf[pk][nc] = sum over row blocks of part[nc][pk][blk], nc in [lo, hi).
__global__ void reduce_parts(const double *part, double *cbuf, int npk, int xblocks, int maxsamp,
                             int lo, int hi)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ns = hi - lo;
    if (idx >= ns * npk) return;
    const int pk = idx / ns, nc = lo + idx % ns;
    double s0 = 0, s1 = 0, s2 = 0;
    const double *p = part + ((long long)nc * npk + pk) * xblocks * 3;
    for (int k = 0; k < xblocks; ++k) {
        s0 += p[3 * k];
        s1 += p[3 * k + 1];
        s2 += p[3 * k + 2];
    }
    double *c = cbuf + ((long long)pk * maxsamp + nc) * 3;
    c[0] = s0;
    c[1] = s1;
    c[2] = s2;
}

namespace {

void axis_from_plan(DevAxis &d, const FftPlan &p, int es, int nl, int lmod, int la, int lb)
{
    d.n = p.n;
    d.ln = ilog2(p.n);
    d.nst = p.nst;
    set_line_geometry(d, es, nl, lmod, la, lb);
    d.tw0 = 0;
    for (int s = 0; s < p.nst; ++s) {
        d.rad[s] = p.rad[s];
        d.lm[s] = p.lm[s];
        d.lL[s] = p.lL[s];
        d.off[s] = p.off[s];
    }
}

// Geometry for the current shape.
void apply_shape(SpoPassPlan &p)
{
    const PassShape &sh = p.shape;
    axis_from_plan(p.axx, p.fpl[0], 1, sh.rows, sh.rows, p.np[0], 0);
    for (int a = 1; a < p.ndim; ++a) axis_from_plan(p.axs[a], p.fpl[a], sh.tile, sh.tile, sh.tile, 1, 0);
    p.xblocks = p.n / (p.np[0] * sh.rows);
}

// When set (per-kernel profiling), the next launches record these events
// at the kernel's own start and end (hipExtLaunchKernelGGL), so that the
// measured duration excludes queue gaps and marker packets.
thread_local hipEvent_t tl_ev_start = nullptr, tl_ev_stop = nullptr;

template <typename K, typename A>
hipError_t launch(K kern, dim3 grid, int nt, size_t lds, hipStream_t s, const A &a)
{
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (tl_ev_start)
        hipExtLaunchKernelGGL(kern, grid, dim3(nt), (std::uint32_t)lds, s, tl_ev_start, tl_ev_stop, 0u,
                              a);
    else
        hipLaunchKernelGGL(kern, grid, dim3(nt), lds, s, a);
    return hipGetLastError();
}

PassArgs base_args(const SpoPassPlan &p)
{
    PassArgs a{};
    a.n = p.n;
    a.npk = p.npk;
    a.np0 = p.np[0];
    a.xblocks = p.xblocks;
    a.x = p.x;
    a.psi0 = p.psi0;
    a.vpot = p.vpot;
    a.vph = static_cast<const double2 *>(p.vph);
    a.mask = p.mask;
    a.cst2 = p.cst2;
    a.part = p.part;
    a.dstep = p.dstep;
    a.b = p.b;
    a.lc = p.lc;
    a.maxsamp = p.maxsamp;
    a.spread = p.shape.spread;
    a.xmap = 0;
    a.ktps = p.kax_pstride;
    return a;
}

size_t rows_lds(const SpoPassPlan &p, const PassShape &sh)
{
    const int cnt = sh.rows * p.np[0];
    return sizeof(double2) * (lds_slots(cnt, true) + 2 * (size_t)cnt + p.ntw[0] + p.np[0] + 2) +
           sizeof(double) * 3 * (sh.rows_nt / 64);
}

size_t strided_lds(const SpoPassPlan &p, int axis, const PassShape &sh)
{
    return sizeof(double2) *
           (lds_slots(sh.tile * p.np[axis], true) + p.ntw[axis] + p.np[axis] + 2);
}

template <typename R>
hipError_t rows(const SpoPassPlan &p, bool inv, bool fwd, hipStream_t s, double b_override = 0)
{
    PassArgs a = base_args(p);
    if (b_override > 0) a.b = b_override;
    a.ax = p.axx;
    a.cnt = p.shape.rows * p.np[0];
    const int nt = p.shape.rows_nt;
    a.ept = a.cnt / nt;
    a.tw = static_cast<const double2 *>(p.tw) + p.tw_off[0];
    a.ntw = p.ntw[0];
    a.kt = static_cast<const double2 *>(p.kax) + p.kax_off[0];
    const bool pht = p.vph != nullptr;
    const int total = p.xblocks * p.npk;
    const int act = p.shape.grid > 0 ? std::min(total, p.shape.grid) : total;
    if (p.shape.rows_pen == 2) {
        a.twr = static_cast<const double2 *>(p.twr) + p.twr_off[0];
        a.knat = static_cast<const double2 *>(p.knat) + p.kax_off[0];
        const dim3 wg(total);
#define ESPEC_WIDE_ROWS2(L_, P_)                                                                \
    if (p.np[0] == L_ && pht == P_) {                                                           \
        if (inv && fwd) return launch(pass_rows_wide<R, L_, true, true, P_>, wg, 64, 0, s, a);  \
        if (inv) return launch(pass_rows_wide<R, L_, true, false, P_>, wg, 64, 0, s, a);        \
        return launch(pass_rows_wide<R, L_, false, true, P_>, wg, 64, 0, s, a);                 \
    }
#define ESPEC_WIDE_ROWS(L_) ESPEC_WIDE_ROWS2(L_, true) ESPEC_WIDE_ROWS2(L_, false)
        ESPEC_WIDE_ROWS(64)
        ESPEC_WIDE_ROWS(128)
        ESPEC_WIDE_ROWS(256)
#undef ESPEC_WIDE_ROWS
#undef ESPEC_WIDE_ROWS2
        return hipErrorInvalidValue;
    }
    if (p.shape.rows_pen == 3) {
        if constexpr (std::is_same<R, float>::value) {
            a.twr_s = static_cast<const float4 *>(p.twr_s) + 2 * p.twr_off[0];
            a.kreg_s = static_cast<const float4 *>(p.kreg_s) + p.kax_off[0];
            const dim3 pg(act);
#define ESPEC_PEN_ROWS_SP2(M_, P_)                                                          \
    if (p.np[0] == 16 * M_ && pht == P_) {                                                  \
        if (inv && fwd) return launch(pass_rows_pen_sp<M_, true, true, P_>, pg, 64, 0, s, a); \
        if (inv) return launch(pass_rows_pen_sp<M_, true, false, P_>, pg, 64, 0, s, a);       \
        return launch(pass_rows_pen_sp<M_, false, true, P_>, pg, 64, 0, s, a);                \
    }
#define ESPEC_PEN_ROWS_SP(M_) ESPEC_PEN_ROWS_SP2(M_, true) ESPEC_PEN_ROWS_SP2(M_, false)
            ESPEC_PEN_ROWS_SP(4)
            ESPEC_PEN_ROWS_SP(8)
            ESPEC_PEN_ROWS_SP(16)
#undef ESPEC_PEN_ROWS_SP
#undef ESPEC_PEN_ROWS_SP2
        }
        return hipErrorInvalidValue;
    }
    if (p.shape.rows_pen) {
        a.twr = static_cast<const double2 *>(p.twr) + p.twr_off[0];
        a.kreg = static_cast<const double2 *>(p.kreg) + p.kax_off[0];
        a.twr_s = static_cast<const float4 *>(p.twr_s) + 2 * p.twr_off[0];
        a.kreg_s = static_cast<const float4 *>(p.kreg_s) + p.kax_off[0];
        const dim3 pg(act);
#define ESPEC_PEN_ROWS2(M_, P_)                                                               \
    if (p.np[0] == 16 * M_ && pht == P_) {                                                    \
        if (inv && fwd) return launch(pass_rows_pen<R, M_, true, true, P_>, pg, 64, 0, s, a); \
        if (inv) return launch(pass_rows_pen<R, M_, true, false, P_>, pg, 64, 0, s, a);       \
        return launch(pass_rows_pen<R, M_, false, true, P_>, pg, 64, 0, s, a);                \
    }
#define ESPEC_PEN_ROWS(M_) ESPEC_PEN_ROWS2(M_, true) ESPEC_PEN_ROWS2(M_, false)
        ESPEC_PEN_ROWS(4)
        ESPEC_PEN_ROWS(8)
        ESPEC_PEN_ROWS(16)
#undef ESPEC_PEN_ROWS
#undef ESPEC_PEN_ROWS2
        return hipErrorInvalidValue;
    }
    const size_t lds = rows_lds(p, p.shape);
    dim3 grid(act * p.shape.spread);
#define ESPEC_ROWS2(NT_, E_, P_)                                                                    \
    if (nt == NT_ && a.ept == E_ && pht == P_) {                                                    \
        if (inv && fwd) return launch(pass_rows<R, NT_, true, true, E_, P_>, grid, NT_, lds, s, a); \
        if (inv) return launch(pass_rows<R, NT_, true, false, E_, P_>, grid, NT_, lds, s, a);       \
        return launch(pass_rows<R, NT_, false, true, E_, P_>, grid, NT_, lds, s, a);                \
    }
#define ESPEC_ROWS(NT_, E_) ESPEC_ROWS2(NT_, E_, true) ESPEC_ROWS2(NT_, E_, false)
    ESPEC_ROWS(64, 1)
    ESPEC_ROWS(64, 2)
    ESPEC_ROWS(64, 4)
    ESPEC_ROWS(64, 8)
    ESPEC_ROWS(64, 16)
    ESPEC_ROWS(256, 1)
    ESPEC_ROWS(256, 2)
    ESPEC_ROWS(256, 4)
    ESPEC_ROWS(256, 8)
    ESPEC_ROWS(256, 16)
#undef ESPEC_ROWS
#undef ESPEC_ROWS2
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t strided(const SpoPassPlan &p, int axis, bool last, hipStream_t s)
{
    PassArgs a = base_args(p);
    a.ax = p.axs[axis];
    a.tw = static_cast<const double2 *>(p.tw) + p.tw_off[axis];
    a.ntw = p.ntw[axis];
    a.kt = static_cast<const double2 *>(p.kax) + p.kax_off[axis];
    a.S = 1;
    for (int k = 0; k < axis; ++k) a.S *= p.np[k];
    a.L = p.np[axis];
    a.ltile = ilog2(p.shape.tile);
    a.cnt = p.shape.tile * a.L;
    const int nt = p.shape.tile_nt;
    a.ept = a.cnt / nt;
    a.last = last ? 1 : 0;
    if (p.shape.tile_pen == 3) {
        a.twr = static_cast<const double2 *>(p.twr) + p.twr_off[axis];
        a.knat = static_cast<const double2 *>(p.knat) + p.kax_off[axis];
        const int total = p.n / 256 * p.npk;  // groups of 256 / L columns
        a.xmap = (p.shape.xmap && total % 8 == 0) ? 1 : 0;
        const dim3 wg(total);
        if (a.L == 64) return launch(pass_strided_wide<R, 64>, wg, 64, 0, s, a);
        if (a.L == 128) return launch(pass_strided_wide<R, 128>, wg, 64, 0, s, a);
        if (a.L == 256) return launch(pass_strided_wide<R, 256>, wg, 64, 0, s, a);
        return hipErrorInvalidValue;
    }
    if (p.shape.tile_pen == 2) {
        a.twr = static_cast<const double2 *>(p.twr) + p.twr_off[axis];
        a.kreg = static_cast<const double2 *>(p.kreg) + p.kax_off[axis];
        a.twr_s = static_cast<const float4 *>(p.twr_s) + 2 * p.twr_off[axis];
        a.kreg_s = static_cast<const float4 *>(p.kreg_s) + p.kax_off[axis];
        const int total = p.n / (a.L * (4096 / a.L)) * p.npk;
        a.xmap = (p.shape.xmap && total % 8 == 0) ? 1 : 0;
        const dim3 tg(total);
        if (a.L == 64) return launch(pass_strided_tile<R, 4>, tg, 256, 0, s, a);
        if (a.L == 128) return launch(pass_strided_tile<R, 8>, tg, 256, 0, s, a);
        if (a.L == 256) return launch(pass_strided_tile<R, 16>, tg, 256, 0, s, a);
        return hipErrorInvalidValue;
    }
    if (p.shape.tile_pen) {
        a.twr = static_cast<const double2 *>(p.twr) + p.twr_off[axis];
        a.kreg = static_cast<const double2 *>(p.kreg) + p.kax_off[axis];
        a.twr_s = static_cast<const float4 *>(p.twr_s) + 2 * p.twr_off[axis];
        a.kreg_s = static_cast<const float4 *>(p.kreg_s) + p.kax_off[axis];
        const int total = p.n / (a.L * (64 / (a.L / 16))) * p.npk;
        const int act = p.shape.grid > 0 ? std::min(total, p.shape.grid) : total;
        a.xmap = (p.shape.xmap && total % 8 == 0) ? 1 : 0;
        const dim3 pg(act);
        const bool pf = p.shape.pen_pf != 0;  // strided passes only
        if (a.L == 64) return launch(pf ? pass_strided_pen<R, 4, true> : pass_strided_pen<R, 4, false>, pg, 64, 0, s, a);
        if (a.L == 128) return launch(pf ? pass_strided_pen<R, 8, true> : pass_strided_pen<R, 8, false>, pg, 64, 0, s, a);
        if (a.L == 256) return launch(pf ? pass_strided_pen<R, 16, true> : pass_strided_pen<R, 16, false>, pg, 64, 0, s, a);
        return hipErrorInvalidValue;
    }
    const int ntiles = p.n / (p.shape.tile * a.L);
    a.xmap = (p.shape.xmap && p.shape.spread == 1 && ntiles % 8 == 0) ? 1 : 0;
    const size_t lds = strided_lds(p, axis, p.shape);
    const int total = ntiles * p.npk;
    const int act = p.shape.grid > 0 ? std::min(total, p.shape.grid) : total;
    dim3 grid(act * p.shape.spread);
#define ESPEC_STRIDED(NT_, E_) \
    if (nt == NT_ && a.ept == E_) return launch(pass_strided<R, NT_, E_>, grid, NT_, lds, s, a);
    ESPEC_STRIDED(64, 1)
    ESPEC_STRIDED(64, 2)
    ESPEC_STRIDED(64, 4)
    ESPEC_STRIDED(64, 8)
    ESPEC_STRIDED(64, 16)
    ESPEC_STRIDED(256, 1)
    ESPEC_STRIDED(256, 2)
    ESPEC_STRIDED(256, 4)
    ESPEC_STRIDED(256, 8)
    ESPEC_STRIDED(256, 16)
#undef ESPEC_STRIDED
    return hipErrorInvalidValue;
}

// LDS bytes of pass_fused2d for this grid: the padded state and the
// row-block sums.
size_t fused_lds(const SpoPassPlan &p)
{
    const size_t elem = p.mixed ? sizeof(float2) : sizeof(double2);
    const int nb = p.np[1] / (64 / (p.np[0] / 16));
    return elem * (size_t)p.np[1] * fz_pitch(p.np[0]) + sizeof(double) * 3 * nb;
}

// Grids with a pass_fused2d instance: 64^2 and 128^2 in complex64, 64^2
// in complex128 (128^2 complex128 needs 272 KiB).
bool fused_grid(const SpoPassPlan &p) { return spo_pass_fused_grid(p.ndim, p.np, p.mixed); }

// The phase table the fused form reads, if any.
const double2 *fused_table(const SpoPassPlan &p)
{
    return static_cast<const double2 *>(p.vph ? p.vph : p.vph_fz);
}

bool fused_ok(const SpoPassPlan &p, const PassShape &s)
{
    if (!s.fused) return true;
    return (s.fused == 4 || s.fused == 8) && fused_grid(p) && !p.mask &&
           fused_table(p) != nullptr && fused_lds(p) <= 160 * 1024;
}

// One launch of pass_fused2d: nsteps steps of every packet from step0.
template <typename R>
hipError_t fused_launch(const SpoPassPlan &p, int nsteps, long long step0, hipStream_t s)
{
    FusedArgs f{};
    f.n = p.n;
    f.npk = p.npk;
    f.nsteps = nsteps;
    f.lc = p.lc;
    f.maxsamp = p.maxsamp;
    f.ktps = p.kax_pstride;
    f.step0 = step0;
    f.x = p.x;
    f.psi0 = p.psi0;
    f.vph = fused_table(p);
    f.cbuf = p.cbuf;
    f.b = p.b;
    for (int a = 0; a < 2; ++a) {
        f.twr[a] = static_cast<const double2 *>(p.twr) + p.twr_off[a];
        f.kreg[a] = static_cast<const double2 *>(p.kreg) + p.kax_off[a];
        f.twr_s[a] = static_cast<const float4 *>(p.twr_s) + 2 * p.twr_off[a];
        f.kreg_s[a] = static_cast<const float4 *>(p.kreg_s) + p.kax_off[a];
    }
    const size_t lds = fused_lds(p);
    const int nw = p.shape.fused;
    const dim3 grid(p.npk);
    if (!f.vph) return hipErrorInvalidValue;
#define ESPEC_FUSED2(M0_, M1_, NW_)                                     \
    if (p.np[0] == 16 * M0_ && p.np[1] == 16 * M1_ && nw == NW_)         \
        return launch(pass_fused2d<R, M0_, M1_, NW_>, grid, 64 * NW_, lds, s, f);
    ESPEC_FUSED2(4, 4, 4)
    ESPEC_FUSED2(4, 4, 8)
    if constexpr (std::is_same<R, float>::value) {
        ESPEC_FUSED2(8, 8, 4)
        ESPEC_FUSED2(8, 8, 8)
    }
#undef ESPEC_FUSED2
    return hipErrorInvalidValue;
}

// A run of the fused shape: launches of about 2^24 point-steps per packet.
template <typename R>
int fused_run(const SpoPassPlan &p, int nsteps, long long step0, hipStream_t s, int *launches)
{
    const int chunk = std::max(1, (1 << 24) / p.n);
    for (int done = 0; done < nsteps;) {
        const int k = std::min(chunk, nsteps - done);
        ESPEC_HIP_TRY(fused_launch<R>(p, k, step0 + done, s));
        ++*launches;
        done += k;
    }
    return ESPEC_OK;
}

template <typename R>
hipError_t one_step(const SpoPassPlan &p, bool inv, bool fwd, hipStream_t s, int *launches)
{
    hipError_t e = rows<R>(p, inv, fwd, s);
    ++*launches;
    if (e != hipSuccess || !fwd) return e;
    if (p.ndim == 1) {
        hipLaunchKernelGGL(add_step, dim3(1), dim3(1), 0, s, p.dstep);
        ++*launches;
        return hipGetLastError();
    }
    for (int ax = 1; ax < p.ndim; ++ax) {
        e = strided<R>(p, ax, ax == p.ndim - 1, s);
        ++*launches;
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

void drop_graph(SpoPassPlan &p)
{
    for (int j = 0; j < SpoPassPlan::kGraphLevels; ++j) {
        if (p.gexec[j]) (void)hipGraphExecDestroy(p.gexec[j]);
        if (p.graph[j]) (void)hipGraphDestroy(p.graph[j]);
        p.gexec[j] = nullptr;
        p.graph[j] = nullptr;
    }
}

// Record 2^j full steps (row pass + strided passes each) into a graph;
// replays advance the device step counter themselves.  Capture records
// without executing, so it costs no steps.  Graphs of 1, 2, 4, ..., 32
// steps let a run of any length replay graphs only (a 20-step run is
// 16 + 2 + 1), so short runs do not fall back to single launches.
template <typename R>
int capture_typed(SpoPassPlan &p, int j)
{
    if (p.gexec[j]) return ESPEC_OK;
    hipStream_t cs;
    ESPEC_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    int dummy = 0;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    for (int k = 0; e == hipSuccess && k < (1 << j); ++k) e = one_step<R>(p, true, true, cs, &dummy);
    hipError_t e2 = hipStreamEndCapture(cs, &p.graph[j]);
    (void)hipStreamDestroy(cs);
    ESPEC_HIP_TRY(e);
    ESPEC_HIP_TRY(e2);
    ESPEC_HIP_TRY(hipGraphInstantiate(&p.gexec[j], p.graph[j], nullptr, nullptr, 0));
    return ESPEC_OK;
}

// The run's kernels launched one by one on s (no step-graph replays): for
// a caller that captures the whole run into a graph of its own, where a
// replay of the plan's step graphs cannot be recorded.
template <typename R>
int enqueue_typed(SpoPassPlan &p, int nsteps, long long step0, hipStream_t s)
{
    int launches = 0;
    if (p.shape.fused) return fused_run<R>(p, nsteps, step0, s, &launches);
    hipLaunchKernelGGL(set_step, dim3(1), dim3(1), 0, s, p.dstep, step0);
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(one_step<R>(p, false, true, s, &launches));
    for (int k = 1; k < nsteps; ++k) ESPEC_HIP_TRY(one_step<R>(p, true, true, s, &launches));
    ESPEC_HIP_TRY(one_step<R>(p, true, false, s, &launches));
    long long lo = step0 / p.lc + 1, hi = (step0 + nsteps) / p.lc + 1;
    lo = std::min<long long>(lo, p.maxsamp);
    hi = std::min<long long>(hi, p.maxsamp);
    if (hi > lo) {
        int cnt = (int)(hi - lo) * p.npk;
        hipLaunchKernelGGL(reduce_parts, dim3((cnt + 255) / 256), dim3(256), 0, s, p.part, p.cbuf,
                           p.npk, p.xblocks, p.maxsamp, (int)lo, (int)hi);
        ESPEC_HIP_TRY(hipGetLastError());
    }
    return ESPEC_OK;
}

template <typename R>
int run_typed(SpoPassPlan &p, int nsteps, long long step0, hipStream_t s, int *launches)
{
    if (p.shape.fused) return fused_run<R>(p, nsteps, step0, s, launches);
    hipLaunchKernelGGL(set_step, dim3(1), dim3(1), 0, s, p.dstep, step0);
    ESPEC_HIP_TRY(hipGetLastError());
    ++*launches;
    // Step 1 of the run: V half phase + forward x, then the other axes.
    ESPEC_HIP_TRY(one_step<R>(p, false, true, s, launches));
    // Steps 2..nsteps: finish the previous step and start the next.  The
    // same launch sequence repeats, so it replays from the step graphs of
    // 32, 16, ..., 1 steps, largest first.
    int left = nsteps - 1;
    while (left > 0) {
        int j = SpoPassPlan::kGraphLevels - 1;
        while ((1 << j) > left) --j;
        const int rc = capture_typed<R>(p, j);
        if (rc) return rc;
        ESPEC_HIP_TRY(hipGraphLaunch(p.gexec[j], s));
        *launches += (1 << j) * (p.ndim == 1 ? 2 : p.ndim);
        left -= 1 << j;
    }
    // Finish the last step: inverse x, V half phase, 1/N, last sample.
    ESPEC_HIP_TRY(one_step<R>(p, true, false, s, launches));
    // Fold the row-block partial sums of this run's samples into cbuf.
    long long lo = step0 / p.lc + 1, hi = (step0 + nsteps) / p.lc + 1;
    lo = std::min<long long>(lo, p.maxsamp);
    hi = std::min<long long>(hi, p.maxsamp);
    if (hi > lo) {
        int cnt = (int)(hi - lo) * p.npk;
        hipLaunchKernelGGL(reduce_parts, dim3((cnt + 255) / 256), dim3(256), 0, s, p.part, p.cbuf,
                           p.npk, p.xblocks, p.maxsamp, (int)lo, (int)hi);
        ESPEC_HIP_TRY(hipGetLastError());
        ++*launches;
    }
    return ESPEC_OK;
}

// ---------------------------------------------------------------- tuner

bool nt_ok(int nt, int cnt)
{
    if (nt != 64 && nt != 256) return false;
    const int e = cnt / nt;
    return cnt % nt == 0 && (e == 1 || e == 2 || e == 4 || e == 8 || e == 16);
}

bool pen_len(int l) { return l == 64 || l == 128 || l == 256; }

bool rows_ok(const SpoPassPlan &p, const PassShape &s)
{
    if (p.mask && s.rows_pen) return false;  // the mask is applied by pass_rows only
    if (s.rows_pen == 2)
        return pen_len(p.np[0]) && s.rows == 256 / p.np[0] && s.rows_nt == 64 &&
               (p.n / p.np[0]) % s.rows == 0 &&
               (p.vph != nullptr || p.phase_max < pencil::kSincosMax);
    if (s.rows_pen > 3 || (s.rows_pen == 3 && !p.mixed)) return false;
    if (s.rows_pen)
        return pen_len(p.np[0]) && s.rows == 1024 / p.np[0] && s.rows_nt == 64 &&
               (p.n / p.np[0]) % s.rows == 0 &&
               (p.vph != nullptr || p.phase_max < pencil::kSincosMax);
    const int cnt = s.rows * p.np[0];
    if (!nt_ok(s.rows_nt, cnt)) return false;
    if (p.ntw[0] > kMaxEpt * s.rows_nt || p.np[0] > kMaxEpt * s.rows_nt) return false;
    return (p.n / p.np[0]) % s.rows == 0 && cnt % s.rows_nt == 0 && cnt / s.rows_nt >= 1 &&
           cnt / s.rows_nt <= kMaxEpt && rows_lds(p, s) <= 160 * 1024;
}

bool tile_ok(const SpoPassPlan &p, const PassShape &s)
{
    if (s.tile_pen == 3) {
        for (int a = 1; a < p.ndim; ++a) {
            int stride = 1;
            for (int k = 0; k < a; ++k) stride *= p.np[k];
            if (!pen_len(p.np[a]) || stride % (256 / p.np[a])) return false;
        }
        return true;
    }
    if (s.tile_pen) {
        if (s.tile_pen > 3) return false;
        // columns per item: 64 / M (one wavefront) or 256 / M (four, tiled)
        const int cols = s.tile_pen == 2 ? 4096 : 1024;
        for (int a = 1; a < p.ndim; ++a) {
            int stride = 1;
            for (int k = 0; k < a; ++k) stride *= p.np[k];
            if (!pen_len(p.np[a]) || stride % (cols / p.np[a])) return false;
        }
        return true;
    }
    for (int a = 1; a < p.ndim; ++a) {
        int stride = 1;
        for (int k = 0; k < a; ++k) stride *= p.np[k];
        const int cnt = s.tile * p.np[a];
        if (!nt_ok(s.tile_nt, cnt)) return false;
        if (p.ntw[a] > kMaxEpt * s.tile_nt || p.np[a] > kMaxEpt * s.tile_nt) return false;
        if (s.tile > stride || cnt % s.tile_nt || cnt / s.tile_nt < 1 || cnt / s.tile_nt > kMaxEpt)
            return false;
        if (sizeof(double2) * (lds_slots(cnt, true) + p.ntw[a] + p.np[a] + 2) > 160 * 1024)
            return false;
    }
    return true;
}

using TuneKey = std::tuple<int, int, int, int, int, int>;
std::mutex g_tune_mu;
std::map<TuneKey, PassShape> g_tuned;

// Duration (ms) per launch of `fn`: the best of three timed groups of
// `reps` launches between two events (the minimum rejects interference).
template <class F>
int time_launches(hipStream_t s, int reps, F fn, float *ms)
{
    hipEvent_t e0, e1;
    ESPEC_HIP_TRY(hipEventCreate(&e0));
    ESPEC_HIP_TRY(hipEventCreate(&e1));
    ESPEC_HIP_TRY(fn());  // warm-up (code object load, LDS attribute)
    float best = 1e30f;
    for (int trial = 0; trial < 3; ++trial) {
        ESPEC_HIP_TRY(hipEventRecord(e0, s));
        for (int k = 0; k < reps; ++k) ESPEC_HIP_TRY(fn());
        ESPEC_HIP_TRY(hipEventRecord(e1, s));
        ESPEC_HIP_TRY(hipEventSynchronize(e1));
        float t = 0;
        ESPEC_HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        best = std::min(best, t / reps);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return ESPEC_OK;
}

// A candidate replaces the incumbent only when clearly faster, so that
// near-ties resolve to the same shape from one run to the next.
inline bool clearly_faster(float ms, float best) { return ms < 0.97f * best; }

// Try the candidate launch shapes of the row pass and of the strided
// passes, keep the fastest of each (timed with HIP events on the plan's
// stream), cache the choice per grid and batch size, and restore psi.
template <typename R>
int tune_typed(SpoPassPlan &p, hipStream_t s)
{
    // The candidates depend on the phase table and the mask as well.
    const TuneKey key{p.ndim, p.np[0], p.np[1], p.np[2], p.npk,
                      p.mixed | (p.vph ? 2 : 0) | (p.mask ? 4 : 0)};
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            p.shape = it->second;
            apply_shape(p);
            return ESPEC_OK;
        }
    }
    const size_t bytes = (size_t)p.n * p.npk * sizeof(V2<R>);
    const int reps = 16;
    PassShape best = p.shape;
    float best_ms = 1e30f;
    // {rows, threads, pencil}
    const int rc_[][3] = {{1, 64, 0},  {2, 64, 0},  {4, 64, 0},  {1, 256, 0},
                          {2, 256, 0}, {4, 256, 0}, {8, 256, 0}, {1024 / p.np[0], 64, 1},
                          {256 / p.np[0], 64, 2}, {1024 / p.np[0], 64, 3}};
    for (auto &c : rc_) {
        PassShape sh = p.shape;
        sh.rows = c[0];
        sh.rows_nt = c[1];
        sh.rows_pen = c[2];
        if (!rows_ok(p, sh)) continue;
        p.shape = sh;
        apply_shape(p);
        ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
        float ms = 0;
        // B = 1/np0 keeps the magnitude of psi over repeated row passes.
        int rc = time_launches(s, reps, [&]() { return rows<R>(p, true, true, s, 1.0 / p.np[0]); }, &ms);
        if (rc) return rc;
        if (clearly_faster(ms, best_ms)) {
            best_ms = ms;
            best.rows = sh.rows;
            best.rows_nt = sh.rows_nt;
            best.rows_pen = sh.rows_pen;
        }
    }
    p.shape.rows = best.rows;
    p.shape.rows_nt = best.rows_nt;
    p.shape.rows_pen = best.rows_pen;
    if (p.ndim > 1) {
        best_ms = 1e30f;
        // {tile, threads, pencil}
        const int tc_[][3] = {{1, 64, 0},  {2, 64, 0},  {4, 64, 0},  {1, 256, 0},  {2, 256, 0},
                              {4, 256, 0}, {8, 256, 0}, {16, 256, 0}, {p.shape.tile, p.shape.tile_nt, 1},
                              {p.shape.tile, p.shape.tile_nt, 2}, {p.shape.tile, p.shape.tile_nt, 3}};
        for (auto &c : tc_) {
            PassShape sh = p.shape;
            sh.tile = c[0];
            sh.tile_nt = c[1];
            sh.tile_pen = c[2];
            if (!tile_ok(p, sh)) continue;
            p.shape = sh;
            apply_shape(p);
            ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
            float ms = 0;
            // Whole steps: the strided kernels differ in how they leave the
            // state in the caches for the row pass that follows.
            int dummy = 0;
            int rc = time_launches(s, reps, [&]() { return one_step<R>(p, true, true, s, &dummy); },
                                   &ms);
            if (rc) return rc;
            if (clearly_faster(ms, best_ms)) {
                best_ms = ms;
                best.tile = sh.tile;
                best.tile_nt = sh.tile_nt;
                best.tile_pen = sh.tile_pen;
            }
        }
        p.shape.tile = best.tile;
        p.shape.tile_nt = best.tile_nt;
        p.shape.tile_pen = best.tile_pen;
    }
    // XCD placement: time whole steps, since what it changes is where the
    // state lives between one pass and the next.
    {
        best_ms = 1e30f;
        int best_spread = 1, best_xmap = 1;
        for (int xm : {1, 0})
            for (int sp : {1, 2, 4, 8}) {
                if ((p.npk > 1 || xm) && sp > 1) continue;
                p.shape.spread = sp;
                p.shape.xmap = xm;
                apply_shape(p);
                ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
                float ms = 0;
                int dummy = 0;
                int rc = time_launches(s, reps, [&]() { return one_step<R>(p, true, true, s, &dummy); },
                                       &ms);
                if (rc) return rc;
                if (clearly_faster(ms, best_ms)) {
                    best_ms = ms;
                    best_spread = sp;
                    best_xmap = xm;
                }
            }
        p.shape.spread = best_spread;
        p.shape.xmap = best_xmap;
    }
    // Workgroups per pass: one per work item, or a grid-stride loop over
    // fewer, each prefetching its next item (pays off at large batches,
    // where short workgroups leave HBM latency exposed).  Multiples of 8
    // keep a workgroup on one XCD for all of its items.
    {
        int items = p.xblocks * p.npk;
        for (int a = 1; a < p.ndim; ++a) {
            const int per = p.shape.tile_pen ? 1024 / p.np[a] : p.shape.tile;
            items = std::max(items, p.n / (per * p.np[a]) * p.npk);
        }
        best_ms = 1e30f;
        int best_grid = 0;
        int best_pf = 0;
        int best_tp = p.shape.tile_pen;
        // The two pencil forms of the strided passes are compared again
        // here, with everything else at its final setting.
        std::vector<int> tps{p.shape.tile_pen};
        if (p.shape.tile_pen && p.ndim > 1)
            for (int tp = 1; tp <= 3; ++tp) {
                PassShape alt = p.shape;
                alt.tile_pen = tp;
                if (tp != p.shape.tile_pen && tile_ok(p, alt)) tps.push_back(tp);
            }
        for (int tp : tps)
        for (int pf : {0, 1})
        for (int gs : {0, 4096, 2048, 1024, 512}) {
            if (gs && (items <= gs || p.shape.spread > 1)) continue;
            if (pf && tp != 1) continue;  // prefetch: one-wavefront pencils only
            p.shape.tile_pen = tp;
            p.shape.grid = gs;
            p.shape.pen_pf = pf;
            apply_shape(p);
            ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
            float ms = 0;
            int dummy = 0;
            int rc = time_launches(s, reps, [&]() { return one_step<R>(p, true, true, s, &dummy); }, &ms);
            if (rc) return rc;
            if (clearly_faster(ms, best_ms)) {
                best_ms = ms;
                best_grid = gs;
                best_pf = pf;
                best_tp = tp;
            }
        }
        p.shape.tile_pen = best_tp;
        p.shape.grid = best_grid;
        p.shape.pen_pf = best_pf;
    }
    // The register row passes compared once more as whole steps, with the
    // strided settings final: timed alone, a row pass finds the state where
    // the previous row pass left it, not where the strided pass does.
    if (p.shape.rows_pen) {
        const PassShape keep = p.shape;
        best_ms = 1e30f;
        PassShape best_sh = keep;
        for (int rp : {keep.rows_pen, 1, 2, 3}) {
            PassShape sh = keep;
            sh.rows_pen = rp;
            sh.rows = rp == 2 ? 256 / p.np[0] : 1024 / p.np[0];
            sh.rows_nt = 64;
            if (!rows_ok(p, sh)) continue;
            p.shape = sh;
            apply_shape(p);
            ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
            float ms = 0;
            int dummy = 0;
            int rc = time_launches(s, reps, [&]() { return one_step<R>(p, true, true, s, &dummy); }, &ms);
            if (rc) return rc;
            if (clearly_faster(ms, best_ms)) {
                best_ms = ms;
                best_sh = sh;
            }
        }
        p.shape = best_sh;
    }
    // The LDS-resident form where the grid has one: launches of 16 steps
    // against the final multi-pass step, per step.  Its samples go past
    // maxsamp, so the tuning writes none.
    if (fused_grid(p)) {
        const PassShape keep = p.shape;
        apply_shape(p);
        ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
        int dummy = 0;
        best_ms = 1e30f;
        int rc = time_launches(s, reps, [&]() { return one_step<R>(p, true, true, s, &dummy); }, &best_ms);
        if (rc) return rc;
        constexpr int kSteps = 16;
        const long long past = (long long)p.maxsamp * p.lc;
        int best_nw = 0;
        for (int nw : {4, 8}) {
            PassShape sh = keep;
            sh.fused = nw;
            if (!fused_ok(p, sh)) continue;
            p.shape = sh;
            ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
            float ms = 0;
            rc = time_launches(s, 2, [&]() { return fused_launch<R>(p, kSteps, past, s); }, &ms);
            if (rc) return rc;
            if (clearly_faster(ms / kSteps, best_ms)) {
                best_ms = ms / kSteps;
                best_nw = nw;
            }
        }
        p.shape = keep;
        p.shape.fused = best_nw;
    }
    apply_shape(p);
    ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
    ESPEC_HIP_TRY(hipStreamSynchronize(s));
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tuned[key] = p.shape;
    return ESPEC_OK;
}

// Per-kernel timing without graphs: events around every row pass and every
// strided pass of steps 2..nsteps.  Algorithmic bytes per launch: the row
// pass reads and writes psi and reads V, plus psi0 at sample steps; a
// strided pass reads and writes psi.
template <typename R>
int profile_typed(SpoPassPlan &p, int nsteps, long long step0, hipStream_t s, double *ms, double *bytes)
{
    if (p.shape.fused) {
        // One launch of nsteps: psi read and written once, V every step,
        // psi0 at every sample step.
        hipEvent_t e0, e1;
        ESPEC_HIP_TRY(hipEventCreate(&e0));
        ESPEC_HIP_TRY(hipEventCreate(&e1));
        tl_ev_start = e0;
        tl_ev_stop = e1;
        const hipError_t e = fused_launch<R>(p, nsteps, step0, s);
        tl_ev_start = tl_ev_stop = nullptr;
        ESPEC_HIP_TRY(e);
        ESPEC_HIP_TRY(hipStreamSynchronize(s));
        float t = 0;
        ESPEC_HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const double vs = sizeof(V2<R>), rs = sizeof(R);
        ms[0] = t;
        bytes[0] = (double)p.n * p.npk * (2 * vs + nsteps * rs + vs * ((double)nsteps / p.lc));
        ms[1] = bytes[1] = 0;
        return ESPEC_OK;
    }
    std::vector<hipEvent_t> ev;
    // Events recorded by the kernel dispatch itself (start, stop).
    auto arm = [&]() -> hipError_t {
        hipEvent_t e0, e1;
        hipError_t r = hipEventCreate(&e0);
        if (r == hipSuccess) r = hipEventCreate(&e1);
        if (r != hipSuccess) return r;
        ev.push_back(e0);
        ev.push_back(e1);
        tl_ev_start = e0;
        tl_ev_stop = e1;
        return hipSuccess;
    };
    auto disarm = [&]() { tl_ev_start = tl_ev_stop = nullptr; };
    int dummy = 0;
    hipLaunchKernelGGL(set_step, dim3(1), dim3(1), 0, s, p.dstep, step0);
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(one_step<R>(p, false, true, s, &dummy));
    std::vector<int> kind;
    hipError_t err = hipSuccess;
    for (int k = 1; k < nsteps && err == hipSuccess; ++k) {
        err = arm();
        if (err == hipSuccess) err = rows<R>(p, true, true, s);
        disarm();
        kind.push_back(0);
        if (p.ndim == 1 && err == hipSuccess) {
            hipLaunchKernelGGL(add_step, dim3(1), dim3(1), 0, s, p.dstep);
            err = hipGetLastError();
        }
        for (int ax = 1; ax < p.ndim && err == hipSuccess; ++ax) {
            err = arm();
            if (err == hipSuccess) err = strided<R>(p, ax, ax == p.ndim - 1, s);
            disarm();
            kind.push_back(1);
        }
    }
    ESPEC_HIP_TRY(err);
    ESPEC_HIP_TRY(one_step<R>(p, true, false, s, &dummy));
    ESPEC_HIP_TRY(hipStreamSynchronize(s));
    double sum[2] = {0, 0};
    int cnt[2] = {0, 0};
    for (size_t i = 0; i < kind.size(); ++i) {
        float t = 0;
        ESPEC_HIP_TRY(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
        sum[kind[i]] += t;
        cnt[kind[i]]++;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    const double vs = sizeof(V2<R>), rs = sizeof(R);
    const double samp = 1.0 / p.lc;  // share of row passes that read psi0
    ms[0] = cnt[0] ? sum[0] / cnt[0] : 0;
    ms[1] = cnt[1] ? sum[1] / cnt[1] : 0;
    bytes[0] = (double)p.n * p.npk * (2 * vs + rs + samp * vs);
    bytes[1] = (double)p.n * p.npk * (2 * vs);
    return ESPEC_OK;
}

}  // namespace

int spo_pass_setup(SpoPassPlan &p, int ndim, const int *np, int npk, bool mixed,
                   const std::vector<FftPlan> &fp, const std::vector<std::vector<double>> &taxes,
                   double cst1, double cst2, void *x, const void *psi0, const void *vpot,
                   const void *vph, const void *vph_fz, const double *mask, double *cbuf,
                   int maxsamp, double phase_max, int tune, hipStream_t s)
{
    p.phase_max = phase_max;
    p.mask = mask;
    p.vph_fz = vph_fz;
    // taxes holds ndim kinetic axes for one shared set of masses, or
    // ndim per packet (packet-major) when the masses differ per packet.
    const int nset = (int)taxes.size() / ndim;
    p.vph = vph;
    p.cst2 = cst2;
    p.ndim = ndim;
    p.n = 1;
    for (int a = 0; a < ndim; ++a) {
        p.np[a] = np[a];
        p.n *= np[a];
        p.fpl[a] = fp[a];
    }
    p.npk = npk;
    p.mixed = mixed;
    p.maxsamp = maxsamp;
    p.x = x;
    p.psi0 = psi0;
    p.vpot = vpot;
    p.cbuf = cbuf;
    p.b = 1.0 / p.n;
    // Default shape (the tuner may replace it): one x line per 64-thread
    // workgroup; strided tiles of 128-byte rows.
    p.shape.rows = 1;
    p.shape.rows_nt = 64;
    p.shape.tile = std::min(mixed ? 16 : 8, np[0]);
    p.shape.tile_nt = 256;
    // Tables: twiddles per axis and per-axis kinetic phases (cos, sin) of
    // cst1 * T_a(k_a) stored at the digit-reversed position of k_a.
    std::vector<double> tw, kax;
    int pstride = 0;
    for (int a = 0; a < ndim; ++a) {
        p.tw_off[a] = (int)(tw.size() / 2);
        p.ntw[a] = (int)fp[a].twc.size();
        for (size_t i = 0; i < fp[a].twc.size(); ++i) {
            tw.push_back(fp[a].twc[i]);
            tw.push_back(fp[a].tws[i]);
        }
        p.kax_off[a] = pstride;
        pstride += np[a];
    }
    p.kax_pstride = nset > 1 ? pstride : 0;
    kax.resize(2 * (size_t)pstride * nset);
    // Pencil tables: exact twiddles exp(2 pi i m / L), and the kinetic
    // phases at the spectral-layout position of each frequency.
    std::vector<double> twr, kreg(kax.size()), knat(kax.size());
    for (int a = 0; a < ndim; ++a) {
        p.twr_off[a] = (int)(twr.size() / 2);
        for (int m = 0; m < np[a]; ++m) {
            twr.push_back(std::cos(kTwoPi * m / np[a]));
            twr.push_back(std::sin(kTwoPi * m / np[a]));
        }
    }
    for (int q = 0; q < nset; ++q)
        for (int a = 0; a < ndim; ++a) {
            const size_t k0 = 2 * ((size_t)q * pstride + p.kax_off[a]);
            for (int k = 0; k < np[a]; ++k) {
                const double th = cst1 * taxes[(size_t)q * ndim + a][k];
                const int pos = fp[a].perm[k];
                kax[k0 + 2 * pos] = std::cos(th);
                kax[k0 + 2 * pos + 1] = std::sin(th);
                knat[k0 + 2 * k] = std::cos(th);
                knat[k0 + 2 * k + 1] = std::sin(th);
            }
            if (pen_len(np[a]))
                for (int pos = 0; pos < np[a]; ++pos) {
                    const double th =
                        cst1 * taxes[(size_t)q * ndim + a][pencil::spectral_index(pos, np[a] / 16)];
                    kreg[k0 + 2 * pos] = std::cos(th);
                    kreg[k0 + 2 * pos + 1] = std::sin(th);
                }
        }
    if (!rows_ok(p, p.shape)) {
        bool found = false;
        for (int r : {1, 2, 4, 8, 16})
            for (int nt : {64, 256}) {
                PassShape sh = p.shape;
                sh.rows = r;
                sh.rows_nt = nt;
                if (!found && rows_ok(p, sh)) {
                    p.shape = sh;
                    found = true;
                }
            }
        if (!found) return ESPEC_ERR_UNSUPPORTED;
    }
    if (ndim > 1 && !tile_ok(p, p.shape)) {
        for (int t : {16, 8, 4, 2, 1})
            for (int nt : {256, 64}) {
                PassShape sh = p.shape;
                sh.tile = t;
                sh.tile_nt = nt;
                if (tile_ok(p, sh)) {
                    p.shape = sh;
                    goto shaped;
                }
            }
        return ESPEC_ERR_UNSUPPORTED;
    }
shaped:
    apply_shape(p);
    const int max_xblocks = p.n / p.np[0];  // rows = 1
    // Split single-precision copies: per axis of length L = 16 M, the
    // forward twiddles W^{j k1} at k1 M + j, then the inverse twiddles
    // W^{(j Q + q) n2} at L + (q M + n2) M + j (Q = 16 / M); the kinetic
    // phases entry for entry.
    std::vector<float> twr_s, kreg_s;
    auto push_split = [](std::vector<float> &v, double c, double sn) {
        const float ch = (float)c, sh = (float)sn;
        v.push_back(ch);
        v.push_back(sh);
        v.push_back((float)(c - (double)ch));
        v.push_back((float)(sn - (double)sh));
    };
    for (int a = 0; a < ndim; ++a) {
        const int L = np[a], M = L / 16, Q = pen_len(L) ? 16 / M : 0;
        auto wpow = [&](long long m) {
            push_split(twr_s, std::cos(kTwoPi * (double)(m % L) / L), std::sin(kTwoPi * (double)(m % L) / L));
        };
        if (!pen_len(L)) {
            for (int i = 0; i < 2 * L; ++i) push_split(twr_s, 1.0, 0.0);
            continue;
        }
        for (int k1 = 0; k1 < 16; ++k1)
            for (int j = 0; j < M; ++j) wpow((long long)j * k1);
        for (int q = 0; q < Q; ++q)
            for (int n2 = 0; n2 < M; ++n2)
                for (int j = 0; j < M; ++j) wpow((long long)(j * Q + q) * n2);
    }
    for (size_t i = 0; i < kreg.size(); i += 2) push_split(kreg_s, kreg[i], kreg[i + 1]);
    DevBuf btw, bk, btwr, bkr, btwf, bkrf, bkn, bpart, bstep;
    ESPEC_HIP_TRY(bkn.upload(knat.data(), sizeof(double) * knat.size()));
    ESPEC_HIP_TRY(btw.upload(tw.data(), sizeof(double) * tw.size()));
    ESPEC_HIP_TRY(bk.upload(kax.data(), sizeof(double) * kax.size()));
    ESPEC_HIP_TRY(btwr.upload(twr.data(), sizeof(double) * twr.size()));
    ESPEC_HIP_TRY(bkr.upload(kreg.data(), sizeof(double) * kreg.size()));
    ESPEC_HIP_TRY(btwf.upload(twr_s.data(), sizeof(float) * twr_s.size()));
    ESPEC_HIP_TRY(bkrf.upload(kreg_s.data(), sizeof(float) * kreg_s.size()));
    ESPEC_HIP_TRY(bpart.alloc(sizeof(double) * 3 * (size_t)maxsamp * npk * max_xblocks));
    ESPEC_HIP_TRY(hipMemset(bpart.p, 0, bpart.bytes));
    ESPEC_HIP_TRY(bstep.alloc(sizeof(long long)));
    p.tw = btw.p;
    btw.p = nullptr;
    p.kax = bk.p;
    bk.p = nullptr;
    p.twr = btwr.p;
    btwr.p = nullptr;
    p.kreg = bkr.p;
    bkr.p = nullptr;
    p.twr_s = btwf.p;
    btwf.p = nullptr;
    p.knat = bkn.p;
    bkn.p = nullptr;
    p.kreg_s = bkrf.p;
    bkrf.p = nullptr;
    p.part = static_cast<double *>(bpart.p);
    bpart.p = nullptr;
    p.dstep = static_cast<long long *>(bstep.p);
    bstep.p = nullptr;
    // A defined step counter before any row pass reads it (the tuner's
    // row passes sample into slot 0).
    hipLaunchKernelGGL(set_step, dim3(1), dim3(1), 0, s, p.dstep, 0LL);
    ESPEC_HIP_TRY(hipGetLastError());
    if (tune) return mixed ? tune_typed<float>(p, s) : tune_typed<double>(p, s);
    return ESPEC_OK;
}

int spo_pass_profile(SpoPassPlan &p, int nsteps, long long step0, int lc, hipStream_t s, double *ms,
                     double *bytes)
{
    p.lc = lc;
    if (nsteps < 2) return ESPEC_ERR_ARG;
    return p.mixed ? profile_typed<float>(p, nsteps, step0, s, ms, bytes)
                   : profile_typed<double>(p, nsteps, step0, s, ms, bytes);
}

int spo_pass_prepare(SpoPassPlan &p, int lc, hipStream_t s, bool prime)
{
    if (lc != p.lc) drop_graph(p);
    p.lc = lc;
    if (p.shape.fused) return ESPEC_OK;  // no step graphs: one launch per chunk of steps
    for (int j = 0; j < SpoPassPlan::kGraphLevels; ++j) {
        const int rc = p.mixed ? capture_typed<float>(p, j) : capture_typed<double>(p, j);
        if (rc) return rc;
    }
    if (prime) {
        // 63 steps in all; the row passes write only row-block partials,
        // which a run rewrites before it reduces them.
        for (int j = 0; j < SpoPassPlan::kGraphLevels; ++j) ESPEC_HIP_TRY(hipGraphLaunch(p.gexec[j], s));
        const size_t bytes = (size_t)p.n * p.npk * (p.mixed ? sizeof(float2) : sizeof(double2));
        ESPEC_HIP_TRY(hipMemcpyAsync(p.x, p.psi0, bytes, hipMemcpyDeviceToDevice, s));
        ESPEC_HIP_TRY(hipStreamSynchronize(s));
    }
    return ESPEC_OK;
}

int spo_pass_run(SpoPassPlan &p, int nsteps, long long step0, int lc, hipStream_t s, int *launches)
{
    if (lc != p.lc) drop_graph(p);  // the captured kernels carry lc
    p.lc = lc;
    *launches = 0;
    if (nsteps <= 0) return ESPEC_OK;
    return p.mixed ? run_typed<float>(p, nsteps, step0, s, launches)
                   : run_typed<double>(p, nsteps, step0, s, launches);
}

int spo_pass_enqueue(SpoPassPlan &p, int nsteps, long long step0, int lc, hipStream_t s)
{
    if (lc != p.lc) drop_graph(p);  // the plan's own step graphs carry lc
    p.lc = lc;
    if (nsteps <= 0) return ESPEC_OK;
    return p.mixed ? enqueue_typed<float>(p, nsteps, step0, s)
                   : enqueue_typed<double>(p, nsteps, step0, s);
}

int spo_pass_set_shape(SpoPassPlan &p, const PassShape &s)
{
    if (!rows_ok(p, s) || (p.ndim > 1 && !tile_ok(p, s)) || !fused_ok(p, s)) return ESPEC_ERR_ARG;
    if (s.spread < 1 || s.spread > 8 || (s.spread & (s.spread - 1)) || s.grid < 0 ||
        (s.grid % 8) != 0)
        return ESPEC_ERR_ARG;
    drop_graph(p);  // the captured launches carry the old shape
    p.shape = s;
    apply_shape(p);
    return ESPEC_OK;
}

bool spo_pass_fused_grid(int ndim, const int *np, bool mixed)
{
    if (ndim != 2) return false;
    if (np[0] == 64 && np[1] == 64) return true;
    return mixed && np[0] == 128 && np[1] == 128;
}

void spo_pass_release(SpoPassPlan &p)
{
    drop_graph(p);
    if (p.tw) (void)hipFree(p.tw);
    if (p.kax) (void)hipFree(p.kax);
    if (p.twr) (void)hipFree(p.twr);
    if (p.kreg) (void)hipFree(p.kreg);
    if (p.twr_s) (void)hipFree(p.twr_s);
    if (p.knat) (void)hipFree(p.knat);
    if (p.kreg_s) (void)hipFree(p.kreg_s);
    if (p.part) (void)hipFree(p.part);
    if (p.dstep) (void)hipFree(p.dstep);
    p = SpoPassPlan{};
}

}  // namespace espec
