// Second-order-differencing propagation (PSODFFT / PSODFFTAUX) on the device.
//
// Step (sodfft.f:54-129):  psi(t+dt) = psi(t-dt) - 2 i dt/hbar H psi(t),
// H psi = V psi + F^-1 T F psi / N, started by two half-strength steps from
// psi0 after which psi(t-dt) is reset to psi0 (psodfft.f:118-128).
//
// LDS-resident kernels (one workgroup per packet, grid up to 2048 points)
// keep psi(t-dt), psi(t) and the transform buffer in LDS in FP64 for every
// step of a launch:
//   sod_lds_fast (ESPEC_EXACT): forward DIF of all axes, T applied in the
//     digit-reversed order, inverse DIT; the half swap of the flattened
//     array around the transform is a translation along the slowest axis
//     and commutes with T, so it is dropped.
//   sod_lds_ref  (ESPEC_REF, FP64): the reference's order: half swap,
//     natural-order transforms per axis with the f32 twiddles, T in natural
//     order, inverse, swap back, and its order of the update arithmetic.
// Larger grids run sod_pass, one launch per axis per step: T = sum_a T_a,
// so each pass transforms the lines of one axis of psi(t), multiplies by
// T_a / n_a and accumulates into an FP64 buffer; the last axis' pass also
// forms psi(t+dt) and the C(t) partial sums (ESPEC_EXACT only).
//
// Absorbers (ABSORB=.TRUE.): .VOPTIC adds dt-scaled W psi(t) to the
// leapfrog sum (sodfft.f:66-72); .SMOOTHW multiplies psi(t+dt) by each
// point's mask.  The reference never reaches its mask loop (sodfft.f:108
// compares the 7 characters TPABSOR(1:7) with the 8-character '.SMOOTHW'),
// so ESPEC_REF propagates .SMOOTHW without absorber.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "api_common.hpp"
#include "axis.hpp"
#include "fft_lds.hpp"

namespace espec {

constexpr int kSodLdsMax = 2048;  // points of an LDS-resident grid

struct SodArgs {
    int ndim, n, npk;
    DevAxis ax[3];
    const double2 *tw;
    const int *perm;
    int perm_off[3];
    void *cur;            // [npk][n] psi(t), storage precision
    void *prev;           // [npk][n] psi(t-dt)
    const void *psi0;     // [npk][n]
    const double *vpot;   // [npk][n]
    const double *tc;     // [n] T: natural order (ref) or digit-reversed (fast)
    double b, cst;        // 1/N, 2 dt / SHFS
    int nsteps, lc, maxsamp;
    long long step0;      // steps done before this launch (the start-up counts as one)
    double *cbuf;         // [npk][maxsamp][3]
    int absorb;           // ESPEC_ABSORB_*
    const double *vabc;   // [npk][n] imaginary potential or mask, or null
};

// new = prev - i h (V c + b x) in the reference's operation order
// (sodfft.f:56-59 and 118-127).
__device__ __forceinline__ double2 sod_update(double2 prev, double2 c, double2 x, double hv,
                                              double hb)
{
    const double wr = prev.x + hv * c.y, wi = prev.y - hv * c.x;
    return {wr + hb * x.y, wi - hb * x.x};
}

// sod_update with an absorber: hw = h W adds hw c to the leapfrog sum
// (sodfft.f:68-69); mk (1 without a mask) scales the result (sodfft.f:112-115).
__device__ __forceinline__ double2 sod_update_abs(double2 prev, double2 c, double2 x, double hv,
                                                  double hb, int absorb, double hw, double mk)
{
    double wr = prev.x + hv * c.y, wi = prev.y - hv * c.x;
    if (absorb == ESPEC_ABSORB_VOPTIC) {
        wr = wr + hw * c.x;
        wi = wi + hw * c.y;
    }
    double2 r = {wr + hb * x.y, wi - hb * x.x};
    if (absorb == ESPEC_ABSORB_SMOOTHW) r = {mk * r.x, mk * r.y};
    return r;
}

// H-application and update, fast path: x <- c, transform, T, inverse.
template <int NT>
__device__ void sod_apply_fast(double2 *x, const double2 *c, const SodArgs &a, const double *tc)
{
    for (int i = threadIdx.x; i < a.n; i += NT) x[i] = c[i];
    __syncthreads();
    for (int d = 0; d < a.ndim; ++d)
        for (int s = 0; s < a.ax[d].nst; ++s) dif_stage<double, NT, 1>(x, a.ax[d], s, a.tw);
    for (int i = threadIdx.x; i < a.n; i += NT) {
        const double t = tc[i];
        x[i] = {t * x[i].x, t * x[i].y};
    }
    __syncthreads();
    for (int d = a.ndim - 1; d >= 0; --d)
        for (int s = a.ax[d].nst - 1; s >= 0; --s) dit_stage<double, NT, -1>(x, a.ax[d], s, a.tw);
}

// Reference path: x <- swap(c), FFT(+1) per axis, T, FFT(-1) per axis
// (sodfft.f:61-117); returns the buffer holding the result (x or y).
template <int NT>
__device__ double2 *sod_apply_ref(double2 *x, double2 *y, const double2 *c, const SodArgs &a,
                                  const double *tc)
{
    const int nh = a.n / 2;
    for (int i = threadIdx.x; i < a.n; i += NT) x[i] = c[i < nh ? i + nh : i - nh];
    __syncthreads();
    for (int d = 0; d < a.ndim; ++d) {
        for (int s = 0; s < a.ax[d].nst; ++s) dif_stage<double, NT, 1>(x, a.ax[d], s, a.tw);
        permute_axis<double, NT>(x, y, a.ax[d], a.perm + a.perm_off[d]);
        double2 *t = x; x = y; y = t;
    }
    for (int i = threadIdx.x; i < a.n; i += NT) {
        const double t = tc[i];
        x[i] = {t * x[i].x, t * x[i].y};
    }
    __syncthreads();
    for (int d = 0; d < a.ndim; ++d) {
        for (int s = 0; s < a.ax[d].nst; ++s) dif_stage<double, NT, -1>(x, a.ax[d], s, a.tw);
        permute_axis<double, NT>(x, y, a.ax[d], a.perm + a.perm_off[d]);
        double2 *t = x; x = y; y = t;
    }
    return x;
}

template <typename S, int NT, bool REF>
__global__ __launch_bounds__(NT) void sod_lds(SodArgs a)
{
    using VS = V2<S>;
    extern __shared__ double smem_d[];
    double2 *c = reinterpret_cast<double2 *>(smem_d);   // psi(t)
    double2 *p = c + a.n;                               // psi(t-dt)
    double2 *x = p + a.n;                               // transform
    double2 *y = x + a.n;                               // permutation scratch (REF)
    double *red = reinterpret_cast<double *>(x + (REF ? 2 : 1) * a.n);
    const int pk = blockIdx.x, n = a.n, nh = n / 2;
    const long long po = (long long)pk * n;
    VS *gc = reinterpret_cast<VS *>(a.cur) + po;
    VS *gp = reinterpret_cast<VS *>(a.prev) + po;
    const VS *g0 = reinterpret_cast<const VS *>(a.psi0) + po;
    const double *vp = a.vpot + po;
    for (int i = threadIdx.x; i < n; i += NT) {
        const VS t = gc[i], q = gp[i];
        c[i] = {(double)t.x, (double)t.y};
        p[i] = {(double)q.x, (double)q.y};
    }
    __syncthreads();
    for (int st = 0; st < a.nsteps; ++st) {
        const long long g = a.step0 + st;
        const int sub = g == 0 ? 2 : 1;      // start-up: two half-strength steps
        const double h = g == 0 ? a.cst / 2.0 : a.cst;
        for (int k = 0; k < sub; ++k) {
            const double2 *r;
            if (REF) r = sod_apply_ref<NT>(x, y, c, a, a.tc);
            else {
                sod_apply_fast<NT>(x, c, a, a.tc);
                r = x;
            }
            for (int i = threadIdx.x; i < n; i += NT) {
                const double2 xv = REF ? r[i < nh ? i + nh : i - nh] : r[i];
                double2 nw;
                if (a.absorb == ESPEC_ABSORB_NONE) {
                    nw = sod_update(p[i], c[i], xv, h * vp[i], h * a.b);
                } else {
                    const double *va = a.vabc + po;
                    const double w = a.absorb == ESPEC_ABSORB_VOPTIC ? h * va[i] : 0.0;
                    const double mk = a.absorb != ESPEC_ABSORB_SMOOTHW ? 1.0 : va[i];
                    nw = sod_update_abs(p[i], c[i], xv, h * vp[i], h * a.b, a.absorb, w, mk);
                }
                p[i] = c[i];
                c[i] = nw;
            }
            __syncthreads();
        }
        if (g == 0) {  // psodfft.f:127-128: psi(t-dt) restarts from psi0
            for (int i = threadIdx.x; i < n; i += NT) {
                const VS t = g0[i];
                p[i] = {(double)t.x, (double)t.y};
            }
            __syncthreads();
        }
        const long long gs = g + 1;
        if (gs % a.lc == 0) {
            double cr = 0, ci = 0, nr = 0;
            for (int i = threadIdx.x; i < n; i += NT) {
                const VS q = g0[i];
                const double2 v = c[i];
                cr += (double)q.x * v.x + (double)q.y * v.y;
                ci += (double)q.x * v.y - (double)q.y * v.x;
                nr += v.x * v.x + v.y * v.y;
            }
            block_sum3<NT>(cr, ci, nr, red);
            const long long nc = gs / a.lc;
            if (threadIdx.x == 0 && nc < a.maxsamp) {
                double *cb = a.cbuf + ((long long)pk * a.maxsamp + nc) * 3;
                cb[0] = cr;
                cb[1] = ci;
                cb[2] = nr;
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += NT) {
        gc[i] = {(S)c[i].x, (S)c[i].y};
        gp[i] = {(S)p[i].x, (S)p[i].y};
    }
}

size_t sod_lds_bytes(int n, bool ref, int nt)
{
    return sizeof(double2) * (size_t)n * (ref ? 4 : 3) + sizeof(double) * 3 * (nt / 64);
}

// ------------------------------------------------------- multi-pass

struct SodPassArgs {
    int n, npk, axis, na, nlow, nhigh, tile;
    DevAxis ax;           // in-LDS geometry of one tile
    const double2 *tw;
    const double *ta;     // [na] T_a / n_a in digit-reversed order
    void *cur, *prev;     // storage precision
    double2 *acc;         // [npk][n] FP64 sum over axes of T_a psi
    const void *psi0;
    const double *vpot;
    double h;             // this step's 2 dt / SHFS (halved in the start-up)
    int first, last;      // pass roles
    int lc, maxsamp, nosample, nblk;
    const long long *dstep;
    double *cbuf, *part;
    unsigned *ticket;
    int absorb;           // ESPEC_ABSORB_*
    const double *vabc;   // [npk][n] or null
};

// One axis of one step for a tile of lines: forward transform of psi(t)
// along the axis, T_a, inverse; accumulate; in the last pass form
// psi(t+dt), shift psi(t) into psi(t-dt) and add the C(t) partial sums.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void sod_pass(SodPassArgs a)
{
    using VS = V2<S>;
    extern __shared__ double smem_d[];
    double2 *x = reinterpret_cast<double2 *>(smem_d);
    __shared__ double red[3 * (NT / 64)];
    __shared__ bool lastblk;
    const int pk = blockIdx.y;
    const long long po = (long long)pk * a.n;
    VS *gc = reinterpret_cast<VS *>(a.cur) + po;
    VS *gp = reinterpret_cast<VS *>(a.prev) + po;
    double2 *gw = a.acc + po;
    const VS *g0 = reinterpret_cast<const VS *>(a.psi0) + po;
    const double *vp = a.vpot + po;
    const int T = a.tile, na = a.na;
    // Lines of this block: axis 0 -> T consecutive rows; axis > 0 -> T
    // consecutive positions of the lower axes at one position of the upper.
    long long base;
    int lstride, estride;  // global step between the tile's lines / elements
    if (a.axis == 0) {
        base = (long long)blockIdx.x * T * na;
        lstride = na;
        estride = 1;
    } else {
        const int lt = a.nlow / T;
        const int hi = blockIdx.x / lt, lo = (blockIdx.x % lt) * T;
        base = (long long)hi * a.nlow * na + lo;
        lstride = 1;
        estride = a.nlow;
    }
    // LDS index of (line t, element k): axis 0 keeps lines contiguous, the
    // others keep the tile's lines adjacent (coalesced global access).
    auto lidx = [&](int t, int k) { return a.axis == 0 ? t * na + k : k * T + t; };
    auto gidx = [&](int t, int k) { return base + (long long)t * lstride + (long long)k * estride; };
    for (int q = threadIdx.x; q < T * na; q += NT) {
        const int t = a.axis == 0 ? q / na : q % T, k = a.axis == 0 ? q % na : q / T;
        const VS v = gc[gidx(t, k)];
        x[lidx(t, k)] = {(double)v.x, (double)v.y};
    }
    __syncthreads();
    for (int s = 0; s < a.ax.nst; ++s) dif_stage<double, NT, 1>(x, a.ax, s, a.tw);
    for (int q = threadIdx.x; q < T * na; q += NT) {
        const int t = a.axis == 0 ? q / na : q % T, k = a.axis == 0 ? q % na : q / T;
        const double tk = a.ta[k];
        const int li = lidx(t, k);
        x[li] = {tk * x[li].x, tk * x[li].y};
    }
    __syncthreads();
    for (int s = a.ax.nst - 1; s >= 0; --s) dit_stage<double, NT, -1>(x, a.ax, s, a.tw);
    double cr = 0, ci = 0, nr = 0;
    const long long g = a.last ? *a.dstep + 1 : 0;
    const bool samp = a.last && !a.nosample && (g % a.lc) == 0;
    for (int q = threadIdx.x; q < T * na; q += NT) {
        const int t = a.axis == 0 ? q / na : q % T, k = a.axis == 0 ? q % na : q / T;
        const long long gi = gidx(t, k);
        double2 r = x[lidx(t, k)];
        if (!a.first) {
            const double2 w = gw[gi];
            r = {w.x + r.x, w.y + r.y};
        }
        if (!a.last) {
            gw[gi] = r;
            continue;
        }
        const VS cv = gc[gi], pv = gp[gi];
        const double2 c = {(double)cv.x, (double)cv.y};
        double2 nw;
        if (a.absorb == ESPEC_ABSORB_NONE) {
            nw = sod_update({(double)pv.x, (double)pv.y}, c, r, a.h * vp[gi], a.h);
        } else {
            const double va = a.vabc[po + gi];
            nw = sod_update_abs({(double)pv.x, (double)pv.y}, c, r, a.h * vp[gi], a.h, a.absorb,
                                a.h * va, va);
        }
        const VS ns = {(S)nw.x, (S)nw.y};
        gp[gi] = cv;
        gc[gi] = ns;
        if (samp) {
            const VS q0 = g0[gi];
            const double ux = (double)ns.x, uy = (double)ns.y;
            cr += (double)q0.x * ux + (double)q0.y * uy;
            ci += (double)q0.x * uy - (double)q0.y * ux;
            nr += ux * ux + uy * uy;
        }
    }
    if (!samp) return;
    block_sum3<NT>(cr, ci, nr, red);
    if (threadIdx.x == 0) {
        double *pp = a.part + ((long long)pk * a.nblk + blockIdx.x) * 3;
        pp[0] = cr;
        pp[1] = ci;
        pp[2] = nr;
        __threadfence();
        const unsigned tk = atomicAdd(&a.ticket[pk], 1u);
        lastblk = tk == (unsigned)a.nblk - 1;
    }
    __syncthreads();
    if (!lastblk || threadIdx.x != 0) return;
    __threadfence();
    double s0 = 0, s1 = 0, s2 = 0;
    for (int b = 0; b < a.nblk; ++b) {
        const double *pp = a.part + ((long long)pk * a.nblk + b) * 3;
        s0 += pp[0];
        s1 += pp[1];
        s2 += pp[2];
    }
    a.ticket[pk] = 0;
    const long long nc = g / a.lc;
    if (nc < a.maxsamp) {
        double *cb = a.cbuf + ((long long)pk * a.maxsamp + nc) * 3;
        cb[0] = s0;
        cb[1] = s1;
        cb[2] = s2;
    }
}

__global__ void sod_step_done(long long *dstep) { *dstep += 1; }
__global__ void sod_set_step(long long *dstep, long long v) { *dstep = v; }

template <typename S>
__global__ __launch_bounds__(256) void sod_copy(V2<S> *dst, const V2<S> *src, long long n)
{
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += gridDim.x * 256LL) dst[i] = src[i];
}

hipError_t launch_spo_sample(const void *x, const void *psi0, int n, int npk, bool mixed,
                             double *cbuf, int maxsamp, int idx, hipStream_t s);

}  // namespace espec

using namespace espec;

struct espec_sod_plan {
    int ndim = 1, np[3] = {1, 1, 1}, n = 0, npk = 0, lc = 1, maxsamp = 0;
    bool mixed = false, ref = false, lds = false;
    int nt = 256;
    long long steps_done = 0;
    hipStream_t stream = nullptr;
    DevBuf cur, prev, psi0, vpot, tc, tw, perm, cbuf, acc, part, ticket, dstep, tax;
    int absorb = ESPEC_ABSORB_NONE;
    DevBuf vabc;  // [npk][n] absorber data (VABC)
    SodArgs la{};
    std::vector<SodPassArgs> passes;  // one per axis
    int pass_grid[3] = {0, 0, 0};
    size_t pass_lds[3] = {0, 0, 0};
    ~espec_sod_plan()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename K>
hipError_t launch_dyn(K kern, dim3 grid, int nt, size_t lds, hipStream_t s, const SodArgs &a)
{
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, grid, dim3(nt), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_sod_lds(espec_sod_plan &P, hipStream_t s)
{
    const size_t lds = sod_lds_bytes(P.n, P.ref, P.nt);
    const dim3 g(P.npk);
    if (P.ref) {
        return P.nt == 64 ? launch_dyn(sod_lds<double, 64, true>, g, 64, lds, s, P.la)
                          : launch_dyn(sod_lds<double, 256, true>, g, 256, lds, s, P.la);
    }
    if (P.mixed)
        return P.nt == 64 ? launch_dyn(sod_lds<float, 64, false>, g, 64, lds, s, P.la)
                          : launch_dyn(sod_lds<float, 256, false>, g, 256, lds, s, P.la);
    return P.nt == 64 ? launch_dyn(sod_lds<double, 64, false>, g, 64, lds, s, P.la)
                      : launch_dyn(sod_lds<double, 256, false>, g, 256, lds, s, P.la);
}

// One multi-pass step with strength h; the start-up's first half step
// takes no sample and does not advance the step counter.
template <typename S>
hipError_t sod_pass_step(espec_sod_plan &P, double h, bool nosample, hipStream_t s)
{
    for (int a = 0; a < P.ndim; ++a) {
        SodPassArgs pa = P.passes[a];
        pa.h = h;
        pa.nosample = nosample ? 1 : 0;
        auto kern = sod_pass<S, 256>;
        if (P.pass_lds[a] > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)P.pass_lds[a]);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kern, dim3(P.pass_grid[a], P.npk), dim3(256), P.pass_lds[a], s, pa);
    }
    if (!nosample) hipLaunchKernelGGL(sod_step_done, dim3(1), dim3(1), 0, s, P.dstep.as<long long>());
    return hipGetLastError();
}

template <typename S>
int sod_run_pass(espec_sod_plan &P, int nsteps)
{
    hipLaunchKernelGGL(sod_set_step, dim3(1), dim3(1), 0, P.stream, P.dstep.as<long long>(),
                       P.steps_done);
    ESPEC_HIP_TRY(hipGetLastError());
    const double cst = P.la.cst;
    int k = 0;
    if (P.steps_done == 0 && nsteps > 0) {
        ESPEC_HIP_TRY(sod_pass_step<S>(P, cst / 2.0, true, P.stream));
        ESPEC_HIP_TRY(sod_pass_step<S>(P, cst / 2.0, false, P.stream));
        const long long tot = (long long)P.n * P.npk;
        hipLaunchKernelGGL((sod_copy<S>), dim3(std::min<long long>(4096, (tot + 255) / 256)),
                           dim3(256), 0, P.stream, (V2<S> *)P.prev.p, (const V2<S> *)P.psi0.p, tot);
        ESPEC_HIP_TRY(hipGetLastError());
        k = 1;
    }
    for (; k < nsteps; ++k) ESPEC_HIP_TRY(sod_pass_step<S>(P, cst, false, P.stream));
    return ESPEC_OK;
}

template <typename S>
int sod_build(espec_sod_plan &P, const double *sh, const double *sm, const double *vpot,
              const double *u, const double *v, const double *vabc)
{
    const int n = P.n, npk = P.npk;
    if (P.absorb != ESPEC_ABSORB_NONE)
        ESPEC_HIP_TRY(P.vabc.upload(vabc, sizeof(double) * (size_t)n * npk));
    std::vector<S> xs = interleave<S>(u, v, (size_t)n * npk);
    ESPEC_HIP_TRY(P.cur.upload(xs.data(), sizeof(S) * xs.size()));
    ESPEC_HIP_TRY(P.prev.upload(xs.data(), sizeof(S) * xs.size()));
    ESPEC_HIP_TRY(P.psi0.upload(xs.data(), sizeof(S) * xs.size()));
    ESPEC_HIP_TRY(P.vpot.upload(vpot, sizeof(double) * (size_t)n * npk));
    ESPEC_HIP_TRY(P.cbuf.alloc(sizeof(double) * 3 * (size_t)P.maxsamp * npk));
    ESPEC_HIP_TRY(hipMemset(P.cbuf.p, 0, P.cbuf.bytes));
    std::vector<FftPlan> fp;
    std::vector<double> tw;
    std::vector<int> perm;
    axis_tables(P.ndim, P.np, P.ref, P.la.ax, fp, tw, perm, P.la.perm_off);
    ESPEC_HIP_TRY(P.tw.upload(tw.data(), sizeof(double) * tw.size()));
    if (P.lds) {
        std::vector<double> tc = kinetic_grid(P.ndim, P.np, sh, sm, P.ref);
        if (!P.ref) tc = scramble(P.ndim, P.np, fp, tc);
        ESPEC_HIP_TRY(P.tc.upload(tc.data(), sizeof(double) * tc.size()));
        ESPEC_HIP_TRY(P.perm.upload(perm.data(), sizeof(int) * perm.size()));
        SodArgs &a = P.la;
        a.ndim = P.ndim;
        a.n = n;
        a.npk = npk;
        a.tw = P.tw.as<double2>();
        a.perm = P.perm.as<int>();
        a.cur = P.cur.p;
        a.prev = P.prev.p;
        a.psi0 = P.psi0.p;
        a.vpot = P.vpot.as<double>();
        a.tc = P.tc.as<double>();
        a.b = 1.0 / n;
        a.lc = P.lc;
        a.maxsamp = P.maxsamp;
        a.cbuf = P.cbuf.as<double>();
        a.absorb = P.absorb;
        a.vabc = P.vabc.as<double>();
        return ESPEC_OK;
    }
    // Multi-pass tables: per axis T_a / n_a in digit-reversed order.
    std::vector<double> tax;
    std::vector<int> tax_off;
    for (int a = 0; a < P.ndim; ++a) {
        std::vector<double> t = kinetic_axis(P.np[a], sh[a], sm[a], false);
        std::vector<double> s(P.np[a]);
        for (int k = 0; k < P.np[a]; ++k) s[fp[a].perm[k]] = t[k] / P.np[a];
        tax_off.push_back((int)tax.size());
        tax.insert(tax.end(), s.begin(), s.end());
    }
    ESPEC_HIP_TRY(P.tax.upload(tax.data(), sizeof(double) * tax.size()));
    ESPEC_HIP_TRY(P.acc.alloc(sizeof(double2) * (size_t)n * npk));
    ESPEC_HIP_TRY(P.dstep.alloc(sizeof(long long)));
    int st = 1, maxblk = 1;
    P.passes.resize(P.ndim);
    for (int a = 0; a < P.ndim; ++a) {
        SodPassArgs &pa = P.passes[a];
        pa = SodPassArgs{};
        const int na = P.np[a];
        pa.n = n;
        pa.npk = npk;
        pa.axis = a;
        pa.na = na;
        pa.nlow = st;
        pa.nhigh = n / (st * na);
        if (a == 0) {
            pa.tile = std::max(1, std::min(pa.nhigh, 2048 / na));
            fill_axis(pa.ax, fp[a], 1, pa.tile, 1, 0, na, 0);
            P.pass_grid[a] = pa.nhigh / pa.tile;
        } else {
            pa.tile = std::max(1, std::min(std::min(16, st), 4096 / na));
            fill_axis(pa.ax, fp[a], pa.tile, pa.tile, pa.tile, 1, 0, 0);
            P.pass_grid[a] = (st / pa.tile) * pa.nhigh;
        }
        P.pass_lds[a] = sizeof(double2) * (size_t)pa.tile * na;
        // The twiddle table of axis a starts at its own offset.
        pa.tw = P.tw.as<double2>() + P.la.ax[a].tw0;
        pa.ta = P.tax.as<double>() + tax_off[a];
        pa.cur = P.cur.p;
        pa.prev = P.prev.p;
        pa.acc = P.acc.as<double2>();
        pa.psi0 = P.psi0.p;
        pa.vpot = P.vpot.as<double>();
        pa.first = a == 0;
        pa.last = a == P.ndim - 1;
        pa.lc = P.lc;
        pa.maxsamp = P.maxsamp;
        pa.nblk = P.pass_grid[a];
        pa.dstep = P.dstep.as<long long>();
        pa.cbuf = P.cbuf.as<double>();
        pa.absorb = P.absorb;
        pa.vabc = P.vabc.as<double>();
        maxblk = std::max(maxblk, P.pass_grid[a]);
        st *= na;
    }
    ESPEC_HIP_TRY(P.part.alloc(sizeof(double) * 3 * (size_t)maxblk * npk));
    ESPEC_HIP_TRY(P.ticket.alloc(sizeof(unsigned) * npk));
    ESPEC_HIP_TRY(hipMemset(P.ticket.p, 0, P.ticket.bytes));
    for (auto &pa : P.passes) {
        pa.part = P.part.as<double>();
        pa.ticket = P.ticket.as<unsigned>();
    }
    return ESPEC_OK;
}

}  // namespace

extern "C" {

typedef struct espec_sod_plan espec_sod_plan;

int espec_sod_plan_create(espec_sod_plan **out, int ndim, const int *np, int npk, double dt,
                          int lc, const double *sh, const double *sm, const double *vpot,
                          const double *u, const double *v, int maxsamp, int flags)
{
    return espec_sod_plan_create_absorb(out, ndim, np, npk, dt, lc, sh, sm, vpot, u, v, maxsamp,
                                        flags, ESPEC_ABSORB_NONE, nullptr);
}

int espec_sod_plan_create_absorb(espec_sod_plan **out, int ndim, const int *np, int npk,
                                 double dt, int lc, const double *sh, const double *sm,
                                 const double *vpot, const double *u, const double *v,
                                 int maxsamp, int flags, int absorb, const double *vabc)
{
    if (!out || ndim < 1 || ndim > 3 || !np || npk < 1 || lc < 1 || !sh || !sm || !vpot || !u ||
        maxsamp < 1 || absorb < ESPEC_ABSORB_NONE || absorb > ESPEC_ABSORB_SMOOTHW ||
        (absorb != ESPEC_ABSORB_NONE && !vabc))
        return ESPEC_ERR_ARG;
    *out = nullptr;
    for (int a = 0; a < ndim; ++a)
        if (!is_pow2(np[a]) || np[a] < 8 || np[a] > 4096) return ESPEC_ERR_ARG;
    std::unique_ptr<espec_sod_plan> P(new espec_sod_plan);
    P->ndim = ndim;
    for (int a = 0; a < ndim; ++a) P->np[a] = np[a];
    P->n = ntot(ndim, np);
    P->npk = npk;
    P->lc = lc;
    P->maxsamp = maxsamp;
    P->mixed = (flags & ESPEC_MIXED) != 0;
    P->ref = (flags & ESPEC_REF) != 0;
    // ESPEC_REF: the reference's .SMOOTHW branch is unreachable (sodfft.f:108).
    P->absorb = (P->ref && absorb == ESPEC_ABSORB_SMOOTHW) ? ESPEC_ABSORB_NONE : absorb;
    if (P->ref && P->mixed) return ESPEC_ERR_UNSUPPORTED;
    // One wavefront per packet up to 256 points; four above (a 1024-point
    // 1-D batch of 260 packets: 22.1 -> 11.1 us per step).
    P->nt = P->n <= 256 ? 64 : 256;
    P->lds = P->n <= kSodLdsMax;
    if (!P->lds && P->ref) return ESPEC_ERR_UNSUPPORTED;
    P->la.cst = 2.0 * dt / kShfs;  // psodfft.f:80
    ESPEC_HIP_TRY(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    int rc = P->mixed ? sod_build<float>(*P, sh, sm, vpot, u, v, vabc)
                      : sod_build<double>(*P, sh, sm, vpot, u, v, vabc);
    if (rc) return rc;
    ESPEC_HIP_TRY(launch_spo_sample(P->cur.p, P->psi0.p, P->n, npk, P->mixed,
                                    P->cbuf.as<double>(), maxsamp, 0, P->stream));
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    *out = P.release();
    return ESPEC_OK;
}

int espec_sod_plan_run(espec_sod_plan *P, int nsteps)
{
    if (!P || nsteps < 0) return ESPEC_ERR_ARG;
    if (nsteps == 0) return ESPEC_OK;
    if (P->lds) {
        const int chunk = std::max(1, (1 << 24) / P->n);
        int done = 0;
        while (done < nsteps) {
            const int k = std::min(chunk, nsteps - done);
            P->la.nsteps = k;
            P->la.step0 = P->steps_done + done;
            ESPEC_HIP_TRY(launch_sod_lds(*P, P->stream));
            done += k;
        }
    } else {
        int rc = P->mixed ? sod_run_pass<float>(*P, nsteps) : sod_run_pass<double>(*P, nsteps);
        if (rc) return rc;
    }
    P->steps_done += nsteps;
    return ESPEC_OK;
}

int espec_sod_plan_sync(espec_sod_plan *P)
{
    if (!P) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    return ESPEC_OK;
}

int espec_sod_plan_get(espec_sod_plan *P, int pk, double *cr, double *ci, double *nrm,
                       int maxsamp, int *nsamp, double *u, double *v)
{
    if (!P || pk < 0 || pk >= P->npk) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    const long long have = P->steps_done / P->lc + 1;
    const int ns = (int)std::min<long long>(have, std::min(maxsamp, P->maxsamp));
    if (nsamp) *nsamp = ns;
    if (ns > 0 && (cr || ci || nrm)) {
        std::vector<double> c((size_t)3 * ns);
        ESPEC_HIP_TRY(hipMemcpy(c.data(), P->cbuf.as<double>() + (size_t)3 * P->maxsamp * pk,
                                sizeof(double) * c.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < ns; ++i) {
            if (cr) cr[i] = c[3 * i];
            if (ci) ci[i] = c[3 * i + 1];
            if (nrm) nrm[i] = c[3 * i + 2];
        }
    }
    if (u || v) {
        const size_t n = P->n;
        if (P->mixed) {
            std::vector<float> h(2 * n);
            ESPEC_HIP_TRY(hipMemcpy(h.data(), P->cur.as<float>() + 2 * n * pk,
                                    sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) {
                if (u) u[i] = h[2 * i];
                if (v) v[i] = h[2 * i + 1];
            }
        } else {
            std::vector<double> h(2 * n);
            ESPEC_HIP_TRY(hipMemcpy(h.data(), P->cur.as<double>() + 2 * n * pk,
                                    sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) {
                if (u) u[i] = h[2 * i];
                if (v) v[i] = h[2 * i + 1];
            }
        }
    }
    return ESPEC_OK;
}

int espec_sod_plan_get_prev(espec_sod_plan *P, int pk, double *u, double *v)
{
    if (!P || pk < 0 || pk >= P->npk || !u || !v) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    const size_t n = P->n;
    if (P->mixed) {
        std::vector<float> h(2 * n);
        ESPEC_HIP_TRY(hipMemcpy(h.data(), P->prev.as<float>() + 2 * n * pk, sizeof(float) * 2 * n,
                                hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            u[i] = h[2 * i];
            v[i] = h[2 * i + 1];
        }
    } else {
        std::vector<double> h(2 * n);
        ESPEC_HIP_TRY(hipMemcpy(h.data(), P->prev.as<double>() + 2 * n * pk,
                                sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            u[i] = h[2 * i];
            v[i] = h[2 * i + 1];
        }
    }
    return ESPEC_OK;
}

int espec_sod_plan_destroy(espec_sod_plan *P)
{
    if (P) {
        (void)hipStreamSynchronize(P->stream);
        delete P;
    }
    return ESPEC_OK;
}

int espec_psodfft(int ndim, const int *np, double dt, double ti, double tf, int lc,
                  const double *sh, const double *sm, double *u1, double *v1, const double *vpot,
                  double *cr, double *ci, double *nrm, int maxsamp, int *nsamp, int flags)
{
    return espec_psodfft_absorb(ndim, np, dt, ti, tf, lc, sh, sm, u1, v1, vpot, cr, ci, nrm,
                                maxsamp, nsamp, flags, ESPEC_ABSORB_NONE, nullptr);
}

int espec_psodfft_absorb(int ndim, const int *np, double dt, double ti, double tf, int lc,
                         const double *sh, const double *sm, double *u1, double *v1,
                         const double *vpot, double *cr, double *ci, double *nrm, int maxsamp,
                         int *nsamp, int flags, int absorb, const double *vabc)
{
    if (!u1 || !v1 || !cr || !ci || maxsamp < 1) return ESPEC_ERR_ARG;
    // The start-up covers TI+DT; the leapfrog loop is DO T=TI+2DT,TF+DT,DT
    // (psodfft.f:131).
    const int nsteps = 1 + fortran_trip(ti + 2 * dt, tf + dt, dt);
    espec_sod_plan *P = nullptr;
    int rc = espec_sod_plan_create_absorb(&P, ndim, np, 1, dt, lc, sh, sm, vpot, u1, v1, maxsamp,
                                          flags, absorb, vabc);
    if (rc) return rc;
    rc = espec_sod_plan_run(P, nsteps);
    if (!rc) rc = espec_sod_plan_get(P, 0, cr, ci, nrm, maxsamp, nsamp, u1, v1);
    espec_sod_plan_destroy(P);
    return rc;
}

}  // extern "C"
