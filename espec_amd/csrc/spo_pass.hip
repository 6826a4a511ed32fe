// Multi-pass split-operator step for grids larger than one workgroup's LDS.
//
// Per step (2-D shown; 3-D adds a z pass like the y pass):
//   row pass  : [inverse x stages of step k]  B e^{-iV dt/2}  (C(t) sample)
//               e^{-iV dt/2}  [forward x stages of step k+1]  e^{-iTx dt}
//               -> state stored with x digit-reversed, y natural
//   y pass    : forward y, e^{-iTy dt}, inverse y on tiles of contiguous
//               x positions (coalesced 128-byte rows), y natural again
// Since e^{-i(Tx+Ty)dt} = e^{-iTx dt} e^{-iTy dt}, the kinetic phase splits
// per axis and each pass keeps one axis of the transform.  The state makes
// one round trip through L2/MALL per pass; the V phases ride on the row
// pass and the kinetic phases on the butterflies.  C(t) partial sums are
// written per row block and reduced after the run (deterministic order).
#include <algorithm>

#include "api_common.hpp"
#include "fft_lds.hpp"
#include "spo_pass.hpp"

namespace espec {

struct PassArgs {
    int n, npk, np0, xblocks;
    DevAxis axx;          // row block: lines = rows, es = 1
    DevAxis axs;          // strided tile: lines = tile columns, es = tile
    const double2 *tw;
    const double2 *kx;    // x kinetic phases (digit-reversed order)
    const double2 *ks;    // strided-axis kinetic phases
    void *x;              // state, storage precision
    const void *psi0;     // initial state, storage precision
    const void *vpot;     // potential, storage precision
    double *part;
    long long *dstep;
    double b, cst2;
    int lc, maxsamp;
    int S, L, ltile, tiles, last;
};

// S: storage precision of psi, psi0 and V in HBM; LDS and arithmetic FP64.
template <typename S, int NT, bool INV, bool FWD>
__global__ __launch_bounds__(NT) void pass_rows(PassArgs a)
{
    using V = double2;
    using VS = V2<S>;
    extern __shared__ double smem_d[];
    V *x = reinterpret_cast<V *>(smem_d);
    const int cnt = a.axx.nl * a.np0;
    double *red = reinterpret_cast<double *>(x + cnt);
    const int pk = blockIdx.y, blk = blockIdx.x;
    const long long po = (long long)pk * a.n + (long long)blk * cnt;
    VS *gx = reinterpret_cast<VS *>(a.x) + po;
    const S *__restrict__ gv = reinterpret_cast<const S *>(a.vpot) + po;
    const VS *__restrict__ g0 = reinterpret_cast<const VS *>(a.psi0) + po;
    const double b = a.b;

    for (int i = threadIdx.x; i < cnt; i += NT) {
        VS t = gx[i];
        x[i] = {(double)t.x, (double)t.y};
    }
    __syncthreads();
    const long long g = INV ? *a.dstep : 0;  // the step this pass completes
    const bool samp = INV && (g % a.lc) == 0;
    if (INV)
        for (int s = a.axx.nst - 1; s >= 1; --s) dit_stage<double, NT, -1>(x, a.axx, s, a.tw);
    double cr = 0, ci = 0, nr = 0;
    fused_outer<double, NT, INV, FWD>(x, a.axx, a.tw, [&](int i, V &v) {
        double sn, cs;
        sincos(a.cst2 * (double)gv[i], &sn, &cs);
        const V w = {cs, sn};
        if (INV) {
            v = cmulc(v, w);
            v.x *= b;
            v.y *= b;
            if (samp) {
                VS p = g0[i];
                cr += (double)p.x * v.x + (double)p.y * v.y;
                ci += (double)p.x * v.y - (double)p.y * v.x;
                nr += v.x * v.x + v.y * v.y;
            }
        }
        if (FWD) v = cmulc(v, w);
    });
    if (FWD) {
        for (int s = 1; s < a.axx.nst; ++s) dif_stage<double, NT, 1>(x, a.axx, s, a.tw);
        for (int i = threadIdx.x; i < cnt; i += NT) x[i] = cmulc(x[i], a.kx[i & (a.np0 - 1)]);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < cnt; i += NT) gx[i] = {(S)x[i].x, (S)x[i].y};
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
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void pass_strided(PassArgs a)
{
    using V = double2;
    using VS = V2<S>;
    extern __shared__ double smem_d[];
    V *x = reinterpret_cast<V *>(smem_d);
    const int tile = 1 << a.ltile;
    const int cnt = tile * a.L;
    const int pk = blockIdx.y, t = blockIdx.x;
    const int spt = a.S >> a.ltile;  // tiles across the lower axes
    const long long base =
        (long long)pk * a.n + (long long)(t / spt) * a.S * a.L + (long long)(t % spt) * tile;
    VS *gx = reinterpret_cast<VS *>(a.x) + base;

    for (int i = threadIdx.x; i < cnt; i += NT) {
        VS v = gx[(long long)(i >> a.ltile) * a.S + (i & (tile - 1))];
        x[i] = {(double)v.x, (double)v.y};
    }
    __syncthreads();
    for (int s = 0; s < a.axs.nst - 1; ++s) dif_stage<double, NT, 1>(x, a.axs, s, a.tw);
    fused_inner<double, NT>(x, a.axs, [&](int i, V &v) { v = cmulc(v, a.ks[i >> a.ltile]); });
    for (int s = a.axs.nst - 2; s >= 0; --s) dit_stage<double, NT, -1>(x, a.axs, s, a.tw);
    for (int i = threadIdx.x; i < cnt; i += NT)
        gx[(long long)(i >> a.ltile) * a.S + (i & (tile - 1))] = {(S)x[i].x, (S)x[i].y};
    if (a.last && t == 0 && pk == 0 && threadIdx.x == 0) *a.dstep += 1;
}

__global__ void set_step(long long *d, long long v) { *d = v; }

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

void axis_from_plan(DevAxis &d, const FftPlan &p, int es, int nl, int lmod, int la, int lb, int tw0)
{
    d.n = p.n;
    d.ln = ilog2(p.n);
    d.nst = p.nst;
    d.es = es;
    d.nl = nl;
    d.lnl = ilog2(nl);
    d.lmod = lmod;
    d.la = la;
    d.lb = lb;
    d.tw0 = tw0;
    for (int s = 0; s < p.nst; ++s) {
        d.rad[s] = p.rad[s];
        d.lm[s] = p.lm[s];
        d.lL[s] = p.lL[s];
        d.off[s] = p.off[s];
    }
}

template <typename K>
hipError_t launch(K kern, dim3 grid, int nt, size_t lds, hipStream_t s, const PassArgs &a)
{
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, grid, dim3(nt), lds, s, a);
    return hipGetLastError();
}

template <typename R>
hipError_t rows(const SpoPassPlan &p, const PassArgs &a, bool inv, bool fwd, hipStream_t s)
{
    const int nt = 64 * p.rows_per_wg;
    size_t lds = sizeof(double2) * p.rows_per_wg * p.np[0] + sizeof(double) * 3 * (nt / 64);
    dim3 grid(p.xblocks, p.npk);
#define ESPEC_ROWS(NT_)                                                                  \
    if (nt == NT_) {                                                                     \
        if (inv && fwd) return launch(pass_rows<R, NT_, true, true>, grid, NT_, lds, s, a); \
        if (inv) return launch(pass_rows<R, NT_, true, false>, grid, NT_, lds, s, a);     \
        return launch(pass_rows<R, NT_, false, true>, grid, NT_, lds, s, a);              \
    }
    ESPEC_ROWS(64)
    ESPEC_ROWS(128)
    ESPEC_ROWS(256)
#undef ESPEC_ROWS
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t strided(const SpoPassPlan &p, PassArgs a, int axis, bool last, hipStream_t s)
{
    a.axs = p.axs[axis];
    a.ks = static_cast<const double2 *>(p.kax) + p.kax_off[axis];
    a.S = 1;
    for (int k = 0; k < axis; ++k) a.S *= p.np[k];
    a.L = p.np[axis];
    a.ltile = ilog2(p.tile);
    a.tiles = p.n / (p.tile * a.L);
    a.last = last ? 1 : 0;
    size_t lds = sizeof(double2) * p.tile * a.L;
    return launch(pass_strided<R, 256>, dim3(a.tiles, p.npk), 256, lds, s, a);
}

template <typename R>
PassArgs base_args(const SpoPassPlan &p)
{
    PassArgs a{};
    a.n = p.n;
    a.npk = p.npk;
    a.np0 = p.np[0];
    a.xblocks = p.xblocks;
    a.axx = p.axx;
    a.tw = static_cast<const double2 *>(p.tw);
    a.kx = static_cast<const double2 *>(p.kax) + p.kax_off[0];
    a.x = p.x;
    a.psi0 = p.psi0;
    a.vpot = p.vpot;
    a.cst2 = p.cst2;
    a.part = p.part;
    a.dstep = p.dstep;
    a.b = p.b;
    a.lc = p.lc;
    a.maxsamp = p.maxsamp;
    return a;
}

template <typename R>
hipError_t one_step(const SpoPassPlan &p, bool inv, bool fwd, hipStream_t s, int *launches)
{
    PassArgs a = base_args<R>(p);
    hipError_t e = rows<R>(p, a, inv, fwd, s);
    ++*launches;
    if (e != hipSuccess || !fwd) return e;
    for (int ax = 1; ax < p.ndim; ++ax) {
        e = strided<R>(p, a, ax, ax == p.ndim - 1, s);
        ++*launches;
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename R>
int run_typed(SpoPassPlan &p, int nsteps, long long step0, hipStream_t s, int *launches)
{
    hipLaunchKernelGGL(set_step, dim3(1), dim3(1), 0, s, p.dstep, step0);
    ESPEC_HIP_TRY(hipGetLastError());
    ++*launches;
    // Step 1 of the run: V half phase + forward x, then the other axes.
    ESPEC_HIP_TRY(one_step<R>(p, false, true, s, launches));
    // Steps 2..nsteps: finish the previous step and start the next.  The
    // same launch sequence repeats, so it is captured once into a graph of
    // graph_steps steps and replayed.
    int left = nsteps - 1;
    const int gsteps = 32;
    if (left >= gsteps) {
        if (!p.gexec) {
            hipStream_t cs;
            ESPEC_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            int dummy = 0;
            hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            for (int k = 0; e == hipSuccess && k < gsteps; ++k) e = one_step<R>(p, true, true, cs, &dummy);
            hipError_t e2 = hipStreamEndCapture(cs, &p.graph);
            (void)hipStreamDestroy(cs);
            ESPEC_HIP_TRY(e);
            ESPEC_HIP_TRY(e2);
            ESPEC_HIP_TRY(hipGraphInstantiate(&p.gexec, p.graph, nullptr, nullptr, 0));
            p.graph_steps = gsteps;
        }
        while (left >= p.graph_steps) {
            ESPEC_HIP_TRY(hipGraphLaunch(p.gexec, s));
            *launches += p.graph_steps * p.ndim;
            left -= p.graph_steps;
        }
    }
    for (; left > 0; --left) ESPEC_HIP_TRY(one_step<R>(p, true, true, s, launches));
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

}  // namespace

int spo_pass_setup(SpoPassPlan &p, int ndim, const int *np, int npk, bool mixed,
                   const std::vector<FftPlan> &fp, const std::vector<std::vector<double>> &taxes,
                   double cst1, double cst2, void *x, const void *psi0, const void *vpot,
                   double *cbuf, int maxsamp)
{
    p.cst2 = cst2;
    p.ndim = ndim;
    p.n = 1;
    for (int a = 0; a < ndim; ++a) {
        p.np[a] = np[a];
        p.n *= np[a];
    }
    p.npk = npk;
    p.mixed = mixed;
    p.maxsamp = maxsamp;
    p.x = x;
    p.psi0 = psi0;
    p.vpot = vpot;
    p.cbuf = cbuf;
    p.b = 1.0 / p.n;
    p.tile = mixed ? 16 : 8;  // 128-byte rows of the strided tiles
    if (p.tile > np[0]) p.tile = np[0];  // a tile never spans past one x row
    p.rows_per_wg = 1;
    p.xblocks = p.n / (p.np[0] * p.rows_per_wg);
    // Tables: twiddles per axis and per-axis kinetic phases (cos, sin) of
    // cst1 * T_a(k_a) stored at the digit-reversed position of k_a.
    std::vector<double> tw, kax;
    for (int a = 0; a < ndim; ++a) {
        int tw0 = (int)(tw.size() / 2);
        for (size_t i = 0; i < fp[a].twc.size(); ++i) {
            tw.push_back(fp[a].twc[i]);
            tw.push_back(fp[a].tws[i]);
        }
        if (a == 0)
            axis_from_plan(p.axx, fp[0], 1, p.rows_per_wg, p.rows_per_wg, p.np[0], 0, tw0);
        else
            axis_from_plan(p.axs[a], fp[a], p.tile, p.tile, p.tile, 1, 0, tw0);
        p.kax_off[a] = (int)(kax.size() / 2);
        size_t k0 = kax.size();
        kax.resize(k0 + 2 * (size_t)np[a]);
        for (int k = 0; k < np[a]; ++k) {
            double th = cst1 * taxes[a][k];
            int pos = fp[a].perm[k];
            kax[k0 + 2 * pos] = std::cos(th);
            kax[k0 + 2 * pos + 1] = std::sin(th);
        }
    }
    DevBuf btw, bk, bpart, bstep;
    ESPEC_HIP_TRY(btw.upload(tw.data(), sizeof(double) * tw.size()));
    ESPEC_HIP_TRY(bk.upload(kax.data(), sizeof(double) * kax.size()));
    ESPEC_HIP_TRY(bpart.alloc(sizeof(double) * 3 * (size_t)maxsamp * npk * p.xblocks));
    ESPEC_HIP_TRY(hipMemset(bpart.p, 0, bpart.bytes));
    ESPEC_HIP_TRY(bstep.alloc(sizeof(long long)));
    p.tw = btw.p;
    btw.p = nullptr;
    p.kax = bk.p;
    bk.p = nullptr;
    p.part = static_cast<double *>(bpart.p);
    bpart.p = nullptr;
    p.dstep = static_cast<long long *>(bstep.p);
    bstep.p = nullptr;
    return ESPEC_OK;
}

int spo_pass_run(SpoPassPlan &p, int nsteps, long long step0, int lc, hipStream_t s, int *launches)
{
    p.lc = lc;
    *launches = 0;
    if (nsteps <= 0) return ESPEC_OK;
    return p.mixed ? run_typed<float>(p, nsteps, step0, s, launches)
                   : run_typed<double>(p, nsteps, step0, s, launches);
}

void spo_pass_release(SpoPassPlan &p)
{
    if (p.gexec) (void)hipGraphExecDestroy(p.gexec);
    if (p.graph) (void)hipGraphDestroy(p.graph);
    if (p.tw) (void)hipFree(p.tw);
    if (p.kax) (void)hipFree(p.kax);
    if (p.part) (void)hipFree(p.part);
    if (p.dstep) (void)hipFree(p.dstep);
    p = SpoPassPlan{};
}

}  // namespace espec
