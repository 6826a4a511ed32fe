// LDS-resident split-operator kernels: one workgroup owns one packet and
// keeps its whole grid in LDS for every step of a launch, so HBM sees the
// state once per launch, not once per step.
//
// Step (spofft.f:112-191):  psi <- B e^{-iV dt/2} F^-1 e^{-iT dt} F e^{-iV dt/2} psi
//
// spo_lds_fast (ESPEC_EXACT, FP64 or mixed): forward DIF stages leave the
// spectrum digit-reversed, the kinetic phase is applied in that order and
// the inverse DIT stages restore natural order, so no permutation pass is
// needed.  The V phases are fused into the outermost butterfly stage (the
// inverse stage that ends step k and the forward stage that starts step
// k+1 touch the same elements), the T phase into the innermost one.
//
// Absorbers: .VOPTIC's damping exp(cst2 W) is folded into the V phase
// table (both half phases); .SMOOTHW's mask multiplies psi after the
// closing half phase (spofft.f:162-176), before the C(t) sample.
//
// spo_lds_ref (ESPEC_REF, FP64): the reference's order of operations,
// including the half-swap of the flattened array around the V phases
// (spofft.f:120-133,178-191), natural-order FFTs per axis with the f32
// twiddles (fft3d.f:119) and the non-unitary kinetic update
// (spofft.f:148-154), so that results match the reference build.
#include "spo_lds.hpp"

#include "fft_lds.hpp"

namespace espec {

// S is the storage precision of psi and psi0 in HBM (float for ESPEC_MIXED,
// double for ESPEC_FP64); LDS, twiddles, phase tables and all arithmetic
// are FP64, so the mixed path rounds psi to complex64 once per launch.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void spo_lds_fast(SpoLdsArgs a)
{
    using V = double2;
    using VS = V2<S>;
    extern __shared__ double smem_d[];
    // LDS: psi | V phases (vph_lds) | kinetic phases, twiddles (tab_lds) | reduction
    V *x = reinterpret_cast<V *>(smem_d);  // padded (lx<true>)
    V *vl = x + lds_slots(a.n, true);
    V *kl = vl + (a.vph_lds ? a.n : 0);
    V *twl = kl + (a.tab_lds ? a.n : 0);
    double *red = reinterpret_cast<double *>(twl + (a.tab_lds ? a.ntw : 0));
    const int pk = blockIdx.x;
    const long long po = (long long)pk * a.n;
    VS *gx = reinterpret_cast<VS *>(a.x) + po;
    const VS *__restrict__ g0 = reinterpret_cast<const VS *>(a.psi0) + po;
    const V *__restrict__ gvg = reinterpret_cast<const V *>(a.vph) + po;
    const V *__restrict__ gkg = reinterpret_cast<const V *>(a.kin) + (a.kin_pp ? po : 0);
    const V *__restrict__ twg = reinterpret_cast<const V *>(a.tw);
    const double *__restrict__ mk = a.mask ? a.mask + po : nullptr;
    const int n = a.n;
    const double b = a.b;
    const DevAxis &ax0 = a.ax[0];
    const DevAxis &axl = a.ax[a.ndim - 1];

    for (int i = threadIdx.x; i < n; i += NT) {
        VS t = gx[i];
        x[lx<true>(i)] = {(double)t.x, (double)t.y};
        if (a.vph_lds) vl[i] = gvg[i];
        if (a.tab_lds) kl[i] = gkg[i];
    }
    if (a.tab_lds)
        for (int j = threadIdx.x; j < a.ntw; j += NT) twl[j] = twg[j];
    __syncthreads();
    const V *gv = a.vph_lds ? vl : gvg;
    const V *gk = a.tab_lds ? kl : gkg;
    const V *tw = a.tab_lds ? twl : twg;

    fused_outer<double, NT, false, true, true>(x, ax0, tw, [&](int i, V &v) { v = cmulc(v, gv[i]); });
    for (int st = 0; st < a.nsteps; ++st) {
        for (int d = 0; d < a.ndim; ++d) {
            const DevAxis &ax = a.ax[d];
            const int s0 = d == 0 ? 1 : 0;
            const int s1 = d == a.ndim - 1 ? ax.nst - 1 : ax.nst;
            for (int s = s0; s < s1; ++s) dif_stage<double, NT, 1, true>(x, ax, s, tw);
        }
        fused_inner<double, NT, true>(x, axl, [&](int i, V &v) { v = cmulc(v, gk[i]); });
        for (int d = a.ndim - 1; d >= 0; --d) {
            const DevAxis &ax = a.ax[d];
            const int s0 = d == 0 ? 1 : 0;
            const int s1 = d == a.ndim - 1 ? ax.nst - 1 : ax.nst;
            for (int s = s1 - 1; s >= s0; --s) dit_stage<double, NT, -1, true>(x, ax, s, tw);
        }
        const long long g = a.step0 + st + 1;
        const bool samp = (g % a.lc) == 0;
        double cr = 0, ci = 0, nr = 0;
        auto post = [&](int i, V &v) {
            V w = gv[i];
            v = cmulc(v, w);
            v.x *= b;
            v.y *= b;
            if (mk) {  // .SMOOTHW (spofft.f:171-174)
                v.x *= mk[i];
                v.y *= mk[i];
            }
            if (samp) {
                VS p = g0[i];
                cr += (double)p.x * v.x + (double)p.y * v.y;
                ci += (double)p.x * v.y - (double)p.y * v.x;
                nr += v.x * v.x + v.y * v.y;
            }
            return w;
        };
        if (st + 1 < a.nsteps)
            fused_outer<double, NT, true, true, true>(x, ax0, tw, [&](int i, V &v) {
                V w = post(i, v);
                v = cmulc(v, w);
            });
        else
            fused_outer<double, NT, true, false, true>(x, ax0, tw, [&](int i, V &v) { post(i, v); });
        if (samp) {
            block_sum3<NT>(cr, ci, nr, red);
            const long long nc = g / a.lc;
            if (threadIdx.x == 0 && nc < a.maxsamp) {
                double *c = a.cbuf + ((long long)pk * a.maxsamp + nc) * 3;
                c[0] = cr;
                c[1] = ci;
                c[2] = nr;
            }
        }
    }
    for (int i = threadIdx.x; i < n; i += NT) {
        const V t = x[lx<true>(i)];
        gx[i] = {(S)t.x, (S)t.y};
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void spo_lds_ref(SpoLdsArgs a)
{
    using V = double2;
    extern __shared__ double smem_d[];
    V *x = reinterpret_cast<V *>(smem_d);
    V *y = x + a.n;
    double *red = reinterpret_cast<double *>(y + a.n);
    const int pk = blockIdx.x;
    const long long po = (long long)pk * a.n;
    V *gx = reinterpret_cast<V *>(a.x) + po;
    const V *__restrict__ g0 = reinterpret_cast<const V *>(a.psi0) + po;
    const V *__restrict__ gv = reinterpret_cast<const V *>(a.vph) + po;
    const V *__restrict__ gk = reinterpret_cast<const V *>(a.kin) + (a.kin_pp ? po : 0);
    const V *__restrict__ tw = reinterpret_cast<const V *>(a.tw);
    const double *__restrict__ mk = a.mask ? a.mask + po : nullptr;
    const int n = a.n, nh = n / 2;
    const double b = a.b;

    for (int i = threadIdx.x; i < n; i += NT) x[i] = gx[i];
    __syncthreads();
    for (int st = 0; st < a.nsteps; ++st) {
        // spofft.f:120-133: swap halves and apply exp(-i cst2 V).
        for (int i = threadIdx.x; i < nh; i += NT) {
            V p = x[i], q = x[i + nh];
            x[i] = cmulc(q, gv[i + nh]);
            x[i + nh] = cmulc(p, gv[i]);
        }
        __syncthreads();
        for (int d = 0; d < a.ndim; ++d) {  // spofft.f:142-146, ISN=+1
            const DevAxis &ax = a.ax[d];
            for (int s = 0; s < ax.nst; ++s) dif_stage<double, NT, 1>(x, ax, s, tw);
            permute_axis<double, NT>(x, y, ax, a.perm + a.perm_off[d]);
            V *t = x; x = y; y = t;
        }
        // spofft.f:148-154: V(I) is formed from the already updated U(I).
        for (int i = threadIdx.x; i < n; i += NT) {
            V v = x[i], w = gk[i];
            double u1 = w.x * v.x + w.y * v.y;
            double v1 = w.x * v.y - w.y * u1;
            x[i] = {u1, v1};
        }
        __syncthreads();
        for (int d = 0; d < a.ndim; ++d) {  // spofft.f:156-160, ISN=-1
            const DevAxis &ax = a.ax[d];
            for (int s = 0; s < ax.nst; ++s) dif_stage<double, NT, -1>(x, ax, s, tw);
            permute_axis<double, NT>(x, y, ax, a.perm + a.perm_off[d]);
            V *t = x; x = y; y = t;
        }
        // spofft.f:178-191: swap back, exp(-i cst2 V) and the 1/N scale.
        for (int i = threadIdx.x; i < nh; i += NT) {
            V p = x[i], q = x[i + nh];
            V w1 = gv[i + nh], w2 = gv[i];
            double c1 = w1.x * b, s1 = w1.y * b, c2 = w2.x * b, s2 = w2.y * b;
            V r1 = {c1 * p.x + s1 * p.y, c1 * p.y - s1 * p.x};
            V r2 = {c2 * q.x + s2 * q.y, c2 * q.y - s2 * q.x};
            if (mk) {  // .SMOOTHW: the mask of the stored point (spofft.f:171-174)
                r1 = {mk[i + nh] * r1.x, mk[i + nh] * r1.y};
                r2 = {mk[i] * r2.x, mk[i] * r2.y};
            }
            x[i + nh] = r1;
            x[i] = r2;
        }
        __syncthreads();
        const long long g = a.step0 + st + 1;
        if ((g % a.lc) == 0) {
            double cr = 0, ci = 0, nr = 0;
            for (int i = threadIdx.x; i < n; i += NT) {
                V p = g0[i], v = x[i];
                cr += p.x * v.x + p.y * v.y;
                ci += p.x * v.y - p.y * v.x;
                nr += v.x * v.x + v.y * v.y;
            }
            block_sum3<NT>(cr, ci, nr, red);
            const long long nc = g / a.lc;
            if (threadIdx.x == 0 && nc < a.maxsamp) {
                double *c = a.cbuf + ((long long)pk * a.maxsamp + nc) * 3;
                c[0] = cr;
                c[1] = ci;
                c[2] = nr;
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += NT) gx[i] = x[i];
}

// <psi0|psi> and |psi|^2 of every packet into sample slot idx.
template <typename R, int NT>
__global__ __launch_bounds__(NT) void spo_sample(const V2<R> *x, const V2<R> *psi0, int n,
                                                 double *cbuf, int maxsamp, int idx)
{
    __shared__ double red[3 * (NT / 64)];
    const long long po = (long long)blockIdx.x * n;
    double cr = 0, ci = 0, nr = 0;
    for (int i = threadIdx.x; i < n; i += NT) {
        V2<R> p = psi0[po + i], v = x[po + i];
        cr += (double)p.x * v.x + (double)p.y * v.y;
        ci += (double)p.x * v.y - (double)p.y * v.x;
        nr += (double)v.x * v.x + (double)v.y * v.y;
    }
    block_sum3<NT>(cr, ci, nr, red);
    if (threadIdx.x == 0 && idx < maxsamp) {
        double *c = cbuf + ((long long)blockIdx.x * maxsamp + idx) * 3;
        c[0] = cr;
        c[1] = ci;
        c[2] = nr;
    }
}

size_t spo_lds_bytes(int n, bool ref, bool vph_lds, int nt, bool tab_lds, int ntw)
{
    // fast kernel: padded state (lds_slots), V phases, kinetic phases + twiddles
    if (ref) return sizeof(double2) * 2 * (size_t)n + sizeof(double) * 3 * (nt / 64);
    size_t v = lds_slots(n, true) + (vph_lds ? (size_t)n : 0);
    if (tab_lds) v += (size_t)n + ntw;
    return sizeof(double2) * v + sizeof(double) * 3 * (nt / 64);
}

template <typename K>
static hipError_t launch_lds(K kern, int grid, int nt, size_t lds, hipStream_t s, const SpoLdsArgs &a)
{
    hipError_t e = hipSuccess;
    if (lds > 64 * 1024)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(nt), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_spo_lds(const SpoLdsArgs &a, bool mixed, bool ref, int nt, hipStream_t s)
{
    size_t lds = spo_lds_bytes(a.n, ref, !ref && a.vph_lds, nt, a.tab_lds, a.ntw);
    if (ref) {
        if (mixed) return hipErrorInvalidValue;
        return nt == 64 ? launch_lds(spo_lds_ref<64>, a.npk, 64, lds, s, a)
                        : launch_lds(spo_lds_ref<256>, a.npk, 256, lds, s, a);
    }
    if (mixed)
        return nt == 64 ? launch_lds(spo_lds_fast<float, 64>, a.npk, 64, lds, s, a)
                        : launch_lds(spo_lds_fast<float, 256>, a.npk, 256, lds, s, a);
    return nt == 64 ? launch_lds(spo_lds_fast<double, 64>, a.npk, 64, lds, s, a)
                    : launch_lds(spo_lds_fast<double, 256>, a.npk, 256, lds, s, a);
}

hipError_t launch_spo_sample(const void *x, const void *psi0, int n, int npk, bool mixed,
                             double *cbuf, int maxsamp, int idx, hipStream_t s)
{
    if (mixed)
        hipLaunchKernelGGL((spo_sample<float, 256>), dim3(npk), dim3(256), 0, s,
                           (const float2 *)x, (const float2 *)psi0, n, cbuf, maxsamp, idx);
    else
        hipLaunchKernelGGL((spo_sample<double, 256>), dim3(npk), dim3(256), 0, s,
                           (const double2 *)x, (const double2 *)psi0, n, cbuf, maxsamp, idx);
    return hipGetLastError();
}

}  // namespace espec
