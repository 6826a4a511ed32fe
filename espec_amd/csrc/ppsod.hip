// Pulse-driven second-order differencing with the finite-difference
// Hamiltonian: PPSOD (ppsod.f:1-548) and its step PSODAUX (sod.f:3-122).
//
// One step is psi(t+dt) = psi(t-dt) - i cst H psi(t) with H the 4th-order
// FD kinetic operator of AU (au1d.f, au2di.f, au3d.f) plus
// VEFF = VP - DM E(T) (chpot.f:58-62).  The two time levels live in two
// device buffers; a step reads psi(t) from one and overwrites psi(t-dt) in
// the other with psi(t+dt), point by point (point i reads psi(t-dt) only
// at i, so the update is in place), and the roles swap.  VEFF is formed
// inside the step from VP, DM and the step's field value, so it costs no
// extra pass.  The time loop is the reference's: CST = 2 DT/SHFS, a
// quarter- and a half-strength start-up step at the fields of (T+DT)/2 and
// T+DT (ppsod.f:246-254), psi(t-dt) reset to psi0, then full steps with
// the field of each step's end time.  Arithmetic is FP64 throughout, as in
// the reference; the order of each update follows sod.f and contraction
// into FMA is off, so one step matches the reference's to rounding of the
// stencil sums.
//
// Samples (C(t) = <psi0|psi>, <psi|psi>, the energies with VEFF and with VP
// as EIGENERG forms them, and up to six projections onto caller vectors,
// ppsod.f:276-330) are block partial sums per sample, reduced in a fixed
// order after the run, so results are deterministic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "api_common.hpp"
#include "common.hpp"
#include "tables.hpp"

namespace espec {

constexpr int kPpsodProj = 6;               // projection vectors (W1..W6, ppsod.f:88-95)
constexpr int kPpsodVals = 5 + 2 * kPpsodProj;  // cr, ci, norm, et, et0, proj re/im
constexpr int kPpsodBlocks = 256;

struct FdGrid {
    int ndim, n, np[3], st[3];
    double shm[3];
};

// (H u)_i of AU with potential value vi at point i: per axis
// shm (u[-2] - 16 u[-1] + 30 u - 16 u[+1] + u[+2]) with neighbours outside
// the grid taken as zero (au1d.f:55-73).
__device__ __forceinline__ double fd_h(const FdGrid &g, const double2 *u, bool im, double vi,
                                       int i)
{
#pragma clang fp contract(off)
    auto at = [&](int k) { return im ? u[k].y : u[k].x; };
    const double ui = at(i);
    double acc = vi * ui;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= g.ndim) break;
        const int s = g.st[d], np = g.np[d];
        const int ia = (i / s) % np;
        double t = 30.0 * ui;
        if (ia >= 2) t += at(i - 2 * s);
        if (ia >= 1) t -= 16.0 * at(i - s);
        if (ia + 1 < np) t -= 16.0 * at(i + s);
        if (ia + 2 < np) t += at(i + 2 * s);
        acc += g.shm[d] * t;
    }
    return acc;
}

// PSODAUX (sod.f:3-122) on interleaved states: prev <- psi(t+dt) from
// prev = psi(t-dt) and cur = psi(t).
__global__ __launch_bounds__(256) void fdsod_step(FdGrid g, double2 *prev, const double2 *cur,
                                                  const double *vp, const double *dm, double e,
                                                  double cst, int absorb, const double *vabc)
{
#pragma clang fp contract(off)
    for (int i = blockIdx.x * 256 + threadIdx.x; i < g.n; i += gridDim.x * 256) {
        const double f = dm[i] * e;
        const double ve = vp[i] - f;  // chpot.f:58-62
        const double hv = fd_h(g, cur, true, ve, i);
        const double hu = fd_h(g, cur, false, ve, i);
        const double2 a = prev[i], b = cur[i];
        double un, vn;
        if (absorb == ESPEC_ABSORB_VOPTIC) {
            const double cw = cst * vabc[i];
            un = a.x + cst * hv - cw * b.x;
            vn = a.y - cst * hu - cw * b.y;
        } else if (absorb == ESPEC_ABSORB_SMOOTHW) {
            un = vabc[i] * (a.x + cst * hv);
            vn = vabc[i] * (a.y - cst * hu);
        } else {
            un = a.x + cst * hv;
            vn = a.y - cst * hu;
        }
        prev[i] = {un, vn};
    }
}

// S2PPSOD's step (s2ppsod.f:260-458, 460-656, 762-975 with TPCRS =
// '.CORD'): PSODAUX on surface A (points [0, n)) and on surface B (the
// next n), then the coordinate coupling with w = w0 XD_i added to the new
// values from the other surface's psi(t):
//   UA += w VB(t), VA -= w UB(t), UB += w VA(t), VB -= w UA(t),
// each sum rounded after PSODAUX's, as the reference's second loop does.
__global__ __launch_bounds__(256) void fdsod2_step(FdGrid g, double2 *prev, const double2 *cur,
                                                   const double *vp, const double *dm, double e,
                                                   double cst, double w0, const double *xd,
                                                   int absorb, const double *vabc)
{
#pragma clang fp contract(off)
    const int n = g.n;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double2 out[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int o = k * n;
            const double f = dm[o + i] * e;
            const double ve = vp[o + i] - f;
            const double hv = fd_h(g, cur + o, true, ve, i);
            const double hu = fd_h(g, cur + o, false, ve, i);
            const double2 a = prev[o + i], b = cur[o + i];
            double un, vn;
            if (absorb == ESPEC_ABSORB_VOPTIC) {
                const double cw = cst * vabc[i];
                un = a.x + cst * hv - cw * b.x;
                vn = a.y - cst * hu - cw * b.y;
            } else if (absorb == ESPEC_ABSORB_SMOOTHW) {
                un = vabc[i] * (a.x + cst * hv);
                vn = vabc[i] * (a.y - cst * hu);
            } else {
                un = a.x + cst * hv;
                vn = a.y - cst * hu;
            }
            out[k] = {un, vn};
        }
        const double w = w0 * xd[i];
        const double2 ca = cur[i], cb = cur[n + i];
        const double wvb = w * cb.y, wub = w * cb.x, wva = w * ca.y, wua = w * ca.x;
        prev[i] = {out[0].x + wvb, out[0].y - wub};
        prev[n + i] = {out[1].x + wva, out[1].y - wua};
    }
}

struct FdSampleArgs {
    FdGrid g;
    const double2 *x;    // psi
    const double2 *x0;   // psi0
    const double *vp, *dm;
    double e;            // field of VEFF
    int energies, npr;
    const double *proj;  // [npr][n]
    double *part;        // [slot][blocks][kPpsodVals]
    int slot;
};

// Block partials of the sample quantities of one state.
__global__ __launch_bounds__(256) void fdsod_sample(FdSampleArgs a)
{
#pragma clang fp contract(off)
    __shared__ double red[4][kPpsodVals];
    double s[kPpsodVals];
#pragma unroll
    for (int k = 0; k < kPpsodVals; ++k) s[k] = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < a.g.n; i += gridDim.x * 256) {
        const double2 p = a.x[i], q = a.x0[i];
        s[0] += q.x * p.x + q.y * p.y;  // ppsod.f: FCORREL(U0,U2) + FCORREL(V0,V2)
        s[1] += q.x * p.y - q.y * p.x;  //          FCORREL(U0,V2) - FCORREL(V0,U2)
        s[2] += p.x * p.x + p.y * p.y;
        if (a.energies) {
            const double f = a.dm[i] * a.e;
            const double ve = a.vp[i] - f;
            s[3] += p.x * fd_h(a.g, a.x, false, ve, i) + p.y * fd_h(a.g, a.x, true, ve, i);
            s[4] += p.x * fd_h(a.g, a.x, false, a.vp[i], i) +
                    p.y * fd_h(a.g, a.x, true, a.vp[i], i);
        }
#pragma unroll
        for (int k = 0; k < kPpsodProj; ++k)
            if (k < a.npr) {
                const double w = a.proj[(long long)k * a.g.n + i];
                s[5 + 2 * k] += w * p.x;
                s[6 + 2 * k] += w * p.y;
            }
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < kPpsodVals; ++k) {
        const double t = wave_sum(s[k]);
        if (lane == 0) red[w][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < kPpsodVals) {
        const double t = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
                         red[3][threadIdx.x];
        a.part[((long long)a.slot * gridDim.x + blockIdx.x) * kPpsodVals + threadIdx.x] = t;
    }
}

// VEFF = VP - DM E with the product rounded first (chpot.f:58-62).
__global__ __launch_bounds__(256) void fdsod_veff(int n, const double *vp, const double *dm,
                                                  double e, double *out)
{
#pragma clang fp contract(off)
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double f = dm[i] * e;
        out[i] = vp[i] - f;
    }
}

// out[slot][k] = sum over blocks in block order.
__global__ __launch_bounds__(256) void fdsod_reduce(const double *part, int blocks, int nslot,
                                                    double *out)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= nslot * kPpsodVals) return;
    const int slot = idx / kPpsodVals, k = idx % kPpsodVals;
    double s = 0;
    for (int b = 0; b < blocks; ++b) s += part[((long long)slot * blocks + b) * kPpsodVals + k];
    out[idx] = s;
}

}  // namespace espec

using namespace espec;

extern "C" {

int espec_ppsod(int ndim, const int *np, double dt, double ti, double tf, int lc,
                const double *shm, double *u, double *v, const double *vp, const double *dm,
                int efc, double e0, double t0, double td, double tp, double omg, double sni,
                int kl, int absorb, const double *vabc, int npr, const double *proj,
                double *uprev, double *vprev, double *veff, double *cr, double *ci,
                double *nrm, double *et, double *et0, double *pr, int maxsamp, int *nsamp,
                int flags)
{
    if (ndim < 1 || ndim > 3 || !np || !shm || !u || !v || !vp || !dm || !cr || !ci ||
        maxsamp < 1 || lc < 1 || !(dt != 0) || efc < ESPEC_PULSE_NONE || efc > ESPEC_PULSE_SIN2 ||
        npr < 0 || npr > kPpsodProj || (npr > 0 && !proj) || absorb < ESPEC_ABSORB_NONE ||
        absorb > ESPEC_ABSORB_SMOOTHW || (absorb != ESPEC_ABSORB_NONE && !vabc))
        return ESPEC_ERR_ARG;
    if (flags & ESPEC_MIXED) return ESPEC_ERR_UNSUPPORTED;  // the FD leapfrog runs in FP64
    FdGrid g{};
    g.ndim = ndim;
    long long n = 1;
    int st = 1;
    for (int a = 0; a < 3; ++a) {
        if (a < ndim && np[a] < 4) return ESPEC_ERR_ARG;  // 5-point stencils
        g.np[a] = a < ndim ? np[a] : 1;
        g.st[a] = st;
        st *= g.np[a];
        n *= g.np[a];
        g.shm[a] = a < ndim ? shm[a] : 0.0;
    }
    if (n > (1LL << 28)) return ESPEC_ERR_ARG;
    g.n = (int)n;
    // The reference's time loop (ppsod.f:80, 246-254, 348-362): fields of
    // the start-up, then NINT((TF+DT/2-T)/DT) steps with T advanced by DT.
    const double shfs = 2.4188843243e-2;
    const double cst = 2.0 * dt / shfs;
    double t = ti;
    std::vector<double> ef, cs;
    ef.push_back(espec_ef(efc, e0, t0, td, tp, omg, sni, kl, (t + dt) / 2.0));
    cs.push_back(cst / 4.0);
    ef.push_back(espec_ef(efc, e0, t0, td, tp, omg, sni, kl, t + dt));
    cs.push_back(cst / 2.0);
    t = ti + dt;
    const long long nf = std::llround((tf + dt / 2.0 - t) / dt);
    for (long long k = 1; k <= nf; ++k) {
        t = t + dt;
        ef.push_back(espec_ef(efc, e0, t0, td, tp, omg, sni, kl, t));
        cs.push_back(cst);
    }
    const int nstep = (int)ef.size() - 1;  // start-up counts as one step
    const int nslot = std::min<long long>(maxsamp, nstep / lc + 1);
    const size_t by = sizeof(double) * g.n;
    std::vector<double> h0(2 * (size_t)g.n);
    for (int i = 0; i < g.n; ++i) {
        h0[2 * i] = u[i];
        h0[2 * i + 1] = v[i];
    }
    DevBuf dpsi0, da, db, dvp, ddm, dva, dpr, dpart, dout;
    ESPEC_HIP_TRY(dpsi0.upload(h0.data(), 2 * by));
    ESPEC_HIP_TRY(da.upload(h0.data(), 2 * by));
    ESPEC_HIP_TRY(db.upload(h0.data(), 2 * by));
    ESPEC_HIP_TRY(dvp.upload(vp, by));
    ESPEC_HIP_TRY(ddm.upload(dm, by));
    if (absorb != ESPEC_ABSORB_NONE) ESPEC_HIP_TRY(dva.upload(vabc, by));
    if (npr > 0) ESPEC_HIP_TRY(dpr.upload(proj, by * npr));
    const int blocks = std::max(1, std::min(kPpsodBlocks, (g.n + 255) / 256));
    ESPEC_HIP_TRY(dpart.alloc(sizeof(double) * kPpsodVals * (size_t)blocks * nslot));
    ESPEC_HIP_TRY(dout.alloc(sizeof(double) * kPpsodVals * (size_t)nslot));
    const bool energies = et || et0;
    FdSampleArgs sa{};
    sa.g = g;
    sa.x0 = dpsi0.as<double2>();
    sa.vp = dvp.as<double>();
    sa.dm = ddm.as<double>();
    sa.energies = energies ? 1 : 0;
    sa.npr = npr;
    sa.proj = npr > 0 ? dpr.as<double>() : nullptr;
    sa.part = dpart.as<double>();
    auto sample = [&](const double2 *x, double e, int slot) -> hipError_t {
        FdSampleArgs s = sa;
        s.x = x;
        s.e = e;
        s.slot = slot;
        hipLaunchKernelGGL(fdsod_sample, dim3(blocks), dim3(256), 0, nullptr, s);
        return hipGetLastError();
    };
    const double *va = absorb != ESPEC_ABSORB_NONE ? dva.as<double>() : nullptr;
    auto step = [&](double2 *prev, const double2 *cur, int k) -> hipError_t {
        hipLaunchKernelGGL(fdsod_step, dim3(blocks), dim3(256), 0, nullptr, g, prev, cur,
                           dvp.as<double>(), ddm.as<double>(), ef[k], cs[k], absorb, va);
        return hipGetLastError();
    };
    // sample 0: psi0 with VP (EIGENERG's VPOT before any CHPOT is VP)
    ESPEC_HIP_TRY(sample(dpsi0.as<double2>(), 0.0, 0));
    double2 *pa = da.as<double2>(), *pb = db.as<double2>();
    // start-up: (prev, cur) = (psi0, psi0) -> quarter step -> half step
    ESPEC_HIP_TRY(step(pa, pb, 0));  // pa = psi(dt/2); prev is now pb = psi0
    ESPEC_HIP_TRY(step(pb, pa, 1));  // pb = psi(dt)
    // psi(t - dt) <- psi0 (ppsod.f:348-351)
    ESPEC_HIP_TRY(hipMemcpy(pa, dpsi0.p, 2 * by, hipMemcpyDeviceToDevice));
    double2 *prev = pa, *cur = pb;
    int slot = 1, mc = 1;
    if (mc == lc) {
        mc = 0;
        if (slot < nslot) ESPEC_HIP_TRY(sample(cur, ef[1], slot));
        ++slot;
    }
    for (int k = 2; k <= nstep; ++k) {
        ESPEC_HIP_TRY(step(prev, cur, k));
        std::swap(prev, cur);
        if (++mc == lc) {
            mc = 0;
            if (slot < nslot) ESPEC_HIP_TRY(sample(cur, ef[k], slot));
            ++slot;
        }
    }
    hipLaunchKernelGGL(fdsod_reduce, dim3((nslot * kPpsodVals + 255) / 256), dim3(256), 0,
                       nullptr, dpart.as<double>(), blocks, nslot, dout.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    std::vector<double> out((size_t)kPpsodVals * nslot), hs(2 * (size_t)g.n);
    ESPEC_HIP_TRY(hipMemcpy(out.data(), dout.p, sizeof(double) * out.size(),
                            hipMemcpyDeviceToHost));
    for (int s = 0; s < nslot; ++s) {
        const double *o = out.data() + (size_t)kPpsodVals * s;
        cr[s] = o[0];
        ci[s] = o[1];
        if (nrm) nrm[s] = o[2];
        if (et) et[s] = o[3];
        if (et0) et0[s] = o[4];
        if (pr)
            for (int k = 0; k < npr; ++k) {
                pr[((size_t)s * npr + k) * 2] = o[5 + 2 * k];
                pr[((size_t)s * npr + k) * 2 + 1] = o[6 + 2 * k];
            }
    }
    if (nsamp) *nsamp = nslot;
    ESPEC_HIP_TRY(hipMemcpy(hs.data(), cur, 2 * by, hipMemcpyDeviceToHost));
    for (int i = 0; i < g.n; ++i) {
        u[i] = hs[2 * i];
        v[i] = hs[2 * i + 1];
    }
    if (uprev || vprev) {
        ESPEC_HIP_TRY(hipMemcpy(hs.data(), prev, 2 * by, hipMemcpyDeviceToHost));
        for (int i = 0; i < g.n; ++i) {
            if (uprev) uprev[i] = hs[2 * i];
            if (vprev) vprev[i] = hs[2 * i + 1];
        }
    }
    if (veff) {  // VPOT as the reference leaves it: the last CHPOT, formed on the device
        DevBuf dve;
        ESPEC_HIP_TRY(dve.alloc(by));
        hipLaunchKernelGGL(fdsod_veff, dim3(blocks), dim3(256), 0, nullptr, g.n, dvp.as<double>(),
                           ddm.as<double>(), ef.back(), dve.as<double>());
        ESPEC_HIP_TRY(hipGetLastError());
        ESPEC_HIP_TRY(hipMemcpy(veff, dve.p, by, hipMemcpyDeviceToHost));
    }
    return ESPEC_OK;
}

// S2PPSOD (s2ppsod.f:1-1143) on the device: the time loop and sampling of
// the oracle's restatement (two surfaces, the '.CORD' coupling, the fields
// only when E0 /= 0), the printed rows as the reference prints them.
int espec_s2ppsod(int ndim, const int *np, double dt, double ti, double tf, int lc,
                  const double *shm, double xi, double sh, double sm1, double aa, double *u,
                  double *v, const double *vp, const double *dm, int efc, double e0, double t0,
                  double td, double tp, double omg, double sni, int kl, int absorb,
                  const double *vabc, int prtcrl, double tol, double *tab, int maxrows,
                  int *nrows, int *info, double *uprev, double *vprev, double *veff, int flags)
{
    if (ndim < 1 || ndim > 3 || !np || !shm || !u || !v || !vp || !dm || lc < 1 ||
        !(dt != 0) || !(sm1 != 0) || efc < ESPEC_PULSE_NONE || efc > ESPEC_PULSE_SIN2 ||
        prtcrl < 0 || prtcrl > 2 || maxrows < 0 || (maxrows > 0 && !tab) ||
        absorb < ESPEC_ABSORB_NONE || absorb > ESPEC_ABSORB_SMOOTHW ||
        (absorb != ESPEC_ABSORB_NONE && !vabc))
        return ESPEC_ERR_ARG;
    if (flags & ESPEC_MIXED) return ESPEC_ERR_UNSUPPORTED;  // FP64 leapfrog
    FdGrid g{};
    g.ndim = ndim;
    long long nn = 1;
    int st = 1;
    for (int a = 0; a < 3; ++a) {
        if (a < ndim && np[a] < 4) return ESPEC_ERR_ARG;
        g.np[a] = a < ndim ? np[a] : 1;
        g.st[a] = st;
        st *= g.np[a];
        nn *= g.np[a];
        g.shm[a] = a < ndim ? shm[a] : 0.0;
    }
    if (nn > (1LL << 27)) return ESPEC_ERR_ARG;
    g.n = (int)nn;
    const int n = g.n;
    const double shfs = 2.4188843243e-2;
    const double cst = 2.0 * dt / shfs, csqc = cst * aa / sm1;
    const bool field = e0 != 0.0;
    // steps: (field, strength, coupling) of the start-up and of the loop
    // DO T=TI+2*DT,TF+DT/2,DT (MAX(INT((m2 - m1 + m3)/m3), 0) trips, T += DT)
    struct Step { double e, cst, w0, t; };
    std::vector<Step> steps;
    auto efv = [&](double t) { return field ? espec_ef(efc, e0, t0, td, tp, omg, sni, kl, t) : 0.0; };
    steps.push_back({efv((ti + dt) / 2.0), cst / 4.0, csqc / 4.0, ti});
    steps.push_back({efv(ti + dt), cst / 2.0, csqc / 2.0, ti + dt});
    const double m1 = ti + 2.0 * dt;
    long long trips = (long long)((tf + dt / 2.0 - m1 + dt) / dt);
    if (trips < 0) trips = 0;
    if (trips > (1LL << 26)) return ESPEC_ERR_ARG;
    double t = m1;
    for (long long k = 0; k < trips; ++k, t = t + dt) steps.push_back({efv(t), cst, csqc, t});
    // printed rows: (after step index, quirk, time, field of VEFF); -1: psi0
    struct Row { int after, quirk; double t, e; };
    std::vector<Row> rows;
    if (prtcrl >= 1) rows.push_back({-1, 1, ti, 0.0});
    if (prtcrl == 2) rows.push_back({1, 2, ti + dt, steps[1].e});
    int mc = lc == 1 ? 0 : 1;
    for (size_t k = 2; k < steps.size(); ++k) {
        if (++mc == lc) mc = 0;
        if (prtcrl == 2 || (prtcrl == 1 && mc == 0))
            rows.push_back({(int)k, 0, steps[k].t, steps[k].e});
    }
    std::vector<double> xd(n);
    double x = xi;
    for (int i = 0; i < n; ++i) {
        x = x + sh;
        xd[i] = x;
    }
    const size_t by = sizeof(double) * n;
    std::vector<double> h0(4 * (size_t)n);
    for (int i = 0; i < 2 * n; ++i) {
        h0[2 * i] = u[i];
        h0[2 * i + 1] = v[i];
    }
    DevBuf dpsi0, da, db, dvp, ddm, dva, dxd, dpart, dout;
    ESPEC_HIP_TRY(dpsi0.upload(h0.data(), 4 * by));
    ESPEC_HIP_TRY(da.upload(h0.data(), 4 * by));
    ESPEC_HIP_TRY(db.upload(h0.data(), 4 * by));
    ESPEC_HIP_TRY(dvp.upload(vp, 2 * by));
    ESPEC_HIP_TRY(ddm.upload(dm, 2 * by));
    ESPEC_HIP_TRY(dxd.upload(xd.data(), by));
    if (absorb != ESPEC_ABSORB_NONE) ESPEC_HIP_TRY(dva.upload(vabc, by));
    const int blocks = std::max(1, std::min(kPpsodBlocks, (n + 255) / 256));
    const int nslot = 2 * (int)std::max<size_t>(rows.size(), 1);
    ESPEC_HIP_TRY(dpart.alloc(sizeof(double) * kPpsodVals * (size_t)blocks * nslot));
    ESPEC_HIP_TRY(dout.alloc(sizeof(double) * kPpsodVals * (size_t)nslot));
    FdSampleArgs sa{};
    sa.g = g;
    sa.energies = 1;
    sa.npr = 0;
    sa.part = dpart.as<double>();
    auto sample = [&](const double2 *xs, double e, int slot) -> hipError_t {
        for (int k = 0; k < 2; ++k) {
            FdSampleArgs a = sa;
            a.x = xs + (size_t)k * n;
            a.x0 = dpsi0.as<double2>() + (size_t)k * n;
            a.vp = dvp.as<double>() + (size_t)k * n;
            a.dm = ddm.as<double>() + (size_t)k * n;
            a.e = e;
            a.slot = 2 * slot + k;
            hipLaunchKernelGGL(fdsod_sample, dim3(blocks), dim3(256), 0, nullptr, a);
        }
        return hipGetLastError();
    };
    const double *va = absorb != ESPEC_ABSORB_NONE ? dva.as<double>() : nullptr;
    double2 *prev = da.as<double2>(), *cur = db.as<double2>();
    size_t ri = 0;
    auto rows_after = [&](int k, const double2 *xs) -> hipError_t {
        while (ri < rows.size() && rows[ri].after == k) {
            hipError_t e = sample(xs, rows[ri].e, (int)ri);
            if (e != hipSuccess) return e;
            ++ri;
        }
        return hipSuccess;
    };
    ESPEC_HIP_TRY(rows_after(-1, dpsi0.as<double2>()));
    for (size_t k = 0; k < steps.size(); ++k) {
        hipLaunchKernelGGL(fdsod2_step, dim3(blocks), dim3(256), 0, nullptr, g, prev, cur,
                           dvp.as<double>(), ddm.as<double>(), steps[k].e, steps[k].cst,
                           steps[k].w0, dxd.as<double>(), absorb, va);
        ESPEC_HIP_TRY(hipGetLastError());
        std::swap(prev, cur);  // cur = psi(t+dt), prev = psi(t)
        ESPEC_HIP_TRY(rows_after((int)k, cur));
        // after the start-up psi(t - dt) restarts from psi0 (s2ppsod.f:753-756)
        if (k == 1) ESPEC_HIP_TRY(hipMemcpy(prev, dpsi0.p, 4 * by, hipMemcpyDeviceToDevice));
    }
    if (!rows.empty()) {
        hipLaunchKernelGGL(fdsod_reduce, dim3((nslot * kPpsodVals + 255) / 256), dim3(256), 0,
                           nullptr, dpart.as<double>(), blocks, nslot, dout.as<double>());
        ESPEC_HIP_TRY(hipGetLastError());
    }
    std::vector<double> out((size_t)kPpsodVals * nslot), hs(4 * (size_t)n);
    if (!rows.empty())
        ESPEC_HIP_TRY(hipMemcpy(out.data(), dout.p, sizeof(double) * out.size(),
                                hipMemcpyDeviceToHost));
    int inf = 0;
    double f0 = 0;
    for (size_t r = 0; r < rows.size(); ++r) {
        const double *a = out.data() + (size_t)kPpsodVals * (2 * r);
        const double *b = a + kPpsodVals;
        const double cta = a[0] * a[0] + a[1] * a[1], ctb = b[0] * b[0] + b[1] * b[1];
        if (rows[r].quirk == 1) f0 = a[2] + b[2];  // F0 = FNTA + FNTB of psi0
        const double fntb = rows[r].quirk == 1 ? b[2] * b[2] : b[2];
        const double fnt = a[2] + fntb;
        if (rows[r].quirk == 0 && absorb == ESPEC_ABSORB_NONE &&
            std::fabs((fnt - f0) / f0) * 1.0e2 > tol)
            inf = 1;
        if ((int)r < maxrows) {
            double *o = tab + 12 * r;
            o[0] = rows[r].t;
            o[1] = a[3] + b[3];
            o[2] = a[4] + b[4];
            o[3] = a[0];
            o[4] = a[1];
            o[5] = cta;
            o[6] = rows[r].quirk == 2 ? ctb : b[0];
            o[7] = b[1];
            o[8] = ctb;
            o[9] = a[2];
            o[10] = fntb;
            o[11] = fnt;
        }
    }
    if (nrows) *nrows = (int)rows.size();
    if (info) *info = inf;
    ESPEC_HIP_TRY(hipMemcpy(hs.data(), cur, 4 * by, hipMemcpyDeviceToHost));
    for (int i = 0; i < 2 * n; ++i) {
        u[i] = hs[2 * i];
        v[i] = hs[2 * i + 1];
    }
    if (uprev || vprev) {
        ESPEC_HIP_TRY(hipMemcpy(hs.data(), prev, 4 * by, hipMemcpyDeviceToHost));
        for (int i = 0; i < 2 * n; ++i) {
            if (uprev) uprev[i] = hs[2 * i];
            if (vprev) vprev[i] = hs[2 * i + 1];
        }
    }
    if (veff) {  // VPOT as the reference leaves it: the last CHPOT when E0 /= 0
        if (field) {
            DevBuf dve;
            ESPEC_HIP_TRY(dve.alloc(2 * by));
            hipLaunchKernelGGL(fdsod_veff, dim3(blocks), dim3(256), 0, nullptr, 2 * n,
                               dvp.as<double>(), ddm.as<double>(), steps.back().e,
                               dve.as<double>());
            ESPEC_HIP_TRY(hipGetLastError());
            ESPEC_HIP_TRY(hipMemcpy(veff, dve.p, 2 * by, hipMemcpyDeviceToHost));
        } else {
            std::copy(vp, vp + 2 * (size_t)n, veff);
        }
    }
    return ESPEC_OK;
}

}  // extern "C"
