// Correlation -> spectrum on the device, and the stand-alone FFT entry.
//
//   espec_fft           FFT(A,B,NTOT,N,NSPAN,ISN) (fft3d.f:1): one variate of
//                       a multivariate transform, one workgroup per line in LDS
//   espec_assemble_corr C(t) on t = -TF..TF from the 2^(MF-1) samples
//                       (espec.f:1546-1566)
//   espec_hwhm          Lorentzian broadening and energy shift (hwhm.f:1)
//   espec_spectrumtd    window, half swap, FFTNT, print order (spectrumtd.f:1)
//   espec_lc_rule       LC and the *CHANGE final-time rule (espec.f:1324-1353)
//
// FFTNT (cfft.f:6-106) is a radix-2 decimation in frequency with twiddle
// exp(-i (ia-1) TWOPI/N) and a final bit reversal; it is reproduced stage by
// stage (same butterflies, same twiddle per butterfly), the early stages
// over HBM while a block is larger than LDS, the rest of each block in LDS.
// In ESPEC_REF mode TWOPI is the single-precision literal of cfft.f:35.
#include <algorithm>
#include <cmath>
#include <vector>

#include "api_common.hpp"
#include "fft_lds.hpp"

namespace espec {

constexpr int kLdsLine = 4096;  // complex128 points of one in-LDS block

// ------------------------------------------------------------- FFTNT

// Stage k of FFTNT over HBM: butterfly g pairs i and i + n2 of its block.
__global__ __launch_bounds__(256) void fftnt_stage_global(double2 *x, int n, int k,
                                                          const double2 *tw)
{
    const int n1 = n >> k, n2 = n1 >> 1;
    for (int g = blockIdx.x * 256 + threadIdx.x; g < n / 2; g += gridDim.x * 256) {
        const int blk = g / n2, j = g - blk * n2;
        const int i = blk * n1 + j, l = i + n2;
        const double2 w = tw[j << k];
        const double2 a = x[i], b = x[l];
        const double xt = a.x - b.x, yt = a.y - b.y;
        x[i] = {a.x + b.x, a.y + b.y};
        x[l] = {w.x * xt + w.y * yt, w.x * yt - w.y * xt};
    }
}

// Stages k0.. of FFTNT for one contiguous block of nb = n >> k0 points in LDS.
__global__ __launch_bounds__(256) void fftnt_stage_lds(double2 *x, int n, int k0, int m,
                                                       const double2 *tw)
{
    __shared__ double2 s[kLdsLine];
    const int nb = n >> k0;
    double2 *g = x + (long long)blockIdx.x * nb;
    for (int i = threadIdx.x; i < nb; i += 256) s[i] = g[i];
    __syncthreads();
    for (int k = k0; k < m; ++k) {
        const int n1 = n >> k, n2 = n1 >> 1;
        for (int q = threadIdx.x; q < nb / 2; q += 256) {
            const int blk = q / n2, j = q - blk * n2;
            const int i = blk * n1 + j, l = i + n2;
            const double2 w = tw[j << k];
            const double2 a = s[i], b = s[l];
            const double xt = a.x - b.x, yt = a.y - b.y;
            s[i] = {a.x + b.x, a.y + b.y};
            s[l] = {w.x * xt + w.y * yt, w.x * yt - w.y * xt};
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nb; i += 256) g[i] = s[i];
}

// FFTNT's input conjugation for ITYPE = -1 (cfft.f:40-44), into the
// interleaved work array.
__global__ __launch_bounds__(256) void fftnt_prep(int n, int itype, const double *x,
                                                  const double *y, double2 *z)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        z[i] = {x[i], itype == -1 ? -y[i] : y[i]};
}

__device__ __forceinline__ int bitrev(int i, int m) { return (int)(__brev((unsigned)i) >> (32 - m)); }

// ------------------------------------------------- spectrum pipeline

struct SpecArgs {
    int n, m, itype, iwin;
    double tf, dt, taux;
};

// spectrumtd.f:65-93: window, half swap; FFTNT's conjugation for ITYPE = -1
// (cfft.f:40-44).  Out of place: y[i] = f(x[(i + n/2) mod n]).
__global__ __launch_bounds__(256) void spec_prep(SpecArgs a, const double *u, const double *v,
                                                 double2 *y)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < a.n; i += gridDim.x * 256) {
        const int src = (i + a.n / 2) & (a.n - 1);
        double ur = u[src], vr = v[src];
        if (a.iwin == 1) {
            const double t = -a.tf + (double)src * a.dt;
            const double c = exp(-(t * t) / a.taux);
            ur = c * ur;
            vr = c * vr;
        }
        y[i] = {ur, a.itype == -1 ? -vr : vr};
    }
}

// Bit reversal of FFTNT and its ITYPE = -1 scaling (cfft.f:86-104).
__global__ __launch_bounds__(256) void fftnt_finish(SpecArgs a, const double2 *x, double *u,
                                                    double *v)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < a.n; i += gridDim.x * 256) {
        const double2 t = x[bitrev(i, a.m)];
        if (a.itype == -1) {
            u[i] = t.x / a.n;
            v[i] = -t.y / a.n;
        } else {
            u[i] = t.x;
            v[i] = t.y;
        }
    }
}

// hwhm.f:30-43: rotate by exp(i DEM t) and damp by exp(-|t| TAL), t = -TF + (i-1) DT.
__global__ __launch_bounds__(256) void hwhm_kernel(int n, double dt, double tf, double dem,
                                                   double ta, double *u, double *v)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double t = -tf + (double)i * dt;
        const double cs = cos(dem * t), sn = sin(dem * t), cstl = exp(-fabs(t) * ta);
        const double ui = u[i], vi = v[i];
        u[i] = cstl * (cs * ui - sn * vi);
        v[i] = cstl * (cs * vi + sn * ui);
    }
}

// espec.f:1552-1566 (1-based there): U(MC+I) = CR(I), U(MC-I+2) = U(MC+I),
// V(MC+I) = CI(I), V(MC-I+2) = -V(MC+I) for I = 1..MC, then U(1) = CR(MC),
// V(1) = -CI(MC).  For I = 1 the mirror lands on MC+1 itself, so V(MC+1)
// ends as -CI(1).
__global__ __launch_bounds__(256) void assemble_kernel(int mc, const double *cr, const double *ci,
                                                       double *u, double *v)
{
    for (int k = blockIdx.x * 256 + threadIdx.x; k < 2 * mc; k += gridDim.x * 256) {
        double uu, vv;
        if (k == 0) {
            uu = cr[mc - 1];
            vv = -ci[mc - 1];
        } else if (k <= mc) {
            uu = cr[mc - k];
            vv = -ci[mc - k];
        } else {
            uu = cr[k - mc];
            vv = ci[k - mc];
        }
        u[k] = uu;
        v[k] = vv;
    }
}

// ------------------------------------------------------ line FFT

struct LineArgs {
    DevAxis ax;
    int stride, nspan, nlines;
    const double2 *tw;
    const int *perm;
};

template <int SIGN>
__global__ __launch_bounds__(256) void fft_lines(LineArgs a, double *re, double *im)
{
    extern __shared__ double2 lds[];
    double2 *x = lds, *y = lds + a.ax.n;
    const int n = a.ax.n;
    for (int line = blockIdx.x; line < a.nlines; line += gridDim.x) {
        const long long base =
            (long long)(line / a.stride) * a.nspan + (line % a.stride);
        for (int k = threadIdx.x; k < n; k += 256) {
            const long long g = base + (long long)k * a.stride;
            x[k] = {re[g], im[g]};
        }
        __syncthreads();
        for (int s = 0; s < a.ax.nst; ++s) dif_stage<double, 256, SIGN>(x, a.ax, s, a.tw);
        permute_axis<double, 256>(x, y, a.ax, a.perm);
        for (int k = threadIdx.x; k < n; k += 256) {
            const long long g = base + (long long)k * a.stride;
            re[g] = y[k].x;
            im[g] = y[k].y;
        }
        __syncthreads();
    }
}

hipError_t run_fftnt(double2 *x, int n, int m, const double2 *tw, hipStream_t s)
{
    int k = 0;
    const int gb = std::max(1, std::min(2048, n / 512));
    for (; (n >> k) > kLdsLine && k < m; ++k)
        hipLaunchKernelGGL(fftnt_stage_global, dim3(gb), dim3(256), 0, s, x, n, k, tw);
    if (k < m)
        hipLaunchKernelGGL(fftnt_stage_lds, dim3(1 << k), dim3(256), 0, s, x, n, k, m, tw);
    return hipGetLastError();
}

// FFT(A,B,NTOT,N,NSPAN,ISN) on device arrays re/im (natural order in and out).
hipError_t run_line_fft(double *re, double *im, int ntot, int n, int nspan, int isn, bool ref,
                        hipStream_t s)
{
    if (n == 1) return hipSuccess;
    FftPlan p = make_fft_plan(n, ref);
    LineArgs la{};
    la.ax.n = n;
    la.ax.ln = ilog2(n);
    la.ax.nst = p.nst;
    set_line_geometry(la.ax, 1, 1, 1, 0, 0);
    la.ax.tw0 = 0;
    for (int st = 0; st < p.nst; ++st) {
        la.ax.rad[st] = p.rad[st];
        la.ax.lm[st] = p.lm[st];
        la.ax.lL[st] = p.lL[st];
        la.ax.off[st] = p.off[st];
    }
    la.stride = nspan / n;
    la.nspan = nspan;
    la.nlines = ntot / n;
    std::vector<double> tw;
    for (size_t i = 0; i < p.twc.size(); ++i) {
        tw.push_back(p.twc[i]);
        tw.push_back(p.tws[i]);
    }
    if (tw.empty()) tw.assign(2, 0.0);
    DevBuf dtw, dperm;
    hipError_t e = dtw.upload(tw.data(), sizeof(double) * tw.size());
    if (e == hipSuccess) e = dperm.upload(p.perm.data(), sizeof(int) * n);
    if (e != hipSuccess) return e;
    la.tw = dtw.as<double2>();
    la.perm = dperm.as<int>();
    const size_t lds = sizeof(double2) * 2 * n;
    auto kern = isn > 0 ? fft_lines<1> : fft_lines<-1>;
    if (lds > 64 * 1024) {
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(std::min(la.nlines, 4096)), dim3(256), lds, s, la, re, im);
    e = hipGetLastError();
    // the tables are freed on return: finish before that
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

std::vector<double> fftnt_twiddles(int n, bool ref)
{
    const double twopi = ref ? (double)(float)kTwoPi : kTwoPi;  // cfft.f:35
    const double p = twopi / n;
    std::vector<double> t((size_t)std::max(2, n));
    for (int j = 0; j < n / 2; ++j) {
        const double a = (double)j * p;
        t[2 * j] = std::cos(a);
        t[2 * j + 1] = std::sin(a);
    }
    return t;
}

}  // namespace espec

using namespace espec;

extern "C" {

int espec_lc_rule(double ti, double *tf, double dt, int mf, int change, int *lc)
{
    if (!tf || !lc || !(dt > 0) || mf < 1 || mf > 30) return ESPEC_ERR_ARG;
    for (;;) {
        const double sc = 2.0 * (*tf - ti) / dt;
        const int nc = (int)(sc / std::pow(2.0, mf));
        if (nc < 0) {
            *lc = 2147483645;
            return ESPEC_ERR_ARG;
        }
        if (std::fabs(sc - nc * std::pow(2.0, mf)) >= 1.0e-10 && change) {
            *tf = nc * std::pow(2.0, mf - 1.0) * dt;
            continue;
        }
        *lc = nc;
        return nc > 0 ? ESPEC_OK : ESPEC_ERR_ARG;
    }
}

int espec_assemble_corr(int mf, const double *cr, const double *ci, double *u, double *v)
{
    if (mf < 1 || mf > 28 || !cr || !ci || !u || !v) return ESPEC_ERR_ARG;
    const int mc = (1 << mf) / 2;
    DevBuf dcr, dci, du, dv;
    ESPEC_HIP_TRY(dcr.upload(cr, sizeof(double) * mc));
    ESPEC_HIP_TRY(dci.upload(ci, sizeof(double) * mc));
    ESPEC_HIP_TRY(du.alloc(sizeof(double) * 2 * mc));
    ESPEC_HIP_TRY(dv.alloc(sizeof(double) * 2 * mc));
    hipLaunchKernelGGL(assemble_kernel, dim3(std::min(1024, (2 * mc + 255) / 256)), dim3(256), 0,
                       nullptr, mc, dcr.as<double>(), dci.as<double>(), du.as<double>(),
                       dv.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(u, du.p, sizeof(double) * 2 * mc, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(v, dv.p, sizeof(double) * 2 * mc, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_hwhm(int mf, double dt, double tf, double tal, double de, double *u, double *v)
{
    if (mf < 1 || mf > 28 || !u || !v) return ESPEC_ERR_ARG;
    const int n = 1 << mf;
    DevBuf du, dv;
    ESPEC_HIP_TRY(du.upload(u, sizeof(double) * n));
    ESPEC_HIP_TRY(dv.upload(v, sizeof(double) * n));
    // hwhm.f:27-28: eV -> rad/fs.
    hipLaunchKernelGGL(hwhm_kernel, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, nullptr,
                       n, dt, tf, -de * 1.5192676, tal * 1.5192676, du.as<double>(),
                       dv.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(u, du.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(v, dv.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_spectrumtd(int iwin, int mf, double dt, double tf, double wp, double de, double *u,
                     double *v, double *e, double *w, double *wi, int flags)
{
    if (mf < 1 || mf > 28 || !u || !v || (iwin == 1 && !(wp > 0 && wp < 1))) return ESPEC_ERR_ARG;
    const int n = 1 << mf;
    const bool ref = (flags & ESPEC_REF) != 0;
    SpecArgs a{};
    a.n = n;
    a.m = mf;
    a.itype = -1;  // spectrumtd.f:57
    a.iwin = iwin == 1 ? 1 : 0;
    a.tf = tf;
    a.dt = dt;
    if (a.iwin) {
        const double tmax = n * dt, lg = std::log(1.0 / wp);
        a.taux = tmax * tmax / (lg * lg);
    }
    std::vector<double> tw = fftnt_twiddles(n, ref);
    DevBuf du, dv, dx, dt_;
    ESPEC_HIP_TRY(du.upload(u, sizeof(double) * n));
    ESPEC_HIP_TRY(dv.upload(v, sizeof(double) * n));
    ESPEC_HIP_TRY(dx.alloc(sizeof(double2) * n));
    ESPEC_HIP_TRY(dt_.upload(tw.data(), sizeof(double) * tw.size()));
    const int gb = std::min(1024, (n + 255) / 256);
    hipLaunchKernelGGL(spec_prep, dim3(gb), dim3(256), 0, nullptr, a, du.as<double>(),
                       dv.as<double>(), dx.as<double2>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(run_fftnt(dx.as<double2>(), n, mf, dt_.as<double2>(), nullptr));
    hipLaunchKernelGGL(fftnt_finish, dim3(gb), dim3(256), 0, nullptr, a, dx.as<double2>(),
                       du.as<double>(), dv.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(u, du.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(v, dv.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    // Printed rows (spectrumtd.f:97-105): second half first; E accumulated.
    double enc = 1.0 / (n * dt) * 1.519829846003e-1;
    enc = enc * 27.2113834;
    double en = -n / 2.0 * enc;
    if (de < 0) de = 0;
    int k = 0;
    for (int h = 0; h < 2; ++h)
        for (int i = h == 0 ? n / 2 : 0; i < (h == 0 ? n : n / 2); ++i, ++k) {
            if (e) e[k] = en + de;
            if (w) w[k] = u[i];
            if (wi) wi[k] = v[i];
            en = en + enc;
        }
    return ESPEC_OK;
}

int espec_fftnt(double *x, double *y, int n, int m, int itype, int flags)
{
    if (!x || !y || m < 0 || m > 28 || n != (1 << m)) return ESPEC_ERR_ARG;
    if (n == 1) return ESPEC_OK;  // cfft.f:38
    const bool ref = (flags & ESPEC_REF) != 0;
    SpecArgs a{};
    a.n = n;
    a.m = m;
    a.itype = itype;
    std::vector<double> tw = fftnt_twiddles(n, ref);
    DevBuf du, dv, dx, dt_;
    ESPEC_HIP_TRY(du.upload(x, sizeof(double) * n));
    ESPEC_HIP_TRY(dv.upload(y, sizeof(double) * n));
    ESPEC_HIP_TRY(dx.alloc(sizeof(double2) * n));
    ESPEC_HIP_TRY(dt_.upload(tw.data(), sizeof(double) * tw.size()));
    const int gb = std::min(1024, (n + 255) / 256);
    hipLaunchKernelGGL(fftnt_prep, dim3(gb), dim3(256), 0, nullptr, n, itype, du.as<double>(),
                       dv.as<double>(), dx.as<double2>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(run_fftnt(dx.as<double2>(), n, m, dt_.as<double2>(), nullptr));
    hipLaunchKernelGGL(fftnt_finish, dim3(gb), dim3(256), 0, nullptr, a, dx.as<double2>(),
                       du.as<double>(), dv.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(x, du.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(y, dv.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_fft(double *a, double *b, int ntot, int n, int nspan, int isn, int flags)
{
    if (!a || !b || n < 1 || ntot < n || nspan < n || nspan % n || ntot % nspan || isn == 0)
        return ESPEC_ERR_ARG;
    if (!is_pow2(n) || n > kLdsLine) return ESPEC_ERR_UNSUPPORTED;
    if (n == 1) return ESPEC_OK;
    DevBuf da, db;
    ESPEC_HIP_TRY(da.upload(a, sizeof(double) * ntot));
    ESPEC_HIP_TRY(db.upload(b, sizeof(double) * ntot));
    ESPEC_HIP_TRY(run_line_fft(da.as<double>(), db.as<double>(), ntot, n, nspan, isn,
                               (flags & ESPEC_REF) != 0, nullptr));
    ESPEC_HIP_TRY(hipMemcpy(a, da.p, sizeof(double) * ntot, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(b, db.p, sizeof(double) * ntot, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

}  // extern "C"
