// Energies and norms on the device: EFFT4 (eigenergfft.f:1-103), EIGENERG
// (eigenerg.f:1-70) and ECNORM (ecnorm.f:1-68).
//
// EFFT4 = <u|V|u> + <v|V|v> + <u|F^-1 T F|u>/N + <v|F^-1 T F|v>/N with the
// transform applied to the half-swapped copy of psi (eigenergfft.f:60-95),
// the line FFTs of espec_fft on each axis, and one FP64 reduction; the
// reductions sum per-block partials in block order (deterministic).
#include <algorithm>
#include <cmath>
#include <vector>

#include "api_common.hpp"
#include "common.hpp"

namespace espec {

hipError_t run_line_fft(double *re, double *im, int ntot, int n, int nspan, int isn, bool ref,
                        hipStream_t s);

constexpr int kRedBlocks = 256;

// part[block] = sum over the block's points of
//   u V u + v V v + b (u wu[sw(i)] + v wv[sw(i)])   (mode 0, EFFT4)
//   u hu                                             (mode 1, EIGENERG / FCORREL)
__global__ __launch_bounds__(256) void energy_parts(int n, int mode, double b, const double *u,
                                                    const double *v, const double *vpot,
                                                    const double *wu, const double *wv,
                                                    double *part)
{
    __shared__ double red[4];
    const int nh = n / 2;
    double acc = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (mode == 0) {
            const int j = i < nh ? i + nh : i - nh;
            acc += u[i] * vpot[i] * u[i] + v[i] * vpot[i] * v[i];
            acc += u[i] * b * wu[j] + v[i] * b * wv[j];
        } else {
            acc += u[i] * wu[i];
        }
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void scale_by(int n, const double *t, double *x, double *y)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        x[i] *= t[i];
        y[i] *= t[i];
    }
}

__global__ __launch_bounds__(256) void half_swap(int n, const double *u, const double *v, double *x,
                                                 double *y)
{
    const int nh = n / 2;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int j = i < nh ? i + nh : i - nh;
        x[i] = u[j];
        y[i] = v[j];
    }
}

__global__ void sum_parts(const double *part, int nb, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int k = 0; k < nb; ++k) s += part[k];
        *out = s;
    }
}

// hu = H u, the 4th-order FD operator (same stencil as the Lanczos kernels).
__global__ __launch_bounds__(256) void fd_apply(int ndim, int n, int np0, int np1, int np2,
                                                double s0, double s1, double s2, const double *vpot,
                                                const double *u, double *hu)
{
    const int npa[3] = {np0, np1, np2};
    const double sha[3] = {s0, s1, s2};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double acc = vpot[i] * u[i];
        int st = 1;
        for (int d = 0; d < ndim; ++d) {
            const int ia = (i / st) % npa[d];
            double t = 30.0 * u[i];
            if (ia >= 2) t += u[i - 2 * st];
            if (ia >= 1) t -= 16.0 * u[i - st];
            if (ia + 1 < npa[d]) t -= 16.0 * u[i + st];
            if (ia + 2 < npa[d]) t += u[i + 2 * st];
            acc += sha[d] * t;
            st *= npa[d];
        }
        hu[i] = acc;
    }
}

hipError_t reduce(int n, int mode, double b, const double *u, const double *v, const double *vp,
                  const double *wu, const double *wv, double *out)
{
    const int nb = std::max(1, std::min(kRedBlocks, (n + 255) / 256));
    DevBuf part, res;
    hipError_t e = part.alloc(sizeof(double) * nb);
    if (e == hipSuccess) e = res.alloc(sizeof(double));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(energy_parts, dim3(nb), dim3(256), 0, nullptr, n, mode, b, u, v, vp, wu, wv,
                       part.as<double>());
    hipLaunchKernelGGL(sum_parts, dim3(1), dim3(64), 0, nullptr, part.as<double>(), nb,
                       res.as<double>());
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, res.p, sizeof(double), hipMemcpyDeviceToHost);
    return e;
}

}  // namespace espec

using namespace espec;

extern "C" {

int espec_efft4(int ndim, const int *np, double b, const double *u, const double *v,
                const double *tc, const double *vpot, double *e, int flags)
{
    if (ndim < 1 || ndim > 3 || !np || !u || !v || !tc || !vpot || !e) return ESPEC_ERR_ARG;
    const int n = ntot(ndim, np);
    for (int a = 0; a < ndim; ++a)
        if (!is_pow2(np[a]) || np[a] > 4096) return ESPEC_ERR_UNSUPPORTED;
    const size_t bytes = sizeof(double) * n;
    DevBuf du, dv, dwu, dwv, dtc, dvp;
    ESPEC_HIP_TRY(du.upload(u, bytes));
    ESPEC_HIP_TRY(dv.upload(v, bytes));
    ESPEC_HIP_TRY(dwu.alloc(bytes));
    ESPEC_HIP_TRY(dwv.alloc(bytes));
    ESPEC_HIP_TRY(dtc.upload(tc, bytes));
    ESPEC_HIP_TRY(dvp.upload(vpot, bytes));
    // eigenergfft.f:60-66: the half-swapped copy that is transformed.
    hipLaunchKernelGGL(half_swap, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, nullptr, n,
                       du.as<double>(), dv.as<double>(), dwu.as<double>(), dwv.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    const bool ref = (flags & ESPEC_REF) != 0;
    int span = 1;
    for (int a = 0; a < ndim; ++a) {
        span *= np[a];
        ESPEC_HIP_TRY(run_line_fft(dwu.as<double>(), dwv.as<double>(), n, np[a], span, 1, ref,
                                   nullptr));
    }
    // T in momentum space (eigenergfft.f:79-82).
    hipLaunchKernelGGL(scale_by, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, nullptr, n,
                       dtc.as<double>(), dwu.as<double>(), dwv.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    span = 1;
    for (int a = 0; a < ndim; ++a) {
        span *= np[a];
        ESPEC_HIP_TRY(run_line_fft(dwu.as<double>(), dwv.as<double>(), n, np[a], span, -1, ref,
                                   nullptr));
    }
    ESPEC_HIP_TRY(reduce(n, 0, b, du.as<double>(), dv.as<double>(), dvp.as<double>(),
                         dwu.as<double>(), dwv.as<double>(), e));
    return ESPEC_OK;
}

int espec_eigenerg(int ndim, const int *np, const double *shm, const double *u,
                   const double *vpot, double *e)
{
    if (ndim < 1 || ndim > 3 || !np || !shm || !u || !vpot || !e) return ESPEC_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (np[a] < 4) return ESPEC_ERR_ARG;
    const int n = ntot(ndim, np);
    const size_t bytes = sizeof(double) * n;
    DevBuf du, dv, dh;
    ESPEC_HIP_TRY(du.upload(u, bytes));
    ESPEC_HIP_TRY(dv.upload(vpot, bytes));
    ESPEC_HIP_TRY(dh.alloc(bytes));
    hipLaunchKernelGGL(fd_apply, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, nullptr, ndim,
                       n, np[0], ndim > 1 ? np[1] : 1, ndim > 2 ? np[2] : 1, shm[0],
                       ndim > 1 ? shm[1] : 0.0, ndim > 2 ? shm[2] : 0.0, dv.as<double>(),
                       du.as<double>(), dh.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(reduce(n, 1, 0.0, du.as<double>(), nullptr, nullptr, dh.as<double>(), nullptr, e));
    return ESPEC_OK;
}

int espec_ecnorm(int n, const double *u, double *out)
{
    if (n < 0 || !u || !out) return ESPEC_ERR_ARG;
    if (n == 0) {
        *out = 0;
        return ESPEC_OK;
    }
    DevBuf du;
    ESPEC_HIP_TRY(du.upload(u, sizeof(double) * n));
    double s = 0;
    ESPEC_HIP_TRY(reduce(n, 1, 0.0, du.as<double>(), nullptr, nullptr, du.as<double>(), nullptr, &s));
    *out = std::sqrt(s);
    return ESPEC_OK;
}

}  // extern "C"
