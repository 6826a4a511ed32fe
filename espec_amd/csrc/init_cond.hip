// Gaussian initial packets (INIT_COND, init_cond.f:1-200; GET_GAUSS,
// gauss.f:1-20) on the device.
//
// Every form of the reference is a product of two 1-D factors over the
// flattened grid (x fastest): a factor (w1, w2) along one axis (the 1-D
// Gaussian times its plane wave for .CALC, or the 1-D eigenvector the
// caller read for .GETR/.GETC) and a Gaussian (ct, st) along the other.
// gauss_axis forms the Gaussian factor of one axis in one workgroup: its
// normalisation is the sequential sum of the reference (thread 0), then
// the lanes form rnorm exp(-ln2 (c/rA)^2) (cos, sin)(k c).  init_product
// writes psi = (w1 + i w2)(ct - i st) at every point, as the reference's
// U = w1 ct + w2 st, V = -w1 st + w2 ct.
#include <cmath>
#include <vector>

#include "api_common.hpp"

namespace espec {
namespace {

constexpr double kLn2Gauss = 0.69314718;  // gauss.f: LN_2 to 8 digits

__device__ __forceinline__ double get_gauss(double ra, double coord, double rnorm)
{
    const double q = coord / ra;
    return rnorm * exp(-(kLn2Gauss * (q * q)));
}

// out[2 i] = cos(k c_i) g_i, out[2 i + 1] = sin(k c_i) g_i with
// c_i = xi + i sh - x0 and g normalised over the axis (init_cond.f:60-74).
__global__ __launch_bounds__(256) void gauss_axis(int np, double xi, double sh, double x0,
                                                  double ra, double rk, double *out)
{
    __shared__ double rnorm;
    if (threadIdx.x == 0) {
        double r = 0;
        for (int i = 0; i < np; ++i) {
            const double g = get_gauss(ra, xi + i * sh - x0, 1.0);
            r = r + g * g;
        }
        rnorm = 1.0 / sqrt(r);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += 256) {
        const double c = xi + i * sh - x0;
        const double g = get_gauss(ra, c, rnorm);
        const double a = rk * c;
        out[2 * i] = cos(a) * g;
        out[2 * i + 1] = sin(a) * g;
    }
}

// psi at (i along the factor's axis, j along the Gaussian's axis);
// eig_fast: the factor runs along x (k = j np0 + i), else along y.
__global__ __launch_bounds__(256) void init_product(int ne, int ng, int eig_fast,
                                                    const double *w, const double *gs,
                                                    int w2_sign, int real_w, double *u, double *v)
{
    const long long n = (long long)ne * ng;
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < n; k += gridDim.x * 256LL) {
        int i, j;
        if (eig_fast) {
            i = (int)(k % ne);
            j = (int)(k / ne);
        } else {
            j = (int)(k % ng);
            i = (int)(k / ng);
        }
        const double w1 = w[2 * i], w2 = real_w ? 0.0 : w2_sign * w[2 * i + 1];
        const double ct = gs[2 * j], st = gs[2 * j + 1];
        if (real_w) {
            u[k] = w1 * ct;
            v[k] = -w1 * st;
        } else {
            u[k] = w1 * ct + w2 * st;
            v[k] = -w1 * st + w2 * ct;
        }
    }
}

}  // namespace
}  // namespace espec

using namespace espec;

extern "C" int espec_init_cond(int ndim, const int *np, int mode, const double *xi,
                               const double *sh, const double *x0, const double *ra,
                               const double *rk, double *u, double *v)
{
    if (!np || !xi || !sh || !x0 || !ra || !rk || !u || !v || ndim < 1 || ndim > 2)
        return ESPEC_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (np[a] < 1 || !(ra[a] != 0.0)) return ESPEC_ERR_ARG;
    if (ndim == 1 && mode != ESPEC_INIT_CALC) return ESPEC_ERR_ARG;
    if (mode < ESPEC_INIT_CALC || mode > ESPEC_INIT_GETC2) return ESPEC_ERR_ARG;
    DevBuf dg, dw, du, dv;
    if (ndim == 1) {  // init_cond.f:57-75: U = cos g, V = sin g
        ESPEC_HIP_TRY(dg.alloc(sizeof(double) * 2 * (size_t)np[0]));
        hipLaunchKernelGGL(gauss_axis, dim3(1), dim3(256), 0, nullptr, np[0], xi[0], sh[0], x0[0],
                           ra[0], rk[0], dg.as<double>());
        ESPEC_HIP_TRY(hipGetLastError());
        std::vector<double> h(2 * (size_t)np[0]);
        ESPEC_HIP_TRY(hipMemcpy(h.data(), dg.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < np[0]; ++i) {
            u[i] = h[2 * i];
            v[i] = h[2 * i + 1];
        }
        return ESPEC_OK;
    }
    // 2-D: the factor's axis and the Gaussian's axis
    const bool eig_y = mode == ESPEC_INIT_GETR2 || mode == ESPEC_INIT_GETC2;
    const int ea = eig_y ? 1 : 0, ga = eig_y ? 0 : 1;
    const int ne = np[ea], ng = np[ga];
    ESPEC_HIP_TRY(dg.alloc(sizeof(double) * 2 * (size_t)ng));
    ESPEC_HIP_TRY(dw.alloc(sizeof(double) * 2 * (size_t)ne));
    hipLaunchKernelGGL(gauss_axis, dim3(1), dim3(256), 0, nullptr, ng, xi[ga], sh[ga], x0[ga],
                       ra[ga], rk[ga], dg.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    int w2_sign = 1;
    const bool real_w = mode == ESPEC_INIT_GETR || mode == ESPEC_INIT_GETR2;
    if (mode == ESPEC_INIT_CALC) {
        // init_cond.f:152-158: W1 = cos g1, W2 = -sin g1
        hipLaunchKernelGGL(gauss_axis, dim3(1), dim3(256), 0, nullptr, ne, xi[ea], sh[ea], x0[ea],
                           ra[ea], rk[ea], dw.as<double>());
        ESPEC_HIP_TRY(hipGetLastError());
        w2_sign = -1;
    } else {  // the caller's 1-D eigenvector in u (and v)
        std::vector<double> h(2 * (size_t)ne);
        for (int i = 0; i < ne; ++i) {
            h[2 * i] = u[i];
            h[2 * i + 1] = real_w ? 0.0 : v[i];
        }
        ESPEC_HIP_TRY(hipMemcpy(dw.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    }
    const size_t n = (size_t)ne * ng;
    ESPEC_HIP_TRY(du.alloc(sizeof(double) * n));
    ESPEC_HIP_TRY(dv.alloc(sizeof(double) * n));
    hipLaunchKernelGGL(init_product, dim3((unsigned)std::min<size_t>(1024, (n + 255) / 256)),
                       dim3(256), 0, nullptr, ne, ng, eig_y ? 0 : 1, dw.as<double>(),
                       dg.as<double>(), w2_sign, real_w ? 1 : 0, du.as<double>(), dv.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(u, du.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(v, dv.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}
