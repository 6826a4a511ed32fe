// Absorbing-boundary data (ABSORBINGBC, absorb.f:1-130) on the device:
// VABC on the grid, x fastest, one thread per point, for the propagators'
// .SMOOTHW mask and .VOPTIC imaginary potential.  The shapes keep the
// reference's formulas, including its quirks: the 1-D .SMOOTHW raises
// (x - AL) to VOI(2) on the left (a growing factor for even VOI(2)), the
// .VOPTIC_KOSLOFF value does not depend on x, and .SMRS is a mask that
// the propagators never apply (they test TPABSOR for .SMOOTHW).
#include <cmath>

#include "api_common.hpp"

namespace espec {
namespace {

struct AbgArgs {
    int ndim, np0, np1, kind;
    double al[2], ar[2], sh[2], xi[2], xf[2], voi[2];
};

__device__ double abg_1d(const AbgArgs &a, double x1)
{
    const bool left = x1 < a.al[0], right = x1 > a.ar[0];
    switch (a.kind) {
    case ESPEC_ABG_SMOOTHW:
        if (left) return exp(a.voi[0] * pow(x1 - a.al[0], a.voi[1]) / (x1 + a.sh[0] - a.xi[0]));
        if (right) return exp(-(a.voi[0] * (x1 - a.ar[0]) / pow(a.xf[0] + a.sh[0] - x1, a.voi[1])));
        return 1.0;
    case ESPEC_ABG_SMRS:
        if (left) return 1.0 - a.voi[0] * ((x1 - a.al[0]) * (x1 - a.al[0]));
        if (right) return 1.0 - a.voi[0] * ((x1 - a.ar[0]) * (x1 - a.ar[0]));
        return 1.0;
    case ESPEC_ABG_VOPTIC_LINEAR:
        if (left) return -a.voi[0] * (x1 - a.al[0]) / (x1 + a.sh[0] - a.xi[0]);
        if (right) return a.voi[0] * (x1 - a.ar[0]) / (a.xf[0] + a.sh[0] - x1);
        return 0.0;
    case ESPEC_ABG_VOPTIC_KOSLOFF:
        if (left) return a.voi[0] / cosh(a.xi[0] - a.al[0] / a.voi[1]);
        if (right) return a.voi[0] / cosh(a.xf[0] - a.al[0] / a.voi[1]);
        return 0.0;
    default:  // ESPEC_ABG_VOPTIC_POLIN
        if (left) return a.voi[0] * pow(x1 - a.al[0], a.voi[1]);
        if (right) return a.voi[0] * pow(x1 - a.ar[0], a.voi[1]);
        return 0.0;
    }
}

// absorb.f:80-100: the 2-D .SMOOTHW mask, tested axis by axis in that order.
__device__ double abg_2d(const AbgArgs &a, double x1, double x2)
{
    if (x1 < a.al[0]) return exp(a.voi[0] * (x1 - a.al[0]) / (x1 + a.sh[0] - a.xi[0]));
    if (x2 < a.al[1]) return exp(a.voi[1] * (x2 - a.al[1]) / (x2 + a.sh[1] - a.xi[1]));
    if (x1 > a.ar[0]) return exp(-a.voi[0] * (x1 - a.ar[0]) / (a.xf[0] + a.sh[0] - x1));
    if (x2 > a.ar[1]) return exp(-a.voi[1] * (x2 - a.ar[1]) / (a.xf[1] + a.sh[1] - x2));
    return 1.0;
}

__global__ __launch_bounds__(256) void absorbingbc_kernel(AbgArgs a, double *vabc)
{
    const long long n = (long long)a.np0 * a.np1;
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < n; k += gridDim.x * 256LL) {
        const int i = (int)(k % a.np0), j = (int)(k / a.np0);
        // X = XI + (I - ONE) SH with the 1-based I (absorb.f:61)
        const double x1 = a.xi[0] + ((i + 1) - 1.0) * a.sh[0];
        if (a.ndim == 1) {
            vabc[k] = abg_1d(a, x1);
        } else {
            const double x2 = a.xi[1] + ((j + 1) - 1.0) * a.sh[1];
            vabc[k] = abg_2d(a, x1, x2);
        }
    }
}

}  // namespace
}  // namespace espec

using namespace espec;

extern "C" int espec_absorbingbc(int ndim, const int *np, int kind, const double *al,
                                 const double *ar, const double *sh, const double *xi,
                                 const double *xf, const double *voi, double *vabc,
                                 double *alpha)
{
    if (!np || !al || !ar || !sh || !xi || !xf || !voi || !vabc || ndim < 1 || ndim > 3 ||
        kind < ESPEC_ABG_SMOOTHW || kind > ESPEC_ABG_VOPTIC_POLIN)
        return ESPEC_ERR_ARG;
    // the reference stops for 3-D and for 2-D shapes other than .SMOOTHW
    if (ndim == 3 || (ndim == 2 && kind != ESPEC_ABG_SMOOTHW)) return ESPEC_ERR_UNSUPPORTED;
    AbgArgs a{};
    a.ndim = ndim;
    a.np0 = np[0];
    a.np1 = ndim == 2 ? np[1] : 1;
    if (a.np0 < 1 || a.np1 < 1) return ESPEC_ERR_ARG;
    a.kind = kind;
    for (int d = 0; d < ndim; ++d) {
        a.al[d] = al[d];
        a.ar[d] = ar[d];
        a.sh[d] = sh[d];
        a.xi[d] = xi[d];
        a.xf[d] = xf[d];
        a.voi[d] = voi[d];
    }
    // The 1-D shapes with an exponent or width in VOI(2) read it
    // (absorb.f:66,72,123,133); the others take one VOI.
    if (ndim == 1 && (kind == ESPEC_ABG_SMOOTHW || kind == ESPEC_ABG_VOPTIC_KOSLOFF ||
                      kind == ESPEC_ABG_VOPTIC_POLIN))
        a.voi[1] = voi[1];
    const size_t n = (size_t)a.np0 * a.np1;
    DevBuf d;
    ESPEC_HIP_TRY(d.alloc(sizeof(double) * n));
    hipLaunchKernelGGL(absorbingbc_kernel, dim3((unsigned)std::min<size_t>(1024, (n + 255) / 256)),
                       dim3(256), 0, nullptr, a, d.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(vabc, d.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    // ALPHA = 1 - 0.1 (XF - AR) with the default-REAL 0.1 (absorb.f:58)
    if (alpha && ndim == 1 && kind == ESPEC_ABG_SMOOTHW)
        *alpha = 1 - (double)0.1f * (xf[0] - ar[0]);
    return ESPEC_OK;
}
