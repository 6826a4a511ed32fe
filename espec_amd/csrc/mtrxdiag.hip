// MTRXDIAG (mtrxdiag.f:1-89): selected eigenpairs IL..IU of the dense
// finite-difference Hamiltonian, on the device.
//
// The reference packs the upper triangle of H (AP1D, ap1d.f; AP2D, ap2d.f)
// and calls DSPEVX('V','I','U') (mtrxdiag.f:82): Householder
// tridiagonalisation, bisection of the selected eigenvalues to ABSTOL and
// inverse iteration.  Here a kernel writes the same matrix, one thread per
// column, into a column-major device array, and rocSOLVER's syevx runs the
// same method (tridiagonal reduction, bisection to ABSTOL, inverse
// iteration, back-transformation) on the device.  This is set-up work
// (initial eigenstates), run once per calculation, not the propagation.
#include <algorithm>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "api_common.hpp"

namespace espec {

// Upper triangle and diagonal of H (column-major, lda = n): ap1d.f:30-46 and
// ap2d.f:34-58 (stencil 30, -16, 1 times SHM per axis, V on the diagonal;
// x neighbours only inside the same grid row).  The strict lower triangle
// is not referenced by syevx with the upper triangle selected.
__global__ __launch_bounds__(256) void fd_dense(int ndim, int np0, int n, double shm0, double shm1,
                                                const double *vpot, double *a)
{
    const int j = blockIdx.x;  // column J (0-based)
    if (j >= n) return;
    const int l = j % np0;     // x index of J (L - 1)
    double *col = a + (long long)j * n;
    for (int i = threadIdx.x; i <= j; i += 256) {
        double v = 0.0;
        if (ndim == 1) {
            if (i == j) v = 30.0 * shm0 + vpot[i];
            else if (i == j - 1) v = -16.0 * shm0;
            else if (i == j - 2) v = shm0;
        } else {
            if (j == i + 2 * np0) v = shm1;
            else if (j == i + np0) v = -16.0 * shm1;
            else if (j == i + 2 && l >= 2) v = shm0;
            else if (j == i + 1 && l >= 1) v = -16.0 * shm0;
            else if (i == j) v = 30.0 * (shm0 + shm1) + vpot[i];
        }
        col[i] = v;
    }
}

}  // namespace espec

using namespace espec;

extern "C" {

int espec_mtrxdiag(int ndim, const int *np, int il, int iu, double abstol, const double *shm,
                   const double *vpot, double *eigvl, double *eigvc, int ldz, int *m)
{
    if ((ndim != 1 && ndim != 2) || !np || !shm || !vpot || !eigvl || !eigvc) return ESPEC_ERR_ARG;
    const int np0 = np[0], np1 = ndim == 2 ? np[1] : 1;
    if (np0 < 1 || np1 < 1) return ESPEC_ERR_ARG;
    const long long nn = (long long)np0 * np1;
    if (nn > 16384) return ESPEC_ERR_UNSUPPORTED;  // dense n x n in FP64: 2 GiB at 16384
    const int n = (int)nn;
    const int ns = iu - il + 1;
    if (il < 1 || iu > n || ns < 1 || ldz < n) return ESPEC_ERR_ARG;
    DevBuf da, dv, dw, dz, difail, dnev, dinfo;
    ESPEC_HIP_TRY(da.alloc(sizeof(double) * (size_t)n * n));
    ESPEC_HIP_TRY(dv.upload(vpot, sizeof(double) * n));
    ESPEC_HIP_TRY(dw.alloc(sizeof(double) * n));
    ESPEC_HIP_TRY(dz.alloc(sizeof(double) * (size_t)n * ns));
    ESPEC_HIP_TRY(difail.alloc(sizeof(int) * n));
    ESPEC_HIP_TRY(dnev.alloc(sizeof(int)));
    ESPEC_HIP_TRY(dinfo.alloc(sizeof(int)));
    hipLaunchKernelGGL(fd_dense, dim3(n), dim3(256), 0, nullptr, ndim, np0, n, shm[0],
                       ndim == 2 ? shm[1] : 0.0, dv.as<double>(), da.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return ESPEC_ERR_HIP;
    const rocblas_status st = rocsolver_dsyevx(
        h, rocblas_evect_original, rocblas_erange_index, rocblas_fill_upper, n, da.as<double>(), n,
        0.0, 0.0, il, iu, abstol, dnev.as<int>(), dw.as<double>(), dz.as<double>(), n,
        difail.as<int>(), dinfo.as<int>());
    (void)hipDeviceSynchronize();
    rocblas_destroy_handle(h);
    if (st != rocblas_status_success) return ESPEC_ERR_HIP;
    int info = 0, nev = 0;
    ESPEC_HIP_TRY(hipMemcpy(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(&nev, dnev.p, sizeof(int), hipMemcpyDeviceToHost));
    nev = std::min(nev, ns);
    if (m) *m = nev;
    ESPEC_HIP_TRY(hipMemcpy(eigvl, dw.p, sizeof(double) * nev, hipMemcpyDeviceToHost));
    for (int k = 0; k < nev; ++k)
        ESPEC_HIP_TRY(hipMemcpy(eigvc + (size_t)k * ldz, dz.as<double>() + (size_t)k * n,
                                sizeof(double) * n, hipMemcpyDeviceToHost));
    return info ? ESPEC_ERR_DIAG : ESPEC_OK;
}

}  // extern "C"
