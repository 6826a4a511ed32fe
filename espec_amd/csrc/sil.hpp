// Short-iterative Lanczos propagation on the device (PLNZ, LNZ, REORT, AU,
// the DSTEVX tridiagonal eigensolve and the DGEMV projections).
#pragma once
#include "common.hpp"

namespace espec {

constexpr int kMaxKrylov = 32;

// Device routine shared by the Lanczos kernels and espec_dstevx: implicit
// QL with Wilkinson shifts on an m x m symmetric tridiagonal (d, e); on
// return d holds the eigenvalues ascending and z (column-major, ld m) the
// eigenvectors.  Returns 0 or the index of a non-converged eigenvalue.
__device__ int tridiag_ql(int m, double *d, double *e, double *z);
// Sturm-count bisection of eigenvalue k (0-based) of (d, e) to the
// termination rule of dstebz.f; e2 = e^2, gl/gu the widened Gershgorin
// interval; returns the midpoint of the final interval.
__device__ double tridiag_bisect_one(int m, const double *d, const double *e2, double pivmin,
                                     double gl, double gu, double atoli, int k);
__device__ void tridiag_bounds(int m, const double *d, const double *e, double abstol, double *e2,
                               double *pivmin, double *gl, double *gu, double *atoli);

}  // namespace espec
