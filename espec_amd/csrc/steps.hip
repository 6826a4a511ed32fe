// Single steps and the Gram-Schmidt pass on the device, for hosts that call
// the reference's finer-grained routines themselves:
//
//   espec_pspofftaux  PSPOFFTAUX (spofft.f:3-203): one split-operator step
//                     with the caller's kinetic grid TC (from DIGT) and
//                     constants CST1 (kinetic) and CST2 (half potential);
//   espec_psodfftaux  PSODFFTAUX (sodfft.f:3-129): one leapfrog step;
//   espec_reort       REORT (reort.f:1-83): one classical Gram-Schmidt pass
//                     of w against m stored vectors.
//
// The propagators (espec_pspofft, espec_psodfft, espec_plnz) keep the whole
// time loop on the device; these entry points serve a host that steps
// itself, one call per step, so each call moves its arrays to the device
// and back.  The FFTs are the multivariate line transforms of espec_fft
// (fft3d.f's definition, f32 twiddles in ESPEC_REF).
//
// REORT's projection r = V^T w runs on the FP64 matrix cores: each
// wavefront takes 16 basis vectors and accumulates
// v_mfma_f64_16x16x4f64 products over its share of the points, four
// points per instruction.  The B operand carries w in all 16 of its
// columns, so every column of the 16x16 result holds r for the 16
// vectors and only the row of each output register matters.
#include <algorithm>
#include <vector>

#include "api_common.hpp"
#include "common.hpp"

namespace espec {

hipError_t run_line_fft(double *re, double *im, int ntot, int n, int nspan, int isn, bool ref,
                        hipStream_t s);

// ------------------------------------------------------------ PSPOFFTAUX

// The half potential phase with the half swap of the flattened array
// (spofft.f:120-133 before the transforms; 178-191 after them, which also
// scales by B and, for .SMOOTHW, multiplies by the mask of the point where
// the value is stored, spofft.f:162-176).  .VOPTIC scales cos and sin of
// each phase by exp(CST2 W) (spofft.f:68-71, 100-103).  One thread per
// pair (i, i + nh).
template <bool POST>
__global__ __launch_bounds__(256) void aux_vhalf(int nh, double cst2, double b, double *u,
                                                 double *v, const double *vpot, int absorb,
                                                 const double *vabc)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nh; i += gridDim.x * 256) {
        const int p1 = POST ? i + nh : i, p2 = POST ? i : i + nh;
        double s1, c1, s2, c2;
        sincos(cst2 * vpot[p1], &s1, &c1);
        sincos(cst2 * vpot[p2], &s2, &c2);
        if (absorb == ESPEC_ABSORB_VOPTIC) {
            const double e1 = exp(cst2 * vabc[p1]), e2 = exp(cst2 * vabc[p2]);
            c1 *= e1;
            s1 *= e1;
            c2 *= e2;
            s2 *= e2;
        }
        const double ui = u[i], vi = v[i], uh = u[i + nh], vh = v[i + nh];
        double ua, va, un, vn;
        if (POST) {
            ua = c1 * b * ui + s1 * b * vi;
            va = c1 * b * vi - s1 * b * ui;
            un = c2 * b * uh + s2 * b * vh;
            vn = c2 * b * vh - s2 * b * uh;
            if (absorb == ESPEC_ABSORB_SMOOTHW) {
                un = vabc[i] * un;
                vn = vabc[i] * vn;
                ua = vabc[i + nh] * ua;
                va = vabc[i + nh] * va;
            }
        } else {
            ua = c1 * ui + s1 * vi;
            va = c1 * vi - s1 * ui;
            un = c2 * uh + s2 * vh;
            vn = c2 * vh - s2 * uh;
        }
        u[i] = un;
        v[i] = vn;
        u[i + nh] = ua;
        v[i + nh] = va;
    }
}

// exp(-i CST1 TC) in momentum space.  REF keeps the reference's update of
// V from the already rotated U (spofft.f:148-154); EXACT rotates.
template <bool REF>
__global__ __launch_bounds__(256) void aux_kinetic(int n, double cst1, const double *tc, double *u,
                                                   double *v)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double s, c;
        sincos(cst1 * tc[i], &s, &c);
        const double ui = u[i], vi = v[i];
        const double un = c * ui + s * vi;
        u[i] = un;
        v[i] = REF ? c * vi - s * un : c * vi - s * ui;
    }
}

// ------------------------------------------------------------ PSODFFTAUX

// psi(t+dt) = psi(t-dt) - 2 i CST H psi(t), H = V + F^-1 TC F: the V part
// and the shift of the time levels (sodfft.f:60-80), .VOPTIC adding
// CST W psi(t) (sodfft.f:66-72).
__global__ __launch_bounds__(256) void sod_pre(int n, double cst, double *u1, const double *u2,
                                               double *v1, const double *v2, const double *vpot,
                                               int absorb, const double *vabc, double *wr,
                                               double *wi)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (absorb == ESPEC_ABSORB_VOPTIC) {
            wr[i] = u1[i] + cst * vpot[i] * v2[i] + cst * vabc[i] * u2[i];
            wi[i] = v1[i] - cst * vpot[i] * u2[i] + cst * vabc[i] * v2[i];
        } else {
            wr[i] = u1[i] + cst * vpot[i] * v2[i];
            wi[i] = v1[i] - cst * vpot[i] * u2[i];
        }
        u1[i] = u2[i];
        v1[i] = v2[i];
    }
}

__global__ __launch_bounds__(256) void half_swap2(int nh, double *u, double *v)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nh; i += gridDim.x * 256) {
        double t = u[i];
        u[i] = u[i + nh];
        u[i + nh] = t;
        t = v[i];
        v[i] = v[i + nh];
        v[i + nh] = t;
    }
}

__global__ __launch_bounds__(256) void scale_tc(int n, const double *tc, double *u, double *v)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        u[i] = tc[i] * u[i];
        v[i] = tc[i] * v[i];
    }
}

// The kinetic part with the swap back (sodfft.f:96-127); the .SMOOTHW mask
// of EXACT mode (the reference's own mask branch is unreachable,
// sodfft.f:108).
__global__ __launch_bounds__(256) void sod_post(int nh, double cst, double b, double *u2,
                                                double *v2, const double *wr, const double *wi,
                                                int mask, const double *vabc)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nh; i += gridDim.x * 256) {
        const double ui = u2[i], vi = v2[i], uh = u2[i + nh], vh = v2[i + nh];
        double ua = wr[i + nh] + cst * b * vi;
        double va = wi[i + nh] - cst * b * ui;
        double un = wr[i] + cst * b * vh;
        double vn = wi[i] - cst * b * uh;
        if (mask) {
            un = vabc[i] * un;
            vn = vabc[i] * vn;
            ua = vabc[i + nh] * ua;
            va = vabc[i + nh] * va;
        }
        u2[i] = un;
        v2[i] = vn;
        u2[i + nh] = ua;
        v2[i + nh] = va;
    }
}

// ------------------------------------------------------------------ REORT

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kReortWaves = 4;      // wavefronts per workgroup (point ranges)
constexpr int kReortBlocks = 64;    // workgroups per group of 16 vectors

// Partial projections of w on 16 basis columns [g16*16, g16*16+16): block
// (x, g16) sums points p = 4 q + k over its share of q with MFMA, wave by
// wave, then writes its 16 partials to part[g16][x][16].  Lane l feeds
// A[l % 16][l / 16] = V[p0 + l / 16][j0 + l % 16] and B[l / 16][*] =
// w[p0 + l / 16]; in the FP64 16x16x4 layout, output register r of lane l
// holds row 4 r + l / 16 (column l % 16).
__global__ __launch_bounds__(64 * kReortWaves) void reort_project(int n, int m, int ld,
                                                                  const double *w,
                                                                  const double *basis,
                                                                  double *part)
{
    __shared__ double red[kReortWaves][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g16 = blockIdx.y, j = g16 * 16 + (lane & 15), k = lane >> 4;
    const bool jok = j < m;
    const double *col = basis + (long long)(jok ? j : 0) * ld;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    const int nq = (n + 3) / 4;
    for (int q = blockIdx.x * kReortWaves + wv; q < nq; q += gridDim.x * kReortWaves) {
        const int p = 4 * q + k;
        const bool pok = p < n;
        const double a = (jok && pok) ? col[p] : 0.0;
        const double bw = pok ? w[p] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bw, acc, 0, 0, 0);
    }
    // Column 0 of the result (lanes 0, 16, 32, 48) holds rows 4 r + l/16.
    if ((lane & 15) == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wv][4 * r + (lane >> 4)] = acc[r];
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < kReortWaves; ++q) s += red[q][threadIdx.x];
        part[((long long)g16 * gridDim.x + blockIdx.x) * 16 + threadIdx.x] = s;
    }
}

// r_j = sum of the block partials in block order; one thread per column.
__global__ __launch_bounds__(256) void reort_finish(int m, int nblk, const double *part,
                                                    double *r)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    const int g16 = j / 16, c = j % 16;
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += part[((long long)g16 * nblk + b) * 16 + c];
    r[j] = s;
}

// w_i <- w_i - r_1 V_i1 - r_2 V_i2 - ... in the reference's order (reort.f:69-78).
__global__ __launch_bounds__(256) void reort_update(int n, int m, int ld, const double *r,
                                                    const double *basis, double *w)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double x = w[i];
        for (int j = 0; j < m; ++j) x = x - r[j] * basis[(long long)j * ld + i];
        w[i] = x;
    }
}

inline dim3 grid_for(int n) { return dim3((unsigned)std::max(1, std::min(1024, (n + 255) / 256))); }

}  // namespace espec

using namespace espec;

extern "C" {

int espec_pspofftaux(int ndim, const int *np, double b, double cst1, double cst2,
                     const double *tc, double *u, double *v, const double *vpot, int absorb,
                     const double *vabc, int flags)
{
    if (ndim < 1 || ndim > 3 || !np || !tc || !u || !v || !vpot) return ESPEC_ERR_ARG;
    if (absorb != ESPEC_ABSORB_NONE && (!vabc || (absorb != ESPEC_ABSORB_VOPTIC &&
                                                  absorb != ESPEC_ABSORB_SMOOTHW)))
        return ESPEC_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (np[a] < 1) return ESPEC_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (!is_pow2(np[a]) || np[a] > 4096) return ESPEC_ERR_UNSUPPORTED;
    const int n = ntot(ndim, np), nh = n / 2;
    const bool ref = (flags & ESPEC_REF) != 0;
    const size_t by = sizeof(double) * n;
    DevBuf du, dv, dt, dp, da;
    ESPEC_HIP_TRY(du.upload(u, by));
    ESPEC_HIP_TRY(dv.upload(v, by));
    ESPEC_HIP_TRY(dt.upload(tc, by));
    ESPEC_HIP_TRY(dp.upload(vpot, by));
    if (absorb) ESPEC_HIP_TRY(da.upload(vabc, by));
    const double *va = absorb ? da.as<double>() : nullptr;
    if (nh > 0) {
        hipLaunchKernelGGL(aux_vhalf<false>, grid_for(nh), dim3(256), 0, nullptr, nh, cst2, b,
                           du.as<double>(), dv.as<double>(), dp.as<double>(), absorb, va);
        ESPEC_HIP_TRY(hipGetLastError());
    }
    int span = 1;
    for (int a = 0; a < ndim; ++a) {
        span *= np[a];
        ESPEC_HIP_TRY(run_line_fft(du.as<double>(), dv.as<double>(), n, np[a], span, 1, ref,
                                   nullptr));
    }
    if (ref)
        hipLaunchKernelGGL(aux_kinetic<true>, grid_for(n), dim3(256), 0, nullptr, n, cst1,
                           dt.as<double>(), du.as<double>(), dv.as<double>());
    else
        hipLaunchKernelGGL(aux_kinetic<false>, grid_for(n), dim3(256), 0, nullptr, n, cst1,
                           dt.as<double>(), du.as<double>(), dv.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    span = 1;
    for (int a = 0; a < ndim; ++a) {
        span *= np[a];
        ESPEC_HIP_TRY(run_line_fft(du.as<double>(), dv.as<double>(), n, np[a], span, -1, ref,
                                   nullptr));
    }
    if (nh > 0) {
        hipLaunchKernelGGL(aux_vhalf<true>, grid_for(nh), dim3(256), 0, nullptr, nh, cst2, b,
                           du.as<double>(), dv.as<double>(), dp.as<double>(), absorb, va);
        ESPEC_HIP_TRY(hipGetLastError());
    }
    ESPEC_HIP_TRY(hipMemcpy(u, du.p, by, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(v, dv.p, by, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_psodfftaux(int ndim, const int *np, double b, double cst, const double *tc,
                     double *u1, double *u2, double *v1, double *v2, const double *vpot,
                     int absorb, const double *vabc, int flags)
{
    if (ndim < 1 || ndim > 3 || !np || !tc || !u1 || !u2 || !v1 || !v2 || !vpot)
        return ESPEC_ERR_ARG;
    if (absorb != ESPEC_ABSORB_NONE && (!vabc || (absorb != ESPEC_ABSORB_VOPTIC &&
                                                  absorb != ESPEC_ABSORB_SMOOTHW)))
        return ESPEC_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (np[a] < 1) return ESPEC_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (!is_pow2(np[a]) || np[a] > 4096) return ESPEC_ERR_UNSUPPORTED;
    const int n = ntot(ndim, np), nh = n / 2;
    const bool ref = (flags & ESPEC_REF) != 0;
    const size_t by = sizeof(double) * n;
    DevBuf d1, d2, e1, e2, dt, dp, da, wr, wi;
    ESPEC_HIP_TRY(d1.upload(u1, by));
    ESPEC_HIP_TRY(d2.upload(u2, by));
    ESPEC_HIP_TRY(e1.upload(v1, by));
    ESPEC_HIP_TRY(e2.upload(v2, by));
    ESPEC_HIP_TRY(dt.upload(tc, by));
    ESPEC_HIP_TRY(dp.upload(vpot, by));
    ESPEC_HIP_TRY(wr.alloc(by));
    ESPEC_HIP_TRY(wi.alloc(by));
    if (absorb) ESPEC_HIP_TRY(da.upload(vabc, by));
    const double *va = absorb ? da.as<double>() : nullptr;
    hipLaunchKernelGGL(sod_pre, grid_for(n), dim3(256), 0, nullptr, n, cst, d1.as<double>(),
                       d2.as<double>(), e1.as<double>(), e2.as<double>(), dp.as<double>(), absorb,
                       va, wr.as<double>(), wi.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    if (nh > 0) {
        hipLaunchKernelGGL(half_swap2, grid_for(nh), dim3(256), 0, nullptr, nh, d2.as<double>(),
                           e2.as<double>());
        ESPEC_HIP_TRY(hipGetLastError());
    }
    int span = 1;
    for (int a = 0; a < ndim; ++a) {
        span *= np[a];
        ESPEC_HIP_TRY(run_line_fft(d2.as<double>(), e2.as<double>(), n, np[a], span, 1, ref,
                                   nullptr));
    }
    hipLaunchKernelGGL(scale_tc, grid_for(n), dim3(256), 0, nullptr, n, dt.as<double>(),
                       d2.as<double>(), e2.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    span = 1;
    for (int a = 0; a < ndim; ++a) {
        span *= np[a];
        ESPEC_HIP_TRY(run_line_fft(d2.as<double>(), e2.as<double>(), n, np[a], span, -1, ref,
                                   nullptr));
    }
    if (nh > 0) {
        const int mask = absorb == ESPEC_ABSORB_SMOOTHW && !ref;
        hipLaunchKernelGGL(sod_post, grid_for(nh), dim3(256), 0, nullptr, nh, cst, b,
                           d2.as<double>(), e2.as<double>(), wr.as<double>(), wi.as<double>(),
                           mask, va);
        ESPEC_HIP_TRY(hipGetLastError());
    }
    ESPEC_HIP_TRY(hipMemcpy(u1, d1.p, by, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(u2, d2.p, by, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(v1, e1.p, by, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(v2, e2.p, by, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_reort(int n, int m, int ld, double *w, const double *basis)
{
    if (n < 0 || m < 0 || !w || (m > 0 && (!basis || ld < n))) return ESPEC_ERR_ARG;
    if (n == 0 || m == 0) return ESPEC_OK;
    DevBuf dw, db, dpart, dr;
    ESPEC_HIP_TRY(dw.upload(w, sizeof(double) * n));
    // the m columns of the caller's (ld x m) array, leading dimension kept
    ESPEC_HIP_TRY(db.upload(basis, sizeof(double) * ((size_t)ld * (m - 1) + n)));
    const int ng = (m + 15) / 16;
    const int nq = (n + 3) / 4;
    const int nblk = std::max(1, std::min(kReortBlocks, (nq + kReortWaves - 1) / kReortWaves));
    ESPEC_HIP_TRY(dpart.alloc(sizeof(double) * 16 * (size_t)nblk * ng));
    ESPEC_HIP_TRY(dr.alloc(sizeof(double) * m));
    hipLaunchKernelGGL(reort_project, dim3(nblk, ng), dim3(64 * kReortWaves), 0, nullptr, n, m,
                       ld, dw.as<double>(), db.as<double>(), dpart.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reort_finish, dim3((m + 255) / 256), dim3(256), 0, nullptr, m, nblk,
                       dpart.as<double>(), dr.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reort_update, grid_for(n), dim3(256), 0, nullptr, n, m, ld,
                       dr.as<double>(), db.as<double>(), dw.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(w, dw.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

}  // extern "C"
