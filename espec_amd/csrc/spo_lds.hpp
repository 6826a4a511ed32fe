// Launch interface of the LDS-resident split-operator kernels.
#pragma once
#include "common.hpp"

namespace espec {

struct SpoLdsArgs {
    int ndim, n, npk;
    DevAxis ax[3];
    const void *tw;        // double2 twiddles, all axes
    const int *perm;       // digit-reversal maps, all axes
    int perm_off[3];
    void *x;               // [npk][n] state (complex64 mixed, complex128 FP64)
    const void *psi0;      // [npk][n] initial state for C(t), same precision
    const void *vph;       // [npk][n] double2 (cos, sin) of cst2 V, one table per packet
    const void *kin;       // double2 (cos, sin) of cst1 T (digit-reversed for the fast kernel):
                           // [n] shared, or [npk][n] when kin_pp
    int kin_pp;            // per-packet kinetic tables
    double b;              // 1/N
    int nsteps, lc, maxsamp;
    long long step0;       // steps done before this launch
    double *cbuf;          // [npk][maxsamp][3] = (Re C, Im C, |psi|^2)
    int vph_lds;           // stage the V phase table in LDS (fast kernel)
    int tab_lds;           // stage kinetic phases and twiddles in LDS (fast kernel)
    int ntw;               // twiddle entries, all axes
    const double *mask;    // [npk][n] .SMOOTHW mask after the closing V half phase, or null
};

size_t spo_lds_bytes(int n, bool ref, bool vph_lds, int nt, bool tab_lds = false, int ntw = 0);
hipError_t launch_spo_lds(const SpoLdsArgs &a, bool mixed, bool ref, int nt, hipStream_t s);
hipError_t launch_spo_sample(const void *x, const void *psi0, int n, int npk, bool mixed,
                             double *cbuf, int maxsamp, int idx, hipStream_t s);

}  // namespace espec
