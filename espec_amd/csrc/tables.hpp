// Host-side tables for the device propagators: FFT stage plans with
// twiddles, digit-reversal maps, kinetic-energy grids and phase tables.
// These are computed once per plan in FP64 on the host and uploaded.
#pragma once
#include <cstdint>
#include <vector>

namespace espec {

constexpr double kShfs = 2.4188843243e-2;  // fs per a.u. of time (pspofft.f:76)
constexpr double kTwoPi = 6.283185307179586476925287;
constexpr int kMaxStages = 24;

// One decimation-in-frequency plan for a power-of-two length n.  Stage s
// works on blocks of length L_s = n / (r_0 ... r_{s-1}) with sub-length
// m_s = L_s / r_s; output q of a butterfly at offset j is multiplied by
// the twiddle tw[off_s + (q-1) m_s + j] (angle q j RAD / L_s for the
// exp(+i) transform).  The radix order is the one Singleton's FFT uses
// (fft3d.f:124-173) so that, in reference mode, the f32-rounded
// RAD = 6.2831853071796 (fft3d.f:119) enters through the same twiddles.
struct FftPlan {
    int n = 0, nst = 0;
    int rad[kMaxStages] = {}, lm[kMaxStages] = {}, lL[kMaxStages] = {},
        off[kMaxStages] = {};
    std::vector<double> twc, tws;  // cos/sin of the exp(+i) twiddles
    std::vector<int> perm;         // perm[k] = position of X[k] after all stages
};

// Radix sequence of Singleton's FFT for n = 2^p: square factors 4 (per 16),
// a square factor 2 when 8 remains, the square-free middle factor, then the
// square factors mirrored.
std::vector<int> singleton_factors(int n);
bool is_pow2(int n);
int ilog2(int n);

// ref = true reproduces the reference twiddles: f32-rounded RAD, and the
// radix-2 stages' reuse of the twiddle of j for m - j via cos -> -cos
// (fft3d.f:206-211).
FftPlan make_fft_plan(int n, bool ref);

// Kinetic grid p^2/(2m) per axis in natural (FFT output) order.  ref = true
// accumulates p with the 10-digit 2*pi = 6.283185307 and flips the sign at
// n/2 as DIGT does for 1-D (dkeft.f:267-282); otherwise the closed form
// p_k = (k - n [k >= n/2]) 2 pi / (n h).
std::vector<double> kinetic_axis(int n, double sh, double sm, bool ref);
// Full flattened grid tc[i] = sum_axes T_a(k_a), x fastest.
std::vector<double> kinetic_grid(int ndim, const int *np, const double *sh,
                                 const double *sm, bool ref);

// Fortran 77 trip count of DO T=M1,M2,M3 (pspofft.f:110).
int fortran_trip(double m1, double m2, double m3);

}  // namespace espec
