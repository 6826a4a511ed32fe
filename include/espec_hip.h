/*
 * espec_hip.h -- C ABI of libespec_hip.so, the MI355X (gfx950) wave-packet
 * propagation library that stands in for eSPec's hot-path subroutines.
 *
 * Every entry point takes plain host pointers and sizes (no device or
 * framework types) so that the Fortran host can bind it with
 * ISO_C_BINDING.  The arrays follow the reference's conventions: psi is
 * split into real and imaginary REAL*8 vectors U, V; multi-dimensional
 * grids are flattened with x fastest (gvpot.f:80); times are in fs and
 * converted with 2.4188843243e-2 fs per a.u. (pspofft.f:76); masses in
 * electron masses; lengths in bohr.
 *
 * Return value: 0 on success (INFO = 0); 1 reproduces the reference's
 * drift flag (pspofft.f:137-142); >= 2 are library errors, see
 * espec_strerror().  No entry point ever exits the process.
 *
 * Flags select the arithmetic:
 *   ESPEC_FP64   complex128 state, FP64 everything;
 *   ESPEC_MIXED  complex64 state in HBM (psi, psi0, Krylov basis), rounded
 *                once per launch or pass; the strided transforms of the
 *                multi-pass split-operator plans compute in FP32 with every
 *                constant split into float hi + lo, everything else in FP64
 *                (row passes and LDS-resident steps, the potential phase,
 *                stencils, norm, C(t), Lanczos alpha/beta and Gram
 *                coefficients): a constant or phase rounded once to float
 *                would scale every step the same way, accumulate
 *                coherently and miss 1e-6 in C(t);
 *   ESPEC_REF    the reference's exact semantics (f32-rounded 2*pi in the
 *                FFT twiddles, the accumulated kinetic grid, the
 *                non-unitary kinetic and Lanczos recombination updates)
 *                so that FP64 results match the reference build;
 *   ESPEC_EXACT  the same algorithms with those defects removed.
 */
#ifndef ESPEC_HIP_H
#define ESPEC_HIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ESPEC_FP64  0x0
#define ESPEC_MIXED 0x1
#define ESPEC_EXACT 0x0
#define ESPEC_REF   0x2
/* Plans only: sm is [npk][ndim], one set of masses per packet (the
 * kinetic phases become per-packet tables). */
#define ESPEC_PER_PACKET_MASS 0x4
/* Split-operator plans only: one pass per axis per step through HBM even
 * when a packet fits one workgroup's LDS (batches of many small packets,
 * where the LDS-resident kernel has one workgroup per packet). */
#define ESPEC_MULTIPASS 0x8
/* Short-iterative Lanczos plans only: launch every kernel of a step
 * (fixed-order, adaptive and trace steps) directly on the plan's stream
 * instead of replaying captured step graphs (same kernels, same results;
 * for kernel tracers that do not follow graph launches). */
#define ESPEC_NO_GRAPH 0x10

enum {
    ESPEC_OK = 0,
    ESPEC_DRIFT = 1,          /* energy/norm drift above TOL (reference INFO=1) */
    ESPEC_ERR_ARG = 2,        /* invalid argument (size, dimension, radix) */
    ESPEC_ERR_HIP = 3,        /* HIP runtime failure */
    ESPEC_ERR_NOMEM = 4,      /* device allocation failed */
    ESPEC_ERR_DIAG = 5,       /* tridiagonal eigensolve failed (sil.f:118-121) */
    ESPEC_ERR_UNSUPPORTED = 6 /* combination not implemented on the device */
};

const char *espec_strerror(int code);
/* Number of HIP devices the library sees (0 when no GPU). */
int espec_device_count(void);
/* Library version string. */
const char *espec_version(void);

/* ---------------------------------------------------------------------
 * Split-operator FFT propagation.  Replaces PSPOFFT (pspofft.f:1-3),
 * PSPOFFTAUX (spofft.f:3-4), DIGT (dkeft.f:204) and the C(t) sampling of
 * FCORREL (fcorrel.f:1) inside the time loop.
 *   ndim, np[ndim]     grid (DIM .1D/.2D/.3D and NP); powers of two
 *   dt, ti, tf         time step, initial and final time in fs; the step
 *                      count is the reference REAL DO trip count of
 *                      DO T=TI+DT,TF+DT,DT (pspofft.f:110)
 *   lc                 C(t) every lc steps (LC argument)
 *   sh[ndim], sm[ndim] grid step (bohr) and mass (m_e) per axis
 *   u1, v1 [N]         in: psi(TI); out: psi after the last step
 *   vpot [N]           potential (hartree)
 *   cr, ci [maxsamp]   out: C(t) samples, sample 0 = C(TI); these are the
 *                      LNZVC(NC,MPR) / LNZVC(NC,MPI) columns of the reference
 *   nrm [maxsamp]      out (may be NULL): sum |psi|^2 at each sample
 *   nsamp              out: samples written
 * ------------------------------------------------------------------- */
int espec_pspofft(int ndim, const int *np, double dt, double ti, double tf,
                  int lc, const double *sh, const double *sm, double *u1,
                  double *v1, const double *vpot, double *cr, double *ci,
                  double *nrm, int maxsamp, int *nsamp, int flags);

/* Absorbing boundaries: PSPOFFT/PSODFFT with ABSORB=.TRUE.; TPABSOR picks
 * the kind and VABC carries its per-point data (absorb.f:1-100):
 *   ESPEC_ABSORB_VOPTIC   vabc = imaginary potential W (hartree, <= 0 to
 *                         absorb): both V half phases are scaled by
 *                         exp(cst2 W) (spofft.f:65-118); the leapfrog adds
 *                         cst W psi(t) (sodfft.f:66-72);
 *   ESPEC_ABSORB_SMOOTHW  vabc = mask in [0, 1]: psi is multiplied by it
 *                         after the closing V half phase (spofft.f:162-176)
 *                         or after the leapfrog update (sodfft.f:108-116).
 *                         The reference's SOD never takes that branch (its
 *                         test compares TPABSOR(1:7) with '.SMOOTHW'), so
 *                         with ESPEC_REF PSODFFT runs without absorber. */
enum {
    ESPEC_ABSORB_NONE = 0,
    ESPEC_ABSORB_VOPTIC = 1,
    ESPEC_ABSORB_SMOOTHW = 2
};
/* Replaces ABSORBINGBC (absorb.f:1-130): the VABC array of an absorber
 * shape on the grid (x fastest), X = XI + (I-1) SH per axis.
 *   kind  ESPEC_ABG_SMOOTHW (.SMOOTHW mask, 1-D and 2-D), ESPEC_ABG_SMRS
 *         (.SMRS mask), ESPEC_ABG_VOPTIC_LINEAR, _KOSLOFF, _POLIN
 *         (imaginary potentials), the last four 1-D only
 *   al, ar, sh, xi, xf [ndim]  the reference's arguments
 *   voi   [ndim], except [2] for the 1-D .SMOOTHW, _KOSLOFF and _POLIN,
 *         which read VOI(2) as an exponent or width
 *   alpha out (may be NULL): 1 - 0.1 (XF - AR) for the 1-D .SMOOTHW
 * ESPEC_ERR_UNSUPPORTED where the reference stops (3-D, 2-D non-SMOOTHW). */
enum {
    ESPEC_ABG_SMOOTHW = 1,
    ESPEC_ABG_SMRS = 2,
    ESPEC_ABG_VOPTIC_LINEAR = 3,
    ESPEC_ABG_VOPTIC_KOSLOFF = 4,
    ESPEC_ABG_VOPTIC_POLIN = 5
};
int espec_absorbingbc(int ndim, const int *np, int kind, const double *al,
                      const double *ar, const double *sh, const double *xi,
                      const double *xf, const double *voi, double *vabc,
                      double *alpha);

/* espec_pspofft with an absorber; vabc [N] (may be NULL for NONE). */
int espec_pspofft_absorb(int ndim, const int *np, double dt, double ti,
                         double tf, int lc, const double *sh, const double *sm,
                         double *u1, double *v1, const double *vpot,
                         double *cr, double *ci, double *nrm, int maxsamp,
                         int *nsamp, int flags, int absorb, const double *vabc);

/* Replaces PSODFFT (psodfft.f:1-3) / PSODFFTAUX (sodfft.f:3-4): second-
 * order-differencing leapfrog with the FFT kinetic operator, including
 * the two half-strength start-up steps (psodfft.f:120-123).  Arguments as
 * espec_pspofft; u1/v1 return psi at the last step. */
int espec_psodfft(int ndim, const int *np, double dt, double ti, double tf,
                  int lc, const double *sh, const double *sm, double *u1,
                  double *v1, const double *vpot, double *cr, double *ci,
                  double *nrm, int maxsamp, int *nsamp, int flags);
/* espec_psodfft with an absorber (sodfft.f:66-72, 108-116); vabc [N]. */
int espec_psodfft_absorb(int ndim, const int *np, double dt, double ti,
                         double tf, int lc, const double *sh, const double *sm,
                         double *u1, double *v1, const double *vpot,
                         double *cr, double *ci, double *nrm, int maxsamp,
                         int *nsamp, int flags, int absorb, const double *vabc);

/* Replaces PSPOFFTAUX(ABSORB,TPABSOR,NDIM,N,B,B1,B2,CST1,CST2,NP,U,V,TC,VPOT,
 * VABC) (spofft.f:3-4): one split-operator step of u + i v with the caller's
 * kinetic grid tc [N] (as DIGT fills it), kinetic constant cst1 (DT/SHFS)
 * and half-potential constant cst2, scaled by b; absorb ESPEC_ABSORB_* with
 * vabc [N] (.VOPTIC potential or .SMOOTHW mask, spofft.f:65-118,162-176).
 * ESPEC_REF keeps the reference's kinetic update and f32 twiddles.  B1, B2
 * are not read by the reference.  Runs on the device. */
int espec_pspofftaux(int ndim, const int *np, double b, double cst1, double cst2,
                     const double *tc, double *u, double *v, const double *vpot, int absorb,
                     const double *vabc, int flags);

/* Replaces PSODFFTAUX(ABSORB,TPABSOR,NDIM,N,B,CST,NP,U1,U2,V1,V2,TC,VPOT,VABC,
 * WRKR,WRKI) (sodfft.f:3-4): one leapfrog step; (u1, v1) = psi(t - dt) on
 * entry and psi(t) on return, (u2, v2) = psi(t) on entry and psi(t + dt) on
 * return.  The work arrays are the library's.  .SMOOTHW masks in
 * ESPEC_EXACT only (the reference's mask branch is unreachable,
 * sodfft.f:108).  Runs on the device. */
int espec_psodfftaux(int ndim, const int *np, double b, double cst, const double *tc,
                     double *u1, double *u2, double *v1, double *v2, const double *vpot,
                     int absorb, const double *vabc, int flags);

/* Replaces REORT(N,M,LMTREORT,MXDCT,WORK,EIGVC) (reort.f:1): one classical
 * Gram-Schmidt pass, r = V^T w on the FP64 matrix cores, then w -= V r in
 * the reference's column order; basis is column-major with leading
 * dimension ld >= n (MXDCT).  Runs on the device. */
int espec_reort(int n, int m, int ld, double *w, const double *basis);

/* Replaces FFTNT(X,Y,N,M,ITYPE) (cfft.f:6): radix-2 DIF transform of
 * n = 2^m points, itype = -1 backward with 1/N; ESPEC_REF keeps the f32
 * TWOPI (cfft.f:35).  Runs on the device. */
int espec_fftnt(double *x, double *y, int n, int m, int itype, int flags);

/* Replaces FFT(A,B,NTOT,N,NSPAN,ISN) (fft3d.f:1): one variate of an in-place
 * multivariate transform, isn > 0 -> exp(+i ...), unnormalised; n a power
 * of two.  Runs on the device. */
int espec_fft(double *a, double *b, int ntot, int n, int nspan, int isn,
              int flags);

/* Replaces DIGT (dkeft.f:204): kinetic grid tc[N] = p^2/(2m) on the FFT
 * momentum grid and b = 1/N (host-side table; ESPEC_REF reproduces the
 * accumulated 1-D grid with 2*pi = 6.283185307). */
int espec_digt(int ndim, const int *np, const double *sh, const double *sm,
               double *tc, double *b, int flags);

/* ---------------------------------------------------------------------
 * Short-iterative Lanczos propagation.  Replaces PLNZ (plnz.f:1-4) with
 * PLNZAUX (sil.f:3-5), LNZ (lnz.f:4-5), REORT (reort.f:1), the DSTEVX
 * tridiagonal eigensolve (dstevx.f:5) and the DGEMV projections.
 *   mp      Krylov order (MP), 2..32
 *   abstol  DSTEVX tolerance: <= 0 selects the QL path (dsteqr.f), > 0 the
 *           bisection path whose eigenvalues are midpoints of Sturm
 *           intervals of width < abstol (dstebz.f)
 *   shm[ndim] FD kinetic prefactor SHM per axis (espec.f:630)
 * ------------------------------------------------------------------- */
int espec_plnz(int ndim, const int *np, int mp, double abstol, double dt,
               double ti, double tf, int lc, const double *shm, double *u,
               double *v, const double *vpot, double *cr, double *ci,
               double *nrm, int maxsamp, int *nsamp, int flags);

/* espec_plnz with PLNZ's ABSORB: absorb ESPEC_ABSORB_SMOOTHW multiplies the
 * state by vabc [N] after every step, before C(t) is sampled, as PLNZAUX
 * does for TPABSOR '.SMOOTHW' (sil.f:171-180); ESPEC_ABSORB_VOPTIC and
 * ESPEC_ABSORB_NONE propagate without absorber (sil.f tests only
 * .SMOOTHW). */
int espec_plnz_absorb(int ndim, const int *np, int mp, double abstol, double dt,
                      double ti, double tf, int lc, const double *shm, double *u,
                      double *v, const double *vpot, double *cr, double *ci,
                      double *nrm, int maxsamp, int *nsamp, int flags, int absorb,
                      const double *vabc);

/* ---------------------------------------------------------------------
 * Pulse-driven short-iterative Lanczos.  Replaces PPLNZ (pplnz.f:1-5):
 * every step of DO T=TI+DT,TF+DT/2,DT (pplnz.f:190) first forms
 * VEFF = VP - DM EF(T) (CHPOT, chpot.f:58-62) with the field at the step's
 * end time T, then takes one PLNZAUX step (sil.f:3-5) with VEFF.
 *   vp, dm [N]           static potential and dipole function
 *   efc .. kl            EF's pulse (ef.f): ESPEC_PULSE_*, E0, T0, TD, TP,
 *                        OMG, SNI, KL
 *   veff [N]             out (may be NULL): the last VEFF, which PPLNZ
 *                        leaves in VPOT
 *   cr, ci, nrm          C(t) as espec_plnz; sample 0 is C(TI) here, where
 *                        the reference stores it only when PRTCRL is .YES
 *                        or .PARTIAL (pplnz.f:130-134)
 * Other arguments as espec_plnz.
 * ------------------------------------------------------------------- */
enum {
    ESPEC_PULSE_NONE = 0,  /* EFC .NONE / .NULL */
    ESPEC_PULSE_GAUS = 1,  /* .GAUS  */
    ESPEC_PULSE_GGAUS = 2, /* .GGAUS */
    ESPEC_PULSE_SIN2 = 3   /* .SIN2  */
};
int espec_pplnz(int ndim, const int *np, int mp, double abstol, double dt,
                double ti, double tf, int lc, const double *shm, double *u,
                double *v, const double *vp, const double *dm, int efc,
                double e0, double t0, double td, double tp, double omg,
                double sni, int kl, double *veff, double *cr, double *ci,
                double *nrm, int maxsamp, int *nsamp, int flags);
/* espec_pplnz with ABSORB: PPLNZ hands ABSORB/TPABSOR to PLNZAUX
 * (pplnz.f:208), which masks the state for .SMOOTHW (sil.f:171-180); as
 * espec_plnz_absorb. */
int espec_pplnz_absorb(int ndim, const int *np, int mp, double abstol, double dt,
                       double ti, double tf, int lc, const double *shm, double *u,
                       double *v, const double *vp, const double *dm, int efc,
                       double e0, double t0, double td, double tp, double omg,
                       double sni, int kl, double *veff, double *cr, double *ci,
                       double *nrm, int maxsamp, int *nsamp, int flags, int absorb,
                       const double *vabc);
/* Replaces S2PPSOD (s2ppsod.f:1-1143): two surfaces A and B (u, v, vp, dm
 * [2N], A first) propagated by the FD leapfrog of PSODAUX (sod.f:3-122)
 * with VEFF = VP - DM E(T) per surface when e0 != 0, coupled by the
 * routine's active coordinate coupling (TPCRS = '.CORD', s2ppsod.f:95):
 * after each step UA += w VB(t), VA -= w UB(t), UB += w VA(t), VB -= w UA(t)
 * with w = c (CST aa / sm1) XD_i, c = 1/4, 1/2, 1 for the quarter, half and
 * full steps and XD_i = xi + (i+1) sh accumulated as the reference does.
 * vabc [N] serves both surfaces (s2ppsod.f passes VABC(1) for both).
 * prtcrl 0 .NO, 1 .PARTIAL, 2 .YES selects the printed rows; tab
 * [maxrows][12] receives them as the reference prints them (T, ET, ET0,
 * CRA, CIA, CTA, CRB, CIB, CTB, FNTA, FNTB, FNT, with its FNTB**2 at TI and
 * CTB in the CRB column after the start-up), *nrows their count, *info 1
 * where the norm drifts by more than tol percent (no absorber).  u, v
 * return psi(TF); uprev, vprev (may be NULL) psi(TF - DT); veff (may be
 * NULL) [2N] VPOT as the reference leaves it.  FP64 only. */
int espec_s2ppsod(int ndim, const int *np, double dt, double ti, double tf, int lc,
                  const double *shm, double xi, double sh, double sm1, double aa, double *u,
                  double *v, const double *vp, const double *dm, int efc, double e0, double t0,
                  double td, double tp, double omg, double sni, int kl, int absorb,
                  const double *vabc, int prtcrl, double tol, double *tab, int maxrows,
                  int *nrows, int *info, double *uprev, double *vprev, double *veff, int flags);

/* ---------------------------------------------------------------------
 * Pulse-driven finite-difference leapfrog.  Replaces PPSOD (ppsod.f:1-548)
 * with its step PSODAUX (sod.f:3-122): psi(t+dt) = psi(t-dt) - i CST H psi(t),
 * CST = 2 DT/SHFS, H the FD operator of AU with VEFF = VP - DM EF(T) at each
 * step's end time; start-up by a quarter- and a half-strength step at the
 * fields of (TI+DT)/2 and TI+DT, then NINT((TF+DT/2-TI-DT)/DT) steps.
 *   u, v            in: psi0; out: psi(TF) (U2, V2)
 *   uprev, vprev    out (may be NULL): psi(TF - DT) (U1, V1)
 *   absorb, vabc    ESPEC_ABSORB_VOPTIC / _SMOOTHW (sod.f:55-120); .SMRS is
 *                   the .SMOOTHW form
 *   npr, proj       0..6 real vectors [npr][N] to project on (W1..W6)
 *   veff            out (may be NULL): the last VEFF (VPOT on return)
 *   cr, ci, nrm     C(t) = <psi0|psi>, <psi|psi> at sample 0 (TI) and then
 *                   every lc steps, the start-up counting as one step (the
 *                   reference stores these when PRTCRL prints)
 *   et, et0         out (may be NULL): <psi|H|psi> with VEFF and with VP
 *                   (EIGENERG of U2 plus V2) at each sample
 *   pr              out (may be NULL): [nsamp][npr][2] (<W,U>, <W,V>)
 * FP64 only (ESPEC_MIXED returns ESPEC_ERR_UNSUPPORTED).  Runs on the
 * device. */
int espec_ppsod(int ndim, const int *np, double dt, double ti, double tf, int lc,
                const double *shm, double *u, double *v, const double *vp,
                const double *dm, int efc, double e0, double t0, double td,
                double tp, double omg, double sni, int kl, int absorb,
                const double *vabc, int npr, const double *proj, double *uprev,
                double *vprev, double *veff, double *cr, double *ci, double *nrm,
                double *et, double *et0, double *pr, int maxsamp, int *nsamp,
                int flags);

/* ---------------------------------------------------------------------
 * Gaussian initial packets.  Replaces INIT_COND (init_cond.f:1-200) with
 * GET_GAUSS (gauss.f:1-20): rnorm exp(-ln2 (c/rA)^2) (cos, sin)(k c) per
 * axis, c = XI + (i-1) SH - X0, normalised over the axis.
 *   mode ESPEC_INIT_CALC   1-D packet (U = cos g, V = sin g), or the 2-D
 *                          product the reference forms (U + iV = the
 *                          complex conjugate of the product of the 1-D
 *                          packets)
 *        ESPEC_INIT_GETR   2-D: the real eigenvector u[0..np0) along x
 *                          times a Gaussian along y (.GETR, .GETR1)
 *        ESPEC_INIT_GETR2  eigenvector u[0..np1) along y, Gaussian along x
 *        ESPEC_INIT_GETC   complex eigenvector u + iv along x (.GETC, .GETC1)
 *        ESPEC_INIT_GETC2  complex eigenvector along y
 *   xi, sh, x0, ra, rk [ndim]  grid start and step, centre, HWHM and wave
 *                              number per axis (a.u.)
 *   u, v [N]               out (in: the eigenvector for the .GET* modes)
 * ------------------------------------------------------------------- */
enum {
    ESPEC_INIT_CALC = 0,
    ESPEC_INIT_GETR = 1,
    ESPEC_INIT_GETR2 = 2,
    ESPEC_INIT_GETC = 3,
    ESPEC_INIT_GETC2 = 4
};
int espec_init_cond(int ndim, const int *np, int mode, const double *xi,
                    const double *sh, const double *x0, const double *ra,
                    const double *rk, double *u, double *v);

/* Replaces EF (ef.f:1-60): the field at time t (fs); NaN for an unknown
 * efc, where the reference stops. */
double espec_ef(int efc, double e0, double t0, double td, double tp,
                double omg, double sni, int kl, double t);

/* Replaces AU(DIM,N,NP,SHM,VPOT,U,WORK,VAR) (au.f:3) -> AU_1D (au1d.f:1),
 * AU_2D (au2di.f:1), AU_3D (au3d.f:1): hu = H u, 4th-order FD kinetic with
 * Dirichlet truncation plus V.  Runs on the device in FP64. */
int espec_au(int ndim, const int *np, const double *shm, const double *vpot,
             const double *u, double *hu);

/* Replaces LNZ (lnz.f:4): m Lanczos steps from start with one classical
 * Gram-Schmidt pass (REORT) against the first nreort vectors.  al[m],
 * bt[m]; basis (n x nreort, column-major) receives the Lanczos vectors. */
int espec_lnz(int ndim, const int *np, int m, int nreort, const double *shm,
              const double *vpot, const double *start, double *al,
              double *bt, double *basis);

/* Replaces the DSTEVX('V','A',...) call of sil.f:82 for an m x m
 * symmetric tridiagonal (d[m], e[m-1]); w ascending, z column-major.
 * Runs on the device (one wavefront). */
int espec_dstevx(int m, const double *d, const double *e, double abstol,
                 double *w, double *z);

/* Replaces LANCZS (lanczs.f:1-3): eigenpairs il..iu (1-based) of the FD
 * Hamiltonian from the RAN start vector (rans.f:3) with an n-step Lanczos
 * and partial reorthogonalisation against nreort vectors; eigenvectors
 * rebuilt by replaying the recurrence once for all il..iu (the reference
 * replays it once per eigenvector).  abstol > 0: eigenvalues are bisection
 * midpoints (DSTEVX('V','I'), dstebz.f).  eigvc is n x (iu-il+1). */
int espec_lanczs(int ndim, const int *np, int il, int iu, int nreort,
                 double abstol, const double *shm, const double *vpot,
                 double *eigvl, double *eigvc);

/* Replaces LANCZSG (lanczsc.f:1-153): states il..iu one at a time by
 * restarted Lanczos: from a RAN vector (NSEED = 1, the first N values
 * discarded), a 40-step recurrence with full Gram-Schmidt (lanczsc.f sets
 * NREORT = 40), DSTEVX('V','I') eigenpairs il..iu of the tridiagonal
 * (bisection to abstol), the Ritz vector of the state as the next start,
 * until the change of the Ritz vector is below abstol, for at most maxint
 * restarts.  ESPEC_REF measures the change as the reference does (its sum
 * reads the 41st component for every J, lanczsc.f:104-109, so it falls
 * below abstol only when grid point 41 is in the state's tail) and returns
 * what the reference returns: eigvl[iu-il+1] the Ritz values of the last
 * restart (as DSTEVX leaves EIGVL) and, for every state but the last, rows
 * 0..39 of its vector replaced by the last restart's tridiagonal
 * eigenvector (DSTEVX writes EIGVC(1:40, 1:M) after those columns were
 * stored).  ESPEC_EXACT sums |VI - s VIAUX| over the grid (s the sign of
 * <VI, VIAUX>) and returns each state's own converged value and vector.
 * eigvc [iu-il+1][N]: the normalised vectors, one row per state.  log (may
 * be NULL): up to logcap records (state, restart, energy, change) as the
 * reference prints them; *nlog receives the count.  ESPEC_ERR_DIAG where
 * the reference stops after maxint restarts.  Runs on the device. */
int espec_lanczsg(int ndim, const int *np, int il, int iu, int maxint, double abstol,
                  const double *shm, const double *vpot, double *eigvl, double *eigvc,
                  double *log, int logcap, int *nlog, int flags);

/* Replaces MTRXDIAG (mtrxdiag.f:1-89): eigenpairs il..iu of the dense FD
 * Hamiltonian that AP1D/AP2D pack (ap1d.f, ap2d.f), by DSPEVX's method on
 * the device: Householder tridiagonalisation, bisection to abstol, inverse
 * iteration (rocSOLVER syevx on a matrix our kernel writes).  1-D and 2-D,
 * N <= 16384.  eigvc column k at eigvc + k ldz; m receives the count. */
int espec_mtrxdiag(int ndim, const int *np, int il, int iu, double abstol,
                   const double *shm, const double *vpot, double *eigvl,
                   double *eigvc, int ldz, int *m);

/* Replaces FCORREL(N,U1,U2) (fcorrel.f:1): sum u1*u2 in FP64 on device. */
int espec_fcorrel(int n, const double *u1, const double *u2, double *out);
/* No counterpart in the reference: espec_fcorrel on device buffers that
 * allocates nothing, for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * u1, u2 [n] and out [1] are device pointers; the call is asynchronous on
 * `stream` (a hipStream_t, NULL for the default stream).  Its block
 * partials and ticket live in a static device array of the library (about
 * 70 KiB per device, held until the process ends), so calls that overlap
 * on one device race; calls in one stream are ordered.  n above
 * ESPEC_NOALLOC_MAX_N returns ESPEC_ERR_ARG.  The sum is bit-identical to
 * espec_fcorrel's (same kernel, same block count). */
#define ESPEC_NOALLOC_MAX_N (1 << 30)
int espec_fcorrel_noalloc(int n, const double *u1, const double *u2, double *out,
                          void *stream);
/* Replaces ECNORM(N,US) (ecnorm.f:1): sqrt(sum u^2). */
int espec_ecnorm(int n, const double *u, double *out);
/* Replaces EFFT4(NDIM,N,NP,B,U,V,WU,WV,TC,VPOT) (eigenergfft.f:1): <psi|V|psi>
 * + <psi|F^-1 TC F|psi> B on the half-swapped copy, FFTs of fft3d.f's
 * definition (ESPEC_REF: f32 twiddles).  tc as filled by DIGT. */
int espec_efft4(int ndim, const int *np, double b, const double *u,
                const double *v, const double *tc, const double *vpot,
                double *e, int flags);
/* Replaces EIGENERG(DIM,N,NP,SHM,U1,VPOT,WORK,VAR) (eigenerg.f:1): <u|H u>
 * with the FD operator of AU. */
int espec_eigenerg(int ndim, const int *np, const double *shm, const double *u,
                   const double *vpot, double *e);

/* ---------------------------------------------------------------------
 * Correlation -> spectrum (espec.f:1546-1594).
 * espec_assemble_corr: espec.f:1546-1566 mirror of 2^(mf-1) samples into
 *   the 2^mf-point C(t) on t = -TF + (i-1)*DTF.
 * espec_hwhm: HWHM (hwhm.f:1).
 * espec_spectrumtd: SPECTRUMTD (spectrumtd.f:1): iwin 1 = .SG window,
 *   0 = none; e/w/wi receive the printed (E, W, FFTi) columns in print
 *   order; the transform is FFTNT (cfft.f:6) on the device.
 * espec_lc_rule: LC and the *CHANGE final-time rule (espec.f:1324-1353).
 * ------------------------------------------------------------------- */
int espec_assemble_corr(int mf, const double *cr, const double *ci,
                        double *u, double *v);
int espec_hwhm(int mf, double dt, double tf, double tal, double de,
               double *u, double *v);
int espec_spectrumtd(int iwin, int mf, double dt, double tf, double wp,
                     double de, double *u, double *v, double *e, double *w,
                     double *wi, int flags);
int espec_lc_rule(double ti, double *tf, double dt, int mf, int change,
                  int *lc);

/* ---------------------------------------------------------------------
 * Device-resident batched split-operator plans (no reference counterpart:
 * they replace the job farm of tools&scrips/parallel.x, which runs one
 * espec.x per packet).  A plan owns device copies of psi, psi0, the phase
 * tables and the C(t) buffer for npk packets on the same grid; run()
 * advances all of them nsteps with inputs already resident in HBM.
 *   vpot  [npk][N], u/v [npk][N]  host arrays copied once at create
 * ------------------------------------------------------------------- */
typedef struct espec_spo_plan espec_spo_plan;
int espec_spo_plan_create(espec_spo_plan **plan, int ndim, const int *np,
                          int npk, double dt, int lc, const double *sh,
                          const double *sm, const double *vpot,
                          const double *u, const double *v, int maxsamp,
                          int flags);
/* As espec_spo_plan_create with an absorber per packet (ABSORB=.TRUE.,
 * spofft.f:65-118, 162-176): vabc [npk][N], see ESPEC_ABSORB_*. */
int espec_spo_plan_create_absorb(espec_spo_plan **plan, int ndim,
                                 const int *np, int npk, double dt, int lc,
                                 const double *sh, const double *sm,
                                 const double *vpot, const double *u,
                                 const double *v, int maxsamp, int flags,
                                 int absorb, const double *vabc);
/* Enqueue nsteps steps on the plan's stream (asynchronous). */
int espec_spo_plan_run(espec_spo_plan *plan, int nsteps);
int espec_spo_plan_sync(espec_spo_plan *plan);
/* Copy out C(t) of packet pk (samples taken so far) and its psi. */
int espec_spo_plan_get(espec_spo_plan *plan, int pk, double *cr, double *ci,
                       double *nrm, int maxsamp, int *nsamp, double *u,
                       double *v);
/* Kernel time (ms) of the last run() measured with HIP events on the
 * plan's stream, and the number of kernel launches it made. */
int espec_spo_plan_last_timing(espec_spo_plan *plan, double *ms,
                               int *launches);
int espec_spo_plan_destroy(espec_spo_plan *plan);
/* Restart a plan from its psi0 (device copy) at step 0, keeping its
 * tables and captured graphs. */
int espec_spo_plan_reset(espec_spo_plan *plan);
/* Per-kernel timing: run nsteps steps with every kernel launched on its
 * own (no graph) between HIP events on the plan's stream.  ms[k] receives
 * the mean duration of kernel kind k and bytes[k] the algorithmic HBM
 * bytes one launch of that kind moves (state read+write, V, psi0 at
 * sample steps): k = 0 the LDS-resident step kernel or the row pass,
 * k = 1 the strided (y, z) passes; nkinds receives the number of kinds. */
int espec_spo_plan_profile(espec_spo_plan *plan, int nsteps, double *ms,
                           double *bytes, int maxkinds, int *nkinds);
/* An uneven batch as one graph (no reference counterpart: it replaces the
 * job farm of tools&scrips/parallel.x).  create() records every kernel of
 * nsteps steps of each plan -- one plan per grid, from each plan's current
 * step -- on its own branch of one HIP graph; run() advances every packet
 * of every plan by nsteps with a single graph launch (the plans' streams
 * wait for it) and returns ESPEC_ERR_ARG unless every plan is still at the
 * step it was captured at (reset() the plans to replay from step 0).
 * sync() waits and gives the graph's duration in ms (HIP events).  The
 * plans must outlive the batch; the graph keeps the launch shapes the
 * plans had at create(). */
typedef struct espec_spo_batch espec_spo_batch;
int espec_spo_batch_create(espec_spo_batch **batch, espec_spo_plan *const *plans,
                           int nplans, int nsteps);
int espec_spo_batch_run(espec_spo_batch *batch);
int espec_spo_batch_sync(espec_spo_batch *batch, double *ms);
int espec_spo_batch_destroy(espec_spo_batch *batch);

/* ---------------------------------------------------------------------
 * Batched short-iterative Lanczos plan: npk packets on one grid (PLNZ's
 * time loop, plnz.f:109-131, without the per-step host round trip).
 * vpot is [npk][n]; u, v [npk][n] (v may be NULL for a real start).
 * Each step runs the two real recurrences of sil.f:70-181 of every packet
 * side by side; C(t) = <psi0|psi(t)> and |psi|^2 every lc steps.
 * ------------------------------------------------------------------- */
typedef struct espec_sil_plan espec_sil_plan;
int espec_sil_plan_create(espec_sil_plan **plan, int ndim, const int *np,
                          int npk, int mp, double abstol, double dt, int lc,
                          const double *shm, const double *vpot,
                          const double *u, const double *v, int maxsamp,
                          int flags);
/* Make every later step of the plan pulse-driven (PPLNZ): step g (1-based
 * over the plan's life) uses VEFF = VPOT - dm E_g with E_g = efield[g-1];
 * dm [npk][N].  Runs past nfield steps return ESPEC_ERR_ARG; adaptive and
 * trace runs of a pulse-driven plan return ESPEC_ERR_UNSUPPORTED. */
int espec_sil_plan_set_pulse(espec_sil_plan *plan, const double *dm,
                             const double *efield, int nfield);
/* The absorber of PLNZAUX for every packet of the plan: ESPEC_ABSORB_SMOOTHW
 * with vabc [npk][N] masks the state after each step (sil.f:171-180);
 * ESPEC_ABSORB_NONE (or _VOPTIC, which PLNZAUX ignores) removes it.
 * Adaptive and trace runs of a masked plan return ESPEC_ERR_UNSUPPORTED. */
int espec_sil_plan_set_absorb(espec_sil_plan *plan, int absorb, const double *vabc);
int espec_sil_plan_run(espec_sil_plan *plan, int nsteps);
int espec_sil_plan_sync(espec_sil_plan *plan);
/* Returns ESPEC_ERR_DIAG if a tridiagonal eigensolve failed to converge
 * (the reference stops with 'Diagonalization error', sil.f:120). */
int espec_sil_plan_get(espec_sil_plan *plan, int pk, double *cr, double *ci,
                       double *nrm, int maxsamp, int *nsamp, double *u,
                       double *v);
int espec_sil_plan_destroy(espec_sil_plan *plan);
/* Restart from psi0 at step 0, keeping captured graphs (no reference
 * counterpart). */
int espec_sil_plan_reset(espec_sil_plan *plan);

/* ---------------------------------------------------------------------
 * Learned Krylov-step controller (no reference counterpart).
 *
 * Adaptive short-iterative Lanczos: before every step a network
 * (8 -> 64 -> 64 -> 2, bfloat16 weights and activations, FP32 sums)
 * maps eight inputs of the previous Krylov run (alpha_1, beta_1, alpha_2,
 * beta_2 times dt, log10 of beta_4 dt and beta_5 dt, log10 of the order-4
 * residual estimate / 10, log10 of the norm drift / 16) to a Krylov order k = ceil(4 + 16 o1) and a step
 * scale s = 2^round(2 o2 - 1) in {1/2, 1, 2} (max k and min s over the
 * packets; after a step that needed the extended space described below,
 * the order is at least one below the one it used).  The step is of order m = k + 1, and it is accepted when, for
 * every real recurrence, the larger of beta0 |y_{m-1} - y_m| (y_j =
 * exp(-i s dt T_j) e1, the order-j approximation in the Krylov basis, zero
 * padded) and the residual estimate beta0 beta_m |[exp(-i s dt T_m)]_{m,1}|
 * is at most tol * s (each half of an s = 1/2 step: tol / 2).  Otherwise
 * the same step is retried with the Krylov space extended to order mp,
 * and when that misses the budget too, the fixed-order step of order m_fix
 * over one dt (or the half step) replaces it, on the same Krylov space
 * (order mp when m_fix misses the budget of one base step).  C(t) is sampled
 * every dt (plans with lc = 1; the exact, not the reference, coefficient
 * rule).  params: the network's kMlpNPar = 4866 FP32 parameters (W1[64][8]
 * b1[64] W2[64][64] b2[64] W3[2][64] b3[2]), or NULL for the fixed-order
 * step every dt.  stats[7]: macro steps, steps at scale 1/2, 1, 2,
 * fallbacks, Lanczos iterations run, host waits.  Orders are capped at
 * mp (the plan's Krylov storage); mp >= 21 lets k reach 20.
 */
int espec_sil_plan_adaptive(espec_sil_plan *plan, const float *params, int ndt,
                            double tol, int m_fix, long long *stats);
/* Diagnostics of the adaptive runs that follow: each network-chosen
 * attempt appends (base step, order, step scale, estimate, budget) to log
 * (cap records of 5 doubles; NULL stops recording); *len receives the
 * number of records written since the previous call. */
int espec_sil_plan_adaptive_log(espec_sil_plan *plan, double *log, int cap, int *len);
/* Training trace: nsteps steps over dt that run mp (>= 21) Lanczos
 * iterations and advance the state with order mprop (5 <= mprop <= mp; a
 * lower order leaves the state as rough as a propagation at that order
 * makes it), recording per step the controller inputs feat[nsteps][npk][8]
 * and lab[nsteps][2 npk][17][3]: for each network order k = 4..20 at
 * dt/2, dt, 2 dt, the estimate the adaptive step of order k + 1 is judged
 * by, for each real recurrence (2 pk: Re psi, 2 pk + 1: Im psi). */
int espec_sil_plan_trace(espec_sil_plan *plan, int nsteps, int mprop, float *feat,
                         double *lab);
/* The network alone: outputs out[n][2] for inputs feat[n][8]; the
 * gradient of one minibatch of 64 rows (loss = mean of squared output
 * errors against t[64][2]); and Adam training on the device over n rows
 * (epochs passes of shuffled 64-row minibatches, one workgroup; params
 * updated in place, loss[epochs] the mean minibatch loss per pass). */
int espec_mlp_infer(const float *params, const float *feat, int n, float *out);
int espec_mlp_grad(const float *params, const float *x, const float *t,
                   float *grad, float *loss);
int espec_mlp_train(float *params, const float *x, const float *t, int n,
                    int epochs, float lr, unsigned seed, float *loss);

/* Launch shape chosen for the plan, shape[13]: 1 if LDS-resident, rows per
 * row-pass workgroup, its threads, strided tile width, its threads, XCD
 * spread, XCD-contiguous tile map, workgroups per pass (0: one per row block
 * or tile; else a grid-stride loop), register-pencil row pass, register-pencil
 * strided passes, prefetch of the next work item in the pencil strided
 * passes, 1 if the row pass reads an FP64 V phase table, and the fused
 * form (0, or 4 or 8 wavefronts per workgroup: a 2-D grid of 64^2, or
 * 128^2 in complex64, runs the row and strided passes' arithmetic on a
 * copy of each packet held in one workgroup's LDS for a whole launch of
 * steps) (no reference counterpart; for reports). */
int espec_spo_plan_shape(espec_spo_plan *plan, int *shape);
/* Pin a multi-pass plan's launch shape (shape[13] as above; shape[0] and
 * shape[11] are ignored) in place of the autotuner's choice, e.g. to test one kernel
 * family.  ESPEC_ERR_ARG if the grid cannot use that shape,
 * ESPEC_ERR_UNSUPPORTED for an LDS-resident plan (no reference counterpart). */
int espec_spo_plan_set_shape(espec_spo_plan *plan, const int *shape);

/* ---------------------------------------------------------------------
 * Batched second-order-differencing plans: npk packets on one grid
 * (PSODFFT's time loop, psodfft.f:109-200).  The first step of the first
 * run is the two half-strength start-up steps; grids above 2048 points run
 * one pass per axis (ESPEC_EXACT only).
 * ------------------------------------------------------------------- */
typedef struct espec_sod_plan espec_sod_plan;
int espec_sod_plan_create(espec_sod_plan **plan, int ndim, const int *np,
                          int npk, double dt, int lc, const double *sh,
                          const double *sm, const double *vpot,
                          const double *u, const double *v, int maxsamp,
                          int flags);
/* As espec_sod_plan_create with an absorber per packet, vabc [npk][N]. */
int espec_sod_plan_create_absorb(espec_sod_plan **plan, int ndim,
                                 const int *np, int npk, double dt, int lc,
                                 const double *sh, const double *sm,
                                 const double *vpot, const double *u,
                                 const double *v, int maxsamp, int flags,
                                 int absorb, const double *vabc);
int espec_sod_plan_run(espec_sod_plan *plan, int nsteps);
int espec_sod_plan_sync(espec_sod_plan *plan);
int espec_sod_plan_get(espec_sod_plan *plan, int pk, double *cr, double *ci,
                       double *nrm, int maxsamp, int *nsamp, double *u,
                       double *v);
/* psi(t-dt) of packet pk (PSODFFT returns it in U1, V1; psodfft.f:200). */
int espec_sod_plan_get_prev(espec_sod_plan *plan, int pk, double *u,
                            double *v);
int espec_sod_plan_destroy(espec_sod_plan *plan);

/* Streaming copy kernel over nbytes, iters times: mean ms per launch and
 * bytes moved per launch (read + write), for bandwidth calibration. */
int espec_bw_probe(size_t nbytes, int iters, double *ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
