/*
 * espec_oracle.h -- CPU FP64 restatement of the eSPec wave-packet hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is
 * compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (libespec_hip.so)
 * never links, loads or falls back to it.
 *
 * Every routine states the reference file:line it restates.  Two modes:
 *   ORC_REF   reproduces the reference's arithmetic quirks: the Singleton
 *             twiddle angles built from the f32-rounded 2*pi (fft3d.f:119)
 *             and the reflected radix-2 twiddles, the accumulated kinetic
 *             grid with 2*pi = 6.283185307 (dkeft.f:241,272-282), the
 *             non-unitary kinetic rotation (spofft.f:152-153), the
 *             cos*sin aliasing of the Lanczos recombination (sil.f:98-99)
 *             and FFTNT's f32 2*pi (cfft.f:35).
 *   ORC_EXACT the same algorithms with those quirks removed.
 * Arrays are split real/imaginary REAL*8 vectors, flattened with x fastest
 * (gvpot.f:80), exactly as the reference passes them.
 */
#ifndef ESPEC_ORACLE_H
#define ESPEC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EXACT = 0, ORC_REF = 1 };
/* Absorbing boundaries (TPABSOR): none, .VOPTIC (vabc = imaginary
 * potential, hartree; spofft.f:65-118, sodfft.f:66-72) and .SMOOTHW
 * (vabc = mask; spofft.f:162-176, sodfft.f:108-116). */
enum { ORC_ABS_NONE = 0, ORC_ABS_VOPTIC = 1, ORC_ABS_SMOOTHW = 2 };
/* Absorber shapes of ABSORBINGBC (absorb.f). */
enum { ORC_ABG_SMOOTHW = 1, ORC_ABG_SMRS = 2, ORC_ABG_VOPTIC_LINEAR = 3,
       ORC_ABG_VOPTIC_KOSLOFF = 4, ORC_ABG_VOPTIC_POLIN = 5 };
/* INIEIGVC choices of INIT_COND (init_cond.f): .CALC, .GETR/.GETR1,
 * .GETR2, .GETC/.GETC1, .GETC2. */
enum { ORC_INIT_CALC = 0, ORC_INIT_GETR = 1, ORC_INIT_GETR2 = 2, ORC_INIT_GETC = 3,
       ORC_INIT_GETC2 = 4 };

/* Atomic unit of time in fs used by every propagator (pspofft.f:76). */
#define ORC_SHFS 2.4188843243e-2

/* fft3d.f:1 FFT(A,B,NTOT,N,NSPAN,ISN), powers of two (and a direct DFT
 * for any other n).  isn > 0 gives exp(+i...), unnormalised. */
void orc_fft(double *a, double *b, int ntot, int n, int nspan, int isn, int mode);
/* Singleton's factor order for n = 2^p (fft3d.f:124-173); returns count. */
int orc_singleton_factors(int n, int *fac);
/* cfft.f:6 FFTNT(X,Y,N,M,ITYPE): itype=-1 is exp(+i...) scaled by 1/N. */
void orc_fftnt(double *x, double *y, int n, int m, int itype, int mode);

/* dkeft.f:204 DIGT: tc = p^2/(2m) on the FFT momentum grid, b = 1/N. */
void orc_digt(int ndim, const int *np, const double *sh, const double *sm,
              double *tc, double *b, int mode);

/* spofft.f:3 PSPOFFTAUX, no absorber: one Strang step in place. */
void orc_spo_step(int ndim, const int *np, double b, double cst1, double cst2,
                  double *u, double *v, const double *tc, const double *vpot,
                  int mode);
/* sodfft.f:3 PSODFFTAUX, no absorber: one leapfrog step in place. */
void orc_sod_step(int ndim, const int *np, double b, double cst, double *u1,
                  double *u2, double *v1, double *v2, const double *tc,
                  const double *vpot, double *wr, double *wi, int mode);

/* PSPOFFTAUX / PSODFFTAUX with an absorber (see ORC_ABS_*). */
void orc_spo_step_abs(int ndim, const int *np, double b, double cst1,
                      double cst2, double *u, double *v, const double *tc,
                      const double *vpot, int absorb, const double *vabc,
                      int mode);
void orc_sod_step_abs(int ndim, const int *np, double b, double cst, double *u1,
                      double *u2, double *v1, double *v2, const double *tc,
                      const double *vpot, int absorb, const double *vabc,
                      double *wr, double *wi, int mode);

/* Trip count of the reference's REAL DO loop DO T=TI+DT,TF+DT,DT. */
int orc_nsteps(double ti, double tf, double dt);

/* pspofft.f:1-150 PSPOFFT with PRTCRL='.NO': C(t) every lc steps into
 * cr/ci (sample 0 = t=TI), returns the sample count.  u1,v1 in/out.
 * If e/nrm are non-NULL they receive EFFT4 energy and norm per sample. */
int orc_pspofft(int ndim, const int *np, double dt, double ti, double tf,
                int lc, const double *sh, const double *sm, double *u1,
                double *v1, const double *vpot, double *cr, double *ci,
                int maxsamp, double *e, double *nrm, int mode);
/* As orc_pspofft with ABSORB=.TRUE. and TPABSOR selected by absorb; vabc
 * [N] is the VABC argument (absorb.f). */
int orc_pspofft_abs(int ndim, const int *np, double dt, double ti, double tf,
                    int lc, const double *sh, const double *sm, double *u1,
                    double *v1, const double *vpot, int absorb,
                    const double *vabc, double *cr, double *ci, int maxsamp,
                    double *e, double *nrm, int mode);
/* psodfft.f:1-211 PSODFFT with PRTCRL='.NO'. */
int orc_psodfft(int ndim, const int *np, double dt, double ti, double tf,
                int lc, const double *sh, const double *sm, double *u1,
                double *v1, const double *vpot, double *cr, double *ci,
                int maxsamp, int mode);
int orc_psodfft_abs(int ndim, const int *np, double dt, double ti, double tf,
                    int lc, const double *sh, const double *sm, double *u1,
                    double *v1, const double *vpot, int absorb,
                    const double *vabc, double *cr, double *ci, int maxsamp,
                    int mode);

/* fcorrel.f:1 / ecnorm.f:1 / eigenergfft.f:1 / eigenerg.f:1. */
double orc_fcorrel(int n, const double *a, const double *b);
double orc_ecnorm(int n, const double *a);
double orc_efft4(int ndim, const int *np, const double *u, const double *v,
                 const double *tc, const double *vpot, int mode);
double orc_eigenerg(int ndim, const int *np, const double *shm,
                    const double *u, const double *vpot);

/* au.f:3 AU -> au1d.f / au2di.f / au3d.f: 4th-order FD H*u with the
 * Dirichlet truncation of the 5-point stencils. */
void orc_au(int ndim, const int *np, const double *shm, const double *vpot,
            const double *u, double *hu);

/* lnz.f:4 LNZ with reort.f:1 REORT: m Lanczos steps from start (not
 * modified), classical Gram-Schmidt against the first nreort vectors.
 * basis is n x nreort (column-major, ld = n).  Returns the number of
 * steps done (0 if the start vector is zero, lnz.f:86). */
int orc_lnz(int ndim, const int *np, int m, int nreort, const double *shm,
            const double *vpot, const double *start, double *al, double *bt,
            double *basis);
/* reort.f:1 REORT: w -= V (V^T w), V = first mcols columns (ld = n). */
void orc_reort(int n, int mcols, const double *basis, double *w);

/* Symmetric tridiagonal eigensolver (the DSTEVX 'V','A' role with
 * ABSTOL <= 0, i.e. dsteqr.f): implicit QL with Wilkinson shifts.
 * d[m] diagonal, e[m-1] off-diagonal; eigenvalues ascending into w,
 * eigenvectors into z (m x m column-major).  Returns 0 or >0 on failure. */
int orc_tridiag_eig(int m, const double *d, const double *e, double *w,
                    double *z);
/* dstebz.f bisection semantics for ABSTOL > 0 (dstevx.f:302-313):
 * eigenvalues only, each the midpoint of its final Sturm interval. */
int orc_tridiag_bisect(int m, const double *d, const double *e, double abstol,
                       double *w);

/* sil.f:3 PLNZAUX (no absorber): one real-split short-iterative Lanczos
 * step of order mp with e^{-i H cst}.  abstol selects the eigensolver as
 * DSTEVX does.  Returns 0 or a nonzero diagonalisation failure. */
int orc_plnz_step(int ndim, const int *np, int mp, double abstol, double cst,
                  const double *shm, const double *vpot, double *u, double *v,
                  int mode);
/* plnz.f:1-156 PLNZ with PRTCRL='.NO'. */
int orc_plnz(int ndim, const int *np, int mp, double abstol, double dt,
             double ti, double tf, int lc, const double *shm, double *u,
             double *v, const double *vpot, double *cr, double *ci,
             int maxsamp, int mode);
/* orc_plnz with the .SMOOTHW mask of PLNZAUX (sil.f:171-180) when mask is
 * not NULL. */
int orc_plnz_abs(int ndim, const int *np, int mp, double abstol, double dt,
                 double ti, double tf, int lc, const double *shm, double *u,
                 double *v, const double *vpot, const double *mask, double *cr,
                 double *ci, int maxsamp, int mode);
/* ef.f EF (efc 0 .NONE/.NULL, 1 .GAUS, 2 .GGAUS, 3 .SIN2) and pplnz.f
 * PPLNZ with PRTCRL='.NO' (veff may be NULL). */
/* init_cond.f INIT_COND (see ORC_INIT_*); u, v [N] out (in: the 1-D
 * eigenvector for the .GET* modes). */
int orc_init_cond(int ndim, const int *np, int mode, const double *xi,
                  const double *sh, const double *x0, const double *ra,
                  const double *rk, double *u, double *v);
/* absorb.f ABSORBINGBC (see ORC_ABG_*); alpha may be NULL. */
int orc_absorbingbc(int ndim, const int *np, int kind, const double *al,
                    const double *ar, const double *sh, const double *xi,
                    const double *xf, const double *voi, double *vabc,
                    double *alpha);
double orc_ef(int efc, double e0, double t0, double td, double tp, double omg,
              double sni, int kl, double t);
int orc_pplnz(int ndim, const int *np, int mp, double abstol, double dt,
              double ti, double tf, int lc, const double *shm, double *u,
              double *v, const double *vp, const double *dm, int efc,
              double e0, double t0, double td, double tp, double omg,
              double sni, int kl, double *veff, double *cr, double *ci,
              int maxsamp, int mode);
/* orc_pplnz with the .SMOOTHW mask of PLNZAUX (pplnz.f:208, sil.f:171-180). */
int orc_pplnz_abs(int ndim, const int *np, int mp, double abstol, double dt,
                  double ti, double tf, int lc, const double *shm, double *u,
                  double *v, const double *vp, const double *dm, int efc,
                  double e0, double t0, double td, double tp, double omg,
                  double sni, int kl, const double *mask, double *veff, double *cr,
                  double *ci, int maxsamp, int mode);

/* rans.f:3 RAN: NSEED=(NSEED*1029+221591) mod 2^20, RAN=NSEED/2^20. */
double orc_ran(int *nseed);
/* lanczs.f:1 LANCZS: eigenpairs il..iu (1-based) of the FD Hamiltonian
 * from the RAN start vector with an n-step Lanczos and partial
 * reorthogonalisation against the first nreort vectors; eigenvectors are
 * rebuilt by replaying the recurrence (lanczs.f:104-177).  eigvc is
 * n x (iu-il+1).  abstol > 0: eigenvalues are dstebz.f bisection
 * midpoints (DSTEVX('V','I'), lanczs.f:90); vectors from QL. */
int orc_lanczs(int ndim, const int *np, int il, int iu, int nreort,
               double abstol, const double *shm, const double *vpot,
               double *eigvl, double *eigvc);

/* espec.f:1324-1353: LC and the *CHANGE final-time rule.  Returns LC and
 * writes the (possibly changed) final time. */
int orc_lc_rule(double ti, double *tf, double dt, int mf, int change);
/* espec.f:1546-1566: mirror C(t) samples (cr/ci, 2^(mf-1) of them) into
 * the 2^mf-point correlation u + i v on t = -TF + (i-1)*DTF. */
void orc_assemble_corr(int mf, const double *cr, const double *ci, double *u,
                       double *v);
/* hwhm.f:1 HWHM: Lorentzian damping and energy shift. */
void orc_hwhm(int mf, double dt, double tf, double tal, double de, double *u,
              double *v);
/* spectrumtd.f:1 SPECTRUMTD: window (iwin 1 = .SG, 0 = none), half swap,
 * FFTNT(-1).  The printed spectrum is e[i], u[i] (=W), v[i] in the
 * reference's print order, written to e_out/w_out/wi_out (length 2^mf). */
void orc_spectrumtd(int iwin, int mf, double dt, double tf, double wp,
                    double de, double *u, double *v, double *e_out,
                    double *w_out, double *wi_out, int mode);

#ifdef __cplusplus
}
#endif

/* PSODAUX (sod.f:3-122): one FD leapfrog step, psi(t+dt) = psi(t-dt)
 * - i cst H psi(t); absorb ORC_ABS_* with vabc. */
void orc_psodaux(int ndim, const int *np, double cst, const double *shm, double *u1,
                 double *u2, double *v1, double *v2, const double *vpot, int absorb,
                 const double *vabc);
/* PPSOD (ppsod.f:1-548) time loop: returns the number of C(t) samples
 * (sample 0 = C(TI), then every LC steps); u, v receive psi(TF). */
int orc_s2ppsod(int ndim, const int *np, double dt, double ti, double tf, int lc,
                const double *shm, double xi, double sh, double sm1, double aa, double *u,
                double *v, const double *vp, const double *dm, int efc, double e0, double t0,
                double td, double tp, double omg, double sni, int kl, int absorb,
                const double *vabc, int prtcrl, double tol, double *tab, int maxrows,
                int *info, double *u1out, double *v1out);
int orc_ppsod(int ndim, const int *np, double dt, double ti, double tf, int lc,
              const double *shm, double *u, double *v, const double *vp, const double *dm,
              int efc, double e0, double t0, double td, double tp, double omg, double sni,
              int kl, int absorb, const double *vabc, double *veff, double *cr, double *ci,
              double *nrm, int maxsamp);
#endif
