/*
 * espec_oracle.c -- CPU FP64 restatement of the eSPec wave-packet hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product); see
 * espec_oracle.h for the contract and the two modes.  Written from the
 * mathematics of each reference routine; the steps that must stay alike
 * (loop order of the split-operator step, sampling rule, Lanczos
 * recurrence order) follow the cited reference lines so that rounding
 * stays within 1e-12 of the reference build.
 */
#include "espec_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double TWOPI_EXACT = 6.283185307179586476925287;

/* fft3d.f:119 RAD=6.2831853071796 is a default-REAL literal: f32. */
static double fft_rad(int mode)
{
    return mode == ORC_REF ? (double)(float)6.2831853071796 : TWOPI_EXACT;
}

static int ntot_of(int ndim, const int *np)
{
    int n = 1;
    for (int i = 0; i < ndim; ++i) n *= np[i];
    return n;
}

/* ------------------------------------------------------------------ FFT */

/* fft3d.f:124-173: square factors of 16 -> 4, then for a remaining 8 a
 * square factor 2, the square-free middle factor, then the square factors
 * mirrored. */
int orc_singleton_factors(int n, int *fac)
{
    int sq[40], ns = 0, k = n, mid = 0, nf = 0;
    while (k % 16 == 0) { sq[ns++] = 4; k /= 16; }
    if (k > 4) {              /* k == 8 */
        sq[ns++] = 2;
        k /= 4;
        mid = k;
    } else if (k != 1) {
        mid = k;
    }
    for (int i = 0; i < ns; ++i) fac[nf++] = sq[i];
    if (mid) fac[nf++] = mid;
    for (int i = ns - 1; i >= 0; --i) fac[nf++] = sq[i];
    return nf;
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* Decimation-in-frequency transform of one contiguous line (re, im) of
 * length n with Singleton's factor order.  Stage s of length L applies
 * the radix-r DFT to x[b + j + p*m], m = L/r, then multiplies output q by
 * the twiddle of angle q*j*RAD/L (fft3d.f:175-177, 420-449).  Singleton's
 * radix-2 stage reuses the twiddle of j for m-j through cos -> -cos
 * (fft3d.f:206-211); with the f32 RAD that reflection is not the same
 * angle, so ORC_REF keeps it.  The digit-reversed result is permuted to
 * natural order (fft3d.f:800-950). */
static void dif_line(double *re, double *im, int n, int isn, int mode,
                     double *tr, double *ti)
{
    int fac[40];
    int nf = orc_singleton_factors(n, fac);
    double rad = fft_rad(mode);
    if (isn < 0) rad = -rad;
    int L = n;
    for (int s = 0; s < nf; ++s) {
        int r = fac[s], m = L / r;
        double phi = rad / (double)L;
        for (int j = 0; j < m; ++j) {
            double c1, s1, c2 = 1, s2 = 0, c3 = 1, s3 = 0;
            if (r == 2 && mode == ORC_REF && 2 * j > m) {
                c1 = -cos((double)(m - j) * phi);
                s1 = sin((double)(m - j) * phi);
            } else {
                c1 = cos((double)j * phi);
                s1 = sin((double)j * phi);
            }
            if (r == 4) {
                c2 = cos(2.0 * j * phi); s2 = sin(2.0 * j * phi);
                c3 = cos(3.0 * j * phi); s3 = sin(3.0 * j * phi);
            }
            for (int b = 0; b < n; b += L) {
                int k0 = b + j;
                if (r == 2) {
                    int k1 = k0 + m;
                    double ar = re[k0] - re[k1], ai = im[k0] - im[k1];
                    re[k0] += re[k1];
                    im[k0] += im[k1];
                    re[k1] = c1 * ar - s1 * ai;
                    im[k1] = s1 * ar + c1 * ai;
                } else { /* r == 4 */
                    int k1 = k0 + m, k2 = k1 + m, k3 = k2 + m;
                    double akpr = re[k0] + re[k2], akmr = re[k0] - re[k2];
                    double akpi = im[k0] + im[k2], akmi = im[k0] - im[k2];
                    double ajpr = re[k1] + re[k3], ajmr = re[k1] - re[k3];
                    double ajpi = im[k1] + im[k3], ajmi = im[k1] - im[k3];
                    double x1r, x1i, x3r, x3i;
                    if (isn >= 0) {
                        x1r = akmr - ajmi; x1i = akmi + ajmr;
                        x3r = akmr + ajmi; x3i = akmi - ajmr;
                    } else {
                        x1r = akmr + ajmi; x1i = akmi - ajmr;
                        x3r = akmr - ajmi; x3i = akmi + ajmr;
                    }
                    double x2r = akpr - ajpr, x2i = akpi - ajpi;
                    re[k0] = akpr + ajpr;
                    im[k0] = akpi + ajpi;
                    re[k1] = x1r * c1 - x1i * s1; im[k1] = x1r * s1 + x1i * c1;
                    re[k2] = x2r * c2 - x2i * s2; im[k2] = x2r * s2 + x2i * c2;
                    re[k3] = x3r * c3 - x3i * s3; im[k3] = x3r * s3 + x3i * c3;
                }
            }
        }
        L = m;
    }
    /* X[k] sits at sum_s q_s * n/(r_1..r_s), k = q_1 + r_1 q_2 + ... */
    for (int k = 0; k < n; ++k) {
        int kk = k, span = n, pos = 0;
        for (int s = 0; s < nf; ++s) {
            span /= fac[s];
            pos += (kk % fac[s]) * span;
            kk /= fac[s];
        }
        tr[k] = re[pos];
        ti[k] = im[pos];
    }
    memcpy(re, tr, sizeof(double) * n);
    memcpy(im, ti, sizeof(double) * n);
}

/* Direct DFT for lengths that are not powers of two (angles from the same
 * RAD); the hot path only uses powers of two (espec.f:260-277). */
static void dft_line(double *re, double *im, int n, int isn, int mode,
                     double *tr, double *ti)
{
    double rad = fft_rad(mode);
    if (isn < 0) rad = -rad;
    for (int k = 0; k < n; ++k) {
        double sr = 0, si = 0;
        for (int j = 0; j < n; ++j) {
            double a = rad * (double)((long)j * k % n) / (double)n;
            double c = cos(a), s = sin(a);
            sr += re[j] * c - im[j] * s;
            si += re[j] * s + im[j] * c;
        }
        tr[k] = sr;
        ti[k] = si;
    }
    memcpy(re, tr, sizeof(double) * n);
    memcpy(im, ti, sizeof(double) * n);
}

/* fft3d.f:1 -- one variate of a multivariate transform: lines of length n
 * whose elements are nspan/n apart. */
void orc_fft(double *a, double *b, int ntot, int n, int nspan, int isn, int mode)
{
    if (n < 2) return;
    int stride = nspan / n;
    double *lr = malloc(sizeof(double) * 4 * n);
    double *li = lr + n, *tr = lr + 2 * n, *ti = lr + 3 * n;
    int nlines = ntot / n;
    for (int l = 0; l < nlines; ++l) {
        int base = (l % stride) + (l / stride) * nspan;
        for (int j = 0; j < n; ++j) {
            lr[j] = a[base + j * stride];
            li[j] = b[base + j * stride];
        }
        if (is_pow2(n)) dif_line(lr, li, n, isn, mode, tr, ti);
        else dft_line(lr, li, n, isn, mode, tr, ti);
        for (int j = 0; j < n; ++j) {
            a[base + j * stride] = lr[j];
            b[base + j * stride] = li[j];
        }
    }
    free(lr);
}

static void fft_all(int ndim, const int *np, double *u, double *v, int isn, int mode)
{
    int n = ntot_of(ndim, np), nf = 1;
    for (int i = 0; i < ndim; ++i) {
        nf *= np[i];
        orc_fft(u, v, n, np[i], nf, isn, mode);
    }
}

/* cfft.f:6-106 FFTNT: radix-2 DIF with exp(-i a), a = (IA-1)*TWOPI/N,
 * bit reversal; itype=-1 conjugates in and out and scales by 1/N.  TWOPI
 * is a default-REAL literal (cfft.f:35). */
void orc_fftnt(double *x, double *y, int n, int m, int itype, int mode)
{
    if (n == 1) return;
    double twopi = mode == ORC_REF ? (double)(float)6.283185307179586476925287
                                   : TWOPI_EXACT;
    if (itype == -1)
        for (int i = 0; i < n; ++i) y[i] = -y[i];
    double p = twopi / n;
    int n2 = n;
    for (int k = 0; k < m; ++k) {
        int n1 = n2;
        n2 /= 2;
        int ie = n / n1, ia = 1;
        for (int j = 0; j < n2; ++j) {
            double a = (double)(ia - 1) * p;
            double c = cos(a), s = sin(a);
            ia += ie;
            for (int i = j; i < n; i += n1) {
                int l = i + n2;
                double xt = x[i] - x[l];
                x[i] = x[i] + x[l];
                double yt = y[i] - y[l];
                y[i] = y[i] + y[l];
                x[l] = c * xt + s * yt;
                y[l] = c * yt - s * xt;
            }
        }
    }
    /* in-place bit reversal */
    for (int i = 0, j = 0; i < n - 1; ++i) {
        if (i < j) {
            double t = x[j]; x[j] = x[i]; x[i] = t;
            t = y[j]; y[j] = y[i]; y[i] = t;
        }
        int k = n / 2;
        while (k <= j) { j -= k; k /= 2; }
        j += k;
    }
    if (itype == -1)
        for (int i = 0; i < n; ++i) {
            y[i] = -y[i] / n;
            x[i] = x[i] / n;
        }
}

/* ------------------------------------------------------- kinetic grid */

/* dkeft.f:204-282.  ORC_REF accumulates p by repeated +DP with the
 * 10-digit 2*pi (dkeft.f:241, 272-282) and flips the sign at N/2; the
 * reference's 2-D/3-D branches do not fill a grid (dkeft.f:283-340), so
 * for ndim > 1 ORC_REF builds each axis with the 1-D recipe and sums.
 * ORC_EXACT uses p_k = (k - N [k >= N/2]) 2 pi / (N h). */
void orc_digt(int ndim, const int *np, const double *sh, const double *sm,
              double *tc, double *b, int mode)
{
    int n = ntot_of(ndim, np);
    *b = 1.0 / n;
    double *ax[3];
    for (int a = 0; a < ndim; ++a) {
        int na = np[a];
        ax[a] = malloc(sizeof(double) * na);
        if (mode == ORC_REF) {
            double dp = 6.283185307 / (na * sh[a]);
            double p1 = 0.0;
            for (int i = 0; i < na / 2; ++i) {
                ax[a][i] = p1 * p1 / (2.0 * sm[a]);
                p1 = p1 + dp;
            }
            p1 = -p1;
            for (int i = na / 2; i < na; ++i) {
                ax[a][i] = p1 * p1 / (2.0 * sm[a]);
                p1 = p1 + dp;
            }
        } else {
            double dp = TWOPI_EXACT / (na * sh[a]);
            for (int i = 0; i < na; ++i) {
                double p = (double)(i < na / 2 ? i : i - na) * dp;
                ax[a][i] = p * p / (2.0 * sm[a]);
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        int r = i;
        double t = 0;
        for (int a = 0; a < ndim; ++a) {
            t += ax[a][r % np[a]];
            r /= np[a];
        }
        tc[i] = t;
    }
    for (int a = 0; a < ndim; ++a) free(ax[a]);
}

/* -------------------------------------------------- propagator steps */

/* spofft.f:120-133 and 178-191: V half-phase fused with the swap of the
 * two halves of the flattened array; post != 0 also scales by b.
 * Absorbers (spofft.f:65-118, 162-176): ORC_ABS_VOPTIC multiplies both
 * half-phases' cosine and sine by exp(cst2 vabc) of the point they act
 * on; ORC_ABS_SMOOTHW multiplies the result of the post half-phase by the
 * mask vabc of the point it is stored at. */
static void vphase_swap(int n, double cst2, double b, int post, double *u,
                        double *v, const double *vpot, int absorb,
                        const double *vabc)
{
    int nh = n / 2;
    for (int i = 0; i < nh; ++i) {
        /* spofft.f:120-133 (pre) / 178-191 (post, reading the swapped
         * positions and scaling by b). */
        int p1 = post ? i + nh : i, p2 = post ? i : i + nh;
        double th1 = cst2 * vpot[p1];
        double c1 = cos(th1), s1 = sin(th1);
        double th2 = cst2 * vpot[p2];
        double c2 = cos(th2), s2 = sin(th2);
        if (absorb == ORC_ABS_VOPTIC) {
            /* spofft.f:68-71, 76-79, 100-103, 108-111 */
            double e1 = exp(cst2 * vabc[p1]), e2 = exp(cst2 * vabc[p2]);
            c1 = c1 * e1;
            s1 = s1 * e1;
            c2 = c2 * e2;
            s2 = s2 * e2;
        }
        double uaux, vaux;
        if (post) {
            uaux = c1 * b * u[i] + s1 * b * v[i];
            vaux = c1 * b * v[i] - s1 * b * u[i];
            double un = c2 * b * u[i + nh] + s2 * b * v[i + nh];
            double vn = c2 * b * v[i + nh] - s2 * b * u[i + nh];
            if (absorb == ORC_ABS_SMOOTHW) {
                /* spofft.f:171-174: the mask of the stored point */
                un = vabc[i] * un;
                vn = vabc[i] * vn;
                uaux = vabc[i + nh] * uaux;
                vaux = vabc[i + nh] * vaux;
            }
            u[i] = un;
            v[i] = vn;
        } else {
            uaux = c1 * u[i] + s1 * v[i];
            vaux = c1 * v[i] - s1 * u[i];
            u[i] = c2 * u[i + nh] + s2 * v[i + nh];
            v[i] = c2 * v[i + nh] - s2 * u[i + nh];
        }
        u[i + nh] = uaux;
        v[i + nh] = vaux;
    }
}

void orc_spo_step_abs(int ndim, const int *np, double b, double cst1,
                      double cst2, double *u, double *v, const double *tc,
                      const double *vpot, int absorb, const double *vabc,
                      int mode)
{
    int n = ntot_of(ndim, np);
    vphase_swap(n, cst2, b, 0, u, v, vpot, absorb, vabc);
    fft_all(ndim, np, u, v, +1, mode);
    for (int i = 0; i < n; ++i) {
        double th = cst1 * tc[i];
        double c = cos(th), s = sin(th);
        if (mode == ORC_REF) {
            /* spofft.f:152-153: V is updated from the new U. */
            u[i] = c * u[i] + s * v[i];
            v[i] = c * v[i] - s * u[i];
        } else {
            double ur = c * u[i] + s * v[i];
            v[i] = c * v[i] - s * u[i];
            u[i] = ur;
        }
    }
    fft_all(ndim, np, u, v, -1, mode);
    vphase_swap(n, cst2, b, 1, u, v, vpot, absorb, vabc);
}

void orc_spo_step(int ndim, const int *np, double b, double cst1, double cst2,
                  double *u, double *v, const double *tc, const double *vpot,
                  int mode)
{
    orc_spo_step_abs(ndim, np, b, cst1, cst2, u, v, tc, vpot, ORC_ABS_NONE, NULL, mode);
}

/* sodfft.f:3-129.  ORC_ABS_VOPTIC adds cst vabc psi(t) to the leapfrog
 * sum (sodfft.f:66-72).  The mask loop of ORC_ABS_SMOOTHW
 * (sodfft.f:108-116) is never reached in the reference: it compares the
 * 7 characters TPABSOR(1:7) with the 8-character '.SMOOTHW', which cannot
 * match, so ORC_REF propagates without absorber; ORC_EXACT multiplies
 * psi(t+dt) by the mask of each point (the loop itself would use the mask
 * of point i for point i + N/2 as well). */
void orc_sod_step_abs(int ndim, const int *np, double b, double cst, double *u1,
                      double *u2, double *v1, double *v2, const double *tc,
                      const double *vpot, int absorb, const double *vabc,
                      double *wr, double *wi, int mode)
{
    int n = ntot_of(ndim, np), nh = n / 2;
    for (int i = 0; i < n; ++i) {
        if (absorb == ORC_ABS_VOPTIC) {
            wr[i] = u1[i] + cst * vpot[i] * v2[i] + cst * vabc[i] * u2[i];
            wi[i] = v1[i] - cst * vpot[i] * u2[i] + cst * vabc[i] * v2[i];
        } else {
            wr[i] = u1[i] + cst * vpot[i] * v2[i];
            wi[i] = v1[i] - cst * vpot[i] * u2[i];
        }
        u1[i] = u2[i];
        v1[i] = v2[i];
    }
    for (int i = 0; i < nh; ++i) {
        double t = u2[i]; u2[i] = u2[i + nh]; u2[i + nh] = t;
        t = v2[i]; v2[i] = v2[i + nh]; v2[i + nh] = t;
    }
    fft_all(ndim, np, u2, v2, +1, mode);
    for (int i = 0; i < n; ++i) {
        u2[i] = tc[i] * u2[i];
        v2[i] = tc[i] * v2[i];
    }
    fft_all(ndim, np, u2, v2, -1, mode);
    for (int i = 0; i < nh; ++i) {
        double uaux = wr[i + nh] + cst * b * v2[i];
        double vaux = wi[i + nh] - cst * b * u2[i];
        double un = wr[i] + cst * b * v2[i + nh];
        double vn = wi[i] - cst * b * u2[i + nh];
        if (absorb == ORC_ABS_SMOOTHW && mode == ORC_EXACT) {
            un = vabc[i] * un;
            vn = vabc[i] * vn;
            uaux = vabc[i + nh] * uaux;
            vaux = vabc[i + nh] * vaux;
        }
        u2[i] = un;
        v2[i] = vn;
        u2[i + nh] = uaux;
        v2[i + nh] = vaux;
    }
}

void orc_sod_step(int ndim, const int *np, double b, double cst, double *u1,
                  double *u2, double *v1, double *v2, const double *tc,
                  const double *vpot, double *wr, double *wi, int mode)
{
    orc_sod_step_abs(ndim, np, b, cst, u1, u2, v1, v2, tc, vpot, ORC_ABS_NONE, NULL, wr, wi,
                     mode);
}

/* Fortran 77 trip count of DO T=M1,M2,M3: MAX(INT((M2-M1+M3)/M3),0). */
static int trip(double m1, double m2, double m3)
{
    double c = (m2 - m1 + m3) / m3;
    int k = (int)c;
    return k > 0 ? k : 0;
}

int orc_nsteps(double ti, double tf, double dt) { return trip(ti + dt, tf + dt, dt); }

double orc_fcorrel(int n, const double *a, const double *b)
{
    double s = 0;
    for (int i = 0; i < n; ++i) s = s + a[i] * b[i];
    return s;
}

double orc_ecnorm(int n, const double *a) { return sqrt(orc_fcorrel(n, a, a)); }

/* eigenergfft.f:1-103 EFFT4: <V> + <T> through a forward/inverse pair on
 * the half-swapped copy, scaled by b = 1/N. */
double orc_efft4(int ndim, const int *np, const double *u, const double *v,
                 const double *tc, const double *vpot, int mode)
{
    int n = ntot_of(ndim, np), nh = n / 2;
    double b = 1.0 / n;
    double su1 = 0, sv1 = 0, su2 = 0, sv2 = 0;
    for (int i = 0; i < n; ++i) {
        su1 = su1 + u[i] * vpot[i] * u[i];
        sv1 = sv1 + v[i] * vpot[i] * v[i];
    }
    double *wu = malloc(sizeof(double) * 2 * n), *wv = wu + n;
    for (int i = 0; i < nh; ++i) {
        wu[i] = u[i + nh]; wv[i] = v[i + nh];
        wu[i + nh] = u[i]; wv[i + nh] = v[i];
    }
    fft_all(ndim, np, wu, wv, +1, mode);
    for (int i = 0; i < n; ++i) { wu[i] *= tc[i]; wv[i] *= tc[i]; }
    fft_all(ndim, np, wu, wv, -1, mode);
    for (int i = 0; i < nh; ++i) {
        su2 = su2 + u[i] * b * wu[i + nh];
        sv2 = sv2 + v[i] * b * wv[i + nh];
        su2 = su2 + u[i + nh] * b * wu[i];
        sv2 = sv2 + v[i + nh] * b * wv[i];
    }
    free(wu);
    return su1 + su2 + sv1 + sv2;
}

static void corr(int n, const double *u0, const double *v0, const double *u,
                 const double *v, double *cr, double *ci)
{
    *cr = orc_fcorrel(n, u0, u) + orc_fcorrel(n, v0, v);
    *ci = orc_fcorrel(n, u0, v) - orc_fcorrel(n, v0, u);
}

/* pspofft.f:72-145 with PRTCRL='.NO' (no print-time energy checks). */
int orc_pspofft_abs(int ndim, const int *np, double dt, double ti, double tf,
                    int lc, const double *sh, const double *sm, double *u1,
                    double *v1, const double *vpot, int absorb,
                    const double *vabc, double *cr, double *ci, int maxsamp,
                    double *e, double *nrm, int mode)
{
    int n = ntot_of(ndim, np);
    double cst1 = dt / ORC_SHFS, cst2 = dt / (2.0 * ORC_SHFS), b;
    double *u0 = malloc(sizeof(double) * 3 * n), *v0 = u0 + n, *tc = u0 + 2 * n;
    memcpy(u0, u1, sizeof(double) * n);
    memcpy(v0, v1, sizeof(double) * n);
    orc_digt(ndim, np, sh, sm, tc, &b, mode);
    int nc = 0, mc = 0;
    if (maxsamp > 0) {
        corr(n, u0, v0, u1, v1, &cr[0], &ci[0]);
        if (e) e[0] = orc_efft4(ndim, np, u1, v1, tc, vpot, mode);
        if (nrm) nrm[0] = orc_fcorrel(n, u1, u1) + orc_fcorrel(n, v1, v1);
    }
    int ns = orc_nsteps(ti, tf, dt);
    for (int s = 0; s < ns; ++s) {
        orc_spo_step_abs(ndim, np, b, cst1, cst2, u1, v1, tc, vpot, absorb, vabc, mode);
        if (++mc == lc) {
            ++nc;
            mc = 0;
            if (nc < maxsamp) {
                corr(n, u0, v0, u1, v1, &cr[nc], &ci[nc]);
                if (e) e[nc] = orc_efft4(ndim, np, u1, v1, tc, vpot, mode);
                if (nrm) nrm[nc] = orc_fcorrel(n, u1, u1) + orc_fcorrel(n, v1, v1);
            }
        }
    }
    free(u0);
    return nc + 1;
}

int orc_pspofft(int ndim, const int *np, double dt, double ti, double tf,
                int lc, const double *sh, const double *sm, double *u1,
                double *v1, const double *vpot, double *cr, double *ci,
                int maxsamp, double *e, double *nrm, int mode)
{
    return orc_pspofft_abs(ndim, np, dt, ti, tf, lc, sh, sm, u1, v1, vpot, ORC_ABS_NONE,
                           NULL, cr, ci, maxsamp, e, nrm, mode);
}

/* psodfft.f:75-200 with PRTCRL='.NO': two half-strength start-up steps
 * from psi0, then leapfrog from TI+2DT. */
int orc_psodfft_abs(int ndim, const int *np, double dt, double ti, double tf,
                    int lc, const double *sh, const double *sm, double *u1,
                    double *v1, const double *vpot, int absorb,
                    const double *vabc, double *cr, double *ci, int maxsamp,
                    int mode)
{
    int n = ntot_of(ndim, np);
    double cst = 2.0 * dt / ORC_SHFS, b;
    double *w = malloc(sizeof(double) * 7 * n);
    double *u0 = w, *v0 = w + n, *u2 = w + 2 * n, *v2 = w + 3 * n, *tc = w + 4 * n;
    double *wr = w + 5 * n, *wi = w + 6 * n;
    memcpy(u0, u1, sizeof(double) * n);
    memcpy(u2, u1, sizeof(double) * n);
    memcpy(v0, v1, sizeof(double) * n);
    memcpy(v2, v1, sizeof(double) * n);
    orc_digt(ndim, np, sh, sm, tc, &b, mode);
    int nc = 0, mc;
    if (maxsamp > 0) corr(n, u0, v0, u2, v2, &cr[0], &ci[0]);
    orc_sod_step_abs(ndim, np, b, cst / 2.0, u1, u2, v1, v2, tc, vpot, absorb, vabc, wr, wi,
                     mode);
    orc_sod_step_abs(ndim, np, b, cst / 2.0, u1, u2, v1, v2, tc, vpot, absorb, vabc, wr, wi,
                     mode);
    mc = 1;
    if (mc == lc) {
        ++nc;
        mc = 0;
        if (nc < maxsamp) corr(n, u0, v0, u2, v2, &cr[nc], &ci[nc]);
    }
    memcpy(u1, u0, sizeof(double) * n);
    memcpy(v1, v0, sizeof(double) * n);
    int ns = trip(ti + 2 * dt, tf + dt, dt);
    for (int s = 0; s < ns; ++s) {
        orc_sod_step_abs(ndim, np, b, cst, u1, u2, v1, v2, tc, vpot, absorb, vabc, wr, wi, mode);
        if (++mc == lc) {
            ++nc;
            mc = 0;
            if (nc < maxsamp) corr(n, u0, v0, u2, v2, &cr[nc], &ci[nc]);
        }
    }
    /* The reference leaves psi(t) in (U2, V2) and psi(t-dt) in (U1, V1);
     * return psi(t) in u1/v1 for the caller. */
    memcpy(u1, u2, sizeof(double) * n);
    memcpy(v1, v2, sizeof(double) * n);
    free(w);
    return nc + 1;
}

int orc_psodfft(int ndim, const int *np, double dt, double ti, double tf,
                int lc, const double *sh, const double *sm, double *u1,
                double *v1, const double *vpot, double *cr, double *ci,
                int maxsamp, int mode)
{
    return orc_psodfft_abs(ndim, np, dt, ti, tf, lc, sh, sm, u1, v1, vpot, ORC_ABS_NONE,
                           NULL, cr, ci, maxsamp, mode);
}

/* --------------------------------------------------- FD Hamiltonian */

/* au1d.f:33-67, au2di.f:1-239, au3d.f:56-113: per axis
 * shm_a (u[-2] - 16 u[-1] + 30 u[0] - 16 u[+1] + u[+2]) plus V u, with
 * neighbours outside the grid taken as zero. */
void orc_au(int ndim, const int *np, const double *shm, const double *vpot,
            const double *u, double *hu)
{
    int n = ntot_of(ndim, np);
    int stride[3] = {1, 1, 1};
    for (int a = 1; a < ndim; ++a) stride[a] = stride[a - 1] * np[a - 1];
    for (int i = 0; i < n; ++i) {
        double acc = vpot[i] * u[i];
        for (int a = 0; a < ndim; ++a) {
            int ia = (i / stride[a]) % np[a], s = stride[a];
            double t = 30.0 * u[i];
            if (ia >= 2) t += u[i - 2 * s];
            if (ia >= 1) t -= 16.0 * u[i - s];
            if (ia + 1 < np[a]) t -= 16.0 * u[i + s];
            if (ia + 2 < np[a]) t += u[i + 2 * s];
            acc += shm[a] * t;
        }
        hu[i] = acc;
    }
}

double orc_eigenerg(int ndim, const int *np, const double *shm,
                    const double *u, const double *vpot)
{
    int n = ntot_of(ndim, np);
    double *w = malloc(sizeof(double) * n);
    orc_au(ndim, np, shm, vpot, u, w);
    double s = orc_fcorrel(n, u, w);
    free(w);
    return s;
}

/* ----------------------------------------------------------- Lanczos */

/* reort.f:54-80: one classical Gram-Schmidt pass. */
void orc_reort(int n, int mcols, const double *basis, double *w)
{
    double rvs[64], *rv = mcols > 64 ? malloc(sizeof(double) * mcols) : rvs;
    for (int j = 0; j < mcols; ++j) {
        double s = 0;
        for (int i = 0; i < n; ++i) s = s + w[i] * basis[(size_t)j * n + i];
        rv[j] = s;
    }
    for (int j = 0; j < mcols; ++j)
        for (int i = 0; i < n; ++i) w[i] = w[i] - rv[j] * basis[(size_t)j * n + i];
    if (rv != rvs) free(rv);
}

/* lnz.f:65-119. */
int orc_lnz(int ndim, const int *np, int m, int nreort, const double *shm,
            const double *vpot, const double *start, double *al, double *bt,
            double *basis)
{
    int n = ntot_of(ndim, np), it = 0;
    double *w = malloc(sizeof(double) * 3 * n), *vi = w + n, *v0 = w + 2 * n;
    memcpy(w, start, sizeof(double) * n);
    memset(vi, 0, sizeof(double) * n);
    memset(basis, 0, sizeof(double) * (size_t)n * nreort);
    double bt0 = orc_ecnorm(n, w);
    if (bt0 == 0) { free(w); return 0; }
    for (int I = 1; I <= m; ++I) {
        for (int j = 0; j < n; ++j) {
            v0[j] = vi[j];
            vi[j] = w[j] / bt0;
            if (I <= nreort) basis[(size_t)(I - 1) * n + j] = vi[j];
        }
        if (I <= nreort) it = I;
        orc_au(ndim, np, shm, vpot, vi, w);
        double s = 0;
        for (int j = 0; j < n; ++j) {
            w[j] = w[j] - bt0 * v0[j];
            s = s + vi[j] * w[j];
        }
        al[I - 1] = s;
        for (int j = 0; j < n; ++j) w[j] = w[j] - al[I - 1] * vi[j];
        orc_reort(n, it, basis, w);
        bt[I - 1] = orc_ecnorm(n, w);
        bt0 = bt[I - 1];
    }
    free(w);
    return m;
}

/* Implicit QL with Wilkinson shifts on a symmetric tridiagonal matrix
 * (the method of dsteqr.f for ABSTOL <= 0), eigenvectors accumulated from
 * the identity, then sorted ascending. */
int orc_tridiag_eig(int m, const double *d_in, const double *e_in, double *w,
                    double *z)
{
    double *d = w, *e = malloc(sizeof(double) * (m + 1));
    memcpy(d, d_in, sizeof(double) * m);
    for (int i = 0; i < m - 1; ++i) e[i] = e_in[i];
    e[m - 1] = 0;
    for (int i = 0; i < m * m; ++i) z[i] = 0;
    for (int i = 0; i < m; ++i) z[i * m + i] = 1;
    for (int l = 0; l < m; ++l) {
        int iter = 0, mm;
        do {
            for (mm = l; mm < m - 1; ++mm) {
                double dd = fabs(d[mm]) + fabs(d[mm + 1]);
                if (fabs(e[mm]) <= 2.2204460492503131e-16 * dd)
                    break;
            }
            if (mm != l) {
                if (++iter > 60) { free(e); return l + 1; }
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot(g, 1.0);
                g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
                double s = 1, c = 1, p = 0;
                int i;
                for (i = mm - 1; i >= l; --i) {
                    double f = s * e[i], bb = c * e[i];
                    r = hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0) {
                        d[i + 1] -= p;
                        e[mm] = 0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * bb;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - bb;
                    for (int k = 0; k < m; ++k) {
                        double zf = z[(i + 1) * m + k];
                        z[(i + 1) * m + k] = s * z[i * m + k] + c * zf;
                        z[i * m + k] = c * z[i * m + k] - s * zf;
                    }
                }
                if (r == 0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[mm] = 0;
            }
        } while (mm != l);
    }
    /* selection sort ascending, carrying vectors */
    for (int i = 0; i < m - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < m; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            double t = d[i]; d[i] = d[k]; d[k] = t;
            for (int r = 0; r < m; ++r) {
                t = z[i * m + r]; z[i * m + r] = z[k * m + r]; z[k * m + r] = t;
            }
        }
    }
    free(e);
    return 0;
}

/* Sturm count: number of eigenvalues < x (dlaebz.f, IJOB=1 recurrence
 * d_j - e_{j-1}^2 / q - x with the pivot guard). */
static int sturm(int m, const double *d, const double *e2, double pivmin, double x)
{
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q <= 0) ++cnt;
    for (int j = 1; j < m; ++j) {
        q = d[j] - e2[j - 1] / q - x;
        if (fabs(q) < pivmin) q = -pivmin;
        if (q <= 0) ++cnt;
    }
    return cnt;
}

/* dstebz.f RANGE='A' for one unreduced block: Gershgorin interval widened
 * as dstebz.f does, bisection of each eigenvalue's interval until its
 * width is below max(abstol, pivmin, reltol*max|end|), midpoint returned.
 * PARITY UNPINNED against dstebz.f's exact interval sequence beyond the
 * tolerance: used for ABSTOL > 0 only. */
int orc_tridiag_bisect(int m, const double *d, const double *e, double abstol,
                       double *w)
{
    const double ulp = 2.2204460492503131e-16, safemn = 2.2250738585072014e-308;
    const double fudge = 2.1, reltol = ulp * 2.0;
    double *e2 = malloc(sizeof(double) * (m + 1));
    double pivmin = 1.0;
    for (int j = 0; j < m - 1; ++j) {
        e2[j] = e[j] * e[j];
        if (e2[j] > pivmin) pivmin = e2[j];
    }
    pivmin = pivmin * safemn;
    double gl = d[0], gu = d[0];
    for (int j = 0; j < m; ++j) {
        double r = (j > 0 ? fabs(e[j - 1]) : 0) + (j < m - 1 ? fabs(e[j]) : 0);
        if (d[j] - r < gl) gl = d[j] - r;
        if (d[j] + r > gu) gu = d[j] + r;
    }
    double tnorm = fmax(fabs(gl), fabs(gu));
    gl = gl - fudge * tnorm * ulp * m - fudge * 2.0 * pivmin;
    gu = gu + fudge * tnorm * ulp * m + fudge * pivmin;
    double atoli = abstol <= 0 ? ulp * tnorm : abstol;
    for (int k = 0; k < m; ++k) {
        double a = gl, bb = gu;
        for (int it = 0; it < 200; ++it) {
            double tmp1 = fabs(bb - a), tmp2 = fmax(fabs(bb), fabs(a));
            if (tmp1 < fmax(atoli, fmax(pivmin, reltol * tmp2))) break;
            double c = 0.5 * (a + bb);
            if (sturm(m, d, e2, pivmin, c) > k) bb = c;
            else a = c;
        }
        w[k] = 0.5 * (a + bb);
    }
    free(e2);
    return 0;
}

/* DGEMV 'T' (dgemv.f): y_j = sum_i A(i,j) x_i. */
static void gemv_t(int rows, int cols, const double *A, int lda, const double *x, double *y)
{
    for (int j = 0; j < cols; ++j) {
        double t = 0;
        for (int i = 0; i < rows; ++i) t = t + A[(size_t)j * lda + i] * x[i];
        y[j] = t;
    }
}

/* DGEMV 'N' (dgemv.f): y = sum_j x_j A(:,j), column by column. */
static void gemv_n(int rows, int cols, const double *A, int lda, const double *x, double *y)
{
    for (int i = 0; i < rows; ++i) y[i] = 0;
    for (int j = 0; j < cols; ++j) {
        double t = x[j];
        for (int i = 0; i < rows; ++i) y[i] = y[i] + t * A[(size_t)j * lda + i];
    }
}

static int tridiag(int mp, const double *al, const double *bt, double abstol,
                   double *lam, double *q)
{
    int info = orc_tridiag_eig(mp, al, bt, lam, q);
    if (info == 0 && abstol > 0) info = orc_tridiag_bisect(mp, al, bt, abstol, lam);
    return info;
}

/* sil.f:70-181 without absorber: Lanczos on Re psi and on Im psi
 * separately, each projected, rotated by exp(-i lambda cst), and
 * recombined.  ORC_REF keeps sil.f:98-99, where -sin multiplies the
 * already cos-scaled coefficient. */
/* sil.f:3-185 with the .SMOOTHW mask of sil.f:171-180 when mask is not
 * NULL: psi = VABC (U + WU) + i VABC (WV + V). */
static int plnz_step_mask(int ndim, const int *np, int mp, double abstol, double cst,
                          const double *shm, const double *vpot, double *u, double *v,
                          const double *mask, int mode);

int orc_plnz_step(int ndim, const int *np, int mp, double abstol, double cst,
                  const double *shm, const double *vpot, double *u, double *v,
                  int mode)
{
    return plnz_step_mask(ndim, np, mp, abstol, cst, shm, vpot, u, v, NULL, mode);
}

static int plnz_step_mask(int ndim, const int *np, int mp, double abstol, double cst,
                          const double *shm, const double *vpot, double *u, double *v,
                          const double *mask, int mode)
{
    int n = ntot_of(ndim, np), info;
    if (mp < 1 || mp > 64) return -2;
    double *basis = malloc(sizeof(double) * (size_t)n * mp);
    double *wv = malloc(sizeof(double) * 2 * n), *wu = wv + n;
    double al[64], bt[64], lam[64], q[64 * 64], c[64], u2[64], ur[64], wvs[64], work[64];

    orc_lnz(ndim, np, mp, mp, shm, vpot, u, al, bt, basis);
    info = tridiag(mp, al, bt, abstol, lam, q);
    if (info) goto done;
    gemv_t(n, mp, basis, n, u, c);
    gemv_t(mp, mp, q, mp, c, u2);
    for (int i = 0; i < mp; ++i) {
        double ang = lam[i] * cst;
        if (mode == ORC_REF) {
            ur[i] = cos(ang) * u2[i];
            wvs[i] = -sin(ang) * ur[i];
        } else {
            ur[i] = cos(ang) * u2[i];
            wvs[i] = -sin(ang) * u2[i];
        }
    }
    gemv_n(mp, mp, q, mp, ur, work);
    gemv_n(n, mp, basis, n, work, u);
    gemv_n(mp, mp, q, mp, wvs, work);
    gemv_n(n, mp, basis, n, work, wv);

    orc_lnz(ndim, np, mp, mp, shm, vpot, v, al, bt, basis);
    info = tridiag(mp, al, bt, abstol, lam, q);
    if (info) goto done;
    gemv_t(n, mp, basis, n, v, c);
    gemv_t(mp, mp, q, mp, c, u2);
    for (int i = 0; i < mp; ++i) {
        double ang = lam[i] * cst;
        wvs[i] = sin(ang) * u2[i];
        ur[i] = cos(ang) * u2[i];
    }
    gemv_n(mp, mp, q, mp, ur, work);
    gemv_n(n, mp, basis, n, work, v);
    gemv_n(mp, mp, q, mp, wvs, work);
    gemv_n(n, mp, basis, n, work, wu);
    for (int i = 0; i < n; ++i) {
        if (mask) {
            u[i] = mask[i] * (u[i] + wu[i]);
            v[i] = mask[i] * (wv[i] + v[i]);
        } else {
            u[i] = u[i] + wu[i];
            v[i] = wv[i] + v[i];
        }
    }
done:
    free(basis);
    free(wv);
    return info;
}

/* plnz.f:72-140 with PRTCRL='.NO'. */
int orc_plnz(int ndim, const int *np, int mp, double abstol, double dt,
             double ti, double tf, int lc, const double *shm, double *u,
             double *v, const double *vpot, double *cr, double *ci,
             int maxsamp, int mode)
{
    return orc_plnz_abs(ndim, np, mp, abstol, dt, ti, tf, lc, shm, u, v, vpot, NULL, cr, ci,
                        maxsamp, mode);
}

int orc_plnz_abs(int ndim, const int *np, int mp, double abstol, double dt,
                 double ti, double tf, int lc, const double *shm, double *u,
                 double *v, const double *vpot, const double *mask, double *cr,
                 double *ci, int maxsamp, int mode)
{
    int n = ntot_of(ndim, np);
    double cst = dt / ORC_SHFS;
    double *u0 = malloc(sizeof(double) * 2 * n), *v0 = u0 + n;
    memcpy(u0, u, sizeof(double) * n);
    memcpy(v0, v, sizeof(double) * n);
    int nc = 0, mc = 0;
    if (maxsamp > 0) corr(n, u0, v0, u, v, &cr[0], &ci[0]);
    int ns = orc_nsteps(ti, tf, dt);
    for (int s = 0; s < ns; ++s) {
        if (plnz_step_mask(ndim, np, mp, abstol, cst, shm, vpot, u, v, mask, mode)) {
            free(u0);
            return -1;
        }
        if (++mc == lc) {
            ++nc;
            mc = 0;
            if (nc < maxsamp) corr(n, u0, v0, u, v, &cr[nc], &ci[nc]);
        }
    }
    free(u0);
    return nc + 1;
}

/* ef.f:1-60 EF: the laser field at time t (fs) in the reference's units.
 * efc: 0 .NONE/.NULL, 1 .GAUS, 2 .GGAUS, 3 .SIN2.  .GGAUS raises the
 * default-REAL (f32) literal 0.693147181 to 1/(2 KL), as the reference
 * does. */
double orc_ef(int efc, double e0, double t0, double td, double tp, double omg,
              double sni, int kl, double t)
{
    const double pi = 3.14159265358979323846;
    const double fat = 2.1132e-9, fate = 2.29371276e+17 * 3.335664e-28, fatt = 1.5192676;
    double ef;
    if (efc == 1) {
        ef = e0 * exp(-(4.0 * 0.693147181 * (t - t0) * (t - t0) / (tp * tp))) *
             cos(omg * t * fatt + sni);
        ef = ef * fate;
    } else if (efc == 2) {
        double a1 = 1.0 / (2.0 * kl);
        double tpa = tp / pow((double)0.693147181f, a1);
        ef = fat * sqrt(e0 * exp(-pow((t - t0) / tpa, 2.0 * kl))) * cos(omg * t * fatt + sni);
    } else if (efc == 3) {
        ef = e0 * cos(pi * (t - td) / tp) * cos(pi * (t - td) / tp) *
             cos(omg * (t - td) * fatt + sni);
        ef = ef * fate;
    } else {
        ef = 0.0;
    }
    return ef;
}

/* pplnz.f:72-300 with PRTCRL='.NO': DO T=TI+DT,TF+DT/2,DT, each step
 * VEFF = VP - DM EF(T) at the step's end time T (chpot.f:58-62) and one
 * PLNZAUX step with VEFF.  veff (may be NULL) receives the last VEFF,
 * which the reference leaves in VPOT. */
int orc_pplnz(int ndim, const int *np, int mp, double abstol, double dt,
              double ti, double tf, int lc, const double *shm, double *u,
              double *v, const double *vp, const double *dm, int efc,
              double e0, double t0, double td, double tp, double omg,
              double sni, int kl, double *veff, double *cr, double *ci,
              int maxsamp, int mode)
{
    return orc_pplnz_abs(ndim, np, mp, abstol, dt, ti, tf, lc, shm, u, v, vp, dm, efc, e0, t0,
                         td, tp, omg, sni, kl, NULL, veff, cr, ci, maxsamp, mode);
}

int orc_pplnz_abs(int ndim, const int *np, int mp, double abstol, double dt,
                  double ti, double tf, int lc, const double *shm, double *u,
                  double *v, const double *vp, const double *dm, int efc,
                  double e0, double t0, double td, double tp, double omg,
                  double sni, int kl, const double *mask, double *veff, double *cr,
                  double *ci, int maxsamp, int mode)
{
    int n = ntot_of(ndim, np);
    double cst = dt / ORC_SHFS;
    double *u0 = malloc(sizeof(double) * 3 * n), *v0 = u0 + n, *ve = u0 + 2 * n;
    memcpy(u0, u, sizeof(double) * n);
    memcpy(v0, v, sizeof(double) * n);
    memcpy(ve, vp, sizeof(double) * n);
    int nc = 0, mc = 0;
    if (maxsamp > 0) corr(n, u0, v0, u, v, &cr[0], &ci[0]);
    int ns = trip(ti + dt, tf + dt / 2, dt);
    double t = ti + dt;
    for (int s = 0; s < ns; ++s, t += dt) {
        double e = orc_ef(efc, e0, t0, td, tp, omg, sni, kl, t);
        for (int i = 0; i < n; ++i) ve[i] = vp[i] - dm[i] * e;
        if (plnz_step_mask(ndim, np, mp, abstol, cst, shm, ve, u, v, mask, mode)) {
            free(u0);
            return -1;
        }
        if (++mc == lc) {
            ++nc;
            mc = 0;
            if (nc < maxsamp) corr(n, u0, v0, u, v, &cr[nc], &ci[nc]);
        }
    }
    if (veff) memcpy(veff, ve, sizeof(double) * n);
    free(u0);
    return nc + 1;
}

/* gauss.f GET_GAUSS: rnorm exp(-ln2 (coord/ra)^2), ln2 to 8 digits. */
static double get_gauss(double ra, double coord, double rnorm)
{
    double q = coord / ra;
    return rnorm * exp(-(0.69314718 * (q * q)));
}

/* 1/sqrt(sum of GET_GAUSS^2) over an axis (init_cond.f:60-66). */
static double gauss_norm(int np, double xi, double sh, double x0, double ra)
{
    double r = 0;
    for (int i = 0; i < np; ++i) {
        double g = get_gauss(ra, xi + i * sh - x0, 1.0);
        r = r + g * g;
    }
    return 1.0 / sqrt(r);
}

/* init_cond.f:1-200 INIT_COND.  mode: ORC_INIT_CALC (1-D or 2-D Gaussian
 * packets), ORC_INIT_GETR/GETR1 (2-D: the real 1-D eigenvector u[0..np1)
 * along axis 1 times a Gaussian along axis 2), ORC_INIT_GETR2 (eigenvector
 * along axis 2, Gaussian along axis 1), ORC_INIT_GETC/GETC1/GETC2 (the same
 * with the complex eigenvector u + i v).  Returns -1 for a combination the
 * reference rejects. */
int orc_init_cond(int ndim, const int *np, int mode, const double *xi,
                  const double *sh, const double *x0, const double *ra,
                  const double *rk, double *u, double *v)
{
    if (ndim == 1 && mode == ORC_INIT_CALC) {
        double rn = gauss_norm(np[0], xi[0], sh[0], x0[0], ra[0]);
        for (int i = 0; i < np[0]; ++i) {
            double c = xi[0] + i * sh[0] - x0[0];
            double g = get_gauss(ra[0], c, rn);
            double a = rk[0] * c;
            u[i] = cos(a) * g;
            v[i] = sin(a) * g;
        }
        return 0;
    }
    if (ndim != 2) return -1;
    /* eig: axis of the given factor (w1, w2); gx: the Gaussian's axis */
    int eig, gx;
    double *w1 = malloc(sizeof(double) * 2 * (np[0] > np[1] ? np[0] : np[1]));
    double *w2 = w1 + (np[0] > np[1] ? np[0] : np[1]);
    if (mode == ORC_INIT_CALC) {
        eig = 0;
        gx = 1;
        double rn1 = gauss_norm(np[0], xi[0], sh[0], x0[0], ra[0]);
        for (int i = 0; i < np[0]; ++i) {
            double c = xi[0] + i * sh[0] - x0[0];
            double g = get_gauss(ra[0], c, rn1);
            double a = rk[0] * c;
            w1[i] = cos(a) * g;
            w2[i] = -sin(a) * g;
        }
    } else if (mode == ORC_INIT_GETR || mode == ORC_INIT_GETC) {
        eig = 0;
        gx = 1;
    } else if (mode == ORC_INIT_GETR2 || mode == ORC_INIT_GETC2) {
        eig = 1;
        gx = 0;
    } else {
        free(w1);
        return -1;
    }
    if (mode != ORC_INIT_CALC)
        for (int i = 0; i < np[eig]; ++i) {
            w1[i] = u[i];
            w2[i] = (mode == ORC_INIT_GETC || mode == ORC_INIT_GETC2) ? v[i] : 0.0;
        }
    double rn = gauss_norm(np[gx], xi[gx], sh[gx], x0[gx], ra[gx]);
    for (int j = 0; j < np[gx]; ++j) {
        double c = xi[gx] + j * sh[gx] - x0[gx];
        double g = get_gauss(ra[gx], c, rn);
        double a = rk[gx] * c;
        double ct = cos(a) * g, st = sin(a) * g;
        for (int i = 0; i < np[eig]; ++i) {
            int k = eig == 0 ? j * np[0] + i : i * np[0] + j;
            if (mode == ORC_INIT_GETR || mode == ORC_INIT_GETR2) {
                u[k] = w1[i] * ct;
                v[k] = -w1[i] * st;
            } else {
                u[k] = w1[i] * ct + w2[i] * st;
                v[k] = -w1[i] * st + w2[i] * ct;
            }
        }
    }
    free(w1);
    return 0;
}

/* absorb.f:1-130 ABSORBINGBC: VABC on the grid (x fastest) for kind
 * ORC_ABG_SMOOTHW (.SMOOTHW mask; 1-D and 2-D), ORC_ABG_SMRS (.SMRS
 * mask), ORC_ABG_VOPTIC_LINEAR, ORC_ABG_VOPTIC_KOSLOFF, ORC_ABG_VOPTIC_POLIN
 * (imaginary potentials; 1-D).  alpha receives 1 - 0.1 (XF - AR) (the
 * default-REAL 0.1) for the 1-D .SMOOTHW.  Returns -1 where the reference
 * stops. */
int orc_absorbingbc(int ndim, const int *np, int kind, const double *al,
                    const double *ar, const double *sh, const double *xi,
                    const double *xf, const double *voi, double *vabc,
                    double *alpha)
{
    if (ndim == 1) {
        for (int i = 0; i < np[0]; ++i) {
            double x1 = xi[0] + (i + 1 - 1.0) * sh[0], v;
            int left = x1 < al[0], right = x1 > ar[0];
            switch (kind) {
            case ORC_ABG_SMOOTHW:
                if (left)
                    v = exp(voi[0] * pow(x1 - al[0], voi[1]) / (x1 + sh[0] - xi[0]));
                else if (right)
                    v = exp(-(voi[0] * (x1 - ar[0]) / pow(xf[0] + sh[0] - x1, voi[1])));
                else
                    v = 1.0;
                break;
            case ORC_ABG_SMRS:
                if (left)
                    v = 1.0 - voi[0] * ((x1 - al[0]) * (x1 - al[0]));
                else if (right)
                    v = 1.0 - voi[0] * ((x1 - ar[0]) * (x1 - ar[0]));
                else
                    v = 1.0;
                break;
            case ORC_ABG_VOPTIC_LINEAR:
                if (left)
                    v = -voi[0] * (x1 - al[0]) / (x1 + sh[0] - xi[0]);
                else if (right)
                    v = voi[0] * (x1 - ar[0]) / (xf[0] + sh[0] - x1);
                else
                    v = 0.0;
                break;
            case ORC_ABG_VOPTIC_KOSLOFF:
                /* absorb.f: the argument of COSH does not depend on X1 */
                if (left)
                    v = voi[0] / cosh(xi[0] - al[0] / voi[1]);
                else if (right)
                    v = voi[0] / cosh(xf[0] - al[0] / voi[1]);
                else
                    v = 0.0;
                break;
            case ORC_ABG_VOPTIC_POLIN:
                if (left)
                    v = voi[0] * pow(x1 - al[0], voi[1]);
                else if (right)
                    v = voi[0] * pow(x1 - ar[0], voi[1]);
                else
                    v = 0.0;
                break;
            default:
                return -1;
            }
            vabc[i] = v;
        }
        if (kind == ORC_ABG_SMOOTHW && alpha) *alpha = 1 - (double)0.1f * (xf[0] - ar[0]);
        return 0;
    }
    if (ndim != 2 || kind != ORC_ABG_SMOOTHW) return -1;
    for (int j = 0; j < np[1]; ++j) {
        double x2 = xi[1] + (j + 1 - 1.0) * sh[1];
        for (int i = 0; i < np[0]; ++i) {
            double x1 = xi[0] + (i + 1 - 1.0) * sh[0], v;
            if (x1 < al[0])
                v = exp(voi[0] * (x1 - al[0]) / (x1 + sh[0] - xi[0]));
            else if (x2 < al[1])
                v = exp(voi[1] * (x2 - al[1]) / (x2 + sh[1] - xi[1]));
            else if (x1 > ar[0])
                v = exp(-voi[0] * (x1 - ar[0]) / (xf[0] + sh[0] - x1));
            else if (x2 > ar[1])
                v = exp(-voi[1] * (x2 - ar[1]) / (xf[1] + sh[1] - x2));
            else
                v = 1.0;
            vabc[(size_t)j * np[0] + i] = v;
        }
    }
    return 0;
}

/* rans.f:32-39. */
double orc_ran(int *nseed)
{
    *nseed = (int)(((long long)*nseed * 1029 + 221591) % 1048576);
    return 1.0 * *nseed / 1048576;
}

/* lanczs.f:76-183.  Eigenvalues: DSTEVX('V','I') (lanczs.f:90) takes the
 * bisection path (dstebz.f) with ABSTOL > 0, whose values are interval
 * midpoints; with ABSTOL <= 0 the QL values.  Vectors: QL's, which the
 * inverse iteration of dstein.f reproduces at either value. */
int orc_lanczs(int ndim, const int *np, int il, int iu, int nreort,
               double abstol, const double *shm, const double *vpot,
               double *eigvl, double *eigvc)
{
    int n = ntot_of(ndim, np), ns = iu - il + 1;
    double *al = malloc(sizeof(double) * 2 * n), *bt = al + n;
    double *start = malloc(sizeof(double) * 6 * n);
    double *wk1 = start + n, *wk2 = start + 2 * n, *v0 = start + 3 * n, *vi = start + 4 * n;
    double *basis = malloc(sizeof(double) * (size_t)n * nreort);
    double *lam = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * (size_t)n * n);
    int seed = 1;
    for (int i = 0; i < n; ++i) start[i] = orc_ran(&seed);
    orc_lnz(ndim, np, n, nreort, shm, vpot, start, al, bt, basis);
    int info = orc_tridiag_eig(n, al, bt, lam, q);
    if (info) goto out;
    if (abstol > 0) {
        double *wb = malloc(sizeof(double) * n);
        orc_tridiag_bisect(n, al, bt, abstol, wb);
        for (int k = il - 1; k < iu; ++k) lam[k] = wb[k];
        free(wb);
    }
    for (int k = 0; k < ns; ++k) {
        double *qk = q + (size_t)(il - 1 + k) * n;
        /* dstein.f normalises each eigenvector so that its largest
         * component is positive. */
        int jm = 0;
        for (int j = 1; j < n; ++j)
            if (fabs(qk[j]) > fabs(qk[jm])) jm = j;
        double sg = qk[jm] < 0 ? -1.0 : 1.0;
        eigvl[k] = lam[il - 1 + k];
        memset(vi, 0, sizeof(double) * n);
        memset(wk2, 0, sizeof(double) * n);
        memset(basis, 0, sizeof(double) * (size_t)n * nreort);
        seed = 1;
        for (int i = 0; i < n; ++i) wk1[i] = orc_ran(&seed);
        double bt0 = orc_ecnorm(n, wk1);
        int it = 0;
        for (int J = 1; J <= n; ++J) {
            for (int i = 0; i < n; ++i) {
                v0[i] = wk2[i];
                wk2[i] = wk1[i] / bt0;
                if (J <= nreort) basis[(size_t)(J - 1) * n + i] = wk2[i];
                vi[i] = vi[i] + sg * qk[J - 1] * wk2[i];
            }
            if (J <= nreort) it = J;
            orc_au(ndim, np, shm, vpot, wk2, wk1);
            double s = 0;
            for (int i = 0; i < n; ++i) {
                wk1[i] = wk1[i] - bt0 * v0[i];
                s = s + wk2[i] * wk1[i];
            }
            for (int i = 0; i < n; ++i) wk1[i] = wk1[i] - s * wk2[i];
            orc_reort(n, it, basis, wk1);
            bt0 = orc_ecnorm(n, wk1);
        }
        memcpy(eigvc + (size_t)k * n, vi, sizeof(double) * n);
    }
out:
    free(al); free(start); free(basis); free(lam); free(q);
    return info;
}

/* -------------------------------------------------------- spectrum */

int orc_lc_rule(double ti, double *tf, double dt, int mf, int change)
{
    for (;;) {
        double sc = 2.0 * (*tf - ti) / dt;
        int nc = (int)(sc / pow(2.0, mf));
        if (nc < 0) return 2147483645;
        if (fabs(sc - nc * pow(2.0, mf)) >= 1.0e-10 && change) {
            *tf = nc * pow(2.0, mf - 1.0) * dt;
            continue;
        }
        return nc;
    }
}

void orc_assemble_corr(int mf, const double *cr, const double *ci, double *u,
                       double *v)
{
    int mc = (1 << mf) / 2;
    /* 1-based indices as in espec.f:1552-1566 */
    double *U = u - 1, *V = v - 1;
    for (int i = 1; i <= 2 * mc; ++i) { U[i] = 0; V[i] = 0; }
    for (int i = 1; i <= mc; ++i) {
        U[mc + i] = U[mc + i] + cr[i - 1];
        U[mc - i + 2] = U[mc + i];
        V[mc + i] = V[mc + i] + ci[i - 1];
        V[mc - i + 2] = -V[mc + i];
    }
    U[1] = U[1] + cr[mc - 1];
    V[1] = V[1] + ci[mc - 1];
    V[1] = -V[1];
}

void orc_hwhm(int mf, double dt, double tf, double tal, double de, double *u,
              double *v)
{
    int n = 1 << mf;
    double dem = -de * 1.5192676, ta = tal * 1.5192676;
    for (int i = 1; i <= n; ++i) {
        double t = -tf + (i - 1.0) * dt;
        double cs = cos(dem * t), sn = sin(dem * t), cstl = exp(-fabs(t) * ta);
        double ui = u[i - 1];
        u[i - 1] = cs * u[i - 1] - sn * v[i - 1];
        v[i - 1] = cs * v[i - 1] + sn * ui;
        u[i - 1] = cstl * u[i - 1];
        v[i - 1] = cstl * v[i - 1];
    }
}

void orc_spectrumtd(int iwin, int mf, double dt, double tf, double wp,
                    double de, double *u, double *v, double *e_out,
                    double *w_out, double *wi_out, int mode)
{
    int n = 1 << mf;
    double enc = 1.0 / (n * dt) * 1.519829846003e-1;
    enc = enc * 27.2113834;
    double en = -n / 2.0 * enc;
    if (de < 0) de = 0;
    if (iwin == 1) {
        double tmax = n * dt;
        double lg = log(1.0 / wp);
        double taux = tmax * tmax / (lg * lg);
        for (int i = 1; i <= n; ++i) {
            double t = -tf + (i - 1.0) * dt;
            double cstl = exp(-(t * t) / taux);
            u[i - 1] = cstl * u[i - 1];
            v[i - 1] = cstl * v[i - 1];
        }
    }
    for (int i = 0; i < n / 2; ++i) {
        double t = u[i]; u[i] = u[n / 2 + i]; u[n / 2 + i] = t;
        t = v[i]; v[i] = v[n / 2 + i]; v[n / 2 + i] = t;
    }
    orc_fftnt(u, v, n, mf, -1, mode);
    int k = 0;
    for (int i = n / 2; i < n; ++i, ++k) {
        e_out[k] = en + de; w_out[k] = u[i]; wi_out[k] = v[i];
        en = en + enc;
    }
    for (int i = 0; i < n / 2; ++i, ++k) {
        e_out[k] = en + de; w_out[k] = u[i]; wi_out[k] = v[i];
        en = en + enc;
    }
}

/* sod.f:3-122 PSODAUX: one finite-difference leapfrog step,
 * psi(t+dt) = psi(t-dt) - i cst H psi(t) with H from AU (the 4th-order FD
 * kinetic operator plus vpot).  (u1, v1) = psi(t-dt) on entry and psi(t)
 * on return, (u2, v2) = psi(t) on entry and psi(t+dt) on return.  The real
 * part is advanced first with H applied to v2; the imaginary part then
 * uses H applied to u1, which by then holds the old u2.  Absorbers
 * (sod.f:55-120): ORC_ABS_VOPTIC subtracts cst vabc psi(t); ORC_ABS_SMOOTHW
 * (and .SMRS) multiplies the new value by vabc. */
void orc_psodaux(int ndim, const int *np, double cst, const double *shm, double *u1,
                 double *u2, double *v1, double *v2, const double *vpot, int absorb,
                 const double *vabc)
{
    int n = ntot_of(ndim, np);
    double *w = malloc(sizeof(double) * n);
    orc_au(ndim, np, shm, vpot, v2, w);
    for (int i = 0; i < n; ++i) {
        double ua = u2[i];
        if (absorb == ORC_ABS_VOPTIC)
            u2[i] = u1[i] + cst * w[i] - cst * vabc[i] * ua;
        else if (absorb == ORC_ABS_SMOOTHW)
            u2[i] = vabc[i] * (u1[i] + cst * w[i]);
        else
            u2[i] = u1[i] + cst * w[i];
        u1[i] = ua;
    }
    orc_au(ndim, np, shm, vpot, u1, w);
    for (int i = 0; i < n; ++i) {
        double va = v2[i];
        if (absorb == ORC_ABS_VOPTIC)
            v2[i] = v1[i] - cst * w[i] - cst * vabc[i] * va;
        else if (absorb == ORC_ABS_SMOOTHW)
            v2[i] = vabc[i] * (v1[i] - cst * w[i]);
        else
            v2[i] = v1[i] - cst * w[i];
        v1[i] = va;
    }
    free(w);
}

static void chpot(int n, const double *vp, const double *dm, double e, double *ve)
{
    for (int i = 0; i < n; ++i) ve[i] = vp[i] - dm[i] * e; /* chpot.f:58-62 */
}

/* ppsod.f:1-548: the pulse-driven FD leapfrog.  CST = 2 DT/SHFS; start-up
 * with a quarter-strength step at the field of (T+DT)/2 and a half-
 * strength step at T+DT (ppsod.f:246-254), psi(t-dt) reset to psi0
 * (ppsod.f:348-351), then NINT((TF+DT/2-T)/DT) full steps with T advanced
 * by DT and the field of each step's end time (ppsod.f:355-362).  C(t)
 * samples: sample 0 is C(TI), then one every LC steps counting the start-
 * up as step 1 (the reference's storage when PRTCRL prints).  nrm the
 * norm <psi|psi> at each sample; veff the last VEFF. */
int orc_ppsod(int ndim, const int *np, double dt, double ti, double tf, int lc,
              const double *shm, double *u, double *v, const double *vp, const double *dm,
              int efc, double e0, double t0, double td, double tp, double omg, double sni,
              int kl, int absorb, const double *vabc, double *veff, double *cr, double *ci,
              double *nrm, int maxsamp)
{
    int n = ntot_of(ndim, np);
    double cst = 2.0 * dt / ORC_SHFS;
    double *u0 = malloc(sizeof(double) * 5 * (size_t)n), *v0 = u0 + n, *u1 = u0 + 2 * n,
           *v1 = u0 + 3 * n, *ve = u0 + 4 * n;
    memcpy(u0, u, sizeof(double) * n);
    memcpy(v0, v, sizeof(double) * n);
    memcpy(u1, u, sizeof(double) * n);
    memcpy(v1, v, sizeof(double) * n);
    memcpy(ve, vp, sizeof(double) * n);
    int nc = 0, mc;
    if (maxsamp > 0) {
        corr(n, u0, v0, u, v, &cr[0], &ci[0]);
        nrm[0] = orc_fcorrel(n, u, u) + orc_fcorrel(n, v, v);
    }
    double t = ti;
    chpot(n, vp, dm, orc_ef(efc, e0, t0, td, tp, omg, sni, kl, (t + dt) / 2.0), ve);
    orc_psodaux(ndim, np, cst / 4.0, shm, u1, u, v1, v, ve, absorb, vabc);
    chpot(n, vp, dm, orc_ef(efc, e0, t0, td, tp, omg, sni, kl, t + dt), ve);
    orc_psodaux(ndim, np, cst / 2.0, shm, u1, u, v1, v, ve, absorb, vabc);
    mc = 1;
    if (mc == lc) {
        mc = 0;
        if (++nc < maxsamp) {
            corr(n, u0, v0, u, v, &cr[nc], &ci[nc]);
            nrm[nc] = orc_fcorrel(n, u, u) + orc_fcorrel(n, v, v);
        }
    }
    memcpy(u1, u0, sizeof(double) * n);
    memcpy(v1, v0, sizeof(double) * n);
    t = ti + dt;
    int nf = (int)lround((tf + dt / 2.0 - t) / dt);
    for (int k = 1; k <= nf; ++k) {
        t = t + dt;
        chpot(n, vp, dm, orc_ef(efc, e0, t0, td, tp, omg, sni, kl, t), ve);
        orc_psodaux(ndim, np, cst, shm, u1, u, v1, v, ve, absorb, vabc);
        if (++mc == lc) {
            mc = 0;
            if (++nc < maxsamp) {
                corr(n, u0, v0, u, v, &cr[nc], &ci[nc]);
                nrm[nc] = orc_fcorrel(n, u, u) + orc_fcorrel(n, v, v);
            }
        }
    }
    if (veff) memcpy(veff, ve, sizeof(double) * n);
    free(u0);
    return nc + 1;
}

/* s2ppsod.f:1-1143: two surfaces A and B coupled by the coordinate
 * operator, each advanced by PSODAUX with its own VEFF = V - DM E(T), with
 * the coupling the routine selects by assignment (TPCRS = '.CORD',
 * s2ppsod.f:95; CSQC = CST AA / SM(1), s2ppsod.f:127): after both
 * surfaces' PSODAUX of a step,
 *   U2A += w V1B, V2A -= w U1B, U2B += w V1A, V2B -= w U1A   (point by point)
 * with w = (CSQC/4) XD, (CSQC/2) XD, CSQC XD for the quarter, half and full
 * steps and XD accumulated from XI(1) by SH(1) before each use, so point i
 * (0-based) sees XI(1) + (i+1) SH(1) (s2ppsod.f:434-444, 632-642,
 * 951-961).  U1/V1 then hold psi(t) (PSODAUX swaps, sod.f:88-121).  The
 * fields are taken at (TI+DT)/2, TI+DT and each T of DO T=TI+2*DT,TF+DT/2,DT
 * (s2ppsod.f:760), and only when E0 /= 0 (s2ppsod.f:260-267, 460-467,
 * 762-775).
 *
 * The printed table (FORMAT 1021) goes to tab[row][12] = (T, ET, ET0, CRA,
 * CIA, CTA, CRB, CIB, CTB, FNTA, FNTB, FNT) as printed:
 *   - at TI when PRTCRL is .YES or .PARTIAL, with FNTB = FNRB**2
 *     (s2ppsod.f:186);
 *   - after the start-up only when PRTCRL is .YES (the .PARTIAL test
 *     compares the reset MC with LC, s2ppsod.f:660-663), with CTB printed in
 *     the CRB column (s2ppsod.f:711);
 *   - in the loop every LC steps for .PARTIAL, every step for .YES; there,
 *     without absorber, INFO = 1 when the total norm has drifted by more
 *     than TOL percent (s2ppsod.f:1030-1034).
 * prtcrl: 0 .NO, 1 .PARTIAL, 2 .YES.  u, v [2n] return U2, V2 (psi(TF));
 * u1, v1 (may be NULL) U1, V1.  Returns the number of rows. */
struct s2ctx {
    int ndim, n, absorb;
    const int *np;
    const double *shm, *vp, *dm, *vabc, *xd;
    double *u0, *v0, *u1, *v1, *ve, *u, *v, *tab;
    double f0, tol;
    int rows, maxrows, *info;
};

static void s2_row(struct s2ctx *c, double t, int quirk)
{
    const int n = c->n;
    double s[2][6];
    for (int k = 0; k < 2; ++k) {
        const int o = k * n;
        double cr, ci;
        corr(n, c->u0 + o, c->v0 + o, c->u + o, c->v + o, &cr, &ci);
        s[k][0] = cr;
        s[k][1] = ci;
        s[k][2] = cr * cr + ci * ci;
        s[k][3] = orc_fcorrel(n, c->u + o, c->u + o) + orc_fcorrel(n, c->v + o, c->v + o);
        s[k][4] = orc_eigenerg(c->ndim, c->np, c->shm, c->u + o, c->ve + o) +
                  orc_eigenerg(c->ndim, c->np, c->shm, c->v + o, c->ve + o);
        s[k][5] = orc_eigenerg(c->ndim, c->np, c->shm, c->u + o, c->vp + o) +
                  orc_eigenerg(c->ndim, c->np, c->shm, c->v + o, c->vp + o);
    }
    const double fntb = quirk == 1 ? s[1][3] * s[1][3] : s[1][3];
    const double fnt = s[0][3] + fntb;
    if (quirk == 0 && c->absorb == ORC_ABS_NONE && fabs((fnt - c->f0) / c->f0) * 1.0e2 > c->tol)
        *c->info = 1;
    if (c->rows < c->maxrows) {
        double *r = c->tab + 12 * (size_t)c->rows;
        r[0] = t;
        r[1] = s[0][4] + s[1][4];
        r[2] = s[0][5] + s[1][5];
        r[3] = s[0][0];
        r[4] = s[0][1];
        r[5] = s[0][2];
        r[6] = quirk == 2 ? s[1][2] : s[1][0];
        r[7] = s[1][1];
        r[8] = s[1][2];
        r[9] = s[0][3];
        r[10] = fntb;
        r[11] = fnt;
    }
    ++c->rows;
}

static void s2_step(struct s2ctx *c, double e, int field, double cstk, double w0)
{
    const int n = c->n;
    for (int k = 0; k < 2; ++k) {
        const int o = k * n;
        if (field) chpot(n, c->vp + o, c->dm + o, e, c->ve + o);
        orc_psodaux(c->ndim, c->np, cstk, c->shm, c->u1 + o, c->u + o, c->v1 + o, c->v + o,
                    c->ve + o, c->absorb, c->vabc);
    }
    double *u = c->u, *v = c->v;
    const double *u1 = c->u1, *v1 = c->v1;
    for (int i = 0; i < n; ++i) {
        const int j = n + i;
        const double w = w0 * c->xd[i];
        u[i] = u[i] + w * v1[j];
        v[i] = v[i] - w * u1[j];
        u[j] = u[j] + w * v1[i];
        v[j] = v[j] - w * u1[i];
    }
}

int orc_s2ppsod(int ndim, const int *np, double dt, double ti, double tf, int lc,
                const double *shm, double xi, double sh, double sm1, double aa, double *u,
                double *v, const double *vp, const double *dm, int efc, double e0, double t0,
                double td, double tp, double omg, double sni, int kl, int absorb,
                const double *vabc, int prtcrl, double tol, double *tab, int maxrows,
                int *info, double *u1out, double *v1out)
{
    const int n = ntot_of(ndim, np), n2 = 2 * n;
    const double cst = 2.0 * dt / ORC_SHFS, csqc = cst * aa / sm1;
    double *buf = malloc(sizeof(double) * 5 * (size_t)n2 + sizeof(double) * n);
    struct s2ctx c;
    c.ndim = ndim;
    c.n = n;
    c.absorb = absorb;
    c.np = np;
    c.shm = shm;
    c.vp = vp;
    c.dm = dm;
    c.vabc = vabc;
    c.u0 = buf;
    c.v0 = buf + n2;
    c.u1 = buf + 2 * n2;
    c.v1 = buf + 3 * n2;
    c.ve = buf + 4 * n2;
    double *xd = buf + 5 * n2;
    c.xd = xd;
    c.u = u;
    c.v = v;
    c.tab = tab;
    c.tol = tol;
    c.rows = 0;
    c.maxrows = maxrows;
    c.info = info;
    memcpy(c.u0, u, sizeof(double) * n2);
    memcpy(c.v0, v, sizeof(double) * n2);
    memcpy(c.u1, u, sizeof(double) * n2);
    memcpy(c.v1, v, sizeof(double) * n2);
    memcpy(c.ve, vp, sizeof(double) * n2);
    double x = xi;
    for (int i = 0; i < n; ++i) {
        x = x + sh;
        xd[i] = x;
    }
    *info = 0;
    c.f0 = (orc_fcorrel(n, u, u) + orc_fcorrel(n, v, v)) +
           (orc_fcorrel(n, u + n, u + n) + orc_fcorrel(n, v + n, v + n));
    const int field = e0 != 0.0;
    double t = ti;
    if (prtcrl >= 1) s2_row(&c, t, 1);
    s2_step(&c, orc_ef(efc, e0, t0, td, tp, omg, sni, kl, (t + dt) / 2.0), field, cst / 4.0,
            csqc / 4.0);
    s2_step(&c, orc_ef(efc, e0, t0, td, tp, omg, sni, kl, t + dt), field, cst / 2.0, csqc / 2.0);
    int mc = 1;
    t = ti + dt;
    if (mc == lc) mc = 0;
    if (prtcrl == 2) s2_row(&c, t, 2);
    memcpy(c.u1, c.u0, sizeof(double) * n2);
    memcpy(c.v1, c.v0, sizeof(double) * n2);
    /* DO T=TI+2*DT,TF+DT/2,DT: MAX(INT((m2 - m1 + m3)/m3), 0) trips, T += DT */
    const double m1 = ti + 2.0 * dt;
    long long trips = (long long)((tf + dt / 2.0 - m1 + dt) / dt);
    if (trips < 0) trips = 0;
    t = m1;
    for (long long k = 0; k < trips; ++k, t = t + dt) {
        s2_step(&c, orc_ef(efc, e0, t0, td, tp, omg, sni, kl, t), field, cst, csqc);
        if (++mc == lc) mc = 0;
        if (prtcrl == 2 || (prtcrl == 1 && mc == 0)) s2_row(&c, t, 0);
    }
    if (u1out) memcpy(u1out, c.u1, sizeof(double) * n2);
    if (v1out) memcpy(v1out, c.v1, sizeof(double) * n2);
    free(buf);
    return c.rows;
}
