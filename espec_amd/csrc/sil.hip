// Short-iterative Lanczos propagation (PLNZ / PLNZAUX / LNZ / REORT / AU)
// on the device.
//
// The reference's step (sil.f:70-181) runs two independent real Lanczos
// recurrences, one on Re psi and one on Im psi, each with one classical
// Gram-Schmidt pass against all previous vectors (lnz.f:88-119,
// reort.f:54-80), diagonalises the m x m tridiagonal (sil.f:82), projects,
// rotates by exp(-i lambda dt) and recombines.  Here the recurrences of
// every packet run side by side as independent "vectors" of one launch
// (blockIdx.y), so a Krylov iteration is three kernels for all of them:
//   lnz_apply   v_i = w/beta, w = H w/beta - beta v_{i-1}, alpha = <v_i, w>
//   lnz_project w -= alpha v_i, r_j = <v_j, w> for j <= i          (CGS, V^T w)
//   lnz_update  w -= sum_j r_j v_j, beta_i = |w|
// Each kernel finishes its reduction in the last workgroup to arrive (a
// ticket counter; partial sums are added in block order, so results are
// deterministic and no workgroup ever waits for another).  After m
// iterations sil_coeffs diagonalises each tridiagonal (implicit QL, or
// Sturm bisection when ABSTOL > 0, as DSTEVX does) and forms the
// coefficients, and sil_combine rebuilds psi and samples C(t).
// Storage of psi and of the Krylov vectors is complex64/FP32 in the mixed
// path; all arithmetic, alpha, beta and the Gram coefficients are FP64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "api_common.hpp"
#include "sil.hpp"

namespace espec {

// ------------------------------------------------------- tridiagonal

__device__ int tridiag_ql(int m, double *d, double *e, double *z)
{
    for (int i = 0; i < m * m; ++i) z[i] = 0;
    for (int i = 0; i < m; ++i) z[i * m + i] = 1;
    e[m - 1] = 0;
    for (int l = 0; l < m; ++l) {
        int iter = 0, mm;
        do {
            for (mm = l; mm < m - 1; ++mm) {
                double dd = fabs(d[mm]) + fabs(d[mm + 1]);
                if (fabs(e[mm]) <= 2.2204460492503131e-16 * dd) break;
            }
            if (mm != l) {
                if (++iter > 60) return l + 1;
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
    for (int i = 0; i < m - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < m; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            double t = d[i];
            d[i] = d[k];
            d[k] = t;
            for (int r = 0; r < m; ++r) {
                t = z[i * m + r];
                z[i * m + r] = z[k * m + r];
                z[k * m + r] = t;
            }
        }
    }
    return 0;
}

__device__ void tridiag_bounds(int m, const double *d, const double *e, double abstol, double *e2,
                               double *pivmin, double *gl, double *gu, double *atoli)
{
    const double ulp = 2.2204460492503131e-16, safemn = 2.2250738585072014e-308, fudge = 2.1;
    double pm = 1.0;
    for (int j = 0; j < m - 1; ++j) {
        e2[j] = e[j] * e[j];
        if (e2[j] > pm) pm = e2[j];
    }
    pm *= safemn;
    double lo = d[0], hi = d[0];
    for (int j = 0; j < m; ++j) {
        double r = (j > 0 ? fabs(e[j - 1]) : 0) + (j < m - 1 ? fabs(e[j]) : 0);
        lo = fmin(lo, d[j] - r);
        hi = fmax(hi, d[j] + r);
    }
    double tn = fmax(fabs(lo), fabs(hi));
    *gl = lo - fudge * tn * ulp * m - fudge * 2.0 * pm;
    *gu = hi + fudge * tn * ulp * m + fudge * pm;
    *pivmin = pm;
    *atoli = abstol <= 0 ? ulp * tn : abstol;
}

__device__ double tridiag_bisect_one(int m, const double *d, const double *e2, double pivmin,
                                     double gl, double gu, double atoli, int k)
{
    const double reltol = 2.0 * 2.2204460492503131e-16;
    double a = gl, b = gu;
    for (int it = 0; it < 200; ++it) {
        double w = fabs(b - a), t = fmax(fabs(b), fabs(a));
        if (w < fmax(atoli, fmax(pivmin, reltol * t))) break;
        double c = 0.5 * (a + b);
        int cnt = 0;
        double q = d[0] - c;
        if (fabs(q) < pivmin) q = -pivmin;
        if (q <= 0) ++cnt;
        for (int j = 1; j < m; ++j) {
            q = d[j] - e2[j - 1] / q - c;
            if (fabs(q) < pivmin) q = -pivmin;
            if (q <= 0) ++cnt;
        }
        if (cnt > k) b = c;
        else a = c;
    }
    return 0.5 * (a + b);
}

// ------------------------------------------------------------ kernels

struct LnzArgs {
    int ndim, n, np[3], st[3];
    double shm[3];
    const double *vpot;   // [npk][n] (or one shared table when vshared)
    int vshared;
    int nvec, m, nb;      // independent vectors, Krylov order, blocks per vector
    void *basis;          // [nvec][m][n] storage precision
    void *w0, *w1;        // [nvec][n] ping-pong residuals
    double *al, *bt;      // [nvec][m]
    double *beta0;        // [nvec] norm of the start vector
    double *rc;           // [nvec][m] Gram coefficients of the current iteration
    double *part;         // [nvec][nb][kMaxKrylov + 2] partial sums
    unsigned *ticket;     // [nvec]
    int *dead;            // [nvec] start vector was zero (lnz.f:86)
    int it;               // current iteration i
    int nreort;           // vectors in the Gram-Schmidt pass: min(i+1, nreort)
};

// H u at point i (au1d.f/au2di.f/au3d.f): per axis the 4th-order stencil
// shm (u[-2] - 16 u[-1] + 30 u - 16 u[+1] + u[+2]) with neighbours outside
// the grid taken as zero, plus V u.
template <typename S>
__device__ __forceinline__ double apply_h(const LnzArgs &a, const S *u, const double *vp, int i)
{
    double acc = vp[i] * (double)u[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= a.ndim) break;
        const int s = a.st[d], np = a.np[d];
        const int ia = (i / s) % np;
        double t = 30.0 * (double)u[i];
        if (ia >= 2) t += (double)u[i - 2 * s];
        if (ia >= 1) t -= 16.0 * (double)u[i - s];
        if (ia + 1 < np) t -= 16.0 * (double)u[i + s];
        if (ia + 2 < np) t += (double)u[i + 2 * s];
        acc += a.shm[d] * t;
    }
    return acc;
}

// Sum k values per thread over the block into red[0..k) (thread 0 valid).
template <int NT>
__device__ void block_sum_k(double *v, int k, double *red)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = 0; j < k; ++j) {
        double x = wave_sum(v[j]);
        if (lane == 0) red[w * (kMaxKrylov + 2) + j] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < k; ++j) {
            double s = 0;
            for (int q = 0; q < NT / 64; ++q) s += red[q * (kMaxKrylov + 2) + j];
            v[j] = s;
        }
}

// Publish k block partials; returns true in the last block to arrive, in
// which tot[0..k) then holds the deterministic (block-ordered) totals.
__device__ bool last_block(const LnzArgs &a, int vec, const double *v, int k, double *tot)
{
    __shared__ bool last;
    double *p = a.part + ((long long)vec * a.nb + blockIdx.x) * (kMaxKrylov + 2);
    if (threadIdx.x == 0) {
        for (int j = 0; j < k; ++j) p[j] = v[j];
        __threadfence();
        unsigned t = atomicAdd(&a.ticket[vec], 1u);
        last = (t == (unsigned)a.nb - 1);
    }
    __syncthreads();
    if (!last) return false;
    if (threadIdx.x == 0) {
        __threadfence();
        for (int j = 0; j < k; ++j) tot[j] = 0;
        for (int b = 0; b < a.nb; ++b) {
            const double *q = a.part + ((long long)vec * a.nb + b) * (kMaxKrylov + 2);
            for (int j = 0; j < k; ++j) tot[j] += q[j];
        }
        a.ticket[vec] = 0;
    }
    return true;
}

// Start: w = start vector of each recurrence (Re psi for even vec, Im psi
// for odd vec, or an explicit start), beta0 = |w|.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_init(LnzArgs a, const V2<S> *x, const S *start)
{
    __shared__ double red[(NT / 64) * (kMaxKrylov + 2)];
    const int vec = blockIdx.y;
    S *w = reinterpret_cast<S *>(a.w0) + (long long)vec * a.n;
    double acc = 0;
    for (int i = blockIdx.x * NT + threadIdx.x; i < a.n; i += a.nb * NT) {
        S val;
        if (start) {
            val = start[(long long)vec * a.n + i];
        } else {
            V2<S> c = x[(long long)(vec >> 1) * a.n + i];
            val = (vec & 1) ? c.y : c.x;
        }
        w[i] = val;
        acc += (double)val * (double)val;
    }
    double v[1] = {acc};
    block_sum_k<NT>(v, 1, red);
    double tot[1];
    if (last_block(a, vec, v, 1, tot) && threadIdx.x == 0) {
        a.beta0[vec] = sqrt(tot[0]);
        a.dead[vec] = tot[0] == 0.0;
    }
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_apply(LnzArgs a)
{
    __shared__ double red[(NT / 64) * (kMaxKrylov + 2)];
    const int vec = blockIdx.y, i = a.it;
    const long long vo = (long long)vec * a.n;
    const S *win = reinterpret_cast<const S *>((i & 1) ? a.w1 : a.w0) + vo;
    S *wout = reinterpret_cast<S *>((i & 1) ? a.w0 : a.w1) + vo;
    S *B = reinterpret_cast<S *>(a.basis) + (long long)vec * a.m * a.n;
    const double *vp = a.vpot + (a.vshared ? 0 : (long long)(vec >> 1) * a.n);
    // A zero start vector leaves the basis at zero (lnz.f:85-86); the
    // recurrence is then carried as zeros so nothing reads a stale beta.
    const bool dead = a.dead[vec];
    const double b0 = dead ? 1.0 : (i == 0 ? a.beta0[vec] : a.bt[vec * a.m + i - 1]);
    const double rb = 1.0 / b0;
    double acc = 0;
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        double vi = 0, hw = 0;
        if (!dead) {
            vi = (double)win[p] / b0;
            hw = apply_h<S>(a, win, vp, p) * rb;
            if (i > 0) hw -= b0 * (double)B[(long long)(i - 1) * a.n + p];
        }
        B[(long long)i * a.n + p] = (S)vi;
        wout[p] = (S)hw;
        acc += vi * hw;
    }
    double v[1] = {acc};
    block_sum_k<NT>(v, 1, red);
    double tot[1];
    if (last_block(a, vec, v, 1, tot) && threadIdx.x == 0) a.al[vec * a.m + i] = tot[0];
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_project(LnzArgs a)
{
    __shared__ double red[(NT / 64) * (kMaxKrylov + 2)];
    const int vec = blockIdx.y, i = a.it;
    const int nj = min(i + 1, a.nreort);
    const long long vo = (long long)vec * a.n;
    S *w = reinterpret_cast<S *>((i & 1) ? a.w0 : a.w1) + vo;
    const S *B = reinterpret_cast<const S *>(a.basis) + (long long)vec * a.m * a.n;
    const double al = a.al[vec * a.m + i];
    double r[kMaxKrylov];
    for (int j = 0; j < nj; ++j) r[j] = 0;
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        const double wv = (double)w[p] - al * (double)B[(long long)i * a.n + p];
        w[p] = (S)wv;
        const double ws = (double)(S)wv;  // the stored value is what REORT reads
        for (int j = 0; j < nj; ++j) r[j] += ws * (double)B[(long long)j * a.n + p];
    }
    block_sum_k<NT>(r, nj, red);
    double tot[kMaxKrylov];
    if (last_block(a, vec, r, nj, tot) && threadIdx.x == 0)
        for (int j = 0; j < nj; ++j) a.rc[vec * a.m + j] = tot[j];
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_update(LnzArgs a)
{
    __shared__ double red[(NT / 64) * (kMaxKrylov + 2)];
    __shared__ double rs[kMaxKrylov];
    const int vec = blockIdx.y, i = a.it;
    const int nj = min(i + 1, a.nreort);
    const long long vo = (long long)vec * a.n;
    S *w = reinterpret_cast<S *>((i & 1) ? a.w0 : a.w1) + vo;
    const S *B = reinterpret_cast<const S *>(a.basis) + (long long)vec * a.m * a.n;
    if (threadIdx.x < nj) rs[threadIdx.x] = a.rc[vec * a.m + threadIdx.x];
    __syncthreads();
    double acc = 0;
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        double wv = (double)w[p];
        for (int j = 0; j < nj; ++j) wv -= rs[j] * (double)B[(long long)j * a.n + p];
        w[p] = (S)wv;
        const double ws = (double)(S)wv;
        acc += ws * ws;
    }
    double v[1] = {acc};
    block_sum_k<NT>(v, 1, red);
    double tot[1];
    if (last_block(a, vec, v, 1, tot) && threadIdx.x == 0) a.bt[vec * a.m + i] = sqrt(tot[0]);
}

struct SilArgs {
    LnzArgs l;
    const void *x0;       // [npk][n] psi at the start of the step (V2<S>)
    double cst;           // dt / SHFS
    double abstol;
    int ref;              // reproduce sil.f:98-99
    double *coef;         // [nvec][2][m]: cos part, sin part in the Krylov basis
    int *info;            // [1] eigensolve failure flag
};

// Projection c_j = <v_j, start>, then in the last block the tridiagonal
// eigensolve and the coefficients of sil.f:87-117 / 135-166.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void sil_coeffs(SilArgs s)
{
    __shared__ double red[(NT / 64) * (kMaxKrylov + 2)];
    __shared__ double d[kMaxKrylov], e[kMaxKrylov], z[kMaxKrylov * kMaxKrylov], e2[kMaxKrylov];
    __shared__ double bnd[4];
    const LnzArgs &a = s.l;
    const int vec = blockIdx.y, m = a.m;
    const S *B = reinterpret_cast<const S *>(a.basis) + (long long)vec * m * a.n;
    const V2<S> *x0 = reinterpret_cast<const V2<S> *>(s.x0) + (long long)(vec >> 1) * a.n;
    double c[kMaxKrylov];
    for (int j = 0; j < m; ++j) c[j] = 0;
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        V2<S> v = x0[p];
        const double u = (vec & 1) ? (double)v.y : (double)v.x;
        for (int j = 0; j < m; ++j) c[j] += (double)B[(long long)j * a.n + p] * u;
    }
    block_sum_k<NT>(c, m, red);
    double tot[kMaxKrylov];
    if (!last_block(a, vec, c, m, tot)) return;
    double *coef = s.coef + (long long)vec * 2 * m;
    if (a.dead[vec]) {
        if (threadIdx.x < 2 * m) coef[threadIdx.x] = 0;
        return;
    }
    if (threadIdx.x == 0) {
        for (int j = 0; j < m; ++j) {
            d[j] = a.al[vec * m + j];
            e[j] = j < m - 1 ? a.bt[vec * m + j] : 0;
        }
        if (s.abstol > 0) {
            double pm, gl, gu, at;
            tridiag_bounds(m, d, e, s.abstol, e2, &pm, &gl, &gu, &at);
            bnd[0] = pm;
            bnd[1] = gl;
            bnd[2] = gu;
            bnd[3] = at;
        }
    }
    __syncthreads();
    // Bisection eigenvalues (ABSTOL > 0, dstebz.f) one per thread, before
    // QL overwrites d.
    double lam_b = 0;
    if (s.abstol > 0 && threadIdx.x < m)
        lam_b = tridiag_bisect_one(m, d, e2, bnd[0], bnd[1], bnd[2], bnd[3], threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        int bad = tridiag_ql(m, d, e, z);
        if (bad) *s.info = bad;
    }
    __syncthreads();
    if (s.abstol > 0 && threadIdx.x < m) d[threadIdx.x] = lam_b;
    __syncthreads();
    if (threadIdx.x == 0) {
        double u2[kMaxKrylov], pc[kMaxKrylov], ps[kMaxKrylov];
        for (int j = 0; j < m; ++j) {  // DGEMV 'T': u2 = Q^T c
            double t = 0;
            for (int i = 0; i < m; ++i) t += z[j * m + i] * tot[i];
            u2[j] = t;
        }
        for (int j = 0; j < m; ++j) {
            const double ang = d[j] * s.cst;
            const double cs = cos(ang), sn = sin(ang);
            if ((vec & 1) == 0) {  // Re psi: sil.f:95-100
                pc[j] = cs * u2[j];
                ps[j] = -sn * (s.ref ? pc[j] : u2[j]);
            } else {               // Im psi: sil.f:143-148
                ps[j] = sn * u2[j];
                pc[j] = cs * u2[j];
            }
        }
        for (int i = 0; i < m; ++i) {  // DGEMV 'N': coefficients = Q * (...)
            double tc = 0, ts = 0;
            for (int j = 0; j < m; ++j) {
                tc += pc[j] * z[j * m + i];
                ts += ps[j] * z[j * m + i];
            }
            coef[i] = tc;
            coef[m + i] = ts;
        }
    }
}

// psi = B_re (cos part) + B_im (sin part) + i (B_re (sin part) + B_im (cos part))
// (sil.f:171-181), and C(t), |psi|^2 at sample steps.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void sil_combine(SilArgs s, V2<S> *x, const V2<S> *psi0,
                                                  double *cbuf, int maxsamp, int lc,
                                                  long long *dstep)
{
    __shared__ double red[(NT / 64) * (kMaxKrylov + 2)];
    __shared__ double cf[4 * kMaxKrylov];
    const LnzArgs &a = s.l;
    const int pk = blockIdx.y, m = a.m;
    const int vr = 2 * pk, vi = 2 * pk + 1;
    if (threadIdx.x < 2 * m) {
        cf[threadIdx.x] = s.coef[(long long)vr * 2 * m + threadIdx.x];
        cf[2 * m + threadIdx.x] = s.coef[(long long)vi * 2 * m + threadIdx.x];
    }
    __syncthreads();
    const S *Br = reinterpret_cast<const S *>(a.basis) + (long long)vr * m * a.n;
    const S *Bi = reinterpret_cast<const S *>(a.basis) + (long long)vi * m * a.n;
    const long long g = *dstep + 1;  // the step this kernel completes
    const bool samp = (g % lc) == 0;
    double acc[3] = {0, 0, 0};
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        double u = 0, wv = 0, v = 0, wu = 0;  // DGEMV 'N' sums, column by column
        for (int j = 0; j < m; ++j) {
            const double br = (double)Br[(long long)j * a.n + p];
            const double bi = (double)Bi[(long long)j * a.n + p];
            u += cf[j] * br;
            wv += cf[m + j] * br;
            v += cf[2 * m + j] * bi;
            wu += cf[3 * m + j] * bi;
        }
        const double nu = u + wu, nv = wv + v;
        V2<S> o;
        o.x = (S)nu;
        o.y = (S)nv;
        x[(long long)pk * a.n + p] = o;
        if (samp) {
            const V2<S> q = psi0[(long long)pk * a.n + p];
            const double ou = (double)o.x, ov = (double)o.y;
            acc[0] += (double)q.x * ou + (double)q.y * ov;
            acc[1] += (double)q.x * ov - (double)q.y * ou;
            acc[2] += ou * ou + ov * ov;
        }
    }
    // Sample reduction (ticket slot of the packet's first vector) and the
    // step counter, advanced once by the last block of packet 0.
    block_sum_k<NT>(acc, 3, red);
    double tot[3];
    if (last_block(a, vr, acc, 3, tot) && threadIdx.x == 0) {
        const long long nc = g / lc;
        if (samp && nc < maxsamp) {
            double *c = cbuf + ((long long)pk * maxsamp + nc) * 3;
            c[0] = tot[0];
            c[1] = tot[1];
            c[2] = tot[2];
        }
    }
}

__global__ void sil_step_done(long long *dstep) { *dstep += 1; }
__global__ void sil_set_step(long long *dstep, long long v) { *dstep = v; }

// -------------------------------------------------------------- plan

}  // namespace espec

using namespace espec;

struct espec_sil_plan {
    int ndim = 1, np[3] = {1, 1, 1}, n = 0, npk = 0, m = 0, lc = 1, maxsamp = 0;
    bool mixed = false, ref = false;
    double cst = 0, abstol = 0;
    long long steps_done = 0;
    int nb = 0, nt = 256;
    hipStream_t stream = nullptr;
    DevBuf x, psi0, vpot, basis, w0, w1, al, bt, beta0, rc, part, ticket, dead, coef, info, cbuf,
        dstep;
    SilArgs sa{};
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    ~espec_sil_plan()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename S>
hipError_t sil_step(espec_sil_plan &P, hipStream_t s)
{
    LnzArgs &l = P.sa.l;
    const int nv = 2 * P.npk;
    dim3 grid(P.nb, nv);
    hipLaunchKernelGGL((lnz_init<S, 256>), grid, dim3(256), 0, s, l, (const V2<S> *)P.x.p,
                       (const S *)nullptr);
    for (int i = 0; i < P.m; ++i) {
        LnzArgs li = l;
        li.it = i;
        hipLaunchKernelGGL((lnz_apply<S, 256>), grid, dim3(256), 0, s, li);
        hipLaunchKernelGGL((lnz_project<S, 256>), grid, dim3(256), 0, s, li);
        hipLaunchKernelGGL((lnz_update<S, 256>), grid, dim3(256), 0, s, li);
    }
    SilArgs sa = P.sa;
    sa.x0 = P.x.p;
    hipLaunchKernelGGL((sil_coeffs<S, 256>), grid, dim3(256), 0, s, sa);
    hipLaunchKernelGGL((sil_combine<S, 256>), dim3(P.nb, P.npk), dim3(256), 0, s, sa,
                       (V2<S> *)P.x.p, (const V2<S> *)P.psi0.p, P.cbuf.as<double>(), P.maxsamp,
                       P.lc, P.dstep.as<long long>());
    hipLaunchKernelGGL(sil_step_done, dim3(1), dim3(1), 0, s, P.dstep.as<long long>());
    return hipGetLastError();
}

template <typename S>
int sil_run_typed(espec_sil_plan &P, int nsteps)
{
    hipLaunchKernelGGL(sil_set_step, dim3(1), dim3(1), 0, P.stream, P.dstep.as<long long>(),
                       P.steps_done);
    ESPEC_HIP_TRY(hipGetLastError());
    const int gsteps = 8;
    int left = nsteps;
    if (left >= gsteps) {
        if (!P.gexec) {
            hipStream_t cs;
            ESPEC_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            for (int k = 0; e == hipSuccess && k < gsteps; ++k) e = sil_step<S>(P, cs);
            hipError_t e2 = hipStreamEndCapture(cs, &P.graph);
            (void)hipStreamDestroy(cs);
            ESPEC_HIP_TRY(e);
            ESPEC_HIP_TRY(e2);
            ESPEC_HIP_TRY(hipGraphInstantiate(&P.gexec, P.graph, nullptr, nullptr, 0));
        }
        while (left >= gsteps) {
            ESPEC_HIP_TRY(hipGraphLaunch(P.gexec, P.stream));
            left -= gsteps;
        }
    }
    for (; left > 0; --left) ESPEC_HIP_TRY(sil_step<S>(P, P.stream));
    P.steps_done += nsteps;
    return ESPEC_OK;
}

}  // namespace

// <psi0|psi> for sample 0 (from spo_lds.hip).
namespace espec {
hipError_t launch_spo_sample(const void *x, const void *psi0, int n, int npk, bool mixed,
                             double *cbuf, int maxsamp, int idx, hipStream_t s);
}

extern "C" {

typedef struct espec_sil_plan espec_sil_plan;

int espec_sil_plan_create(espec_sil_plan **out, int ndim, const int *np, int npk, int mp,
                          double abstol, double dt, int lc, const double *shm, const double *vpot,
                          const double *u, const double *v, int maxsamp, int flags)
{
    if (!out || ndim < 1 || ndim > 3 || !np || npk < 1 || mp < 2 || mp > kMaxKrylov || lc < 1 ||
        !shm || !vpot || !u || maxsamp < 1)
        return ESPEC_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<espec_sil_plan> P(new espec_sil_plan);
    P->ndim = ndim;
    P->n = 1;
    for (int a = 0; a < ndim; ++a) {
        if (np[a] < 1) return ESPEC_ERR_ARG;
        P->np[a] = np[a];
        P->n *= np[a];
    }
    for (int a = 0; a < ndim; ++a)
        if (np[a] < 4) return ESPEC_ERR_ARG;  // the 5-point stencils need 4 points (au1d.f)
    P->npk = npk;
    P->m = mp;
    P->lc = lc;
    P->maxsamp = maxsamp;
    P->mixed = (flags & ESPEC_MIXED) != 0;
    P->ref = (flags & ESPEC_REF) != 0;
    P->cst = dt / kShfs;
    P->abstol = abstol;
    const int n = P->n, nv = 2 * npk;
    P->nb = std::max(1, std::min(256, (n + 255) / 256));
    ESPEC_HIP_TRY(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    const size_t ss = P->mixed ? sizeof(float) : sizeof(double);
    if (P->mixed) {
        std::vector<float> xs((size_t)2 * n * npk);
        for (size_t i = 0; i < (size_t)n * npk; ++i) {
            xs[2 * i] = (float)u[i];
            xs[2 * i + 1] = v ? (float)v[i] : 0.f;
        }
        ESPEC_HIP_TRY(P->x.upload(xs.data(), sizeof(float) * xs.size()));
        ESPEC_HIP_TRY(P->psi0.upload(xs.data(), sizeof(float) * xs.size()));
    } else {
        std::vector<double> xs((size_t)2 * n * npk);
        for (size_t i = 0; i < (size_t)n * npk; ++i) {
            xs[2 * i] = u[i];
            xs[2 * i + 1] = v ? v[i] : 0.0;
        }
        ESPEC_HIP_TRY(P->x.upload(xs.data(), sizeof(double) * xs.size()));
        ESPEC_HIP_TRY(P->psi0.upload(xs.data(), sizeof(double) * xs.size()));
    }
    ESPEC_HIP_TRY(P->vpot.upload(vpot, sizeof(double) * (size_t)n * npk));
    ESPEC_HIP_TRY(P->basis.alloc(ss * (size_t)nv * mp * n));
    ESPEC_HIP_TRY(P->w0.alloc(ss * (size_t)nv * n));
    ESPEC_HIP_TRY(P->w1.alloc(ss * (size_t)nv * n));
    ESPEC_HIP_TRY(P->al.alloc(sizeof(double) * nv * mp));
    ESPEC_HIP_TRY(P->bt.alloc(sizeof(double) * nv * mp));
    ESPEC_HIP_TRY(P->beta0.alloc(sizeof(double) * nv));
    ESPEC_HIP_TRY(P->rc.alloc(sizeof(double) * nv * mp));
    ESPEC_HIP_TRY(P->part.alloc(sizeof(double) * (size_t)nv * P->nb * (kMaxKrylov + 2)));
    ESPEC_HIP_TRY(P->ticket.alloc(sizeof(unsigned) * nv));
    ESPEC_HIP_TRY(hipMemset(P->ticket.p, 0, P->ticket.bytes));
    ESPEC_HIP_TRY(P->dead.alloc(sizeof(int) * nv));
    ESPEC_HIP_TRY(P->coef.alloc(sizeof(double) * nv * 2 * mp));
    ESPEC_HIP_TRY(P->info.alloc(sizeof(int)));
    ESPEC_HIP_TRY(hipMemset(P->info.p, 0, sizeof(int)));
    ESPEC_HIP_TRY(P->cbuf.alloc(sizeof(double) * 3 * (size_t)maxsamp * npk));
    ESPEC_HIP_TRY(hipMemset(P->cbuf.p, 0, P->cbuf.bytes));
    ESPEC_HIP_TRY(P->dstep.alloc(sizeof(long long)));
    LnzArgs &l = P->sa.l;
    l.ndim = ndim;
    l.n = n;
    int st = 1;
    for (int a = 0; a < 3; ++a) {
        l.np[a] = a < ndim ? np[a] : 1;
        l.st[a] = st;
        st *= l.np[a];
        l.shm[a] = a < ndim ? shm[a] : 0.0;
    }
    l.vpot = P->vpot.as<double>();
    l.vshared = 0;
    l.nvec = nv;
    l.m = mp;
    l.nb = P->nb;
    l.basis = P->basis.p;
    l.w0 = P->w0.p;
    l.w1 = P->w1.p;
    l.al = P->al.as<double>();
    l.bt = P->bt.as<double>();
    l.beta0 = P->beta0.as<double>();
    l.rc = P->rc.as<double>();
    l.part = P->part.as<double>();
    l.ticket = P->ticket.as<unsigned>();
    l.dead = P->dead.as<int>();
    l.nreort = mp;
    P->sa.cst = P->cst;
    P->sa.abstol = abstol;
    P->sa.ref = P->ref ? 1 : 0;
    P->sa.coef = P->coef.as<double>();
    P->sa.info = P->info.as<int>();
    ESPEC_HIP_TRY(launch_spo_sample(P->x.p, P->psi0.p, n, npk, P->mixed, P->cbuf.as<double>(),
                                    maxsamp, 0, P->stream));
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    *out = P.release();
    return ESPEC_OK;
}

int espec_sil_plan_run(espec_sil_plan *P, int nsteps)
{
    if (!P || nsteps < 0) return ESPEC_ERR_ARG;
    if (nsteps == 0) return ESPEC_OK;
    return P->mixed ? sil_run_typed<float>(*P, nsteps) : sil_run_typed<double>(*P, nsteps);
}

int espec_sil_plan_get(espec_sil_plan *P, int pk, double *cr, double *ci, double *nrm, int maxsamp,
                       int *nsamp, double *u, double *v)
{
    if (!P || pk < 0 || pk >= P->npk) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    int info = 0;
    ESPEC_HIP_TRY(hipMemcpy(&info, P->info.p, sizeof(int), hipMemcpyDeviceToHost));
    long long have = P->steps_done / P->lc + 1;
    int ns = (int)std::min<long long>(have, std::min(maxsamp, P->maxsamp));
    if (nsamp) *nsamp = ns;
    if (ns > 0 && (cr || ci || nrm)) {
        std::vector<double> c((size_t)3 * ns);
        ESPEC_HIP_TRY(hipMemcpy(c.data(), P->cbuf.as<double>() + (size_t)3 * P->maxsamp * pk,
                                sizeof(double) * c.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < ns; ++i) {
            if (cr) cr[i] = c[3 * i];
            if (ci) ci[i] = c[3 * i + 1];
            if (nrm) nrm[i] = c[3 * i + 2];
        }
    }
    if (u || v) {
        const size_t n = P->n;
        if (P->mixed) {
            std::vector<float> h(2 * n);
            ESPEC_HIP_TRY(hipMemcpy(h.data(), P->x.as<float>() + 2 * n * pk, sizeof(float) * 2 * n,
                                    hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) {
                if (u) u[i] = h[2 * i];
                if (v) v[i] = h[2 * i + 1];
            }
        } else {
            std::vector<double> h(2 * n);
            ESPEC_HIP_TRY(hipMemcpy(h.data(), P->x.as<double>() + 2 * n * pk,
                                    sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) {
                if (u) u[i] = h[2 * i];
                if (v) v[i] = h[2 * i + 1];
            }
        }
    }
    return info ? ESPEC_ERR_DIAG : ESPEC_OK;
}

int espec_sil_plan_sync(espec_sil_plan *P)
{
    if (!P) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    return ESPEC_OK;
}

int espec_sil_plan_destroy(espec_sil_plan *P)
{
    if (P) {
        (void)hipStreamSynchronize(P->stream);
        delete P;
    }
    return ESPEC_OK;
}

int espec_plnz(int ndim, const int *np, int mp, double abstol, double dt, double ti, double tf,
               int lc, const double *shm, double *u, double *v, const double *vpot, double *cr,
               double *ci, double *nrm, int maxsamp, int *nsamp, int flags)
{
    if (!u || !v || !cr || !ci || maxsamp < 1) return ESPEC_ERR_ARG;
    const int nsteps = fortran_trip(ti + dt, tf + dt, dt);  // plnz.f:109
    espec_sil_plan *P = nullptr;
    int rc = espec_sil_plan_create(&P, ndim, np, 1, mp, abstol, dt, lc, shm, vpot, u, v, maxsamp,
                                   flags);
    if (rc) return rc;
    rc = espec_sil_plan_run(P, nsteps);
    if (!rc) rc = espec_sil_plan_get(P, 0, cr, ci, nrm, maxsamp, nsamp, u, v);
    espec_sil_plan_destroy(P);
    return rc;
}

}  // extern "C"
