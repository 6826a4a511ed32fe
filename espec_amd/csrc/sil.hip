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
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "api_common.hpp"
#include "mlp.hpp"
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

// Register read of lane i's value (i wave-uniform).
__device__ __forceinline__ double lane_get(double v, int i)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffLL), i);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), i);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The QL of tridiag_ql run by one wavefront with the tridiagonal spread
// over the lanes: lane i holds d_i and e_i, so every operand of the
// rotation chain is a register read of one lane (readlane) rather than an
// LDS round trip, and the split point of each sweep comes from one ballot
// instead of a scan.  Lane k applies each rotation to row k of Z, which it
// keeps in LDS (zr[j 64 + k], kMaxKrylov * 64 entries per wavefront): the
// row's loads are issued before the rotation's scalar chain, which hides
// their latency.  The same operations in the same order as tridiag_ql, so
// the results are those of the scalar form bit for bit.  On entry
// d and e (LDS, m entries) hold the matrix; on return d holds the
// eigenvalues in ascending order and z[j m + k] component k of eigenvector
// j.  Returns l + 1 when eigenvalue l has not converged after 60
// iterations (d and z are then written unsorted), else 0.
__device__ int tridiag_ql_wave(int m, double *d, double *e, double *z, double *zr)
{
    constexpr double eps = 2.2204460492503131e-16;
    const int lane = threadIdx.x & 63;
    double dv = lane < m ? d[lane] : 0.0;
    double ev = lane < m - 1 ? e[lane] : 0.0;  // e[m-1] = 0
    for (int j = 0; j < m; ++j) zr[j * 64 + lane] = j == lane ? 1.0 : 0.0;
    int info = 0;
    for (int l = 0; l < m && !info; ++l) {
        int iter = 0;
        for (;;) {
            // mm: the first index >= l whose off-diagonal is negligible
            const double dn = __shfl_down(dv, 1, 64);
            const bool neg = lane >= l && lane < m - 1 && fabs(ev) <= eps * (fabs(dv) + fabs(dn));
            const unsigned long long bal = __ballot(neg);
            const int mm = bal ? __ffsll((long long)bal) - 1 : m - 1;
            if (mm == l) break;
            if (++iter > 60) {
                info = l + 1;
                break;
            }
            const double dl = lane_get(dv, l), el = lane_get(ev, l);
            double g = (lane_get(dv, l + 1) - dl) / (2.0 * el);
            double r = hypot(g, 1.0);
            g = lane_get(dv, mm) - dl + el / (g + (g >= 0 ? fabs(r) : -fabs(r)));
            double s = 1, c = 1, p = 0;
            int i;
            for (i = mm - 1; i >= l; --i) {
                const double zi = zr[i * 64 + lane], zf = zr[(i + 1) * 64 + lane];
                const double ei = lane_get(ev, i);
                const double f = s * ei, bb = c * ei;
                r = hypot(f, g);
                if (lane == i + 1) ev = r;
                if (r == 0) {
                    if (lane == i + 1) dv -= p;
                    if (lane == mm) ev = 0;
                    break;
                }
                s = f / r;
                c = g / r;
                g = lane_get(dv, i + 1) - p;
                r = (lane_get(dv, i) - g) * s + 2.0 * c * bb;
                p = s * r;
                if (lane == i + 1) dv = g + p;
                g = c * r - bb;
                zr[(i + 1) * 64 + lane] = s * zi + c * zf;
                zr[i * 64 + lane] = c * zi - s * zf;
            }
            if (r == 0 && i >= l) continue;
            if (lane == l) {
                dv -= p;
                ev = g;
            }
            if (lane == mm) ev = 0;
        }
    }
    // Ascending order: eigenvalue j (lane j) and eigenvector j go to the
    // rank of d_j (ties, which an unreduced tridiagonal does not have,
    // keep index order).
    int rank = lane;
    if (!info) {
        rank = 0;
        for (int q = 0; q < m; ++q) {
            const double dq = lane_get(dv, q);
            rank += (dq < dv || (dq == dv && q < lane)) ? 1 : 0;
        }
    }
    if (lane < m) d[rank] = dv;
    for (int j = 0; j < m; ++j) {
        const int rj = __builtin_amdgcn_readlane(rank, j);
        if (lane < m) z[rj * m + lane] = zr[j * 64 + lane];
    }
    return info;
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

constexpr int kChunk = 32;            // Gram coefficients summed per block pass
constexpr int kPart = kChunk + 2;     // partial-sum slots per block
constexpr int kMaxReort = 512;        // widest Gram-Schmidt window (LMTREORT = 300, param.h:9)

struct LnzArgs {
    int ndim, n, np[3], st[3];
    double shm[3];
    const double *vpot;   // potential of recurrence v: vpot + (v >> vshift) * n
    int vshift;
    int nvec, m, nb;      // independent recurrences, Krylov order, blocks per recurrence
    int ncols, nchunk;    // stored basis columns (NREORT), Gram chunks of kChunk
    void *basis;          // [nvec][ncols][n] storage precision
    void *vring;          // [2][nvec][n] v_i and v_{i-1}
    void *w0, *w1;        // [nvec][n] ping-pong residuals
    double *al, *bt;      // [nvec][m]
    double *beta0;        // [nvec] norm of the start vector
    double *rc;           // [nvec][ncols] Gram coefficients of the current iteration
    double *part;         // [nvec][nchunk][nb][kPart] partial sums (last-block reductions)
    double *pa, *pb, *pr; // iteration partials: alpha [nvec][nb][kPart], |w|^2 [nvec][nb][kPart],
                          // Gram coefficients [nvec][nchunk][nb][kPart]
    int bfrom;            // lnz_apply takes beta_{i-1} from bt (1) or from pb (0)
    unsigned *ticket;     // [nvec][nchunk]
    int *dead;            // [nvec] start vector was zero (lnz.f:85-86)
    int it;               // current iteration i (0-based)
    int nreort;           // Gram window: columns j < min(i+1, nreort)
    const double *acoef;  // [nacc][m] replay coefficients (LANCZS), or null
    double *acc;          // [nacc][n] replay accumulators
    int nacc;
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

// Sum k <= K values per thread over the block (thread 0 receives the
// sums).  The loops run to the array's compile-time length with the count
// as a predicate, so the array stays in registers (a runtime-indexed array
// would live in scratch memory).
template <int NT, int K>
__device__ __forceinline__ void block_sum_k(double (&v)[K], int k, double *red)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (j < k) {
            const double x = wave_sum(v[j]);
            if (lane == 0) red[w * kPart + j] = x;
        }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < k) {
                double s = 0;
                for (int q = 0; q < NT / 64; ++q) s += red[q * kPart + j];
                v[j] = s;
            }
    }
}

// Publish k block partials of (recurrence vec, chunk c); returns true in
// the last block to arrive, whose thread 0 then holds in tot[0..k) the
// totals.  No block waits.  The last block sums the partials in parallel
// (block b's go to thread b, nb <= NT) and then in the fixed order of
// block_sum_k, so results are deterministic.
template <int NT, int K>
__device__ bool last_block(const LnzArgs &a, int vec, int c, const double (&v)[K], int k,
                           double *tot, double *red)
{
    __shared__ bool last;
    const long long slot = (long long)vec * a.nchunk + c;
    double *p = a.part + (slot * a.nb + blockIdx.x) * kPart;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < k) p[j] = v[j];
        __threadfence();
        unsigned t = atomicAdd(&a.ticket[slot], 1u);
        last = (t == (unsigned)a.nb - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double s[K];
    const double *q = a.part + (slot * a.nb + threadIdx.x) * kPart;
    const bool mine = (int)threadIdx.x < a.nb;
#pragma unroll
    for (int j = 0; j < K; ++j) s[j] = (mine && j < k) ? q[j] : 0.0;
    block_sum_k<NT>(s, k, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < k) tot[j] = s[j];
        a.ticket[slot] = 0;
    }
    return true;
}

// Total over the nb block partials p[b kPart] that the previous kernel
// wrote, in the fixed order of last_block (block b's to thread b, then
// block_sum_k), so every block of the consuming kernel gets the same bits.
// The Lanczos iteration passes its reductions this way: a kernel boundary
// orders the partials, where a last-block reduction needs a device-scope
// fence in every block (an L2 write-back on each XCD).
template <int NT>
__device__ double sum_parts1(const LnzArgs &a, const double *p, double *red)
{
    __shared__ double sh;
    double v[1] = {(int)threadIdx.x < a.nb ? p[threadIdx.x * kPart] : 0.0};
    block_sum_k<NT>(v, 1, red);
    if (threadIdx.x == 0) sh = v[0];
    __syncthreads();
    const double t = sh;
    __syncthreads();  // red and sh are reused by the caller
    return t;
}

// Totals of k values over the nb block partials p[b kPart + j] that the
// previous kernel wrote (put_parts), in last_block's fixed order (block
// b's to thread b, then block_sum_k); thread 0 receives them in tot.
template <int NT, int K>
__device__ __forceinline__ void sum_parts_k(const LnzArgs &a, const double *p, int k, double *tot,
                                            double *red)
{
    double v[K];
    const double *q = p + (long long)threadIdx.x * kPart;
    const bool mine = (int)threadIdx.x < a.nb;
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] = (mine && j < k) ? q[j] : 0.0;
    block_sum_k<NT>(v, k, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < k) tot[j] = v[j];
    }
}

// Publish this block's k partials (thread 0 holds them after block_sum_k).
template <int K>
__device__ __forceinline__ void put_parts(double *p, const double (&v)[K], int k)
{
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < k) p[(long long)blockIdx.x * kPart + j] = v[j];
    }
}

// Start: w = start vector of each recurrence (Re psi for even vec, Im psi
// for odd vec, or an explicit start); |w|^2 partials for lnz_apply, which
// forms beta0 = |w| (lnz.f:75-86).
template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_init(LnzArgs a, const V2<S> *x, const S *start)
{
    __shared__ double red[(NT / 64) * kPart];
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
    put_parts(a.pb + (long long)vec * a.nb * kPart, v, 1);
}

// lnz.f:89-108: v_i = w / beta, w = H v_i - beta v_{i-1}, alpha_i = <v_i, w>.
// One recurrence's vectors: the residual ping-pong w[2], the ring of v_i
// and v_{i-1}, and its basis [ncols][n] (global memory, or LDS in
// lnz_block_lds).
template <typename S> struct VecView {
    S *w0, *w1;      // residual of even / odd iterations' input
    S *r0, *r1;      // v_i of even / odd iterations
    S *basis;
    // selects, not a runtime index: an indexed pointer array would live in
    // scratch memory
    __device__ S *w(int k) const { return k ? w1 : w0; }
    __device__ S *ring(int k) const { return k ? r1 : r0; }
};

template <typename S>
__device__ __forceinline__ VecView<S> global_view(const LnzArgs &a, int vec)
{
    const long long vo = (long long)vec * a.n;
    VecView<S> v;
    v.w0 = reinterpret_cast<S *>(a.w0) + vo;
    v.w1 = reinterpret_cast<S *>(a.w1) + vo;
    v.r0 = reinterpret_cast<S *>(a.vring) + vo;
    v.r1 = reinterpret_cast<S *>(a.vring) + ((long long)a.nvec + vec) * a.n;
    v.basis = reinterpret_cast<S *>(a.basis) + (long long)vec * a.ncols * a.n;
    return v;
}

// Iteration i's three steps are device functions (i, the chunk c of the
// projection and the recurrence's vectors as arguments) so that lnz_block
// can run a whole run of iterations in one launch; the kernels below wrap
// them.
template <typename S, int NT>
__device__ __forceinline__ void apply_it(const LnzArgs &a, const VecView<S> &V, int i, int bfrom,
                                         double *red)
{
    const int vec = blockIdx.y;
    const S *win = V.w(i & 1);
    S *wout = V.w((i & 1) ^ 1);
    S *vcur = V.ring(i & 1);
    const S *vprev = V.ring((i + 1) & 1);
    S *B = V.basis;
    const double *vp = a.vpot + ((long long)(vec >> a.vshift)) * a.n;
    // beta_{i-1} (beta0 at i = 0) from the partials of the kernel before,
    // or from bt when this launch extends an earlier run.  A zero start
    // vector leaves the basis at zero (lnz.f:85-86); the recurrence is then
    // carried as zeros so nothing reads a stale beta.
    bool dead;
    double bprev;
    if (i > 0 && bfrom) {
        dead = a.dead[vec];
        bprev = a.bt[vec * a.m + i - 1];
    } else {
        const double t = sum_parts1<NT>(a, a.pb + (long long)vec * a.nb * kPart, red);
        bprev = sqrt(t);
        if (i == 0) {
            dead = t == 0.0;
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                a.beta0[vec] = bprev;
                a.dead[vec] = dead;
            }
        } else {
            dead = a.dead[vec];
            if (blockIdx.x == 0 && threadIdx.x == 0) a.bt[vec * a.m + i - 1] = bprev;
        }
    }
    const double b0 = dead ? 1.0 : bprev;
    const double rb = 1.0 / b0;
    const bool store = i < a.ncols;
    double acc = 0;
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        double vi = 0, hw = 0;
        if (!dead) {
            vi = (double)win[p] / b0;
            hw = apply_h<S>(a, win, vp, p) * rb;
            if (i > 0) hw -= b0 * (double)vprev[p];
        }
        vcur[p] = (S)vi;
        if (store) B[(long long)i * a.n + p] = (S)vi;
        wout[p] = (S)hw;
        acc += vi * hw;
        for (int k = 0; k < a.nacc; ++k) a.acc[(long long)k * a.n + p] += a.acoef[k * a.m + i] * vi;
    }
    double v[1] = {acc};
    block_sum_k<NT>(v, 1, red);
    put_parts(a.pa + (long long)vec * a.nb * kPart, v, 1);
}

// lnz.f:110-112 and reort.f:54-67: r_j = <v_j, w - alpha v_i> for the
// columns of the Gram window, kChunk columns per grid.z slice (the residual
// itself is updated by lnz_update, so every slice reads the same w).
template <typename S, int NT>
__device__ __forceinline__ void project_it(const LnzArgs &a, const VecView<S> &V, int i, int c,
                                           double *red)
{
    const int vec = blockIdx.y;
    const int nj = min(i + 1, a.nreort);
    const int j0 = c * kChunk;
    if (j0 >= nj) return;
    const int jn = min(kChunk, nj - j0);
    const S *w = V.w((i & 1) ^ 1);
    const S *vcur = V.ring(i & 1);
    const S *B = V.basis + (long long)j0 * a.n;
    const double al = sum_parts1<NT>(a, a.pa + (long long)vec * a.nb * kPart, red);
    if (c == 0 && blockIdx.x == 0 && threadIdx.x == 0) a.al[vec * a.m + i] = al;
    double r[kChunk];
    for (int j = 0; j < kChunk; ++j) r[j] = 0;
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        const double ws = (double)(S)((double)w[p] - al * (double)vcur[p]);
#pragma unroll
        for (int j = 0; j < kChunk; ++j)
            if (j < jn) r[j] += ws * (double)B[(long long)j * a.n + p];
    }
    block_sum_k<NT>(r, jn, red);
    put_parts(a.pr + ((long long)vec * a.nchunk + c) * a.nb * kPart, r, jn);
}

// w <- (w - alpha v_i) - sum_j r_j v_j (reort.f:69-78), beta_i = |w|.
template <typename S, int NT>
__device__ __forceinline__ void update_it(const LnzArgs &a, const VecView<S> &V, int i, double *red,
                                          double *rs)
{
    const int vec = blockIdx.y;
    const int nj = min(i + 1, a.nreort);
    S *w = V.w((i & 1) ^ 1);
    const S *vcur = V.ring(i & 1);
    const S *B = V.basis;
    const double al = a.al[vec * a.m + i];
    // Gram coefficient j: wavefront j mod (NT/64) sums the nb partials of
    // lnz_project, lanes in a fixed order, then a wavefront shuffle tree.
    // A wavefront takes eight of its columns at a time, so that their
    // loads and shuffle trees overlap.
    {
        constexpr int NW = NT / 64, G = 8;
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int j0 = wv; j0 < nj; j0 += G * NW) {
            double x[G];
#pragma unroll
            for (int q = 0; q < G; ++q) x[q] = 0;
            for (int b = lane; b < a.nb; b += 64) {
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    const int j = j0 + q * NW;
                    if (j < nj)
                        x[q] += a.pr[((long long)vec * a.nchunk + j / kChunk) * a.nb * kPart +
                                     (long long)b * kPart + j % kChunk];
                }
            }
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const int j = j0 + q * NW;
                if (j < nj) {
                    const double t = wave_sum(x[q]);
                    if (lane == 0) rs[j] = t;
                }
            }
        }
    }
    __syncthreads();
    double acc = 0;
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        double wv = (double)(S)((double)w[p] - al * (double)vcur[p]);
        for (int j = 0; j < nj; ++j) wv -= rs[j] * (double)B[(long long)j * a.n + p];
        w[p] = (S)wv;
        const double ws = (double)(S)wv;
        acc += ws * ws;
    }
    double v[1] = {acc};
    block_sum_k<NT>(v, 1, red);
    put_parts(a.pb + (long long)vec * a.nb * kPart, v, 1);
}

// beta of the last iteration of a run (no lnz_apply follows to form it).
template <int NT>
__device__ __forceinline__ void finish_it(const LnzArgs &a, int i, double *red)
{
    const int vec = blockIdx.y;
    const double t = sum_parts1<NT>(a, a.pb + (long long)vec * a.nb * kPart, red);
    if (threadIdx.x == 0) a.bt[vec * a.m + i] = sqrt(t);
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_apply(LnzArgs a)
{
    __shared__ double red[(NT / 64) * kPart];
    apply_it<S, NT>(a, global_view<S>(a, blockIdx.y), a.it, a.bfrom, red);
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_project(LnzArgs a)
{
    __shared__ double red[(NT / 64) * kPart];
    project_it<S, NT>(a, global_view<S>(a, blockIdx.y), a.it, blockIdx.z, red);
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_update(LnzArgs a)
{
    __shared__ double red[(NT / 64) * kPart];
    __shared__ double rs[kMaxReort];
    update_it<S, NT>(a, global_view<S>(a, blockIdx.y), a.it, red, rs);
}

template <int NT>
__global__ __launch_bounds__(NT) void lnz_finish(LnzArgs a)
{
    __shared__ double red[(NT / 64) * kPart];
    finish_it<NT>(a, a.it, red);
}

// Iterations i0 .. i1-1 of every recurrence in one launch, one workgroup
// per recurrence, for grids of one block (nb = 1, n <= 256): the same
// steps, sums and order as the kernels above, with a workgroup barrier
// where those have a kernel boundary (the partials of a one-block grid are
// read back by the thread that wrote them).  bfrom as in lnz_apply for
// iteration i0; with fin, also beta of iteration i1-1 (lnz_finish).
template <typename S, int NT>
__device__ __forceinline__ void block_iterations(const LnzArgs &a, const VecView<S> &V, int i0,
                                                 int i1, double *red, double *rs)
{
    for (int i = i0; i < i1; ++i) {
        apply_it<S, NT>(a, V, i, i == i0 ? a.bfrom : 0, red);
        __syncthreads();
        const int nj = min(i + 1, a.nreort);
        for (int c = 0; c * kChunk < nj; ++c) {
            project_it<S, NT>(a, V, i, c, red);
            __syncthreads();  // thread 0 may still read red in the last sum
        }
        update_it<S, NT>(a, V, i, red, rs);
        __syncthreads();
    }
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_block(LnzArgs a, int i0, int i1, int fin)
{
    __shared__ double red[(NT / 64) * kPart];
    __shared__ double rs[kMaxReort];
    block_iterations<S, NT>(a, global_view<S>(a, blockIdx.y), i0, i1, red, rs);
    if (fin) finish_it<NT>(a, i1 - 1, red);
}

// LDS bytes of lnz_block_lds for grid size n and ncols basis columns.
size_t lnz_lds_bytes(int n, int ncols, size_t elem) { return (size_t)(4 + ncols) * n * elem; }

// lnz_block with the recurrence's vectors (residuals, v ring, basis) held
// in LDS for the whole run: they are loaded once, every step reads and
// writes them there instead of making a round trip to L2 per phase, and
// they are stored back at the end, so that the kernels that follow (and a
// later extension of the run) find them in global memory as lnz_block
// leaves them.  Same arithmetic and order as lnz_block, bit for bit.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void lnz_block_lds(LnzArgs a, int i0, int i1, int fin)
{
    __shared__ double red[(NT / 64) * kPart];
    __shared__ double rs[kMaxReort];
    extern __shared__ double smem_l[];
    const int n = a.n;
    const VecView<S> G = global_view<S>(a, blockIdx.y);
    S *base = reinterpret_cast<S *>(smem_l);
    VecView<S> L;
    L.w0 = base;
    L.w1 = base + n;
    L.r0 = base + 2 * n;
    L.r1 = base + 3 * n;
    L.basis = base + 4 * n;
    const long long nb = (long long)a.ncols * n;
    for (int p = threadIdx.x; p < n; p += NT) {
        L.w0[p] = G.w0[p];
        L.w1[p] = G.w1[p];
        L.r0[p] = G.r0[p];
        L.r1[p] = G.r1[p];
    }
    for (long long p = threadIdx.x; p < nb; p += NT) L.basis[p] = G.basis[p];
    __syncthreads();
    block_iterations<S, NT>(a, L, i0, i1, red, rs);
    for (int p = threadIdx.x; p < n; p += NT) {
        G.w0[p] = L.w0[p];
        G.w1[p] = L.w1[p];
        G.r0[p] = L.r0[p];
        G.r1[p] = L.r1[p];
    }
    for (long long p = threadIdx.x; p < nb; p += NT) G.basis[p] = L.basis[p];
    if (fin) finish_it<NT>(a, i1 - 1, red);
}

struct SilArgs {
    LnzArgs l;
    const void *x0;       // [npk][n] psi at the start of the step (V2<S>)
    double cst;           // dt / SHFS
    double abstol;
    int ref;              // reproduce sil.f:98-99
    double *coef;         // [nvec][2][m]: cos part, sin part in the Krylov basis
    int *info;            // [1] eigensolve failure flag
    const double *mask;   // [npk][n] .SMOOTHW mask applied to the new state, or null
};

// Block partials of the projection c_j = <v_j, start> (DGEMV 'T',
// sil.f:86-87) of the m basis vectors of start vector `start` (Re psi for
// even recurrences, Im psi for odd ones), into part slot (vec, 0).
template <typename S, int NT>
__device__ __forceinline__ void project_start(const LnzArgs &a, const void *start, int m,
                                              double *red)
{
    const int vec = blockIdx.y;
    const S *B = reinterpret_cast<const S *>(a.basis) + (long long)vec * a.ncols * a.n;
    const V2<S> *x0 = reinterpret_cast<const V2<S> *>(start) + (long long)(vec >> 1) * a.n;
    double c[kMaxKrylov];
    for (int j = 0; j < kMaxKrylov; ++j) c[j] = 0;
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        V2<S> v = x0[p];
        const double u = (vec & 1) ? (double)v.y : (double)v.x;
#pragma unroll
        for (int j = 0; j < kMaxKrylov; ++j)
            if (j < m) c[j] += (double)B[(long long)j * a.n + p] * u;
    }
    block_sum_k<NT>(c, m, red);
    put_parts(a.part + (long long)vec * a.nchunk * a.nb * kPart, c, m);
}

// The projection's block partials (sil_coeffs sums them: the kernel
// boundary orders them, so no block needs a device-scope fence).
template <typename S, int NT>
__global__ __launch_bounds__(NT) void sil_project(SilArgs s)
{
    __shared__ double red[(NT / 64) * kPart];
    project_start<S, NT>(s.l, s.x0, s.l.m, red);
}

// One workgroup per recurrence: the projection's total, the tridiagonal
// eigensolve (sil.f:82) and the coefficients of sil.f:88-117 / 136-166.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void sil_coeffs(SilArgs s)
{
    __shared__ double red[(NT / 64) * kPart];
    __shared__ double d[kMaxKrylov], e[kMaxKrylov], z[kMaxKrylov * kMaxKrylov], e2[kMaxKrylov];
    __shared__ double zr[kMaxKrylov * 64];
    __shared__ double bnd[4];
    const LnzArgs &a = s.l;
    const int vec = blockIdx.y, m = a.m;
    double tot[kMaxKrylov];
    sum_parts_k<NT, kMaxKrylov>(a, a.part + (long long)vec * a.nchunk * a.nb * kPart, m, tot, red);
    double *coef = s.coef + (long long)vec * 2 * m;
    if (a.dead[vec]) {
        if (threadIdx.x < 2 * m) coef[threadIdx.x] = 0;
        return;
    }
    __shared__ double tq[kMaxKrylov], pcs[kMaxKrylov], pss[kMaxKrylov];
    if (threadIdx.x == 0) {
        for (int j = 0; j < m; ++j) {
            d[j] = a.al[vec * m + j];
            e[j] = j < m - 1 ? a.bt[vec * m + j] : 0;
        }
#pragma unroll
        for (int j = 0; j < kMaxKrylov; ++j)  // constant indices keep tot in registers
            if (j < m) tq[j] = tot[j];
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
    if (threadIdx.x < 64) {
        const int bad = tridiag_ql_wave(m, d, e, z, zr);
        if (bad && threadIdx.x == 0) *s.info = bad;
    }
    __syncthreads();
    if (s.abstol > 0 && threadIdx.x < m) d[threadIdx.x] = lam_b;
    __syncthreads();
    // Thread j takes eigenvalue j, then thread i row i of the products;
    // every sum runs in ascending order, as the serial DGEMVs do.
    const int k = threadIdx.x;
    if (k < m) {
        double u2 = 0;  // DGEMV 'T': (Q^T c)_k
        for (int i = 0; i < m; ++i) u2 += z[k * m + i] * tq[i];
        const double ang = d[k] * s.cst;
        const double cs = cos(ang), sn = sin(ang);
        if ((vec & 1) == 0) {  // Re psi: sil.f:95-100
            pcs[k] = cs * u2;
            pss[k] = -sn * (s.ref ? cs * u2 : u2);
        } else {               // Im psi: sil.f:143-148
            pss[k] = sn * u2;
            pcs[k] = cs * u2;
        }
    }
    __syncthreads();
    if (k < m) {  // DGEMV 'N': coefficients = Q * (...)
        double tc = 0, ts = 0;
        for (int j = 0; j < m; ++j) {
            tc += pcs[j] * z[j * m + k];
            ts += pss[j] * z[j * m + k];
        }
        coef[k] = tc;
        coef[m + k] = ts;
    }
}

// psi = B_re (cos part) + B_im (sin part) + i (B_re (sin part) + B_im (cos part))
// (sil.f:171-181), and C(t), |psi|^2 at sample steps.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void sil_combine(SilArgs s, V2<S> *x, const V2<S> *psi0,
                                                  double *cbuf, int maxsamp, int lc,
                                                  const long long *dstep)
{
    __shared__ double red[(NT / 64) * kPart];
    __shared__ double cf[4 * kMaxKrylov];
    const LnzArgs &a = s.l;
    const int pk = blockIdx.y, m = a.m;
    const int vr = 2 * pk, vi = 2 * pk + 1;
    if (threadIdx.x < 2 * m) {
        cf[threadIdx.x] = s.coef[(long long)vr * 2 * m + threadIdx.x];
        cf[2 * m + threadIdx.x] = s.coef[(long long)vi * 2 * m + threadIdx.x];
    }
    __syncthreads();
    const S *Br = reinterpret_cast<const S *>(a.basis) + (long long)vr * a.ncols * a.n;
    const S *Bi = reinterpret_cast<const S *>(a.basis) + (long long)vi * a.ncols * a.n;
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
        double nu = u + wu, nv = wv + v;
        if (s.mask) {  // sil.f:171-175: U = VABC (U + WU), V = VABC (WV + V)
            const double mk = s.mask[(long long)pk * a.n + p];
            nu = mk * nu;
            nv = mk * nv;
        }
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
    block_sum_k<NT>(acc, 3, red);
    double tot[3];
    if (last_block<NT>(a, vr, 0, acc, 3, tot, red) && threadIdx.x == 0) {
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

// hu = H u (AU) for one vector.
__global__ __launch_bounds__(256) void au_kernel(LnzArgs a, const double *u, double *hu)
{
    for (int p = blockIdx.x * 256 + threadIdx.x; p < a.n; p += gridDim.x * 256)
        hu[p] = apply_h<double>(a, u, a.vpot, p);
}

// FCORREL: sum u1 u2 over n, block-ordered two-level sum.
__global__ __launch_bounds__(256) void dot_kernel(LnzArgs a, const double *u1, const double *u2,
                                                  double *out)
{
    __shared__ double red[4 * kPart];
    double acc = 0;
    for (int p = blockIdx.x * 256 + threadIdx.x; p < a.n; p += a.nb * 256) acc += u1[p] * u2[p];
    double v[1] = {acc};
    block_sum_k<256>(v, 1, red);
    double tot[1];
    if (last_block<256>(a, 0, 0, v, 1, tot, red) && threadIdx.x == 0) *out = tot[0];
}

// Scratch of espec_fcorrel_noalloc, which allocates nothing: the block
// partials of dot_kernel (at most kScrBlocks blocks of kPart slots) and
// its ticket counter, in one array per device for the life of the process.
// Host code reaches it only through the address hipGetSymbolAddress gives
// for the current device (scratch_address); the array's own name is a host
// stand-in, not a device address.
constexpr int kScrBlocks = 256;  // blocks_for() never exceeds this
constexpr size_t kScrTicket = sizeof(double) * kScrBlocks * kPart;
constexpr size_t kScrBytes = kScrTicket + 256;
static __device__ __align__(256) unsigned char scratch[kScrBytes];

__global__ void scratch_clear_ticket(unsigned *ticket) { *ticket = 0u; }
__global__ void scratch_zero_out(double *out) { *out = 0.0; }

// LANCZSG's Ritz vector (lanczsc.f:100-106): out_J = sum_I z_I v_I(J) over
// the m stored Lanczos vectors, in the reference's order.
__global__ __launch_bounds__(256) void ritz_vector(int n, int m, const double *basis,
                                                   const double *z, double *out)
{
#pragma clang fp contract(off)
    for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s = s + z[i] * basis[(long long)i * n + j];
        out[j] = s;
    }
}

// sum |a - s b| with s = sign <a, b> in one workgroup (fixed order):
// LANCZSG's convergence measure with the sign of b matched to a.
__global__ __launch_bounds__(1024) void diff_sum(int n, const double *a, const double *b,
                                                 double *out)
{
    __shared__ double red[3][16];
    double d0 = 0, d1 = 0, ab = 0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        d0 += fabs(a[i] - b[i]);
        d1 += fabs(a[i] + b[i]);
        ab += a[i] * b[i];
    }
    d0 = wave_sum(d0);
    d1 = wave_sum(d1);
    ab = wave_sum(ab);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = d0;
        red[1][w] = d1;
        red[2][w] = ab;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, t = 0;
        for (int k = 0; k < 16; ++k) {
            s0 += red[0][k];
            s1 += red[1][k];
            t += red[2][k];
        }
        out[0] = t >= 0 ? s0 : s1;
        out[1] = t;
    }
}

// a <- a / |a| (ECNORM, ecnorm.f) in one workgroup.
__global__ __launch_bounds__(1024) void normalise(int n, double *a)
{
    __shared__ double red[16];
    __shared__ double nrm;
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 1024) s += a[i] * a[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < 16; ++k) t += red[k];
        nrm = sqrt(t);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 1024) a[i] = a[i] / nrm;
}

// DSTEVX('V','A') for one tridiagonal: QL vectors, QL or bisection values.
__global__ __launch_bounds__(64) void dstevx_kernel(int m, double *d, double *e, double *e2,
                                                    double abstol, double *w, double *z,
                                                    int *info)
{
    __shared__ double bnd[4];
    if (threadIdx.x == 0) {
        if (abstol > 0) {
            double pm, gl, gu, at;
            tridiag_bounds(m, d, e, abstol, e2, &pm, &gl, &gu, &at);
            bnd[0] = pm;
            bnd[1] = gl;
            bnd[2] = gu;
            bnd[3] = at;
        }
    }
    __syncthreads();
    if (abstol > 0)
        for (int k = threadIdx.x; k < m; k += 64)
            w[k] = tridiag_bisect_one(m, d, e2, bnd[0], bnd[1], bnd[2], bnd[3], k);
    __syncthreads();
    if (threadIdx.x == 0) {
        *info = tridiag_ql(m, d, e, z);
        if (abstol <= 0)
            for (int k = 0; k < m; ++k) w[k] = d[k];
    }
}

// LANCZS eigenvectors of the n x n Lanczos tridiagonal for the selected
// indices: eigenvalue k by bisection (to ABSTOL when > 0, as dstebz.f in
// DSTEVX('V','I'), lanczs.f:90; else to rounding), vector by inverse iteration with a
// partially pivoted LU of T - lambda I (the method of dstein.f), scaled to
// unit norm with its largest component positive (dstein.f), written as the
// replay coefficients coef[k][J].  One thread per selected eigenvalue.
__global__ __launch_bounds__(64) void lanczs_select(int n, const double *d, const double *e,
                                                    double *e2, int il, int ns, double abstol,
                                                    double *eigvl, double *coef, double *work,
                                                    int *info)
{
    __shared__ double bnd[4];
    if (threadIdx.x == 0) {
        double pm, gl, gu, at;
        tridiag_bounds(n, d, e, abstol, e2, &pm, &gl, &gu, &at);
        bnd[0] = pm;
        bnd[1] = gl;
        bnd[2] = gu;
        bnd[3] = at;
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k >= ns) return;
    const double lam = tridiag_bisect_one(n, d, e2, bnd[0], bnd[1], bnd[2], bnd[3], il - 1 + k);
    eigvl[k] = lam;
    double *u0 = work + (size_t)k * 6 * n, *u1 = u0 + n, *u2 = u1 + n, *lm = u2 + n, *x = lm + n,
           *pv = x + n;
    const double eps = 2.2204460492503131e-16;
    double tnorm = 0;
    for (int j = 0; j < n; ++j)
        tnorm = fmax(tnorm, fabs(d[j]) + (j > 0 ? fabs(e[j - 1]) : 0) + (j < n - 1 ? fabs(e[j]) : 0));
    const double tiny = eps * fmax(tnorm, 1e-300);
    // LU of T - lam I with partial pivoting: row k of U is (u0, u1, u2) at
    // columns k, k+1, k+2.
    double a0 = d[0] - lam, a1 = n > 1 ? e[0] : 0, a2 = 0;
    for (int j = 0; j < n - 1; ++j) {
        const double b0 = e[j], b1 = d[j + 1] - lam, b2 = j + 1 < n - 1 ? e[j + 1] : 0;
        if (fabs(a0) >= fabs(b0)) {
            const double l = a0 != 0 ? b0 / a0 : 0;
            u0[j] = a0 != 0 ? a0 : tiny;
            u1[j] = a1;
            u2[j] = a2;
            lm[j] = l;
            pv[j] = 0;
            a0 = b1 - l * a1;
            a1 = b2 - l * a2;
        } else {
            const double l = a0 / b0;
            u0[j] = b0;
            u1[j] = b1;
            u2[j] = b2;
            lm[j] = l;
            pv[j] = 1;
            const double na0 = a1 - l * b1, na1 = a2 - l * b2;
            a0 = na0;
            a1 = na1;
        }
        a2 = 0;
    }
    u0[n - 1] = a0 != 0 ? a0 : tiny;
    for (int j = 0; j < n; ++j) x[j] = 1.0 + 0.5 * sin(1.2345 * j + 0.5 * k);
    for (int itr = 0; itr < 5; ++itr) {
        for (int j = 0; j < n - 1; ++j) {  // apply P and L^-1
            if (pv[j] != 0) {
                const double t = x[j];
                x[j] = x[j + 1];
                x[j + 1] = t;
            }
            x[j + 1] -= lm[j] * x[j];
        }
        for (int j = n - 1; j >= 0; --j) {  // U^-1
            double t = x[j];
            if (j + 1 < n) t -= u1[j] * x[j + 1];
            if (j + 2 < n) t -= u2[j] * x[j + 2];
            x[j] = t / u0[j];
        }
        double nr = 0;
        for (int j = 0; j < n; ++j) nr += x[j] * x[j];
        nr = sqrt(nr);
        if (!(nr > 0) || !isfinite(nr)) {
            atomicExch(info, k + 1);
            return;
        }
        for (int j = 0; j < n; ++j) x[j] /= nr;
    }
    int jm = 0;
    for (int j = 1; j < n; ++j)
        if (fabs(x[j]) > fabs(x[jm])) jm = j;
    const double sg = x[jm] < 0 ? -1.0 : 1.0;
    for (int j = 0; j < n; ++j) coef[(size_t)k * n + j] = sg * x[j];
}

// ------------------------------------------------ adaptive Krylov steps
// Steps whose Krylov order m and time step tau are chosen per step by the
// controller network (mlp.hip), judged by the a-posteriori estimate of the
// Krylov exponential's error
//   err(m, tau) = beta0 beta_m |[exp(-i tau T_m)]_{m,1}|
// (beta0 = |start|, beta_m = |w| after iteration m), with the fixed-order
// step as fallback.  Exact arithmetic (no sil.f:98-99 coefficient reuse).
// State updates go to a second buffer and are adopted only when the step
// is accepted.  No reference counterpart.

constexpr int kLabM0 = 4, kLabM = 17;  // label orders 4..20
// The step runs one order beyond the network's: its error estimate judges
// the network's order against the next one (the inputs come from the
// previous step, and a rejected step costs its fallback).
constexpr int kMarginM = 1;
constexpr int kLabT = 3;               // label steps tau/2, tau, 2 tau

struct CtlArgs {
    LnzArgs l;          // l.m: stored Krylov columns (stride of al, bt)
    const void *x0;     // [npk][n] state at the start of the step
    int mact;           // Krylov order of this step (5 <= mact <= l.m)
    int nsub;           // coefficient sets: tau (1), or tau and 2 tau (2)
    double tau;         // time step of set 0, a.u.
    double tfeat;       // base time step for the controller inputs, a.u.
    double *coef;       // [nvec][2 sets][2][l.m]: cos part, sin part
    double *est;        // [nvec] error estimate of the step
    double *est4;       // [nvec] err(4, tfeat)
    int *info;
};

// err(m, tau) from the eigenpairs (lam, z column-major, ld m) of T_m.
__device__ double krylov_est(int m, const double *lam, const double *z, double bm, double b0,
                             double tau)
{
    double sr = 0, si = 0;
    for (int j = 0; j < m; ++j) {
        const double w = z[j * m + m - 1] * z[j * m];
        sr += w * cos(lam[j] * tau);
        si -= w * sin(lam[j] * tau);
    }
    return b0 * bm * sqrt(sr * sr + si * si);
}

// As leading_ql, by one whole wavefront on LDS arrays (zr: the
// wavefront's rows of Z, kMaxKrylov * 64 entries).
__device__ int leading_ql_wave(const LnzArgs &a, int vec, int m, double *d, double *e, double *z,
                               double *zr)
{
    for (int j = 0; j < m; ++j) {  // every lane stores the same values
        d[j] = a.al[vec * a.m + j];
        e[j] = j < m - 1 ? a.bt[vec * a.m + j] : 0;
    }
    return tridiag_ql_wave(m, d, e, z, zr);
}

// Leading m x m tridiagonal of recurrence vec into (d, e); QL in place.
__device__ int leading_ql(const LnzArgs &a, int vec, int m, double *d, double *e, double *z)
{
    for (int j = 0; j < m; ++j) {
        d[j] = a.al[vec * a.m + j];
        e[j] = j < m - 1 ? a.bt[vec * a.m + j] : 0;
    }
    return tridiag_ql(m, d, e, z);
}

// exp(-i t T) e1 in the Krylov basis from the eigenpairs (lam, z column-
// major, ld m) of T: y[i] = sum_j z[j][i] z[j][0] exp(-i lam_j t).
__device__ void krylov_e1(int m, const double *lam, const double *z, double t, double *yr,
                          double *yi)
{
    for (int i = 0; i < m; ++i) {
        double sr = 0, si = 0;
        for (int j = 0; j < m; ++j) {
            const double w = z[j * m + i] * z[j * m];
            sr += w * cos(lam[j] * t);
            si -= w * sin(lam[j] * t);
        }
        yr[i] = sr;
        yi[i] = si;
    }
}

// beta0 |y_m - y_M| for orders m < M (y_m padded with zeros): the error of
// the order-m approximation measured against the order-M one.
__device__ double krylov_diff(int m, const double *lm, const double *zm, int M, const double *lM,
                              const double *zM, double b0, double t)
{
    double ar[kMaxKrylov], ai[kMaxKrylov], br[kMaxKrylov], bi[kMaxKrylov];
    krylov_e1(m, lm, zm, t, ar, ai);
    krylov_e1(M, lM, zM, t, br, bi);
    double s = 0;
    for (int i = 0; i < M; ++i) {
        const double dr = br[i] - (i < m ? ar[i] : 0.0), di = bi[i] - (i < m ? ai[i] : 0.0);
        s += dr * dr + di * di;
    }
    return b0 * sqrt(s);
}

// Block partials of the projection onto the first mact basis vectors
// (as sil_project); ctl_coeffs sums them.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void ctl_project(CtlArgs c)
{
    __shared__ double red[(NT / 64) * kPart];
    project_start<S, NT>(c.l, c.x0, c.mact, red);
}

// One workgroup per recurrence: the projection's total, eigensolves of
// T_mact and T_{mact-1} (two wavefronts), the coefficient sets of order
// mact, and the step's error estimate: the larger of beta0 |y_{mact-1} -
// y_mact| at the step's end and the residual estimate of order mact.  Also
// err(4, tfeat), a controller input.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void ctl_coeffs(CtlArgs c)
{
    __shared__ double red[(NT / 64) * kPart];
    __shared__ double d[2][kMaxKrylov], ee[2][kMaxKrylov], z[2][kMaxKrylov * kMaxKrylov];
    __shared__ double zr[2][kMaxKrylov * 64];
    __shared__ int bad;
    const LnzArgs &a = c.l;
    const int vec = blockIdx.y, m = c.mact, ms = a.m;
    double tot[kMaxKrylov];
    sum_parts_k<NT, kMaxKrylov>(a, a.part + (long long)vec * a.nchunk * a.nb * kPart, m, tot, red);
    double *coef = c.coef + (long long)vec * 4 * ms;
    if (a.dead[vec]) {
        for (int i = threadIdx.x; i < 4 * ms; i += NT) coef[i] = 0;
        if (threadIdx.x == 0) {
            c.est[vec] = 0;
            c.est4[vec] = 0;
        }
        return;
    }
    const double b0 = a.beta0[vec];
    __shared__ double tq_[kMaxKrylov], u2[kMaxKrylov];
    __shared__ double pc[2][kMaxKrylov], ps[2][kMaxKrylov];
    __shared__ double ce[2][kMaxKrylov], se[2][kMaxKrylov];  // cos, sin(lam_j tend): T_m, T_{m-1}
    __shared__ double yr[2][kMaxKrylov], yi[2][kMaxKrylov];  // exp(-i tend T) e1: T_m, T_{m-1}
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < kMaxKrylov; ++j)  // constant indices keep tot in registers
            if (j < m) tq_[j] = tot[j];
    if (threadIdx.x < 64) {  // wavefront 0: T_m; wavefront 1: T_{m-1}
        const int r = leading_ql_wave(a, vec, m, d[0], ee[0], z[0], zr[0]);
        if (threadIdx.x == 0) bad = r;
    } else if (threadIdx.x < 128) {
        leading_ql_wave(a, vec, m - 1, d[1], ee[1], z[1], zr[1]);
    } else if (threadIdx.x == 128) {
        double d4[4], e4[4], z4[16];
        leading_ql(a, vec, 4, d4, e4, z4);
        c.est4[vec] = krylov_est(4, d4, z4, a.bt[vec * ms + 3], b0, c.tfeat);
    }
    __syncthreads();
    // Wavefront 0 forms the coefficients and the estimate, lane j taking
    // eigenvalue j (lanes 32 + j those of T_{m-1}) and then row i of the
    // products; every sum keeps the ascending order of the serial form.
    const double *dd = d[0], *zz = z[0];
    const double tend = c.nsub * c.tau;
    const int k = threadIdx.x;
    if (k < m) {
        double t = 0;  // (Q^T c)_k
        for (int i = 0; i < m; ++i) t += zz[k * m + i] * tq_[i];
        u2[k] = t;
        for (int q = 0; q < c.nsub; ++q) {
            const double tq = (q + 1) * c.tau;
            const double cs = cos(dd[k] * tq), sn = sin(dd[k] * tq);
            pc[q][k] = cs * t;
            ps[q][k] = ((vec & 1) ? sn : -sn) * t;
        }
        ce[0][k] = cos(dd[k] * tend);
        se[0][k] = sin(dd[k] * tend);
    } else if (k >= 32 && k < 32 + m - 1) {
        ce[1][k - 32] = cos(d[1][k - 32] * tend);
        se[1][k - 32] = sin(d[1][k - 32] * tend);
    }
    __syncthreads();
    if (k < m) {
        for (int q = 0; q < c.nsub; ++q) {  // Q (...)
            double tc = 0, ts = 0;
            for (int j = 0; j < m; ++j) {
                tc += pc[q][j] * zz[j * m + k];
                ts += ps[q][j] * zz[j * m + k];
            }
            double *cq = coef + q * 2 * ms;
            cq[k] = tc;
            cq[ms + k] = ts;
        }
        double sr = 0, si = 0;  // y_m[k]
        for (int j = 0; j < m; ++j) {
            const double w = zz[j * m + k] * zz[j * m];
            sr += w * ce[0][j];
            si -= w * se[0][j];
        }
        yr[0][k] = sr;
        yi[0][k] = si;
    } else if (k >= 32 && k < 32 + m - 1) {
        const int i = k - 32, m1 = m - 1;
        const double *z1 = z[1];
        double sr = 0, si = 0;  // y_{m-1}[i]
        for (int j = 0; j < m1; ++j) {
            const double w = z1[j * m1 + i] * z1[j * m1];
            sr += w * ce[1][j];
            si -= w * se[1][j];
        }
        yr[1][i] = sr;
        yi[1][i] = si;
    }
    __syncthreads();
    if (k == 0) {
        if (bad) *c.info = bad;
        // residual estimate of order m (krylov_est) and beta0 |y_{m-1} - y_m|
        double sr = 0, si = 0;
        for (int j = 0; j < m; ++j) {
            const double w = zz[j * m + m - 1] * zz[j * m];
            sr += w * ce[0][j];
            si -= w * se[0][j];
        }
        const double res = b0 * a.bt[vec * ms + m - 1] * sqrt(sr * sr + si * si);
        double sq = 0;
        for (int i = 0; i < m; ++i) {
            const double dr = yr[0][i] - (i < m - 1 ? yr[1][i] : 0.0);
            const double di = yi[0][i] - (i < m - 1 ? yi[1][i] : 0.0);
            sq += dr * dr + di * di;
        }
        c.est[vec] = fmax(res, b0 * sqrt(sq));
    }
}

// Rebuild psi at tau (and 2 tau) into xout; C(t) and |psi|^2 of each set
// into samples g0 (and g0 + 1) when emit.
template <typename S, int NT>
__global__ __launch_bounds__(NT) void ctl_combine(CtlArgs c, V2<S> *xout, const V2<S> *psi0,
                                                  double *cbuf, int maxsamp,
                                                  const long long *clock, int emit)
{
    const long long g0 = *clock + 1;  // the first sample this step completes
    __shared__ double red[(NT / 64) * kPart];
    __shared__ double cf[2][4 * kMaxKrylov];
    const LnzArgs &a = c.l;
    const int pk = blockIdx.y, m = c.mact, ms = a.m, ns = c.nsub;
    const int vr = 2 * pk, vi = 2 * pk + 1;
    for (int q = 0; q < ns; ++q)
        for (int i = threadIdx.x; i < 2 * m; i += NT) {
            const int part = i / m, j = i % m;  // part 0: cos, 1: sin
            cf[q][part * m + j] = c.coef[(long long)vr * 4 * ms + q * 2 * ms + part * ms + j];
            cf[q][2 * m + part * m + j] = c.coef[(long long)vi * 4 * ms + q * 2 * ms + part * ms + j];
        }
    __syncthreads();
    const S *Br = reinterpret_cast<const S *>(a.basis) + (long long)vr * a.ncols * a.n;
    const S *Bi = reinterpret_cast<const S *>(a.basis) + (long long)vi * a.ncols * a.n;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int p = blockIdx.x * NT + threadIdx.x; p < a.n; p += a.nb * NT) {
        double u[2] = {0, 0}, wv[2] = {0, 0}, v[2] = {0, 0}, wu[2] = {0, 0};
        for (int j = 0; j < m; ++j) {
            const double br = (double)Br[(long long)j * a.n + p];
            const double bi = (double)Bi[(long long)j * a.n + p];
            for (int q = 0; q < 2; ++q) {
                if (q >= ns) break;
                u[q] += cf[q][j] * br;
                wv[q] += cf[q][m + j] * br;
                v[q] += cf[q][2 * m + j] * bi;
                wu[q] += cf[q][3 * m + j] * bi;
            }
        }
        const V2<S> q0 = psi0[(long long)pk * a.n + p];
        for (int q = 0; q < 2; ++q) {
            if (q >= ns) break;
            V2<S> o;
            o.x = (S)(u[q] + wu[q]);
            o.y = (S)(wv[q] + v[q]);
            if (q == ns - 1) xout[(long long)pk * a.n + p] = o;
            const double ou = (double)o.x, ov = (double)o.y;
            acc[3 * q] += (double)q0.x * ou + (double)q0.y * ov;
            acc[3 * q + 1] += (double)q0.x * ov - (double)q0.y * ou;
            acc[3 * q + 2] += ou * ou + ov * ov;
        }
    }
    block_sum_k<NT>(acc, 3 * ns, red);
    double tot[6];
    if (last_block<NT>(a, vr, 0, acc, 3 * ns, tot, red) && threadIdx.x == 0 && emit)
        for (int q = 0; q < ns; ++q) {
            const long long g = g0 + q;
            if (g < maxsamp) {
                double *cb = cbuf + ((long long)pk * maxsamp + g) * 3;
                cb[0] = tot[3 * q];
                cb[1] = tot[3 * q + 1];
                cb[2] = tot[3 * q + 2];
            }
        }
}

// Controller inputs of each packet from its dominant recurrence (the
// larger of |Re psi|, |Im psi|): alpha_1, beta_1, alpha_2, beta_2 times the
// base step; log10 of beta_4 and beta_5 times the base step (a rough
// component of the state, amplified by H^j, dominates the Krylov vectors
// from about the fourth on); log10 err(4, tfeat) / 10 and
// log10 |1 - |psi|^2| / 16.
__global__ void ctl_feat(CtlArgs c, float *feat, int npk)
{
    const int pk = blockIdx.x * blockDim.x + threadIdx.x;
    if (pk >= npk) return;
    const LnzArgs &a = c.l;
    const double br = a.beta0[2 * pk], bi = a.beta0[2 * pk + 1];
    const int v = bi > br ? 2 * pk + 1 : 2 * pk;
    float *f = feat + (long long)pk * kMlpIn;
    for (int j = 0; j < 2; ++j) {
        f[2 * j] = (float)(a.al[v * a.m + j] * c.tfeat);
        f[2 * j + 1] = (float)(a.bt[v * a.m + j] * c.tfeat);
    }
    f[4] = (float)log10(fmax(a.bt[v * a.m + 3] * c.tfeat, 1e-30));
    f[5] = (float)log10(fmax(a.bt[v * a.m + 4] * c.tfeat, 1e-30));
    f[6] = (float)(log10(c.est4[v] + 1e-300) / 10.0);
    f[7] = (float)(log10(fabs(br * br + bi * bi - 1.0) + 1e-16) / 16.0);
}

// Training labels of a trace step (mp = l.m iterations): for each network
// order k = 4..20 and t = tau/2, tau, 2 tau, the estimate the adaptive
// step of order k + 1 would be judged by, max(residual estimate of order
// k + 1, beta0 |y_k - y_{k+1}|); one thread per k (needs mp >= 21).
__global__ __launch_bounds__(64) void ctl_labels(CtlArgs c, double *lab)
{
    const LnzArgs &a = c.l;
    const int vec = blockIdx.x, k = threadIdx.x, m = kLabM0 + k;
    if (k >= kLabM || m + 1 > a.m) return;
    double *o = lab + ((long long)vec * kLabM + k) * kLabT;
    if (a.dead[vec]) {
        for (int t = 0; t < kLabT; ++t) o[t] = 0;
        return;
    }
    constexpr int L = kLabM0 + kLabM;  // largest order used, 21
    double d0[L], z0[L * L], d1[L], z1[L * L], e[L];
    if (leading_ql(a, vec, m, d0, e, z0) || leading_ql(a, vec, m + 1, d1, e, z1)) {
        for (int t = 0; t < kLabT; ++t) o[t] = 1e300;
        return;
    }
    const double tt[kLabT] = {0.5 * c.tau, c.tau, 2.0 * c.tau};
    const double b0 = a.beta0[vec];
    for (int t = 0; t < kLabT; ++t)
        o[t] = fmax(krylov_est(m + 1, d1, z1, a.bt[vec * a.m + m], b0, tt[t]),
                    krylov_diff(m, d0, z0, m + 1, d1, z1, b0, tt[t]));
}

// Working storage of nvec Lanczos recurrences with ncols stored columns.
struct LnzBufs {
    DevBuf basis, vring, w0, w1, al, bt, beta0, rc, part, ticket, dead;
    hipError_t alloc(LnzArgs &l, size_t ss)
    {
        const size_t nv = l.nvec, n = l.n;
        hipError_t e;
        if ((e = basis.alloc(ss * nv * l.ncols * n))) return e;
        if ((e = hipMemset(basis.p, 0, basis.bytes))) return e;
        if ((e = vring.alloc(ss * 2 * nv * n))) return e;
        if ((e = w0.alloc(ss * nv * n))) return e;
        if ((e = w1.alloc(ss * nv * n))) return e;
        if ((e = al.alloc(sizeof(double) * nv * l.m))) return e;
        if ((e = bt.alloc(sizeof(double) * nv * l.m))) return e;
        if ((e = beta0.alloc(sizeof(double) * nv))) return e;
        if ((e = rc.alloc(sizeof(double) * nv * l.ncols))) return e;
        const size_t np1 = nv * l.nchunk * l.nb * kPart;  // one region of partials
        if ((e = part.alloc(sizeof(double) * 4 * np1))) return e;
        if ((e = ticket.alloc(sizeof(unsigned) * nv * l.nchunk))) return e;
        if ((e = hipMemset(ticket.p, 0, ticket.bytes))) return e;
        if ((e = dead.alloc(sizeof(int) * nv))) return e;
        l.basis = basis.p;
        l.vring = vring.p;
        l.w0 = w0.p;
        l.w1 = w1.p;
        l.al = al.as<double>();
        l.bt = bt.as<double>();
        l.beta0 = beta0.as<double>();
        l.rc = rc.as<double>();
        l.part = part.as<double>();
        l.pa = l.part + np1;
        l.pb = l.part + 2 * np1;
        l.pr = l.part + 3 * np1;
        l.bfrom = 0;
        l.ticket = ticket.as<unsigned>();
        l.dead = dead.as<int>();
        return hipSuccess;
    }
};

}  // namespace espec

using namespace espec;

struct espec_sil_plan {
    int ndim = 1, np[3] = {1, 1, 1}, n = 0, npk = 0, m = 0, lc = 1, maxsamp = 0;
    bool mixed = false, ref = false;
    bool nograph = false;  // ESPEC_NO_GRAPH: launch each step's kernels directly
    double cst = 0, abstol = 0;
    long long steps_done = 0;
    int nb = 0;
    hipStream_t stream = nullptr;
    DevBuf x, psi0, vpot, coef, info, cbuf, dstep;
    // pulse-driven steps (PPLNZ): dipole [npk][n], field per step, V - DM E
    DevBuf dm, efield, veff;
    DevBuf mask;  // .SMOOTHW VABC [npk][n] (PLNZAUX with ABSORB), when set
    bool pulse = false;
    long long nfield = 0;
    LnzBufs lb;
    SilArgs sa{};
    // step graphs of 8 steps and of one step (runs replay graphs unless nograph)
    hipGraphExec_t gexec = nullptr, gexec1 = nullptr;
    hipGraph_t graph = nullptr, graph1 = nullptr;
    // adaptive and trace runs: second state buffer, coefficient sets,
    // estimates, controller inputs/outputs and its bfloat16 weights
    DevBuf xalt, ccoef, cest, cest4, cfeat, cout, cwb;
    std::map<unsigned long long, hipGraphExec_t> cgraphs;  // adaptive/trace step graphs
    double *hpin = nullptr;  // pinned: estimates [nvec], then outputs [npk][2] (float)
    double *log = nullptr;   // optional per-attempt record of the next adaptive run
    int log_cap = 0, log_len = 0;
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (gexec1) (void)hipGraphExecDestroy(gexec1);
        if (graph1) (void)hipGraphDestroy(graph1);
        gexec = gexec1 = nullptr;
        graph = graph1 = nullptr;
    }
    ~espec_sil_plan()
    {
        drop_graph();
        for (auto &g : cgraphs) (void)hipGraphExecDestroy(g.second);
        if (hpin) (void)hipHostFree(hpin);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Grid geometry of an FD operator on an ndim grid (x fastest).
bool set_grid(LnzArgs &l, int ndim, const int *np, const double *shm)
{
    if (ndim < 1 || ndim > 3 || !np || !shm) return false;
    l.ndim = ndim;
    long long n = 1;
    int st = 1;
    for (int a = 0; a < 3; ++a) {
        if (a < ndim && np[a] < 4) return false;  // 5-point stencils (au1d.f)
        l.np[a] = a < ndim ? np[a] : 1;
        l.st[a] = st;
        st *= l.np[a];
        n *= l.np[a];
        l.shm[a] = a < ndim ? shm[a] : 0.0;
    }
    if (n > (1LL << 30)) return false;
    l.n = (int)n;
    return true;
}

int blocks_for(int n) { return std::max(1, std::min(256, (n + 255) / 256)); }

// The m iterations of LNZ for every recurrence of l, on stream s.
// Iterations i0 .. i1-1 and the last beta of a one-block grid in one
// launch: in LDS when the recurrence's vectors fit beside the kernel's
// static arrays within the default 64 KiB (no attribute call, which a
// stream capture may not record), else on global memory.
constexpr size_t kBlockLdsMax = 56 * 1024;

template <typename S>
hipError_t launch_block(const LnzArgs &lb, int i0, int i1, hipStream_t s)
{
    const size_t lds = lnz_lds_bytes(lb.n, lb.ncols, sizeof(S));
    if (lds <= kBlockLdsMax)
        hipLaunchKernelGGL((lnz_block_lds<S, 256>), dim3(1, lb.nvec), dim3(256), lds, s, lb, i0,
                           i1, 1);
    else
        hipLaunchKernelGGL((lnz_block<S, 256>), dim3(1, lb.nvec), dim3(256), 0, s, lb, i0, i1, 1);
    return hipGetLastError();
}

template <typename S>
hipError_t lnz_iterations(const LnzArgs &l, hipStream_t s)
{
    if (l.nb == 1) {  // one launch for all m iterations and the last beta
        LnzArgs lb = l;
        lb.bfrom = 0;
        return launch_block<S>(lb, 0, l.m, s);
    }
    dim3 grid(l.nb, l.nvec), gridp(l.nb, l.nvec, l.nchunk);
    for (int i = 0; i < l.m; ++i) {
        LnzArgs li = l;
        li.it = i;
        li.bfrom = 0;
        hipLaunchKernelGGL((lnz_apply<S, 256>), grid, dim3(256), 0, s, li);
        hipLaunchKernelGGL((lnz_project<S, 256>), gridp, dim3(256), 0, s, li);
        hipLaunchKernelGGL((lnz_update<S, 256>), grid, dim3(256), 0, s, li);
    }
    LnzArgs lf = l;
    lf.it = l.m - 1;
    hipLaunchKernelGGL(lnz_finish<256>, dim3(1, l.nvec), dim3(256), 0, s, lf);
    return hipGetLastError();
}

// Effective potential of a pulse-driven step (chpot.f:58-62):
// VEFF = VP - DM EF(T) with the field at the step's end time; the step
// that completes step count g + 1 reads efield[g] (g = *dstep).
__global__ __launch_bounds__(256) void sil_field(const double *vp, const double *dm,
                                                 const double *efield, const long long *dstep,
                                                 double *veff, long long n)
{
    // The product is rounded before the difference, as CHPOT's host code
    // does, so that VEFF matches the reference bit for bit.
#pragma clang fp contract(off)
    const double e = efield[*dstep];
    for (long long p = blockIdx.x * 256LL + threadIdx.x; p < n; p += gridDim.x * 256LL) {
        const double f = dm[p] * e;
        veff[p] = vp[p] - f;
    }
}

template <typename S>
hipError_t sil_step(espec_sil_plan &P, hipStream_t s)
{
    const LnzArgs &l = P.sa.l;
    dim3 grid(P.nb, l.nvec);
    if (P.pulse) {
        const long long tot = (long long)P.n * P.npk;
        hipLaunchKernelGGL(sil_field, dim3((unsigned)std::min<long long>(1024, (tot + 255) / 256)),
                           dim3(256), 0, s, P.vpot.as<double>(), P.dm.as<double>(),
                           P.efield.as<double>(), P.dstep.as<long long>(), P.veff.as<double>(),
                           tot);
    }
    hipLaunchKernelGGL((lnz_init<S, 256>), grid, dim3(256), 0, s, l, (const V2<S> *)P.x.p,
                       (const S *)nullptr);
    hipError_t e = lnz_iterations<S>(l, s);
    if (e != hipSuccess) return e;
    SilArgs sa = P.sa;
    sa.x0 = P.x.p;
    hipLaunchKernelGGL((sil_project<S, 256>), grid, dim3(256), 0, s, sa);
    hipLaunchKernelGGL((sil_coeffs<S, 256>), dim3(1, l.nvec), dim3(256), 0, s, sa);
    hipLaunchKernelGGL((sil_combine<S, 256>), dim3(P.nb, P.npk), dim3(256), 0, s, sa,
                       (V2<S> *)P.x.p, (const V2<S> *)P.psi0.p, P.cbuf.as<double>(), P.maxsamp,
                       P.lc, P.dstep.as<long long>());
    hipLaunchKernelGGL(sil_step_done, dim3(1), dim3(1), 0, s, P.dstep.as<long long>());
    return hipGetLastError();
}

template <typename S>
int sil_capture(espec_sil_plan &P, int steps, hipGraph_t *g, hipGraphExec_t *x)
{
    if (*x) return ESPEC_OK;
    hipStream_t cs;
    ESPEC_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    for (int k = 0; e == hipSuccess && k < steps; ++k) e = sil_step<S>(P, cs);
    hipError_t e2 = hipStreamEndCapture(cs, g);
    (void)hipStreamDestroy(cs);
    ESPEC_HIP_TRY(e);
    ESPEC_HIP_TRY(e2);
    ESPEC_HIP_TRY(hipGraphInstantiate(x, *g, nullptr, nullptr, 0));
    return ESPEC_OK;
}

template <typename S>
int sil_run_typed(espec_sil_plan &P, int nsteps)
{
    hipLaunchKernelGGL(sil_set_step, dim3(1), dim3(1), 0, P.stream, P.dstep.as<long long>(),
                       P.steps_done);
    ESPEC_HIP_TRY(hipGetLastError());
    if (P.nograph) {
        for (int k = 0; k < nsteps; ++k) ESPEC_HIP_TRY(sil_step<S>(P, P.stream));
        P.steps_done += nsteps;
        return ESPEC_OK;
    }
    const int gsteps = 8;
    int left = nsteps;
    if (left >= gsteps) {
        const int rc = sil_capture<S>(P, gsteps, &P.graph, &P.gexec);
        if (rc) return rc;
        while (left >= gsteps) {
            ESPEC_HIP_TRY(hipGraphLaunch(P.gexec, P.stream));
            left -= gsteps;
        }
    }
    if (left > 0) {
        const int rc = sil_capture<S>(P, 1, &P.graph1, &P.gexec1);
        if (rc) return rc;
        for (; left > 0; --left) ESPEC_HIP_TRY(hipGraphLaunch(P.gexec1, P.stream));
    }
    P.steps_done += nsteps;
    return ESPEC_OK;
}

int ctl_setup(espec_sil_plan &P)
{
    if (P.ccoef.p) return ESPEC_OK;
    const int nv = 2 * P.npk;
    const size_t ss = P.mixed ? sizeof(float) : sizeof(double);
    ESPEC_HIP_TRY(P.xalt.alloc(2 * ss * (size_t)P.n * P.npk));
    ESPEC_HIP_TRY(P.ccoef.alloc(sizeof(double) * nv * 4 * P.m));
    ESPEC_HIP_TRY(P.cest.alloc(sizeof(double) * nv));
    ESPEC_HIP_TRY(P.cest4.alloc(sizeof(double) * nv));
    ESPEC_HIP_TRY(P.cfeat.alloc(sizeof(float) * kMlpIn * P.npk));
    ESPEC_HIP_TRY(P.cout.alloc(sizeof(float) * kMlpOut * P.npk));
    ESPEC_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&P.hpin),
                                sizeof(double) * (nv + P.npk), hipHostMallocDefault));
    return ESPEC_OK;
}

CtlArgs ctl_args(espec_sil_plan &P, int mact, int nsub, double tau)
{
    CtlArgs c{};
    c.l = P.sa.l;
    c.x0 = P.x.p;
    c.mact = mact;
    c.nsub = nsub;
    c.tau = tau;
    c.tfeat = P.cst;
    c.coef = P.ccoef.as<double>();
    c.est = P.cest.as<double>();
    c.est4 = P.cest4.as<double>();
    c.info = P.info.as<int>();
    return c;
}

// One Krylov (sub)step from P.x: Lanczos iterations have..iters-1 (from
// P.x when have = 0), coefficients of order mact, estimates, controller
// inputs (and its outputs when wb is given) and the new state in P.xalt,
// sampled (when emit) into the sample after the device clock and the next.
template <typename S>
hipError_t ctl_launches(espec_sil_plan &P, const CtlArgs &c, const uint16_t *wb, int emit, int have,
                        int iters, hipStream_t s)
{
    const LnzArgs &l = c.l;
    const dim3 grid(P.nb, l.nvec), gridp(P.nb, l.nvec, l.nchunk);
    if (iters > have) {
        if (have == 0)
            hipLaunchKernelGGL((lnz_init<S, 256>), grid, dim3(256), 0, s, l, (const V2<S> *)P.x.p,
                               (const S *)nullptr);
        if (l.nb == 1) {  // one launch (extending: beta_{have-1} is in bt)
            LnzArgs lb = l;
            lb.bfrom = have > 0;
            hipError_t e = launch_block<S>(lb, have, iters, s);
            if (e != hipSuccess) return e;
        } else {
            for (int i = have; i < iters; ++i) {
                LnzArgs li = l;
                li.it = i;
                li.bfrom = i == have && have > 0;  // extending: beta_{have-1} is in bt
                hipLaunchKernelGGL((lnz_apply<S, 256>), grid, dim3(256), 0, s, li);
                hipLaunchKernelGGL((lnz_project<S, 256>), gridp, dim3(256), 0, s, li);
                hipLaunchKernelGGL((lnz_update<S, 256>), grid, dim3(256), 0, s, li);
            }
            LnzArgs lf = l;
            lf.it = iters - 1;
            hipLaunchKernelGGL(lnz_finish<256>, dim3(1, l.nvec), dim3(256), 0, s, lf);
        }
    }
    hipLaunchKernelGGL((ctl_project<S, 256>), grid, dim3(256), 0, s, c);
    hipLaunchKernelGGL((ctl_coeffs<S, 256>), dim3(1, grid.y), dim3(256), 0, s, c);
    hipLaunchKernelGGL(ctl_feat, dim3((P.npk + 63) / 64), dim3(64), 0, s, c, P.cfeat.as<float>(),
                       P.npk);
    if (wb) {
        hipError_t e = launch_mlp_infer(wb, P.cfeat.as<float>(), P.cout.as<float>(), P.npk, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((ctl_combine<S, 256>), dim3(P.nb, P.npk), dim3(256), 0, s, c,
                       (V2<S> *)P.xalt.p, (const V2<S> *)P.psi0.p, P.cbuf.as<double>(), P.maxsamp,
                       P.dstep.as<long long>(), emit);
    return hipGetLastError();
}

// The step's launches replay from a graph captured on first use (unless
// the plan has ESPEC_NO_GRAPH, which launches them directly): one
// graph per (order, iterations already run, coefficient sets, step, sample
// flag, network, which state buffer is current); the sample index comes
// from the device clock set just before.
template <typename S>
hipError_t ctl_step(espec_sil_plan &P, const CtlArgs &c, const uint16_t *wb, long long g0, int emit,
                    int have = 0, int iters = 0)
{
    const hipStream_t s = P.stream;
    iters = std::max(iters, c.mact);
    hipLaunchKernelGGL(sil_set_step, dim3(1), dim3(1), 0, s, P.dstep.as<long long>(), g0 - 1);
    const int tcode = c.tau == P.cst ? 0 : (c.tau == 0.5 * P.cst ? 1 : -1);
    if (tcode < 0 || P.nograph) return ctl_launches<S>(P, c, wb, emit, have, iters, s);
    const unsigned long long key = (unsigned long long)c.mact | ((unsigned long long)have << 8) |
                                   ((unsigned long long)iters << 32) |
                                   ((unsigned long long)c.nsub << 16) |
                                   ((unsigned long long)tcode << 20) |
                                   ((unsigned long long)emit << 22) |
                                   ((unsigned long long)(wb != nullptr) << 23) |
                                   ((unsigned long long)(P.x.p < P.xalt.p) << 24);
    auto it = P.cgraphs.find(key);
    if (it == P.cgraphs.end()) {
        hipStream_t cs;
        hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
        if (e != hipSuccess) return e;
        hipGraph_t g = nullptr;
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) e = ctl_launches<S>(P, c, wb, emit, have, iters, cs);
        hipError_t e2 = hipStreamEndCapture(cs, &g);
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess) return e;
        if (e2 != hipSuccess) return e2;
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return e;
        it = P.cgraphs.emplace(key, x).first;
    }
    return hipGraphLaunch(it->second, s);
}

// Estimates (and controller outputs) of the last step to the host.
int ctl_fetch(espec_sil_plan &P, bool outputs)
{
    const int nv = 2 * P.npk;
    ESPEC_HIP_TRY(hipMemcpyAsync(P.hpin, P.cest.p, sizeof(double) * nv, hipMemcpyDeviceToHost,
                                 P.stream));
    if (outputs)
        ESPEC_HIP_TRY(hipMemcpyAsync(P.hpin + nv, P.cout.p, sizeof(float) * kMlpOut * P.npk,
                                     hipMemcpyDeviceToHost, P.stream));
    ESPEC_HIP_TRY(hipStreamSynchronize(P.stream));
    return ESPEC_OK;
}

// Adopt the state the last step wrote to P.xalt.
void ctl_accept(espec_sil_plan &P) { std::swap(P.x.p, P.xalt.p); }

// stats: [0] macro steps, [1..3] steps at scale 1/2, 1, 2, [4] fallbacks,
// [5] sum of Krylov orders over all (sub)steps taken, [6] host waits.
template <typename S>
int adaptive_typed(espec_sil_plan &P, const uint16_t *wb, int ndt, double tol, int m_fix,
                   long long *stats)
{
    const double dt = P.cst;
    const int nv = 2 * P.npk;
    int m = m_fix, l2s = 0;  // Krylov order and log2 of the step scale
    int used = 0;            // order the last accepted attempt needed beyond the network's
    long long clock = P.steps_done;
    int left = ndt;
    long long st[7] = {0, 0, 0, 0, 0, 0, 0};
    auto decide = [&]() {
        if (!wb) return;
        const float *o = reinterpret_cast<const float *>(P.hpin + nv);
        int mm = kLabM0, ll = 1;
        for (int pk = 0; pk < P.npk; ++pk) {
            const int mq = (int)std::ceil(kLabM0 + 16.0f * o[2 * pk]) + kMarginM;  // iterations
            const int lq = (int)std::lround(2.0f * o[2 * pk + 1] - 1.0f);
            mm = std::max(mm, mq);
            ll = std::min(ll, lq);
        }
        // Feedback floor: a step that needed the extended space leaves the
        // next one at most one order below the order it used.
        m = std::min(std::max(std::max(mm, kLabM0 + 1), used - 1), P.m);
        l2s = std::min(std::max(ll, -1), 1);
    };
    while (left > 0) {
        if (l2s > 0 && left < 2) l2s = 0;
        const int subs = l2s < 0 ? 2 : 1, nsub = l2s > 0 ? 2 : 1;
        const double tau = l2s < 0 ? 0.5 * dt : dt;
        const double budget = tol * nsub / subs;
        int adv = nsub;
        st[0]++;
        st[2 + l2s]++;
        for (int h = 0; h < subs; ++h) {
            const int emit = h == subs - 1;
            ESPEC_HIP_TRY(ctl_step<S>(P, ctl_args(P, m, nsub, tau), wb, clock + 1, emit));
            int rc = ctl_fetch(P, wb != nullptr);
            if (rc) return rc;
            st[5] += m;
            st[6]++;
            double e = 0;
            for (int v = 0; v < nv; ++v) e = std::max(e, P.hpin[v]);
            if (P.log && P.log_len < P.log_cap) {  // (step, order, scale, estimate, budget)
                double *r = P.log + 5 * (size_t)P.log_len++;
                r[0] = (double)clock;
                r[1] = m;
                r[2] = subs == 2 ? 0.5 : nsub;
                r[3] = e;
                r[4] = budget;
            }
            int have = m;
            used = 0;
            if (wb && !(e <= budget) && m < P.m) {
                // First the same step with the Krylov space extended to the
                // plan's largest order.
                ESPEC_HIP_TRY(ctl_step<S>(P, ctl_args(P, P.m, nsub, tau), wb, clock + 1, emit, m));
                rc = ctl_fetch(P, true);
                if (rc) return rc;
                st[5] += P.m - m;
                st[6]++;
                have = P.m;
                e = 0;
                for (int v = 0; v < nv; ++v) e = std::max(e, P.hpin[v]);
                if (e <= budget) used = P.m;
            }
            if (wb && !(e <= budget)) {
                // Fallback: the fixed-order step over tau (one base step
                // when the rejected step covered two).  Its Krylov space is
                // the rejected step's, extended to order m_fix if shorter;
                // when that step misses the budget of one base step too,
                // to the plan's largest order.
                const double b1 = tol / subs;
                int mf = m_fix;
                for (;;) {
                    ESPEC_HIP_TRY(
                        ctl_step<S>(P, ctl_args(P, mf, 1, tau), wb, clock + 1, emit, have));
                    rc = ctl_fetch(P, true);
                    if (rc) return rc;
                    st[5] += std::max(0, mf - have);
                    st[6]++;
                    have = std::max(have, mf);
                    e = 0;
                    for (int v = 0; v < nv; ++v) e = std::max(e, P.hpin[v]);
                    if (e <= b1 || mf >= P.m) break;
                    mf = P.m;
                }
                st[4]++;
                adv = 1;
            }
            ctl_accept(P);
        }
        if (l2s < 0) adv = 1;
        clock += adv;
        left -= adv;
        decide();
    }
    P.steps_done = clock;
    if (stats)
        for (int i = 0; i < 7; ++i) stats[i] = st[i];
    int info = 0;
    ESPEC_HIP_TRY(hipMemcpy(&info, P.info.p, sizeof(int), hipMemcpyDeviceToHost));
    return info ? ESPEC_ERR_DIAG : ESPEC_OK;
}

template <typename S>
int trace_typed(espec_sil_plan &P, int nsteps, int mprop, float *feat, double *lab)
{
    const int nv = 2 * P.npk;
    const size_t nf = (size_t)P.npk * kMlpIn, nl = (size_t)nv * kLabM * kLabT;
    DevBuf df, dl;
    ESPEC_HIP_TRY(df.alloc(sizeof(float) * nf * nsteps));
    ESPEC_HIP_TRY(dl.alloc(sizeof(double) * nl * nsteps));
    long long clock = P.steps_done;
    for (int k = 0; k < nsteps; ++k) {
        // mp iterations for the labels; the state advances with order mprop
        const CtlArgs c = ctl_args(P, mprop, 1, P.cst);
        ESPEC_HIP_TRY(ctl_step<S>(P, c, nullptr, clock + 1, 1, 0, P.m));
        hipLaunchKernelGGL(ctl_labels, dim3(nv), dim3(64), 0, P.stream, c, dl.as<double>() + nl * k);
        ESPEC_HIP_TRY(hipGetLastError());
        ESPEC_HIP_TRY(hipMemcpyAsync(df.as<float>() + nf * k, P.cfeat.p, sizeof(float) * nf,
                                     hipMemcpyDeviceToDevice, P.stream));
        ctl_accept(P);  // later launches read the new state; the queue keeps the order
        ++clock;
    }
    ESPEC_HIP_TRY(hipStreamSynchronize(P.stream));
    ESPEC_HIP_TRY(hipMemcpy(feat, df.p, sizeof(float) * nf * nsteps, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(lab, dl.p, sizeof(double) * nl * nsteps, hipMemcpyDeviceToHost));
    P.steps_done = clock;
    int info = 0;
    ESPEC_HIP_TRY(hipMemcpy(&info, P.info.p, sizeof(int), hipMemcpyDeviceToHost));
    return info ? ESPEC_ERR_DIAG : ESPEC_OK;
}

}  // namespace

// <psi0|psi> for sample 0 (spo_lds.hip).
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
    if (!out || npk < 1 || mp < 2 || mp > kMaxKrylov || lc < 1 || !vpot || !u || maxsamp < 1)
        return ESPEC_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<espec_sil_plan> P(new espec_sil_plan);
    LnzArgs &l = P->sa.l;
    if (!set_grid(l, ndim, np, shm)) return ESPEC_ERR_ARG;
    P->ndim = ndim;
    for (int a = 0; a < ndim; ++a) P->np[a] = np[a];
    P->n = l.n;
    P->npk = npk;
    P->m = mp;
    P->lc = lc;
    P->maxsamp = maxsamp;
    P->mixed = (flags & ESPEC_MIXED) != 0;
    P->ref = (flags & ESPEC_REF) != 0;
    P->nograph = (flags & ESPEC_NO_GRAPH) != 0;
    P->cst = dt / kShfs;
    P->abstol = abstol;
    const int n = P->n, nv = 2 * npk;
    P->nb = blocks_for(n);
    ESPEC_HIP_TRY(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    const size_t ss = P->mixed ? sizeof(float) : sizeof(double);
    if (P->mixed) {
        std::vector<float> xs = interleave<float>(u, v, (size_t)n * npk);
        ESPEC_HIP_TRY(P->x.upload(xs.data(), sizeof(float) * xs.size()));
        ESPEC_HIP_TRY(P->psi0.upload(xs.data(), sizeof(float) * xs.size()));
    } else {
        std::vector<double> xs = interleave<double>(u, v, (size_t)n * npk);
        ESPEC_HIP_TRY(P->x.upload(xs.data(), sizeof(double) * xs.size()));
        ESPEC_HIP_TRY(P->psi0.upload(xs.data(), sizeof(double) * xs.size()));
    }
    ESPEC_HIP_TRY(P->vpot.upload(vpot, sizeof(double) * (size_t)n * npk));
    l.vpot = P->vpot.as<double>();
    l.vshift = 1;
    l.nvec = nv;
    l.m = mp;
    l.nb = P->nb;
    l.ncols = mp;
    l.nchunk = 1;
    l.nreort = mp;
    l.acoef = nullptr;
    l.acc = nullptr;
    l.nacc = 0;
    ESPEC_HIP_TRY(P->lb.alloc(l, ss));
    ESPEC_HIP_TRY(P->coef.alloc(sizeof(double) * nv * 2 * mp));
    ESPEC_HIP_TRY(P->info.alloc(sizeof(int)));
    ESPEC_HIP_TRY(hipMemset(P->info.p, 0, sizeof(int)));
    ESPEC_HIP_TRY(P->cbuf.alloc(sizeof(double) * 3 * (size_t)maxsamp * npk));
    ESPEC_HIP_TRY(hipMemset(P->cbuf.p, 0, P->cbuf.bytes));
    ESPEC_HIP_TRY(P->dstep.alloc(sizeof(long long)));
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
    if (P->pulse && P->steps_done + nsteps > P->nfield) return ESPEC_ERR_ARG;
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

int espec_sil_plan_adaptive(espec_sil_plan *P, const float *params, int ndt, double tol,
                            int m_fix, long long *stats)
{
    if (!P || ndt < 0 || !(tol > 0) || m_fix < kLabM0 || m_fix > P->m || P->lc != 1 || P->ref)
        return ESPEC_ERR_ARG;
    if (P->pulse || P->sa.mask) return ESPEC_ERR_UNSUPPORTED;
    if (ndt == 0) return ESPEC_OK;
    int rc = ctl_setup(*P);
    if (rc) return rc;
    P->drop_graph();  // the state buffers trade places
    const uint16_t *wb = nullptr;
    if (params) {
        std::vector<uint16_t> h(kMlpNPar);
        for (int i = 0; i < kMlpNPar; ++i) h[i] = f2bf(params[i]);
        if (!P->cwb.p) ESPEC_HIP_TRY(P->cwb.alloc(sizeof(uint16_t) * kMlpNPar));
        ESPEC_HIP_TRY(hipMemcpy(P->cwb.p, h.data(), sizeof(uint16_t) * h.size(),
                                hipMemcpyHostToDevice));
        wb = P->cwb.as<uint16_t>();
    }
    return P->mixed ? adaptive_typed<float>(*P, wb, ndt, tol, m_fix, stats)
                    : adaptive_typed<double>(*P, wb, ndt, tol, m_fix, stats);
}

int espec_sil_plan_reset(espec_sil_plan *P)
{
    if (!P) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipMemcpyAsync(P->x.p, P->psi0.p, P->x.bytes, hipMemcpyDeviceToDevice, P->stream));
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    P->steps_done = 0;  // sample 0, C(psi0, psi0), is already in place
    return ESPEC_OK;
}

int espec_sil_plan_adaptive_log(espec_sil_plan *P, double *log, int cap, int *len)
{
    if (!P || cap < 0) return ESPEC_ERR_ARG;
    if (len) *len = P->log_len;
    P->log = log;
    P->log_cap = log ? cap : 0;
    P->log_len = 0;
    return ESPEC_OK;
}

int espec_sil_plan_trace(espec_sil_plan *P, int nsteps, int mprop, float *feat, double *lab)
{
    if (!P || nsteps < 1 || !feat || !lab || P->m < kLabM0 + kLabM || P->lc != 1 || P->ref ||
        mprop < kLabM0 + 1 || mprop > P->m)
        return ESPEC_ERR_ARG;
    if (P->pulse || P->sa.mask) return ESPEC_ERR_UNSUPPORTED;
    int rc = ctl_setup(*P);
    if (rc) return rc;
    P->drop_graph();
    return P->mixed ? trace_typed<float>(*P, nsteps, mprop, feat, lab)
                    : trace_typed<double>(*P, nsteps, mprop, feat, lab);
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

int espec_sil_plan_set_pulse(espec_sil_plan *P, const double *dm, const double *efield,
                             int nfield)
{
    if (!P || !dm || !efield || nfield < 0) return ESPEC_ERR_ARG;
    const size_t tot = (size_t)P->n * P->npk;
    ESPEC_HIP_TRY(P->dm.upload(dm, sizeof(double) * tot));
    ESPEC_HIP_TRY(P->efield.upload(efield, sizeof(double) * (size_t)(nfield > 0 ? nfield : 1)));
    if (!P->veff.p) ESPEC_HIP_TRY(P->veff.alloc(sizeof(double) * tot));
    P->nfield = nfield;
    P->pulse = true;
    P->sa.l.vpot = P->veff.as<double>();
    P->drop_graph();  // the step's launches change
    return ESPEC_OK;
}

int espec_sil_plan_set_absorb(espec_sil_plan *P, int absorb, const double *vabc)
{
    if (!P || absorb < ESPEC_ABSORB_NONE || absorb > ESPEC_ABSORB_SMOOTHW ||
        (absorb == ESPEC_ABSORB_SMOOTHW && !vabc))
        return ESPEC_ERR_ARG;
    // PLNZAUX tests TPABSOR(1:8) = '.SMOOTHW' only (sil.f:171): .VOPTIC has
    // no effect on the Lanczos step.
    if (absorb == ESPEC_ABSORB_SMOOTHW) {
        ESPEC_HIP_TRY(P->mask.upload(vabc, sizeof(double) * (size_t)P->n * P->npk));
        P->sa.mask = P->mask.as<double>();
    } else {
        P->sa.mask = nullptr;
    }
    P->drop_graph();  // the captured combine carries the mask pointer
    return ESPEC_OK;
}

int espec_plnz_absorb(int ndim, const int *np, int mp, double abstol, double dt, double ti,
                      double tf, int lc, const double *shm, double *u, double *v,
                      const double *vpot, double *cr, double *ci, double *nrm, int maxsamp,
                      int *nsamp, int flags, int absorb, const double *vabc)
{
    if (!u || !v || !cr || !ci || maxsamp < 1) return ESPEC_ERR_ARG;
    const int nsteps = fortran_trip(ti + dt, tf + dt, dt);  // plnz.f:109
    espec_sil_plan *P = nullptr;
    int rc = espec_sil_plan_create(&P, ndim, np, 1, mp, abstol, dt, lc, shm, vpot, u, v, maxsamp,
                                   flags);
    if (rc) return rc;
    rc = espec_sil_plan_set_absorb(P, absorb, vabc);
    if (!rc) rc = espec_sil_plan_run(P, nsteps);
    if (!rc) rc = espec_sil_plan_get(P, 0, cr, ci, nrm, maxsamp, nsamp, u, v);
    espec_sil_plan_destroy(P);
    return rc;
}

double espec_ef(int efc, double e0, double t0, double td, double tp, double omg, double sni,
                int kl, double t)
{
    // ef.f: FAT, FATE (V/cm -> a.u./D) and FATT (eV -> rad/fs) as the
    // reference writes them; .GGAUS takes the default-REAL (f32) 0.693147181.
    const double pi = 3.14159265358979323846;
    const double fat = 2.1132e-9, fate = 2.29371276e+17 * 3.335664e-28, fatt = 1.5192676;
    switch (efc) {
    case ESPEC_PULSE_NONE:
        return 0.0;
    case ESPEC_PULSE_GAUS:
        return e0 * std::exp(-(4.0 * 0.693147181 * (t - t0) * (t - t0) / (tp * tp))) *
               std::cos(omg * t * fatt + sni) * fate;
    case ESPEC_PULSE_GGAUS: {
        const double a1 = 1.0 / (2.0 * kl);
        const double tpa = tp / std::pow((double)0.693147181f, a1);
        return fat * std::sqrt(e0 * std::exp(-std::pow((t - t0) / tpa, 2.0 * kl))) *
               std::cos(omg * t * fatt + sni);
    }
    case ESPEC_PULSE_SIN2:
        return e0 * std::cos(pi * (t - td) / tp) * std::cos(pi * (t - td) / tp) *
               std::cos(omg * (t - td) * fatt + sni) * fate;
    default:
        return std::nan("");
    }
}

int espec_pplnz_absorb(int ndim, const int *np, int mp, double abstol, double dt, double ti,
                       double tf, int lc, const double *shm, double *u, double *v,
                       const double *vp, const double *dm, int efc, double e0, double t0,
                       double td, double tp, double omg, double sni, int kl, double *veff,
                       double *cr, double *ci, double *nrm, int maxsamp, int *nsamp, int flags,
                       int absorb, const double *vabc)
{
    if (!u || !v || !vp || !dm || !cr || !ci || maxsamp < 1 || efc < ESPEC_PULSE_NONE ||
        efc > ESPEC_PULSE_SIN2)
        return ESPEC_ERR_ARG;
    // DO T=TI+DT,TF+DT/2,DT (pplnz.f:190), T advanced by DT each pass.
    const int nsteps = fortran_trip(ti + dt, tf + dt / 2, dt);
    std::vector<double> ef((size_t)std::max(nsteps, 1));
    double t = ti + dt;
    for (int k = 0; k < nsteps; ++k, t += dt) ef[k] = espec_ef(efc, e0, t0, td, tp, omg, sni, kl, t);
    espec_sil_plan *P = nullptr;
    int rc = espec_sil_plan_create(&P, ndim, np, 1, mp, abstol, dt, lc, shm, vp, u, v, maxsamp,
                                   flags);
    if (rc) return rc;
    rc = espec_sil_plan_set_pulse(P, dm, ef.data(), nsteps);
    if (!rc) rc = espec_sil_plan_set_absorb(P, absorb, vabc);
    if (!rc) rc = espec_sil_plan_run(P, nsteps);
    if (!rc) rc = espec_sil_plan_get(P, 0, cr, ci, nrm, maxsamp, nsamp, u, v);
    // VPOT as the reference leaves it: the VEFF of the last step, which the
    // device formed (sil_field), or VP when no step ran.
    if (!rc && veff) {
        const size_t by = sizeof(double) * (size_t)P->n;
        if (nsteps > 0) {
            hipError_t e = hipMemcpy(veff, P->veff.p, by, hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = ESPEC_ERR_HIP;
        } else {
            std::memcpy(veff, vp, by);
        }
    }
    espec_sil_plan_destroy(P);
    return rc;
}

int espec_pplnz(int ndim, const int *np, int mp, double abstol, double dt, double ti, double tf,
                int lc, const double *shm, double *u, double *v, const double *vp,
                const double *dm, int efc, double e0, double t0, double td, double tp,
                double omg, double sni, int kl, double *veff, double *cr, double *ci,
                double *nrm, int maxsamp, int *nsamp, int flags)
{
    return espec_pplnz_absorb(ndim, np, mp, abstol, dt, ti, tf, lc, shm, u, v, vp, dm, efc, e0,
                              t0, td, tp, omg, sni, kl, veff, cr, ci, nrm, maxsamp, nsamp, flags,
                              ESPEC_ABSORB_NONE, nullptr);
}

int espec_au(int ndim, const int *np, const double *shm, const double *vpot, const double *u,
             double *hu)
{
    LnzArgs l{};
    if (!vpot || !u || !hu || !set_grid(l, ndim, np, shm)) return ESPEC_ERR_ARG;
    const size_t bytes = sizeof(double) * l.n;
    DevBuf dv, du, dh;
    ESPEC_HIP_TRY(dv.upload(vpot, bytes));
    ESPEC_HIP_TRY(du.upload(u, bytes));
    ESPEC_HIP_TRY(dh.alloc(bytes));
    l.vpot = dv.as<double>();
    hipLaunchKernelGGL(au_kernel, dim3(std::min(1024, (l.n + 255) / 256)), dim3(256), 0, nullptr,
                       l, du.as<double>(), dh.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(hu, dh.p, bytes, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_fcorrel(int n, const double *u1, const double *u2, double *out)
{
    if (n < 0 || !u1 || !u2 || !out) return ESPEC_ERR_ARG;
    if (n == 0) {
        *out = 0;
        return ESPEC_OK;
    }
    LnzArgs l{};
    l.n = n;
    l.nb = blocks_for(n);
    l.nchunk = 1;
    DevBuf a, b, part, ticket, res;
    ESPEC_HIP_TRY(a.upload(u1, sizeof(double) * n));
    ESPEC_HIP_TRY(b.upload(u2, sizeof(double) * n));
    ESPEC_HIP_TRY(part.alloc(sizeof(double) * l.nb * kPart));
    ESPEC_HIP_TRY(ticket.alloc(sizeof(unsigned)));
    ESPEC_HIP_TRY(hipMemset(ticket.p, 0, sizeof(unsigned)));
    ESPEC_HIP_TRY(res.alloc(sizeof(double)));
    l.part = part.as<double>();
    l.ticket = ticket.as<unsigned>();
    hipLaunchKernelGGL(dot_kernel, dim3(l.nb), dim3(256), 0, nullptr, l, a.as<double>(),
                       b.as<double>(), res.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(out, res.p, sizeof(double), hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
static int scratch_address(void **out)
{
    static std::mutex mu;
    static std::map<int, void *> addr;
    int dev = 0;
    ESPEC_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = addr.find(dev);
    if (it == addr.end()) {
        void *p = nullptr;
        ESPEC_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return ESPEC_ERR_HIP;
        it = addr.emplace(dev, p).first;
    }
    *out = it->second;
    return ESPEC_OK;
}

int espec_fcorrel_noalloc(int n, const double *u1, const double *u2, double *out, void *stream)
{
    if (n < 0 || !u1 || !u2 || !out) return ESPEC_ERR_ARG;
    if (n > ESPEC_NOALLOC_MAX_N) {
        std::fprintf(stderr, "espec_fcorrel_noalloc: n = %d exceeds the maximum %d\n", n,
                     ESPEC_NOALLOC_MAX_N);
        return ESPEC_ERR_ARG;
    }
    void *scr = nullptr;
    const int rc = scratch_address(&scr);
    if (rc) return rc;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) {
        hipLaunchKernelGGL(scratch_zero_out, dim3(1), dim3(1), 0, s, out);
        ESPEC_HIP_TRY(hipGetLastError());
        return ESPEC_OK;
    }
    LnzArgs l{};
    l.n = n;
    l.nb = blocks_for(n);
    l.nchunk = 1;
    if (l.nb > kScrBlocks) return ESPEC_ERR_ARG;
    l.part = static_cast<double *>(scr);
    l.ticket = reinterpret_cast<unsigned *>(static_cast<char *>(scr) + kScrTicket);
    // The ticket starts at zero (a kernel, not a memset, so that the call
    // can be captured into a graph); the last block leaves it at zero too.
    hipLaunchKernelGGL(scratch_clear_ticket, dim3(1), dim3(1), 0, s, l.ticket);
    hipLaunchKernelGGL(dot_kernel, dim3(l.nb), dim3(256), 0, s, l, u1, u2, out);
    ESPEC_HIP_TRY(hipGetLastError());
    return ESPEC_OK;
}

int espec_lnz(int ndim, const int *np, int m, int nreort, const double *shm, const double *vpot,
              const double *start, double *al, double *bt, double *basis)
{
    LnzArgs l{};
    if (m < 1 || nreort < 1 || nreort > kMaxReort || !vpot || !start || !al || !bt ||
        !set_grid(l, ndim, np, shm))
        return ESPEC_ERR_ARG;
    DevBuf dv, ds;
    ESPEC_HIP_TRY(dv.upload(vpot, sizeof(double) * l.n));
    ESPEC_HIP_TRY(ds.upload(start, sizeof(double) * l.n));
    l.vpot = dv.as<double>();
    l.vshift = 0;
    l.nvec = 1;
    l.m = m;
    l.nb = blocks_for(l.n);
    l.ncols = nreort;
    l.nchunk = (nreort + kChunk - 1) / kChunk;
    l.nreort = nreort;
    LnzBufs bufs;
    ESPEC_HIP_TRY(bufs.alloc(l, sizeof(double)));
    hipLaunchKernelGGL((lnz_init<double, 256>), dim3(l.nb, 1), dim3(256), 0, nullptr, l,
                       (const double2 *)nullptr, ds.as<double>());
    ESPEC_HIP_TRY(hipGetLastError());
    // A zero start runs as zeros (the first lnz_apply sets the flag) and
    // leaves AL, BT untouched (lnz.f:86).
    ESPEC_HIP_TRY(lnz_iterations<double>(l, nullptr));
    int dead = 0;
    ESPEC_HIP_TRY(hipMemcpy(&dead, bufs.dead.p, sizeof(int), hipMemcpyDeviceToHost));
    if (!dead) {
        ESPEC_HIP_TRY(hipMemcpy(al, bufs.al.p, sizeof(double) * m, hipMemcpyDeviceToHost));
        ESPEC_HIP_TRY(hipMemcpy(bt, bufs.bt.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    }
    if (basis)
        ESPEC_HIP_TRY(hipMemcpy(basis, bufs.basis.p, sizeof(double) * (size_t)nreort * l.n,
                                hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_dstevx(int m, const double *d, const double *e, double abstol, double *w, double *z)
{
    if (m < 1 || m > 4096 || !d || !w || (m > 1 && !e)) return ESPEC_ERR_ARG;
    std::vector<double> ee(m, 0.0);
    for (int i = 0; i + 1 < m; ++i) ee[i] = e[i];
    DevBuf dd, de, de2, dw, dz, di;
    ESPEC_HIP_TRY(dd.upload(d, sizeof(double) * m));
    ESPEC_HIP_TRY(de.upload(ee.data(), sizeof(double) * m));
    ESPEC_HIP_TRY(de2.alloc(sizeof(double) * m));
    ESPEC_HIP_TRY(dw.alloc(sizeof(double) * m));
    ESPEC_HIP_TRY(dz.alloc(sizeof(double) * (size_t)m * m));
    ESPEC_HIP_TRY(di.alloc(sizeof(int)));
    hipLaunchKernelGGL(dstevx_kernel, dim3(1), dim3(64), 0, nullptr, m, dd.as<double>(),
                       de.as<double>(), de2.as<double>(), abstol, dw.as<double>(),
                       dz.as<double>(), di.as<int>());
    ESPEC_HIP_TRY(hipGetLastError());
    int info = 0;
    ESPEC_HIP_TRY(hipMemcpy(&info, di.p, sizeof(int), hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(w, dw.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (z) ESPEC_HIP_TRY(hipMemcpy(z, dz.p, sizeof(double) * (size_t)m * m, hipMemcpyDeviceToHost));
    return info ? ESPEC_ERR_DIAG : ESPEC_OK;
}

int espec_lanczsg(int ndim, const int *np, int il, int iu, int maxint, double abstol,
                  const double *shm, const double *vpot, double *eigvl, double *eigvc,
                  double *log, int logcap, int *nlog, int flags)
{
    LnzArgs l{};
    if (!vpot || !eigvl || !eigvc || maxint < 1 || !set_grid(l, ndim, np, shm))
        return ESPEC_ERR_ARG;
    const int m = 40;  // lanczsc.f:79 sets NREORT = 40 whatever the caller asked
    const int n = l.n, ns = iu - il + 1;
    if (n <= m || il < 1 || iu > m || ns < 1 || ns > 64) return ESPEC_ERR_ARG;
    const bool ref = (flags & ESPEC_REF) != 0;
    DevBuf dv, dvi, daux, dcoef, de2, dw, dwork, dinfo, dsum;
    ESPEC_HIP_TRY(dv.upload(vpot, sizeof(double) * n));
    ESPEC_HIP_TRY(dvi.alloc(sizeof(double) * n));
    ESPEC_HIP_TRY(daux.alloc(sizeof(double) * n));
    ESPEC_HIP_TRY(dcoef.alloc(sizeof(double) * (size_t)ns * m));
    ESPEC_HIP_TRY(de2.alloc(sizeof(double) * m));
    ESPEC_HIP_TRY(dw.alloc(sizeof(double) * ns));
    ESPEC_HIP_TRY(dwork.alloc(sizeof(double) * (size_t)ns * 6 * m));
    ESPEC_HIP_TRY(dinfo.alloc(sizeof(int)));
    ESPEC_HIP_TRY(dsum.alloc(sizeof(double) * 2));
    l.vpot = dv.as<double>();
    l.vshift = 0;
    l.nvec = 1;
    l.m = m;
    l.nb = blocks_for(n);
    l.ncols = m;
    l.nchunk = (m + kChunk - 1) / kChunk;
    l.nreort = m;
    LnzBufs bufs;
    ESPEC_HIP_TRY(bufs.alloc(l, sizeof(double)));
    // RAN with NSEED = 1 (lanczsc.f:78-82): the first N values are drawn
    // and overwritten, each state then starts from the next N.
    long long seed = 1;
    auto ran = [&]() {
        seed = (seed * 1029 + 221591) % 1048576;
        return (double)seed / 1048576.0;
    };
    std::vector<double> h(n), wstate(ns);
    for (int i = 0; i < n; ++i) (void)ran();
    int nl = 0;
    for (int k = 0; k < ns; ++k) {
        for (int i = 0; i < n; ++i) h[i] = ran();
        ESPEC_HIP_TRY(hipMemcpy(dvi.p, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        int kk = 0;
        double sum = 0;
        do {
            ++kk;
            ESPEC_HIP_TRY(hipMemcpy(daux.p, dvi.p, sizeof(double) * n, hipMemcpyDeviceToDevice));
            // LNZ(M = NREORT = 40) from VI (lanczsc.f:96), DSTEVX('V','I') (:98)
            hipLaunchKernelGGL((lnz_init<double, 256>), dim3(l.nb, 1), dim3(256), 0, nullptr, l,
                               (const double2 *)nullptr, dvi.as<double>());
            ESPEC_HIP_TRY(lnz_iterations<double>(l, nullptr));
            ESPEC_HIP_TRY(hipMemset(dinfo.p, 0, sizeof(int)));
            hipLaunchKernelGGL(lanczs_select, dim3(1), dim3(64), 0, nullptr, m, l.al, l.bt,
                               de2.as<double>(), il, ns, abstol, dw.as<double>(),
                               dcoef.as<double>(), dwork.as<double>(), dinfo.as<int>());
            // the Ritz vector of state k (lanczsc.f:100-106)
            hipLaunchKernelGGL(ritz_vector, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0,
                               nullptr, n, m, bufs.basis.as<double>(),
                               dcoef.as<double>() + (size_t)k * m, dvi.as<double>());
            ESPEC_HIP_TRY(hipGetLastError());
            int info = 0;
            ESPEC_HIP_TRY(hipMemcpy(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost));
            if (info) return ESPEC_ERR_DIAG;
            if (ref) {
                // lanczsc.f:104-109 sums ABS(VI(I) - VIAUX(I)) with I = NREORT + 1,
                // the index the finished inner loop leaves: VI(41) is still zero
                // for J <= 40 and final afterwards.
                double a41 = 0, v41 = 0;
                ESPEC_HIP_TRY(hipMemcpy(&a41, daux.as<double>() + m, sizeof(double),
                                        hipMemcpyDeviceToHost));
                ESPEC_HIP_TRY(hipMemcpy(&v41, dvi.as<double>() + m, sizeof(double),
                                        hipMemcpyDeviceToHost));
                sum = 0;
                for (int j = 1; j <= n; ++j) sum = sum + std::fabs((j <= m ? 0.0 : v41) - a41);
            } else {
                hipLaunchKernelGGL(diff_sum, dim3(1), dim3(1024), 0, nullptr, n, dvi.as<double>(),
                                   daux.as<double>(), dsum.as<double>());
                ESPEC_HIP_TRY(hipGetLastError());
                ESPEC_HIP_TRY(hipMemcpy(&sum, dsum.p, sizeof(double), hipMemcpyDeviceToHost));
            }
            double wk = 0;
            ESPEC_HIP_TRY(hipMemcpy(&wk, dw.as<double>() + k, sizeof(double),
                                    hipMemcpyDeviceToHost));
            wstate[k] = wk;
            if (log && nl < logcap) {
                double *r = log + 4 * (size_t)nl;
                r[0] = k + 1;
                r[1] = kk;
                r[2] = wk;
                r[3] = sum;
            }
            ++nl;
        } while (sum > abstol && kk <= maxint);
        if (nlog) *nlog = nl;
        if (kk >= maxint) return ESPEC_ERR_DIAG;  // "didn't achieve the convergence"
        // EIGVC(:,K) = VI / ECNORM(N, VI) (lanczsc.f:128-131)
        hipLaunchKernelGGL(normalise, dim3(1), dim3(1024), 0, nullptr, n, dvi.as<double>());
        ESPEC_HIP_TRY(hipGetLastError());
        ESPEC_HIP_TRY(hipMemcpy(eigvc + (size_t)k * n, dvi.p, sizeof(double) * n,
                                hipMemcpyDeviceToHost));
    }
    if (nlog) *nlog = nl;
    if (!ref) {
        // each state's value from the restart that converged it
        for (int k = 0; k < ns; ++k) eigvl[k] = wstate[k];
        return ESPEC_OK;
    }
    // EIGVL as the last DSTEVX left it, and DSTEVX writes EIGVC(1:40, 1:M)
    // on every restart, so the columns of earlier states end with rows 1..40
    // replaced by the last restart's tridiagonal eigenvectors.
    ESPEC_HIP_TRY(hipMemcpy(eigvl, dw.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
    for (int k = 0; k + 1 < ns; ++k)
        ESPEC_HIP_TRY(hipMemcpy(eigvc + (size_t)k * n, dcoef.as<double>() + (size_t)k * m,
                                sizeof(double) * m, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_lanczs(int ndim, const int *np, int il, int iu, int nreort, double abstol,
                 const double *shm, const double *vpot, double *eigvl, double *eigvc)
{
    LnzArgs l{};
    if (nreort < 1 || nreort > kMaxReort || !vpot || !eigvl || !eigvc ||
        !set_grid(l, ndim, np, shm))
        return ESPEC_ERR_ARG;
    const int n = l.n, ns = iu - il + 1;
    if (il < 1 || iu > n || ns < 1 || ns > 64) return ESPEC_ERR_ARG;
    if (nreort > n) nreort = n;
    // RAN start vector, NSEED = 1 (lanczs.f:98-101, rans.f:32-39).
    std::vector<double> start(n);
    long long seed = 1;
    for (int i = 0; i < n; ++i) {
        seed = (seed * 1029 + 221591) % 1048576;
        start[i] = (double)seed / 1048576.0;
    }
    DevBuf dv, ds, dacc, dcoef, de2, dw, dwork, dinfo;
    ESPEC_HIP_TRY(dv.upload(vpot, sizeof(double) * n));
    ESPEC_HIP_TRY(ds.upload(start.data(), sizeof(double) * n));
    l.vpot = dv.as<double>();
    l.vshift = 0;
    l.nvec = 1;
    l.m = n;
    l.nb = blocks_for(n);
    l.ncols = nreort;
    l.nchunk = (nreort + kChunk - 1) / kChunk;
    l.nreort = nreort;
    LnzBufs bufs;
    ESPEC_HIP_TRY(bufs.alloc(l, sizeof(double)));
    // 1) n-step recurrence for the tridiagonal (lanczs.f:102-104).
    hipLaunchKernelGGL((lnz_init<double, 256>), dim3(l.nb, 1), dim3(256), 0, nullptr, l,
                       (const double2 *)nullptr, ds.as<double>());
    ESPEC_HIP_TRY(lnz_iterations<double>(l, nullptr));
    // 2) selected eigenpairs of the tridiagonal (lanczs.f:105-108).
    ESPEC_HIP_TRY(dcoef.alloc(sizeof(double) * (size_t)ns * n));
    ESPEC_HIP_TRY(de2.alloc(sizeof(double) * n));
    ESPEC_HIP_TRY(dw.alloc(sizeof(double) * ns));
    ESPEC_HIP_TRY(dwork.alloc(sizeof(double) * (size_t)ns * 6 * n));
    ESPEC_HIP_TRY(dinfo.alloc(sizeof(int)));
    ESPEC_HIP_TRY(hipMemset(dinfo.p, 0, sizeof(int)));
    hipLaunchKernelGGL(lanczs_select, dim3(1), dim3(64), 0, nullptr, n, l.al, l.bt,
                       de2.as<double>(), il, ns, abstol, dw.as<double>(), dcoef.as<double>(),
                       dwork.as<double>(), dinfo.as<int>());
    ESPEC_HIP_TRY(hipGetLastError());
    // 3) replay the same recurrence, accumulating sum_J q_k[J] v_J for every
    //    selected k in one pass (lanczs.f:109-176 replays once per k).
    ESPEC_HIP_TRY(dacc.alloc(sizeof(double) * (size_t)ns * n));
    ESPEC_HIP_TRY(hipMemset(dacc.p, 0, dacc.bytes));
    ESPEC_HIP_TRY(hipMemset(bufs.basis.p, 0, bufs.basis.bytes));
    LnzArgs r = l;
    r.acoef = dcoef.as<double>();
    r.acc = dacc.as<double>();
    r.nacc = ns;
    hipLaunchKernelGGL((lnz_init<double, 256>), dim3(r.nb, 1), dim3(256), 0, nullptr, r,
                       (const double2 *)nullptr, ds.as<double>());
    ESPEC_HIP_TRY(lnz_iterations<double>(r, nullptr));
    int info = 0;
    ESPEC_HIP_TRY(hipMemcpy(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(eigvl, dw.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(eigvc, dacc.p, sizeof(double) * (size_t)ns * n, hipMemcpyDeviceToHost));
    return info ? ESPEC_ERR_DIAG : ESPEC_OK;
}

}  // extern "C"
