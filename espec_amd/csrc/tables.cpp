#include "tables.hpp"

#include <cmath>

namespace espec {

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

std::vector<int> singleton_factors(int n)
{
    std::vector<int> sq, f;
    int k = n, mid = 0;
    while (k % 16 == 0) {
        sq.push_back(4);
        k /= 16;
    }
    if (k > 4) {  // 8 left: one more square factor 2, then 2 in the middle
        sq.push_back(2);
        mid = k / 4;
    } else if (k != 1) {
        mid = k;
    }
    f = sq;
    if (mid) f.push_back(mid);
    for (auto it = sq.rbegin(); it != sq.rend(); ++it) f.push_back(*it);
    return f;
}

FftPlan make_fft_plan(int n, bool ref)
{
    FftPlan p;
    p.n = n;
    std::vector<int> fac = singleton_factors(n);
    p.nst = (int)fac.size();
    // The default-REAL literal of fft3d.f:119 rounds to f32.
    const double rad = ref ? (double)(float)6.2831853071796 : kTwoPi;
    int L = n, off = 0;
    for (int s = 0; s < p.nst; ++s) {
        int r = fac[s], m = L / r;
        p.rad[s] = r;
        p.lm[s] = ilog2(m);
        p.lL[s] = ilog2(L);
        p.off[s] = off;
        const double phi = rad / (double)L;
        for (int q = 1; q < r; ++q)
            for (int j = 0; j < m; ++j) {
                double c, sn;
                if (r == 2 && ref && 2 * j > m) {
                    c = -std::cos((double)(m - j) * phi);
                    sn = std::sin((double)(m - j) * phi);
                } else {
                    c = std::cos((double)q * j * phi);
                    sn = std::sin((double)q * j * phi);
                }
                p.twc.push_back(c);
                p.tws.push_back(sn);
            }
        off += (r - 1) * m;
        L = m;
    }
    p.perm.resize(n);
    for (int k = 0; k < n; ++k) {
        int kk = k, span = n, pos = 0;
        for (int s = 0; s < p.nst; ++s) {
            span /= fac[s];
            pos += (kk % fac[s]) * span;
            kk /= fac[s];
        }
        p.perm[k] = pos;
    }
    return p;
}

std::vector<double> kinetic_axis(int n, double sh, double sm, bool ref)
{
    std::vector<double> t(n);
    if (ref) {
        const double dp = 6.283185307 / (n * sh);
        double p1 = 0.0;
        for (int i = 0; i < n / 2; ++i) {
            t[i] = p1 * p1 / (2.0 * sm);
            p1 = p1 + dp;
        }
        p1 = -p1;
        for (int i = n / 2; i < n; ++i) {
            t[i] = p1 * p1 / (2.0 * sm);
            p1 = p1 + dp;
        }
    } else {
        const double dp = kTwoPi / (n * sh);
        for (int i = 0; i < n; ++i) {
            double p = (double)(i < n / 2 ? i : i - n) * dp;
            t[i] = p * p / (2.0 * sm);
        }
    }
    return t;
}

std::vector<double> kinetic_grid(int ndim, const int *np, const double *sh,
                                 const double *sm, bool ref)
{
    int n = 1;
    for (int a = 0; a < ndim; ++a) n *= np[a];
    std::vector<std::vector<double>> ax(ndim);
    for (int a = 0; a < ndim; ++a) ax[a] = kinetic_axis(np[a], sh[a], sm[a], ref);
    std::vector<double> tc(n);
    for (int i = 0; i < n; ++i) {
        int r = i;
        double t = 0;
        for (int a = 0; a < ndim; ++a) {
            t += ax[a][r % np[a]];
            r /= np[a];
        }
        tc[i] = t;
    }
    return tc;
}

int fortran_trip(double m1, double m2, double m3)
{
    double c = (m2 - m1 + m3) / m3;
    int k = (int)c;
    return k > 0 ? k : 0;
}

}  // namespace espec
