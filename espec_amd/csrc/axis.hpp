// Host-side setup of device FFT axis descriptors (DevAxis) from a stage
// plan and a line geometry.
#pragma once
#include "api_common.hpp"
#include "common.hpp"

namespace espec {

inline void fill_axis(DevAxis &d, const FftPlan &p, int es, int nl, int lmod, int la, int lb,
                      int tw0)
{
    d.n = p.n;
    d.ln = ilog2(p.n);
    d.nst = p.nst;
    set_line_geometry(d, es, nl, lmod, la, lb);
    d.tw0 = tw0;
    for (int s = 0; s < p.nst; ++s) {
        d.rad[s] = p.rad[s];
        d.lm[s] = p.lm[s];
        d.lL[s] = p.lL[s];
        d.off[s] = p.off[s];
    }
}

// Line geometry of axis a of a whole ndim grid held in LDS (x fastest).
inline void axis_geometry(int ndim, const int *np, int a, int &es, int &nl, int &lmod, int &la,
                          int &lb)
{
    int n = ntot(ndim, np);
    es = 1;
    for (int k = 0; k < a; ++k) es *= np[k];
    nl = n / np[a];
    lmod = es;          // lines inside one slab of the lower axes
    la = 1;
    lb = es * np[a];    // jump over the transformed axis
}

// Twiddles of all axes as interleaved (cos, sin) with each axis' offset
// recorded in its DevAxis, and the digit-reversal maps.
inline void axis_tables(int ndim, const int *np, bool ref, DevAxis *ax, std::vector<FftPlan> &fp,
                        std::vector<double> &tw, std::vector<int> &perm, int *perm_off)
{
    fp.clear();
    tw.clear();
    perm.clear();
    for (int a = 0; a < ndim; ++a) {
        fp.push_back(make_fft_plan(np[a], ref));
        int es, nl, lmod, la, lb;
        axis_geometry(ndim, np, a, es, nl, lmod, la, lb);
        fill_axis(ax[a], fp[a], es, nl, lmod, la, lb, (int)(tw.size() / 2));
        for (size_t i = 0; i < fp[a].twc.size(); ++i) {
            tw.push_back(fp[a].twc[i]);
            tw.push_back(fp[a].tws[i]);
        }
        perm_off[a] = (int)perm.size();
        perm.insert(perm.end(), fp[a].perm.begin(), fp[a].perm.end());
    }
    if (tw.empty()) tw.assign(2, 0.0);
}

// Natural-order table t[k] placed where the fused kernels find frequency
// k after the forward DIF stages of every axis: position sum_a perm_a[k_a] * stride_a.
inline std::vector<double> scramble(int ndim, const int *np, const std::vector<FftPlan> &fp,
                                    const std::vector<double> &t)
{
    const int n = ntot(ndim, np);
    std::vector<double> s(n);
    for (int k = 0; k < n; ++k) {
        int r = k, pos = 0, stride = 1;
        for (int a = 0; a < ndim; ++a) {
            pos += fp[a].perm[r % np[a]] * stride;
            r /= np[a];
            stride *= np[a];
        }
        s[pos] = t[k];
    }
    return s;
}

}  // namespace espec
