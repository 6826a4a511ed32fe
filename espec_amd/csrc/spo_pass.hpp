// Multi-pass split-operator propagation for grids that do not fit in one
// workgroup's LDS (256^2 complex64, 128^3): one pass per axis per step.
#pragma once
#include <vector>

#include "common.hpp"

namespace espec {

struct SpoPassPlan {
    int ndim = 0, np[3] = {1, 1, 1}, n = 0, npk = 0, mixed = 0, maxsamp = 0;
    int rows_per_wg = 1;   // x-lines per workgroup of the row pass
    int tile = 16;         // contiguous elements per strided-pass tile
    int xblocks = 0;       // row-pass workgroups per packet
    DevAxis axx{};         // LDS geometry of a row-pass block
    DevAxis axs[3]{};      // LDS geometry of strided-pass tiles (axes 1, 2)
    void *x = nullptr;
    const void *psi0 = nullptr, *vpot = nullptr;  // storage precision
    double *cbuf = nullptr;
    void *tw = nullptr;    // double2 twiddles, all axes
    void *kax = nullptr;   // double2 per-axis kinetic phases in digit-reversed order
    int kax_off[3] = {};
    double *part = nullptr;   // [maxsamp][npk][xblocks][3] row-pass partial sums
    long long *dstep = nullptr;  // steps completed, advanced on the device
    double b = 1, cst2 = 0;
    int lc = 1;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
};

int spo_pass_setup(SpoPassPlan &p, int ndim, const int *np, int npk, bool mixed,
                   const std::vector<FftPlan> &fp, const std::vector<std::vector<double>> &taxes,
                   double cst1, double cst2, void *x, const void *psi0, const void *vpot,
                   double *cbuf, int maxsamp);
int spo_pass_run(SpoPassPlan &p, int nsteps, long long step0, int lc, hipStream_t s,
                 int *launches);
void spo_pass_release(SpoPassPlan &p);

}  // namespace espec
