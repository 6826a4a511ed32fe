// Multi-pass split-operator propagation for grids that do not fit in one
// workgroup's LDS (256^2, 128^3): one pass per axis per step.
#pragma once
#include <vector>

#include "common.hpp"

namespace espec {

// Launch shape of the passes (chosen by the autotuner).
struct PassShape {
    int rows = 1, rows_nt = 64;    // x lines per row-pass workgroup, its threads
    int tile = 16, tile_nt = 256;  // contiguous elements per strided tile, its threads
    int spread = 1;                // active workgroups at linear ids 0, spread, 2 spread, ...
    int xmap = 1;                  // contiguous strided tiles per XCD
    int grid = 0;                  // workgroups per pass (0: one per work item)
    int rows_pen = 0;              // row pass on register pencils (rows = 64 / (np0 / 16)),
                                   // 2: one 256-point line per wavefront (rows = 1),
                                   // 3: as 1 with FP32 transforms (complex64 storage)
    int tile_pen = 0;              // strided passes on register pencils: 1 one wavefront per
                                   // item, 2 four wavefronts sharing an LDS-staged tile,
                                   // 3 one 256-point column per wavefront
    int pen_pf = 0;                // pencil strided passes prefetch their next work item
    int fused = 0;                 // 4 or 8: 2-D grids that fit one workgroup's LDS run
                                   // whole launches of steps LDS-resident (pass_fused2d),
                                   // one packet per workgroup of that many wavefronts
};

struct SpoPassPlan {
    int ndim = 0, np[3] = {1, 1, 1}, n = 0, npk = 0, mixed = 0, maxsamp = 0;
    PassShape shape;
    int xblocks = 0;       // row-pass workgroups per packet
    DevAxis axx{};         // LDS geometry of a row-pass block (twiddles at 0)
    DevAxis axs[3]{};      // LDS geometry of strided-pass tiles (axes 1, 2)
    FftPlan fpl[3];        // host copies, to rebuild geometry when the shape changes
    void *x = nullptr;
    const void *psi0 = nullptr, *vpot = nullptr;  // storage precision
    const void *vph = nullptr;  // FP64 V phase table, or null for V + sincos
    const void *vph_fz = nullptr;  // the same table for the fused form only, where the
                                   // multi-pass kernels form the phase from V
    const double *mask = nullptr;  // .SMOOTHW mask [npk][n] (forces the LDS-stage row pass)
    double *cbuf = nullptr;
    void *tw = nullptr;    // double2 twiddles, axis a at tw_off[a], ntw[a] entries
    int tw_off[3] = {}, ntw[3] = {};
    void *kax = nullptr;   // double2 per-axis kinetic phases in digit-reversed order
    void *twr = nullptr;   // double2 exp(2 pi i m / np[a]), axis a at twr_off[a] (pencils)
    int twr_off[3] = {};
    void *kreg = nullptr;  // as kax, in the pencils' spectral-layout order
    void *twr_s = nullptr;   // split float4 twiddles, 2 np[a] per axis at 2 twr_off[a]
    void *kreg_s = nullptr;  // split float4 copy of kreg
    void *knat = nullptr;    // as kax, in natural frequency order (wide lines)
    int kax_off[3] = {};   // offset of axis a within one packet's set
    int kax_pstride = 0;   // entries between packets' sets (0: one shared set)
    double *part = nullptr;      // [maxsamp][npk][max xblocks][3] row-pass partial sums
    long long *dstep = nullptr;  // steps completed, advanced on the device
    double b = 1, cst2 = 0;
    double phase_max = 0;  // max |cst2 V| over the batch (range of the pencils' sincos)
    int lc = 1;
    // Step graphs of 2^j full steps, j = 0..kGraphLevels-1 (captured by
    // spo_pass_prepare); a run replays them largest first.
    static constexpr int kGraphLevels = 6;
    hipGraphExec_t gexec[kGraphLevels] = {};
    hipGraph_t graph[kGraphLevels] = {};
};

int spo_pass_setup(SpoPassPlan &p, int ndim, const int *np, int npk, bool mixed,
                   const std::vector<FftPlan> &fp, const std::vector<std::vector<double>> &taxes,
                   double cst1, double cst2, void *x, const void *psi0, const void *vpot,
                   const void *vph, const void *vph_fz, const double *mask, double *cbuf,
                   int maxsamp, double phase_max, int tune, hipStream_t s);
// True if the grid has an LDS-resident fused form (PassShape::fused).
bool spo_pass_fused_grid(int ndim, const int *np, bool mixed);
int spo_pass_run(SpoPassPlan &p, int nsteps, long long step0, int lc, hipStream_t s,
                 int *launches);
// spo_pass_run's kernels launched one by one on s, for capture into a
// caller's graph.
int spo_pass_enqueue(SpoPassPlan &p, int nsteps, long long step0, int lc, hipStream_t s);
int spo_pass_profile(SpoPassPlan &p, int nsteps, long long step0, int lc, hipStream_t s, double *ms,
                     double *bytes);
void spo_pass_release(SpoPassPlan &p);
// Replace the tuned launch shape (ESPEC_ERR_ARG if the grid cannot use it).
int spo_pass_set_shape(SpoPassPlan &p, const PassShape &s);
// Capture the per-step launch graphs for sampling interval lc ahead of the
// first run (a run captures them itself when needed).  With prime (only
// while psi equals psi0), replay each once on s and restore psi from
// psi0: a graph's first replay costs more than the next ones, and a short
// timed run should not pay for it.
int spo_pass_prepare(SpoPassPlan &p, int lc, hipStream_t s, bool prime);

}  // namespace espec
