// Split-operator plans and the PSPOFFT / DIGT entry points.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "api_common.hpp"
#include "axis.hpp"
#include "spo_lds.hpp"
#include "spo_pass.hpp"

using namespace espec;

struct espec_spo_plan {
    int ndim = 1, np[3] = {1, 1, 1}, n = 0, npk = 0, lc = 1, maxsamp = 0, flags = 0;
    bool per_packet_mass = false;
    int nt = 64;
    bool mixed = false, ref = false, lds = true;
    double dt = 0, b = 1;
    long long steps_done = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf x, psi0, vph, kin, tw, perm, cbuf;
    DevBuf vst;  // multi-pass: the potential in storage precision
    int absorb = ESPEC_ABSORB_NONE;
    DevBuf mask;  // .SMOOTHW: FP64 [npk][n] mask applied after the closing V half phase
    SpoLdsArgs la{};
    SpoPassPlan pp{};  // multi-pass plan for grids that do not fit in LDS
    float last_ms = 0;
    int last_launches = 0;

    ~espec_spo_plan()
    {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
        spo_pass_release(pp);
    }
};

namespace {

// S is the storage precision of psi (float: ESPEC_MIXED); every table the
// kernels multiply by (twiddles, V and T phases) is FP64.
template <typename S>
int build_tables(espec_spo_plan &P, const double *sh, const double *sm, const double *vpot,
                 const double *u, const double *v, const double *vabc)
{
    const int n = P.n, npk = P.npk;
    const double cst1 = P.dt / kShfs, cst2 = P.dt / (2.0 * kShfs);
    // V phase tables of all packets; .VOPTIC scales each entry by the
    // damping exp(cst2 VABC) of its point, as the reference scales DCOS
    // and DSIN (spofft.f:68-71, 100-103), so both half phases absorb.
    auto vphase = [&]() {
        std::vector<double> vph((size_t)2 * n * npk);
        for (int p = 0; p < npk; ++p) {
            std::vector<double> t = phase_table<double>(vpot + (size_t)p * n, n, cst2);
            if (P.absorb == ESPEC_ABSORB_VOPTIC)
                for (int i = 0; i < n; ++i) {
                    const double e = std::exp(cst2 * vabc[(size_t)p * n + i]);
                    t[2 * i] *= e;
                    t[2 * i + 1] *= e;
                }
            std::memcpy(vph.data() + (size_t)2 * n * p, t.data(), sizeof(double) * 2 * n);
        }
        return vph;
    };
    // .SMOOTHW: the mask multiplies psi after the closing half phase
    // (spofft.f:162-176).
    if (P.absorb == ESPEC_ABSORB_SMOOTHW)
        ESPEC_HIP_TRY(P.mask.upload(vabc, sizeof(double) * (size_t)n * npk));
    std::vector<FftPlan> fp;
    std::vector<double> tw;
    std::vector<int> perm;
    axis_tables(P.ndim, P.np, P.ref, P.la.ax, fp, tw, perm, P.la.perm_off);
    std::vector<S> xs((size_t)2 * n * npk);
    for (int p = 0; p < npk; ++p) {
        std::vector<S> t = interleave<S>(u + (size_t)p * n, v ? v + (size_t)p * n : nullptr, n);
        std::memcpy(xs.data() + (size_t)2 * n * p, t.data(), sizeof(S) * 2 * n);
    }
    ESPEC_HIP_TRY(P.x.upload(xs.data(), sizeof(S) * xs.size()));
    ESPEC_HIP_TRY(P.psi0.upload(xs.data(), sizeof(S) * xs.size()));
    ESPEC_HIP_TRY(P.cbuf.alloc(sizeof(double) * 3 * (size_t)P.maxsamp * npk));
    ESPEC_HIP_TRY(hipMemset(P.cbuf.p, 0, P.cbuf.bytes));
    if (!P.lds) {
        // Multi-pass: the potential itself in storage precision; the
        // kernels form exp(-i cst2 V) in FP64.
        std::vector<S> vs((size_t)n * npk);
        for (size_t i = 0; i < vs.size(); ++i) vs[i] = (S)vpot[i];
        ESPEC_HIP_TRY(P.vst.upload(vs.data(), sizeof(S) * vs.size()));
        // FP64 phase table while it stays small enough to live in the MALL
        // (then the pass is latency-bound and a table load beats sincos),
        // and always with double-precision storage: there the row pass
        // cannot hold V, psi0 and the state across the transform without
        // spilling, and the table removes V and the sincos from it.
        // .VOPTIC always uses the table: its damping rides on it.
        const bool table = !P.mixed || P.absorb == ESPEC_ABSORB_VOPTIC ||
                           (size_t)n * npk * sizeof(double2) <= (size_t)64 << 20;
        // The LDS-resident fused form moves no state through HBM within a
        // launch, so there a table load beats the sincos at any batch size:
        // grids with that form get a table for it alone.
        const bool fz_table = !table && spo_pass_fused_grid(P.ndim, P.np, P.mixed);
        if (table || fz_table) {
            std::vector<double> vph = vphase();
            ESPEC_HIP_TRY(P.vph.upload(vph.data(), sizeof(double) * vph.size()));
        }
        std::vector<std::vector<double>> taxes;
        const int nset = P.per_packet_mass ? npk : 1;
        for (int q = 0; q < nset; ++q)
            for (int a = 0; a < P.ndim; ++a)
                taxes.push_back(kinetic_axis(P.np[a], sh[a], sm[(size_t)q * P.ndim + a], false));
        double vmax = 0;
        for (size_t i = 0; i < (size_t)n * npk; ++i) vmax = std::max(vmax, std::fabs(vpot[i]));
        return spo_pass_setup(P.pp, P.ndim, P.np, npk, P.mixed, fp, taxes, cst1, cst2, P.x.p,
                              P.psi0.p, P.vst.p, table ? P.vph.p : nullptr,
                              fz_table ? P.vph.p : nullptr, P.mask.as<double>(),
                              P.cbuf.as<double>(), P.maxsamp, std::fabs(cst2) * vmax, 1,
                              P.stream);
    }
    // Kinetic phase: natural order for the reference kernel; for the fused
    // kernel, position of frequency k is sum_a perm_a[k_a] * stride_a.
    const int nset = P.per_packet_mass ? npk : 1;
    std::vector<double> kin;
    for (int q = 0; q < nset; ++q) {
        std::vector<double> tc =
            kinetic_grid(P.ndim, P.np, sh, sm + (size_t)q * P.ndim, P.ref);
        std::vector<double> tcs = P.ref ? tc : scramble(P.ndim, P.np, fp, tc);
        std::vector<double> k = phase_table<double>(tcs.data(), n, cst1);
        kin.insert(kin.end(), k.begin(), k.end());
    }
    P.la.kin_pp = nset > 1 ? 1 : 0;
    P.la.mask = P.mask.as<double>();
    std::vector<double> vph = vphase();
    ESPEC_HIP_TRY(P.vph.upload(vph.data(), sizeof(double) * vph.size()));
    ESPEC_HIP_TRY(P.kin.upload(kin.data(), sizeof(double) * kin.size()));
    ESPEC_HIP_TRY(P.tw.upload(tw.data(), sizeof(double) * tw.size()));
    ESPEC_HIP_TRY(P.perm.upload(perm.data(), sizeof(int) * perm.size()));
    return ESPEC_OK;
}

}  // namespace

extern "C" {

const char *espec_strerror(int code)
{
    switch (code) {
    case ESPEC_OK: return "ok";
    case ESPEC_DRIFT: return "energy or norm drift above TOL";
    case ESPEC_ERR_ARG: return "invalid argument";
    case ESPEC_ERR_HIP: return "HIP runtime error";
    case ESPEC_ERR_NOMEM: return "device allocation failed";
    case ESPEC_ERR_DIAG: return "tridiagonal eigensolve failed";
    case ESPEC_ERR_UNSUPPORTED: return "unsupported on the device";
    default: return "unknown error";
    }
}

int espec_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

const char *espec_version(void) { return "espec-mi355x 0.1 (gfx950)"; }

int espec_digt(int ndim, const int *np, const double *sh, const double *sm, double *tc,
               double *b, int flags)
{
    if (ndim < 1 || ndim > 3 || !np || !sh || !sm || !tc) return ESPEC_ERR_ARG;
    std::vector<double> t = kinetic_grid(ndim, np, sh, sm, (flags & ESPEC_REF) != 0);
    std::memcpy(tc, t.data(), sizeof(double) * t.size());
    if (b) *b = 1.0 / ntot(ndim, np);
    return ESPEC_OK;
}

int espec_spo_plan_create(espec_spo_plan **out, int ndim, const int *np, int npk, double dt,
                          int lc, const double *sh, const double *sm, const double *vpot,
                          const double *u, const double *v, int maxsamp, int flags)
{
    return espec_spo_plan_create_absorb(out, ndim, np, npk, dt, lc, sh, sm, vpot, u, v, maxsamp,
                                        flags, ESPEC_ABSORB_NONE, nullptr);
}

int espec_spo_plan_create_absorb(espec_spo_plan **out, int ndim, const int *np, int npk,
                                 double dt, int lc, const double *sh, const double *sm,
                                 const double *vpot, const double *u, const double *v,
                                 int maxsamp, int flags, int absorb, const double *vabc)
{
    if (!out || ndim < 1 || ndim > 3 || !np || npk < 1 || lc < 1 || !sh || !sm || !vpot || !u ||
        maxsamp < 1 || absorb < ESPEC_ABSORB_NONE || absorb > ESPEC_ABSORB_SMOOTHW ||
        (absorb != ESPEC_ABSORB_NONE && !vabc))
        return ESPEC_ERR_ARG;
    *out = nullptr;
    for (int a = 0; a < ndim; ++a)
        if (!is_pow2(np[a]) || np[a] < 8 || np[a] > 65536) return ESPEC_ERR_ARG;
    std::unique_ptr<espec_spo_plan> P(new espec_spo_plan);
    P->ndim = ndim;
    for (int a = 0; a < ndim; ++a) P->np[a] = np[a];
    P->n = ntot(ndim, np);
    P->npk = npk;
    P->dt = dt;
    P->lc = lc;
    P->maxsamp = maxsamp;
    P->flags = flags;
    P->mixed = (flags & ESPEC_MIXED) != 0;
    P->ref = (flags & ESPEC_REF) != 0;
    P->per_packet_mass = (flags & ESPEC_PER_PACKET_MASS) != 0;
    P->absorb = absorb;
    if (P->ref && P->mixed) return ESPEC_ERR_UNSUPPORTED;
    P->b = 1.0 / P->n;
    // One wavefront per packet up to 256 points; four above (a 1024-point
    // 1-D batch of 260 packets: 15.4 -> 7.9 us per step).
    P->nt = P->n <= 256 ? 64 : 256;
    // LDS-resident when the FP64 grid fits in one workgroup's LDS; the V
    // phase table is staged next to it when that fits too.
    const size_t lds_cap = 160 * 1024 - 1024;
    P->lds = !(flags & ESPEC_MULTIPASS) && spo_lds_bytes(P->n, P->ref, false, P->nt) <= lds_cap;
    P->la.vph_lds = (!P->ref && spo_lds_bytes(P->n, false, true, P->nt) <= lds_cap) ? 1 : 0;
    {
        int ntw = 0;
        for (int a = 0; a < ndim; ++a) ntw += (int)make_fft_plan(np[a], P->ref).twc.size();
        P->la.ntw = ntw;
        P->la.tab_lds = (P->la.vph_lds &&
                         spo_lds_bytes(P->n, false, true, P->nt, true, ntw) <= lds_cap) ? 1 : 0;
    }
    if (!P->lds && P->ref) return ESPEC_ERR_UNSUPPORTED;
    ESPEC_HIP_TRY(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    ESPEC_HIP_TRY(hipEventCreate(&P->ev0));
    ESPEC_HIP_TRY(hipEventCreate(&P->ev1));
    int rc = P->mixed ? build_tables<float>(*P, sh, sm, vpot, u, v, vabc)
                      : build_tables<double>(*P, sh, sm, vpot, u, v, vabc);
    if (rc) return rc;
    if (!P->lds) {
        rc = spo_pass_prepare(P->pp, lc, P->stream, true);
        if (rc) return rc;
    }
    SpoLdsArgs &a = P->la;
    a.ndim = ndim;
    a.n = P->n;
    a.npk = npk;
    a.tw = P->tw.p;
    a.perm = P->perm.as<int>();
    a.x = P->x.p;
    a.psi0 = P->psi0.p;
    a.vph = P->vph.p;
    a.kin = P->kin.p;
    a.b = P->b;
    a.lc = lc;
    a.maxsamp = maxsamp;
    a.cbuf = P->cbuf.as<double>();
    // C(TI): sample 0 (pspofft.f:86-90).
    ESPEC_HIP_TRY(launch_spo_sample(P->x.p, P->psi0.p, P->n, npk, P->mixed, a.cbuf, maxsamp, 0,
                                    P->stream));
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    *out = P.release();
    return ESPEC_OK;
}

int espec_spo_plan_run(espec_spo_plan *P, int nsteps)
{
    if (!P || nsteps < 0) return ESPEC_ERR_ARG;
    if (nsteps == 0) return ESPEC_OK;
    ESPEC_HIP_TRY(hipEventRecord(P->ev0, P->stream));
    int launches = 0;
    if (P->lds) {
        // Keep single launches short: about 2^24 point-steps per packet.
        const int chunk = std::max(1, (1 << 24) / P->n);
        int done = 0;
        while (done < nsteps) {
            int k = std::min(chunk, nsteps - done);
            P->la.nsteps = k;
            P->la.step0 = P->steps_done + done;
            ESPEC_HIP_TRY(launch_spo_lds(P->la, P->mixed, P->ref, P->nt, P->stream));
            ++launches;
            done += k;
        }
    } else {
        int rc = spo_pass_run(P->pp, nsteps, P->steps_done, P->lc, P->stream, &launches);
        if (rc) return rc;
    }
    ESPEC_HIP_TRY(hipEventRecord(P->ev1, P->stream));
    P->steps_done += nsteps;
    P->last_launches = launches;
    return ESPEC_OK;
}

int espec_spo_plan_sync(espec_spo_plan *P)
{
    if (!P) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    float ms = 0;
    if (hipEventElapsedTime(&ms, P->ev0, P->ev1) == hipSuccess) P->last_ms = ms;
    return ESPEC_OK;
}

int espec_spo_plan_last_timing(espec_spo_plan *P, double *ms, int *launches)
{
    if (!P) return ESPEC_ERR_ARG;
    if (ms) *ms = P->last_ms;
    if (launches) *launches = P->last_launches;
    return ESPEC_OK;
}

int espec_spo_plan_get(espec_spo_plan *P, int pk, double *cr, double *ci, double *nrm,
                       int maxsamp, int *nsamp, double *u, double *v)
{
    if (!P || pk < 0 || pk >= P->npk) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
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
            ESPEC_HIP_TRY(hipMemcpy(h.data(), P->x.as<double>() + 2 * n * pk, sizeof(double) * 2 * n,
                                    hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) {
                if (u) u[i] = h[2 * i];
                if (v) v[i] = h[2 * i + 1];
            }
        }
    }
    return ESPEC_OK;
}

int espec_spo_plan_destroy(espec_spo_plan *P)
{
    if (P) {
        (void)hipStreamSynchronize(P->stream);
        delete P;
    }
    return ESPEC_OK;
}

int espec_spo_plan_reset(espec_spo_plan *P)
{
    if (!P) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipMemcpyAsync(P->x.p, P->psi0.p, P->x.bytes, hipMemcpyDeviceToDevice, P->stream));
    ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
    P->steps_done = 0;
    return ESPEC_OK;
}

int espec_spo_plan_shape(espec_spo_plan *P, int *shape)
{
    if (!P || !shape) return ESPEC_ERR_ARG;
    const PassShape &s = P->pp.shape;
    shape[0] = P->lds ? 1 : 0;
    shape[1] = s.rows;
    shape[2] = s.rows_nt;
    shape[3] = s.tile;
    shape[4] = s.tile_nt;
    shape[5] = s.spread;
    shape[6] = s.xmap;
    shape[7] = s.grid;
    shape[8] = s.rows_pen;
    shape[9] = s.tile_pen;
    shape[10] = s.pen_pf;
    shape[11] = P->lds ? 0 : (P->pp.vph != nullptr || (s.fused && P->pp.vph_fz != nullptr));
    shape[12] = P->lds ? 0 : s.fused;
    return ESPEC_OK;
}

int espec_spo_plan_set_shape(espec_spo_plan *P, const int *shape)
{
    if (!P || !shape) return ESPEC_ERR_ARG;
    if (P->lds) return ESPEC_ERR_UNSUPPORTED;  // one fixed kernel shape
    PassShape s;
    s.rows = shape[1];
    s.rows_nt = shape[2];
    s.tile = shape[3];
    s.tile_nt = shape[4];
    s.spread = shape[5];
    s.xmap = shape[6];
    s.grid = shape[7];
    s.rows_pen = shape[8];
    s.tile_pen = shape[9];
    s.pen_pf = shape[10];
    s.fused = shape[12];
    const int rc = spo_pass_set_shape(P->pp, s);
    return rc ? rc : spo_pass_prepare(P->pp, P->lc, P->stream, P->steps_done == 0);
}

int espec_spo_plan_profile(espec_spo_plan *P, int nsteps, double *ms, double *bytes, int maxkinds,
                           int *nkinds)
{
    if (!P || !ms || !bytes || !nkinds || maxkinds < 1 || nsteps < 2) return ESPEC_ERR_ARG;
    double m[2] = {0, 0}, b[2] = {0, 0};
    int kinds = 1;
    if (P->lds) {
        hipEvent_t e0, e1;
        ESPEC_HIP_TRY(hipEventCreate(&e0));
        ESPEC_HIP_TRY(hipEventCreate(&e1));
        P->la.nsteps = nsteps;
        P->la.step0 = P->steps_done;
        ESPEC_HIP_TRY(hipEventRecord(e0, P->stream));
        ESPEC_HIP_TRY(launch_spo_lds(P->la, P->mixed, P->ref, P->nt, P->stream));
        ESPEC_HIP_TRY(hipEventRecord(e1, P->stream));
        ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
        float t = 0;
        ESPEC_HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        P->steps_done += nsteps;
        // One launch reads psi and writes it back once, reads the V phase
        // table once and psi0 at every sample step.
        const double vs = P->mixed ? 8.0 : 16.0;
        m[0] = t;
        b[0] = (double)P->n * P->npk * (2 * vs + 16.0 + vs * ((double)nsteps / P->lc));
    } else {
        int rc = spo_pass_profile(P->pp, nsteps, P->steps_done, P->lc, P->stream, m, b);
        if (rc) return rc;
        P->steps_done += nsteps;
        kinds = (P->ndim > 1 && !P->pp.shape.fused) ? 2 : 1;
    }
    for (int k = 0; k < kinds && k < maxkinds; ++k) {
        ms[k] = m[k];
        bytes[k] = b[k];
    }
    *nkinds = kinds < maxkinds ? kinds : maxkinds;
    return ESPEC_OK;
}

}  // extern "C"

// A batch of split-operator plans run as one graph: each plan's whole run
// of nsteps (every kernel of it, launched one by one) on its own branch,
// forked from and joined to one capture stream, so that one hipGraphLaunch
// advances every packet of every plan.  The captured kernels carry each
// plan's step count at capture, so the graph replays that range of steps.
struct espec_spo_batch {
    std::vector<espec_spo_plan *> plans;
    std::vector<long long> step0;
    int nsteps = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr, t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> ready;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float last_ms = 0;
    ~espec_spo_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : ready) (void)hipEventDestroy(e);
        if (done) (void)hipEventDestroy(done);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Plan P's run of nsteps from step0, every kernel launched on s.
int plan_enqueue(espec_spo_plan *P, int nsteps, long long step0, hipStream_t s)
{
    if (!P->lds) return spo_pass_enqueue(P->pp, nsteps, step0, P->lc, s);
    const int chunk = std::max(1, (1 << 24) / P->n);
    SpoLdsArgs la = P->la;
    for (int done = 0; done < nsteps;) {
        const int k = std::min(chunk, nsteps - done);
        la.nsteps = k;
        la.step0 = step0 + done;
        ESPEC_HIP_TRY(launch_spo_lds(la, P->mixed, P->ref, P->nt, s));
        done += k;
    }
    return ESPEC_OK;
}

int batch_capture(espec_spo_batch &B)
{
    const int np = (int)B.plans.size();
    std::vector<hipStream_t> br(np, nullptr);
    std::vector<hipEvent_t> joins(np, nullptr);
    hipEvent_t fork = nullptr;
    hipStream_t cs = nullptr;
    int rc = ESPEC_OK;
    auto fail = [&](hipError_t e) {
        if (e != hipSuccess && rc == ESPEC_OK) rc = ESPEC_ERR_HIP;
        return e != hipSuccess;
    };
    if (!fail(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) &&
        !fail(hipEventCreateWithFlags(&fork, hipEventDisableTiming))) {
        for (int i = 0; i < np && rc == ESPEC_OK; ++i) {
            if (!fail(hipStreamCreateWithFlags(&br[i], hipStreamNonBlocking)))
                fail(hipEventCreateWithFlags(&joins[i], hipEventDisableTiming));
        }
    }
    if (rc == ESPEC_OK && !fail(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal))) {
        fail(hipEventRecord(fork, cs));
        for (int i = 0; i < np && rc == ESPEC_OK; ++i) {
            if (fail(hipStreamWaitEvent(br[i], fork, 0))) break;
            const int r = plan_enqueue(B.plans[i], B.nsteps, B.step0[i], br[i]);
            if (r) {
                rc = r;
                break;
            }
            if (fail(hipEventRecord(joins[i], br[i]))) break;
            fail(hipStreamWaitEvent(cs, joins[i], 0));
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc == ESPEC_OK && !fail(e)) {
            B.graph = g;
            fail(hipGraphInstantiate(&B.exec, g, nullptr, nullptr, 0));
        } else if (g) {
            (void)hipGraphDestroy(g);
        }
    }
    for (int i = 0; i < np; ++i) {
        if (joins[i]) (void)hipEventDestroy(joins[i]);
        if (br[i]) (void)hipStreamDestroy(br[i]);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (cs) (void)hipStreamDestroy(cs);
    return rc;
}

}  // namespace

extern "C" {

int espec_spo_batch_create(espec_spo_batch **out, espec_spo_plan *const *plans, int nplans,
                           int nsteps)
{
    if (!out || !plans || nplans < 1 || nsteps < 1) return ESPEC_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<espec_spo_batch> B(new espec_spo_batch);
    B->nsteps = nsteps;
    for (int i = 0; i < nplans; ++i) {
        espec_spo_plan *P = plans[i];
        if (!P) return ESPEC_ERR_ARG;
        for (espec_spo_plan *Q : B->plans)
            if (Q == P) return ESPEC_ERR_ARG;  // each plan once
        ESPEC_HIP_TRY(hipStreamSynchronize(P->stream));
        B->plans.push_back(P);
        B->step0.push_back(P->steps_done);
    }
    ESPEC_HIP_TRY(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    ESPEC_HIP_TRY(hipEventCreateWithFlags(&B->done, hipEventDisableTiming));
    ESPEC_HIP_TRY(hipEventCreate(&B->t0));
    ESPEC_HIP_TRY(hipEventCreate(&B->t1));
    for (int i = 0; i < nplans; ++i) {
        hipEvent_t e = nullptr;
        ESPEC_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        B->ready.push_back(e);
    }
    const int rc = batch_capture(*B);
    if (rc) return rc;
    *out = B.release();
    return ESPEC_OK;
}

int espec_spo_batch_run(espec_spo_batch *B)
{
    if (!B || !B->exec) return ESPEC_ERR_ARG;
    for (size_t i = 0; i < B->plans.size(); ++i)
        if (B->plans[i]->steps_done != B->step0[i]) return ESPEC_ERR_ARG;  // not the captured range
    // the graph starts after every plan's pending work and the plans' own
    // streams (get, sync, later runs) continue after the graph
    for (size_t i = 0; i < B->plans.size(); ++i) {
        ESPEC_HIP_TRY(hipEventRecord(B->ready[i], B->plans[i]->stream));
        ESPEC_HIP_TRY(hipStreamWaitEvent(B->stream, B->ready[i], 0));
    }
    ESPEC_HIP_TRY(hipEventRecord(B->t0, B->stream));
    ESPEC_HIP_TRY(hipGraphLaunch(B->exec, B->stream));
    ESPEC_HIP_TRY(hipEventRecord(B->t1, B->stream));
    ESPEC_HIP_TRY(hipEventRecord(B->done, B->stream));
    for (espec_spo_plan *P : B->plans) {
        ESPEC_HIP_TRY(hipStreamWaitEvent(P->stream, B->done, 0));
        P->steps_done += B->nsteps;
    }
    return ESPEC_OK;
}

int espec_spo_batch_sync(espec_spo_batch *B, double *ms)
{
    if (!B) return ESPEC_ERR_ARG;
    ESPEC_HIP_TRY(hipStreamSynchronize(B->stream));
    float t = 0;
    if (hipEventElapsedTime(&t, B->t0, B->t1) == hipSuccess) B->last_ms = t;
    if (ms) *ms = B->last_ms;
    return ESPEC_OK;
}

int espec_spo_batch_destroy(espec_spo_batch *B)
{
    delete B;
    return ESPEC_OK;
}

int espec_pspofft(int ndim, const int *np, double dt, double ti, double tf, int lc,
                  const double *sh, const double *sm, double *u1, double *v1, const double *vpot,
                  double *cr, double *ci, double *nrm, int maxsamp, int *nsamp, int flags)
{
    return espec_pspofft_absorb(ndim, np, dt, ti, tf, lc, sh, sm, u1, v1, vpot, cr, ci, nrm,
                                maxsamp, nsamp, flags, ESPEC_ABSORB_NONE, nullptr);
}

int espec_pspofft_absorb(int ndim, const int *np, double dt, double ti, double tf, int lc,
                         const double *sh, const double *sm, double *u1, double *v1,
                         const double *vpot, double *cr, double *ci, double *nrm, int maxsamp,
                         int *nsamp, int flags, int absorb, const double *vabc)
{
    if (!u1 || !v1 || !cr || !ci || maxsamp < 1) return ESPEC_ERR_ARG;
    // DO T=TI+DT,TF+DT,DT (pspofft.f:110).
    const int nsteps = fortran_trip(ti + dt, tf + dt, dt);
    espec_spo_plan *P = nullptr;
    int rc = espec_spo_plan_create_absorb(&P, ndim, np, 1, dt, lc, sh, sm, vpot, u1, v1, maxsamp,
                                          flags, absorb, vabc);
    if (rc) return rc;
    rc = espec_spo_plan_run(P, nsteps);
    if (!rc) rc = espec_spo_plan_get(P, 0, cr, ci, nrm, maxsamp, nsamp, u1, v1);
    espec_spo_plan_destroy(P);
    return rc;
}

}  // extern "C"
