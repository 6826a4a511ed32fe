"""espec_amd -- MI355X (gfx950) wave-packet propagation behind eSPec's interfaces.

Python host mirror of the C ABI in include/espec_hip.h.  Each function keeps
the name, argument meaning and error behaviour of the eSPec subroutine it
replaces (PSPOFFT, PSODFFT, PLNZ, FFT, AU, LNZ, DSTEVX, LANCZS, FCORREL,
HWHM, SPECTRUMTD, ...): arrays are split real/imaginary float64 vectors,
times in fs, and a nonzero INFO is raised as ``EspecError``.

The library is loaded from espec_amd/libespec_hip.so (built in-tree by
``python -m espec_amd.build``).  There is no CPU fallback: if the library or
a GPU is missing, calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libespec_hip.so")

FP64, MIXED = 0x0, 0x1
EXACT, REF = 0x0, 0x2
PER_PACKET_MASS = 0x4
MULTIPASS = 0x8
NO_GRAPH = 0x10  # Lanczos plans: launch each step's kernels directly (no graph replays)
# Absorbing boundaries (TPABSOR with ABSORB=.TRUE.; spofft.f:65-118, 162-176,
# sodfft.f:66-72, 108-116): VABC is an imaginary potential or a mask.
ABSORB_NONE, ABSORB_VOPTIC, ABSORB_SMOOTHW = 0, 1, 2
# Laser pulse shapes of EF (ef.f): EFC .NONE/.NULL, .GAUS, .GGAUS, .SIN2.
PULSE_NONE, PULSE_GAUS, PULSE_GGAUS, PULSE_SIN2 = 0, 1, 2, 3
# INIEIGVC forms of INIT_COND (init_cond.f): .CALC, .GETR/.GETR1, .GETR2,
# .GETC/.GETC1, .GETC2.
INIT_CALC, INIT_GETR, INIT_GETR2, INIT_GETC, INIT_GETC2 = 0, 1, 2, 3, 4
# Absorber shapes of ABSORBINGBC (absorb.f).
ABG_SMOOTHW, ABG_SMRS, ABG_VOPTIC_LINEAR, ABG_VOPTIC_KOSLOFF, ABG_VOPTIC_POLIN = 1, 2, 3, 4, 5
# Return codes of the C ABI (INFO of the drop-ins).
ESPEC_OK, ESPEC_ERR_ARG, ESPEC_ERR_HIP, ESPEC_ERR_NOMEM = 0, 2, 3, 4
ESPEC_ERR_DIAG, ESPEC_ERR_UNSUPPORTED = 5, 6
SHFS = 2.4188843243e-2  # fs per atomic unit of time (pspofft.f:76)
AMU = 1836.152667539    # electron masses per amu (espec.f:614)
A0A = 0.5291772083      # angstrom per bohr (espec.f:67)

_lib = None


class EspecError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        msg = _load().espec_strerror(code).decode() if _lib else str(code)
        super().__init__(f"{where}: INFO={code} ({msg})")


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} not built; run `python -m espec_amd.build`")
    L = C.CDLL(LIB_PATH)
    d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    vp, dp, ip = C.c_void_p, C.c_double, C.c_int
    P = C.POINTER
    L.espec_strerror.restype = C.c_char_p
    L.espec_strerror.argtypes = [ip]
    L.espec_version.restype = C.c_char_p
    L.espec_device_count.restype = ip
    prop = [ip, i, dp, dp, dp, ip, d, d, d, d, d, d, d, vp, ip, P(ip), ip]
    L.espec_pspofft.argtypes = prop
    L.espec_digt.argtypes = [ip, i, d, d, d, P(dp), ip]
    L.espec_spo_plan_create.argtypes = [P(vp), ip, i, ip, dp, ip, d, d, d, d, vp, ip, ip]
    L.espec_spo_plan_create_absorb.argtypes = L.espec_spo_plan_create.argtypes + [ip, vp]
    L.espec_pspofft_absorb.argtypes = prop + [ip, vp]
    L.espec_spo_plan_run.argtypes = [vp, ip]
    L.espec_spo_plan_sync.argtypes = [vp]
    L.espec_spo_plan_get.argtypes = [vp, ip, vp, vp, vp, ip, P(ip), vp, vp]
    L.espec_spo_plan_last_timing.argtypes = [vp, P(dp), P(ip)]
    L.espec_spo_plan_destroy.argtypes = [vp]
    L.espec_spo_plan_reset.argtypes = [vp]
    L.espec_spo_plan_profile.argtypes = [vp, ip, d, d, ip, P(ip)]
    L.espec_spo_plan_shape.argtypes = [vp, i]
    L.espec_spo_plan_set_shape.argtypes = [vp, i]
    f = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    fp_or_none = C.c_void_p
    L.espec_sil_plan_adaptive.argtypes = [vp, fp_or_none, ip, dp, ip, P(C.c_longlong)]
    L.espec_sil_plan_trace.argtypes = [vp, ip, ip, f, d]
    L.espec_sil_plan_adaptive_log.argtypes = [vp, C.c_void_p, ip, P(ip)]
    L.espec_sil_plan_reset.argtypes = [vp]
    L.espec_mlp_infer.argtypes = [f, f, ip, f]
    L.espec_mlp_grad.argtypes = [f, f, f, f, P(C.c_float)]
    L.espec_mlp_train.argtypes = [f, f, f, ip, ip, C.c_float, C.c_uint, f]
    L.espec_bw_probe.argtypes = [C.c_size_t, ip, P(dp), P(dp)]
    L.espec_plnz.argtypes = [ip, i, ip, dp, dp, dp, dp, ip, d, d, d, d, d, d, vp, ip, P(ip), ip]
    L.espec_plnz_absorb.argtypes = L.espec_plnz.argtypes + [ip, vp]
    L.espec_pspofftaux.argtypes = [ip, i, dp, dp, dp, d, d, d, d, ip, vp, ip]
    L.espec_psodfftaux.argtypes = [ip, i, dp, dp, d, d, d, d, d, d, ip, vp, ip]
    L.espec_reort.argtypes = [ip, ip, ip, d, d]
    L.espec_lanczsg.argtypes = [ip, i, ip, ip, ip, dp, d, d, d, d, vp, ip, P(ip), ip]
    L.espec_mtrxdiag.argtypes = [ip, i, ip, ip, dp, d, d, d, d, ip, P(ip)]
    L.espec_ppsod.argtypes = [ip, i, dp, dp, dp, ip, d, d, d, d, d, ip, dp, dp, dp, dp, dp, dp,
                              ip, ip, vp, ip, vp, vp, vp, vp, d, d, vp, vp, vp, vp, ip, P(ip),
                              ip]
    L.espec_s2ppsod.argtypes = [ip, i, dp, dp, dp, ip, d, dp, dp, dp, dp, d, d, d, d, ip, dp,
                                dp, dp, dp, dp, dp, ip, ip, vp, ip, dp, vp, ip, P(ip), P(ip),
                                vp, vp, vp, ip]
    L.espec_fftnt.argtypes = [d, d, ip, ip, ip, ip]
    L.espec_sil_plan_set_absorb.argtypes = [vp, ip, vp]
    L.espec_sil_plan_create.argtypes = [P(vp), ip, i, ip, ip, dp, dp, ip, d, d, d, vp, ip, ip]
    L.espec_sil_plan_set_pulse.argtypes = [vp, d, d, ip]
    L.espec_ef.argtypes = [ip, dp, dp, dp, dp, dp, dp, ip, dp]
    L.espec_init_cond.argtypes = [ip, i, ip, d, d, d, d, d, d, d]
    L.espec_absorbingbc.argtypes = [ip, i, ip, d, d, d, d, d, d, d, P(dp)]
    L.espec_ef.restype = dp
    L.espec_pplnz.argtypes = [ip, i, ip, dp, dp, dp, dp, ip, d, d, d, d, d, ip, dp, dp, dp, dp,
                              dp, dp, ip, vp, d, d, vp, ip, P(ip), ip]
    L.espec_sil_plan_run.argtypes = [vp, ip]
    L.espec_sil_plan_sync.argtypes = [vp]
    L.espec_sil_plan_get.argtypes = [vp, ip, vp, vp, vp, ip, P(ip), vp, vp]
    L.espec_sil_plan_destroy.argtypes = [vp]
    L.espec_psodfft.argtypes = prop
    L.espec_psodfft_absorb.argtypes = prop + [ip, vp]
    L.espec_sod_plan_create_absorb.argtypes = L.espec_spo_plan_create.argtypes + [ip, vp]
    L.espec_sod_plan_create.argtypes = L.espec_spo_plan_create.argtypes
    L.espec_sod_plan_run.argtypes = [vp, ip]
    L.espec_sod_plan_sync.argtypes = [vp]
    L.espec_sod_plan_get.argtypes = L.espec_spo_plan_get.argtypes
    L.espec_sod_plan_destroy.argtypes = [vp]
    L.espec_fft.argtypes = [d, d, ip, ip, ip, ip, ip]
    L.espec_au.argtypes = [ip, i, d, d, d, d]
    L.espec_lnz.argtypes = [ip, i, ip, ip, d, d, d, d, d, vp]
    L.espec_dstevx.argtypes = [ip, d, d, dp, d, vp]
    L.espec_lanczs.argtypes = [ip, i, ip, ip, ip, dp, d, d, d, d]
    L.espec_fcorrel.argtypes = [ip, d, d, P(dp)]
    L.espec_assemble_corr.argtypes = [ip, d, d, d, d]
    L.espec_hwhm.argtypes = [ip, dp, dp, dp, dp, d, d]
    L.espec_spectrumtd.argtypes = [ip, ip, dp, dp, dp, dp, d, d, d, d, d, ip]
    L.espec_lc_rule.argtypes = [dp, P(dp), dp, ip, ip, P(ip)]
    L.espec_ecnorm.argtypes = [ip, d, P(dp)]
    L.espec_efft4.argtypes = [ip, i, dp, d, d, d, d, P(dp), ip]
    L.espec_eigenerg.argtypes = [ip, i, d, d, d, P(dp)]
    _lib = L
    return L


def lib():
    return _load()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _np3(np_):
    a = np.ones(3, dtype=np.int32)
    a[: len(np_)] = np_
    return a


def _v3(x, fill=1.0):
    a = np.full(3, fill, dtype=np.float64)
    a[: len(x)] = x
    return a


def _check(rc, where):
    if rc != 0:
        raise EspecError(rc, where)


def nsteps(ti, tf, dt):
    """Trip count of DO T=TI+DT,TF+DT,DT (pspofft.f:110)."""
    c = ((tf + dt) - (ti + dt) + dt) / dt
    return max(int(c), 0)


def bw_probe(nbytes=1 << 30, iters=20):
    """Streaming copy kernel: (mean ms per launch, bytes moved per launch)."""
    ms, by = C.c_double(), C.c_double()
    _check(_load().espec_bw_probe(nbytes, iters, C.byref(ms), C.byref(by)), "bw_probe")
    return ms.value, by.value


def device_count():
    return _load().espec_device_count()


def pspofft(np_, dt, ti, tf, lc, sh, sm, u1, v1, vpot, flags=FP64 | EXACT, absorb=ABSORB_NONE,
            vabc=None):
    """PSPOFFT (pspofft.f:1-150): returns dict(cr, ci, norm, u, v).

    cr/ci are the LNZVC(NC,MPR)/LNZVC(NC,MPI) columns: sample 0 is C(TI),
    then one sample every lc steps.  absorb/vabc: ABSORB=.TRUE. with
    TPABSOR .VOPTIC or .SMOOTHW and its VABC array (spofft.f:65-118, 162-176)."""
    L = _load()
    maxs = nsteps(ti, tf, dt) // lc + 1
    cr, ci, nr = np.zeros(maxs), np.zeros(maxs), np.zeros(maxs)
    u, v = _f64(u1).copy(), _f64(v1).copy()
    ns = C.c_int(0)
    va = _f64(vabc) if vabc is not None else None
    rc = L.espec_pspofft_absorb(len(np_), _np3(np_), dt, ti, tf, lc, _v3(sh), _v3(sm), u, v,
                                _f64(vpot), cr, ci, nr.ctypes.data, maxs, C.byref(ns), flags,
                                absorb, va.ctypes.data if va is not None else None)
    _check(rc, "PSPOFFT")
    k = ns.value
    return dict(cr=cr[:k], ci=ci[:k], norm=nr[:k], u=u, v=v)


def plnz(np_, mp, abstol, dt, ti, tf, lc, shm, u1, v1, vpot, flags=FP64 | EXACT,
         absorb=ABSORB_NONE, vabc=None):
    """PLNZ (plnz.f:1-140, step sil.f:70-181): returns dict(cr, ci, norm, u, v).

    shm is the FD kinetic prefactor per axis (espec.f:630); mp the Krylov
    order; abstol the DSTEVX tolerance (<= 0: QL, > 0: bisection).
    absorb=ABSORB_SMOOTHW with the mask vabc: PLNZAUX's .SMOOTHW absorber
    (sil.f:171-180)."""
    L = _load()
    maxs = nsteps(ti, tf, dt) // lc + 1
    cr, ci, nr = np.zeros(maxs), np.zeros(maxs), np.zeros(maxs)
    u, v = _f64(u1).copy(), _f64(v1).copy()
    ns = C.c_int(0)
    va = None if vabc is None else _f64(vabc)
    rc = L.espec_plnz_absorb(len(np_), _np3(np_), mp, abstol, dt, ti, tf, lc, _v3(shm, 0.0), u,
                             v, _f64(vpot), cr, ci, nr.ctypes.data, maxs, C.byref(ns), flags,
                             absorb, None if va is None else va.ctypes.data)
    _check(rc, "PLNZ")
    k = ns.value
    return dict(cr=cr[:k], ci=ci[:k], norm=nr[:k], u=u, v=v)


def absorbingbc(np_, kind, al, ar, sh, xi, xf, voi):
    """ABSORBINGBC (absorb.f): (vabc, alpha) on the device."""
    va, alpha = np.zeros(int(np.prod(np_))), C.c_double(0.0)
    _check(_load().espec_absorbingbc(len(np_), _np3(np_), kind, _v3(al, 0.0), _v3(ar, 0.0),
                                     _v3(sh), _v3(xi, 0.0), _v3(xf, 0.0), _v3(voi, 0.0), va,
                                     C.byref(alpha)), "ABSORBINGBC")
    return va, alpha.value


def init_cond(np_, mode, xi, sh, x0, ra, rk, u=None, v=None):
    """INIT_COND (init_cond.f): Gaussian initial packet (u, v) on the
    device; u/v carry the 1-D eigenvector for the .GET* modes."""
    n = int(np.prod(np_))
    uu, vv = np.zeros(n), np.zeros(n)
    if u is not None:
        uu[:len(u)] = u
    if v is not None:
        vv[:len(v)] = v
    _check(_load().espec_init_cond(len(np_), _np3(np_), mode, _v3(xi, 0.0), _v3(sh),
                                   _v3(x0, 0.0), _v3(ra), _v3(rk, 0.0), uu, vv), "INIT_COND")
    return uu, vv


def ef(efc, e0, t0, td, tp, omg, sni, kl, t):
    """EF (ef.f): the laser field at time t (fs) for pulse shape efc."""
    return _load().espec_ef(efc, e0, t0, td, tp, omg, sni, kl, t)


def pplnz(np_, mp, abstol, dt, ti, tf, lc, shm, u1, v1, vp, dm, pulse, flags=FP64 | EXACT):
    """PPLNZ (pplnz.f:1-340) time loop: PLNZ steps with VEFF = VP - DM EF(T)
    (chpot.f:58-62) at each step's end time.  pulse = (efc, e0, t0, td, tp,
    omg, sni, kl).  Returns dict(cr, ci, norm, u, v, veff); veff is the
    last VEFF, which the reference leaves in VPOT."""
    L = _load()
    maxs = nsteps(ti, tf - dt / 2, dt) // lc + 1
    cr, ci, nr = np.zeros(maxs), np.zeros(maxs), np.zeros(maxs)
    u, v = _f64(u1).copy(), _f64(v1).copy()
    ve = np.zeros(u.size)
    ns = C.c_int(0)
    efc, e0, t0, td, tp, omg, sni, kl = pulse
    rc = L.espec_pplnz(len(np_), _np3(np_), mp, abstol, dt, ti, tf, lc, _v3(shm, 0.0), u, v,
                       _f64(vp), _f64(dm), efc, e0, t0, td, tp, omg, sni, kl, ve.ctypes.data,
                       cr, ci, nr.ctypes.data, maxs, C.byref(ns), flags)
    _check(rc, "PPLNZ")
    k = ns.value
    return dict(cr=cr[:k], ci=ci[:k], norm=nr[:k], u=u, v=v, veff=ve)


def ppsod(np_, dt, ti, tf, lc, shm, u1, v1, vp, dm, pulse, flags=FP64 | EXACT,
          absorb=ABSORB_NONE, vabc=None, proj=None, energies=False):
    """PPSOD (ppsod.f:1-548, step PSODAUX sod.f:3-122): the pulse-driven FD
    leapfrog on the device.  pulse = (efc, e0, t0, td, tp, omg, sni, kl).
    Returns dict(cr, ci, norm, u, v, uprev, vprev, veff[, et, et0][, proj])."""
    L = _load()
    n = int(np.prod(np_))
    nf = int(np.floor((tf + dt / 2 - (ti + dt)) / dt + 0.5))
    maxs = (nf + 1) // lc + 1
    cr, ci, nr = np.zeros(maxs), np.zeros(maxs), np.zeros(maxs)
    u, v = _f64(u1).copy(), _f64(v1).copy()
    up, vq, ve = np.zeros(n), np.zeros(n), np.zeros(n)
    et, et0 = (np.zeros(maxs), np.zeros(maxs)) if energies else (None, None)
    w = None if proj is None else np.ascontiguousarray(proj, dtype=np.float64).reshape(-1, n)
    npr = 0 if w is None else w.shape[0]
    pr = np.zeros((maxs, max(npr, 1), 2))
    va = None if vabc is None else _f64(vabc)
    ns = C.c_int(0)
    efc, e0, t0, td, tp, omg, sni, kl = pulse
    ptr = lambda a: None if a is None else a.ctypes.data
    rc = L.espec_ppsod(len(np_), _np3(np_), dt, ti, tf, lc, _v3(shm, 0.0), u, v, _f64(vp),
                       _f64(dm), efc, e0, t0, td, tp, omg, sni, kl, absorb, ptr(va), npr, ptr(w),
                       up.ctypes.data, vq.ctypes.data, ve.ctypes.data, cr, ci, nr.ctypes.data,
                       ptr(et), ptr(et0), pr.ctypes.data if npr else None, maxs, C.byref(ns),
                       flags)
    _check(rc, "PPSOD")
    k = ns.value
    out = dict(cr=cr[:k], ci=ci[:k], norm=nr[:k], u=u, v=v, uprev=up, vprev=vq, veff=ve)
    if energies:
        out.update(et=et[:k], et0=et0[:k])
    if npr:
        out["proj"] = pr[:k, :npr]
    return out


def s2ppsod(np_, dt, ti, tf, lc, shm, xi, sh, sm1, aa, u1, v1, vp, dm, pulse, prtcrl=2,
            tol=1e30, flags=FP64 | EXACT, absorb=ABSORB_NONE, vabc=None):
    """S2PPSOD (s2ppsod.f:1-1143): two surfaces coupled by the coordinate
    operator, FD leapfrog on the device.  u1, v1, vp, dm [2n] (surface A,
    then B); pulse = (efc, e0, t0, td, tp, omg, sni, kl); prtcrl 0 .NO,
    1 .PARTIAL, 2 .YES.  Returns dict(tab [rows][12] as printed, info, u, v,
    uprev, vprev, veff)."""
    L = _load()
    n = int(np.prod(np_))
    maxrows = int((tf - ti) / dt) + 4
    tab = np.zeros(12 * maxrows)
    u, v = _f64(u1).copy(), _f64(v1).copy()
    up, vq, ve = np.zeros(2 * n), np.zeros(2 * n), np.zeros(2 * n)
    va = None if vabc is None else _f64(vabc)
    nr, inf = C.c_int(0), C.c_int(0)
    efc, e0, t0, td, tp, omg, sni, kl = pulse
    rc = L.espec_s2ppsod(len(np_), _np3(np_), dt, ti, tf, lc, _v3(shm, 0.0), xi, sh, sm1, aa, u,
                         v, _f64(vp), _f64(dm), efc, e0, t0, td, tp, omg, sni, kl, absorb,
                         None if va is None else va.ctypes.data, prtcrl, tol, tab.ctypes.data,
                         maxrows, C.byref(nr), C.byref(inf), up.ctypes.data, vq.ctypes.data,
                         ve.ctypes.data, flags)
    _check(rc, "S2PPSOD")
    k = min(nr.value, maxrows)
    return dict(tab=tab[:12 * k].reshape(k, 12), info=inf.value, u=u, v=v, uprev=up, vprev=vq,
                veff=ve)


def psodfft(np_, dt, ti, tf, lc, sh, sm, u1, v1, vpot, flags=FP64 | EXACT, absorb=ABSORB_NONE,
            vabc=None):
    """PSODFFT (psodfft.f:1-200, step sodfft.f): returns dict(cr, ci, norm, u, v).

    Sample 0 is C(TI); the two half-strength start-up steps count as the
    first step, then one sample every lc leapfrog steps.  absorb/vabc as
    in pspofft (sodfft.f:66-72, 108-116)."""
    L = _load()
    nst = 1 + nsteps(ti + dt, tf, dt)
    maxs = nst // lc + 2
    cr, ci, nr = np.zeros(maxs), np.zeros(maxs), np.zeros(maxs)
    u, v = _f64(u1).copy(), _f64(v1).copy()
    ns = C.c_int(0)
    va = _f64(vabc) if vabc is not None else None
    rc = L.espec_psodfft_absorb(len(np_), _np3(np_), dt, ti, tf, lc, _v3(sh), _v3(sm), u, v,
                                _f64(vpot), cr, ci, nr.ctypes.data, maxs, C.byref(ns), flags,
                                absorb, va.ctypes.data if va is not None else None)
    _check(rc, "PSODFFT")
    k = ns.value
    return dict(cr=cr[:k], ci=ci[:k], norm=nr[:k], u=u, v=v)


def fft(a, b, ntot, n, nspan, isn, flags=REF):
    """FFT(A,B,NTOT,N,NSPAN,ISN) (fft3d.f:1) on the device: returns (a, b)."""
    a, b = _f64(a).copy(), _f64(b).copy()
    _check(_load().espec_fft(a, b, ntot, n, nspan, isn, flags), "FFT")
    return a, b


def pspofftaux(np_, b, cst1, cst2, tc, u, v, vpot, flags=REF, absorb=ABSORB_NONE, vabc=None):
    """PSPOFFTAUX (spofft.f:3-203): one split-operator step on the device
    with the caller's kinetic grid tc; returns (u, v)."""
    u, v = _f64(u).copy(), _f64(v).copy()
    va = None if vabc is None else _f64(vabc)
    _check(_load().espec_pspofftaux(len(np_), _np3(np_), b, cst1, cst2, _f64(tc), u, v,
                                    _f64(vpot), absorb, None if va is None else va.ctypes.data,
                                    flags), "PSPOFFTAUX")
    return u, v


def psodfftaux(np_, b, cst, tc, u1, u2, v1, v2, vpot, flags=REF, absorb=ABSORB_NONE,
               vabc=None):
    """PSODFFTAUX (sodfft.f:3-129): one leapfrog step; returns (u1, u2, v1, v2)."""
    u1, u2, v1, v2 = (_f64(a).copy() for a in (u1, u2, v1, v2))
    va = None if vabc is None else _f64(vabc)
    _check(_load().espec_psodfftaux(len(np_), _np3(np_), b, cst, _f64(tc), u1, u2, v1, v2,
                                    _f64(vpot), absorb, None if va is None else va.ctypes.data,
                                    flags), "PSODFFTAUX")
    return u1, u2, v1, v2


def reort(basis, w):
    """REORT (reort.f:1-83): w - V (V^T w) for the columns basis[j] (rows
    of the [m][n] array are the vectors); returns the new w."""
    basis = np.ascontiguousarray(basis, dtype=np.float64)
    w = _f64(w).copy()
    m, n = basis.shape
    _check(_load().espec_reort(n, m, n, w, basis), "REORT")
    return w


def lanczsg(np_, il, iu, maxint, abstol, shm, vpot, flags=REF, logcap=4096):
    """LANCZSG (lanczsc.f:1-153) on the device: (eigvl, eigvc[ns][n], log
    [(state, restart, energy, change)]).  Raises EspecError with INFO=5
    where the reference stops without convergence."""
    n = int(np.prod(np_))
    ns = iu - il + 1
    ev, vc = np.zeros(ns), np.zeros(ns * n)
    lg = np.zeros(4 * logcap)
    nl = C.c_int(0)
    rc = _load().espec_lanczsg(len(np_), _np3(np_), il, iu, maxint, abstol, _v3(shm, 0.0),
                               _f64(vpot), ev, vc, lg.ctypes.data, logcap, C.byref(nl), flags)
    log = lg[:4 * min(nl.value, logcap)].reshape(-1, 4)
    if rc:
        err = EspecError(rc, "LANCZSG")
        err.log = log
        raise err
    return ev, vc.reshape(ns, n), log


def mtrxdiag(np_, il, iu, abstol, shm, vpot):
    """MTRXDIAG (mtrxdiag.f:1-89) on the device: (eigvl, eigvc[ns][n])."""
    n = int(np.prod(np_))
    ns = iu - il + 1
    ev, vc = np.zeros(ns), np.zeros(ns * n)
    m = C.c_int(0)
    _check(_load().espec_mtrxdiag(len(np_), _np3(np_), il, iu, abstol, _v3(shm, 0.0), _f64(vpot),
                                  ev, vc, n, C.byref(m)), "MTRXDIAG")
    return ev[:m.value], vc.reshape(ns, n)[:m.value]


def fftnt(x, y, m, itype, flags=REF):
    """FFTNT(X,Y,N,M,ITYPE) (cfft.f:6-106) on the device: returns (x, y)."""
    x, y = _f64(x).copy(), _f64(y).copy()
    _check(_load().espec_fftnt(x, y, len(x), m, itype, flags), "FFTNT")
    return x, y


def au(np_, shm, vpot, u):
    """AU (au.f:3): H u with the 4th-order FD kinetic operator."""
    hu = np.zeros(int(np.prod(np_)))
    _check(_load().espec_au(len(np_), _np3(np_), _v3(shm, 0.0), _f64(vpot), _f64(u), hu), "AU")
    return hu


def lnz(np_, m, nreort, shm, vpot, start):
    """LNZ (lnz.f:4): (al[m], bt[m], basis[nreort][n])."""
    n = int(np.prod(np_))
    al, bt = np.zeros(m), np.zeros(m)
    basis = np.zeros(nreort * n)
    _check(_load().espec_lnz(len(np_), _np3(np_), m, nreort, _v3(shm, 0.0), _f64(vpot),
                             _f64(start), al, bt, basis.ctypes.data), "LNZ")
    return al, bt, basis.reshape(nreort, n)


def dstevx(d, e, abstol=0.0):
    """DSTEVX('V','A') (dstevx.f:5): (w, z) with eigenvectors as columns of z."""
    m = len(d)
    w, z = np.zeros(m), np.zeros(m * m)
    _check(_load().espec_dstevx(m, _f64(d), _f64(np.append(_f64(e), 0.0)), abstol, w,
                                z.ctypes.data), "DSTEVX")
    return w, z.reshape(m, m).T


def lanczs(np_, il, iu, nreort, shm, vpot, abstol=0.0):
    """LANCZS (lanczs.f:1): (eigvl[iu-il+1], eigvc[iu-il+1][n])."""
    n = int(np.prod(np_))
    ns = iu - il + 1
    ev, vc = np.zeros(ns), np.zeros(ns * n)
    _check(_load().espec_lanczs(len(np_), _np3(np_), il, iu, nreort, abstol, _v3(shm, 0.0),
                                _f64(vpot), ev, vc), "LANCZS")
    return ev, vc.reshape(ns, n)


def fcorrel(u1, u2):
    """FCORREL (fcorrel.f:1): sum u1 u2."""
    out = C.c_double()
    u1, u2 = _f64(u1), _f64(u2)
    _check(_load().espec_fcorrel(len(u1), u1, u2, C.byref(out)), "FCORREL")
    return out.value


def ecnorm(u):
    """ECNORM (ecnorm.f:1): sqrt(sum u^2)."""
    out = C.c_double()
    u = _f64(u)
    _check(_load().espec_ecnorm(len(u), u, C.byref(out)), "ECNORM")
    return out.value


def efft4(np_, b, u, v, tc, vpot, flags=REF):
    """EFFT4 (eigenergfft.f:1): <psi|H|psi> with the FFT kinetic term."""
    out = C.c_double()
    _check(_load().espec_efft4(len(np_), _np3(np_), b, _f64(u), _f64(v), _f64(tc), _f64(vpot),
                               C.byref(out), flags), "EFFT4")
    return out.value


def eigenerg(np_, shm, u, vpot):
    """EIGENERG (eigenerg.f:1): <u|H u> with the FD operator."""
    out = C.c_double()
    _check(_load().espec_eigenerg(len(np_), _np3(np_), _v3(shm, 0.0), _f64(u), _f64(vpot),
                                  C.byref(out)), "EIGENERG")
    return out.value


def assemble_corr(mf, cr, ci):
    """C(t) on t = -TF + (i-1) DTF from 2^(mf-1) samples (espec.f:1546-1566)."""
    n = 1 << mf
    u, v = np.zeros(n), np.zeros(n)
    _check(_load().espec_assemble_corr(mf, _f64(cr), _f64(ci), u, v), "assemble_corr")
    return u, v


def hwhm(mf, dt, tf, tal, de, u, v):
    """HWHM (hwhm.f:1): returns (u, v)."""
    u, v = _f64(u).copy(), _f64(v).copy()
    _check(_load().espec_hwhm(mf, dt, tf, tal, de, u, v), "HWHM")
    return u, v


def spectrumtd(iwin, mf, dt, tf, wp, de, u, v, flags=REF):
    """SPECTRUMTD (spectrumtd.f:1): (E, W, FFTi) in print order."""
    n = 1 << mf
    u, v = _f64(u).copy(), _f64(v).copy()
    e, w, wi = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(_load().espec_spectrumtd(iwin, mf, dt, tf, wp, de, u, v, e, w, wi, flags),
           "SPECTRUMTD")
    return e, w, wi


def lc_rule(ti, tf, dt, mf, change=True):
    """LC and the adjusted TF (espec.f:1324-1353)."""
    t, lc = C.c_double(tf), C.c_int()
    _check(_load().espec_lc_rule(ti, C.byref(t), dt, mf, int(change), C.byref(lc)), "lc_rule")
    return lc.value, t.value


def digt(np_, sh, sm, flags=EXACT):
    """DIGT (dkeft.f:204): kinetic grid p^2/2m and B = 1/N."""
    L = _load()
    tc = np.zeros(int(np.prod(np_)))
    b = C.c_double()
    _check(L.espec_digt(len(np_), _np3(np_), _v3(sh), _v3(sm), tc, C.byref(b), flags), "DIGT")
    return tc, b.value


class _Plan:
    """Shared handle logic of the device-resident plans (prefix selects the
    espec_<prefix>_plan_* entry points)."""

    _prefix = ""

    def _fn(self, name):
        return getattr(self._L, f"espec_{self._prefix}_plan_{name}")

    def _setup(self, np_, vpot, u, v, maxsamp):
        self._L = _load()
        vpot = _f64(vpot)
        n = int(np.prod(np_))
        self.n, self.np = n, list(np_)
        self.npk = vpot.size // n
        self.maxsamp = maxsamp
        self._h = C.c_void_p()
        self._vpot = vpot.ravel()
        self._u = _f64(u).ravel()
        self._v = _f64(v).ravel() if v is not None else None
        return self._v.ctypes.data if self._v is not None else None

    def run(self, nsteps):
        _check(self._fn("run")(self._h, nsteps), f"{self._prefix}_plan_run")

    def sync(self):
        _check(self._fn("sync")(self._h), f"{self._prefix}_plan_sync")

    def get(self, pk=0, psi=True):
        cr, ci, nr = np.zeros(self.maxsamp), np.zeros(self.maxsamp), np.zeros(self.maxsamp)
        u, v = np.zeros(self.n), np.zeros(self.n)
        ns = C.c_int()
        _check(self._fn("get")(self._h, pk, cr.ctypes.data, ci.ctypes.data, nr.ctypes.data,
                               self.maxsamp, C.byref(ns), u.ctypes.data if psi else None,
                               v.ctypes.data if psi else None), f"{self._prefix}_plan_get")
        k = ns.value
        out = dict(cr=cr[:k], ci=ci[:k], norm=nr[:k])
        if psi:
            out.update(u=u, v=v)
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._fn("destroy")(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpoBatchGraph:
    """Split-operator plans advanced together by one graph launch (no
    reference counterpart: replaces the job farm of tools&scrips/parallel.x).
    The graph records every kernel of `nsteps` steps of each plan from its
    current step, one branch per plan; run() replays it once and is valid
    while every plan is still at the step it was captured at (reset() the
    plans to replay from step 0)."""

    def __init__(self, plans, nsteps):
        self._L = _load()
        self.plans = list(plans)
        self.nsteps = nsteps
        self._h = C.c_void_p()
        arr = (C.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        _check(self._L.espec_spo_batch_create(C.byref(self._h), arr, len(self.plans), nsteps),
               "spo_batch_create")

    def run(self):
        _check(self._L.espec_spo_batch_run(self._h), "spo_batch_run")

    def sync(self):
        """Wait for the last run; its graph's duration in ms (HIP events)."""
        ms = C.c_double()
        _check(self._L.espec_spo_batch_sync(self._h, C.byref(ms)), "spo_batch_sync")
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            self._L.espec_spo_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SilPlan(_Plan):
    """Device-resident batch of short-iterative Lanczos packets on one grid
    (PLNZ's time loop, plnz.f:109-131, for npk packets at once)."""

    _prefix = "sil"

    def __init__(self, np_, mp, abstol, dt, lc, shm, vpot, u, v=None, maxsamp=1,
                 flags=FP64 | EXACT):
        vp = self._setup(np_, vpot, u, v, maxsamp)
        rc = self._L.espec_sil_plan_create(C.byref(self._h), len(np_), _np3(np_), self.npk, mp,
                                           abstol, dt, lc, _v3(shm, 0.0), self._vpot, self._u,
                                           vp, maxsamp, flags)
        _check(rc, "sil_plan_create")
        self.mp = mp

    def set_pulse(self, dm, efield):
        """Pulse-driven steps from now on (PPLNZ): step g of the plan uses
        VPOT - dm * efield[g-1]; dm shaped like vpot."""
        self._dm = _f64(dm).ravel()
        self._ef = _f64(efield).ravel()
        _check(self._L.espec_sil_plan_set_pulse(self._h, self._dm, self._ef, self._ef.size),
               "sil_plan_set_pulse")

    def set_absorb(self, absorb, vabc=None):
        """PLNZAUX's .SMOOTHW mask (sil.f:171-180) for every packet, vabc
        [npk][n]; ABSORB_NONE removes it."""
        self._va = None if vabc is None else _f64(vabc)
        _check(self._L.espec_sil_plan_set_absorb(
            self._h, absorb, None if self._va is None else self._va.ctypes.data),
            "sil_plan_set_absorb")

    def adaptive(self, ndt, params=None, tol=1e-10, m_fix=10):
        """Advance ndt base steps with the learned controller choosing the
        Krylov order and step scale (params: its 4866 FP32 parameters), or
        with the fixed-order step every dt when params is None (no reference
        counterpart).  Returns the run statistics."""
        st = (C.c_longlong * 7)()
        pp = None
        if params is not None:
            self._params = np.ascontiguousarray(params, dtype=np.float32)
            assert self._params.size == MLP_NPAR
            pp = self._params.ctypes.data
        _check(self._L.espec_sil_plan_adaptive(self._h, pp, ndt, tol, m_fix, st),
               "sil_plan_adaptive")
        keys = ("steps", "scale_half", "scale_1", "scale_2", "fallbacks", "krylov_sum", "waits")
        return dict(zip(keys, list(st)))

    def adaptive_log(self, cap=None):
        """Record the next adaptive run's attempts (cap records), or, with
        cap None, stop and return the records [n][5] = (base step, order,
        scale, estimate, budget)."""
        n = C.c_int()
        if cap is not None:
            self._log = np.zeros((cap, 5))
            _check(self._L.espec_sil_plan_adaptive_log(self._h, self._log.ctypes.data, cap,
                                                       C.byref(n)), "adaptive_log")
            return None
        _check(self._L.espec_sil_plan_adaptive_log(self._h, None, 0, C.byref(n)), "adaptive_log")
        return self._log[:n.value].copy()

    def reset(self):
        """Restart from psi0 at step 0 (captured graphs kept)."""
        _check(self._L.espec_sil_plan_reset(self._h), "sil_plan_reset")

    def trace(self, nsteps, mprop=None):
        """nsteps steps of mp (>= 20) iterations advancing with order mprop
        (default mp): controller inputs [nsteps][npk][8] and error estimates
        [nsteps][2 npk][17][3] of orders 4..20 at dt/2, dt, 2 dt."""
        feat = np.zeros((nsteps, self.npk, 8), dtype=np.float32)
        lab = np.zeros((nsteps, 2 * self.npk, 17, 3), dtype=np.float64)
        _check(self._L.espec_sil_plan_trace(self._h, nsteps, mprop or self.mp, feat, lab),
               "sil_plan_trace")
        return feat, lab


class SodPlan(_Plan):
    """Device-resident batch of second-order-differencing packets on one
    grid (PSODFFT's time loop for npk packets at once); the first run step
    is the two half-strength start-up steps."""

    _prefix = "sod"

    def __init__(self, np_, dt, lc, sh, sm, vpot, u, v=None, maxsamp=1, flags=FP64 | EXACT,
                 absorb=ABSORB_NONE, vabc=None):
        vp = self._setup(np_, vpot, u, v, maxsamp)
        self._vabc = _f64(vabc).ravel() if vabc is not None else None
        rc = self._L.espec_sod_plan_create_absorb(
            C.byref(self._h), len(np_), _np3(np_), self.npk, dt, lc, _v3(sh), _v3(sm),
            self._vpot, self._u, vp, maxsamp, flags, absorb,
            self._vabc.ctypes.data if self._vabc is not None else None)
        _check(rc, "sod_plan_create")


MLP_NPAR = 4866  # controller network parameters: W1[64][8] b1 W2[64][64] b2 W3[2][64] b3


def mlp_infer(params, feat):
    """Controller network outputs [n][2] for inputs [n][8] (bfloat16 on the device)."""
    L = _load()
    feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, 8)
    out = np.zeros((feat.shape[0], 2), dtype=np.float32)
    _check(L.espec_mlp_infer(np.ascontiguousarray(params, dtype=np.float32), feat,
                             feat.shape[0], out), "mlp_infer")
    return out


def mlp_grad(params, x, t):
    """Gradient and loss of one 64-row minibatch on the device."""
    L = _load()
    g = np.zeros(MLP_NPAR, dtype=np.float32)
    loss = C.c_float()
    _check(L.espec_mlp_grad(np.ascontiguousarray(params, dtype=np.float32),
                            np.ascontiguousarray(x, dtype=np.float32).reshape(64, 8),
                            np.ascontiguousarray(t, dtype=np.float32).reshape(64, 2), g,
                            C.byref(loss)), "mlp_grad")
    return g, loss.value


def mlp_train(params, x, t, epochs, lr=1e-3, seed=1):
    """Adam training on the device; returns (new params, per-epoch loss)."""
    L = _load()
    p = np.array(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 8)
    t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 2)
    loss = np.zeros(epochs, dtype=np.float32)
    _check(L.espec_mlp_train(p, x, t, x.shape[0], epochs, lr, seed, loss), "mlp_train")
    return p, loss


_SHAPE_KEYS = ("lds", "rows", "rows_nt", "tile", "tile_nt", "spread", "xmap", "grid", "rows_pen",
               "tile_pen", "pen_pf", "vph", "fused")


class SpoPlan(_Plan):
    """Device-resident batch of split-operator packets on one grid (no
    reference counterpart: replaces the job farm of tools&scrips/parallel.x)."""

    _prefix = "spo"

    def __init__(self, np_, dt, lc, sh, sm, vpot, u, v=None, maxsamp=1, flags=MIXED,
                 absorb=ABSORB_NONE, vabc=None):
        """sm: masses per axis, or an [npk][ndim] array of per-packet masses;
        absorb/vabc: an absorber per packet, vabc shaped like vpot."""
        vp = self._setup(np_, vpot, u, v, maxsamp)
        sm = np.asarray(sm, dtype=np.float64)
        if sm.ndim == 2:
            flags |= PER_PACKET_MASS
            self._sm = np.ascontiguousarray(sm).ravel()
        else:
            self._sm = _v3(sm)
        self._vabc = _f64(vabc).ravel() if vabc is not None else None
        rc = self._L.espec_spo_plan_create_absorb(
            C.byref(self._h), len(np_), _np3(np_), self.npk, dt, lc, _v3(sh), self._sm,
            self._vpot, self._u, vp, maxsamp, flags, absorb,
            self._vabc.ctypes.data if self._vabc is not None else None)
        _check(rc, "spo_plan_create")

    def timing(self):
        ms, nl = C.c_double(), C.c_int()
        _check(self._L.espec_spo_plan_last_timing(self._h, C.byref(ms), C.byref(nl)), "timing")
        return ms.value, nl.value

    def shape(self):
        """Launch shape: dict(lds, rows, rows_nt, tile, tile_nt, spread, xmap, grid,
        rows_pen, tile_pen, pen_pf, vph, fused)."""
        a = np.zeros(len(_SHAPE_KEYS), dtype=np.int32)
        _check(self._L.espec_spo_plan_shape(self._h, a), "spo_plan_shape")
        return dict(zip(_SHAPE_KEYS, a.tolist()))

    def set_shape(self, **kw):
        """Pin the multi-pass launch shape: the current shape with the given
        fields replaced (keys as returned by shape())."""
        sh = self.shape()
        sh.update(kw)
        a = np.array([sh[k] for k in _SHAPE_KEYS], dtype=np.int32)
        _check(self._L.espec_spo_plan_set_shape(self._h, a), "spo_plan_set_shape")

    def reset(self):
        """Restart from psi0 at step 0 (tables and captured graphs kept)."""
        _check(self._L.espec_spo_plan_reset(self._h), "spo_plan_reset")

    def profile(self, nsteps):
        """Per-kernel HIP-event timing: list of (mean ms, algorithmic bytes)
        per kernel kind (0: step kernel or row pass, 1: strided passes)."""
        ms, by = np.zeros(4), np.zeros(4)
        k = C.c_int()
        _check(self._L.espec_spo_plan_profile(self._h, nsteps, ms, by, 4, C.byref(k)),
               "spo_plan_profile")
        return [(ms[i], by[i]) for i in range(k.value)]
