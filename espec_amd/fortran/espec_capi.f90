! ISO_C_BINDING interfaces to libespec_hip.so (include/espec_hip.h) and the
! small helpers the drop-in subroutines share.  Arrays cross the boundary
! as plain addresses (assumed-size REAL*8 / INTEGER arrays); no descriptor
! or framework type is involved.
module espec_capi
  use iso_c_binding
  implicit none

  integer, parameter :: ESPEC_FP64 = 0, ESPEC_MIXED = 1, ESPEC_EXACT = 0, ESPEC_REF = 2
  integer, parameter :: ESPEC_OK = 0, ESPEC_ERR_ARG = 2, ESPEC_ERR_DIAG = 5, &
                        ESPEC_ERR_UNSUPPORTED = 6
  integer, parameter :: ESPEC_ABSORB_NONE = 0, ESPEC_ABSORB_VOPTIC = 1, ESPEC_ABSORB_SMOOTHW = 2
  integer, parameter :: ESPEC_ABG_SMOOTHW = 1, ESPEC_ABG_SMRS = 2, ESPEC_ABG_VOPTIC_LINEAR = 3, &
                        ESPEC_ABG_VOPTIC_KOSLOFF = 4, ESPEC_ABG_VOPTIC_POLIN = 5

  interface
    function espec_strerror(code) bind(C, name='espec_strerror') result(p)
      import :: c_int, c_ptr
      integer(c_int), value :: code
      type(c_ptr) :: p
    end function

    function espec_digt(ndim, np, sh, sm, tc, b, flags) bind(C, name='espec_digt') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, flags
      integer(c_int) :: np(*)
      real(c_double) :: sh(*), sm(*), tc(*), b
      integer(c_int) :: rc
    end function

    function espec_spo_plan_create(plan, ndim, np, npk, dt, lc, sh, sm, vpot, u, v, maxsamp, &
                                   flags) bind(C, name='espec_spo_plan_create') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr) :: plan
      integer(c_int), value :: ndim, npk, lc, maxsamp, flags
      integer(c_int) :: np(*)
      real(c_double), value :: dt
      real(c_double) :: sh(*), sm(*), vpot(*), u(*), v(*)
      integer(c_int) :: rc
    end function

    function espec_spo_plan_create_absorb(plan, ndim, np, npk, dt, lc, sh, sm, vpot, u, v, &
                                          maxsamp, flags, absorb, vabc) &
        bind(C, name='espec_spo_plan_create_absorb') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr) :: plan
      integer(c_int), value :: ndim, npk, lc, maxsamp, flags, absorb
      integer(c_int) :: np(*)
      real(c_double), value :: dt
      real(c_double) :: sh(*), sm(*), vpot(*), u(*), v(*), vabc(*)
      integer(c_int) :: rc
    end function

    function espec_sod_plan_create(plan, ndim, np, npk, dt, lc, sh, sm, vpot, u, v, maxsamp, &
                                   flags) bind(C, name='espec_sod_plan_create') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr) :: plan
      integer(c_int), value :: ndim, npk, lc, maxsamp, flags
      integer(c_int) :: np(*)
      real(c_double), value :: dt
      real(c_double) :: sh(*), sm(*), vpot(*), u(*), v(*)
      integer(c_int) :: rc
    end function

    function espec_sod_plan_create_absorb(plan, ndim, np, npk, dt, lc, sh, sm, vpot, u, v, &
                                          maxsamp, flags, absorb, vabc) &
        bind(C, name='espec_sod_plan_create_absorb') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr) :: plan
      integer(c_int), value :: ndim, npk, lc, maxsamp, flags, absorb
      integer(c_int) :: np(*)
      real(c_double), value :: dt
      real(c_double) :: sh(*), sm(*), vpot(*), u(*), v(*), vabc(*)
      integer(c_int) :: rc
    end function

    function espec_sil_plan_set_pulse(plan, dm, efield, nfield) &
        bind(C, name='espec_sil_plan_set_pulse') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: plan
      real(c_double) :: dm(*), efield(*)
      integer(c_int), value :: nfield
      integer(c_int) :: rc
    end function

    function espec_absorbingbc(ndim, np, kind, al, ar, sh, xi, xf, voi, vabc, alpha) &
        bind(C, name='espec_absorbingbc') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, kind
      integer(c_int) :: np(*)
      real(c_double) :: al(*), ar(*), sh(*), xi(*), xf(*), voi(*), vabc(*), alpha
      integer(c_int) :: rc
    end function

    function espec_ef(efc, e0, t0, td, tp, omg, sni, kl, t) bind(C, name='espec_ef') result(r)
      import :: c_int, c_double
      integer(c_int), value :: efc, kl
      real(c_double), value :: e0, t0, td, tp, omg, sni, t
      real(c_double) :: r
    end function

    function espec_sil_plan_create(plan, ndim, np, npk, mp, abstol, dt, lc, shm, vpot, u, v, &
                                   maxsamp, flags) bind(C, name='espec_sil_plan_create') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr) :: plan
      integer(c_int), value :: ndim, npk, mp, lc, maxsamp, flags
      integer(c_int) :: np(*)
      real(c_double), value :: abstol, dt
      real(c_double) :: shm(*), vpot(*), u(*), v(*)
      integer(c_int) :: rc
    end function

    ! run / get / destroy share one signature across the three plan kinds
    function espec_spo_plan_run(plan, nsteps) bind(C, name='espec_spo_plan_run') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: nsteps
      integer(c_int) :: rc
    end function
    function espec_sod_plan_run(plan, nsteps) bind(C, name='espec_sod_plan_run') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: nsteps
      integer(c_int) :: rc
    end function
    function espec_sil_plan_run(plan, nsteps) bind(C, name='espec_sil_plan_run') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: nsteps
      integer(c_int) :: rc
    end function

    function espec_spo_plan_get(plan, pk, cr, ci, nrm, maxsamp, nsamp, u, v) &
        bind(C, name='espec_spo_plan_get') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: pk, maxsamp
      real(c_double) :: cr(*), ci(*), nrm(*), u(*), v(*)
      integer(c_int) :: nsamp, rc
    end function
    function espec_sod_plan_get(plan, pk, cr, ci, nrm, maxsamp, nsamp, u, v) &
        bind(C, name='espec_sod_plan_get') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: pk, maxsamp
      real(c_double) :: cr(*), ci(*), nrm(*), u(*), v(*)
      integer(c_int) :: nsamp, rc
    end function
    function espec_sil_plan_get(plan, pk, cr, ci, nrm, maxsamp, nsamp, u, v) &
        bind(C, name='espec_sil_plan_get') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: pk, maxsamp
      real(c_double) :: cr(*), ci(*), nrm(*), u(*), v(*)
      integer(c_int) :: nsamp, rc
    end function
    function espec_sod_plan_get_prev(plan, pk, u, v) bind(C, name='espec_sod_plan_get_prev') &
        result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: pk
      real(c_double) :: u(*), v(*)
      integer(c_int) :: rc
    end function

    function espec_spo_plan_destroy(plan) bind(C, name='espec_spo_plan_destroy') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int) :: rc
    end function
    function espec_sod_plan_destroy(plan) bind(C, name='espec_sod_plan_destroy') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int) :: rc
    end function
    function espec_sil_plan_destroy(plan) bind(C, name='espec_sil_plan_destroy') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
      integer(c_int) :: rc
    end function

    function espec_sil_plan_set_absorb(plan, absorb, vabc) &
        bind(C, name='espec_sil_plan_set_absorb') result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: plan
      integer(c_int), value :: absorb
      real(c_double) :: vabc(*)
      integer(c_int) :: rc
    end function

    function espec_pspofftaux(ndim, np, b, cst1, cst2, tc, u, v, vpot, absorb, vabc, flags) &
        bind(C, name='espec_pspofftaux') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, absorb, flags
      integer(c_int) :: np(*)
      real(c_double), value :: b, cst1, cst2
      real(c_double) :: tc(*), u(*), v(*), vpot(*), vabc(*)
      integer(c_int) :: rc
    end function

    function espec_psodfftaux(ndim, np, b, cst, tc, u1, u2, v1, v2, vpot, absorb, vabc, flags) &
        bind(C, name='espec_psodfftaux') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, absorb, flags
      integer(c_int) :: np(*)
      real(c_double), value :: b, cst
      real(c_double) :: tc(*), u1(*), u2(*), v1(*), v2(*), vpot(*), vabc(*)
      integer(c_int) :: rc
    end function

    function espec_reort(n, m, ld, w, basis) bind(C, name='espec_reort') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: n, m, ld
      real(c_double) :: w(*), basis(*)
      integer(c_int) :: rc
    end function

    function espec_fftnt(x, y, n, m, itype, flags) bind(C, name='espec_fftnt') result(rc)
      import :: c_int, c_double
      real(c_double) :: x(*), y(*)
      integer(c_int), value :: n, m, itype, flags
      integer(c_int) :: rc
    end function

    function espec_plnz_absorb(ndim, np, mp, abstol, dt, ti, tf, lc, shm, u, v, vpot, cr, ci, &
                               nrm, maxsamp, nsamp, flags, absorb, vabc) &
        bind(C, name='espec_plnz_absorb') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, mp, lc, maxsamp, flags, absorb
      integer(c_int) :: np(*), nsamp
      real(c_double), value :: abstol, dt, ti, tf
      real(c_double) :: shm(*), u(*), v(*), vpot(*), cr(*), ci(*), nrm(*), vabc(*)
      integer(c_int) :: rc
    end function

    function espec_lnz(ndim, np, m, nreort, shm, vpot, start, al, bt, basis) &
        bind(C, name='espec_lnz') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, m, nreort
      integer(c_int) :: np(*)
      real(c_double) :: shm(*), vpot(*), start(*), al(*), bt(*), basis(*)
      integer(c_int) :: rc
    end function

    function espec_ppsod(ndim, np, dt, ti, tf, lc, shm, u, v, vp, dm, efc, e0, t0, td, tp, &
                         omg, sni, kl, absorb, vabc, npr, proj, uprev, vprev, veff, cr, ci, &
                         nrm, et, et0, pr, maxsamp, nsamp, flags) &
        bind(C, name='espec_ppsod') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, lc, efc, kl, absorb, npr, maxsamp, flags
      integer(c_int) :: np(*), nsamp
      real(c_double), value :: dt, ti, tf, e0, t0, td, tp, omg, sni
      real(c_double) :: shm(*), u(*), v(*), vp(*), dm(*), vabc(*), proj(*), uprev(*), vprev(*)
      real(c_double) :: veff(*), cr(*), ci(*), nrm(*), et(*), et0(*), pr(*)
      integer(c_int) :: rc
    end function

    function espec_fft(a, b, ntot, n, nspan, isn, flags) bind(C, name='espec_fft') result(rc)
      import :: c_int, c_double
      real(c_double) :: a(*), b(*)
      integer(c_int), value :: ntot, n, nspan, isn, flags
      integer(c_int) :: rc
    end function

    function espec_au(ndim, np, shm, vpot, u, hu) bind(C, name='espec_au') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim
      integer(c_int) :: np(*)
      real(c_double) :: shm(*), vpot(*), u(*), hu(*)
      integer(c_int) :: rc
    end function

    function espec_fcorrel(n, u1, u2, out) bind(C, name='espec_fcorrel') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double) :: u1(*), u2(*), out
      integer(c_int) :: rc
    end function

    function espec_ecnorm(n, u, out) bind(C, name='espec_ecnorm') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double) :: u(*), out
      integer(c_int) :: rc
    end function

    function espec_efft4(ndim, np, b, u, v, tc, vpot, e, flags) bind(C, name='espec_efft4') &
        result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, flags
      integer(c_int) :: np(*)
      real(c_double), value :: b
      real(c_double) :: u(*), v(*), tc(*), vpot(*), e
      integer(c_int) :: rc
    end function

    function espec_eigenerg(ndim, np, shm, u, vpot, e) bind(C, name='espec_eigenerg') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim
      integer(c_int) :: np(*)
      real(c_double) :: shm(*), u(*), vpot(*), e
      integer(c_int) :: rc
    end function

    function espec_lanczs(ndim, np, il, iu, nreort, abstol, shm, vpot, eigvl, eigvc) &
        bind(C, name='espec_lanczs') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, il, iu, nreort
      integer(c_int) :: np(*)
      real(c_double), value :: abstol
      real(c_double) :: shm(*), vpot(*), eigvl(*), eigvc(*)
      integer(c_int) :: rc
    end function

    function espec_s2ppsod(ndim, np, dt, ti, tf, lc, shm, xi, sh, sm1, aa, u, v, vp, dm, efc, &
                           e0, t0, td, tp, omg, sni, kl, absorb, vabc, prtcrl, tol, tab, &
                           maxrows, nrows, info, uprev, vprev, veff, flags) &
        bind(C, name='espec_s2ppsod') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, lc, efc, kl, absorb, prtcrl, maxrows, flags
      integer(c_int) :: np(*), nrows, info
      real(c_double), value :: dt, ti, tf, xi, sh, sm1, aa, e0, t0, td, tp, omg, sni, tol
      real(c_double) :: shm(*), u(*), v(*), vp(*), dm(*), vabc(*), tab(*), uprev(*), vprev(*), &
                        veff(*)
      integer(c_int) :: rc
    end function

    function espec_lanczsg(ndim, np, il, iu, maxint, abstol, shm, vpot, eigvl, eigvc, &
                           log, logcap, nlog, flags) bind(C, name='espec_lanczsg') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, il, iu, maxint, logcap, flags
      integer(c_int) :: np(*), nlog
      real(c_double), value :: abstol
      real(c_double) :: shm(*), vpot(*), eigvl(*), eigvc(*), log(*)
      integer(c_int) :: rc
    end function

    function espec_mtrxdiag(ndim, np, il, iu, abstol, shm, vpot, eigvl, eigvc, ldz, m) &
        bind(C, name='espec_mtrxdiag') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: ndim, il, iu, ldz
      integer(c_int) :: np(*), m
      real(c_double), value :: abstol
      real(c_double) :: shm(*), vpot(*), eigvl(*), eigvc(*)
      integer(c_int) :: rc
    end function

    function espec_hwhm(mf, dt, tf, tal, de, u, v) bind(C, name='espec_hwhm') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: mf
      real(c_double), value :: dt, tf, tal, de
      real(c_double) :: u(*), v(*)
      integer(c_int) :: rc
    end function

    function espec_spectrumtd(iwin, mf, dt, tf, wp, de, u, v, e, w, wi, flags) &
        bind(C, name='espec_spectrumtd') result(rc)
      import :: c_int, c_double
      integer(c_int), value :: iwin, mf, flags
      real(c_double), value :: dt, tf, wp, de
      real(c_double) :: u(*), v(*), e(*), w(*), wi(*)
      integer(c_int) :: rc
    end function
  end interface

contains

  ! DIM keyword -> number of axes of the supported operators (.1D/.2D/.3D);
  ! 0 for the coupled-surface variants (.2DC, .2DCT, ...), not on the device.
  integer function dim_of(dim)
    character(len=*), intent(in) :: dim
    dim_of = 0
    if (len(dim) < 3) return
    if (len_trim(dim) > 3) then
      if (dim(4:4) /= ' ') return
    end if
    select case (dim(1:3))
    case ('.1D')
      dim_of = 1
    case ('.2D')
      dim_of = 2
    case ('.3D')
      dim_of = 3
    end select
  end function

  ! Arithmetic of the device path, from the environment: ESPEC_MODE=REF
  ! (default: the reference's own rounding quirks, FP64), EXACT (FP64,
  ! quirks removed) or MIXED (complex64 state, FP64 arithmetic).
  integer function mode_flags()
    character(len=16) :: val
    integer :: ln, st
    call get_environment_variable('ESPEC_MODE', val, ln, st)
    mode_flags = ESPEC_FP64 + ESPEC_REF
    if (st /= 0) return
    select case (val(1:max(ln, 1)))
    case ('EXACT', 'exact')
      mode_flags = ESPEC_FP64 + ESPEC_EXACT
    case ('MIXED', 'mixed')
      mode_flags = ESPEC_MIXED + ESPEC_EXACT
    end select
  end function

  ! Absorber kind of ABSORB/TPABSOR as PSPOFFTAUX and PSODFFTAUX test it
  ! (spofft.f:65,162; sodfft.f:66,108): ESPEC_ABSORB_NONE when ABSORB is
  ! false or TPABSOR names neither .VOPTIC nor .SMOOTHW.
  integer function absorb_of(absorb, tpabsor)
    logical, intent(in) :: absorb
    character(len=*), intent(in) :: tpabsor
    absorb_of = 0
    if (.not. absorb) return
    if (starts(tpabsor, '.VOPTIC')) absorb_of = 1
    if (starts(tpabsor, '.SMOOTHW')) absorb_of = 2
  end function

  ! Pulse shape code of EFC as EF tests it (ef.f:40-58); -1 where the
  ! reference stops ("Desired laser pulse wasn't found").
  integer function pulse_of(efc)
    character(len=*), intent(in) :: efc
    pulse_of = -1
    if (starts(efc, '.GAUS')) then
      pulse_of = 1
    else if (starts(efc, '.GGAUS')) then
      pulse_of = 2
    else if (starts(efc, '.SIN2')) then
      pulse_of = 3
    else if (starts(efc, '.NONE') .or. starts(efc, '.NULL')) then
      pulse_of = 0
    end if
  end function

  logical function starts(s, p)
    character(len=*), intent(in) :: s, p
    starts = .false.
    if (len(s) >= len(p)) starts = s(1:len(p)) == p
  end function

  ! Trip count of DO T=M1,M2,M3 (pspofft.f:110).
  integer function trip(m1, m2, m3)
    real(8), intent(in) :: m1, m2, m3
    trip = max(int((m2 - m1 + m3) / m3), 0)
  end function

  ! C(t) = <psi0|psi>, |psi|^2 terms on the device (FCORREL, fcorrel.f:1).
  subroutine corr(n, u0, v0, u, v, cr, ci, fnr, fni)
    integer, intent(in) :: n
    real(8) :: u0(*), v0(*), u(*), v(*)
    real(8), intent(out) :: cr, ci, fnr, fni
    real(8) :: a, b
    integer :: rc
    rc = espec_fcorrel(n, u0, u, a)
    rc = espec_fcorrel(n, v0, v, b)
    cr = a + b
    rc = espec_fcorrel(n, u0, v, a)
    rc = espec_fcorrel(n, v0, u, b)
    ci = a - b
    rc = espec_fcorrel(n, u, u, a)
    rc = espec_fcorrel(n, v, v, b)
    fnr = a + b
    rc = espec_fcorrel(n, u, v, a)
    rc = espec_fcorrel(n, v, u, b)
    fni = a - b
  end subroutine

end module espec_capi
