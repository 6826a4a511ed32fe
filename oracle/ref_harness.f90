! Test-only C-callable harness around the eSPec reference Fortran.
!
! This file is compiled together with the reference sources, where they
! lie under the reference checkout, into oracle/_ref/libespec_ref.so (see
! oracle/Makefile).  It is checker infrastructure: only the tests and the
! CPU-baseline leg of bench.py load it, never the product library.
!
! Each wrapper turns plain C arguments into the reference's Fortran
! 77 calling convention (character flags, LNZVC column offsets, work
! arrays) and calls the reference subroutine unchanged.
module espec_ref_harness
  use iso_c_binding
  implicit none
contains

  subroutine dimstr(ndim, s)
    integer(c_int), intent(in) :: ndim
    character(len=8), intent(out) :: s
    if (ndim == 1) then
      s = '.1D     '
    else if (ndim == 2) then
      s = '.2D     '
    else
      s = '.3D     '
    end if
  end subroutine dimstr

  ! PSPOFFT (pspofft.f:1-150).  C(t) samples go to cr/ci (length >= nsamp).
  subroutine ref_pspofft(ndim, np, dt, ti, tf, lc, sh, sm, u1, v1, vpot, &
                         cr, ci, nsamp, info) bind(C, name='ref_pspofft')
    integer(c_int), value :: ndim, lc, nsamp
    integer(c_int) :: np(3)
    real(c_double), value :: dt, ti, tf
    real(c_double) :: sh(3), sm(3), u1(*), v1(*), vpot(*), cr(*), ci(*)
    integer(c_int) :: info
    character(len=8) :: dims
    integer :: n, mxdct, i
    real(8), allocatable :: tc(:), vabc(:), lnzvc(:,:), w1(:), w2(:)
    real(8) :: shm(3), xi(3), xf(3), var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    mxdct = max(n, nsamp + 2)
    allocate(tc(n), vabc(n), lnzvc(mxdct, 2), w1(n), w2(n))
    lnzvc = 0d0
    vabc = 1d0
    shm = 0d0; xi = 0d0; xf = 0d0; var = 0d0
    call dimstr(ndim, dims)
    call PSPOFFT(.false., dims, '.NO     ', '.NONE   ', info, lc, 0, 2, 1, &
                 mxdct, n, dt, ti, tf, 1d30, np, shm, sh, sm, u1, v1, xi, xf, &
                 tc, vpot, vabc, lnzvc, w1, w2, var)
    do i = 1, nsamp
      cr(i) = lnzvc(i, 1)
      ci(i) = lnzvc(i, 2)
    end do
    deallocate(tc, vabc, lnzvc, w1, w2)
  end subroutine ref_pspofft

  ! PSODFFT (psodfft.f:1-211).
  subroutine ref_psodfft(ndim, np, dt, ti, tf, lc, sh, sm, u1, v1, vpot, &
                         cr, ci, nsamp, info) bind(C, name='ref_psodfft')
    integer(c_int), value :: ndim, lc, nsamp
    integer(c_int) :: np(3)
    real(c_double), value :: dt, ti, tf
    real(c_double) :: sh(3), sm(3), u1(*), v1(*), vpot(*), cr(*), ci(*)
    integer(c_int) :: info
    character(len=8) :: dims
    integer :: n, mxdct, i
    real(8), allocatable :: tc(:), vabc(:), lnzvc(:,:), u2(:), v2(:), wr(:), wi(:)
    real(8) :: shm(3), xi(3), xf(3), var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    mxdct = max(n, nsamp + 2)
    allocate(tc(n), vabc(n), lnzvc(mxdct, 2), u2(n), v2(n), wr(n), wi(n))
    lnzvc = 0d0
    vabc = 1d0
    shm = 0d0; xi = 0d0; xf = 0d0; var = 0d0
    call dimstr(ndim, dims)
    call PSODFFT(.false., dims, '.NO     ', '.NONE   ', info, lc, 0, 2, 1, &
                 mxdct, n, dt, ti, tf, 1d30, np, shm, sh, sm, u1, u2, v1, v2, &
                 xi, xf, tc, vpot, vabc, wr, wi, var, lnzvc)
    do i = 1, nsamp
      cr(i) = lnzvc(i, 1)
      ci(i) = lnzvc(i, 2)
    end do
    deallocate(tc, vabc, lnzvc, u2, v2, wr, wi)
  end subroutine ref_psodfft

  ! TPABSOR of absorber kind 1 (.VOPTIC) or 2 (.SMOOTHW).
  subroutine absstr(kind, s)
    integer(c_int), value :: kind
    character(len=8) :: s
    s = '.NONE   '
    if (kind == 1) s = '.VOPTIC '
    if (kind == 2) s = '.SMOOTHW'
  end subroutine absstr

  ! PSPOFFT with ABSORB=.TRUE., TPABSOR from kind and the caller's VABC.
  subroutine ref_pspofft_abs(ndim, np, dt, ti, tf, lc, sh, sm, u1, v1, vpot, &
                             kind, vabc, cr, ci, nsamp, info) bind(C, name='ref_pspofft_abs')
    integer(c_int), value :: ndim, lc, nsamp, kind
    integer(c_int) :: np(3)
    real(c_double), value :: dt, ti, tf
    real(c_double) :: sh(3), sm(3), u1(*), v1(*), vpot(*), vabc(*), cr(*), ci(*)
    integer(c_int) :: info
    character(len=8) :: dims, tpab
    integer :: n, mxdct, i
    real(8), allocatable :: tc(:), lnzvc(:,:), w1(:), w2(:)
    real(8) :: shm(3), xi(3), xf(3), var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    mxdct = max(n, nsamp + 2)
    allocate(tc(n), lnzvc(mxdct, 2), w1(n), w2(n))
    lnzvc = 0d0
    shm = 0d0; xi = 0d0; xf = 0d0; var = 0d0
    call dimstr(ndim, dims)
    call absstr(kind, tpab)
    call PSPOFFT(.true., dims, '.NO     ', tpab, info, lc, 0, 2, 1, &
                 mxdct, n, dt, ti, tf, 1d30, np, shm, sh, sm, u1, v1, xi, xf, &
                 tc, vpot, vabc, lnzvc, w1, w2, var)
    do i = 1, nsamp
      cr(i) = lnzvc(i, 1)
      ci(i) = lnzvc(i, 2)
    end do
    deallocate(tc, lnzvc, w1, w2)
  end subroutine ref_pspofft_abs

  ! PSODFFT with ABSORB=.TRUE., TPABSOR from kind and the caller's VABC.
  subroutine ref_psodfft_abs(ndim, np, dt, ti, tf, lc, sh, sm, u1, v1, vpot, &
                             kind, vabc, cr, ci, nsamp, info) bind(C, name='ref_psodfft_abs')
    integer(c_int), value :: ndim, lc, nsamp, kind
    integer(c_int) :: np(3)
    real(c_double), value :: dt, ti, tf
    real(c_double) :: sh(3), sm(3), u1(*), v1(*), vpot(*), vabc(*), cr(*), ci(*)
    integer(c_int) :: info
    character(len=8) :: dims, tpab
    integer :: n, mxdct, i
    real(8), allocatable :: tc(:), lnzvc(:,:), u2(:), v2(:), wr(:), wi(:)
    real(8) :: shm(3), xi(3), xf(3), var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    mxdct = max(n, nsamp + 2)
    allocate(tc(n), lnzvc(mxdct, 2), u2(n), v2(n), wr(n), wi(n))
    lnzvc = 0d0
    shm = 0d0; xi = 0d0; xf = 0d0; var = 0d0
    call dimstr(ndim, dims)
    call absstr(kind, tpab)
    call PSODFFT(.true., dims, '.NO     ', tpab, info, lc, 0, 2, 1, &
                 mxdct, n, dt, ti, tf, 1d30, np, shm, sh, sm, u1, u2, v1, v2, &
                 xi, xf, tc, vpot, vabc, wr, wi, var, lnzvc)
    do i = 1, nsamp
      cr(i) = lnzvc(i, 1)
      ci(i) = lnzvc(i, 2)
    end do
    deallocate(tc, lnzvc, u2, v2, wr, wi)
  end subroutine ref_psodfft_abs

  ! PLNZ (plnz.f:1-156); Krylov order mp, DSTEVX tolerance abstol.
  subroutine ref_plnz(ndim, np, mp, abstol, dt, ti, tf, lc, shm, u, v, vpot, &
                      cr, ci, nsamp, info) bind(C, name='ref_plnz')
    integer(c_int), value :: ndim, mp, lc, nsamp
    integer(c_int) :: np(3)
    real(c_double), value :: abstol, dt, ti, tf
    real(c_double) :: shm(3), u(*), v(*), vpot(*), cr(*), ci(*)
    integer(c_int) :: info
    character(len=8) :: dims
    integer :: n, mxdct, i
    integer, allocatable :: iwork(:)
    real(8), allocatable :: vabc(:), lnzvc(:,:), wu(:), wv(:), wp(:), work(:)
    real(8) :: xi(3), xf(3), var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    mxdct = max(n, nsamp + 2)
    allocate(iwork(5*mp + 16), vabc(n), lnzvc(mxdct, mp + 2), wu(n), wv(n), &
             wp(n), work(max(n, 5*mp + 16)))
    lnzvc = 0d0
    vabc = 1d0
    xi = 0d0; xf = 0d0; var = 0d0
    call dimstr(ndim, dims)
    call PLNZ(.false., .false., dims, '.NO     ', '.NONE   ', info, lc, mp, &
              mp, mp + 2, mp + 1, mxdct, n, abstol, dt, ti, tf, 1d30, np, &
              iwork, shm, vpot, vabc, u, v, xi, xf, wu, wv, wp, work, var, lnzvc)
    do i = 1, nsamp
      cr(i) = lnzvc(i, mp + 1)
      ci(i) = lnzvc(i, mp + 2)
    end do
    deallocate(iwork, vabc, lnzvc, wu, wv, wp, work)
  end subroutine ref_plnz

  ! PLNZ with ABSORB and TPABSOR '.SMOOTHW': PLNZAUX multiplies each step's
  ! state by the mask (sil.f:171-180).
  subroutine ref_plnz_smoothw(ndim, np, mp, abstol, dt, ti, tf, lc, shm, u, v, vpot, &
                              mask, cr, ci, nsamp, info) bind(C, name='ref_plnz_smoothw')
    integer(c_int), value :: ndim, mp, lc, nsamp
    integer(c_int) :: np(3)
    real(c_double), value :: abstol, dt, ti, tf
    real(c_double) :: shm(3), u(*), v(*), vpot(*), mask(*), cr(*), ci(*)
    integer(c_int) :: info
    character(len=8) :: dims
    integer :: n, mxdct, i
    integer, allocatable :: iwork(:)
    real(8), allocatable :: lnzvc(:,:), wu(:), wv(:), wp(:), work(:)
    real(8) :: xi(3), xf(3), var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    mxdct = max(n, nsamp + 2)
    allocate(iwork(5*mp + 16), lnzvc(mxdct, mp + 2), wu(n), wv(n), &
             wp(n), work(max(n, 5*mp + 16)))
    lnzvc = 0d0
    xi = 0d0; xf = 0d0; var = 0d0
    call dimstr(ndim, dims)
    call PLNZ(.true., .false., dims, '.NO     ', '.SMOOTHW', info, lc, mp, &
              mp, mp + 2, mp + 1, mxdct, n, abstol, dt, ti, tf, 1d30, np, &
              iwork, shm, vpot, mask, u, v, xi, xf, wu, wv, wp, work, var, lnzvc)
    do i = 1, nsamp
      cr(i) = lnzvc(i, mp + 1)
      ci(i) = lnzvc(i, mp + 2)
    end do
    deallocate(iwork, lnzvc, wu, wv, wp, work)
  end subroutine ref_plnz_smoothw

  ! EFC string of pulse kind 0..3 (ef.f).
  subroutine efcstr(efc, s)
    integer(c_int), value :: efc
    character(len=8) :: s
    s = '.NONE   '
    if (efc == 1) s = '.GAUS   '
    if (efc == 2) s = '.GGAUS  '
    if (efc == 3) s = '.SIN2   '
  end subroutine efcstr

  ! EF (ef.f): the field at time t.
  function ref_ef(efc, e0, t0, td, tp, omg, sni, kl, t) bind(C, name='ref_ef') result(r)
    integer(c_int), value :: efc, kl
    real(c_double), value :: e0, t0, td, tp, omg, sni, t
    real(c_double) :: r
    character(len=8) :: s
    integer :: k
    real(8) :: a0, a1, a2, a3, a4, a5, a6, EF
    external EF
    call efcstr(efc, s)
    k = kl
    a0 = e0; a1 = t0; a2 = td; a3 = tp; a4 = omg; a5 = sni; a6 = t
    r = EF(s, a0, a1, a2, a3, a4, a5, k, a6)
  end function ref_ef

  ! PPLNZ (pplnz.f:1-340) with PRTCRL='.NO', no projections (NPR = 0);
  ! veff receives VPOT as the reference leaves it (the last VEFF).
  subroutine ref_pplnz(ndim, np, mp, abstol, dt, ti, tf, lc, shm, u, v, vp, dm, &
                       efc, e0, t0, td, tp, omg, sni, kl, veff, cr, ci, nsamp, info) &
      bind(C, name='ref_pplnz')
    integer(c_int), value :: ndim, mp, lc, nsamp, efc, kl
    integer(c_int) :: np(3)
    real(c_double), value :: abstol, dt, ti, tf, e0, t0, td, tp, omg, sni
    real(c_double) :: shm(3), u(*), v(*), vp(*), dm(*), veff(*), cr(*), ci(*)
    integer(c_int) :: info
    character(len=8) :: dims, efcs
    integer :: n, mxdct, i, nd(3)
    integer, allocatable :: iwork(:)
    real(8), allocatable :: vabc(:), lnzvc(:,:), wu(:), wv(:), wp(:), work(:), vpot(:)
    real(8) :: xi(3), xf(3), xp(3), sh(3), var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    mxdct = max(n, nsamp + 2)
    allocate(iwork(5*mp + 16), vabc(n), lnzvc(mxdct, mp + 2), wu(n), wv(n), &
             wp(n), work(max(n, 5*mp + 16)), vpot(n))
    lnzvc = 0d0
    vabc = 1d0
    vpot = vp(1:n)
    xi = 0d0; xf = 0d0; xp = 0d0; sh = 0d0; var = 0d0; nd = 0
    call dimstr(ndim, dims)
    call efcstr(efc, efcs)
    call PPLNZ(.false., .false., .false., .false., .false., dims, efcs, '.NO     ', &
               '.NONE   ', info, lc, mp, mp, mp + 2, mp + 1, mxdct, n, 1, kl, 0, 0, &
               abstol, dt, ti, tf, 1d30, e0, t0, td, tp, omg, sni, np, nd, iwork, sh, &
               shm, vpot, vabc, u, v, xi, xf, xp, dm, wu, wv, wp, work, var, lnzvc)
    do i = 1, nsamp
      cr(i) = lnzvc(i, mp + 1)
      ci(i) = lnzvc(i, mp + 2)
    end do
    veff(1:n) = vpot(1:n)
    deallocate(iwork, vabc, lnzvc, wu, wv, wp, work, vpot)
  end subroutine ref_pplnz

  ! PPSOD (ppsod.f:1-548) with PRTCRL='.PARTIAL' (the reference stores its
  ! C(t) samples only when it prints), no projections (NPR = 0); absorb 0
  ! none, 1 .VOPTIC, 2 .SMOOTHW with vabc.  u, v receive U2, V2 (psi(TF)),
  ! veff the last VEFF.
  subroutine ref_ppsod(ndim, np, dt, ti, tf, lc, shm, u, v, vp, dm, efc, e0, t0, td, tp, &
                       omg, sni, kl, absorb, vabcin, veff, cr, ci, nsamp, info) &
      bind(C, name='ref_ppsod')
    integer(c_int), value :: ndim, lc, nsamp, efc, kl, absorb
    integer(c_int) :: np(3)
    real(c_double), value :: dt, ti, tf, e0, t0, td, tp, omg, sni
    real(c_double) :: shm(3), u(*), v(*), vp(*), dm(*), vabcin(*), veff(*), cr(*), ci(*)
    integer(c_int) :: info
    character(len=8) :: dims, efcs, tpab
    integer :: n, mxdct, i, nd(3)
    real(8), allocatable :: vabc(:), lnzvc(:,:), u1(:), u2(:), v1(:), v2(:), work(:), vpot(:)
    real(8) :: xi(3), xf(3), xp(3), sh(3), var(4), range(5, 7)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    mxdct = max(n, nsamp + 2)
    allocate(vabc(n), lnzvc(mxdct, 4), u1(n), u2(n), v1(n), v2(n), work(n), vpot(n))
    lnzvc = 0d0
    vabc = 1d0
    if (absorb /= 0) vabc = vabcin(1:n)
    tpab = '.NONE   '
    if (absorb == 1) tpab = '.VOPTIC '
    if (absorb == 2) tpab = '.SMOOTHW'
    vpot = vp(1:n)
    u1 = u(1:n); v1 = v(1:n); u2 = 0d0; v2 = 0d0
    xi = 0d0; xf = 0d0; xp = 0d0; sh = 0d0; var = 0d0; nd = 0; range = 0d0
    call dimstr(ndim, dims)
    call efcstr(efc, efcs)
    call PPSOD(absorb /= 0, .false., .false., .false., dims, efcs, '.PARTIAL', tpab, info, &
               lc, 2, 4, 3, mxdct, n, 1000000, 0, 0, 1, e0, t0, td, tp, omg, sni, kl, dt, &
               ti, tf, 1d30, np, nd, sh, shm, u1, u2, v1, v2, xi, xp, xf, dm, vpot, vabc, &
               work, var, lnzvc, .false., 0, range, .false.)
    do i = 1, nsamp
      cr(i) = lnzvc(i, 3)
      ci(i) = lnzvc(i, 4)
    end do
    u(1:n) = u2
    v(1:n) = v2
    veff(1:n) = vpot(1:n)
    deallocate(vabc, lnzvc, u1, u2, v1, v2, work, vpot)
  end subroutine ref_ppsod

  ! INIT_COND (init_cond.f): mode 0 .CALC, 1 .GETR, 2 .GETR2, 3 .GETC, 4 .GETC2.
  subroutine ref_init_cond(ndim, np, mode, xi, sh, x0, ra, rk, u, v) &
      bind(C, name='ref_init_cond')
    integer(c_int), value :: ndim, mode
    integer(c_int) :: np(3)
    real(c_double) :: xi(3), sh(3), x0(3), ra(3), rk(3), u(*), v(*)
    character(len=8) :: dims, ini
    integer :: n
    real(8), allocatable :: w1(:), w2(:)
    real(8) :: var(4)
    n = max(np(1), np(2), np(3)) + 1
    allocate(w1(n), w2(n))
    var = 0d0
    call dimstr(ndim, dims)
    select case (mode)
    case (1)
      ini = '.GETR   '
    case (2)
      ini = '.GETR2  '
    case (3)
      ini = '.GETC   '
    case (4)
      ini = '.GETC2  '
    case default
      ini = '.CALC   '
    end select
    call INIT_COND(dims, ini, np, rk, ra, sh, xi, x0, u, v, w1, w2, var)
    deallocate(w1, w2)
  end subroutine ref_init_cond

  ! ABSORBINGBC (absorb.f): kind 1 .SMOOTHW, 2 .SMRS, 3 .VOPTIC_LINEAR,
  ! 4 .VOPTIC_KOSLOFF, 5 .VOPTIC_POLIN.
  subroutine ref_absorbingbc(ndim, np, kind, al, ar, sh, xi, xf, voi, vabc, alpha) &
      bind(C, name='ref_absorbingbc')
    integer(c_int), value :: ndim, kind
    integer(c_int) :: np(3)
    real(c_double) :: al(3), ar(3), sh(3), xi(3), xf(3), voi(3), vabc(*), alpha
    character(len=8) :: dims
    character(len=16) :: tp
    integer :: n, i
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    call dimstr(ndim, dims)
    select case (kind)
    case (1)
      tp = '.SMOOTHW        '
    case (2)
      tp = '.SMRS           '
    case (3)
      tp = '.VOPTIC_LINEAR  '
    case (4)
      tp = '.VOPTIC_KOSLOFF '
    case default
      tp = '.VOPTIC_POLIN   '
    end select
    call ABSORBINGBC(dims, tp, n, np, al, ar, sh, xi, xf, voi, vabc, alpha)
  end subroutine ref_absorbingbc

  ! FFT (fft3d.f:1): Singleton multivariate in-place transform.
  subroutine ref_fft(a, b, ntot, n, nspan, isn) bind(C, name='ref_fft')
    real(c_double) :: a(*), b(*)
    integer(c_int), value :: ntot, n, nspan, isn
    call FFT(a, b, ntot, n, nspan, isn)
  end subroutine ref_fft

  ! FFTNT (cfft.f:6): radix-2 transform used by the spectrum.
  subroutine ref_fftnt(x, y, n, m, itype) bind(C, name='ref_fftnt')
    real(c_double) :: x(*), y(*)
    integer(c_int), value :: n, m, itype
    call FFTNT(x, y, n, m, itype)
  end subroutine ref_fftnt

  ! DIGT (dkeft.f:204): kinetic-energy grid.
  subroutine ref_digt(ndim, np, sh, sm, tc, b) bind(C, name='ref_digt')
    integer(c_int), value :: ndim
    integer(c_int) :: np(3)
    real(c_double) :: sh(3), sm(3), tc(*), b
    character(len=8) :: dims
    integer :: nd, n, i
    real(8) :: b1, b2
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    call dimstr(ndim, dims)
    call DIGT(dims, nd, n, b, b1, b2, np, sh, sm, tc)
  end subroutine ref_digt

  ! PSPOFFTAUX (spofft.f:3): one split-operator step, no absorber.
  subroutine ref_spo_step(ndim, np, b, cst1, cst2, u, v, tc, vpot) &
      bind(C, name='ref_spo_step')
    integer(c_int), value :: ndim
    integer(c_int) :: np(3)
    real(c_double), value :: b, cst1, cst2
    real(c_double) :: u(*), v(*), tc(*), vpot(*)
    integer :: n, i
    real(8) :: vabc(1)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    call PSPOFFTAUX(.false., '.NONE   ', ndim, n, b, b, b, cst1, cst2, np, &
                    u, v, tc, vpot, vabc)
  end subroutine ref_spo_step

  ! AU (au.f:3): 4th-order finite-difference H*u.
  subroutine ref_au(ndim, np, shm, vpot, u, hu) bind(C, name='ref_au')
    integer(c_int), value :: ndim
    integer(c_int) :: np(3)
    real(c_double) :: shm(3), vpot(*), u(*), hu(*)
    character(len=8) :: dims
    integer :: n, i
    real(8) :: var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    var = 0d0
    call dimstr(ndim, dims)
    call AU(dims, n, np, shm, vpot, u, hu, var)
  end subroutine ref_au

  ! EFFT4 (eigenergfft.f:1): energy through the FFT kinetic operator.
  function ref_efft4(ndim, np, u, v, tc, vpot) result(e) bind(C, name='ref_efft4')
    integer(c_int), value :: ndim
    integer(c_int) :: np(3)
    real(c_double) :: u(*), v(*), tc(*), vpot(*)
    real(c_double) :: e
    integer :: n, i
    real(8), allocatable :: wu(:), wv(:)
    real(8) :: b
    real(8), external :: EFFT4
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    b = 1d0/n
    allocate(wu(n), wv(n))
    e = EFFT4(ndim, n, np, b, u, v, wu, wv, tc, vpot)
    deallocate(wu, wv)
  end function ref_efft4

  ! LNZ (lnz.f:4) followed by nothing: returns alpha, beta and the basis.
  subroutine ref_lnz(ndim, np, m, nreort, shm, vpot, start, al, bt, basis, ldb) &
      bind(C, name='ref_lnz')
    integer(c_int), value :: ndim, m, nreort, ldb
    integer(c_int) :: np(3)
    real(c_double) :: shm(3), vpot(*), start(*), al(*), bt(*), basis(ldb, *)
    character(len=8) :: dims
    integer :: n, i
    real(8), allocatable :: v0(:), vi(:), work(:)
    real(8) :: var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    allocate(v0(n), vi(n), work(n))
    vi(1:n) = start(1:n)
    var = 0d0
    call dimstr(ndim, dims)
    call LNZ(.false., dims, n, m, nreort, nreort, ldb, np, shm, vpot, v0, vi, &
             work, al, bt, var, basis)
    deallocate(v0, vi, work)
  end subroutine ref_lnz

  ! DSTEVX (dstevx.f): all eigenpairs of the m x m tridiagonal (d, e).
  subroutine ref_dstevx(m, d, e, abstol, w, z, info) bind(C, name='ref_dstevx')
    integer(c_int), value :: m
    real(c_double), value :: abstol
    real(c_double) :: d(*), e(*), w(*), z(m, *)
    integer(c_int) :: info
    integer :: nf, il, iu
    integer, allocatable :: iwork(:)
    real(8), allocatable :: work(:)
    real(8) :: vl, vu
    ! The vendored DSTEVX drops LAPACK's IFAIL argument (dstevx.f:5-6).
    allocate(iwork(5*m), work(5*m))
    vl = 0d0; vu = 0d0; il = 1; iu = m
    call DSTEVX('V', 'A', m, d, e, vl, vu, il, iu, abstol, nf, w, z, m, work, &
                iwork, info)
    deallocate(iwork, work)
  end subroutine ref_dstevx

  ! HWHM (hwhm.f:1) Lorentzian damping of the assembled correlation.
  subroutine ref_hwhm(m, dt, tf, tal, de, u, v) bind(C, name='ref_hwhm')
    integer(c_int), value :: m
    real(c_double), value :: dt, tf, tal, de
    real(c_double) :: u(*), v(*)
    real(8) :: t2
    t2 = tal
    call HWHM(m, dt, tf, t2, de, u, v)
  end subroutine ref_hwhm

  ! SPECTRUMTD (spectrumtd.f:1) with the .SG window (iwin=1) or none (0).
  subroutine ref_spectrumtd(iwin, m, dt, tf, wp, de, u, v) &
      bind(C, name='ref_spectrumtd')
    integer(c_int), value :: iwin, m
    real(c_double), value :: dt, tf, wp, de
    real(c_double) :: u(*), v(*)
    real(8) :: de2
    de2 = de
    if (iwin == 1) then
      call SPECTRUMTD('.SG     ', m, dt, tf, wp, de2, u, v)
    else
      call SPECTRUMTD('.NONE   ', m, dt, tf, wp, de2, u, v)
    end if
  end subroutine ref_spectrumtd

  ! RAN (rans.f:3): the linear congruential start-vector generator.
  function ref_ran(nseed) result(r) bind(C, name='ref_ran')
    integer(c_int) :: nseed
    real(c_double) :: r
    real(8), external :: RAN
    r = RAN(nseed)
  end function ref_ran

  ! LANCZS (lanczs.f:1): eigenstates il..iu of H from the RAN start vector.
  subroutine ref_lanczs(ndim, np, il, iu, nreort, abstol, shm, vpot, eigvl, &
                        eigvc, ldv, info) bind(C, name='ref_lanczs')
    integer(c_int), value :: ndim, il, iu, nreort, ldv
    integer(c_int) :: np(3)
    real(c_double), value :: abstol
    real(c_double) :: shm(3), vpot(*), eigvl(*), eigvc(ldv, *)
    integer(c_int) :: info
    character(len=8) :: dims
    integer :: n, i, nseed, nr
    integer, allocatable :: iwk(:)
    real(8), allocatable :: al(:), bt(:), v0(:), vi(:), wk1(:), wk2(:), &
                            lnzvc(:,:), lanc(:,:), ev(:)
    real(8) :: var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    allocate(iwk(5*n + 16), al(n), bt(n), v0(n), vi(n), wk1(5*n), wk2(5*n), &
             lnzvc(ldv, max(nreort, iu - il + 1)), lanc(ldv, max(nreort, iu - il + 1)), &
             ev(n))
    var = 0d0
    nseed = 1
    nr = nreort
    call dimstr(ndim, dims)
    ! lanczs.f's LNZVC dummy is Krylov storage and its EIGVC dummy the
    ! tridiagonal eigenvectors (espec.f:918-920 passes them swapped).
    call LANCZS(.true., dims, il, iu, info, nr, nr, n, nseed, ldv, abstol, iwk, &
                np, al, bt, ev, shm, v0, vi, vpot, wk1, wk2, var, lnzvc, lanc)
    do i = 1, iu - il + 1
      eigvl(i) = ev(i)
      eigvc(1:n, i) = lnzvc(1:n, i)
    end do
    deallocate(iwk, al, bt, v0, vi, wk1, wk2, lnzvc, lanc, ev)
  end subroutine ref_lanczs

  ! LANCZSG (lanczsc.f:1-153): restarted 40-step Lanczos, eigenstates il..iu.
  ! The per-restart log goes to unit 6 (FORMAT 1031-1033), flushed on return.
  ! The reference STOPs when a state does not converge within maxint.
  subroutine ref_lanczsg(ndim, np, il, iu, maxint, abstol, shm, vpot, eigvl, &
                         eigvc, ldv, info) bind(C, name='ref_lanczsg')
    integer(c_int), value :: ndim, il, iu, maxint, ldv
    integer(c_int) :: np(3)
    real(c_double), value :: abstol
    real(c_double) :: shm(3), vpot(*), eigvl(*), eigvc(ldv, *)
    integer(c_int) :: info
    character(len=8) :: dims
    integer :: n, i, nseed, nr, inf
    integer, allocatable :: iwk(:)
    real(8), allocatable :: al(:), bt(:), v0(:), vi(:), wk1(:), wk2(:), &
                            lnzvc(:,:), ev(:), zz(:,:)
    real(8) :: var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    allocate(iwk(5*n + 256), al(n + 64), bt(n + 64), v0(n), vi(n + 64), wk1(5*n + 256), &
             wk2(5*n + 256), lnzvc(ldv, max(40, iu - il + 1)), &
             zz(ldv, max(40, iu - il + 1)), ev(n + 64))
    var = 0d0
    nseed = 1
    nr = 40
    inf = 0
    call dimstr(ndim, dims)
    ! espec.f:913-915 passes its EIGVC as LANCZSG's LNZVC (Krylov storage,
    ! the eigenvectors on return) and LANCZ as LANCZSG's EIGVC.
    call LANCZSG(.true., dims, il, iu, inf, nr, nr, n, nseed, ldv, abstol, iwk, &
                 maxint, np, al, bt, ev, shm, v0, vi, vpot, wk1, wk2, var, lnzvc, zz)
    flush(6)
    info = inf
    do i = 1, iu - il + 1
      eigvl(i) = ev(i)
      eigvc(1:n, i) = lnzvc(1:n, i)
    end do
    deallocate(iwk, al, bt, v0, vi, wk1, wk2, lnzvc, zz, ev)
  end subroutine ref_lanczsg

  ! MTRXDIAG (mtrxdiag.f:1-89): dense DSPEVX of the packed FD Hamiltonian.
  subroutine ref_mtrxdiag(ndim, np, il, iu, abstol, shm, vpot, eigvl, eigvc, ldv, &
                          info) bind(C, name='ref_mtrxdiag')
    integer(c_int), value :: ndim, il, iu, ldv
    integer(c_int) :: np(3)
    real(c_double), value :: abstol
    real(c_double) :: shm(3), vpot(*), eigvl(*), eigvc(ldv, *)
    integer(c_int) :: info
    character(len=8) :: dims
    integer :: n, i, inf
    integer, allocatable :: iwk(:)
    real(8), allocatable :: work(:), wk(:), ev(:), z(:,:)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    allocate(iwk(5*n), work(8*n), wk(8*n), ev(n), z(ldv, 20))
    inf = 0
    call dimstr(ndim, dims)
    call MTRXDIAG(dims, il, iu, inf, ldv, n, abstol, iwk, np, ev, shm, vpot, work, wk, z)
    info = inf
    do i = 1, iu - il + 1
      eigvl(i) = ev(i)
      eigvc(1:n, i) = z(1:n, i)
    end do
    deallocate(iwk, work, wk, ev, z)
  end subroutine ref_mtrxdiag

  ! S2PPSOD (s2ppsod.f:1-1143): two coupled surfaces, u/v/vp/dm [2n]
  ! (surface A then B), vabc [n]; prtcrl 0 .NO, 1 .PARTIAL, 2 .YES.  The
  ! table goes to unit 6 (flushed on return); u, v receive U2, V2 and
  ! u1o, v1o U1, V1.  No files: PRTVEFF, PRTEIGVC2 and PRTPULS are false.
  subroutine ref_s2ppsod(ndim, np, dt, ti, tf, lc, shm, xi1, sh1, sm1, aa, u, v, vp, dm, &
                         efc, e0, t0, td, tp, omg, sni, kl, absorb, vabcin, prtcrl, tol, &
                         u1o, v1o, info) bind(C, name='ref_s2ppsod')
    integer(c_int), value :: ndim, lc, efc, kl, absorb, prtcrl
    integer(c_int) :: np(3)
    real(c_double), value :: dt, ti, tf, xi1, sh1, sm1, aa, e0, t0, td, tp, omg, sni, tol
    real(c_double) :: shm(3), u(*), v(*), vp(*), dm(*), vabcin(*), u1o(*), v1o(*)
    integer(c_int) :: info
    character(len=8) :: dims, efcs, tpab, prt
    integer :: n, i, nd(3), inf
    real(8), allocatable :: vabc(:), lnzvc(:,:), u1(:), u2(:), v1(:), v2(:), work(:), vpot(:), &
                            dmm(:)
    real(8) :: xi(3), xf(3), xp(3), sh(3), sm(3), var(4)
    n = np(1)
    do i = 2, ndim
      n = n*np(i)
    end do
    allocate(vabc(n), lnzvc(n, 4), u1(2*n), u2(2*n), v1(2*n), v2(2*n), work(2*n), &
             vpot(2*n), dmm(2*n))
    lnzvc = 0d0
    vabc = 1d0
    if (absorb /= 0) vabc = vabcin(1:n)
    tpab = '.NONE   '
    if (absorb == 1) tpab = '.VOPTIC '
    if (absorb == 2) tpab = '.SMOOTHW'
    prt = '.NO     '
    if (prtcrl == 1) prt = '.PARTIAL'
    if (prtcrl == 2) prt = '.YES    '
    vpot = vp(1:2*n)
    dmm = dm(1:2*n)
    u1 = u(1:2*n); v1 = v(1:2*n); u2 = 0d0; v2 = 0d0
    xi = 0d0; xf = 0d0; xp = 0d0; sh = 0d0; sm = 1d0; var = 0d0; nd = 0
    xi(1) = xi1
    sh(1) = sh1
    sm(1) = sm1
    xf(1) = xi1 + (np(1) - 1)*sh1
    inf = 0
    call dimstr(ndim, dims)
    call efcstr(efc, efcs)
    call S2PPSOD(absorb /= 0, .false., .false., .false., dims, efcs, prt, tpab, inf, lc, 2, 4, &
                 3, n, n, 1000000, 0, 0, 1, kl, e0, t0, td, tp, omg, sni, dt, ti, tf, tol, &
                 1d0, 0d0, aa, np, nd, sh, shm, u1, u2, v1, v2, xi, xp, xf, sm, dmm, vpot, &
                 vabc, work, var, lnzvc)
    flush(6)
    info = inf
    u(1:2*n) = u2
    v(1:2*n) = v2
    u1o(1:2*n) = u1
    v1o(1:2*n) = v1
    deallocate(vabc, lnzvc, u1, u2, v1, v2, work, vpot, dmm)
  end subroutine ref_s2ppsod

end module espec_ref_harness
