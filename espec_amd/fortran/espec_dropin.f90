! Drop-in replacements for eSPec's hot-path subroutines.  Each keeps the
! name, argument list and meaning of the reference routine it replaces
! (cited per routine) so that linking libespec_f.so in place of those
! object files moves the work to the GPU; the arithmetic runs in
! libespec_hip.so through the ISO_C_BINDING interfaces of espec_capi.
!
! Differences a caller can see:
!   - no routine STOPs: failures come back in INFO (>= 2, espec_strerror)
!     or, for routines without INFO, as a message on stdout;
!   - the coupled-surface DIM variants return INFO = 6 (not on the
!     device); absorbing boundaries (ABSORB with TPABSOR .VOPTIC or
!     .SMOOTHW and VABC) run on the device in PSPOFFT and PSODFFT, and
!     PLNZ and PPLNZ apply the .SMOOTHW mask of PLNZAUX (sil.f:171-180);
!   - ESPEC_MODE in the environment selects the arithmetic (REF: the
!     reference's own rounding, the default; EXACT; MIXED).

! pspofft.f:1-150
subroutine PSPOFFT(ABSORB, DIM, PRTCRL, TPABSOR, INFO, LC, MP, MPI, MPR, MXDCT, N, DT, TI, TF, &
                   TOL, NP, SHM, SH, SM, U1, V1, XI, XF, TC, VPOT, VABC, LNZVC, WRK1, WRK2, VAR)
  use iso_c_binding
  use espec_capi
  implicit none
  logical :: ABSORB
  character(len=*) :: DIM, PRTCRL, TPABSOR
  integer :: INFO, LC, MP, MPI, MPR, MXDCT, N
  real(8) :: DT, TI, TF, TOL
  integer :: NP(*)
  real(8) :: SHM(*), SH(*), SM(*), U1(*), V1(*), XI(*), XF(*), TC(*), VPOT(*), VABC(*)
  real(8) :: LNZVC(MXDCT, *), WRK1(*), WRK2(*), VAR(*)
  type(c_ptr) :: plan
  integer :: ndim, flags, nst, maxs, ns, rc, k, i, step, chunk, kind
  real(8) :: b, t, cr, ci, ct, et, e0, fnr, fni, fnt, f0
  real(8), allocatable :: u0(:), v0(:), crs(:), cis(:), nrs(:)
  logical :: every, prt

  INFO = 0
  ndim = dim_of(DIM)
  if (ndim == 0) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  kind = absorb_of(ABSORB, TPABSOR)
  flags = mode_flags()
  rc = espec_digt(ndim, NP, SH, SM, TC, b, flags)
  nst = trip(TI + DT, TF + DT, DT)
  maxs = nst / LC + 1
  allocate(u0(N), v0(N), crs(maxs), cis(maxs), nrs(maxs))
  u0 = U1(1:N)
  v0 = V1(1:N)
  rc = espec_spo_plan_create_absorb(plan, ndim, NP, 1, DT, LC, SH, SM, VPOT, U1, V1, maxs, &
                                    flags, kind, VABC)
  if (rc /= 0) then
    INFO = rc
    return
  end if
  every = starts(PRTCRL, '.YES')
  prt = every .or. starts(PRTCRL, '.PARTIAL')
  if (prt) then
    t = TI
    call corr(N, u0, v0, U1, V1, cr, ci, fnr, fni)
    ct = sqrt(cr**2 + ci**2)
    rc = espec_efft4(ndim, NP, b, U1, V1, TC, VPOT, et, flags)
    e0 = et
    fnt = sqrt(fnr**2 + fni**2)
    f0 = fnt
    write(*,*)
    write(*,*) 'Partial auto-correlation function:'
    write(*,*) '=================================='
    write(*,1011) '*t/fs*', '*E/a.u.*', '*Cr(t)*', '*Ci(t)*', '*|C(t)|*', '*|C|*'
    write(*,1021) t, et, cr, ci, ct, fnt
    chunk = LC
    if (every) chunk = 1
    step = 0
    do while (step < nst)
      k = min(chunk, nst - step)
      rc = espec_spo_plan_run(plan, k)
      if (rc /= 0) exit
      do i = 1, k
        t = t + DT
      end do
      step = step + k
      if (every .or. mod(step, LC) == 0) then
        rc = espec_spo_plan_get(plan, 0, crs, cis, nrs, maxs, ns, U1, V1)
        call corr(N, u0, v0, U1, V1, cr, ci, fnr, fni)
        ct = sqrt(cr**2 + ci**2)
        rc = espec_efft4(ndim, NP, b, U1, V1, TC, VPOT, et, flags)
        fnt = sqrt(fnr**2 + fni**2)
        ! pspofft.f:137-142: with ABSORB the norm is checked as well
        if (abs((et - e0) / e0) * 1.0d2 > TOL) INFO = 1
        if (ABSORB .and. abs((fnt - f0) / f0) * 1.0d2 > TOL) INFO = 1
        write(*,1021) t, et, cr, ci, ct, fnt
      end if
    end do
  else
    rc = espec_spo_plan_run(plan, nst)
  end if
  if (rc == 0) rc = espec_spo_plan_get(plan, 0, crs, cis, nrs, maxs, ns, U1, V1)
  if (rc /= 0) then
    INFO = rc
  else
    do k = 1, ns
      LNZVC(k, MPR) = crs(k)
      LNZVC(k, MPI) = cis(k)
    end do
  end if
  rc = espec_spo_plan_destroy(plan)
  deallocate(u0, v0, crs, cis, nrs)
1011 format(/,7X,A6,7X,A8,8X,A7,7X,A7,6X,A8,6X,A5)
1021 format(1X,F14.8,3X,E12.6,2(4X,F10.7),4X,F10.8,4X,F8.6)
end subroutine PSPOFFT

! psodfft.f:1-211: psi(t-dt) returns in U1, V1 and psi(t) in U2, V2.
subroutine PSODFFT(ABSORB, DIM, PRTCRL, TPABSOR, INFO, LC, MP, MPI, MPR, MXDCT, N, DT, TI, TF, &
                   TOL, NP, SHM, SH, SM, U1, U2, V1, V2, XI, XF, TC, VPOT, VABC, WRKR, WRKI, VAR, &
                   LNZVC)
  use iso_c_binding
  use espec_capi
  implicit none
  logical :: ABSORB
  character(len=*) :: DIM, PRTCRL, TPABSOR
  integer :: INFO, LC, MP, MPI, MPR, MXDCT, N
  real(8) :: DT, TI, TF, TOL
  integer :: NP(*)
  real(8) :: SHM(*), SH(*), SM(*), U1(*), U2(*), V1(*), V2(*), XI(*), XF(*), TC(*), VPOT(*)
  real(8) :: VABC(*), WRKR(*), WRKI(*), VAR(*), LNZVC(MXDCT, *)
  type(c_ptr) :: plan
  integer :: ndim, flags, nst, maxs, ns, rc, k, i, step, chunk, kind
  real(8) :: b, t, cr, ci, ct, et, e0, fnr, fni, fnt, f0
  real(8), allocatable :: u0(:), v0(:), crs(:), cis(:), nrs(:)
  logical :: every, prt

  INFO = 0
  ndim = dim_of(DIM)
  if (ndim == 0) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  kind = absorb_of(ABSORB, TPABSOR)
  flags = mode_flags()
  rc = espec_digt(ndim, NP, SH, SM, TC, b, flags)
  ! the start-up (two half steps) is step 1; then DO T=TI+2DT,TF+DT,DT
  nst = 1 + trip(TI + 2 * DT, TF + DT, DT)
  maxs = nst / LC + 1
  allocate(u0(N), v0(N), crs(maxs), cis(maxs), nrs(maxs))
  u0 = U1(1:N)
  v0 = V1(1:N)
  U2(1:N) = u0
  V2(1:N) = v0
  rc = espec_sod_plan_create_absorb(plan, ndim, NP, 1, DT, LC, SH, SM, VPOT, U1, V1, maxs, &
                                    flags, kind, VABC)
  if (rc /= 0) then
    INFO = rc
    return
  end if
  every = starts(PRTCRL, '.YES')
  prt = every .or. starts(PRTCRL, '.PARTIAL')
  call corr(N, u0, v0, U2, V2, cr, ci, fnr, fni)
  rc = espec_efft4(ndim, NP, b, U2, V2, TC, VPOT, et, flags)
  e0 = et
  f0 = sqrt(fnr**2 + fni**2)
  if (prt) then
    t = TI
    ct = sqrt(cr**2 + ci**2)
    write(*,*)
    write(*,*) 'Partial auto-correlation function:'
    write(*,*) '=================================='
    write(*,1011) '*t/fs*', '*E/a.u.*', '*Cr(t)*', '*Ci(t)*', '*|C(t)|*', '*|C|*'
    write(*,1021) t, et, cr, ci, ct, f0
    chunk = LC
    if (every) chunk = 1
    step = 0
    do while (step < nst)
      k = min(chunk, nst - step)
      rc = espec_sod_plan_run(plan, k)
      if (rc /= 0) exit
      if (step == 0) then
        t = TI + DT
        do i = 2, k
          t = t + DT
        end do
      else
        do i = 1, k
          t = t + DT
        end do
      end if
      step = step + k
      if (every .or. mod(step, LC) == 0) then
        rc = espec_sod_plan_get(plan, 0, crs, cis, nrs, maxs, ns, U2, V2)
        call corr(N, u0, v0, U2, V2, cr, ci, fnr, fni)
        ct = sqrt(cr**2 + ci**2)
        rc = espec_efft4(ndim, NP, b, U2, V2, TC, VPOT, et, flags)
        fnt = sqrt(fnr**2 + fni**2)
        if (step > 1) then  ! psodfft.f:178-183 checks inside the leapfrog loop
          if (drift(et, e0, fnt, f0, TOL, .not. ABSORB)) INFO = 1
        end if
        write(*,1021) t, et, cr, ci, ct, fnt
      end if
    end do
  else
    rc = espec_sod_plan_run(plan, nst)
  end if
  if (rc == 0) rc = espec_sod_plan_get(plan, 0, crs, cis, nrs, maxs, ns, U2, V2)
  if (rc == 0) rc = espec_sod_plan_get_prev(plan, 0, U1, V1)
  if (rc /= 0) then
    INFO = rc
  else
    do k = 1, ns
      LNZVC(k, MPR) = crs(k)
      LNZVC(k, MPI) = cis(k)
    end do
    ! psodfft.f:194-204: final drift check
    call corr(N, u0, v0, U2, V2, cr, ci, fnr, fni)
    rc = espec_efft4(ndim, NP, b, U2, V2, TC, VPOT, et, flags)
    fnt = sqrt(fnr**2 + fni**2)
    if (drift(et, e0, fnt, f0, TOL, .not. ABSORB)) INFO = 1
  end if
  rc = espec_sod_plan_destroy(plan)
  deallocate(u0, v0, crs, cis, nrs)
1011 format(/,7X,A6,7X,A8,8X,A7,7X,A7,6X,A8,6X,A5)
1021 format(1X,F14.8,3X,E12.6,2(4X,F10.7),4X,F10.8,4X,F8.6)
contains
  ! psodfft.f:178-183, 201-206: energy drift, and the norm's only without
  ! an absorber.
  logical function drift(et, e0, fnt, f0, tol, norm)
    real(8), intent(in) :: et, e0, fnt, f0, tol
    logical, intent(in) :: norm
    drift = abs((et - e0) / e0) * 1.0d2 > tol
    if (norm) drift = drift .or. abs((fnt - f0) / f0) * 1.0d2 > tol
  end function
end subroutine PSODFFT

! plnz.f:1-156
subroutine PLNZ(ABSORB, CHANGE, DIM, PRTCRL, TPABSOR, INFO, LC, LMTREORT, MP, MPI, MPR, MXDCT, &
                N, ABSTOL, DT, TI, TF, TOL, NP, IWORK, SHM, VPOT, VABC, U, V, XI, XF, WU, WV, WP, &
                WORK, VAR, LNZVC)
  use iso_c_binding
  use espec_capi
  implicit none
  logical :: ABSORB, CHANGE
  character(len=*) :: DIM, PRTCRL, TPABSOR
  integer :: INFO, LC, LMTREORT, MP, MPI, MPR, MXDCT, N
  real(8) :: ABSTOL, DT, TI, TF, TOL
  integer :: NP(*), IWORK(*)
  real(8) :: SHM(*), VPOT(*), VABC(*), U(*), V(*), XI(*), XF(*), WU(*), WV(*), WP(*), WORK(*)
  real(8) :: VAR(*), LNZVC(MXDCT, *)
  type(c_ptr) :: plan
  integer :: ndim, flags, nst, maxs, ns, rc, k, i, step, chunk
  real(8) :: t, cr, ci, ct, er, ei, et, e0, fnr, fni, fnt, f0
  real(8), allocatable :: u0(:), v0(:), crs(:), cis(:), nrs(:)
  logical :: every, prt

  INFO = 0
  ndim = dim_of(DIM)
  if (ndim == 0) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  flags = mode_flags()
  nst = trip(TI + DT, TF + DT, DT)
  maxs = nst / LC + 1
  allocate(u0(N), v0(N), crs(maxs), cis(maxs), nrs(maxs))
  u0 = U(1:N)
  v0 = V(1:N)
  rc = espec_sil_plan_create(plan, ndim, NP, 1, MP, ABSTOL, DT, LC, SHM, VPOT, U, V, maxs, flags)
  if (rc /= 0) then
    INFO = rc
    return
  end if
  ! PLNZAUX masks the state for ABSORB with .SMOOTHW (sil.f:171-180)
  if (absorb_of(ABSORB, TPABSOR) == ESPEC_ABSORB_SMOOTHW) then
    rc = espec_sil_plan_set_absorb(plan, ESPEC_ABSORB_SMOOTHW, VABC)
    if (rc /= 0) then
      INFO = rc
      rc = espec_sil_plan_destroy(plan)
      return
    end if
  end if
  every = starts(PRTCRL, '.YES')
  prt = every .or. starts(PRTCRL, '.PARTIAL')
  call corr(N, u0, v0, U, V, cr, ci, fnr, fni)
  rc = espec_eigenerg(ndim, NP, SHM, U, VPOT, er)
  rc = espec_eigenerg(ndim, NP, SHM, V, VPOT, ei)
  et = er + ei
  e0 = et
  f0 = fnr
  if (prt) then
    t = TI
    ct = cr**2 + ci**2
    write(*,*)
    write(*,*) 'Partial auto-correlation function:'
    write(*,*) '=================================='
    write(*,1011) '*t*', '*Energy*', '*Cr(t)*', '*Ci(t)*', '*|C(t)|^2*', '*|C|*'
    write(*,1021) t, et, cr, ci, ct, f0
    chunk = LC
    if (every) chunk = 1
    step = 0
    do while (step < nst)
      k = min(chunk, nst - step)
      rc = espec_sil_plan_run(plan, k)
      if (rc /= 0) exit
      do i = 1, k
        t = t + DT
      end do
      step = step + k
      if (every .or. mod(step, LC) == 0) then
        rc = espec_sil_plan_get(plan, 0, crs, cis, nrs, maxs, ns, U, V)
        if (rc /= 0) exit
        call corr(N, u0, v0, U, V, cr, ci, fnr, fni)
        ct = cr**2 + ci**2
        rc = espec_eigenerg(ndim, NP, SHM, U, VPOT, er)
        rc = espec_eigenerg(ndim, NP, SHM, V, VPOT, ei)
        et = er + ei
        fnt = fnr
        if (abs((et - e0) / e0) * 1.0d2 > TOL .or. abs((fnt - f0) / f0) * 1.0d2 > TOL) INFO = 1
        write(*,1021) t, et, cr, ci, ct, fnt
      end if
    end do
  else
    rc = espec_sil_plan_run(plan, nst)
  end if
  if (rc == 0) rc = espec_sil_plan_get(plan, 0, crs, cis, nrs, maxs, ns, U, V)
  if (rc /= 0) then
    if (rc == 5) write(*,'(A)') '<<<>>> Diagonalization error <<<>>>'
    INFO = rc
  else
    do k = 1, ns
      LNZVC(k, MPR) = crs(k)
      LNZVC(k, MPI) = cis(k)
    end do
    rc = espec_eigenerg(ndim, NP, SHM, U, VPOT, er)
    rc = espec_eigenerg(ndim, NP, SHM, V, VPOT, ei)
    et = er + ei
    call corr(N, u0, v0, U, V, cr, ci, fnr, fni)
    if (abs((et - e0) / e0) * 1.0d2 > TOL .or. abs((fnr - f0) / f0) * 1.0d2 > TOL) INFO = 1
  end if
  rc = espec_sil_plan_destroy(plan)
  deallocate(u0, v0, crs, cis, nrs)
1011 format(/,7X,A6,7X,A8,8X,A7,7X,A7,4X,A10,6X,A5)
1021 format(1X,F14.8,3X,E12.6,2(4X,F10.7),4X,F10.8,4X,F8.6)
end subroutine PLNZ

! pplnz.f:1-340: PLNZ steps with VEFF = VP - DM EF(T) (chpot.f:58-62) at
! each step's end time.  The movie, potential and pulse files (PRTVEFF,
! PRTEIGVC2, PRTPULS, PRTCRL '.LASTWP') are not written: INFO = 6.
subroutine PPLNZ(ABSORB, CHANGE, PRTVEFF, PRTEIGVC2, PRTPULS, DIM, EFC, PRTCRL, TPABSOR, INFO, &
                 LC, LMTREORT, MP, MPI, MPR, MXDCT, N, NSHOT, KL, NPR, KP, ABSTOL, DT, TI, TF, &
                 TOL, E0, T0, TD, TP, OMG, SNI, NP, ND, IWORK, SH, SHM, VPOT, VABC, U, V, XI, &
                 XF, XP, DM, WU, WV, WP, WORK, VAR, LNZVC)
  use iso_c_binding
  use espec_capi
  implicit none
  logical :: ABSORB, CHANGE, PRTVEFF, PRTEIGVC2, PRTPULS
  character(len=*) :: DIM, EFC, PRTCRL, TPABSOR
  integer :: INFO, LC, LMTREORT, MP, MPI, MPR, MXDCT, N, NSHOT, KL, NPR, KP
  real(8) :: ABSTOL, DT, TI, TF, TOL, E0, T0, TD, TP, OMG, SNI
  integer :: NP(*), ND(*), IWORK(*)
  real(8) :: SH(*), SHM(*), VPOT(*), VABC(*), U(*), V(*), XI(*), XF(*), XP(*), DM(*)
  real(8) :: WU(*), WV(*), WP(*), WORK(*), VAR(*), LNZVC(MXDCT, *)
  type(c_ptr) :: plan
  integer :: ndim, flags, nst, maxs, ns, rc, k, i, j, step, chunk, efk, nc
  real(8) :: t, cr, ci, er, ei, et, et0, fnr, fni, fnt, f0, ef
  real(8) :: ct(7)
  real(8), allocatable :: u0(:), v0(:), vp(:), crs(:), cis(:), nrs(:), efield(:)
  logical :: every, prt

  INFO = 0
  ndim = dim_of(DIM)
  efk = pulse_of(EFC)
  if (ndim == 0 .or. PRTVEFF .or. PRTEIGVC2 .or. PRTPULS .or. starts(PRTCRL, '.LASTWP')) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  if (efk < 0) then
    write(*,'(A)') '<<<>>> Desired laser pulse wasn''t found. <<<>>>'
    INFO = ESPEC_ERR_ARG
    return
  end if
  flags = mode_flags()
  nst = trip(TI + DT, TF + DT / 2, DT)  ! pplnz.f:190
  maxs = nst / LC + 1
  allocate(u0(N), v0(N), vp(N), crs(maxs), cis(maxs), nrs(maxs), efield(max(nst, 1)))
  u0 = U(1:N)
  v0 = V(1:N)
  vp = VPOT(1:N)
  t = TI
  do k = 1, nst
    t = t + DT
    efield(k) = espec_ef(efk, E0, T0, TD, TP, OMG, SNI, KL, t)
  end do
  rc = espec_sil_plan_create(plan, ndim, NP, 1, MP, ABSTOL, DT, LC, SHM, vp, U, V, maxs, flags)
  if (rc == 0) rc = espec_sil_plan_set_pulse(plan, DM, efield, nst)
  ! PPLNZ hands ABSORB to PLNZAUX (pplnz.f:208), which masks for .SMOOTHW
  if (rc == 0 .and. absorb_of(ABSORB, TPABSOR) == ESPEC_ABSORB_SMOOTHW) &
    rc = espec_sil_plan_set_absorb(plan, ESPEC_ABSORB_SMOOTHW, VABC)
  if (rc /= 0) then
    INFO = rc
    return
  end if
  every = starts(PRTCRL, '.YES')
  prt = every .or. starts(PRTCRL, '.PARTIAL')
  t = TI
  step = 0
  if (prt) then  ! pplnz.f:100-160
    call project(0)
    rc = espec_eigenerg(ndim, NP, SHM, U, VPOT, er)
    rc = espec_eigenerg(ndim, NP, SHM, V, VPOT, ei)
    et = er + ei
    et0 = et
    f0 = fnr**2
    write(*,*)
    if (ndim == 1) then
      write(*,*) 'Auto-correlation function:'
    else
      write(*,*) 'Partial auto-correlation function:'
    end if
    write(*,*) '=================================='
    write(*,1011) ('*|P', j - 1, '|^2*', j = 1, NPR), '*|c', 0, '|^2*'
    write(*,1021) t, et, et0, cr, ci, (ct(j), j = 1, NPR), f0
    chunk = LC
    if (every) chunk = 1
    do while (step < nst)
      k = min(chunk, nst - step)
      rc = espec_sil_plan_run(plan, k)
      if (rc /= 0) exit
      do i = 1, k
        t = t + DT
      end do
      step = step + k
      if (every .or. mod(step, LC) == 0) then
        rc = espec_sil_plan_get(plan, 0, crs, cis, nrs, maxs, ns, U, V)
        if (rc /= 0) exit
        ef = efield(step)
        do i = 1, N
          VPOT(i) = vp(i) - DM(i) * ef
        end do
        call project(1)
        rc = espec_eigenerg(ndim, NP, SHM, U, VPOT, er)
        rc = espec_eigenerg(ndim, NP, SHM, V, VPOT, ei)
        et = er + ei
        rc = espec_eigenerg(ndim, NP, SHM, U, vp, er)
        rc = espec_eigenerg(ndim, NP, SHM, V, vp, ei)
        et0 = er + ei
        fnt = fnr
        if (.not. ABSORB .and. abs((fnt - f0) / f0) * 1.0d2 > TOL) INFO = 1
        write(*,1021) t, et, et0, cr, ci, (ct(j), j = 1, NPR), fnt
      end if
    end do
  else
    rc = espec_sil_plan_run(plan, nst)
  end if
  if (rc == 0) rc = espec_sil_plan_get(plan, 0, crs, cis, nrs, maxs, ns, U, V)
  if (rc /= 0) then
    if (rc == 5) write(*,'(A)') '<<<>>> Diagonalization error <<<>>>'
    INFO = rc
  else
    ! C(TI) is stored only with PRTCRL .YES/.PARTIAL (pplnz.f:130-134)
    nc = 1
    if (.not. prt) nc = 2
    do k = nc, ns
      LNZVC(k, MPR) = crs(k)
      LNZVC(k, MPI) = cis(k)
    end do
    ! VPOT is left holding the last step's VEFF (chpot.f:58-62)
    if (nst > 0) then
      ef = efield(nst)
      do i = 1, N
        VPOT(i) = vp(i) - DM(i) * ef
      end do
    end if
  end if
  rc = espec_sil_plan_destroy(plan)
  deallocate(u0, v0, vp, crs, cis, nrs, efield)
1011 format(/,8X,'*t/fs*',5X,'*ET/a.u.*',6X,'*E0/a.u.*',6X,'*Pr(t)*',4X,'*Pi(t)*',3X, &
            7(A3,I1,A4,3X))
1021 format(1X,F14.6,3X,2(E12.6,3X),2(F8.5,3X),10(F8.6,3X))
contains
  ! C(t), the populations |<W_k|psi>|^2 of LNZVC(:,KP+k) and the norm
  ! (pplnz.f:101-140); pass 0 is the initial state.
  subroutine project(pass)
    integer, intent(in) :: pass
    integer :: q, ii
    real(8) :: pr, pi
    do q = 2, min(NPR, 7)
      pr = 0d0
      pi = 0d0
      do ii = 1, N
        pr = pr + LNZVC(ii, KP + q - 1) * U(ii)
      end do
      do ii = 1, N
        pi = pi + LNZVC(ii, KP + q - 1) * V(ii)
      end do
      ct(q) = pr**2 + pi**2
    end do
    call corr(N, u0, v0, U, V, cr, ci, fnr, fni)
    ct(1) = cr**2 + ci**2
    if (pass == 0) then
      LNZVC(1, MPR) = cr
      LNZVC(1, MPI) = ci
    end if
  end subroutine
end subroutine PPLNZ

! lanczs.f:1-190: eigenvectors into EIGVC(:,1..IU-IL+1) and LNZVC(:,1..IU-IL+1).
subroutine LANCZS(CHANGE, DIM, IL, IU, INFO, NREORT, LMTREORT, N, NSEED, MXDCT, ABSTOL, IWK, &
                  NP, AL, BT, EIGVL, SHM, V0, VI, VPOT, WK1, WK2, VAR, LNZVC, EIGVC)
  use espec_capi
  implicit none
  logical :: CHANGE
  character(len=*) :: DIM
  integer :: IL, IU, INFO, NREORT, LMTREORT, N, NSEED, MXDCT
  real(8) :: ABSTOL
  integer :: IWK(*), NP(*)
  real(8) :: AL(*), BT(*), EIGVL(*), SHM(*), V0(*), VI(*), VPOT(*), WK1(*), WK2(*), VAR(*)
  real(8) :: LNZVC(MXDCT, *), EIGVC(MXDCT, *)
  integer :: ndim, ns, k, nc, rc
  character(len=2) :: chnc, chaux
  real(8), allocatable :: vc(:)

  INFO = 0
  ndim = dim_of(DIM)
  if (ndim == 0) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  if (NREORT > LMTREORT) then  ! lanczs.f:68-74
    write(*,*) 'Error NREORT > LMTREORT'
    write(*,*) 'NREORT=', NREORT, '  LMTREORT=', LMTREORT
    NREORT = LMTREORT
    write(*,*) 'Changing NREORT to correct value!'
    write(*,*) 'NREORT=', NREORT, '  LMTREORT=', LMTREORT
  end if
  NSEED = 1
  ns = IU - IL + 1
  allocate(vc(N * ns))
  rc = espec_lanczs(ndim, NP, IL, IU, NREORT, ABSTOL, SHM, VPOT, EIGVL, vc)
  if (rc /= 0) INFO = rc
  nc = IL - 2
  do k = 1, ns
    nc = nc + 1
    write(chnc, '(I2)') nc
    if (chnc(2:2) == '1') then
      chaux = 'st'
    else if (chnc(2:2) == '2') then
      chaux = 'nd'
    else if (chnc(2:2) == '3') then
      chaux = 'rd'
    else
      chaux = 'th'
    end if
    if (nc < 10 .and. INFO == 0) then
      write(*,1031) nc, chaux, EIGVL(k)
    else if (INFO == 0) then
      write(*,1032) nc, chaux, EIGVL(k)
    else
      write(*,1033) nc, chaux
    end if
    EIGVC(1:N, k) = vc((k - 1) * N + 1:k * N)
    LNZVC(1:N, k) = vc((k - 1) * N + 1:k * N)
  end do
  deallocate(vc)
1031 format(4X,I1,A2,1X,'state has been calculating =',E16.8,'/a.u..')
1032 format(3X,I2,A2,1X,'state has been calculating =',E16.8,'/a.u..')
1033 format(7X,I2,A2,1X,'state hasn`t been calculating correctly!')
end subroutine LANCZS

! lanczsc.f:1-153: restarted 40-step Lanczos for eigenstates IL..IU on the
! device (NREORT is set to 40 as the reference does).  The per-restart lines
! (FORMAT 1031-1033) are printed after the device returns; a state that does
! not converge within MAXINT restarts stops the program as the reference does.
! As at espec.f:913-915, LNZVC holds the eigenvectors on return and EIGVC
! the eigenvector columns, whose rows 1..40 hold for earlier states what the
! last restart's DSTEVX wrote in ESPEC_MODE=REF.
subroutine LANCZSG(CHANGE, DIM, IL, IU, INFO, NREORT, LMTREORT, N, NSEED, MXDCT, ABSTOL, IWK, &
                   MAXINT, NP, AL, BT, EIGVL, SHM, V0, VI, VPOT, WK1, WK2, VAR, LNZVC, EIGVC)
  use espec_capi
  implicit none
  logical :: CHANGE
  character(len=*) :: DIM
  integer :: IL, IU, INFO, NREORT, LMTREORT, N, NSEED, MXDCT, MAXINT
  real(8) :: ABSTOL
  integer :: IWK(*), NP(*)
  real(8) :: AL(*), BT(*), EIGVL(*), SHM(*), V0(*), VI(*), VPOT(*), WK1(*), WK2(*), VAR(*)
  real(8) :: LNZVC(MXDCT, *), EIGVC(MXDCT, *)
  integer :: ndim, ns, k, rc, nl, logcap, kk
  real(8), allocatable :: vc(:), lg(:)

  INFO = 0
  ndim = dim_of(DIM)
  if (ndim == 0 .or. ndim == 3) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  if (NREORT > LMTREORT) then  ! lanczsc.f:69-75
    write(*,*) 'Error NREORT > LMTREORT'
    write(*,*) 'NREORT=', NREORT, '  LMTREORT=', LMTREORT
    NREORT = LMTREORT
    write(*,*) 'Changing NREORT to correct value!'
    write(*,*) 'NREORT=', NREORT, '  LMTREORT=', LMTREORT
  end if
  NSEED = 1
  NREORT = 40
  ns = IU - IL + 1
  logcap = ns * (MAXINT + 2)
  allocate(vc(N * ns), lg(4 * logcap))
  nl = 0
  rc = espec_lanczsg(ndim, NP, IL, IU, MAXINT, ABSTOL, SHM, VPOT, EIGVL, vc, lg, logcap, nl, &
                     mode_flags())
  do k = 1, min(nl, logcap)
    kk = nint(lg(4 * k - 2))
    if (kk < 10) then
      write(*,1031) kk, lg(4 * k - 1), lg(4 * k)
    else if (kk < 100) then
      write(*,1032) kk, lg(4 * k - 1), lg(4 * k)
    else if (kk < 1000) then
      write(*,1033) kk, lg(4 * k - 1), lg(4 * k)
    else
      write(*,*) '<<<>>> Too many interactions <<<>>>'
      write(*,*) 'Stopping program!'
      stop
    end if
  end do
  if (rc == ESPEC_ERR_DIAG) then
    write(*,*) '<<<>>> Eigenvector didn''t achieve the ', 'convergence <<<>>>'
    write(*,*) 'Stopping eSPec!'
    stop
  else if (rc /= 0) then
    write(*,*) '<<<>>> LANCZSG failed on the device, code ', rc
    stop
  end if
  do k = 1, ns
    EIGVC(1:N, k) = vc((k - 1) * N + 1:k * N)
    LNZVC(1:N, k) = vc((k - 1) * N + 1:k * N)
  end do
  deallocate(vc, lg)
1031 format(5X,'Interaction',1X,I1,';',1X,'Energy =',E16.8,'/a.u.; ','WF converg. =',E16.8,';')
1032 format(5X,'Interaction',1X,I2,';',1X,'Energy =',E16.8,'/a.u.; ','WF converg. =',E16.8,';')
1033 format(5X,'Interaction',1X,I3,';',1X,'Energy =',E16.8,'/a.u.; ','WF converg. =',E16.8,';')
end subroutine LANCZSG

! mtrxdiag.f:1-89: eigenpairs IL..IU of the dense FD Hamiltonian (the matrix
! of ap1d.f / ap2d.f) by a device syevx: bisection to ABSTOL, inverse
! iteration.  Eigenvectors agree with DSPEVX's up to sign.  DIM .2DCT and
! .3D return INFO = ESPEC_ERR_UNSUPPORTED (the reference builds no matrix
! for .3D); N above 16384 returns it too (the dense matrix alone is 2 GiB).
subroutine MTRXDIAG(DIM, IL, IU, INFO, MXDCT, N, ABSTOL, IWORK, NP, EIGVL, SHM, VPOT, WORK, WK, &
                    EIGVC)
  use espec_capi
  implicit none
  character(len=*) :: DIM
  integer :: IL, IU, INFO, MXDCT, N
  real(8) :: ABSTOL
  integer :: IWORK(*), NP(*)
  real(8) :: EIGVL(*), SHM(*), VPOT(*), WORK(*), WK(*)
  real(8) :: EIGVC(MXDCT, *)
  integer :: ndim, m
  INFO = 0
  ndim = dim_of(DIM)
  if (ndim == 0 .or. ndim == 3) then
    write(*,'(A)') '<<<>>> Matrix generation error <<<>>>'
    write(*,*) DIM
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  m = 0
  INFO = espec_mtrxdiag(ndim, NP, IL, IU, ABSTOL, SHM, VPOT, EIGVL, EIGVC, MXDCT, m)
end subroutine MTRXDIAG

! fft3d.f:1
! absorb.f:1-130: VABC of TPABSOR on the grid, on the device (N is not
! read; the grid is NP).  DIM and TPABSOR are matched on their leading
! characters in the reference's order (absorb.f:62-156), so '.SMOOTHW_au'
! or a TPABSOR that carries a trailing comment select the same shape.
! Where the reference STOPs (a 1-D shape it does not know, the 2-D
! .VOPTIC, 3-D, an unknown DIM) the drop-in stops too, as it does when
! the device fails; a 2-D TPABSOR that is neither .SMOOTHW nor .VOPTIC
! leaves VABC as it was, as the reference does.
subroutine ABSORBINGBC(DIM, TPABSOR, N, NP, AL, AR, SH, XI, XF, VOI, VABC, ALPHA)
  use espec_capi
  implicit none
  character(len=*) :: DIM, TPABSOR
  integer :: N
  integer :: NP(*)
  real(8) :: AL(*), AR(*), SH(*), XI(*), XF(*), VOI(*), VABC(*), ALPHA
  integer :: rc, kind, nd
  kind = 0
  if (starts(DIM, '.1D')) then
    nd = 1
    if (starts(TPABSOR, '.SMOOTHW')) then
      kind = ESPEC_ABG_SMOOTHW
    else if (starts(TPABSOR, '.SMRS')) then
      kind = ESPEC_ABG_SMRS
    else if (starts(TPABSOR, '.VOPTIC_LINEAR')) then
      kind = ESPEC_ABG_VOPTIC_LINEAR
    else if (starts(TPABSOR, '.VOPTIC_KOSLOFF')) then
      kind = ESPEC_ABG_VOPTIC_KOSLOFF
    else if (starts(TPABSOR, '.VOPTIC_POLIN')) then
      kind = ESPEC_ABG_VOPTIC_POLIN
    else
      write(*,*)
      error stop 'ABSORBINGBC: no 1-D absorber of this TPABSOR'
    end if
  else if (starts(DIM, '.2D')) then
    nd = 2
    if (starts(TPABSOR, '.SMOOTHW')) then
      kind = ESPEC_ABG_SMOOTHW
    else if (starts(TPABSOR, '.VOPTIC')) then
      write(*,*) '<<<>>> Absorbing bondary conditions .VOPTIC '
      write(*,*) '          is not implemente in the code <<<>>>'
      error stop
    else
      return
    end if
  else if (starts(DIM, '.3D')) then
    write(*,*) '<<<>>> Absorbing bondary conditions is not '
    write(*,*) '      implemente for the 3-dimensional code <<<>>>'
    error stop
  else
    write(*,*) '<<<>>> Dimension error <<<>>>'
    error stop
  end if
  rc = espec_absorbingbc(nd, NP, kind, AL, AR, SH, XI, XF, VOI, VABC, ALPHA)
  if (rc /= 0) then
    write(*,'(A,I3)') '<<<>>> ABSORBINGBC on the device failed, code ', rc
    error stop
  end if
end subroutine ABSORBINGBC

subroutine FFT(A, B, NTOT, N, NSPAN, ISN)
  use espec_capi
  implicit none
  integer :: NTOT, N, NSPAN, ISN
  real(8) :: A(*), B(*)
  integer :: rc
  rc = espec_fft(A, B, NTOT, N, NSPAN, ISN, mode_flags())
  if (rc /= 0) write(*,'(A,I3)') '<<<>>> FFT on the device failed, code ', rc
end subroutine FFT

! dkeft.f:1-144 (the 1-D grid, and per-axis sums for 2-D/3-D)
subroutine DIGT(DIM, NDIM, N, B, B1, B2, NP, SH, SM, TC)
  use espec_capi
  implicit none
  character(len=*) :: DIM
  integer :: NDIM, N
  real(8) :: B, B1, B2
  integer :: NP(*)
  real(8) :: SH(*), SM(*), TC(*)
  integer :: rc
  NDIM = dim_of(DIM)
  if (NDIM == 0) return
  rc = espec_digt(NDIM, NP, SH, SM, TC, B, mode_flags())
  B1 = SH(1)
  B2 = 1.0d0 / (NP(1) * SH(1))
end subroutine DIGT

! au.f:3-84
subroutine AU(DIM, N, NP, SHM, VPOT, U, WORK, VAR)
  use espec_capi
  implicit none
  character(len=*) :: DIM
  integer :: N, NP(*)
  real(8) :: SHM(*), VPOT(*), U(*), WORK(*), VAR(*)
  integer :: rc
  rc = espec_au(dim_of(DIM), NP, SHM, VPOT, U, WORK)
  if (rc /= 0) write(*,'(A,I3)') '<<<>>> AU on the device failed, code ', rc
end subroutine AU

! au1d.f:1
subroutine AU_1D(N, SHM, VPOT, US, AUS)
  use espec_capi
  implicit none
  integer :: N
  real(8) :: SHM(*), VPOT(*), US(*), AUS(*)
  integer :: rc, np(1)
  np(1) = N
  rc = espec_au(1, np, SHM, VPOT, US, AUS)
end subroutine AU_1D

! au2di.f:1
subroutine AU_2D(NP, SHM, VPOT, US, AUS)
  use espec_capi
  implicit none
  integer :: NP(*)
  real(8) :: SHM(*), VPOT(*), US(*), AUS(*)
  integer :: rc
  rc = espec_au(2, NP, SHM, VPOT, US, AUS)
end subroutine AU_2D

! au3d.f:1
subroutine AU_3D(NP, SHM, VPOT, US, AUS)
  use espec_capi
  implicit none
  integer :: NP(*)
  real(8) :: SHM(*), VPOT(*), US(*), AUS(*)
  integer :: rc
  rc = espec_au(3, NP, SHM, VPOT, US, AUS)
end subroutine AU_3D

! fcorrel.f:1-68
real(8) function FCORREL(N, U1, U2)
  use espec_capi
  implicit none
  integer :: N
  real(8) :: U1(*), U2(*)
  integer :: rc
  rc = espec_fcorrel(N, U1, U2, FCORREL)
end function FCORREL

! ecnorm.f:1-68
real(8) function ECNORM(N, US)
  use espec_capi
  implicit none
  integer :: N
  real(8) :: US(*)
  integer :: rc
  rc = espec_ecnorm(N, US, ECNORM)
end function ECNORM

! eigenergfft.f:1-103
real(8) function EFFT4(NDIM, N, NP, B, U, V, WU, WV, TC, VPOT)
  use espec_capi
  implicit none
  integer :: NDIM, N, NP(*)
  real(8) :: B, U(*), V(*), WU(*), WV(*), TC(*), VPOT(*)
  integer :: rc
  rc = espec_efft4(NDIM, NP, B, U, V, TC, VPOT, EFFT4, mode_flags())
end function EFFT4

! eigenerg.f:1-70
real(8) function EIGENERG(DIM, N, NP, SH, U1, VPOT, WORK, VAR)
  use espec_capi
  implicit none
  character(len=*) :: DIM
  integer :: N, NP(*)
  real(8) :: SH(*), U1(*), VPOT(*), WORK(*), VAR(*)
  integer :: rc
  rc = espec_eigenerg(dim_of(DIM), NP, SH, U1, VPOT, EIGENERG)
end function EIGENERG

! rans.f:3-42
real(8) function RAN(NSEED)
  implicit none
  integer :: NSEED
  NSEED = int(mod(int(NSEED, 8) * 1029_8 + 221591_8, 1048576_8))
  RAN = dble(NSEED) / 1048576.0d0
end function RAN

! hwhm.f:1-92
subroutine HWHM(M, DT, TF, TAL, DE, U, V)
  use espec_capi
  implicit none
  integer :: M
  real(8) :: DT, TF, TAL, DE, U(*), V(*)
  integer :: rc
  rc = espec_hwhm(M, DT, TF, TAL, DE, U, V)
end subroutine HWHM

! spectrumtd.f:1-116, including its printed table.
subroutine SPECTRUMTD(TPWIND, M, DT, TF, WP, DE, U, V)
  use espec_capi
  implicit none
  character(len=*) :: TPWIND
  integer :: M
  real(8) :: DT, TF, WP, DE, U(*), V(*)
  integer :: iwin, rc, n, i
  real(8), allocatable :: e(:), w(:), wi(:)
  if (starts(TPWIND, '.SG')) then
    iwin = 1
  else if (starts(TPWIND, '.NONE') .or. starts(TPWIND, '.NULL')) then
    iwin = 0
  else
    write(*,'(A)') '<<<>>> Windowing function error <<<>>>'
    return
  end if
  n = 2**M
  allocate(e(n), w(n), wi(n))
  rc = espec_spectrumtd(iwin, M, DT, TF, WP, DE, U, V, e, w, wi, mode_flags())
  if (rc == 0) then
    write(*,*)
    write(*,*) 'Spectrum:'
    write(*,*) '========='
    write(*,1011)
    do i = 1, n
      write(*,1021) e(i), w(i), w(i), wi(i)
    end do
  end if
  deallocate(e, w, wi)
1011 format(5X,'*E (eV)*',11X,'*W*',11X,'*FFTr*',10X,'*FFTi*')
1021 format(1X,4(F12.7,4X))
end subroutine SPECTRUMTD

! sil.f:3-185: one short-iterative Lanczos step of U + iV over CST (a.u.),
! on the device (espec_plnz_absorb with TI = TF = 0, DT = CST SHFS, one
! step), with PLNZAUX's .SMOOTHW mask for ABSORB.  A failed eigensolve
! returns INFO = 5 where the reference stops ("Diagonalization error").
! LNZVC, IWORK and the work arrays are the library's business: the
! Krylov basis is not left in LNZVC.
subroutine PLNZAUX(ABSORB, CHANGE, DIM, TPABSOR, INFO, LMTREORT, MP, MXDCT, N, ABSTOL, CST, &
                   NP, IWORK, SHM, VPOT, VABC, U, V, WU, WV, WP, WORK, VAR, LNZVC)
  use espec_capi
  implicit none
  logical :: ABSORB, CHANGE
  character(len=*) :: DIM, TPABSOR
  integer :: INFO, LMTREORT, MP, MXDCT, N
  real(8) :: ABSTOL, CST
  integer :: NP(*), IWORK(*)
  real(8) :: SHM(*), VPOT(*), VABC(*), U(*), V(*), WU(*), WV(*), WP(*), WORK(*), VAR(*)
  real(8) :: LNZVC(MXDCT, *)
  real(8), parameter :: SHFS = 2.4188843243d-2
  real(8) :: cr(2), ci(2), nrm(2)
  integer :: ndim, ns, rc
  INFO = 0
  ndim = dim_of(DIM)
  if (ndim == 0) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  rc = espec_plnz_absorb(ndim, NP, MP, ABSTOL, CST * SHFS, 0.0d0, 0.0d0, 1, SHM, U, V, VPOT, &
                         cr, ci, nrm, 2, ns, mode_flags(), absorb_of(ABSORB, TPABSOR), VABC)
  if (rc /= 0) then
    if (rc == 5) write(*,'(A)') '<<<>>> Diagonalization error <<<>>>'
    INFO = rc
  end if
end subroutine PLNZAUX

! lnz.f:4-129: M Lanczos steps from VI with one classical Gram-Schmidt pass
! against the first NREORT vectors, on the device (espec_lnz).  AL, BT and
! LNZVC(:,1..NREORT) as the reference leaves them (columns past M zero);
! with M <= NREORT, V0 and VI hold the last two Lanczos vectors as in the
! reference.  WORK is scratch.
subroutine LNZ(CHANGE, DIM, N, M, NREORT, LMTREORT, MXDCT, NP, SHM, VPOT, V0, VI, WORK, AL, BT, &
               VAR, LNZVC)
  use espec_capi
  implicit none
  logical :: CHANGE
  character(len=*) :: DIM
  integer :: N, M, NREORT, LMTREORT, MXDCT
  integer :: NP(*)
  real(8) :: SHM(*), VPOT(*), V0(*), VI(*), WORK(*), AL(*), BT(*), VAR(*), LNZVC(MXDCT, *)
  real(8), allocatable :: basis(:)
  integer :: ndim, rc, j, nre
  if (NREORT > LMTREORT) then  ! lnz.f:64-73
    write(*,*) 'Error NREORT > LMTREORT'
    write(*,*) 'NREORT=', NREORT, '  LMTREORT=', LMTREORT
    if (CHANGE) then
      NREORT = LMTREORT
      write(*,*) 'Modifing NREORT to correct value!'
      write(*,*) 'NREORT=', NREORT, '  LMTREORT=', LMTREORT
    end if
  end if
  ndim = dim_of(DIM)
  if (ndim == 0 .or. M < 1 .or. NREORT < 1) then
    write(*,'(A)') '<<<>>> LNZ: not on the device for this DIM/M/NREORT'
    return
  end if
  nre = NREORT
  allocate(basis(N * nre))
  rc = espec_lnz(ndim, NP, M, nre, SHM, VPOT, VI, AL, BT, basis)
  if (rc /= 0) then
    write(*,'(A,I3)') '<<<>>> LNZ on the device failed, code ', rc
    deallocate(basis)
    return
  end if
  do j = 1, nre
    LNZVC(1:N, j) = basis((j - 1) * N + 1:j * N)
  end do
  if (M <= nre) then
    VI(1:N) = basis((M - 1) * N + 1:M * N)
    if (M > 1) then
      V0(1:N) = basis((M - 2) * N + 1:(M - 1) * N)
    else
      V0(1:N) = 0.0d0
    end if
  end if
  deallocate(basis)
end subroutine LNZ

! reort.f:1-83: WORK <- WORK - V (V^T WORK) over the first M columns of
! EIGVC (leading dimension MXDCT), on the device (espec_reort).
subroutine REORT(N, M, LMTREORT, MXDCT, WORK, EIGVC)
  use espec_capi
  implicit none
  integer :: N, M, LMTREORT, MXDCT
  real(8) :: WORK(*), EIGVC(MXDCT, *)
  integer :: rc
  rc = espec_reort(N, M, MXDCT, WORK, EIGVC)
  if (rc /= 0) write(*,'(A,I3)') '<<<>>> REORT on the device failed, code ', rc
end subroutine REORT

! spofft.f:3-203: one split-operator step with the caller's TC (from DIGT),
! on the device (espec_pspofftaux).
subroutine PSPOFFTAUX(ABSORB, TPABSOR, NDIM, N, B, B1, B2, CST1, CST2, NP, U, V, TC, VPOT, VABC)
  use espec_capi
  implicit none
  logical :: ABSORB
  character(len=*) :: TPABSOR
  integer :: NDIM, N
  real(8) :: B, B1, B2, CST1, CST2
  integer :: NP(*)
  real(8) :: U(*), V(*), TC(*), VPOT(*), VABC(*)
  integer :: rc
  rc = espec_pspofftaux(NDIM, NP, B, CST1, CST2, TC, U, V, VPOT, absorb_of(ABSORB, TPABSOR), &
                        VABC, mode_flags())
  if (rc /= 0) write(*,'(A,I3)') '<<<>>> PSPOFFTAUX on the device failed, code ', rc
end subroutine PSPOFFTAUX

! sodfft.f:3-129: one leapfrog step (U1,V1: psi(t-dt) -> psi(t); U2,V2:
! psi(t) -> psi(t+dt)), on the device (espec_psodfftaux); WRKR and WRKI are
! not used.
subroutine PSODFFTAUX(ABSORB, TPABSOR, NDIM, N, B, CST, NP, U1, U2, V1, V2, TC, VPOT, VABC, &
                      WRKR, WRKI)
  use espec_capi
  implicit none
  logical :: ABSORB
  character(len=*) :: TPABSOR
  integer :: NDIM, N
  real(8) :: B, CST
  integer :: NP(*)
  real(8) :: U1(*), U2(*), V1(*), V2(*), TC(*), VPOT(*), VABC(*), WRKR(*), WRKI(*)
  integer :: rc
  rc = espec_psodfftaux(NDIM, NP, B, CST, TC, U1, U2, V1, V2, VPOT, absorb_of(ABSORB, TPABSOR), &
                        VABC, mode_flags())
  if (rc /= 0) write(*,'(A,I3)') '<<<>>> PSODFFTAUX on the device failed, code ', rc
end subroutine PSODFFTAUX

! cfft.f:6-106: radix-2 transform of N = 2**M points on the device
! (espec_fftnt); ITYPE = -1 backward with 1/N.
subroutine FFTNT(X, Y, N, M, ITYPE)
  use espec_capi
  implicit none
  integer :: N, M, ITYPE
  real(8) :: X(*), Y(*)
  integer :: rc
  rc = espec_fftnt(X, Y, N, M, ITYPE, mode_flags())
  if (rc /= 0) write(*,'(A,I3)') '<<<>>> FFTNT on the device failed, code ', rc
end subroutine FFTNT

! s2ppsod.f:1-1143: two surfaces coupled by the coordinate operator (the
! routine's active TPCRS = '.CORD'), FD leapfrog with a pulse, on the
! device.  U1, V1, VPOT, DM hold surface A in 1..N and B in N+1..2N.  The
! routine's own diagnostic lines and its table (FORMAT 1011/1021, with the
! reference's FNTB**2 at TI and CTB in the CRB column after the start-up)
! are printed; U2, V2 return psi(TF), U1, V1 psi(TF - DT) and VPOT the last
! VEFF.  The files (PRTVEFF, PRTEIGVC2, PRTPULS, .LASTWP) are not written:
! INFO = 6.  An unknown EFC gives INFO = 2; an unknown TPABSOR with ABSORB
! stops as PSODAUX does.
subroutine S2PPSOD(ABSORB, PRTVEFF, PRTEIGVC2, PRTPULS, DIM, EFC, PRTCRL, TPABSOR, INFO, LC, &
                   MP, MPI, MPR, MXDCT, N, NSHOT, NPR, KP, NISPG, KL, E0, T0, TD, TP, OMG, SNI, &
                   DT, TI, TF, TOL, DELQ, QC, AA, NP, ND, SH, SHM, U1, U2, V1, V2, XI, XP, XF, &
                   SM, DM, VPOT, VABC, WORK, VAR, LNZVC)
  use espec_capi
  implicit none
  logical :: ABSORB, PRTVEFF, PRTEIGVC2, PRTPULS
  character(len=*) :: DIM, EFC, PRTCRL, TPABSOR
  integer :: INFO, LC, MP, MPI, MPR, MXDCT, N, NSHOT, NPR, KP, NISPG, KL
  real(8) :: DT, TI, TF, TOL, E0, T0, TD, TP, OMG, SNI, DELQ, QC, AA
  integer :: NP(*), ND(*)
  real(8) :: SH(*), SHM(*), U1(*), U2(*), V1(*), V2(*), XI(*), XP(*), XF(*), SM(*), DM(*)
  real(8) :: VPOT(*), VABC(*), WORK(*), VAR(*), LNZVC(MXDCT, *)
  integer :: ndim, efk, kind, prt, maxrows, nrows, inf, rc, r
  real(8) :: cst, csqc
  real(8), allocatable :: vp(:), tab(:)

  INFO = 0
  ndim = dim_of(DIM)
  efk = pulse_of(EFC)
  cst = 2.0d0 * DT / 2.4188843243d-2
  csqc = cst * AA / SM(1)
  ! s2ppsod.f:130-132: the routine's own diagnostic output
  write(*,*) 'oi', csqc, 0d0, 0d0
  write(*,*) DELQ, QC, AA
  write(*,*) SH(1), (XF(1) - XI(1)) / (N - 1), SHM(1), csqc, SHM(1) * 30, AA
  flush(6)
  if (ndim == 0 .or. PRTVEFF .or. PRTEIGVC2 .or. PRTPULS .or. starts(PRTCRL, '.LASTWP')) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  if (efk < 0) then
    write(*,'(A)') '<<<>>> Desired laser pulse wasn''t found. <<<>>>'
    INFO = ESPEC_ERR_ARG
    return
  end if
  kind = ESPEC_ABSORB_NONE
  if (ABSORB) then  ! sod.f:55-120
    if (starts(TPABSOR, '.VOPTIC')) then
      kind = ESPEC_ABSORB_VOPTIC
    else if (starts(TPABSOR, '.SMOOTHW') .or. starts(TPABSOR, '.SMRS')) then
      kind = ESPEC_ABSORB_SMOOTHW
    else
      write(*,*) '<<<>>> SOD propagation error <<<>>>'
      error stop
    end if
  end if
  prt = 0
  if (starts(PRTCRL, '.PARTIAL')) prt = 1
  if (starts(PRTCRL, '.YES')) prt = 2
  maxrows = int((TF - TI) / DT) + 4
  allocate(vp(2 * N), tab(12 * maxrows))
  vp = VPOT(1:2 * N)
  ! psi0 comes in U1, V1 (s2ppsod.f:143-149); U2, V2 carry it to the device
  U2(1:2 * N) = U1(1:2 * N)
  V2(1:2 * N) = V1(1:2 * N)
  nrows = 0
  inf = 0
  rc = espec_s2ppsod(ndim, NP, DT, TI, TF, LC, SHM, XI(1), SH(1), SM(1), AA, U2, V2, vp, DM, &
                     efk, E0, T0, TD, TP, OMG, SNI, KL, kind, VABC, prt, TOL, tab, maxrows, &
                     nrows, inf, U1, V1, VPOT, ESPEC_FP64)
  deallocate(vp)
  if (rc /= 0) then
    INFO = rc
    deallocate(tab)
    return
  end if
  INFO = inf
  if (prt > 0) then
    write(*,*)
    if (DIM == '.1D') then
      write(*,*) 'Auto-correlation function:'
    else
      write(*,*) 'Partial auto-correlation function:'
    end if
    write(*,*) '=================================='
    write(*,1011)
    do r = 1, min(nrows, maxrows)
      write(*,1021) tab(12 * (r - 1) + 1:12 * r)
    end do
  end if
  deallocate(tab)
  flush(6)
1011 format(/,8X,'*t/fs*',6X,'*ET/a.u.*',6X,'*E0/a.u.*',6X,'*CIA*',6X,'*CRA*',6X,'*CTA*',8X, &
            '*CIB*',6X,'*CRB*',6X,'*CTB*',8X,'*FTA*',8X,'*FTB*',8X,'*FT*')
1021 format(1X,F14.6,3X,2(E12.6,3X),2(F8.5,3X),E10.4,3X,2(F8.5,3X),E10.4,3X,2(E10.4,3X),F8.6)
end subroutine S2PPSOD

! ppsod.f:1-548: the pulse-driven finite-difference leapfrog (PSODAUX,
! sod.f:3-122) on the device (espec_ppsod), with the reference's table
! (FORMAT 1011/1012/1021: t, ET, E0, Pr, Pi, |c|^2, the populations of
! LNZVC(:,KP+1..KP+NPR) and the norm), every printed quantity formed on the
! device.  LNZVC(:,MPR/MPI) receives C(t) where the reference stores it:
! C(TI) and the first step's sample only when PRTCRL prints (so with
! PRTCRL '.NO' and LC = 1 nothing is stored, ppsod.f:276-296).  U2, V2 end
! as psi(TF), U1, V1 as psi(TF - DT) and VPOT as the last VEFF.  The movie,
! potential and pulse files (PRTVEFF, PRTEIGVC2, PRTPULS) return INFO = 6;
! a TPABSOR that PSODAUX does not know stops, as the reference does.
subroutine PPSOD(ABSORB, PRTVEFF, PRTEIGVC2, PRTPULS, DIM, EFC, PRTCRL, TPABSOR, INFO, LC, MP, &
                 MPI, MPR, MXDCT, N, NSHOT, NPR, KP, NISPG, E0, T0, TD, TP, OMG, SNI, KL, DT, &
                 TI, TF, TOL, NP, ND, SH, SHM, U1, U2, V1, V2, XI, XP, XF, DM, VPOT, VABC, &
                 WORK, VAR, LNZVC, PRTREGION, NREG, RANGE, PRTONLYREIM)
  use espec_capi
  implicit none
  logical :: ABSORB, PRTVEFF, PRTEIGVC2, PRTPULS, PRTREGION, PRTONLYREIM
  character(len=*) :: DIM, EFC, PRTCRL, TPABSOR
  integer :: INFO, LC, MP, MPI, MPR, MXDCT, N, NSHOT, NPR, KP, NISPG, KL, NREG
  real(8) :: DT, TI, TF, TOL, E0, T0, TD, TP, OMG, SNI
  integer :: NP(*), ND(*)
  real(8) :: SH(*), SHM(*), U1(*), U2(*), V1(*), V2(*), XI(*), XP(*), XF(*), DM(*), VPOT(*)
  real(8) :: VABC(*), WORK(*), VAR(*), LNZVC(MXDCT, *), RANGE(5, 7)
  integer :: ndim, efk, kind, npj, nf, nstep, lcd, maxs, ns, rc, s, j, step, nc, m
  real(8) :: t, f0, ct(7)
  real(8), allocatable :: vp(:), w(:), cr(:), ci(:), nr(:), et(:), et0(:), pr(:)
  logical :: every, prt

  INFO = 0
  ndim = dim_of(DIM)
  efk = pulse_of(EFC)
  if (ndim == 0 .or. PRTVEFF .or. PRTEIGVC2 .or. PRTPULS) then
    INFO = ESPEC_ERR_UNSUPPORTED
    return
  end if
  if (efk < 0) then
    write(*,'(A)') '<<<>>> Desired laser pulse wasn''t found. <<<>>>'
    INFO = ESPEC_ERR_ARG
    return
  end if
  kind = ESPEC_ABSORB_NONE
  if (ABSORB) then  ! sod.f:55-120
    if (starts(TPABSOR, '.VOPTIC')) then
      kind = ESPEC_ABSORB_VOPTIC
    else if (starts(TPABSOR, '.SMOOTHW') .or. starts(TPABSOR, '.SMRS')) then
      kind = ESPEC_ABSORB_SMOOTHW
    else
      write(*,*) '<<<>>> SOD propagation error <<<>>>'
      error stop
    end if
  end if
  every = starts(PRTCRL, '.YES')
  prt = every .or. starts(PRTCRL, '.PARTIAL')
  npj = 0
  if (prt) npj = max(0, min(NPR, 6))
  nf = nint((TF + DT / 2 - (TI + DT)) / DT)  ! ppsod.f:355
  nstep = max(nf, 0) + 1                     ! the start-up counts as one step
  lcd = LC
  if (every) lcd = 1
  maxs = nstep / lcd + 1
  allocate(vp(N), w(max(1, N * npj)), cr(maxs), ci(maxs), nr(maxs), et(maxs), et0(maxs), &
           pr(max(1, 2 * npj * maxs)))
  vp = VPOT(1:N)
  do j = 1, npj
    w((j - 1) * N + 1:j * N) = LNZVC(1:N, KP + j)
  end do
  ! psi0 comes in U1, V1 (ppsod.f:85-92); U2, V2 carry it to the device
  U2(1:N) = U1(1:N)
  V2(1:N) = V1(1:N)
  rc = espec_ppsod(ndim, NP, DT, TI, TF, lcd, SHM, U2, V2, vp, DM, efk, E0, T0, TD, TP, OMG, &
                   SNI, KL, kind, VABC, npj, w, U1, V1, VPOT, cr, ci, nr, et, et0, pr, maxs, &
                   ns, ESPEC_FP64)
  if (rc /= 0) then
    INFO = rc
    deallocate(vp, w, cr, ci, nr, et, et0, pr)
    return
  end if
  deallocate(vp)
  ! LNZVC(:,MPR/MPI): sample s sits at step s*lcd; the reference stores
  ! the steps that are multiples of LC, C(TI) and step 1 only when printing.
  nc = 1
  do s = 0, ns - 1
    step = s * lcd
    if (mod(step, LC) /= 0) cycle
    if (.not. prt .and. (s == 0 .or. LC == 1)) cycle
    nc = step / LC + 1
    LNZVC(nc, MPR) = cr(s + 1)
    LNZVC(nc, MPI) = ci(s + 1)
  end do
  if (prt) then
    f0 = nr(1)
    write(*,*)
    if (DIM == '.1D') then
      write(*,*) 'Auto-correlation function:'
    else
      write(*,*) 'Partial auto-correlation function:'
    end if
    write(*,*) '=================================='
    if (NPR + KP - 1 > 10) then
      write(*,1012) '*|P', NISPG - 1, '|^2*', ('*|P', j - 1 + KP, '|^2*', j = 1, NPR), &
                    '*|c', NISPG - 1, '|^2*'
    else
      write(*,1011) '*|P', NISPG - 1, '|^2*', ('*|P', j - 1 + KP, '|^2*', j = 1, NPR), &
                    ' |c', NISPG - 1, '|^2 '
    end if
    t = TI  ! T = TI + DT after the start-up, then T = T + DT (ppsod.f:256, 358)
    m = 0
    do s = 0, ns - 1
      do while (m < s * lcd)
        if (m == 0) then
          t = TI + DT
        else
          t = t + DT
        end if
        m = m + 1
      end do
      ct(1) = cr(s + 1)**2 + ci(s + 1)**2
      do j = 1, npj
        ct(j + 1) = pr(2 * (s * npj + j) - 1)**2 + pr(2 * (s * npj + j))**2
      end do
      ! ppsod.f:456-460: the norm check of the loop's printed rows
      if (s >= 2 .or. (s == 1 .and. lcd > 1)) then
        if (.not. ABSORB .and. abs((nr(s + 1) - f0) / f0) * 1.0d2 > TOL) INFO = 1
      end if
      write(*,1021) t, et(s + 1), et0(s + 1), cr(s + 1), ci(s + 1), (ct(j), j = 1, NPR + 1), &
                    nr(s + 1)
    end do
    flush(6)
  end if
  deallocate(w, cr, ci, nr, et, et0, pr)
1011 format(/,8X,'*t/fs*',5X,'*ET/a.u.*',6X,'*E0/a.u.*',6X,'*Pr(t)*',4X,'*Pi(t)*',3X, &
            8(A3,I1,A4,3X))
1012 format(/,8X,'*t/fs*',5X,'*ET/a.u.*',6X,'*E0/a.u.*',6X,'*Pr(t)*',4X,'*Pi(t)*',3X, &
            8(A3,I2,A4,3X))
1021 format(1X,F14.6,3X,2(E12.6,3X),2(F8.5,3X),10(F8.6,3X))
end subroutine PPSOD
