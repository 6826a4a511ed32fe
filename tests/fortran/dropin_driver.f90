! Host program for the drop-in test: calls PSPOFFT / PSODFFT / PLNZ / PPLNZ /
! LANCZSG / MTRXDIAG with the argument lists eSPec's main program uses
! (espec.f:906-915, 1373-1376, 1474-1477, 1499-1502) and the LC/TF rule of
! espec.f:1324-1353, against whatever library provides those symbols (here
! libespec_f.so).
!
! stdin:  method ('PFSPO', 'PFSOD', 'PSIL', 'PPSIL', 'LANCZSG', 'MTRXDIAG'),
!         DIM keyword, N
!         TI TF DT MF MP ABSTOL TOL
!         NP(1:3) SH(1:3) SM(1:3) SHM(1:3)
!         N lines: VPOT U V
! stdout: the routine's own partial-correlation table, then lines
!         'C ' NC Re Im for every stored sample and 'INFO ' INFO.
program dropin_driver
  implicit none
  character(len=8) :: method, dim
  integer :: n, mf, mp, lc, info, i, mxdct, mpr, mpi, nsamp
  integer :: np(3)
  real(8) :: ti, tf, dt, abstol, tol, sc
  real(8) :: sh(3), sm(3), shm(3)
  real(8), allocatable :: vpot(:), u(:), v(:), u2(:), v2(:), tc(:), w1(:), w2(:), w3(:), w4(:)
  real(8), allocatable :: lnzvc(:, :), var(:), xi(:), xf(:), vabc(:), xp(:), dm(:)
  integer, allocatable :: iwork(:)
  integer :: nd(3)

  read(*,*) method, dim, n
  read(*,*) ti, tf, dt, mf, mp, abstol, tol
  read(*,*) np, sh, sm, shm
  allocate(vpot(n), u(n), v(n), u2(n), v2(n), tc(n), w1(n), w2(n), w3(n), w4(n), var(8), &
           xi(3), xf(3), vabc(n), iwork(5 * n + 16), xp(3), dm(n))
  do i = 1, n
    read(*,*) vpot(i), u(i), v(i)
  end do
  if (trim(method) == 'LANCZSG' .or. trim(method) == 'MTRXDIAG') then
    call eigenstates()
    stop
  end if
  ! espec.f:1324-1353 with *CHANGE .YES: LC from 2^MF samples, TF adjusted
  do
    sc = 2.0d0 * (tf - ti) / dt
    lc = int(sc / 2.0d0**mf)
    if (abs(sc - lc * 2.0d0**mf) >= 1.0d-10) then
      tf = lc * 2.0d0**(mf - 1) * dt
      cycle
    end if
    exit
  end do
  mxdct = max(n, 2**mf)
  mpr = mp + 1
  mpi = mp + 2
  allocate(lnzvc(mxdct, mp + 2))
  lnzvc = 0
  xi = 0
  xf = 0
  vabc = 1
  var = 0
  xp = 0
  dm = 0
  nd = 0
  info = 0
  select case (trim(method))
  case ('PFSPO')
    call PSPOFFT(.false., dim, '.PARTIAL', '.NONE', info, lc, mp, mpi, mpr, mxdct, n, dt, ti, &
                 tf, tol, np, shm, sh, sm, u, v, xi, xf, tc, vpot, vabc, lnzvc, w1, w2, var)
  case ('PFSOD')
    call PSODFFT(.false., dim, '.PARTIAL', '.NONE', info, lc, mp, mpi, mpr, mxdct, n, dt, ti, &
                 tf, tol, np, shm, sh, sm, u, u2, v, v2, xi, xf, tc, vpot, vabc, w1, w2, var, &
                 lnzvc)
  case ('PSIL')
    call PLNZ(.false., .true., dim, '.PARTIAL', '.NONE', info, lc, 10, mp, mpi, mpr, mxdct, n, &
              abstol, dt, ti, tf, tol, np, iwork, shm, vpot, vabc, u, v, xi, xf, w1, w2, w3, w4, &
              var, lnzvc)
  case ('PPSIL')  ! espec.f:1509-1518 with no laser (EFC .NONE), one population
    call PPLNZ(.false., .true., .false., .false., .false., dim, '.NONE', '.PARTIAL', '.NONE', &
               info, lc, 10, mp, mpi, mpr, mxdct, n, 1, 1, 1, 0, abstol, dt, ti, tf, tol, &
               0d0, 0d0, 0d0, 1d0, 0d0, 0d0, np, nd, iwork, sh, shm, vpot, vabc, u, v, xi, xf, &
               xp, dm, w1, w2, w3, w4, var, lnzvc)
  end select
  nsamp = 2**(mf - 1) + 1
  do i = 1, nsamp
    write(*,'(A,I8,2ES26.17)') 'C ', i, lnzvc(i, mpr), lnzvc(i, mpi)
  end do
  write(*,'(A,I4)') 'INFO ', info

contains

  ! LANCZSG / MTRXDIAG with espec.f:906-915's argument lists; for these
  ! methods MF is MAXINT and MP is IU (IL = 1).  Prints 'E ' K EIGVL(K) and
  ! 'V ' K I EIGVC(I,K) (LANCZSG: the LNZVC dummy, espec's EIGVC).
  subroutine eigenstates()
    real(8), allocatable :: ev(:), egv(:, :), lz(:, :), al(:), bt(:)
    integer :: k, j, nreort, nseed, ns
    ns = mp
    allocate(ev(max(n, 64)), egv(n, max(40, ns)), lz(n, max(40, ns)), al(64), bt(64))
    egv = 0
    lz = 0
    var = 0
    nreort = 40
    nseed = 1
    info = 0
    if (trim(method) == 'LANCZSG') then
      call LANCZSG(.true., dim, 1, ns, info, nreort, 40, n, nseed, n, abstol, iwork, mf, np, &
                   al, bt, ev, shm, w1, w2, vpot, w3, w4, var, lz, egv)
    else
      call MTRXDIAG(dim, 1, ns, info, n, n, abstol, iwork, np, ev, shm, vpot, w1, w2, lz)
    end if
    do k = 1, ns
      write(*,'(A,I4,ES26.17)') 'E ', k, ev(k)
      do j = 1, n
        write(*,'(A,I4,I8,ES26.17)') 'V ', k, j, lz(j, k)
      end do
    end do
    write(*,'(A,I4)') 'INFO ', info
  end subroutine eigenstates
end program dropin_driver
