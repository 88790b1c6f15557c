!-----------------------------------------------------------------------
! Host program in the shape of ratestate.f90:307-328 calling the GPU through
! the ucy_fortran module: reads a rectangle mesh and friction (stream file),
! builds K and the displacement kernels of the observation points (buildO),
! sets the initial state and takes nsteps RKQSm steps, writing time, dt, the
! state and the observation-point displacement of the final state
! (exportPoints).  Used by tests/test_fortran_binding.py.
!-----------------------------------------------------------------------
PROGRAM fortran_driver
  USE ucy_fortran
  IMPLICIT NONE
  INTEGER :: ns, nsteps, i, n, u, nobs
  REAL*8 :: lambda, mu, t0, time, dt_try, dt_done, dt_next
  REAL*8, ALLOCATABLE :: x(:,:), length(:), width(:), strike(:), dip(:)
  REAL*8, ALLOCATABLE :: Vpl(:), rake(:), mu0(:), sig(:), a(:), b(:), L(:), Vo(:), damping(:)
  REAL*8, ALLOCATABLE :: y(:), dydt(:), yscal(:), ycand(:), yerr(:), ytmp(:)
  REAL*8, ALLOCATABLE :: tlog(:), dtlog(:)
  REAL*8, ALLOCATABLE :: xobs(:,:), d(:), df(:), dl(:)
  CHARACTER(LEN=256) :: fin, fout

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN (NEWUNIT=u, FILE=TRIM(fin), ACCESS="STREAM", FORM="UNFORMATTED", STATUS="OLD")
  READ (u) ns, nsteps, lambda, mu
  ALLOCATE(x(3,ns), length(ns), width(ns), strike(ns), dip(ns))
  ALLOCATE(Vpl(ns), rake(ns), mu0(ns), sig(ns), a(ns), b(ns), L(ns), Vo(ns), damping(ns))
  READ (u) x, length, width, strike, dip, Vpl, rake, mu0, sig, a, b, L, Vo, damping
  READ (u) nobs
  ALLOCATE(xobs(3,nobs), d(3*nobs), df(3*nobs), dl(3*nobs))
  READ (u) xobs
  CLOSE (u)

  CALL ucy_init(0)
  CALL ucy_setup_rectangles(lambda, mu, ns, x, length, width, strike, dip, &
                            Vpl, rake, mu0, sig, a, b, L, Vo, damping)
  CALL buildO_gpu(nobs, xobs)
  n = 7 * ns
  ALLOCATE(y(n), dydt(n), yscal(n), ycand(n), yerr(n), ytmp(n), tlog(nsteps), dtlog(nsteps))
  CALL initStateVector_gpu(n, y, 0)

  ! main loop of ratestate.f90:307-328 (RK time origin t0 = 0, :317)
  time = 0._8
  dt_try = 1.d-3
  DO i = 1, nsteps
     CALL odefun_gpu(n, time, y, dydt)
     yscal(:) = ABS(y(:)) + ABS(dt_try * dydt(:)) + 1.d-30
     t0 = 0._8
     CALL RKQSm_gpu(n, t0, y, dydt, yscal, ycand, yerr, ytmp, dt_try, dt_done, dt_next, odefun_gpu)
     time = time + dt_done
     tlog(i) = time
     dtlog(i) = dt_done
     dt_try = dt_next
  END DO
  CALL exportPoints_gpu(n, y, 3 * nobs, d, df, dl)
  CALL ucy_finalize()

  OPEN (NEWUNIT=u, FILE=TRIM(fout), ACCESS="STREAM", FORM="UNFORMATTED", STATUS="REPLACE")
  WRITE (u) tlog, dtlog, y, d, df, dl
  CLOSE (u)
END PROGRAM fortran_driver
