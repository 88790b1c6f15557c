!-----------------------------------------------------------------------
! Host program in the shape of viscouscycles.f90:413-434 calling the GPU
! through the ucy_fortran module with a mixed mesh: rectangles, triangles
! and vertical strain volumes with their rheology, in the reference's array
! shapes (stream file).  It writes the element frames the library computes
! (computeReferenceSystem*), then takes nsteps RKQSm steps from the
! viscouscycles initial state and writes time, dt and the final state.
! Used by tests/test_fortran_binding.py.
!-----------------------------------------------------------------------
PROGRAM fortran_driver_viscous
  USE ucy_fortran
  IMPLICIT NONE
  INTEGER :: nRect, nTri, nVe, nVol, nsteps, np, n, i, u
  REAL*8 :: lambda, mu, t0, time, dt_try, dt_done, dt_next
  REAL*8, ALLOCATABLE :: x(:,:), length(:), width(:), strike(:), dip(:), vert(:,:)
  INTEGER, ALLOCATABLE :: i1(:), i2(:), i3(:)
  REAL*8, ALLOCATABLE :: Vpl(:), rake(:), mu0(:), sig(:), a(:), b(:), L(:), Vo(:), damping(:)
  REAL*8, ALLOCATABLE :: vx(:,:), vlength(:), vthickness(:), vwidth(:), vstrike(:)
  REAL*8, ALLOCATABLE :: g0(:), npow(:), e(:,:), Gm(:), Q(:), R(:), To(:)
  REAL*8, ALLOCATABLE :: fr(:,:,:), ft(:,:,:), fv(:,:,:)
  REAL*8, ALLOCATABLE :: y(:), dydt(:), yscal(:), ycand(:), yerr(:), ytmp(:), tlog(:), dtlog(:)
  CHARACTER(LEN=256) :: fin, fout

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN (NEWUNIT=u, FILE=TRIM(fin), ACCESS="STREAM", FORM="UNFORMATTED", STATUS="OLD")
  READ (u) nRect, nTri, nVe, nVol, nsteps, lambda, mu
  np = nRect + nTri
  ALLOCATE(x(3,nRect), length(nRect), width(nRect), strike(nRect), dip(nRect))
  ALLOCATE(i1(nTri), i2(nTri), i3(nTri), vert(3,nVe))
  ALLOCATE(Vpl(np), rake(np), mu0(np), sig(np), a(np), b(np), L(np), Vo(np), damping(np))
  ALLOCATE(vx(3,nVol), vlength(nVol), vthickness(nVol), vwidth(nVol), vstrike(nVol))
  ALLOCATE(g0(nVol), npow(nVol), e(6,nVol), Gm(nVol), Q(nVol), R(nVol), To(nVol))
  READ (u) x, length, width, strike, dip, i1, i2, i3, vert
  READ (u) Vpl, rake, mu0, sig, a, b, L, Vo, damping
  READ (u) vx, vlength, vthickness, vwidth, vstrike, g0, npow, e, Gm, Q, R, To
  CLOSE (u)

  ! the frames, as the reference computes them before buildG
  ALLOCATE(fr(3,nRect,4), ft(3,nTri,4), fv(3,nVol,4))
  CALL computeReferenceSystemOkada92_gpu(nRect, x, length, width, strike, dip, &
       fr(:,:,1), fr(:,:,2), fr(:,:,3), fr(:,:,4))
  CALL computeReferenceSystemNikkhoo15_gpu(nVe, vert, nTri, i1, i2, i3, &
       ft(:,:,1), ft(:,:,2), ft(:,:,3), ft(:,:,4))
  CALL computeReferenceSystemVerticalStrainVolume_gpu(nVol, vx, vlength, vwidth, vstrike, &
       fv(:,:,1), fv(:,:,2), fv(:,:,3), fv(:,:,4))

  CALL ucy_init(0)
  CALL ucy_setup_mesh(lambda, mu, nRect, x, length, width, strike, dip, nTri, i1, i2, i3, &
                      nVe, vert, nVol, vx, vlength, vthickness, vwidth, vstrike)
  CALL ucy_setup_friction(np, Vpl, rake, mu0, sig, a, b, L, Vo, damping)
  CALL ucy_setup_rheology(nVol, g0, npow, e, Gm, Q, R, To)
  CALL buildG_gpu(0)
  n = 7 * np + 12 * nVol
  ALLOCATE(y(n), dydt(n), yscal(n), ycand(n), yerr(n), ytmp(n), tlog(nsteps), dtlog(nsteps))
  CALL initStateVector_gpu(n, y, 1)

  ! main loop of viscouscycles.f90:413-434 (RK time origin t0 = 0)
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
  CALL ucy_finalize()

  OPEN (NEWUNIT=u, FILE=TRIM(fout), ACCESS="STREAM", FORM="UNFORMATTED", STATUS="REPLACE")
  WRITE (u) fr, ft, fv, tlog, dtlog, y
  CLOSE (u)
END PROGRAM fortran_driver_viscous
