!-----------------------------------------------------------------------
!> ucy_fortran -- iso_c_binding layer over libucy (include/ucy.h) for the
!! fortran/3d host program of jdpmoore/unicycle.
!!
!! The interfaces bind the C entry points one to one.  The wrappers below
!! keep the reference's argument lists so that the call sites of
!! ratestate.f90 / viscouscycles.f90 change only by name:
!!   odefun(n,time,y,dydt)              ratestate.f90:624-630
!!   RKQSm(n,t,y,dydt,yscal,ycand,yerr,ytmp,htry,hdid,hnext,odefun)
!!                                      ode45.f90:51-62
!!   RKCKm(n,t,yin,dydt,h,yout,yerr,ytmp,odefun)
!!                                      ode45.f90:116-124
!!   computeTractionKernelsOkada92(...) okada92.f90:368-377 (and the other
!!                                      Green's-function routines)
!!   buildO(in,layout,O,Of,Ol)          greens.f90:302 (O stays on the GPU)
!!   computeReferenceSystemOkada92 / Nikkhoo15 / VerticalStrainVolume
!!                                      okada92.f90:59, nikkhoo15.f90:50,
!!                                      strainvolume.f90:135
!! and ucy_setup_mesh / ucy_setup_friction / ucy_setup_rheology / buildG_gpu
!! take the mesh of any of the 3-D programs (rectangles, triangles, strain
!! volumes with their rheology) in the reference's array shapes.
!!   exportPoints()                     ratestate.f90:414-486 (d = O u)
!! A nonzero status from the library becomes the reference's STOP with the
!! library's message (the library itself never aborts).
!-----------------------------------------------------------------------
MODULE ucy_fortran

  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE

  PRIVATE :: ucy_check

  !> the context of this process (one GPU, as one MPI rank of the reference)
  TYPE(C_PTR), SAVE :: ucy_ctx = C_NULL_PTR
  !> maximum slip velocity of the last odefun (ratestate.f90 module vMax)
  REAL*8, SAVE :: ucy_vmax = 0._8
  !> RKQSm module parameters (ode45.f90:40-43)
  REAL*8, SAVE :: ucy_epsilon = 1.d-6
  REAL*8, SAVE :: ucy_maximum_time_step = HUGE(1._8)

  INTEGER(C_INT), PARAMETER :: UCY_PREC_F64 = 0

  INTERFACE
     FUNCTION ucy_create(device, flags) BIND(C, NAME="ucy_create")
       IMPORT :: C_PTR, C_INT
       INTEGER(C_INT), VALUE :: device, flags
       TYPE(C_PTR) :: ucy_create
     END FUNCTION ucy_create

     SUBROUTINE ucy_destroy(ctx) BIND(C, NAME="ucy_destroy")
       IMPORT :: C_PTR
       TYPE(C_PTR), VALUE :: ctx
     END SUBROUTINE ucy_destroy

     FUNCTION ucy_last_error(ctx) BIND(C, NAME="ucy_last_error")
       IMPORT :: C_PTR
       TYPE(C_PTR), VALUE :: ctx
       TYPE(C_PTR) :: ucy_last_error
     END FUNCTION ucy_last_error

     FUNCTION ucy_set_elastic(ctx, lambda, mu) BIND(C, NAME="ucy_set_elastic")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       REAL(C_DOUBLE), VALUE :: lambda, mu
       INTEGER(C_INT) :: ucy_set_elastic
     END FUNCTION ucy_set_elastic

     FUNCTION ucy_set_rectangles(ctx, ns, x, length, width, strike, dip) &
          BIND(C, NAME="ucy_set_rectangles")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: ns
       REAL(C_DOUBLE), INTENT(IN) :: x(3,*), length(*), width(*), strike(*), dip(*)
       INTEGER(C_INT) :: ucy_set_rectangles
     END FUNCTION ucy_set_rectangles

     FUNCTION ucy_set_triangles(ctx, ns, i1, i2, i3, nVe, v) BIND(C, NAME="ucy_set_triangles")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: ns, nVe
       INTEGER(C_INT), INTENT(IN) :: i1(*), i2(*), i3(*)
       REAL(C_DOUBLE), INTENT(IN) :: v(3,*)
       INTEGER(C_INT) :: ucy_set_triangles
     END FUNCTION ucy_set_triangles

     FUNCTION ucy_set_volumes(ctx, ns, x, length, thickness, width, strike) &
          BIND(C, NAME="ucy_set_volumes")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: ns
       REAL(C_DOUBLE), INTENT(IN) :: x(3,*), length(*), thickness(*), width(*), strike(*)
       INTEGER(C_INT) :: ucy_set_volumes
     END FUNCTION ucy_set_volumes

     FUNCTION ucy_set_friction(ctx, np, Vpl, rake, mu0, sig, a, b, L, Vo, damping) &
          BIND(C, NAME="ucy_set_friction")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: np
       REAL(C_DOUBLE), INTENT(IN) :: Vpl(*), rake(*), mu0(*), sig(*), a(*), b(*), L(*), &
                                     Vo(*), damping(*)
       INTEGER(C_INT) :: ucy_set_friction
     END FUNCTION ucy_set_friction

     FUNCTION ucy_set_rheology(ctx, nv, gammadot0, npower, edot_bg, Gm, Q, R, To) &
          BIND(C, NAME="ucy_set_rheology")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: nv
       REAL(C_DOUBLE), INTENT(IN) :: gammadot0(*), npower(*), edot_bg(6,*), Gm(*), Q(*), &
                                     R(*), To(*)
       INTEGER(C_INT) :: ucy_set_rheology
     END FUNCTION ucy_set_rheology

     FUNCTION ucy_reference_system_okada92(ns, x, L, W, strike, dip, sv, dv, nv, xc) &
          BIND(C, NAME="ucy_reference_system_okada92")
       IMPORT :: C_INT, C_DOUBLE
       INTEGER(C_INT), VALUE :: ns
       REAL(C_DOUBLE), INTENT(IN) :: x(3,*), L(*), W(*), strike(*), dip(*)
       REAL(C_DOUBLE), INTENT(OUT) :: sv(3,*), dv(3,*), nv(3,*), xc(3,*)
       INTEGER(C_INT) :: ucy_reference_system_okada92
     END FUNCTION ucy_reference_system_okada92

     FUNCTION ucy_reference_system_nikkhoo15(nVe, v, ns, i1, i2, i3, sv, dv, nv, xc) &
          BIND(C, NAME="ucy_reference_system_nikkhoo15")
       IMPORT :: C_INT, C_DOUBLE
       INTEGER(C_INT), VALUE :: nVe, ns
       REAL(C_DOUBLE), INTENT(IN) :: v(3,*)
       INTEGER(C_INT), INTENT(IN) :: i1(*), i2(*), i3(*)
       REAL(C_DOUBLE), INTENT(OUT) :: sv(3,*), dv(3,*), nv(3,*), xc(3,*)
       INTEGER(C_INT) :: ucy_reference_system_nikkhoo15
     END FUNCTION ucy_reference_system_nikkhoo15

     FUNCTION ucy_reference_system_strainvolume(ns, x, L, W, strike, sv, dv, nv, xc) &
          BIND(C, NAME="ucy_reference_system_strainvolume")
       IMPORT :: C_INT, C_DOUBLE
       INTEGER(C_INT), VALUE :: ns
       REAL(C_DOUBLE), INTENT(IN) :: x(3,*), L(*), W(*), strike(*)
       REAL(C_DOUBLE), INTENT(OUT) :: sv(3,*), dv(3,*), nv(3,*), xc(3,*)
       INTEGER(C_INT) :: ucy_reference_system_strainvolume
     END FUNCTION ucy_reference_system_strainvolume

     FUNCTION ucy_build_kernel(ctx, precision) BIND(C, NAME="ucy_build_kernel")
       IMPORT :: C_PTR, C_INT
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: precision
       INTEGER(C_INT) :: ucy_build_kernel
     END FUNCTION ucy_build_kernel

     FUNCTION ucy_init_state(ctx, variant, y) BIND(C, NAME="ucy_init_state")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: variant
       REAL(C_DOUBLE), INTENT(OUT) :: y(*)
       INTEGER(C_INT) :: ucy_init_state
     END FUNCTION ucy_init_state

     FUNCTION ucy_odefun(ctx, n, time, y, dydt, vmax) BIND(C, NAME="ucy_odefun")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: n
       REAL(C_DOUBLE), VALUE :: time
       REAL(C_DOUBLE), INTENT(IN) :: y(*)
       REAL(C_DOUBLE), INTENT(OUT) :: dydt(*)
       REAL(C_DOUBLE), INTENT(OUT) :: vmax
       INTEGER(C_INT) :: ucy_odefun
     END FUNCTION ucy_odefun

     FUNCTION ucy_rkck(ctx, n, t, yin, dydt, h, yout, yerr) BIND(C, NAME="ucy_rkck")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: n
       REAL(C_DOUBLE), VALUE :: t, h
       REAL(C_DOUBLE), INTENT(IN) :: yin(*), dydt(*)
       REAL(C_DOUBLE), INTENT(OUT) :: yout(*), yerr(*)
       INTEGER(C_INT) :: ucy_rkck
     END FUNCTION ucy_rkck

     FUNCTION ucy_rkqs(ctx, n, t, y, dydt, yscal, htry, hdid, hnext, epsilon, max_step) &
          BIND(C, NAME="ucy_rkqs")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: n
       REAL(C_DOUBLE), INTENT(INOUT) :: t, y(*), dydt(*)
       REAL(C_DOUBLE), INTENT(IN) :: yscal(*)
       REAL(C_DOUBLE), VALUE :: htry, epsilon, max_step
       REAL(C_DOUBLE), INTENT(OUT) :: hdid, hnext
       INTEGER(C_INT) :: ucy_rkqs
     END FUNCTION ucy_rkqs

     FUNCTION ucy_okada92_traction_kernels(ctx, x, sv, dv, nv, ns, y, L, W, strike, dip, &
          s, d, t, G, lambda, ts, td, tn) BIND(C, NAME="ucy_okada92_traction_kernels")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: ns
       REAL(C_DOUBLE), INTENT(IN) :: x(3), sv(3), dv(3), nv(3), y(3,*), L(*), W(*), &
                                     strike(*), dip(*)
       REAL(C_DOUBLE), VALUE :: s, d, t, G, lambda
       REAL(C_DOUBLE), INTENT(OUT) :: ts(*), td(*), tn(*)
       INTEGER(C_INT) :: ucy_okada92_traction_kernels
     END FUNCTION ucy_okada92_traction_kernels

     FUNCTION ucy_build_displacement(ctx, npts, x) BIND(C, NAME="ucy_build_displacement")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT), VALUE :: npts
       REAL(C_DOUBLE), INTENT(IN) :: x(3,*)
       INTEGER(C_INT) :: ucy_build_displacement
     END FUNCTION ucy_build_displacement

     FUNCTION ucy_displacement(ctx, y, d, df, dl) BIND(C, NAME="ucy_displacement")
       IMPORT :: C_PTR, C_INT, C_DOUBLE
       TYPE(C_PTR), VALUE :: ctx
       REAL(C_DOUBLE), INTENT(IN) :: y(*)
       REAL(C_DOUBLE), INTENT(OUT) :: d(*), df(*), dl(*)
       INTEGER(C_INT) :: ucy_displacement
     END FUNCTION ucy_displacement

     FUNCTION ucy_strlen(s) BIND(C, NAME="strlen")
       IMPORT :: C_PTR, C_SIZE_T
       TYPE(C_PTR), VALUE :: s
       INTEGER(C_SIZE_T) :: ucy_strlen
     END FUNCTION ucy_strlen
  END INTERFACE

CONTAINS

  !> STOP with the library's message on a nonzero status
  SUBROUTINE ucy_check(status, where)
    INTEGER(C_INT), INTENT(IN) :: status
    CHARACTER(LEN=*), INTENT(IN) :: where
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    TYPE(C_PTR) :: p
    INTEGER :: k, n
    IF (0 .EQ. status) RETURN
    p = ucy_last_error(ucy_ctx)
    WRITE (0,'("error in ",A,": status ",I0)') where, status
    IF (C_ASSOCIATED(p)) THEN
       n = INT(ucy_strlen(p))
       CALL C_F_POINTER(p, msg, [n])
       WRITE (0,'(1000A)') (msg(k), k=1,n)
    END IF
    STOP 1
  END SUBROUTINE ucy_check

  !> open the context on the GPU of this process
  SUBROUTINE ucy_init(device)
    INTEGER, INTENT(IN) :: device
    ucy_ctx = ucy_create(INT(device, C_INT), 0_C_INT)
    IF (.NOT. C_ASSOCIATED(ucy_ctx)) THEN
       WRITE (0,'("error: no GPU context on device ",I0)') device
       STOP 1
    END IF
  END SUBROUTINE ucy_init

  SUBROUTINE ucy_finalize()
    IF (C_ASSOCIATED(ucy_ctx)) CALL ucy_destroy(ucy_ctx)
    ucy_ctx = C_NULL_PTR
  END SUBROUTINE ucy_finalize

  !> mesh, friction and buildG (greens.f90:73-275) for rectangular patches
  SUBROUTINE ucy_setup_rectangles(lambda, mu, ns, x, length, width, strike, dip, &
                                  Vpl, rake, mu0, sig, a, b, L, Vo, damping)
    REAL*8, INTENT(IN) :: lambda, mu
    INTEGER, INTENT(IN) :: ns
    REAL*8, DIMENSION(3,ns), INTENT(IN) :: x
    REAL*8, DIMENSION(ns), INTENT(IN) :: length, width, strike, dip
    REAL*8, DIMENSION(ns), INTENT(IN) :: Vpl, rake, mu0, sig, a, b, L, Vo, damping
    CALL ucy_check(ucy_set_elastic(ucy_ctx, lambda, mu), "ucy_set_elastic")
    CALL ucy_check(ucy_set_rectangles(ucy_ctx, INT(ns, C_INT), x, length, width, strike, dip), &
                   "ucy_set_rectangles")
    CALL ucy_check(ucy_set_friction(ucy_ctx, INT(ns, C_INT), Vpl, rake, mu0, sig, a, b, L, Vo, &
                   damping), "ucy_set_friction")
    CALL ucy_check(ucy_build_kernel(ucy_ctx, UCY_PREC_F64), "buildG")
  END SUBROUTINE ucy_setup_rectangles

  !> the elements of any 3-D program (ratestate.f90:1128-1396, viscouscycles.f90:
  !! 1730-1820): rectangles x(3,nRect) (upper corners), length, width,
  !! strike, dip; triangles i1, i2, i3 (1-based) on vertices(3,nVe); vertical
  !! strain volumes vx(3,nVol) (upper corners), length, thickness, width,
  !! strike; angles in radians.  A count of zero skips its arrays.
  SUBROUTINE ucy_setup_mesh(lambda, mu, nRect, x, length, width, strike, dip, &
                            nTri, i1, i2, i3, nVe, vertices, &
                            nVol, vx, vlength, vthickness, vwidth, vstrike)
    REAL*8, INTENT(IN) :: lambda, mu
    INTEGER, INTENT(IN) :: nRect, nTri, nVe, nVol
    REAL*8, DIMENSION(3,*), INTENT(IN) :: x, vertices, vx
    REAL*8, DIMENSION(*), INTENT(IN) :: length, width, strike, dip
    INTEGER, DIMENSION(*), INTENT(IN) :: i1, i2, i3
    REAL*8, DIMENSION(*), INTENT(IN) :: vlength, vthickness, vwidth, vstrike
    INTEGER(C_INT), ALLOCATABLE :: j1(:), j2(:), j3(:)
    CALL ucy_check(ucy_set_elastic(ucy_ctx, lambda, mu), "ucy_set_elastic")
    IF (0 .LT. nRect) &
       CALL ucy_check(ucy_set_rectangles(ucy_ctx, INT(nRect, C_INT), x, length, width, strike, &
                      dip), "ucy_set_rectangles")
    IF (0 .LT. nTri) THEN
       ALLOCATE(j1(nTri), j2(nTri), j3(nTri))
       j1 = INT(i1(1:nTri), C_INT)
       j2 = INT(i2(1:nTri), C_INT)
       j3 = INT(i3(1:nTri), C_INT)
       CALL ucy_check(ucy_set_triangles(ucy_ctx, INT(nTri, C_INT), j1, j2, j3, &
                      INT(nVe, C_INT), vertices), "ucy_set_triangles")
       DEALLOCATE(j1, j2, j3)
    END IF
    IF (0 .LT. nVol) &
       CALL ucy_check(ucy_set_volumes(ucy_ctx, INT(nVol, C_INT), vx, vlength, vthickness, vwidth, &
                      vstrike), "ucy_set_volumes")
  END SUBROUTINE ucy_setup_mesh

  !> friction of every patch, rectangles then triangles (the decks' friction
  !! sections, ratestate.f90:1180-1216 and 1326-1362)
  SUBROUTINE ucy_setup_friction(np, Vpl, rake, mu0, sig, a, b, L, Vo, damping)
    INTEGER, INTENT(IN) :: np
    REAL*8, DIMENSION(np), INTENT(IN) :: Vpl, rake, mu0, sig, a, b, L, Vo, damping
    CALL ucy_check(ucy_set_friction(ucy_ctx, INT(np, C_INT), Vpl, rake, mu0, sig, a, b, L, Vo, &
                   damping), "ucy_set_friction")
  END SUBROUTINE ucy_setup_friction

  !> the volumes' nonlinear Maxwell and thermal properties (viscouscycles.f90:
  !! 1821-1930) with the loading strain rate e(6,nv) (e11, e12, e13, e22, e23, e33)
  SUBROUTINE ucy_setup_rheology(nv, gammadot0, npower, e, Gm, Q, R, To)
    INTEGER, INTENT(IN) :: nv
    REAL*8, DIMENSION(nv), INTENT(IN) :: gammadot0, npower, Gm, Q, R, To
    REAL*8, DIMENSION(6,nv), INTENT(IN) :: e
    CALL ucy_check(ucy_set_rheology(ucy_ctx, INT(nv, C_INT), gammadot0, npower, e, Gm, Q, R, &
                   To), "ucy_set_rheology")
  END SUBROUTINE ucy_setup_rheology

  !> buildG (greens.f90:73-275): K assembled and kept on the GPU
  SUBROUTINE buildG_gpu(precision)
    INTEGER, INTENT(IN) :: precision
    CALL ucy_check(ucy_build_kernel(ucy_ctx, INT(precision, C_INT)), "buildG")
  END SUBROUTINE buildG_gpu

  !> computeReferenceSystemOkada92 (okada92.f90:59-87)
  SUBROUTINE computeReferenceSystemOkada92_gpu(ns, x, L, W, strike, dip, sv, dv, nv, xc)
    INTEGER, INTENT(IN) :: ns
    REAL*8, DIMENSION(3,ns), INTENT(IN) :: x
    REAL*8, DIMENSION(ns), INTENT(IN) :: L, W, strike, dip
    REAL*8, DIMENSION(3,ns), INTENT(OUT) :: sv, dv, nv, xc
    CALL ucy_check(ucy_reference_system_okada92(INT(ns, C_INT), x, L, W, strike, dip, sv, dv, nv, &
                   xc), "computeReferenceSystemOkada92")
  END SUBROUTINE computeReferenceSystemOkada92_gpu

  !> computeReferenceSystemNikkhoo15 (nikkhoo15.f90:50-101, stuart97.f90:44-95)
  SUBROUTINE computeReferenceSystemNikkhoo15_gpu(nVe, v, ns, i1, i2, i3, sv, dv, nv, xc)
    INTEGER, INTENT(IN) :: nVe, ns
    REAL*8, DIMENSION(3,nVe), INTENT(IN) :: v
    INTEGER, DIMENSION(ns), INTENT(IN) :: i1, i2, i3
    REAL*8, DIMENSION(3,ns), INTENT(OUT) :: sv, dv, nv, xc
    CALL ucy_check(ucy_reference_system_nikkhoo15(INT(nVe, C_INT), v, INT(ns, C_INT), &
                   INT(i1, C_INT), INT(i2, C_INT), INT(i3, C_INT), sv, dv, nv, xc), &
                   "computeReferenceSystemNikkhoo15")
  END SUBROUTINE computeReferenceSystemNikkhoo15_gpu

  !> computeReferenceSystemVerticalStrainVolume (strainvolume.f90:135-163)
  SUBROUTINE computeReferenceSystemVerticalStrainVolume_gpu(ns, x, L, W, strike, sv, dv, nv, xc)
    INTEGER, INTENT(IN) :: ns
    REAL*8, DIMENSION(3,ns), INTENT(IN) :: x
    REAL*8, DIMENSION(ns), INTENT(IN) :: L, W, strike
    REAL*8, DIMENSION(3,ns), INTENT(OUT) :: sv, dv, nv, xc
    CALL ucy_check(ucy_reference_system_strainvolume(INT(ns, C_INT), x, L, W, strike, sv, dv, nv, &
                   xc), "computeReferenceSystemVerticalStrainVolume")
  END SUBROUTINE computeReferenceSystemVerticalStrainVolume_gpu

  !> initStateVector (ratestate.f90:548-610; variant 1: viscouscycles.f90:710-820)
  SUBROUTINE initStateVector_gpu(n, y, variant)
    INTEGER, INTENT(IN) :: n, variant
    REAL*8, DIMENSION(n), INTENT(OUT) :: y
    CALL ucy_check(ucy_init_state(ucy_ctx, INT(variant, C_INT), y), "initStateVector")
  END SUBROUTINE initStateVector_gpu

  !> odefun(n,time,y,dydt) of ratestate.f90:624; sets ucy_vmax
  SUBROUTINE odefun_gpu(n, time, y, dydt)
    INTEGER, INTENT(IN) :: n
    REAL*8, INTENT(IN) :: time
    REAL*8, DIMENSION(n), INTENT(IN) :: y
    REAL*8, DIMENSION(n), INTENT(INOUT) :: dydt
    CALL ucy_check(ucy_odefun(ucy_ctx, INT(n, C_INT), time, y, dydt, ucy_vmax), "odefun")
  END SUBROUTINE odefun_gpu

  !> RKCKm of ode45.f90:116 (ytmp and odefun unused: the stages run on the GPU)
  SUBROUTINE RKCKm_gpu(n, t, yin, dydt, h, yout, yerr, ytmp, odefun)
    INTEGER, INTENT(IN) :: n
    REAL*8, INTENT(IN) :: t, h
    REAL*8, DIMENSION(n), INTENT(IN) :: yin, dydt
    REAL*8, DIMENSION(n), INTENT(OUT) :: yout, yerr
    REAL*8, DIMENSION(n), INTENT(INOUT) :: ytmp
    EXTERNAL :: odefun
    CALL ucy_check(ucy_rkck(ucy_ctx, INT(n, C_INT), t, yin, dydt, h, yout, yerr), "RKCKm")
  END SUBROUTINE RKCKm_gpu

  !> RKQSm of ode45.f90:51 (ycand, yerr, ytmp and odefun unused)
  SUBROUTINE RKQSm_gpu(n, t, y, dydt, yscal, ycand, yerr, ytmp, htry, hdid, hnext, odefun)
    INTEGER, INTENT(IN) :: n
    REAL*8, INTENT(INOUT) :: t
    REAL*8, DIMENSION(n), INTENT(INOUT) :: y, dydt
    REAL*8, DIMENSION(n), INTENT(IN) :: yscal
    REAL*8, DIMENSION(n), INTENT(INOUT) :: ycand, yerr, ytmp
    REAL*8, INTENT(IN) :: htry
    REAL*8, INTENT(OUT) :: hdid, hnext
    EXTERNAL :: odefun
    CALL ucy_check(ucy_rkqs(ucy_ctx, INT(n, C_INT), t, y, dydt, yscal, htry, hdid, hnext, &
                   ucy_epsilon, ucy_maximum_time_step), "RKQSm")
  END SUBROUTINE RKQSm_gpu

  !> computeTractionKernelsOkada92 of okada92.f90:368 on the GPU
  SUBROUTINE computeTractionKernelsOkada92_gpu(x, sv, dv, nv, ns, y, L, W, strike, dip, &
                                               s, d, t, G, lambda, ts, td, tn)
    REAL*8, DIMENSION(3), INTENT(IN) :: x, sv, dv, nv
    INTEGER, INTENT(IN) :: ns
    REAL*8, DIMENSION(3,ns), INTENT(IN) :: y
    REAL*8, DIMENSION(ns), INTENT(IN) :: L, W, strike, dip
    REAL*8, INTENT(IN) :: s, d, t, G, lambda
    REAL*8, DIMENSION(ns), INTENT(OUT) :: ts, td, tn
    CALL ucy_check(ucy_okada92_traction_kernels(ucy_ctx, x, sv, dv, nv, INT(ns, C_INT), y, L, &
                   W, strike, dip, s, d, t, G, lambda, ts, td, tn), &
                   "computeTractionKernelsOkada92")
  END SUBROUTINE computeTractionKernelsOkada92_gpu

  !> buildO of greens.f90:302 for nObs observation points x(3,nObs) (north,
  !! east, down): the displacement kernels are assembled and kept on the GPU
  SUBROUTINE buildO_gpu(nObs, x)
    INTEGER, INTENT(IN) :: nObs
    REAL*8, DIMENSION(3,nObs), INTENT(IN) :: x
    CALL ucy_check(ucy_build_displacement(ucy_ctx, INT(nObs, C_INT), x), "buildO")
  END SUBROUTINE buildO_gpu

  !> the product of exportPoints (ratestate.f90:414-486; viscouscycles.f90:
  !! 523-625): d = O u, df = Of u, dl = Ol u, u gathered from the state y
  SUBROUTINE exportPoints_gpu(n, y, nd, d, df, dl)
    INTEGER, INTENT(IN) :: n, nd
    REAL*8, DIMENSION(n), INTENT(IN) :: y
    REAL*8, DIMENSION(nd), INTENT(OUT) :: d, df, dl
    CALL ucy_check(ucy_displacement(ucy_ctx, y, d, df, dl), "exportPoints")
  END SUBROUTINE exportPoints_gpu

END MODULE ucy_fortran
