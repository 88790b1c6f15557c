!-----------------------------------------------------------------------
! ref_harness.f90 -- C-callable wrappers around the reference's own
! Green's-function module procedures, for the oracle/_ref checker build.
!
! TEST INFRASTRUCTURE ONLY.  Compiled together with the reference sources
! (jdpmoore/unicycle fortran/3d/src: dc3d.f, okada92.f90, nikkhoo15.f90,
! StrainInHalfSpace.c, DispInHalfSpace.c, strainvolume.f90; and
! fortran/2d/antiplane/src/antiplane.f90), read where they lie, into
! oracle/_ref/libucyref.so.  Nothing here is product code.
!
! Each wrapper forwards its arguments unchanged to the reference routine
! named in its comment; only the calling convention changes (BIND(C),
! scalars by value).
!-----------------------------------------------------------------------
MODULE ucy_ref_harness
  USE ISO_C_BINDING
  IMPLICIT NONE
CONTAINS

  ! okada92.f90:59 computeReferenceSystemOkada92
  SUBROUTINE ref_refsys_okada92(ns,x,L,W,strike,dip,sv,dv,nv,xc) BIND(C,NAME="ref_refsys_okada92")
    USE okada92
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3,ns),L(ns),W(ns),strike(ns),dip(ns)
    REAL(C_DOUBLE), INTENT(OUT) :: sv(3,ns),dv(3,ns),nv(3,ns),xc(3,ns)
    CALL computeReferenceSystemOkada92(ns,x,L,W,strike,dip,sv,dv,nv,xc)
  END SUBROUTINE

  ! okada92.f90:368 computeTractionKernelsOkada92
  SUBROUTINE ref_traction_okada92(x,sv,dv,nv,ns,y,L,W,strike,dip,s,d,t,G,lambda,ts,td,tn) &
       BIND(C,NAME="ref_traction_okada92")
    USE okada92
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),L(ns),W(ns),strike(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: s,d,t,G,lambda
    REAL(C_DOUBLE), INTENT(OUT) :: ts(ns),td(ns),tn(ns)
    CALL computeTractionKernelsOkada92(x,sv,dv,nv,ns,y,L,W,strike,dip,s,d,t,G,lambda,ts,td,tn)
  END SUBROUTINE

  ! okada92.f90:424 computeStressKernelsOkada92
  SUBROUTINE ref_stress_okada92(x,sv,dv,nv,ns,y,L,W,strike,dip,s,d,t,G,lambda, &
       s11,s12,s13,s22,s23,s33) BIND(C,NAME="ref_stress_okada92")
    USE okada92
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),L(ns),W(ns),strike(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: s,d,t,G,lambda
    REAL(C_DOUBLE), INTENT(OUT) :: s11(ns),s12(ns),s13(ns),s22(ns),s23(ns),s33(ns)
    CALL computeStressKernelsOkada92(x,sv,dv,nv,ns,y,L,W,strike,dip,s,d,t,G,lambda, &
         s11,s12,s13,s22,s23,s33)
  END SUBROUTINE

  ! okada92.f90:109 computeStressOkada92 (one receiver, one source)
  SUBROUTINE ref_stress_point_okada92(x1,x2,x3,q1,q2,q3,L,W,strike,dipd,s,d,t,G,lambda,sig) &
       BIND(C,NAME="ref_stress_point_okada92")
    USE okada92
    REAL(C_DOUBLE), VALUE :: x1,x2,x3,q1,q2,q3,L,W,strike,dipd,s,d,t,G,lambda
    REAL(C_DOUBLE), INTENT(OUT) :: sig(6)
    CALL computeStressOkada92(x1,x2,x3,q1,q2,q3,L,W,strike,dipd,s,d,t,G,lambda, &
         sig(1),sig(2),sig(3),sig(4),sig(5),sig(6))
  END SUBROUTINE

  ! okada92.f90:230 computeDisplacementOkada92 (one point, one source)
  SUBROUTINE ref_displacement_okada92(x1,x2,x3,q1,q2,q3,L,W,strike,dipd,s,d,t,G,lambda,u) &
       BIND(C,NAME="ref_displacement_okada92")
    USE okada92
    REAL(C_DOUBLE), VALUE :: x1,x2,x3,q1,q2,q3,L,W,strike,dipd,s,d,t,G,lambda
    REAL(C_DOUBLE), INTENT(OUT) :: u(3)
    CALL computeDisplacementOkada92(x1,x2,x3,q1,q2,q3,L,W,strike,dipd,s,d,t,G,lambda, &
         u(1),u(2),u(3))
  END SUBROUTINE

  ! dc3d.f:1 DC3D (REAL*4 interface)
  SUBROUTINE ref_dc3d(alpha,x,y,z,depth,dip,al1,al2,aw1,aw2,d1,d2,d3,u,iret) BIND(C,NAME="ref_dc3d")
    REAL(C_FLOAT), VALUE :: alpha,x,y,z,depth,dip,al1,al2,aw1,aw2,d1,d2,d3
    REAL(C_FLOAT), INTENT(OUT) :: u(12)
    INTEGER(C_INT), INTENT(OUT) :: iret
    REAL*4 :: a4,x4,y4,z4,dep4,dip4,al14,al24,aw14,aw24,s1,s2,s3
    INTEGER :: ir
    EXTERNAL :: dc3d
    a4=alpha; x4=x; y4=y; z4=z; dep4=depth; dip4=dip
    al14=al1; al24=al2; aw14=aw1; aw24=aw2; s1=d1; s2=d2; s3=d3
    CALL dc3d(a4,x4,y4,z4,dep4,dip4,al14,al24,aw14,aw24,s1,s2,s3, &
              u(1),u(2),u(3),u(4),u(5),u(6),u(7),u(8),u(9),u(10),u(11),u(12),ir)
    iret=ir
  END SUBROUTINE

  ! nikkhoo15.f90:50 computeReferenceSystemNikkhoo15
  SUBROUTINE ref_refsys_nikkhoo15(nVe,v,ns,i1,i2,i3,sv,dv,nv,xc) BIND(C,NAME="ref_refsys_nikkhoo15")
    USE nikkhoo15
    INTEGER(C_INT), VALUE :: nVe,ns
    REAL(C_DOUBLE), INTENT(IN) :: v(3,nVe)
    INTEGER(C_INT), INTENT(IN) :: i1(ns),i2(ns),i3(ns)
    REAL(C_DOUBLE), INTENT(OUT) :: sv(3,ns),dv(3,ns),nv(3,ns),xc(3,ns)
    INTEGER :: j1(ns),j2(ns),j3(ns)
    j1=i1; j2=i2; j3=i3
    CALL computeReferenceSystemNikkhoo15(nVe,v,ns,j1,j2,j3,sv,dv,nv,xc)
  END SUBROUTINE

  ! nikkhoo15.f90:187 computeTractionKernelsNikkhoo15
  SUBROUTINE ref_traction_nikkhoo15(x,sv,dv,nv,ns,i1,i2,i3,nVe,vert,s,d,t,G,lambda,ts,td,tn) &
       BIND(C,NAME="ref_traction_nikkhoo15")
    USE nikkhoo15
    INTEGER(C_INT), VALUE :: ns,nVe
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),vert(3,nVe)
    INTEGER(C_INT), INTENT(IN) :: i1(ns),i2(ns),i3(ns)
    REAL(C_DOUBLE), VALUE :: s,d,t,G,lambda
    REAL(C_DOUBLE), INTENT(OUT) :: ts(ns),td(ns),tn(ns)
    INTEGER :: j1(ns),j2(ns),j3(ns)
    j1=i1; j2=i2; j3=i3
    CALL computeTractionKernelsNikkhoo15(x,sv,dv,nv,ns,j1,j2,j3,nVe,vert,s,d,t,G,lambda,ts,td,tn)
  END SUBROUTINE

  ! nikkhoo15.f90:250 computeStressKernelsNikkhoo15
  SUBROUTINE ref_stress_nikkhoo15(x,sv,dv,nv,ns,i1,i2,i3,nVe,vert,s,d,t,G,lambda, &
       s11,s12,s13,s22,s23,s33) BIND(C,NAME="ref_stress_nikkhoo15")
    USE nikkhoo15
    INTEGER(C_INT), VALUE :: ns,nVe
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),vert(3,nVe)
    INTEGER(C_INT), INTENT(IN) :: i1(ns),i2(ns),i3(ns)
    REAL(C_DOUBLE), VALUE :: s,d,t,G,lambda
    REAL(C_DOUBLE), INTENT(OUT) :: s11(ns),s12(ns),s13(ns),s22(ns),s23(ns),s33(ns)
    INTEGER :: j1(ns),j2(ns),j3(ns)
    j1=i1; j2=i2; j3=i3
    CALL computeStressKernelsNikkhoo15(x,sv,dv,nv,ns,j1,j2,j3,nVe,vert,s,d,t,G,lambda, &
         s11,s12,s13,s22,s23,s33)
  END SUBROUTINE

  ! nikkhoo15.f90:324 computeDisplacementNikkhoo15 (one point, one triangle)
  SUBROUTINE ref_displacement_nikkhoo15(x1,x2,x3,P1,P2,P3,s,d,t,G,lambda,u) &
       BIND(C,NAME="ref_displacement_nikkhoo15")
    USE nikkhoo15
    REAL(C_DOUBLE), VALUE :: x1,x2,x3,s,d,t,G,lambda
    REAL(C_DOUBLE), INTENT(IN) :: P1(3),P2(3),P3(3)
    REAL(C_DOUBLE), INTENT(OUT) :: u(3)
    CALL computeDisplacementNikkhoo15(x1,x2,x3,P1,P2,P3,s,d,t,G,lambda,u(1),u(2),u(3))
  END SUBROUTINE

  ! strainvolume.f90:135 computeReferenceSystemVerticalStrainVolume
  SUBROUTINE ref_refsys_strainvolume(ns,x,L,W,strike,sv,dv,nv,xc) BIND(C,NAME="ref_refsys_strainvolume")
    USE strainvolume
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3,ns),L(ns),W(ns),strike(ns)
    REAL(C_DOUBLE), INTENT(OUT) :: sv(3,ns),dv(3,ns),nv(3,ns),xc(3,ns)
    CALL computeReferenceSystemVerticalStrainVolume(ns,x,L,W,strike,sv,dv,nv,xc)
  END SUBROUTINE

  ! strainvolume.f90:4671 computeTractionKernelsVerticalStrainVolume
  SUBROUTINE ref_traction_strainvolume(x,sv,dv,nv,ns,y,L,T,W,strike,e11,e12,e13,e22,e23,e33, &
       G,lambda,ts,td,tn) BIND(C,NAME="ref_traction_strainvolume")
    USE strainvolume
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),L(ns),T(ns),W(ns),strike(ns)
    REAL(C_DOUBLE), VALUE :: e11,e12,e13,e22,e23,e33,G,lambda
    REAL(C_DOUBLE), INTENT(OUT) :: ts(ns),td(ns),tn(ns)
    CALL computeTractionKernelsVerticalStrainVolume(x,sv,dv,nv,ns,y,L,T,W,strike, &
         e11,e12,e13,e22,e23,e33,G,lambda,ts,td,tn)
  END SUBROUTINE

  ! strainvolume.f90:4743 computeStressKernelsVerticalStrainVolume
  SUBROUTINE ref_stress_strainvolume(x,sv,dv,nv,ns,y,L,T,W,strike,e11,e12,e13,e22,e23,e33, &
       G,lambda,s11,s12,s13,s22,s23,s33) BIND(C,NAME="ref_stress_strainvolume")
    USE strainvolume
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),L(ns),T(ns),W(ns),strike(ns)
    REAL(C_DOUBLE), VALUE :: e11,e12,e13,e22,e23,e33,G,lambda
    REAL(C_DOUBLE), INTENT(OUT) :: s11(ns),s12(ns),s13(ns),s22(ns),s23(ns),s33(ns)
    CALL computeStressKernelsVerticalStrainVolume(x,sv,dv,nv,ns,y,L,T,W,strike, &
         e11,e12,e13,e22,e23,e33,G,lambda,s11,s12,s13,s22,s23,s33)
  END SUBROUTINE

  ! strainvolume.f90:215 computeDisplacementVerticalStrainVolume (for the
  ! finite-difference consistency check of the stress routine)
  SUBROUTINE ref_displacement_strainvolume(x1,x2,x3,q1,q2,q3,L,T,W,theta, &
       e11,e12,e13,e22,e23,e33,G,nu,u1,u2,u3) BIND(C,NAME="ref_displacement_strainvolume")
    USE strainvolume
    REAL(C_DOUBLE), VALUE :: x1,x2,x3,q1,q2,q3,L,T,W,theta,e11,e12,e13,e22,e23,e33,G,nu
    REAL(C_DOUBLE), INTENT(OUT) :: u1,u2,u3
    CALL computeDisplacementVerticalStrainVolume(x1,x2,x3,q1,q2,q3,L,T,W,theta, &
         e11,e12,e13,e22,e23,e33,G,nu,u1,u2,u3)
  END SUBROUTINE

  ! fortran/2d/antiplane/src/antiplane.f90:131 computeStressAntiplane
  SUBROUTINE ref_stress_antiplane(x2,x3,q2,q3,W,dip,G,s12,s13) BIND(C,NAME="ref_stress_antiplane")
    USE antiplane
    REAL(C_DOUBLE), VALUE :: x2,x3,q2,q3,W,dip,G
    REAL(C_DOUBLE), INTENT(OUT) :: s12,s13
    CALL computeStressAntiplane(x2,x3,q2,q3,W,dip,G,s12,s13)
  END SUBROUTINE

  ! antiplane.f90:352 computeTractionKernelsAntiplane
  SUBROUTINE ref_traction_antiplane(x,sv,dv,nv,ns,y,W,dip,G,ts) BIND(C,NAME="ref_traction_antiplane")
    USE antiplane
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: G
    REAL(C_DOUBLE), INTENT(OUT) :: ts(ns)
    CALL computeTractionKernelsAntiplane(x,sv,dv,nv,ns,y,W,dip,G,ts)
  END SUBROUTINE

  ! antiplane.f90:81 computeReferenceSystemAntiplane
  SUBROUTINE ref_refsys_antiplane(ns,x,W,dip,sv,dv,nv,xc) BIND(C,NAME="ref_refsys_antiplane")
    USE antiplane
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3,ns),W(ns),dip(ns)
    REAL(C_DOUBLE), INTENT(OUT) :: sv(3,ns),dv(3,ns),nv(3,ns),xc(3,ns)
    CALL computeReferenceSystemAntiplane(ns,x,W,dip,sv,dv,nv,xc)
  END SUBROUTINE

  ! antiplane.f90:479 computeStressStrainVolumeAntiplane
  SUBROUTINE ref_stress_sv_antiplane(x2,x3,q2,q3,T,W,dip,e12,e13,G,s12,s13) &
       BIND(C,NAME="ref_stress_sv_antiplane")
    USE antiplane
    REAL(C_DOUBLE), VALUE :: x2,x3,q2,q3,T,W,dip,e12,e13,G
    REAL(C_DOUBLE), INTENT(OUT) :: s12,s13
    CALL computeStressStrainVolumeAntiplane(x2,x3,q2,q3,T,W,dip,e12,e13,G,s12,s13)
  END SUBROUTINE

  ! antiplane.f90:213 computeDisplacementStrainVolumeAntiplane
  SUBROUTINE ref_displacement_sv_antiplane(x2,x3,q2,q3,T,W,dip,e12,e13,u1) &
       BIND(C,NAME="ref_displacement_sv_antiplane")
    USE antiplane
    REAL(C_DOUBLE), VALUE :: x2,x3,q2,q3,T,W,dip,e12,e13
    REAL(C_DOUBLE), INTENT(OUT) :: u1
    CALL computeDisplacementStrainVolumeAntiplane(x2,x3,q2,q3,T,W,dip,e12,e13,u1)
  END SUBROUTINE

  ! antiplane.f90:543 computeTractionKernelsStrainVolumeAntiplane
  SUBROUTINE ref_traction_sv_antiplane(x,sv,dv,nv,ns,y,T,W,dip,e12,e13,G,ts) &
       BIND(C,NAME="ref_traction_sv_antiplane")
    USE antiplane
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),T(ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: e12,e13,G
    REAL(C_DOUBLE), INTENT(OUT) :: ts(ns)
    CALL computeTractionKernelsStrainVolumeAntiplane(x,sv,dv,nv,ns,y,T,W,dip,e12,e13,G,ts)
  END SUBROUTINE

  ! antiplane.f90:600 computeStressKernelsStrainVolumeAntiplane
  SUBROUTINE ref_stress_kernels_sv_antiplane(x,sv,dv,nv,ns,y,T,W,dip,e12,e13,G,s12,s13) &
       BIND(C,NAME="ref_stress_kernels_sv_antiplane")
    USE antiplane
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),T(ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: e12,e13,G
    REAL(C_DOUBLE), INTENT(OUT) :: s12(ns),s13(ns)
    CALL computeStressKernelsStrainVolumeAntiplane(x,sv,dv,nv,ns,y,T,W,dip,e12,e13,G,s12,s13)
  END SUBROUTINE

  ! antiplane.f90:401 computeStressKernelsAntiplane
  SUBROUTINE ref_stress_kernels_antiplane(x,sv,dv,nv,ns,y,W,dip,G,s12,s13) &
       BIND(C,NAME="ref_stress_kernels_antiplane")
    USE antiplane
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: G
    REAL(C_DOUBLE), INTENT(OUT) :: s12(ns),s13(ns)
    CALL computeStressKernelsAntiplane(x,sv,dv,nv,ns,y,W,dip,G,s12,s13)
  END SUBROUTINE

  ! fortran/2d/planestrain/src/computeStressStrainVolumePlaneStrain.f90:53
  SUBROUTINE ref_stress_sv_planestrain(r2,x3,q2,q3,T,W,dip,e22,e23,e33,G,lam,s22,s23,s33) &
       BIND(C,NAME="ref_stress_sv_planestrain")
    REAL(C_DOUBLE), VALUE :: r2,x3,q2,q3,T,W,dip,e22,e23,e33,G,lam
    REAL(C_DOUBLE), INTENT(OUT) :: s22,s23,s33
    EXTERNAL :: computeStressStrainVolumePlaneStrain
    CALL computeStressStrainVolumePlaneStrain(r2,x3,q2,q3,T,W,dip,e22,e23,e33,G,lam,s22,s23,s33)
  END SUBROUTINE

  ! computeDisplacementStrainVolumePlaneStrain.f90:54
  SUBROUTINE ref_displacement_sv_planestrain(r2,x3,q2,q3,T,W,dip,e22,e23,e33,G,lam,u2,u3) &
       BIND(C,NAME="ref_displacement_sv_planestrain")
    REAL(C_DOUBLE), VALUE :: r2,x3,q2,q3,T,W,dip,e22,e23,e33,G,lam
    REAL(C_DOUBLE), INTENT(OUT) :: u2,u3
    EXTERNAL :: computeDisplacementStrainVolumePlaneStrain
    CALL computeDisplacementStrainVolumePlaneStrain(r2,x3,q2,q3,T,W,dip,e22,e23,e33,G,lam,u2,u3)
  END SUBROUTINE

  ! planestrain.f90:105 computeStressPlaneStrain
  SUBROUTINE ref_stress_planestrain(x2,x3,q2,q3,W,dip,s,G,lam,s22,s23,s33) &
       BIND(C,NAME="ref_stress_planestrain")
    USE planestrain
    REAL(C_DOUBLE), VALUE :: x2,x3,q2,q3,W,dip,s,G,lam
    REAL(C_DOUBLE), INTENT(OUT) :: s22,s23,s33
    CALL computeStressPlaneStrain(x2,x3,q2,q3,W,dip,s,G,lam,s22,s23,s33)
  END SUBROUTINE

  ! planestrain.f90:146 computeDisplacementPlaneStrain
  SUBROUTINE ref_displacement_planestrain(x2,x3,q2,q3,W,dip,s,G,lam,u2,u3) &
       BIND(C,NAME="ref_displacement_planestrain")
    USE planestrain
    REAL(C_DOUBLE), VALUE :: x2,x3,q2,q3,W,dip,s,G,lam
    REAL(C_DOUBLE), INTENT(OUT) :: u2,u3
    CALL computeDisplacementPlaneStrain(x2,x3,q2,q3,W,dip,s,G,lam,u2,u3)
  END SUBROUTINE

  ! planestrain.f90:266 computeTractionKernelsPlaneStrain
  SUBROUTINE ref_traction_planestrain(x,sv,dv,nv,ns,y,W,dip,G,lam,td,tn) &
       BIND(C,NAME="ref_traction_planestrain")
    USE planestrain
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: G,lam
    REAL(C_DOUBLE), INTENT(OUT) :: td(ns),tn(ns)
    CALL computeTractionKernelsPlaneStrain(x,sv,dv,nv,ns,y,W,dip,G,lam,td,tn)
  END SUBROUTINE

  ! planestrain.f90:314 computeStressKernelsPlaneStrain (its own argument
  ! order: lambda before G)
  SUBROUTINE ref_stress_kernels_planestrain(x,sv,dv,nv,ns,y,W,dip,lam,G,s22,s23,s33) &
       BIND(C,NAME="ref_stress_kernels_planestrain")
    USE planestrain
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: lam,G
    REAL(C_DOUBLE), INTENT(OUT) :: s22(ns),s23(ns),s33(ns)
    CALL computeStressKernelsPlaneStrain(x,sv,dv,nv,ns,y,W,dip,lam,G,s22,s23,s33)
  END SUBROUTINE

  ! planestrain.f90:375 computeTractionKernelsStrainVolumePlaneStrain
  SUBROUTINE ref_traction_sv_planestrain(x,sv,dv,nv,ns,y,T,W,dip,e22,e23,e33,G,lam,td,tn) &
       BIND(C,NAME="ref_traction_sv_planestrain")
    USE planestrain
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),T(ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: e22,e23,e33,G,lam
    REAL(C_DOUBLE), INTENT(OUT) :: td(ns),tn(ns)
    CALL computeTractionKernelsStrainVolumePlaneStrain(x,sv,dv,nv,ns,y,T,W,dip,e22,e23,e33, &
         G,lam,td,tn)
  END SUBROUTINE

  ! planestrain.f90:435 computeStressKernelsStrainVolumePlaneStrain
  SUBROUTINE ref_stress_kernels_sv_planestrain(x,sv,dv,nv,ns,y,T,W,dip,e22,e23,e33,G,lam, &
       s22,s23,s33) BIND(C,NAME="ref_stress_kernels_sv_planestrain")
    USE planestrain
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),sv(3),dv(3),nv(3),y(3,ns),T(ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: e22,e23,e33,G,lam
    REAL(C_DOUBLE), INTENT(OUT) :: s22(ns),s23(ns),s33(ns)
    CALL computeStressKernelsStrainVolumePlaneStrain(x,sv,dv,nv,ns,y,T,W,dip,e22,e23,e33, &
         G,lam,s22,s23,s33)
  END SUBROUTINE

  ! planestrain.f90:221 computeDisplacementKernelsStrainVolumePlaneStrain
  SUBROUTINE ref_displacement_kernels_sv_planestrain(x,ns,y,T,W,dip,e22,e23,e33,G,lam,u2,u3) &
       BIND(C,NAME="ref_displacement_kernels_sv_planestrain")
    USE planestrain
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3),y(3,ns),T(ns),W(ns),dip(ns)
    REAL(C_DOUBLE), VALUE :: e22,e23,e33,G,lam
    REAL(C_DOUBLE), INTENT(OUT) :: u2(ns),u3(ns)
    CALL computeDisplacementKernelsStrainVolumePlaneStrain(x,ns,y,T,W,dip,e22,e23,e33, &
         G,lam,u2,u3)
  END SUBROUTINE

  ! planestrain.f90:55 computeReferenceSystemPlaneStrain
  SUBROUTINE ref_refsys_planestrain(ns,x,W,dip,sv,dv,nv,xc) BIND(C,NAME="ref_refsys_planestrain")
    USE planestrain
    INTEGER(C_INT), VALUE :: ns
    REAL(C_DOUBLE), INTENT(IN) :: x(3,ns),W(ns),dip(ns)
    REAL(C_DOUBLE), INTENT(OUT) :: sv(3,ns),dv(3,ns),nv(3,ns),xc(3,ns)
    CALL computeReferenceSystemPlaneStrain(ns,x,W,dip,sv,dv,nv,xc)
  END SUBROUTINE

END MODULE ucy_ref_harness
