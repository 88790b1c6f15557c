// ps_jint.h -- displacement of a dipping strain volume (shear zone) in a
// half space under plane strain, Barbot, Moore & Lambert (2017, BSSA
// 107(2)), as used by jdpmoore/unicycle fortran/2d/planestrain/src/
// computeDisplacementStrainVolumePlaneStrain.f90 (IU2, IU3 and the eight
// integrals I223 .. I332) and, for line dislocations, planestrain.f90:105-164
// (a zone 1e-4 of its width thick with the shear strain -s / (2 T)).
//
// The reference's stress routine (computeStressStrainVolumePlaneStrain.f90)
// uses hand-differentiated integrals; here the displacement gradient is the
// exact derivative of the displacement integrals, formed by forward-mode
// differentiation with respect to the observation point (x2, x3).  The
// integrals keep the reference's operand order term by term, so the
// value-only evaluation rounds as the reference does.
//
// T provides +, -, *, / with T and double, .val(), and ps_sqrt, ps_log and
// ps_atan2 (atan2 of T and a double).  The observation point enters as
// x2 = r2 - q2 and x3; the source's q3 and dip (through its sines and
// cosines of 1..4 dip, formed once with the math library as the reference
// forms SIN(k dip)) and Poisson's ratio are plain doubles.
#pragma once

#if defined(__HIPCC__)
#define PS_FN __host__ __device__ __forceinline__
#else
#define PS_FN inline
#endif

namespace psj {

struct Geom {
    double q3, nu;
    double s1, c1, s2, c2, s3, c3, s4, c4;   // SIN(k dip), COS(k dip), k = 1..4
};

template <class T>
struct Point {
    T x2, x3;   // r2 - q2 and x3
};

// computeDisplacementStrainVolumePlaneStrain.f90:144-216
template <class T>
PS_FN T y2f(const Geom &g, double y2p, double y3p)
{
    return T(y2p * g.s1 + y3p * g.c1);
}
template <class T>
PS_FN T p2f(const Point<T> &P, const Geom &g, double y2p)
{
    return (P.x3 - g.q3) * g.c1 - g.s1 * P.x2 + y2p;
}
template <class T>
PS_FN T p3f(const Point<T> &P, const Geom &g, double y3p)
{
    return -((P.x3 - g.q3) * g.s1) - g.c1 * P.x2 + y3p;
}
template <class T>
PS_FN T p2pf(const Point<T> &P, const Geom &g, double y2p)
{
    return -((g.q3 + P.x3) * g.c1) - g.s1 * P.x2 + y2p;
}
template <class T>
PS_FN T p3pf(const Point<T> &P, const Geom &g, double y3p)
{
    return (g.q3 + P.x3) * g.s1 - g.c1 * P.x2 + y3p;
}
// squared distances to the source point and to its image, in the reference's
// expanded form
template <class T>
PS_FN T r12pf(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3;
    return q3 * q3 + x2 * x2 + (-2.0) * q3 * x3 + x3 * x3 + y2p * y2p + y3p * y3p
           + (-2.0) * q3 * y2p * g.c1 + 2.0 * x3 * y2p * g.c1 + (-2.0) * x2 * y3p * g.c1
           + (-2.0) * x2 * y2p * g.s1 + 2.0 * q3 * y3p * g.s1 + (-2.0) * x3 * y3p * g.s1;
}
template <class T>
PS_FN T r22pf(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3;
    return q3 * q3 + x2 * x2 + 2.0 * q3 * x3 + x3 * x3 + y2p * y2p + y3p * y3p
           + (-2.0) * q3 * y2p * g.c1 + (-2.0) * x3 * y2p * g.c1 + (-2.0) * x2 * y3p * g.c1
           + (-2.0) * x2 * y2p * g.s1 + 2.0 * q3 * y3p * g.s1 + 2.0 * x3 * y3p * g.s1;
}
// r2p^-2 with r2p the image distance formed from y2, y3 (:165-170)
template <class T>
PS_FN T inv_r2p2(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const double y2 = y2p * g.s1 + y3p * g.c1, y3 = -y2p * g.c1 + y3p * g.s1 + g.q3;
    const T a = P.x3 + y3, b = P.x2 - y2;
    const T r = ps_sqrt(a * a + b * b);
    return 1.0 / (r * r);
}

// I223 (:222-261)
template <class T>
PS_FN T I223(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3, nu = g.nu;
    const T p2 = p2f(P, g, y2p), p3 = p3f(P, g, y3p), p2p = p2pf(P, g, y2p), p3p = p3pf(P, g, y3p);
    return (-1.0) * p2 * ps_atan2(p3 / p2, 1.0) * (3.0 - 4.0 * nu + g.c2)
           + (1.0 / 4.0) * ps_log(r12pf(P, g, y2p, y3p))
                 * (-6.0 * y3p + 8.0 * y3p * nu + x2 * (5.0 - 8.0 * nu) * g.c1 + x2 * g.c3
                    - 2.0 * (q3 - x3) * (4.0 - 4.0 * nu + g.c2) * g.s1 + 2.0 * y2p * g.s2)
           + (1.0 / 2.0) * ps_atan2(-p3p / p2p, 1.0)
                 * (2.0 * y2p * (5.0 + 4.0 * nu * (-3.0 + 2.0 * nu))
                    - (13.0 * q3 + 11.0 * x3 - 28.0 * (q3 + x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.c1
                    - 2.0 * y2p * (-3.0 + 4.0 * nu) * g.c2
                    - (3.0 * q3 + 5.0 * x3 - 4.0 * (q3 + x3) * nu) * g.c3
                    - x2 * (7.0 + 4.0 * nu * (-5.0 + 4.0 * nu)) * g.s1 - x2 * (3.0 - 4.0 * nu) * g.s3)
           + (1.0 / 4.0) * ps_log(r22pf(P, g, y2p, y3p))
                 * (-2.0 * y3p * (5.0 + 4.0 * nu * (-3.0 + 2.0 * nu))
                    + x2 * (7.0 + 4.0 * nu * (-5.0 + 4.0 * nu)) * g.c1 + x2 * (3.0 - 4.0 * nu) * g.c3
                    - (13.0 * q3 + 11.0 * x3 - 28.0 * (q3 + x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.s1
                    + 2.0 * y2p * (3.0 - 4.0 * nu) * g.s2
                    + (-3.0 * q3 - 5.0 * x3 + 4.0 * (q3 + x3) * nu) * g.s3)
           + inv_r2p2(P, g, y2p, y3p) * x3
                 * (-2.0 * q3 * x2 * g.c1 + (3.0 * x2 * y2p + q3 * y3p - x3 * y3p) * g.c2
                    - 2.0 * y2p * y3p * g.c3 + (-x2 * y2p + (q3 + x3) * y3p) * g.c4
                    + (-(q3 * q3) + x2 * x2 + x3 * x3) * g.s1
                    + (3.0 * q3 * y2p + x3 * y2p - x2 * y3p) * g.s2
                    - (x2 * x2 + (q3 + x3) * (q3 + x3) + 2.0 * (y2p * y2p)) * g.s3
                    + ((q3 + x3) * y2p + x2 * y3p) * g.s4);
}

// I222 (:267-307)
template <class T>
PS_FN T I222(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3, nu = g.nu;
    const T p2 = p2f(P, g, y2p), p3 = p3f(P, g, y3p), p2p = p2pf(P, g, y2p), p3p = p3pf(P, g, y3p);
    return p3 * ps_atan2(p2 / p3, 1.0) * (-3.0 + 4.0 * nu + g.c2)
           + (1.0 / 4.0) * ps_log(r12pf(P, g, y2p, y3p))
                 * (-6.0 * y2p + 8.0 * y2p * nu - 2.0 * (q3 - x3) * g.c1 * (-4.0 + 4.0 * nu + g.c2)
                    - 2.0 * x2 * (-2.0 + 4.0 * nu + g.c2) * g.s1 + 2.0 * y3p * g.s2)
           + (1.0 / 4.0) * ps_log(r22pf(P, g, y2p, y3p))
                 * (-2.0 * y2p * (5.0 + 4.0 * nu * (-3.0 + 2.0 * nu))
                    + (13.0 * q3 + 11.0 * x3 - 28.0 * (q3 + x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.c1
                    + (-3.0 * q3 - 5.0 * x3 + 4.0 * (q3 + x3) * nu) * g.c3
                    + x2 * (7.0 + 4.0 * nu * (-5.0 + 4.0 * nu)) * g.s1
                    + 2.0 * y3p * (3.0 - 4.0 * nu) * g.s2 + x2 * (-3.0 + 4.0 * nu) * g.s3)
           + (1.0 / 2.0) * ps_atan2(p2p / p3p, 1.0)
                 * (-2.0 * y3p * (5.0 + 4.0 * nu * (-3.0 + 2.0 * nu))
                    - x2 * (-7.0 + 4.0 * (5.0 - 4.0 * nu) * nu) * g.c1
                    - 2.0 * y3p * (-3.0 + 4.0 * nu) * g.c2 - x2 * (3.0 - 4.0 * nu) * g.c3
                    - (13.0 * q3 + 11.0 * x3) * g.s1 - 4.0 * (q3 + x3) * nu * (-7.0 + 4.0 * nu) * g.s1
                    - (-3.0 * q3 - 5.0 * x3 + 4.0 * (q3 + x3) * nu) * g.s3)
           + inv_r2p2(P, g, y2p, y3p) * x3
                 * (-(-(q3 * q3) + x2 * x2 + x3 * x3) * g.c1
                    + (-q3 * y2p + x3 * y2p + 3.0 * x2 * y3p) * g.c2
                    - (x2 * x2 + (q3 + x3) * (q3 + x3) + 2.0 * (y3p * y3p)) * g.c3
                    + ((q3 + x3) * y2p + x2 * y3p) * g.c4 - 2.0 * q3 * x2 * g.s1
                    + (x2 * y2p + (3.0 * q3 + x3) * y3p) * g.s2 - 2.0 * y2p * y3p * g.s3
                    + (x2 * y2p - (q3 + x3) * y3p) * g.s4);
}

// I233 (:313-345)
template <class T>
PS_FN T I233(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3, nu = g.nu;
    const double y2 = y2p * g.s1 + y3p * g.c1, y3 = -y2p * g.c1 + y3p * g.s1 + g.q3;
    const T p2 = p2f(P, g, y2p), p3 = p3f(P, g, y3p), p2p = p2pf(P, g, y2p), p3p = p3pf(P, g, y3p);
    return 4.0 * y3p * (-1.0 + nu) * (-1.0 + 2.0 * nu) * ps_atan2((x2 - y2) / (x3 + y3), 1.0)
           + (-1.0 / 2.0) * p2 * g.c2 * ps_log(r12pf(P, g, y2p, y3p))
           + 4.0 * (-1.0 + nu) * (-1.0 + 2.0 * nu) * ps_atan2(p2p * (1.0 / p3p), 1.0)
                 * (x2 * g.c1 - (q3 + x3) * g.s1)
           - p2 * ps_atan2((1.0 / p2) * p3, 1.0) * g.s2
           + ps_atan2((-1.0) * (1.0 / p2p) * p3p, 1.0) * g.s1
                 * (2.0 * y2p * (3.0 + (-4.0) * nu) * g.c1
                    + (-3.0 * q3 - x3 + 4.0 * (q3 + x3) * nu) * g.c2
                    + (-3.0 + 4.0 * nu) * (q3 - x3 + x2 * g.s2))
           + 0.25 * ps_log(r22pf(P, g, y2p, y3p))
                 * (-8.0 * y2p * (-1.0 + nu) * (-1.0 + 2.0 * nu)
                    + (11.0 * q3 + x3 - 4.0 * (7.0 * q3 + 3.0 * x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.c1
                    + 2.0 * y2p * (-3.0 + 4.0 * nu) * g.c2
                    + (3.0 * q3 + x3 - 4.0 * (q3 + x3) * nu) * g.c3
                    + x2 * (5.0 + 4.0 * nu * (-5.0 + 4.0 * nu)) * g.s1 + x2 * (3.0 - 4.0 * nu) * g.s3)
           + inv_r2p2(P, g, y2p, y3p) * x3
                 * ((-(q3 * q3) + x2 * x2 + x3 * x3) * g.c1
                    + (3.0 * q3 * y2p + x3 * y2p - x2 * y3p) * g.c2
                    + (-1.0) * (x2 * x2 + (q3 + x3) * (q3 + x3) + 2.0 * (y2p * y2p)) * g.c3
                    + ((q3 + x3) * y2p + x2 * y3p) * g.c4 + 2.0 * q3 * x2 * g.s1
                    + (-3.0 * x2 * y2p - q3 * y3p + x3 * y3p) * g.s2 + 2.0 * y2p * y3p * g.s3
                    + (x2 * y2p - (q3 + x3) * y3p) * g.s4);
}

// I232 (:351-384)
template <class T>
PS_FN T I232(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3, nu = g.nu;
    const T p3 = p3f(P, g, y3p), p2 = p2f(P, g, y2p), p2p = p2pf(P, g, y2p), p3p = p3pf(P, g, y3p);
    return 4.0 * y2p * (-1.0 + nu) * (-1.0 + 2.0 * nu)
               * ps_atan2((x2 - y3p * g.c1 - y2p * g.s1) / (q3 + x3 - y2p * g.c1 + y3p * g.s1), 1.0)
           - (1.0 / 2.0) * p3 * g.c2 * ps_log(r12pf(P, g, y2p, y3p))
           - 4.0 * (-1.0 + nu) * (-1.0 + 2.0 * nu) * ps_atan2((1.0 / p2p) * p3p, 1.0)
                 * ((q3 + x3) * g.c1 + x2 * g.s1)
           + p3 * ps_atan2(p2 * (1.0 / p3), 1.0) * g.s2
           - ps_atan2(p2p * (1.0 / p3p), 1.0) * g.c1
                 * ((3.0 * q3 + x3 - 4.0 * (q3 + x3) * nu) * g.c2
                    + (-3.0 + 4.0 * nu) * (q3 - x3 + 2.0 * y3p * g.s1 - x2 * g.s2))
           + 0.25 * ps_log(r22pf(P, g, y2p, y3p))
                 * (8.0 * y3p * (-1.0 + nu) * (-1.0 + 2.0 * nu)
                    + x2 * (-5.0 + 4.0 * (5.0 - 4.0 * nu) * nu) * g.c1
                    + 2.0 * y3p * (-3.0 + 4.0 * nu) * g.c2 + x2 * (3.0 - 4.0 * nu) * g.c3
                    + (11.0 * q3 + x3 - 4.0 * (7.0 * q3 + 3.0 * x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.s1
                    + (-3.0 * q3 - x3 + 4.0 * (q3 + x3) * nu) * g.s3)
           + inv_r2p2(P, g, y2p, y3p) * x3
                 * (-2.0 * q3 * x2 * g.c1 + (x2 * y2p + 3.0 * q3 * y3p + x3 * y3p) * g.c2
                    - 2.0 * y2p * y3p * g.c3 + (x2 * y2p - (q3 + x3) * y3p) * g.c4
                    + (-(q3 * q3) + x2 * x2 + x3 * x3) * g.s1
                    + (q3 * y2p - x3 * y2p - 3.0 * x2 * y3p) * g.s2
                    + (x2 * x2 + (q3 + x3) * (q3 + x3) + 2.0 * (y3p * y3p)) * g.s3
                    - ((q3 + x3) * y2p + x2 * y3p) * g.s4);
}

// I323 (:390-426)
template <class T>
PS_FN T I323(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3, nu = g.nu;
    const T p2 = p2f(P, g, y2p), p3 = p3f(P, g, y3p), p2p = p2pf(P, g, y2p), p3p = p3pf(P, g, y3p);
    return -4.0 * y3p * (-1.0 + nu) * (-1.0 + 2.0 * nu)
               * ps_atan2((x2 - y3p * g.c1 - y2p * g.s1) / (q3 + x3 + (-1.0) * y2p * g.c1 + y3p * g.s1), 1.0)
           - (1.0 / 2.0) * p2 * g.c2 * ps_log(r12pf(P, g, y2p, y3p))
           - 4.0 * (-1.0 + nu) * (-1.0 + 2.0 * nu) * ps_atan2(p2p * (1.0 / p3p), 1.0)
                 * (x2 * g.c1 - (q3 + x3) * g.s1)
           - p2 * ps_atan2((1.0 / p2) * p3, 1.0) * g.s2
           + ps_atan2((-1.0) * (1.0 / p2p) * p3p, 1.0) * g.s1
                 * (2.0 * y2p * (3.0 + (-4.0) * nu) * g.c1
                    + ((-3.0) * q3 + (-5.0) * x3 + 4.0 * (q3 + x3) * nu) * g.c2
                    + ((-3.0) + 4.0 * nu) * (q3 + (-1.0) * x3 + x2 * g.s2))
           + 0.25 * ps_log(r22pf(P, g, y2p, y3p))
                 * (8.0 * y2p * (-1.0 + nu) * (-1.0 + 2.0 * nu)
                    - (x3 * (19.0 + 4.0 * nu * (-9.0 + 4.0 * nu)) + q3 * (5.0 + 4.0 * nu * (-5.0 + 4.0 * nu))) * g.c1
                    + 2.0 * y2p * (-3.0 + 4.0 * nu) * g.c2
                    + (3.0 * q3 + 5.0 * x3 - 4.0 * (q3 + x3) * nu) * g.c3
                    + x2 * (-11.0 + 4.0 * (7.0 - 4.0 * nu) * nu) * g.s1 + x2 * (3.0 - 4.0 * nu) * g.s3)
           + inv_r2p2(P, g, y2p, y3p) * x3
                 * ((q3 * q3 - x2 * x2 - x3 * x3) * g.c1
                    + (-((3.0 * q3 + x3) * y2p) + x2 * y3p) * g.c2
                    + (x2 * x2 + (q3 + x3) * (q3 + x3) + 2.0 * (y2p * y2p)) * g.c3
                    - ((q3 + x3) * y2p + x2 * y3p) * g.c4 - 2.0 * q3 * x2 * g.s1
                    + (3.0 * x2 * y2p + q3 * y3p - x3 * y3p) * g.s2 - 2.0 * y2p * y3p * g.s3
                    + (-x2 * y2p + (q3 + x3) * y3p) * g.s4);
}

// I322 (:432-466)
template <class T>
PS_FN T I322(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3, nu = g.nu;
    const T p3 = p3f(P, g, y3p), p2 = p2f(P, g, y2p), p2p = p2pf(P, g, y2p), p3p = p3pf(P, g, y3p);
    return -4.0 * y2p * ((-1.0) + nu) * ((-1.0) + 2.0 * nu)
               * ps_atan2((x2 + (-1.0) * y3p * g.c1 + (-1.0) * y2p * g.s1)
                              * (1.0 / (q3 + x3 + (-1.0) * y2p * g.c1 + y3p * g.s1)), 1.0)
           - (0.5) * p3 * g.c2 * ps_log(r12pf(P, g, y2p, y3p))
           + 4.0 * ((-1.0) + nu) * ((-1.0) + 2.0 * nu) * ps_atan2((1.0 / p2p) * p3p, 1.0)
                 * ((q3 + x3) * g.c1 + x2 * g.s1)
           + ps_atan2(p2p * (1.0 / p3p), 1.0) * g.c1
                 * ((-3.0 * q3 - 5.0 * x3 + 4.0 * (q3 + x3) * nu) * g.c2
                    - (-3.0 + 4.0 * nu) * (q3 - x3 + 2.0 * (y3p - x2 * g.c1) * g.s1))
           + p3 * ps_atan2(p2 / p3, 1.0) * g.s2
           + 0.25 * ps_log(r22pf(P, g, y2p, y3p))
                 * (-8.0 * y3p * (-1.0 + nu) * (-1.0 + 2.0 * nu)
                    + x2 * (11.0 + 4.0 * nu * (-7.0 + 4.0 * nu)) * g.c1
                    + 2.0 * y3p * (-3.0 + 4.0 * nu) * g.c2 + x2 * (3.0 - 4.0 * nu) * g.c3
                    - (5.0 * q3 + 19.0 * x3 - 4.0 * (5.0 * q3 + 9.0 * x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.s1
                    + (-3.0 * q3 - 5.0 * x3 + 4.0 * (q3 + x3) * nu) * g.s3)
           + inv_r2p2(P, g, y2p, y3p) * x3
                 * (2.0 * q3 * x2 * g.c1 - (x2 * y2p + 3.0 * q3 * y3p + x3 * y3p) * g.c2
                    + 2.0 * y2p * y3p * g.c3 + (-x2 * y2p + (q3 + x3) * y3p) * g.c4
                    - (-(q3 * q3) + x2 * x2 + x3 * x3) * g.s1
                    + (-q3 * y2p + x3 * y2p + 3.0 * x2 * y3p) * g.s2
                    - (x2 * x2 + (q3 + x3) * (q3 + x3) + 2.0 * (y3p * y3p)) * g.s3
                    + ((q3 + x3) * y2p + x2 * y3p) * g.s4);
}

// I333 (:472-510)
template <class T>
PS_FN T I333(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3, nu = g.nu;
    const T p2 = p2f(P, g, y2p), p3 = p3f(P, g, y3p), p2p = p2pf(P, g, y2p), p3p = p3pf(P, g, y3p);
    return p2 * ps_atan2(p3 / p2, 1.0) * ((-3.0) + 4.0 * nu + g.c2)
           + (1.0 / 4.0) * ps_log(r12pf(P, g, y2p, y3p))
                 * (-6.0 * y3p + 8.0 * y3p * nu + x2 * (7.0 - 8.0 * nu) * g.c1 - x2 * g.c3
                    + 2.0 * (q3 - x3) * (-2.0 + 4.0 * nu + g.c2) * g.s1 - 2.0 * y2p * g.s2)
           + 0.5 * ps_atan2(-p3p / p2p, 1.0)
                 * (2.0 * y2p * (5.0 + 4.0 * nu * (-3.0 + 2.0 * nu))
                    - (7.0 * q3 + 5.0 * x3 - 20.0 * (q3 + x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.c1
                    + 2.0 * y2p * (-3.0 + 4.0 * nu) * g.c2
                    + (3.0 * q3 + x3 - 4.0 * (q3 + x3) * nu) * g.c3
                    + x2 * (-13.0 + 4.0 * (7.0 - 4.0 * nu) * nu) * g.s1 + x2 * (3.0 - 4.0 * nu) * g.s3)
           + 0.25 * ps_log(r22pf(P, g, y2p, y3p))
                 * (-2.0 * y3p * (5.0 + 4.0 * nu * (-3.0 + 2.0 * nu))
                    + x2 * (13.0 + 4.0 * nu * (-7.0 + 4.0 * nu)) * g.c1 + x2 * (-3.0 + 4.0 * nu) * g.c3
                    - (7.0 * q3 + 5.0 * x3 - 20.0 * (q3 + x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.s1
                    + 2.0 * y2p * (-3.0 + 4.0 * nu) * g.s2
                    + (3.0 * q3 + x3 - 4.0 * (q3 + x3) * nu) * g.s3)
           + inv_r2p2(P, g, y2p, y3p) * x3
                 * (-2.0 * q3 * x2 * g.c1 + (3.0 * x2 * y2p + q3 * y3p - x3 * y3p) * g.c2
                    - 2.0 * y2p * y3p * g.c3 + (-x2 * y2p + (q3 + x3) * y3p) * g.c4
                    + (-(q3 * q3) + x2 * x2 + x3 * x3) * g.s1
                    + (3.0 * q3 * y2p + x3 * y2p - x2 * y3p) * g.s2
                    - (x2 * x2 + (q3 + x3) * (q3 + x3) + 2.0 * (y2p * y2p)) * g.s3
                    + ((q3 + x3) * y2p + x2 * y3p) * g.s4);
}

// I332 (:516-556)
template <class T>
PS_FN T I332(const Point<T> &P, const Geom &g, double y2p, double y3p)
{
    const T &x2 = P.x2, &x3 = P.x3;
    const double q3 = g.q3, nu = g.nu;
    const T p3 = p3f(P, g, y3p), p2 = p2f(P, g, y2p), p2p = p2pf(P, g, y2p), p3p = p3pf(P, g, y3p);
    return -p3 * ps_atan2(p2 / p3, 1.0) * (3.0 - 4.0 * nu + g.c2)
           + (1.0 / 4.0) * ps_log(r12pf(P, g, y2p, y3p))
                 * (-6.0 * y2p + 8.0 * y2p * nu + 2.0 * (q3 - x3) * g.c1 * (2.0 - 4.0 * nu + g.c2)
                    + 2.0 * x2 * (4.0 - 4.0 * nu + g.c2) * g.s1 - 2.0 * y3p * g.s2)
           + 0.25 * ps_log(r22pf(P, g, y2p, y3p))
                 * ((-2.0) * y2p * (5.0 + 4.0 * nu * ((-3.0) + 2.0 * nu))
                    + (7.0 * q3 + 5.0 * x3 + (-20.0) * (q3 + x3) * nu + 16.0 * (q3 + x3) * (nu * nu)) * g.c1
                    + (3.0 * q3 + x3 + (-4.0) * (q3 + x3) * nu) * g.c3
                    + x2 * (13.0 + 4.0 * nu * ((-7.0) + 4.0 * nu)) * g.s1
                    + 2.0 * y3p * ((-3.0) + 4.0 * nu) * g.s2 + x2 * (3.0 + (-4.0) * nu) * g.s3)
           + 0.5 * ps_atan2(p2p / p3p, 1.0)
                 * (-2.0 * y3p * (5.0 + 4.0 * nu * (-3.0 + 2.0 * nu))
                    - x2 * (-13.0 + 4.0 * (7.0 - 4.0 * nu) * nu) * g.c1
                    - 2.0 * y3p * (3.0 - 4.0 * nu) * g.c2 - x2 * ((-3.0) + 4.0 * nu) * g.c3
                    - (7.0 * q3 + 5.0 * x3) * g.s1 - 4.0 * (q3 + x3) * nu * ((-5.0) + 4.0 * nu) * g.s1
                    - (3.0 * q3 + x3 + (-4.0) * (q3 + x3) * nu) * g.s3)
           + inv_r2p2(P, g, y2p, y3p) * x3
                 * ((-1.0) * ((-1.0) * (q3 * q3) + x2 * x2 + x3 * x3) * g.c1
                    + ((-1.0) * q3 * y2p + x3 * y2p + 3.0 * x2 * y3p) * g.c2
                    - (x2 * x2 + (q3 + x3) * (q3 + x3) + 2.0 * (y3p * y3p)) * g.c3
                    + ((q3 + x3) * y2p + x2 * y3p) * g.c4 + (-2.0) * q3 * x2 * g.s1
                    + (x2 * y2p + (3.0 * q3 + x3) * y3p) * g.s2 - 2.0 * y2p * y3p * g.s3
                    + (x2 * y2p + (-1.0) * (q3 + x3) * y3p) * g.s4);
}

// IU2, IU3 (:112-142): the displacement integrands for the strain (e22,
// e23, e33) of the zone in the unprimed frame; lam, G the moduli
struct Strain {
    double e22, e23, e33, lam, G;
};

template <class T>
PS_FN void IU(const Point<T> &P, const Geom &g, const Strain &e, double y2p, double y3p, T &u2,
              T &u3)
{
    const double pi = 3.141592653589793115997963468544185161;
    const double epskk = e.e22 + e.e33;
    const double A = e.lam * epskk + 2.0 * e.G * e.e22, B = 2.0 * e.G * e.e23,
                 C = e.lam * epskk + 2.0 * e.G * e.e33;
    const double f = 1.0 / (8.0 * pi * e.G * (1.0 - g.nu));
    const T i223 = I223(P, g, y2p, y3p), i222 = I222(P, g, y2p, y3p), i323 = I323(P, g, y2p, y3p),
            i322 = I322(P, g, y2p, y3p);
    u2 = f * (g.s1 * (A * i223 + B * (i222 + i323) + C * i322)
              + g.c1 * (A * i222 + B * (i322 - i223) - C * i323));
    const T i233 = I233(P, g, y2p, y3p), i232 = I232(P, g, y2p, y3p), i333 = I333(P, g, y2p, y3p),
            i332 = I332(P, g, y2p, y3p);
    u3 = f * (g.s1 * (A * i233 + B * (i232 + i333) + C * i332)
              + g.c1 * (A * i232 + B * (i332 - i233) - C * i333));
}

// the zone's strain (e22', e23', e33' in its own frame) in the unprimed
// frame (computeDisplacementStrainVolumePlaneStrain.f90:90-93)
PS_FN Strain rotate(const Geom &g, double e22p, double e23p, double e33p, double lam, double G)
{
    Strain s;
    s.e22 = +g.s1 * (+e22p * g.s1 + e23p * g.c1) + g.c1 * (+e23p * g.s1 + e33p * g.c1);
    s.e23 = +g.s1 * (-e22p * g.c1 + e23p * g.s1) + g.c1 * (-e23p * g.c1 + e33p * g.s1);
    s.e33 = -g.c1 * (-e22p * g.c1 + e23p * g.s1) + g.s1 * (-e23p * g.c1 + e33p * g.s1);
    s.lam = lam;
    s.G = G;
    return s;
}

// displacement (u2, u3) of a zone of thickness Tz and width W (:99-101)
template <class T>
PS_FN void displacement(const Point<T> &P, const Geom &g, const Strain &e, double Tz, double W,
                        T &u2, T &u3)
{
    T a2, a3, b2, b3, c2, c3, d2, d3;
    IU(P, g, e, Tz / 2, W, a2, a3);
    IU(P, g, e, -Tz / 2, W, b2, b3);
    IU(P, g, e, -Tz / 2, 0.0, c2, c3);
    IU(P, g, e, Tz / 2, 0.0, d2, d3);
    u2 = a2 - b2 + c2 - d2;
    u3 = a3 - b3 + c3 - d3;
}

// Omega and S of omega.f90, s.f90, heaviside.f90
PS_FN double heaviside(double x) { return 0.0 < x ? 1.0 : 0.0; }
PS_FN double omega(double x) { return heaviside(x + 0.5) - heaviside(x - 0.5); }
PS_FN double boxcar_s(double x) { return omega(x - 0.5); }

}  // namespace psj
