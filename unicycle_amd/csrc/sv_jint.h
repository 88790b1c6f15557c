// sv_jint.h -- displacement J-integrals of a vertical strain volume (cuboid)
// in a half space, Barbot, Moore & Lambert (2017, BSSA 107(2)), as used by
// jdpmoore/unicycle fortran/3d/src/strainvolume.f90:215-991
// (computeDisplacementVerticalStrainVolume and its J functions).
//
// The stress routine of the reference (strainvolume.f90:1042-4640) uses the
// hand-differentiated J functions; here the displacement gradient is the
// exact derivative of the displacement J-integrals, formed by forward-mode
// differentiation with respect to the observation point (x1', x2', x3).
// The formulas are templated on the number type T, which provides +, -, *, /
// with T and double and the functions sv_sqrt, sv_atan2, sv_atan3 (atan of
// y/x, +-pi/2 at x = 0, strainvolume.f90:91-100), sv_xlogy (x log y with
// 0 log y = 0, strainvolume.f90:107-116; its derivative keeps that rule for
// the derivative of x) and sv_acoth (macros.f90:7).  T also exposes
// .d(j), the derivative along x1', x2', x3.
//
// Every J is returned without the common factor 1/(16 pi (1-nu) G).
#pragma once

#if defined(__HIPCC__)
#define SV_FN __host__ __device__ __forceinline__
#else
#define SV_FN inline
#endif

namespace svj {

// one corner (y1, y2, y3) of the cuboid seen from (x1', x2', x3)
template <class T>
struct Corner {
    T a, b, c, p;      // x1'-y1, x2'-y2, x3-y3, x3+y3
    T R1, R2;          // distances to the source point and to its image
    T x1, x2, x3;      // observation point (x1', x2', x3)
    double y1, y2, y3;
};

template <class T>
SV_FN Corner<T> corner(const T &x1, const T &x2, const T &x3, double y1, double y2, double y3)
{
    Corner<T> q{x1 - y1, x2 - y2, x3 - y3, x3 + y3, x1, x1, x1, x2, x3, y1, y2, y3};
    q.R1 = sv_sqrt(q.a * q.a + q.b * q.b + q.c * q.c);
    q.R2 = sv_sqrt(q.a * q.a + q.b * q.b + q.p * q.p);
    return q;
}

// the pairs (a, b) / (b, a) recur with x1' and x2' exchanged; the functions
// below name them u (first) and w (second) where both forms exist

// J1112 (and J2212 with the horizontal roles exchanged); strainvolume.f90:344, 620
template <class T>
SV_FN T j_x12(const Corner<T> &q, const T &u, const T &w, double nu)
{
    const double m1 = nu - 1.0, m2 = 2.0 * nu - 1.0;
    const T &x3 = q.x3, &c = q.c, &p = q.p;
    const double y3 = q.y3;
    const T a = q.a;  // the single-angle terms are functions of x1'-y1 in both
    return 2.0 * x3 * u * w * y3 / (q.R2 * (u * u + p * p)) - 4.0 * m1 * m2 * p * sv_atan3(w, u)
           - x3 * sv_atan2(x3, a) - 3.0 * x3 * sv_atan2(3.0 * x3, a)
           + 4.0 * nu * x3 * sv_atan2(-nu * x3, a)
           + 4.0 * m1 * m2 * p * sv_atan2(-(q.R2 * u), w * p)
           - 4.0 * m1 * c * sv_atan2(q.R1 * c, u * w)
           + 3.0 * y3 * sv_atan2(T(-3.0 * y3), a) - y3 * sv_atan2(T(y3), a)
           - 4.0 * nu * y3 * sv_atan2(T(nu * y3), a)
           - 4.0 * m1 * p * sv_atan2(q.R2 * p, u * w);
}

template <class T>
SV_FN T J1112(const Corner<T> &q, double nu)
{
    const double m1 = nu - 1.0;
    return j_x12(q, q.a, q.b, nu) + sv_xlogy(-(4.0 * nu - 3.0) * q.a, q.R1 + q.b)
           + sv_xlogy((5.0 + 4.0 * nu * (2.0 * nu - 3.0)) * q.a, q.R2 + q.b)
           + sv_xlogy(-4.0 * m1 * q.b, q.R1 + q.a) + sv_xlogy(-4.0 * m1 * q.b, q.R2 + q.a);
}

template <class T>
SV_FN T J2212(const Corner<T> &q, double nu)
{
    const double m1 = nu - 1.0;
    return j_x12(q, q.b, q.a, nu) + sv_xlogy(-4.0 * m1 * q.a, q.R1 + q.b)
           + sv_xlogy(-4.0 * m1 * q.a, q.R2 + q.b) + sv_xlogy(-(4.0 * nu - 3.0) * q.b, q.R1 + q.a)
           + sv_xlogy((5.0 + 4.0 * nu * (2.0 * nu - 3.0)) * q.b, q.R2 + q.a);
}

// the bracket shared by J1113, J2113 (= J1213), J2123 (= J1223), J2223
template <class T>
SV_FN T q_bracket(const Corner<T> &q, double nu)
{
    const double k = (nu - 1.0) * (2.0 * nu - 1.0);
    const T s = q.a * q.a + q.b * q.b;
    return -k * q.R2 * q.R2 * q.p + k * q.R2 * q.y3 * (2.0 * q.x3 + q.y3) + q.x3 * (s + q.x3 * q.p);
}

// J1113 (u = a, the horizontal coordinate x = x2') and J2223 (u = b, x = x1');
// strainvolume.f90:373, 683
template <class T>
SV_FN T j_x13(const Corner<T> &q, const T &u, const T &w, const T &xw, double yw, double nu)
{
    const double m1 = nu - 1.0;
    const T s = q.a * q.a + q.b * q.b;
    return 2.0 * u * q_bracket(q, nu) / (q.R2 * s) + xw * sv_atan2(-xw, u)
           - 3.0 * xw * sv_atan2(3.0 * xw, u) + 4.0 * nu * xw * sv_atan2(-nu * xw, u)
           - 4.0 * m1 * w * sv_atan2(q.R1 * w, u * q.c) + 4.0 * m1 * w * sv_atan2(q.R2 * w, u * q.p)
           + 3.0 * yw * sv_atan2(T(-3.0 * yw), u) - yw * sv_atan2(T(yw), u)
           - 4.0 * nu * yw * sv_atan2(T(nu * yw), u)
           + sv_xlogy(-(4.0 * nu - 3.0) * u, q.R1 + q.c)
           + sv_xlogy(-(3.0 - 6.0 * nu + 4.0 * nu * nu) * u, q.R2 + q.p)
           + sv_xlogy(-4.0 * m1 * q.c, q.R1 + u) + sv_xlogy(4.0 * m1 * q.p, q.R2 + u);
}

template <class T>
SV_FN T J1113(const Corner<T> &q, double nu) { return j_x13(q, q.a, q.b, q.x2, q.y2, nu); }
template <class T>
SV_FN T J2223(const Corner<T> &q, double nu) { return j_x13(q, q.b, q.a, q.x1, q.y1, nu); }

// J1123 (u = a, w = b, horizontal x1', y1) and J2213 (u = b, w = a, x2', y2);
// strainvolume.f90:402, 649
template <class T>
SV_FN T j_x23(const Corner<T> &q, const T &u, const T &w, const T &xu, double yu, double nu)
{
    const double k = (nu - 1.0) * (2.0 * nu - 1.0);
    const T &x3 = q.x3, &p = q.p;
    const double y3 = q.y3;
    const T u2 = u * u, w2 = w * w, x32 = x3 * x3, up = u2 + p * p;
    const T poly = x3 * ((x32 + u2) * (x32 + u2 + w2) + x3 * (3.0 * x32 + 2.0 * u2 + w2) * y3
                         + 3.0 * x32 * (y3 * y3) + x3 * (y3 * y3 * y3))
                   - k * q.R2 * q.R2 * p * up + k * q.R2 * y3 * (2.0 * x3 + y3) * up;
    return -2.0 * w * poly / (q.R2 * (u2 + w2) * up) + 2.0 * k * u * sv_atan3(u, w)
           + xu * sv_atan2(-xu, w) - 3.0 * xu * sv_atan2(3.0 * xu, w)
           + 4.0 * nu * xu * sv_atan2(-nu * xu, w) + 3.0 * yu * sv_atan2(T(-3.0 * yu), w)
           - yu * sv_atan2(T(yu), w) - 4.0 * nu * yu * sv_atan2(T(nu * yu), w)
           + 2.0 * (2.0 * nu - 1.0) * u * sv_atan2(-(q.R1 * u), w * q.c)
           + 2.0 * (1.0 - 2.0 * nu) * (1.0 - 2.0 * nu) * u * sv_atan2(-(q.R2 * u), w * p)
           + sv_xlogy(-2.0 * x3, q.R2 - w) + sv_xlogy(-(4.0 * nu - 3.0) * w, q.R1 + q.c)
           + sv_xlogy(-(3.0 - 6.0 * nu + 4.0 * nu * nu) * w, q.R2 + p)
           + sv_xlogy(-(4.0 * nu - 3.0) * q.c, q.R1 + w)
           + sv_xlogy(-(5.0 + 4.0 * nu * (2.0 * nu - 3.0)) * p, q.R2 + w);
}

template <class T>
SV_FN T J1123(const Corner<T> &q, double nu) { return j_x23(q, q.a, q.b, q.x1, q.y1, nu); }
template <class T>
SV_FN T J2213(const Corner<T> &q, double nu) { return j_x23(q, q.b, q.a, q.x2, q.y2, nu); }

// J2112 = J1212; strainvolume.f90:436, 560
template <class T>
SV_FN T J2112(const Corner<T> &q, double nu)
{
    return -q.R1 + (1.0 + 8.0 * (nu - 1.0) * nu) * q.R2 - 2.0 * q.x3 * q.y3 / q.R2
           + sv_xlogy(-4.0 * (nu - 1.0) * (2.0 * nu - 1.0) * q.p, q.R2 + q.p);
}

// J2113 = J1213 (u = b) and J2123 = J1223 (u = a); strainvolume.f90:454, 475
template <class T>
SV_FN T j_q(const Corner<T> &q, const T &u, double nu)
{
    return 2.0 * u * q_bracket(q, nu) / (q.R2 * (q.a * q.a + q.b * q.b))
           + sv_xlogy(-(4.0 * nu * nu - 2.0 * nu - 1.0) * u, q.R2 + q.p) + sv_xlogy(-u, q.R1 + q.c);
}

template <class T>
SV_FN T J2113(const Corner<T> &q, double nu) { return j_q(q, q.b, nu); }
template <class T>
SV_FN T J2123(const Corner<T> &q, double nu) { return j_q(q, q.a, nu); }

// J3112 (u = a, w = b) and J3212 (u = b, w = a), sign included;
// strainvolume.f90:496, 712
template <class T>
SV_FN T j_3x12(const Corner<T> &q, const T &u, const T &w, double nu)
{
    const double k = (nu - 1.0) * (2.0 * nu - 1.0);
    const T &x3 = q.x3, &p = q.p;
    const double y3 = q.y3;
    return -(-2.0 * x3 * w * y3 * p / (q.R2 * (u * u + p * p)) + 4.0 * k * u * sv_atan3(u, w)
             + 4.0 * k * u * sv_atan2(-(q.R2 * u), w * p) + sv_xlogy(-4.0 * k * w, q.R2 + p)
             + sv_xlogy(q.c, q.R1 + w)
             + sv_xlogy(-x3 - 7.0 * y3 - 8.0 * nu * nu * p + 8.0 * nu * (x3 + 2.0 * y3), q.R2 + w));
}

template <class T>
SV_FN T J3112(const Corner<T> &q, double nu) { return j_3x12(q, q.a, q.b, nu); }
template <class T>
SV_FN T J3212(const Corner<T> &q, double nu) { return j_3x12(q, q.b, q.a, nu); }

// J3113 = J3223; strainvolume.f90:519, 757
template <class T>
SV_FN T J3113(const Corner<T> &q, double nu)
{
    const T &x3 = q.x3, &p = q.p;
    const double y3 = q.y3;
    return -(q.R1 + (-1.0 - 8.0 * (nu - 1.0) * nu) * q.R2 - 2.0 * x3 * y3 / q.R2
             + 2.0 * (4.0 * nu - 3.0) * x3 * sv_acoth(p / q.R2)
             + sv_xlogy(2.0 * (3.0 * x3 + 2.0 * y3 - 6.0 * nu * p + 4.0 * nu * nu * p), q.R2 + p));
}

// J3123 (u = a, w = b) and J3213 (u = b, w = a), sign included;
// strainvolume.f90:538, 735
template <class T>
SV_FN T j_3x23(const Corner<T> &q, const T &u, const T &w, double nu)
{
    const double k = (nu - 1.0) * (2.0 * nu - 1.0);
    const T &x3 = q.x3, &p = q.p;
    const double y3 = q.y3;
    return -(2.0 * x3 * u * w * y3 / (q.R2 * (u * u + p * p)) + 4.0 * k * p * sv_atan3(w, u)
             + 4.0 * (2.0 * nu - 1.0) * (nu * x3 + (nu - 1.0) * y3) * sv_atan2(q.R2 * u, w * p)
             + sv_xlogy(u, q.R1 + w) + sv_xlogy(-(1.0 + 8.0 * (nu - 1.0) * nu) * u, q.R2 + w));
}

template <class T>
SV_FN T J3123(const Corner<T> &q, double nu) { return j_3x23(q, q.a, q.b, nu); }
template <class T>
SV_FN T J3213(const Corner<T> &q, double nu) { return j_3x23(q, q.b, q.a, nu); }

// J1312 (u = a, w = b) and J2312 (u = b, w = a), sign included;
// strainvolume.f90:776, 840
template <class T>
SV_FN T j_x312(const Corner<T> &q, const T &u, const T &w, double nu)
{
    const double k = (nu - 1.0) * (2.0 * nu - 1.0);
    const T &x3 = q.x3, &p = q.p;
    const double y3 = q.y3;
    return -(2.0 * x3 * w * y3 * p / (q.R2 * (u * u + p * p)) - 4.0 * k * u * sv_atan3(u, w)
             + 4.0 * k * u * sv_atan2(q.R2 * u, w * p) + sv_xlogy(4.0 * k * w, q.R2 + p)
             + sv_xlogy(q.c, q.R1 + w)
             + sv_xlogy((7.0 + 8.0 * (nu - 2.0) * nu) * x3 + y3 + 8.0 * (nu - 1.0) * nu * y3, q.R2 + w));
}

template <class T>
SV_FN T J1312(const Corner<T> &q, double nu) { return j_x312(q, q.a, q.b, nu); }
template <class T>
SV_FN T J2312(const Corner<T> &q, double nu) { return j_x312(q, q.b, q.a, nu); }

// J1313 = J2323; strainvolume.f90:799, 885
template <class T>
SV_FN T J1313(const Corner<T> &q, double nu)
{
    const T &x3 = q.x3, &p = q.p;
    const double y3 = q.y3;
    return -(q.R1 + ((7.0 + 8.0 * (nu - 2.0) * nu) * q.R2 * q.R2 + 2.0 * x3 * y3) / q.R2
             + 2.0 * (4.0 * nu - 3.0) * x3 * sv_acoth(p / q.R2)
             + sv_xlogy(2.0 * (-3.0 * x3 - 2.0 * y3 + 6.0 * nu * p - 4.0 * nu * nu * p), q.R2 + p));
}

// J1323 (u = a, w = b) and J2313 (u = b, w = a), sign included;
// strainvolume.f90:818, 863
template <class T>
SV_FN T j_x323(const Corner<T> &q, const T &u, const T &w, double nu)
{
    const double k = (nu - 1.0) * (2.0 * nu - 1.0);
    const T &x3 = q.x3, &p = q.p;
    const double y3 = q.y3;
    return -(-2.0 * x3 * u * w * y3 / (q.R2 * (u * u + p * p)) - 4.0 * k * p * sv_atan3(w, u)
             - 4.0 * (nu - 1.0) * (-3.0 * x3 - y3 + 2.0 * nu * p) * sv_atan2(q.R2 * u, w * p)
             + sv_xlogy(u, q.R1 + w) + sv_xlogy((7.0 + 8.0 * (nu - 2.0) * nu) * u, q.R2 + w));
}

template <class T>
SV_FN T J1323(const Corner<T> &q, double nu) { return j_x323(q, q.a, q.b, nu); }
template <class T>
SV_FN T J2313(const Corner<T> &q, double nu) { return j_x323(q, q.b, q.a, nu); }

// J3312; strainvolume.f90:904
template <class T>
SV_FN T J3312(const Corner<T> &q, double nu)
{
    const T &x3 = q.x3, &a = q.a, &b = q.b, &c = q.c, &p = q.p;
    const double y3 = q.y3, nu2 = nu * nu, g = (1.0 - 2.0 * nu) * (1.0 - 2.0 * nu);
    const double c5 = 5.0 + 4.0 * nu * (2.0 * nu - 3.0), c3 = 4.0 * nu - 3.0;
    return 2.0 * x3 * a * b * y3 * (a * a + b * b + 2.0 * (p * p))
               / (q.R2 * (a * a + p * p) * (b * b + p * p))
           - 3.0 * x3 * sv_atan2(3.0 * x3, a) - 5.0 * x3 * sv_atan2(5.0 * x3, b)
           + 12.0 * nu * x3 * sv_atan2(-3.0 * nu * x3, b) + 4.0 * nu * x3 * sv_atan2(-nu * x3, a)
           - 8.0 * nu2 * x3 * sv_atan2(nu2 * x3, b) + 3.0 * y3 * sv_atan2(T(-3.0 * y3), a)
           - 5.0 * y3 * sv_atan2(T(5.0 * y3), b) + 12.0 * nu * y3 * sv_atan2(T(-3.0 * nu * y3), b)
           - 4.0 * nu * y3 * sv_atan2(T(nu * y3), a) - 8.0 * nu2 * y3 * sv_atan2(T(nu2 * y3), b)
           + 2.0 * (2.0 * nu - 1.0) * c * sv_atan2(-(q.R1 * c), a * b)
           + 2.0 * g * p * sv_atan2(q.R2 * p, a * b) + sv_xlogy(-c3 * a, q.R1 + b)
           + sv_xlogy(c5 * a, q.R2 + b) + sv_xlogy(-c3 * b, q.R1 + a) + sv_xlogy(c5 * b, q.R2 + a);
}

// J3313 (u = a, w = b, horizontal x2', y2) and J3323 (u = b, w = a, x1', y1);
// strainvolume.f90:935, 965
template <class T>
SV_FN T j_33x(const Corner<T> &q, const T &u, const T &w, const T &xw, double yw, double nu)
{
    const T &x3 = q.x3, &p = q.p;
    const double y3 = q.y3, nu2 = nu * nu, m1 = nu - 1.0;
    return 2.0 * x3 * u * y3 * p / (q.R2 * (w * w + p * p)) + 5.0 * xw * sv_atan2(-5.0 * xw, u)
           - 3.0 * xw * sv_atan2(3.0 * xw, u) + 4.0 * nu * xw * sv_atan2(-nu * xw, u)
           - 12.0 * nu * xw * sv_atan2(3.0 * nu * xw, u) + 8.0 * nu2 * xw * sv_atan2(-nu2 * xw, u)
           - 4.0 * m1 * w * sv_atan2(q.R1 * w, u * q.c) - 8.0 * m1 * m1 * w * sv_atan2(q.R2 * w, u * p)
           + 3.0 * yw * sv_atan2(T(-3.0 * yw), u) - 5.0 * yw * sv_atan2(T(5.0 * yw), u)
           + 12.0 * nu * yw * sv_atan2(T(-3.0 * nu * yw), u) - 4.0 * nu * yw * sv_atan2(T(nu * yw), u)
           - 8.0 * nu2 * yw * sv_atan2(T(nu2 * yw), u) + sv_xlogy(-4.0 * x3, q.R2 - u)
           + sv_xlogy(-4.0 * m1 * u, q.R1 + q.c) + sv_xlogy(-8.0 * m1 * m1 * u, q.R2 + p)
           + sv_xlogy(-(4.0 * nu - 3.0) * q.c, q.R1 + u)
           + sv_xlogy(-7.0 * x3 - 5.0 * y3 + 12.0 * nu * p - 8.0 * nu2 * p, q.R2 + u);
}

template <class T>
SV_FN T J3313(const Corner<T> &q, double nu) { return j_33x(q, q.a, q.b, q.x2, q.y2, nu); }
template <class T>
SV_FN T J3323(const Corner<T> &q, double nu) { return j_33x(q, q.b, q.a, q.x1, q.y1, nu); }

// Displacement-gradient accumulators of one source for the six unit
// eigenstrains: for displacement component i and derivative direction j,
//   g[i][k][j]  k = 0,1,2: the J multiplying (lambda tr + 2G eps_kk), kk = 11,22,33
//               k = 3,4,5: the J pair multiplying 2G eps_12, 2G eps_13, 2G eps_23
// (strainvolume.f90:296-337).
// u[i][k] accumulates the values (the displacement J sums themselves).
struct Grad {
    double g[3][6][3];
    double u[3][6];
};

// number types that carry no derivatives (value-only evaluation for the
// displacement) specialise this to false; their .d(j) is never called
template <class T>
struct has_derivative {
    static constexpr bool value = true;
};

template <class T>
SV_FN void add(Grad &G, int i, int k, double s, const T &v)
{
    if constexpr (has_derivative<T>::value) {
        G.g[i][k][0] += s * v.d(0);
        G.g[i][k][1] += s * v.d(1);
        G.g[i][k][2] += s * v.d(2);
    }
    G.u[i][k] += s * v.val();
}

// one corner with sign s (+1/-1) of the 8-corner sum (strainvolume.f90:255-260)
template <class T>
SV_FN void accumulate(Grad &G, const Corner<T> &q, double nu, double s)
{
    const T A = J2112(q, nu), B = J2113(q, nu), C = J2123(q, nu), D = J3113(q, nu),
            E = J1313(q, nu);
    // u1: diag J1123, J1213 (= B), J1312; pairs J1223 (= C) + J1113, J1323 + J1112,
    //     J1212 (= A) + J1313 (= E)
    add(G, 0, 0, s, J1123(q, nu));
    add(G, 0, 1, s, B);
    add(G, 0, 2, s, J1312(q, nu));
    add(G, 0, 3, s, C + J1113(q, nu));
    add(G, 0, 4, s, J1323(q, nu) + J1112(q, nu));
    add(G, 0, 5, s, A + E);
    // u2: diag J2123 (= C), J2213, J2312; pairs J2223 + J2113 (= B),
    //     J2323 (= E) + J2112 (= A), J2212 + J2313
    add(G, 1, 0, s, C);
    add(G, 1, 1, s, J2213(q, nu));
    add(G, 1, 2, s, J2312(q, nu));
    add(G, 1, 3, s, J2223(q, nu) + B);
    add(G, 1, 4, s, E + A);
    add(G, 1, 5, s, J2212(q, nu) + J2313(q, nu));
    // u3: diag J3123, J3213, J3312; pairs J3223 (= D) + J3113 (= D),
    //     J3323 + J3112, J3212 + J3313
    add(G, 2, 0, s, J3123(q, nu));
    add(G, 2, 1, s, J3213(q, nu));
    add(G, 2, 2, s, J3312(q, nu));
    add(G, 2, 3, s, D + D);
    add(G, 2, 4, s, J3323(q, nu) + J3112(q, nu));
    add(G, 2, 5, s, J3212(q, nu) + J3313(q, nu));
}

// The same sums restricted to displacement component I (8 or 9 of the 22
// J-integrals; every accumulator gets the same operations as in accumulate)
template <int I, class T>
SV_FN void accumulate_component(Grad &G, const Corner<T> &q, double nu, double s)
{
    if constexpr (I == 0) {
        const T A = J2112(q, nu), B = J2113(q, nu), C = J2123(q, nu), E = J1313(q, nu);
        add(G, 0, 0, s, J1123(q, nu));
        add(G, 0, 1, s, B);
        add(G, 0, 2, s, J1312(q, nu));
        add(G, 0, 3, s, C + J1113(q, nu));
        add(G, 0, 4, s, J1323(q, nu) + J1112(q, nu));
        add(G, 0, 5, s, A + E);
    } else if constexpr (I == 1) {
        const T A = J2112(q, nu), B = J2113(q, nu), C = J2123(q, nu), E = J1313(q, nu);
        add(G, 1, 0, s, C);
        add(G, 1, 1, s, J2213(q, nu));
        add(G, 1, 2, s, J2312(q, nu));
        add(G, 1, 3, s, J2223(q, nu) + B);
        add(G, 1, 4, s, E + A);
        add(G, 1, 5, s, J2212(q, nu) + J2313(q, nu));
    } else {
        const T D = J3113(q, nu);
        add(G, 2, 0, s, J3123(q, nu));
        add(G, 2, 1, s, J3213(q, nu));
        add(G, 2, 2, s, J3312(q, nu));
        add(G, 2, 3, s, D + D);
        add(G, 2, 4, s, J3323(q, nu) + J3112(q, nu));
        add(G, 2, 5, s, J3212(q, nu) + J3313(q, nu));
    }
}

// The 8 corners (y1 in {0, L}, y2 in {+-T/2}, y3 in {q3, q3+W}) with the
// signs of strainvolume.f90:255-260.  seed(v, j) makes a T with value v and
// unit derivative along direction j.
template <class T, class Seed>
SV_FN void gradients(Grad &G, double x1p, double x2p, double x3, double q3, double L, double Th,
                     double W, double nu, Seed seed)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 6; k++) {
            G.u[i][k] = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) G.g[i][k][j] = 0.0;
        }
    const T X1 = seed(x1p, 0), X2 = seed(x2p, 1), X3 = seed(x3, 2);
    const double h = Th / 2.0;
    // corner n: y1 = L for n < 4 else 0; (y2, y3, sign) cycle through
    // (h, q3+W, +), (-h, q3+W, -), (-h, q3, +), (h, q3, -), negated for y1 = 0
    for (int n = 0; n < 8; n++) {
        const int m = n & 3;
        const double y1 = n < 4 ? L : 0.0;
        const double y2 = (m == 0 || m == 3) ? h : -h;
        const double y3 = m < 2 ? q3 + W : q3;
        const double sg = ((m & 1) ? -1.0 : 1.0) * (n < 4 ? 1.0 : -1.0);
        accumulate(G, corner(X1, X2, X3, y1, y2, y3), nu, sg);
    }
}

// gradients() for displacement component I only (G.g[I] and G.u[I])
template <int I, class T, class Seed>
SV_FN void gradients_component(Grad &G, double x1p, double x2p, double x3, double q3, double L,
                               double Th, double W, double nu, Seed seed)
{
#pragma unroll
    for (int k = 0; k < 6; k++) {
        G.u[I][k] = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) G.g[I][k][j] = 0.0;
    }
    const T X1 = seed(x1p, 0), X2 = seed(x2p, 1), X3 = seed(x3, 2);
    const double h = Th / 2.0;
    for (int n = 0; n < 8; n++) {
        const int m = n & 3;
        const double y1 = n < 4 ? L : 0.0;
        const double y2 = (m == 0 || m == 3) ? h : -h;
        const double y3 = m < 2 ? q3 + W : q3;
        const double sg = ((m & 1) ? -1.0 : 1.0) * (n < 4 ? 1.0 : -1.0);
        accumulate_component<I>(G, corner(X1, X2, X3, y1, y2, y3), nu, sg);
    }
}

// displacement of the primed eigenstrain epsp from the accumulated J sums,
// rotated back from the primed frame (strainvolume.f90:262-268)
SV_FN void disp_from_sums(const Grad &Gr, const double epsp[6], double G, double nu, double ct,
                          double st, double u[3])
{
    const double lambda = G * 2.0 * nu / (1.0 - 2.0 * nu);
    const double K = 1.0 / (16.0 * 3.141592653589793115997963468544185161 * (1.0 - nu) * G);
    const double tr = epsp[0] + epsp[3] + epsp[5];
    const double cf[6] = {lambda * tr + 2.0 * G * epsp[0], lambda * tr + 2.0 * G * epsp[3],
                          lambda * tr + 2.0 * G * epsp[5], 2.0 * G * epsp[1], 2.0 * G * epsp[2],
                          2.0 * G * epsp[4]};
    double up[3];
    for (int i = 0; i < 3; i++) {
        double acc = 0.0;
        for (int k = 0; k < 6; k++) acc += cf[k] * Gr.u[i][k];
        up[i] = K * acc;
    }
    u[0] = up[0] * ct - up[1] * st;
    u[1] = up[0] * st + up[1] * ct;
    u[2] = up[2];
}

// Displacement (NED u1 u2 u3) at (x1, x2, x3) for the primed eigenstrain
// epsp (11 12 13 22 23 33); strainvolume.f90:215-268
template <class T, class Seed>
SV_FN void displacement(double x1, double x2, double x3, double q1, double q2, double q3,
                        double L, double Th, double W, double theta, const double epsp[6],
                        double G, double nu, Seed seed, double u[3])
{
    const double ct = cos(theta), st = sin(theta);
    const double x1p = (x1 - q1) * ct + (x2 - q2) * st;
    const double x2p = -(x1 - q1) * st + (x2 - q2) * ct;
    Grad Gr;
    gradients<T>(Gr, x1p, x2p, x3, q3, L, Th, W, nu, seed);
    disp_from_sums(Gr, epsp, G, nu, ct, st, u);
}

// The same for the six unit eigenstrains at once (one 8-corner sum):
// out[e] = displacement of unit eigenstrain e (buildO, greens.f90:423-448)
template <class T, class Seed>
SV_FN void displacement6(double x1, double x2, double x3, double q1, double q2, double q3,
                         double L, double Th, double W, double theta, double G, double nu,
                         Seed seed, double out[6][3])
{
    const double ct = cos(theta), st = sin(theta);
    const double x1p = (x1 - q1) * ct + (x2 - q2) * st;
    const double x2p = -(x1 - q1) * st + (x2 - q2) * ct;
    Grad Gr;
    gradients<T>(Gr, x1p, x2p, x3, q3, L, Th, W, nu, seed);
#pragma unroll
    for (int e = 0; e < 6; e++) {
        double ep[6] = {0, 0, 0, 0, 0, 0};
        ep[e] = 1.0;
        disp_from_sums(Gr, ep, G, nu, ct, st, out[e]);
    }
}

// boxcar functions of strainvolume.f90:53-84
SV_FN double sv_omega(double x) { return (x + 0.5 > 0.0 ? 1.0 : 0.0) - (x - 0.5 > 0.0 ? 1.0 : 0.0); }
SV_FN double sv_s(double x) { return sv_omega(x - 0.5); }

// Stress (NED: s11 s12 s13 s22 s23 s33) at (x1, x2, x3) for each of the six
// unit eigenstrains (11 12 13 22 23 33, primed frame) of the volume with
// upper corner (q1, q2, q3), length L, thickness T, width W, strike theta;
// strainvolume.f90:1042-1138.  The parameter checks (STOP) are the caller's.
// SPLIT: the displacement gradient one component at a time (three 8-corner
// passes, 26 J evaluations instead of 22, a third of the accumulators live;
// the same value for every accumulator)
template <class T, class Seed, bool SPLIT = false>
SV_FN void stress6(double x1, double x2, double x3, double q1, double q2, double q3, double L,
                   double Th, double W, double theta, double G, double nu, Seed seed,
                   double out[6][6])
{
    const double lambda = G * 2.0 * nu / (1.0 - 2.0 * nu);
    const double ct = cos(theta), st = sin(theta);
    const double x1p = (x1 - q1) * ct + (x2 - q2) * st;
    const double x2p = -(x1 - q1) * st + (x2 - q2) * ct;
    Grad Gr;
    if constexpr (SPLIT) {
        gradients_component<0, T>(Gr, x1p, x2p, x3, q3, L, Th, W, nu, seed);
        gradients_component<1, T>(Gr, x1p, x2p, x3, q3, L, Th, W, nu, seed);
        gradients_component<2, T>(Gr, x1p, x2p, x3, q3, L, Th, W, nu, seed);
    } else {
        gradients<T>(Gr, x1p, x2p, x3, q3, L, Th, W, nu, seed);
    }
    const double K = 1.0 / (16.0 * 3.141592653589793115997963468544185161 * (1.0 - nu) * G);
    const double inside = sv_s(x1p / L) * sv_omega(x2p / Th) * sv_s((x3 - q3) / W);
#pragma unroll
    for (int e = 0; e < 6; e++) {
        // unit eigenstrain e in the primed frame
        double ep[6] = {0, 0, 0, 0, 0, 0};
        ep[e] = 1.0;
        const double tr = ep[0] + ep[3] + ep[5];
        const double cf[6] = {lambda * tr + 2.0 * G * ep[0], lambda * tr + 2.0 * G * ep[3],
                              lambda * tr + 2.0 * G * ep[5], 2.0 * G * ep[1], 2.0 * G * ep[2],
                              2.0 * G * ep[4]};
        double du[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < 6; k++)
                    if (cf[k] != 0.0) acc += cf[k] * Gr.g[i][k][j];
                du[i][j] = K * acc;
            }
        const double e11p = du[0][0], e22p = du[1][1], e33 = du[2][2];
        const double e12p = (du[0][1] + du[1][0]) / 2.0, e13p = (du[0][2] + du[2][0]) / 2.0,
                     e23p = (du[1][2] + du[2][1]) / 2.0;
        double e11 = (ct * e11p - st * e12p) * ct - (ct * e12p - st * e22p) * st;
        double e12 = (ct * e11p - st * e12p) * st + (ct * e12p - st * e22p) * ct;
        double e13 = ct * e13p - st * e23p;
        double e22 = (st * e11p + ct * e12p) * st + (st * e12p + ct * e22p) * ct;
        double e23 = st * e13p + ct * e23p;
        double e3 = e33;
        if (inside != 0.0) {
            const double q11 = (ct * ep[0] - st * ep[1]) * ct - (ct * ep[1] - st * ep[3]) * st;
            const double q12 = (ct * ep[0] - st * ep[1]) * st + (ct * ep[1] - st * ep[3]) * ct;
            const double q13 = ct * ep[2] - st * ep[4];
            const double q22 = (st * ep[0] + ct * ep[1]) * st + (st * ep[1] + ct * ep[3]) * ct;
            const double q23 = st * ep[2] + ct * ep[4];
            e11 -= q11 * inside;
            e12 -= q12 * inside;
            e13 -= q13 * inside;
            e22 -= q22 * inside;
            e23 -= q23 * inside;
            e3 -= ep[5] * inside;
        }
        const double ekk = e11 + e22 + e3;
        out[e][0] = lambda * ekk + 2.0 * G * e11;
        out[e][1] = 2.0 * G * e12;
        out[e][2] = 2.0 * G * e13;
        out[e][3] = lambda * ekk + 2.0 * G * e22;
        out[e][4] = 2.0 * G * e23;
        out[e][5] = lambda * ekk + 2.0 * G * e3;
    }
}

}  // namespace svj
