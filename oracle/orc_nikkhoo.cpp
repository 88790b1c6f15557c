/*
 * orc_nikkhoo.cpp -- oracle restatement of the triangular-dislocation
 * (Nikkhoo & Walter 2015) stress kernels of jdpmoore/unicycle.
 *
 * TEST INFRASTRUCTURE ONLY (see ucy_oracle.h).  Build with -ffp-contract=off.
 *
 * Follows fortran/3d/src/StrainInHalfSpace.c:42-173 (straininhalfspace_:
 * main + image + free-surface-harmonic parts) and nikkhoo15.f90:50-304.
 * The strain of each part is formed as the exact gradient of the same
 * part's displacement field (DispInHalfSpace.c:153-616: TDdispFS,
 * TDdisp_HarFunc, TDSetupD, AngSetupFSC, AngDisDisp, AngDisDispFSC), by
 * forward-mode differentiation with respect to the observation point,
 * instead of the reference's hand-differentiated strain expressions; both
 * are the same mathematical field.  Reference quirks kept:
 *   - triangles with all vertices at z = 0: the image part's yz and zz
 *     strain entries change sign (StrainInHalfSpace.c:162-165; the C code
 *     stores (xx,xy,xz,yy,yz,zz), so indices 4,5 are yz,zz);
 *   - NaN on the triangle sides (trimode 0, StrainInHalfSpace.c:307-315);
 *   - horizontal-triangle strike fallback (StrainInHalfSpace.c:222-227);
 *   - free-surface correction skipped for vertical sides (beta ~ 0 or pi,
 *     StrainInHalfSpace.c:615-617).
 */
#include <cmath>
#include <cstring>
#include "ucy_oracle.h"

namespace {

/* value and gradient with respect to the observation point (x, y, z) */
struct D3 {
    double v, g[3];
    D3() : v(0), g{0, 0, 0} {}
    D3(double c) : v(c), g{0, 0, 0} {}
};
inline D3 mk(double v, double a, double b, double c)
{
    D3 r;
    r.v = v;
    r.g[0] = a;
    r.g[1] = b;
    r.g[2] = c;
    return r;
}
inline D3 operator+(const D3 &a, const D3 &b) { return mk(a.v + b.v, a.g[0] + b.g[0], a.g[1] + b.g[1], a.g[2] + b.g[2]); }
inline D3 operator-(const D3 &a, const D3 &b) { return mk(a.v - b.v, a.g[0] - b.g[0], a.g[1] - b.g[1], a.g[2] - b.g[2]); }
inline D3 operator-(const D3 &a) { return mk(-a.v, -a.g[0], -a.g[1], -a.g[2]); }
inline D3 operator*(const D3 &a, const D3 &b)
{
    return mk(a.v * b.v, a.g[0] * b.v + a.v * b.g[0], a.g[1] * b.v + a.v * b.g[1], a.g[2] * b.v + a.v * b.g[2]);
}
inline D3 operator/(const D3 &a, const D3 &b)
{
    const double q = a.v / b.v;
    return mk(q, (a.g[0] - q * b.g[0]) / b.v, (a.g[1] - q * b.g[1]) / b.v, (a.g[2] - q * b.g[2]) / b.v);
}
inline D3 operator+(const D3 &a, double b) { return mk(a.v + b, a.g[0], a.g[1], a.g[2]); }
inline D3 operator+(double b, const D3 &a) { return a + b; }
inline D3 operator-(const D3 &a, double b) { return mk(a.v - b, a.g[0], a.g[1], a.g[2]); }
inline D3 operator-(double b, const D3 &a) { return mk(b - a.v, -a.g[0], -a.g[1], -a.g[2]); }
inline D3 operator*(const D3 &a, double b) { return mk(a.v * b, a.g[0] * b, a.g[1] * b, a.g[2] * b); }
inline D3 operator*(double b, const D3 &a) { return a * b; }
inline D3 operator/(const D3 &a, double b) { return mk(a.v / b, a.g[0] / b, a.g[1] / b, a.g[2] / b); }
inline D3 operator/(double b, const D3 &a)
{
    const double q = b / a.v;
    const double s = -q / a.v;
    return mk(q, s * a.g[0], s * a.g[1], s * a.g[2]);
}
inline D3 dsqrt(const D3 &a)
{
    const double s = std::sqrt(a.v), d = 0.5 / s;
    return mk(s, d * a.g[0], d * a.g[1], d * a.g[2]);
}
inline D3 dlog(const D3 &a) { return mk(std::log(a.v), a.g[0] / a.v, a.g[1] / a.v, a.g[2] / a.v); }
inline D3 datan(const D3 &a)
{
    const double d = 1.0 / (1.0 + a.v * a.v);
    return mk(std::atan(a.v), d * a.g[0], d * a.g[1], d * a.g[2]);
}
inline D3 datan2(const D3 &y, const D3 &x)
{
    const double r2 = x.v * x.v + y.v * y.v;
    return mk(std::atan2(y.v, x.v), (x.v * y.g[0] - y.v * x.g[0]) / r2,
              (x.v * y.g[1] - y.v * x.g[1]) / r2, (x.v * y.g[2] - y.v * x.g[2]) / r2);
}
inline D3 dnan()
{
    return mk(NAN, NAN, NAN, NAN);
}

const double kPi = 3.14159265358979323846;

/* unit normal, strike and dip of a TD (DispInHalfSpace.c:186-214); the
 * image flip of the horizontal-TD strike applies in the full-space routine */
void td_frame(const double P1[3], const double P2[3], const double P3[3], bool image_flip,
              double nv[3], double sv[3], double dv[3])
{
    const double a[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    const double b[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
    nv[0] = a[1] * b[2] - a[2] * b[1];
    nv[1] = a[2] * b[0] - a[0] * b[2];
    nv[2] = a[0] * b[1] - a[1] * b[0];
    double t = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    for (int k = 0; k < 3; k++) nv[k] = nv[k] / t;
    /* strike = eZ x n */
    sv[0] = -nv[1];
    sv[1] = nv[0];
    sv[2] = 0.0 * nv[1] - 0.0 * nv[0];
    sv[0] = 0.0 * nv[2] - 1.0 * nv[1];
    sv[1] = 1.0 * nv[0] - 0.0 * nv[2];
    t = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
    if (t == 0.0) {
        sv[0] = 0.0 * nv[2];
        sv[1] = 1.0 * nv[2];
        sv[2] = 0.0 * nv[2];
        if (image_flip && P1[2] > 0.0)
            for (int k = 0; k < 3; k++) sv[k] = -sv[k];
    }
    t = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
    for (int k = 0; k < 3; k++) sv[k] = sv[k] / t;
    dv[0] = nv[1] * sv[2] - nv[2] * sv[1];
    dv[1] = nv[2] * sv[0] - nv[0] * sv[2];
    dv[2] = nv[0] * sv[1] - nv[1] * sv[0];
    t = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    for (int k = 0; k < 3; k++) dv[k] = dv[k] / t;
}

/* barycentric configuration choice, DispInHalfSpace.c:389-417 (values only) */
int tri_mode(double x, double y, double z, double p1a, double p1b, double p2a, double p2b,
             double p3a, double p3b)
{
    const double den = (p2b - p3b) * (p1a - p3a) + (p3a - p2a) * (p1b - p3b);
    const double a = ((p2b - p3b) * (x - p3a) + (p3a - p2a) * (y - p3b)) / den;
    const double b = ((p3b - p1b) * (x - p3a) + (p1a - p3a) * (y - p3b)) / den;
    const double c = 1.0 - a - b;
    int m = 1;
    if (a <= 0.0 && b > c && c > a) m = -1;
    else if (b <= 0.0 && c > a && a > b) m = -1;
    else if (c <= 0.0 && a > b && b > c) m = -1;
    else if (a == 0.0 && b >= 0.0 && c >= 0.0) m = 0;
    else if (a >= 0.0 && b == 0.0 && c >= 0.0) m = 0;
    else if (a >= 0.0 && b >= 0.0 && c == 0.0) m = 0;
    if (m == 0 && z != 0.0) m = 1;
    return m;
}

/* angular dislocation in a full space, without the Burgers-function term
 * (Comninou & Dundurs 1975; DispInHalfSpace.c:541-571) */
void angdis_disp(D3 x, D3 y, D3 z, double alpha, double bx, double by, double bz, double nu,
                 D3 u[3])
{
    const double ca = std::cos(alpha), sa = std::sin(alpha);
    const D3 eta = y * ca - z * sa;
    D3 zeta = y * sa + z * ca;
    const D3 r = dsqrt(x * x + y * y + z * z);
    if (zeta.v > r.v) zeta = r;
    if (z.v > r.v) z = r;
    const double k = 1.0 / (8.0 * kPi * (1.0 - nu)), n12 = 1.0 - 2.0 * nu;
    const D3 rz = r - z, rzt = r - zeta;
    const D3 Lz = dlog(rz), Lzt = dlog(rzt);
    const D3 ir_rz = 1.0 / (r * rz), ir_rzt = 1.0 / (r * rzt), i_rzt = 1.0 / rzt;
    const D3 ux = bx * k * (x * y * ir_rz - x * eta * ir_rzt);
    const D3 vx = bx * k * (eta * sa * i_rzt - y * eta * ir_rzt + y * y * ir_rz + n12 * (ca * Lzt - Lz));
    const D3 wx = bx * k * (eta * ca * i_rzt - y / r - eta * z * ir_rzt - n12 * sa * Lzt);
    const D3 uy = by * k * (x * x * ca * ir_rzt - x * x * ir_rz - n12 * (ca * Lzt - Lz));
    const D3 vy = by * k * x * (y * ca * ir_rzt - sa * ca * i_rzt - y * ir_rz);
    const D3 wy = by * k * x * (z * ca * ir_rzt - ca * ca * i_rzt + 1.0 / r);
    const D3 uz = bz * k * sa * (n12 * Lzt - x * x * ir_rzt);
    const D3 vz = bz * k * sa * x * (sa * i_rzt - y * ir_rzt);
    const D3 wz = bz * k * sa * x * (ca * i_rzt - z * ir_rzt);
    u[0] = ux + uy + uz;
    u[1] = vx + vy + vz;
    u[2] = wx + wy + wz;
}

/* DispInHalfSpace.c:419-460 */
void td_setup(D3 x, D3 y, D3 z, double alpha, double bx, double by, double bz, double nu,
              const double vert[3], const double side[3], D3 out[3])
{
    const double a00 = side[2], a01 = -side[1], a10 = side[1], a11 = side[2];
    const D3 ty = y - vert[1], tz = z - vert[2];
    const D3 y1 = a00 * ty + a01 * tz, z1 = a10 * ty + a11 * tz;
    const double by1 = a00 * by + a01 * bz, bz1 = a10 * by + a11 * bz;
    D3 d[3];
    angdis_disp(x, y1, z1, -kPi + alpha, bx, by1, bz1, nu, d);
    out[0] = d[0];
    out[1] = a00 * d[1] + a10 * d[2];
    out[2] = a01 * d[1] + a11 * d[2];
}

/* full-space TD displacement, DispInHalfSpace.c:153-300 */
void td_disp_fs(D3 X, D3 Y, D3 Z, const double P1[3], const double P2[3], const double P3[3],
                double by, double bz, double bx, double nu, D3 u[3])
{
    double nv[3], sv[3], dv[3];
    td_frame(P1, P2, P3, true, nv, sv, dv);
    const double A[3][3] = {{nv[0], nv[1], nv[2]}, {sv[0], sv[1], sv[2]}, {dv[0], dv[1], dv[2]}};
    const D3 dX = X - P2[0], dY = Y - P2[1], dZ = Z - P2[2];
    const D3 x = A[0][0] * dX + A[0][1] * dY + A[0][2] * dZ;
    const D3 y = A[1][0] * dX + A[1][1] * dY + A[1][2] * dZ;
    const D3 z = A[2][0] * dX + A[2][1] * dY + A[2][2] * dZ;
    double p1[3], p3[3];
    const double p2[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++) {
        p1[i] = A[i][0] * (P1[0] - P2[0]) + A[i][1] * (P1[1] - P2[1]) + A[i][2] * (P1[2] - P2[2]);
        p3[i] = A[i][0] * (P3[0] - P2[0]) + A[i][1] * (P3[1] - P2[1]) + A[i][2] * (P3[2] - P2[2]);
    }
    double e12[3], e13[3], e23[3];
    const double n12 = std::sqrt((p2[0] - p1[0]) * (p2[0] - p1[0]) + (p2[1] - p1[1]) * (p2[1] - p1[1]) + (p2[2] - p1[2]) * (p2[2] - p1[2]));
    const double n13 = std::sqrt((p3[0] - p1[0]) * (p3[0] - p1[0]) + (p3[1] - p1[1]) * (p3[1] - p1[1]) + (p3[2] - p1[2]) * (p3[2] - p1[2]));
    const double n23 = std::sqrt((p3[0] - p2[0]) * (p3[0] - p2[0]) + (p3[1] - p2[1]) * (p3[1] - p2[1]) + (p3[2] - p2[2]) * (p3[2] - p2[2]));
    for (int i = 0; i < 3; i++) {
        e12[i] = (p2[i] - p1[i]) / n12;
        e13[i] = (p3[i] - p1[i]) / n13;
        e23[i] = (p3[i] - p2[i]) / n23;
    }
    const double Aang = std::acos(e12[0] * e13[0] + e12[1] * e13[1] + e12[2] * e13[2]);
    const double Bang = std::acos(-e12[0] * e23[0] - e12[1] * e23[1] - e12[2] * e23[2]);
    const double Cang = std::acos(e23[0] * e13[0] + e23[1] * e13[1] + e23[2] * e13[2]);
    const int mode = tri_mode(y.v, z.v, x.v, p1[1], p1[2], p2[1], p2[2], p3[1], p3[2]);
    D3 d[3];
    if (mode == 0) {
        d[0] = d[1] = d[2] = dnan();
    } else {
        D3 d1[3], d2[3], d3[3];
        double m13[3], m12[3], m23[3];
        for (int i = 0; i < 3; i++) { m13[i] = -e13[i]; m12[i] = -e12[i]; m23[i] = -e23[i]; }
        if (mode == 1) {
            td_setup(x, y, z, Aang, bx, by, bz, nu, p1, m13, d1);
            td_setup(x, y, z, Bang, bx, by, bz, nu, p2, e12, d2);
            td_setup(x, y, z, Cang, bx, by, bz, nu, p3, e23, d3);
        } else {
            td_setup(x, y, z, Aang, bx, by, bz, nu, p1, e13, d1);
            td_setup(x, y, z, Bang, bx, by, bz, nu, p2, m12, d2);
            td_setup(x, y, z, Cang, bx, by, bz, nu, p3, m23, d3);
        }
        for (int i = 0; i < 3; i++) d[i] = d1[i] + d2[i] + d3[i];
    }
    /* Burgers function: solid angle of the TD seen from the point */
    const D3 a0 = -x, a1 = p1[1] - y, a2 = p1[2] - z;
    const D3 b0 = -x, b1 = -y, b2 = -z;
    const D3 c0 = -x, c1 = p3[1] - y, c2 = p3[2] - z;
    const D3 na = dsqrt(a0 * a0 + a1 * a1 + a2 * a2), nb = dsqrt(b0 * b0 + b1 * b1 + b2 * b2),
             nc = dsqrt(c0 * c0 + c1 * c1 + c2 * c2);
    const D3 trip = a0 * (b1 * c2 - b2 * c1) - a1 * (b0 * c2 - b2 * c0) + a2 * (b0 * c1 - b1 * c0);
    const D3 den = na * nb * nc + (a0 * b0 + a1 * b1 + a2 * b2) * nc + (a0 * c0 + a1 * c1 + a2 * c2) * nb
                   + (b0 * c0 + b1 * c1 + b2 * c2) * na;
    const D3 Fi = -2.0 * datan2(trip, den) / (4.0 * kPi);
    d[0] = d[0] + bx * Fi;
    d[1] = d[1] + by * Fi;
    d[2] = d[2] + bz * Fi;
    for (int i = 0; i < 3; i++) u[i] = A[0][i] * d[0] + A[1][i] * d[1] + A[2][i] * d[2];
}

/* harmonic (free-surface) correction of one angular dislocation in a
 * half-space, DispInHalfSpace.c:577-606 (Nikkhoo & Walter 2015) */
void angdis_disp_fsc(D3 y1, D3 y2, D3 y3, double beta, double b1, double b2, double b3,
                     double nu, double a, D3 v[3])
{
    const double sb = std::sin(beta), cb = std::cos(beta), cot = cb / sb;
    const double k = 1.0 / (4.0 * kPi * (1.0 - nu)), N1 = 1.0 - 2.0 * nu, nu1 = 1.0 - nu;
    const D3 y3b = y3 + 2.0 * a;
    const D3 z1b = y1 * cb + y3b * sb, z3b = -y1 * sb + y3b * cb;
    const D3 rb = dsqrt(y1 * y1 + y2 * y2 + y3b * y3b);
    const D3 Fib = 2.0 * datan(-y2 / (-(rb + y3b) * (std::cos(beta / 2.0) / std::sin(beta / 2.0)) + y1));
    const D3 W6 = rb + y3b, W7 = rb + z3b, ya = y3b - a;
    const D3 ar = a / rb, ar2 = a / (rb * rb), ar3 = a / (rb * rb * rb);
    const D3 L6 = dlog(W6), L7 = dlog(W7);
    const D3 q = rb * cb + y3b;          /* rb cos(beta) + y3b */
    const D3 s1 = rb * sb - y1;          /* rb sin(beta) - y1 */
    const double c2 = cot * cot;
    /* strike-parallel Burgers component b1 */
    const D3 v1a = -2.0 * nu1 * N1 * Fib * c2 + N1 * y2 / W6 * ((N1 - ar) * cot - y1 / W6 * (nu + ar))
                   + N1 * y2 * cb * cot / W7 * (cb + ar) + a * y2 * ya * cot / (rb * rb * rb)
                   + y2 * ya / (rb * W6) * (-N1 * cot + y1 / W6 * (2.0 * nu + ar) + a * y1 / (rb * rb))
                   + y2 * ya / (rb * W7) * (cb / W7 * (q * (N1 * cb - ar) * cot + 2.0 * nu1 * s1 * cb)
                                            - a * y3b * cb * cot / (rb * rb));
    const D3 v2a = N1 * ((2.0 * nu1 * c2 - nu) * L6 - (2.0 * nu1 * c2 + N1) * cb * L7)
                   - N1 / W6 * (y1 * cot * (N1 - ar) + nu * y3b - a + y2 * y2 / W6 * (nu + ar))
                   - N1 * z1b * cot / W7 * (cb + ar) - a * y1 * ya * cot / (rb * rb * rb)
                   + ya / W6 * (-2.0 * nu + 1.0 / rb * (N1 * y1 * cot - a) + y2 * y2 / (rb * W6) * (2.0 * nu + ar)
                                + a * y2 * y2 / (rb * rb * rb))
                   + ya / W7 * (cb * cb - 1.0 / rb * (N1 * z1b * cot + a * cb) + a * y3b * z1b * cot / (rb * rb * rb)
                                - 1.0 / (rb * W7) * (y2 * y2 * cb * cb - a * z1b * cot / rb * q));
    const D3 v3a = 2.0 * nu1 * (N1 * Fib * cot + y2 / W6 * (2.0 * nu + ar) - y2 * cb / W7 * (cb + ar))
                   + y2 * ya / rb * (2.0 * nu / W6 + ar2)
                   + y2 * ya * cb / (rb * W7) * (N1 - q / W7 * (cb + ar) - a * y3b / (rb * rb));
    /* dip-parallel Burgers component b2 */
    const D3 v1b = N1 * ((2.0 * nu1 * c2 + nu) * L6 - (2.0 * nu1 * c2 + 1.0) * cb * L7)
                   + N1 / W6 * (-N1 * y1 * cot + nu * y3b - a + a * y1 * cot / rb + y1 * y1 / W6 * (nu + ar))
                   - N1 * cot / W7 * (z1b * cb - a * s1 / (rb * cb)) - a * y1 * ya * cot / (rb * rb * rb)
                   + ya / W6 * (2.0 * nu + 1.0 / rb * (N1 * y1 * cot + a) - y1 * y1 / (rb * W6) * (2.0 * nu + ar)
                                - a * y1 * y1 / (rb * rb * rb))
                   + ya * cot / W7 * (-cb * sb + a * y1 * y3b / (rb * rb * rb * cb)
                                      + s1 / rb * (2.0 * nu1 * cb - q / W7 * (1.0 + a / (rb * cb))));
    const D3 v2b = 2.0 * nu1 * N1 * Fib * c2 + N1 * y2 / W6 * (-(N1 - ar) * cot + y1 / W6 * (nu + ar))
                   - N1 * y2 * cot / W7 * (1.0 + a / (rb * cb)) - a * y2 * ya * cot / (rb * rb * rb)
                   + y2 * ya / (rb * W6) * (N1 * cot - 2.0 * nu * y1 / W6 - a * y1 / rb * (1.0 / rb + 1.0 / W6))
                   + y2 * ya * cot / (rb * W7) * (-2.0 * nu1 * cb + q / W7 * (1.0 + a / (rb * cb))
                                                  + a * y3b / (rb * rb * cb));
    const D3 v3b = -2.0 * nu1 * N1 * cot * (L6 - cb * L7) - 2.0 * nu1 * y1 / W6 * (2.0 * nu + ar)
                   + 2.0 * nu1 * z1b / W7 * (cb + ar) + ya / rb * (N1 * cot - 2.0 * nu * y1 / W6 - a * y1 / (rb * rb))
                   - ya / W7 * (cb * sb + q * cot / rb * (2.0 * nu1 * cb - q / W7)
                                + a / rb * (sb - y3b * z1b / (rb * rb) - z1b * q / (rb * W7)));
    /* normal Burgers component b3 */
    const D3 v1c = N1 * (y2 / W6 * (1.0 + ar) - y2 * cb / W7 * (cb + ar)) - y2 * ya / rb * (ar2 + 1.0 / W6)
                   + y2 * ya * cb / (rb * W7) * (q / W7 * (cb + ar) + a * y3b / (rb * rb));
    const D3 v2c = N1 * (-sb * L7 - y1 / W6 * (1.0 + ar) + z1b / W7 * (cb + ar)) + y1 * ya / rb * (ar2 + 1.0 / W6)
                   - ya / W7 * (sb * (cb - ar) + z1b / rb * (1.0 + a * y3b / (rb * rb))
                                - 1.0 / (rb * W7) * (y2 * y2 * cb * sb - a * z1b / rb * q));
    const D3 v3c = 2.0 * nu1 * Fib + 2.0 * nu1 * (y2 * sb / W7 * (cb + ar))
                   + y2 * ya * sb / (rb * W7) * (1.0 + q / W7 * (cb + ar) + a * y3b / (rb * rb));
    (void)ar3;
    v[0] = b1 * k * v1a + b2 * k * v1b + b3 * k * v1c;
    v[1] = b1 * k * v2a + b2 * k * v2b + b3 * k * v2c;
    v[2] = b1 * k * v3a + b2 * k * v3b + b3 * k * v3c;
}

/* DispInHalfSpace.c:462-539 */
void ang_setup_fsc(D3 X, D3 Y, D3 Z, double bX, double bY, double bZ, const double PA[3],
                   const double PB[3], double nu, D3 out[3])
{
    const double side[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
    const double len = std::sqrt(side[0] * side[0] + side[1] * side[1] + side[2] * side[2]);
    const double beta = std::acos(-side[2] / len);
    const double eps = 2.220446049250313e-16;
    if (std::fabs(beta) < eps || std::fabs(kPi - beta) < eps) {
        out[0] = out[1] = out[2] = D3(0.0);
        return;
    }
    double ey1[3] = {side[0], side[1], 0.0};
    const double n1 = std::sqrt(ey1[0] * ey1[0] + ey1[1] * ey1[1]);
    ey1[0] /= n1;
    ey1[1] /= n1;
    const double ey3[3] = {0.0, 0.0, -1.0};
    const double ey2[3] = {ey3[1] * ey1[2] - ey3[2] * ey1[1], ey3[2] * ey1[0] - ey3[0] * ey1[2],
                           ey3[0] * ey1[1] - ey3[1] * ey1[0]};
    /* columns ey1, ey2, ey3 */
    const double A[3][3] = {{ey1[0], ey2[0], ey3[0]}, {ey1[1], ey2[1], ey3[1]}, {ey1[2], ey2[2], ey3[2]}};
    const D3 dx = X - PA[0], dy = Y - PA[1], dz = Z - PA[2];
    D3 yA[3], yB[3];
    for (int i = 0; i < 3; i++) yA[i] = A[i][0] * dx + A[i][1] * dy + A[i][2] * dz;
    double sA[3], bA[3];
    for (int i = 0; i < 3; i++) {
        sA[i] = A[i][0] * side[0] + A[i][1] * side[1] + A[i][2] * side[2];
        bA[i] = A[i][0] * bX + A[i][1] * bY + A[i][2] * bZ;
    }
    for (int i = 0; i < 3; i++) yB[i] = yA[i] - sA[i];
    const bool conf1 = beta * yA[0].v >= 0.0;
    const double bet = conf1 ? -kPi + beta : beta;
    D3 vA[3], vB[3];
    angdis_disp_fsc(yA[0], yA[1], yA[2], bet, bA[0], bA[1], bA[2], nu, -PA[2], vA);
    angdis_disp_fsc(yB[0], yB[1], yB[2], bet, bA[0], bA[1], bA[2], nu, -PB[2], vB);
    const D3 d0 = vB[0] - vA[0], d1 = vB[1] - vA[1], d2 = vB[2] - vA[2];
    /* rows ey1, ey2, ey3 */
    const double At[3][3] = {{ey1[0], ey1[1], ey1[2]}, {ey2[0], ey2[1], ey2[2]}, {ey3[0], ey3[1], ey3[2]}};
    for (int i = 0; i < 3; i++) out[i] = At[i][0] * d0 + At[i][1] * d1 + At[i][2] * d2;
}

/* harmonic function part, DispInHalfSpace.c:306-368 */
void td_disp_harfunc(D3 X, D3 Y, D3 Z, const double P1[3], const double P2[3], const double P3[3],
                     double by, double bz, double bx, double nu, D3 u[3])
{
    double nv[3], sv[3], dv[3];
    td_frame(P1, P2, P3, false, nv, sv, dv);
    /* slip from TDCS to EFCS: columns normal, strike, dip */
    const double b[3] = {nv[0] * bx + sv[0] * by + dv[0] * bz, nv[1] * bx + sv[1] * by + dv[1] * bz,
                         nv[2] * bx + sv[2] * by + dv[2] * bz};
    D3 d1[3], d2[3], d3[3];
    ang_setup_fsc(X, Y, Z, b[0], b[1], b[2], P1, P2, nu, d1);
    ang_setup_fsc(X, Y, Z, b[0], b[1], b[2], P2, P3, nu, d2);
    ang_setup_fsc(X, Y, Z, b[0], b[1], b[2], P3, P1, nu, d3);
    for (int i = 0; i < 3; i++) u[i] = d1[i] + d2[i] + d3[i];
}

/* symmetric gradient in the reference's order (xx, xy, xz, yy, yz, zz) */
void strain_of(const D3 u[3], double e[6])
{
    e[0] = u[0].g[0];
    e[1] = (u[0].g[1] + u[1].g[0]) / 2.0;
    e[2] = (u[0].g[2] + u[2].g[0]) / 2.0;
    e[3] = u[1].g[1];
    e[4] = (u[1].g[2] + u[2].g[1]) / 2.0;
    e[5] = u[2].g[2];
}

void hooke(const double e[6], double mu, double lambda, double s[6])
{
    const double tr = e[0] + e[3] + e[5];
    s[0] = 2.0 * mu * e[0] + lambda * tr;
    s[3] = 2.0 * mu * e[3] + lambda * tr;
    s[5] = 2.0 * mu * e[5] + lambda * tr;
    s[1] = 2.0 * mu * e[1];
    s[2] = 2.0 * mu * e[2];
    s[4] = 2.0 * mu * e[4];
}

}  // namespace

/* straininhalfspace_ (StrainInHalfSpace.c:42-173); ENU coordinates */
extern "C" void orc_strain_in_halfspace(double stress[6], double strain[6], double X, double Y,
                                        double Z, const double P1[3], const double P2[3],
                                        const double P3[3], double Ss, double Ds, double Ts,
                                        double mu, double lambda)
{
    const double nu = lambda / (lambda + mu) / 2.0;
    const D3 x = mk(X, 1, 0, 0), y = mk(Y, 0, 1, 0), z = mk(Z, 0, 0, 1);
    const double M1[3] = {P1[0], P1[1], -P1[2]}, M2[3] = {P2[0], P2[1], -P2[2]},
                 M3[3] = {P3[0], P3[1], -P3[2]};
    D3 ums[3], ufs[3], uis[3];
    td_disp_fs(x, y, z, P1, P2, P3, Ss, Ds, Ts, nu, ums);
    td_disp_harfunc(x, y, z, P1, P2, P3, Ss, Ds, Ts, nu, ufs);
    td_disp_fs(x, y, z, M1, M2, M3, Ss, Ds, Ts, nu, uis);
    double ems[6], efs[6], eis[6], sms[6], sfs[6], sis[6];
    strain_of(ums, ems);
    strain_of(ufs, efs);
    strain_of(uis, eis);
    hooke(ems, mu, lambda, sms);
    hooke(efs, mu, lambda, sfs);
    hooke(eis, mu, lambda, sis);
    if (M1[2] == 0.0 && M2[2] == 0.0 && M3[2] == 0.0) {
        sis[4] = -sis[4];
        sis[5] = -sis[5];
        eis[4] = -eis[4];
        eis[5] = -eis[5];
    }
    for (int i = 0; i < 6; i++) {
        stress[i] = sms[i] + sis[i] + sfs[i];
        strain[i] = ems[i] + eis[i] + efs[i];
    }
}

/* dispinhalfspace_ (DispInHalfSpace.c:42-152): main + harmonic + image parts
 * summed in that order; for a triangle with all vertices at z = 0 the image
 * part's vertical component and then the total change sign. ENU. */
extern "C" void orc_disp_in_halfspace(double u[3], double X, double Y, double Z,
                                      const double P1[3], const double P2[3],
                                      const double P3[3], double Ss, double Ds, double Ts,
                                      double nu)
{
    const D3 x(X), y(Y), z(Z);
    const double M1[3] = {P1[0], P1[1], -P1[2]}, M2[3] = {P2[0], P2[1], -P2[2]},
                 M3[3] = {P3[0], P3[1], -P3[2]};
    D3 ums[3], ufs[3], uis[3];
    td_disp_fs(x, y, z, P1, P2, P3, Ss, Ds, Ts, nu, ums);
    td_disp_harfunc(x, y, z, P1, P2, P3, Ss, Ds, Ts, nu, ufs);
    td_disp_fs(x, y, z, M1, M2, M3, Ss, Ds, Ts, nu, uis);
    const bool surface = M1[2] == 0.0 && M2[2] == 0.0 && M3[2] == 0.0;
    if (surface) uis[2].v = -uis[2].v;
    for (int i = 0; i < 3; i++) {
        u[i] = ums[i].v + ufs[i].v + uis[i].v;
        if (surface) u[i] = -u[i];
    }
}

/* computeDisplacementNikkhoo15 (nikkhoo15.f90:324-367): NED point and
 * vertices to ENU, dip slip negated, result back to NED */
extern "C" void orc_displacement_nikkhoo15(double x1, double x2, double x3, const double P1[3],
                                           const double P2[3], const double P3[3], double ss,
                                           double d, double t, double G, double lambda,
                                           double u[3])
{
    const double nu = lambda / (lambda + G) / 2.0;
    const double Q1[3] = {P1[1], P1[0], -P1[2]}, Q2[3] = {P2[1], P2[0], -P2[2]},
                 Q3[3] = {P3[1], P3[0], -P3[2]};
    double w[3];
    orc_disp_in_halfspace(w, x2, x1, -x3, Q1, Q2, Q3, ss, -d, t, nu);
    u[0] = w[1];
    u[1] = w[0];
    u[2] = -w[2];
}

/* computeReferenceSystemNikkhoo15 (nikkhoo15.f90:50-101); 1-based indices */
extern "C" void orc_refsys_nikkhoo15(int nVe, const double *v, int ns, const int *i1,
                                     const int *i2, const int *i3, double *sv, double *dv,
                                     double *nv, double *xc)
{
    (void)nVe;
    for (int j = 0; j < ns; j++) {
        const double *A = v + 3 * (i1[j] - 1), *B = v + 3 * (i2[j] - 1), *C = v + 3 * (i3[j] - 1);
        double n[3] = {(B[1] - A[1]) * (C[2] - A[2]) - (B[2] - A[2]) * (C[1] - A[1]),
                       (B[2] - A[2]) * (C[0] - A[0]) - (B[0] - A[0]) * (C[2] - A[2]),
                       (B[0] - A[0]) * (C[1] - A[1]) - (B[1] - A[1]) * (C[0] - A[0])};
        const double norm = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; k++) n[k] = n[k] / norm;
        if (0.0 < n[2])
            for (int k = 0; k < 3; k++) n[k] = -n[k];
        for (int k = 0; k < 3; k++) nv[3 * j + k] = n[k];
        sv[3 * j] = std::cos(std::atan2(n[0], n[1]));
        sv[3 * j + 1] = -std::sin(std::atan2(n[0], n[1]));
        sv[3 * j + 2] = 0.0;
        const double *S = sv + 3 * j;
        dv[3 * j] = n[1] * S[2] - n[2] * S[1];
        dv[3 * j + 1] = n[2] * S[0] - n[0] * S[2];
        dv[3 * j + 2] = n[0] * S[1] - n[1] * S[0];
        for (int k = 0; k < 3; k++) xc[3 * j + k] = (A[k] + B[k] + C[k]) / 3.0;
    }
}

/* computeStressNikkhoo15 (nikkhoo15.f90:120-163): NED <-> ENU */
static void stress_nikkhoo15(double x1, double x2, double x3, const double *P1, const double *P2,
                             const double *P3, double ss, double d, double t, double G,
                             double lambda, double S[6])
{
    const double Q1[3] = {P1[1], P1[0], -P1[2]}, Q2[3] = {P2[1], P2[0], -P2[2]},
                 Q3[3] = {P3[1], P3[0], -P3[2]};
    double s[6], e[6];
    orc_strain_in_halfspace(s, e, x2, x1, -x3, Q1, Q2, Q3, ss, -d, t, G, lambda);
    S[0] = s[3];
    S[1] = s[1];
    S[2] = -s[4];
    S[3] = s[0];
    S[4] = -s[2];
    S[5] = s[5];
}

extern "C" void orc_project_traction(const double nv[3], const double sv[3], const double dv[3],
                                     const double S[6], double *ts, double *td, double *tn);
extern "C" void orc_project_volume_frame(const double sv[3], const double dv[3],
                                         const double nv[3], const double S[6], double out[6]);

/* computeTractionKernelsNikkhoo15 (nikkhoo15.f90:187-225) */
extern "C" void orc_traction_kernels_nikkhoo15(const double x[3], const double sv[3],
                                               const double dv[3], const double nv[3], int ns,
                                               const int *i1, const int *i2, const int *i3,
                                               int nVe, const double *vert, double s, double d,
                                               double t, double G, double lambda, double *ts,
                                               double *td, double *tn)
{
    (void)nVe;
    for (int i = 0; i < ns; i++) {
        double S[6];
        stress_nikkhoo15(x[0], x[1], x[2], vert + 3 * (i1[i] - 1), vert + 3 * (i2[i] - 1),
                         vert + 3 * (i3[i] - 1), s, d, t, G, lambda, S);
        orc_project_traction(nv, sv, dv, S, &ts[i], &td[i], &tn[i]);
    }
}

/* computeStressKernelsNikkhoo15 (nikkhoo15.f90:250-304) */
extern "C" void orc_stress_kernels_nikkhoo15(const double x[3], const double sv[3],
                                             const double dv[3], const double nv[3], int ns,
                                             const int *i1, const int *i2, const int *i3, int nVe,
                                             const double *vert, double s, double d, double t,
                                             double G, double lambda, double *s11, double *s12,
                                             double *s13, double *s22, double *s23, double *s33)
{
    (void)nVe;
    for (int i = 0; i < ns; i++) {
        double S[6], R[6];
        stress_nikkhoo15(x[0], x[1], x[2], vert + 3 * (i1[i] - 1), vert + 3 * (i2[i] - 1),
                         vert + 3 * (i3[i] - 1), s, d, t, G, lambda, S);
        orc_project_volume_frame(sv, dv, nv, S, R);
        s11[i] = R[0]; s12[i] = R[1]; s13[i] = R[2];
        s22[i] = R[3]; s23[i] = R[4]; s33[i] = R[5];
    }
}
