// tri_dev.h -- triangular dislocation (Nikkhoo & Walter 2015) stress for one
// (receiver, source) pair per lane, strike-slip and dip-slip at once.
//
// Follows jdpmoore/unicycle fortran/3d/src/StrainInHalfSpace.c:42-173
// (straininhalfspace_: main + image + free-surface harmonic parts) and
// nikkhoo15.f90:120-163.  Each part's strain is the exact gradient, by
// forward-mode differentiation with respect to the observation point, of the
// same part's displacement field (DispInHalfSpace.c:153-616), evaluated in
// binary64.  The displacement is linear in the Burgers vector, so every
// angular-dislocation term is formed once as a 3x3 basis and combined for the
// two unit slips.  Source-only geometry (frames, vertex angles, side
// transforms) is precomputed on the host with the C library (TriSource).
#pragma once
#include <hip/hip_runtime.h>

namespace ucy {

struct G3 {  // value and gradient w.r.t. the observation point
    double v, x, y, z;
};
__device__ __forceinline__ G3 gmk(double v, double x, double y, double z) { return {v, x, y, z}; }
__device__ __forceinline__ G3 gc(double c) { return {c, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ G3 operator+(G3 a, G3 b) { return {a.v + b.v, a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ G3 operator-(G3 a, G3 b) { return {a.v - b.v, a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ G3 operator-(G3 a) { return {-a.v, -a.x, -a.y, -a.z}; }
__device__ __forceinline__ G3 operator*(G3 a, G3 b)
{
    return {a.v * b.v, a.x * b.v + a.v * b.x, a.y * b.v + a.v * b.y, a.z * b.v + a.v * b.z};
}
__device__ __forceinline__ G3 operator/(G3 a, G3 b)
{
    const double q = a.v / b.v;
    return {q, (a.x - q * b.x) / b.v, (a.y - q * b.y) / b.v, (a.z - q * b.z) / b.v};
}
__device__ __forceinline__ G3 operator+(G3 a, double b) { return {a.v + b, a.x, a.y, a.z}; }
__device__ __forceinline__ G3 operator+(double b, G3 a) { return {a.v + b, a.x, a.y, a.z}; }
__device__ __forceinline__ G3 operator-(G3 a, double b) { return {a.v - b, a.x, a.y, a.z}; }
__device__ __forceinline__ G3 operator-(double b, G3 a) { return {b - a.v, -a.x, -a.y, -a.z}; }
__device__ __forceinline__ G3 operator*(G3 a, double b) { return {a.v * b, a.x * b, a.y * b, a.z * b}; }
__device__ __forceinline__ G3 operator*(double b, G3 a) { return {a.v * b, a.x * b, a.y * b, a.z * b}; }
__device__ __forceinline__ G3 operator/(G3 a, double b) { return {a.v / b, a.x / b, a.y / b, a.z / b}; }
__device__ __forceinline__ G3 operator/(double b, G3 a)
{
    const double q = b / a.v, s = -q / a.v;
    return {q, s * a.x, s * a.y, s * a.z};
}
__device__ __forceinline__ G3 gsqrt(G3 a)
{
    const double s = sqrt(a.v), d = 0.5 / s;
    return {s, d * a.x, d * a.y, d * a.z};
}
__device__ __forceinline__ G3 glog(G3 a) { return {log(a.v), a.x / a.v, a.y / a.v, a.z / a.v}; }
__device__ __forceinline__ G3 gatan(G3 a)
{
    const double d = 1.0 / (1.0 + a.v * a.v);
    return {atan(a.v), d * a.x, d * a.y, d * a.z};
}
__device__ __forceinline__ G3 gatan2(G3 y, G3 x)
{
    const double r2 = x.v * x.v + y.v * y.v;
    return {atan2(y.v, x.v), (x.v * y.x - y.v * x.x) / r2, (x.v * y.y - y.v * x.y) / r2,
            (x.v * y.z - y.v * x.z) / r2};
}

// full-space TD (main or image) geometry, DispInHalfSpace.c:153-240
struct TdFull {
    double A[3][3];       // rows: normal, strike, dip (TDCS axes in ENU)
    double P2[3];         // origin of the TDCS (second vertex)
    double p1y, p1z, p3y, p3z;
    double vy[3], vz[3];  // vertex y,z in the TDCS (p1, p2 = 0, p3)
    double s1[3][2];      // configuration I side vectors (y,z): -e13, e12, e23
    double s2[3][2];      // configuration II: e13, -e12, -e23
    double ca[3], sa[3];  // cos/sin(-pi + vertex angle)
};

// one TD side of the harmonic correction, DispInHalfSpace.c:462-539
struct TdSide {
    int skip;             // vertical side (beta ~ 0 or pi)
    int pad;
    double A[3][3];       // ADCS axes (ey1 ey2 ey3 as columns)
    double PA[3];
    double sA[3];         // side vector in ADCS
    double aA, aB;        // -PA_z, -PB_z
    double beta;
    double trig[2][4];    // for angle -pi+beta (conf. I) and beta: sin, cos, cot, cot(angle/2)
};

struct TriSource {
    TdFull full[2];       // main, image
    TdSide side[3];
    double bfs[2][3];     // EFCS slip of the two computed directions (harmonic part)
    double btd[2][3];     // TDCS Burgers vectors (tensile, strike, dip) of the two
                          // directions; used by the general-slip form only
    int surface;          // all vertices at z = 0 (image yz/zz sign quirk)
    int pad;
};

struct TriMedium {
    double mu, lambda, nu;
};

// barycentric configuration choice, DispInHalfSpace.c:389-417
__device__ __forceinline__ int tri_mode_d(double x, double y, double z, double p1a, double p1b,
                                          double p3a, double p3b)
{
    const double p2a = 0.0, p2b = 0.0;
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

// angular dislocation without the Burgers-function term (DispInHalfSpace.c:
// 541-571): basis B[i][j] = displacement component i for unit Burgers
// component j (x, y, z) in the angular-dislocation frame
__device__ __forceinline__ void angdis_basis(G3 x, G3 y, G3 z, double ca, double sa, double nu,
                                             G3 B[3][3])
{
    const G3 eta = y * ca - z * sa;
    G3 zeta = y * sa + z * ca;
    const G3 r = gsqrt(x * x + y * y + z * z);
    if (zeta.v > r.v) zeta = r;
    if (z.v > r.v) z = r;
    const double k = 1.0 / (8.0 * 3.14159265358979323846 * (1.0 - nu)), n12 = 1.0 - 2.0 * nu;
    const G3 rz = r - z, rzt = r - zeta;
    const G3 Lz = glog(rz), Lzt = glog(rzt);
    const G3 irz = 1.0 / (r * rz), irzt = 1.0 / (r * rzt), izt = 1.0 / rzt;
    B[0][0] = k * (x * y * irz - x * eta * irzt);
    B[1][0] = k * (eta * sa * izt - y * eta * irzt + y * y * irz + n12 * (ca * Lzt - Lz));
    B[2][0] = k * (eta * ca * izt - y / r - eta * z * irzt - n12 * sa * Lzt);
    B[0][1] = k * (x * x * ca * irzt - x * x * irz - n12 * (ca * Lzt - Lz));
    B[1][1] = k * x * (y * ca * irzt - sa * ca * izt - y * irz);
    B[2][1] = k * x * (z * ca * irzt - ca * ca * izt + 1.0 / r);
    B[0][2] = k * sa * (n12 * Lzt - x * x * irzt);
    B[1][2] = k * sa * x * (sa * izt - y * irzt);
    B[2][2] = k * sa * x * (ca * izt - z * irzt);
}

// full-space TD displacement for the two TDCS slips bs, bd (x,y,z = tensile,
// strike, dip components), DispInHalfSpace.c:153-300
__device__ __forceinline__ void td_full2(const TdFull &T, G3 X, G3 Y, G3 Z, double nu,
                                         const double bs[3], const double bd[3], G3 us[3], G3 ud[3])
{
    const G3 dX = X - T.P2[0], dY = Y - T.P2[1], dZ = Z - T.P2[2];
    const G3 x = T.A[0][0] * dX + T.A[0][1] * dY + T.A[0][2] * dZ;
    const G3 y = T.A[1][0] * dX + T.A[1][1] * dY + T.A[1][2] * dZ;
    const G3 z = T.A[2][0] * dX + T.A[2][1] * dY + T.A[2][2] * dZ;
    const int mode = tri_mode_d(y.v, z.v, x.v, T.p1y, T.p1z, T.p3y, T.p3z);
    G3 ds[3], dd[3];
    if (mode == 0) {
        const double q = __builtin_nan("");
        for (int i = 0; i < 3; i++) ds[i] = dd[i] = gmk(q, q, q, q);
    } else {
        for (int i = 0; i < 3; i++) ds[i] = dd[i] = gc(0.0);
        for (int v = 0; v < 3; v++) {
            const double s_1 = mode == 1 ? T.s1[v][0] : T.s2[v][0];
            const double s_2 = mode == 1 ? T.s1[v][1] : T.s2[v][1];
            // TDSetupD (DispInHalfSpace.c:419-460)
            const G3 ty = y - T.vy[v], tz = z - T.vz[v];
            const G3 y1 = s_2 * ty + (-s_1) * tz, z1 = s_1 * ty + s_2 * tz;
            G3 B[3][3];
            angdis_basis(x, y1, z1, T.ca[v], T.sa[v], nu, B);
            const double bys = s_2 * bs[1] + (-s_1) * bs[2], bzs = s_1 * bs[1] + s_2 * bs[2];
            const double byd = s_2 * bd[1] + (-s_1) * bd[2], bzd = s_1 * bd[1] + s_2 * bd[2];
            G3 es[3], ed[3];
            for (int i = 0; i < 3; i++) {
                es[i] = B[i][0] * bs[0] + B[i][1] * bys + B[i][2] * bzs;
                ed[i] = B[i][0] * bd[0] + B[i][1] * byd + B[i][2] * bzd;
            }
            ds[0] = ds[0] + es[0];
            ds[1] = ds[1] + (s_2 * es[1] + s_1 * es[2]);
            ds[2] = ds[2] + ((-s_1) * es[1] + s_2 * es[2]);
            dd[0] = dd[0] + ed[0];
            dd[1] = dd[1] + (s_2 * ed[1] + s_1 * ed[2]);
            dd[2] = dd[2] + ((-s_1) * ed[1] + s_2 * ed[2]);
        }
    }
    // Burgers function: solid angle of the TD (DispInHalfSpace.c:276-290)
    const G3 a0 = -x, a1 = T.p1y - y, a2 = T.p1z - z;
    const G3 b0 = -x, b1 = -y, b2 = -z;
    const G3 c0 = -x, c1 = T.p3y - y, c2 = T.p3z - z;
    const G3 na = gsqrt(a0 * a0 + a1 * a1 + a2 * a2), nb = gsqrt(b0 * b0 + b1 * b1 + b2 * b2),
             nc = gsqrt(c0 * c0 + c1 * c1 + c2 * c2);
    const G3 trip = a0 * (b1 * c2 - b2 * c1) - a1 * (b0 * c2 - b2 * c0) + a2 * (b0 * c1 - b1 * c0);
    const G3 den = na * nb * nc + (a0 * b0 + a1 * b1 + a2 * b2) * nc + (a0 * c0 + a1 * c1 + a2 * c2) * nb
                   + (b0 * c0 + b1 * c1 + b2 * c2) * na;
    const G3 Fi = -2.0 * gatan2(trip, den) / (4.0 * 3.14159265358979323846);
    for (int i = 0; i < 3; i++) {
        ds[i] = ds[i] + bs[i] * Fi;
        dd[i] = dd[i] + bd[i] * Fi;
    }
    for (int i = 0; i < 3; i++) {
        us[i] = T.A[0][i] * ds[0] + T.A[1][i] * ds[1] + T.A[2][i] * ds[2];
        ud[i] = T.A[0][i] * dd[0] + T.A[1][i] * dd[1] + T.A[2][i] * dd[2];
    }
}

// harmonic correction of an angular dislocation (DispInHalfSpace.c:577-606;
// Nikkhoo & Walter 2015), basis over the three ADCS Burgers components
__device__ __forceinline__ void fsc_basis(G3 y1, G3 y2, G3 y3, const double tr[4], double nu,
                                          double a, G3 B[3][3])
{
    const double sb = tr[0], cb = tr[1], cot = tr[2], cth = tr[3];
    const double k = 1.0 / (4.0 * 3.14159265358979323846 * (1.0 - nu)), N1 = 1.0 - 2.0 * nu,
                 nu1 = 1.0 - nu, c2 = cot * cot;
    const G3 y3b = y3 + 2.0 * a;
    const G3 z1b = y1 * cb + y3b * sb, z3b = -y1 * sb + y3b * cb;
    const G3 rb = gsqrt(y1 * y1 + y2 * y2 + y3b * y3b);
    const G3 Fib = 2.0 * gatan(-y2 / (-(rb + y3b) * cth + y1));
    const G3 W6 = rb + y3b, W7 = rb + z3b, ya = y3b - a;
    const G3 irb = 1.0 / rb, ar = a * irb, ir2 = irb * irb, ar2 = a * ir2, ir3 = ir2 * irb;
    const G3 L6 = glog(W6), L7 = glog(W7);
    const G3 q = rb * cb + y3b, s1 = rb * sb - y1;
    const G3 iW6 = 1.0 / W6, iW7 = 1.0 / W7;
    B[0][0] = k * (-2.0 * nu1 * N1 * Fib * c2 + N1 * y2 * iW6 * ((N1 - ar) * cot - y1 * iW6 * (nu + ar))
                   + N1 * y2 * cb * cot * iW7 * (cb + ar) + a * y2 * ya * cot * ir3
                   + y2 * ya * irb * iW6 * (-N1 * cot + y1 * iW6 * (2.0 * nu + ar) + a * y1 * ir2)
                   + y2 * ya * irb * iW7 * (cb * iW7 * (q * (N1 * cb - ar) * cot + 2.0 * nu1 * s1 * cb)
                                             - a * y3b * cb * cot * ir2));
    B[1][0] = k * (N1 * ((2.0 * nu1 * c2 - nu) * L6 - (2.0 * nu1 * c2 + N1) * cb * L7)
                   - N1 * iW6 * (y1 * cot * (N1 - ar) + nu * y3b - a + y2 * y2 * iW6 * (nu + ar))
                   - N1 * z1b * cot * iW7 * (cb + ar) - a * y1 * ya * cot * ir3
                   + ya * iW6 * (-2.0 * nu + irb * (N1 * y1 * cot - a) + y2 * y2 * irb * iW6 * (2.0 * nu + ar)
                                 + a * y2 * y2 * ir3)
                   + ya * iW7 * (cb * cb - irb * (N1 * z1b * cot + a * cb) + a * y3b * z1b * cot * ir3
                                 - irb * iW7 * (y2 * y2 * cb * cb - a * z1b * cot * irb * q)));
    B[2][0] = k * (2.0 * nu1 * (N1 * Fib * cot + y2 * iW6 * (2.0 * nu + ar) - y2 * cb * iW7 * (cb + ar))
                   + y2 * ya * irb * (2.0 * nu * iW6 + ar2)
                   + y2 * ya * cb * irb * iW7 * (N1 - q * iW7 * (cb + ar) - a * y3b * ir2));
    B[0][1] = k * (N1 * ((2.0 * nu1 * c2 + nu) * L6 - (2.0 * nu1 * c2 + 1.0) * cb * L7)
                   + N1 * iW6 * (-N1 * y1 * cot + nu * y3b - a + a * y1 * cot * irb + y1 * y1 * iW6 * (nu + ar))
                   - N1 * cot * iW7 * (z1b * cb - a * s1 * irb / cb) - a * y1 * ya * cot * ir3
                   + ya * iW6 * (2.0 * nu + irb * (N1 * y1 * cot + a) - y1 * y1 * irb * iW6 * (2.0 * nu + ar)
                                 - a * y1 * y1 * ir3)
                   + ya * cot * iW7 * (-cb * sb + a * y1 * y3b * ir3 / cb
                                       + s1 * irb * (2.0 * nu1 * cb - q * iW7 * (1.0 + a * irb / cb))));
    B[1][1] = k * (2.0 * nu1 * N1 * Fib * c2 + N1 * y2 * iW6 * (-(N1 - ar) * cot + y1 * iW6 * (nu + ar))
                   - N1 * y2 * cot * iW7 * (1.0 + a * irb / cb) - a * y2 * ya * cot * ir3
                   + y2 * ya * irb * iW6 * (N1 * cot - 2.0 * nu * y1 * iW6 - a * y1 * irb * (irb + iW6))
                   + y2 * ya * cot * irb * iW7 * (-2.0 * nu1 * cb + q * iW7 * (1.0 + a * irb / cb)
                                                  + a * y3b * ir2 / cb));
    B[2][1] = k * (-2.0 * nu1 * N1 * cot * (L6 - cb * L7) - 2.0 * nu1 * y1 * iW6 * (2.0 * nu + ar)
                   + 2.0 * nu1 * z1b * iW7 * (cb + ar) + ya * irb * (N1 * cot - 2.0 * nu * y1 * iW6 - a * y1 * ir2)
                   - ya * iW7 * (cb * sb + q * cot * irb * (2.0 * nu1 * cb - q * iW7)
                                 + ar * (sb - y3b * z1b * ir2 - z1b * q * irb * iW7)));
    B[0][2] = k * (N1 * (y2 * iW6 * (1.0 + ar) - y2 * cb * iW7 * (cb + ar)) - y2 * ya * irb * (ar2 + iW6)
                   + y2 * ya * cb * irb * iW7 * (q * iW7 * (cb + ar) + a * y3b * ir2));
    B[1][2] = k * (N1 * (-sb * L7 - y1 * iW6 * (1.0 + ar) + z1b * iW7 * (cb + ar)) + y1 * ya * irb * (ar2 + iW6)
                   - ya * iW7 * (sb * (cb - ar) + z1b * irb * (1.0 + a * y3b * ir2)
                                 - irb * iW7 * (y2 * y2 * cb * sb - a * z1b * irb * q)));
    B[2][2] = k * (2.0 * nu1 * Fib + 2.0 * nu1 * (y2 * sb * iW7 * (cb + ar))
                   + y2 * ya * sb * irb * iW7 * (1.0 + q * iW7 * (cb + ar) + a * y3b * ir2));
}

// harmonic part for both EFCS slips (DispInHalfSpace.c:306-368, 462-539)
__device__ __forceinline__ void td_fsc2(const TriSource &S, G3 X, G3 Y, G3 Z, double nu, G3 us[3],
                                        G3 ud[3])
{
    for (int i = 0; i < 3; i++) us[i] = ud[i] = gc(0.0);
    for (int sd = 0; sd < 3; sd++) {
        const TdSide &T = S.side[sd];
        if (T.skip) continue;
        const G3 dx = X - T.PA[0], dy = Y - T.PA[1], dz = Z - T.PA[2];
        G3 yA[3], yB[3];
        for (int i = 0; i < 3; i++) {
            yA[i] = T.A[i][0] * dx + T.A[i][1] * dy + T.A[i][2] * dz;
            yB[i] = yA[i] - T.sA[i];
        }
        const int cfg = (T.beta * yA[0].v >= 0.0) ? 0 : 1;
        G3 BA[3][3], BB[3][3];
        fsc_basis(yA[0], yA[1], yA[2], T.trig[cfg], nu, T.aA, BA);
        fsc_basis(yB[0], yB[1], yB[2], T.trig[cfg], nu, T.aB, BB);
        double bsA[3], bdA[3];
        for (int i = 0; i < 3; i++) {
            bsA[i] = T.A[i][0] * S.bfs[0][0] + T.A[i][1] * S.bfs[0][1] + T.A[i][2] * S.bfs[0][2];
            bdA[i] = T.A[i][0] * S.bfs[1][0] + T.A[i][1] * S.bfs[1][1] + T.A[i][2] * S.bfs[1][2];
        }
        G3 ws[3], wd[3];
        for (int i = 0; i < 3; i++) {
            const G3 d0 = BB[i][0] - BA[i][0], d1 = BB[i][1] - BA[i][1], d2 = BB[i][2] - BA[i][2];
            ws[i] = bsA[0] * d0 + bsA[1] * d1 + bsA[2] * d2;
            wd[i] = bdA[0] * d0 + bdA[1] * d1 + bdA[2] * d2;
        }
        // back to EFCS with the transposed axes (rows ey1, ey2, ey3)
        for (int i = 0; i < 3; i++) {
            us[i] = us[i] + (T.A[0][i] * ws[0] + T.A[1][i] * ws[1] + T.A[2][i] * ws[2]);
            ud[i] = ud[i] + (T.A[0][i] * wd[0] + T.A[1][i] * wd[1] + T.A[2][i] * wd[2]);
        }
    }
}

__device__ __forceinline__ void sym_strain(const G3 u[3], double e[6])
{
    e[0] = u[0].x;
    e[1] = (u[0].y + u[1].x) / 2.0;
    e[2] = (u[0].z + u[2].x) / 2.0;
    e[3] = u[1].y;
    e[4] = (u[1].z + u[2].y) / 2.0;
    e[5] = u[2].z;
}

__device__ __forceinline__ void hooke6(const double e[6], double mu, double lambda, double s[6])
{
    const double tr = e[0] + e[3] + e[5];
    s[0] = 2.0 * mu * e[0] + lambda * tr;
    s[3] = 2.0 * mu * e[3] + lambda * tr;
    s[5] = 2.0 * mu * e[5] + lambda * tr;
    s[1] = 2.0 * mu * e[1];
    s[2] = 2.0 * mu * e[2];
    s[4] = 2.0 * mu * e[4];
}

// NED stress (s11,s12,s13,s22,s23,s33) at NED point (x1,x2,x3) for two slip
// vectors of the source: straininhalfspace_ in ENU with the conversions of
// nikkhoo15.f90:130-161.  GEN = false: unit strike (Ss) and unit dip (Sd),
// with the Burgers vectors as compile-time constants; GEN = true: the two
// TDCS Burgers vectors stored in the source (any strike / dip / opening).
template <bool GEN = false>
__device__ __forceinline__ void tri_stress2(const TriMedium &M, const TriSource &S, double x1,
                                            double x2, double x3, double Ss[6], double Sd[6])
{
    const G3 X = gmk(x2, 1, 0, 0), Y = gmk(x1, 0, 1, 0), Z = gmk(-x3, 0, 0, 1);
    // TDCS Burgers vectors (tensile, strike, dip): strike unit -> (0,1,0);
    // dip unit d=1 enters as Ds=-d (nikkhoo15.f90:142) -> (0,0,-1)
    const double cs[3] = {0.0, 1.0, 0.0}, cd[3] = {0.0, 0.0, -1.0};
    const double *bs = GEN ? S.btd[0] : cs, *bd = GEN ? S.btd[1] : cd;
    double stress[2][6];
    for (int k = 0; k < 2; k++)
        for (int i = 0; i < 6; i++) stress[k][i] = 0.0;
    G3 us[3], ud[3];
    double e[6], s[6];
    // main dislocation
    td_full2(S.full[0], X, Y, Z, M.nu, bs, bd, us, ud);
    double ms[6], md[6];
    sym_strain(us, e);
    hooke6(e, M.mu, M.lambda, ms);
    sym_strain(ud, e);
    hooke6(e, M.mu, M.lambda, md);
    // image dislocation
    td_full2(S.full[1], X, Y, Z, M.nu, bs, bd, us, ud);
    double is_[6], id[6];
    sym_strain(us, e);
    hooke6(e, M.mu, M.lambda, is_);
    sym_strain(ud, e);
    hooke6(e, M.mu, M.lambda, id);
    if (S.surface) {  // StrainInHalfSpace.c:162-165
        is_[4] = -is_[4];
        is_[5] = -is_[5];
        id[4] = -id[4];
        id[5] = -id[5];
    }
    // harmonic correction
    td_fsc2(S, X, Y, Z, M.nu, us, ud);
    double fs[6], fd[6];
    sym_strain(us, e);
    hooke6(e, M.mu, M.lambda, fs);
    sym_strain(ud, e);
    hooke6(e, M.mu, M.lambda, fd);
    for (int i = 0; i < 6; i++) {
        s[i] = ms[i] + is_[i] + fs[i];
        stress[0][i] = s[i];
        stress[1][i] = md[i] + id[i] + fd[i];
    }
    // ENU (xx,xy,xz,yy,yz,zz) -> NED (nikkhoo15.f90:156-161)
    double *out[2] = {Ss, Sd};
    for (int k = 0; k < 2; k++) {
        out[k][0] = stress[k][3];
        out[k][1] = stress[k][1];
        out[k][2] = -stress[k][4];
        out[k][3] = stress[k][0];
        out[k][4] = -stress[k][2];
        out[k][5] = stress[k][5];
    }
}

// NED displacement at NED point (x1, x2, x3) for unit strike slip (us) and
// unit dip slip (ud): computeDisplacementNikkhoo15 (nikkhoo15.f90:324-367)
// over dispinhalfspace_ (DispInHalfSpace.c:42-152).  The main, harmonic and
// image parts are summed in that order; with every vertex at the surface the
// image part's vertical component and then the total change sign.  Constant
// seeds: only the values of the G3 arithmetic are used.
__device__ __forceinline__ void tri_disp2(const TriMedium &M, const TriSource &S, double x1,
                                          double x2, double x3, double us[3], double ud[3])
{
    const G3 X = gc(x2), Y = gc(x1), Z = gc(-x3);
    const double cs[3] = {0.0, 1.0, 0.0}, cd[3] = {0.0, 0.0, -1.0};
    G3 ms[3], md[3], fs[3], fd[3], is_[3], id[3];
    td_full2(S.full[0], X, Y, Z, M.nu, cs, cd, ms, md);
    td_fsc2(S, X, Y, Z, M.nu, fs, fd);
    td_full2(S.full[1], X, Y, Z, M.nu, cs, cd, is_, id);
    double ws[3], wd[3];
    for (int i = 0; i < 3; i++) {
        const bool flip = S.surface && i == 2;
        ws[i] = ms[i].v + fs[i].v + (flip ? -is_[i].v : is_[i].v);
        wd[i] = md[i].v + fd[i].v + (flip ? -id[i].v : id[i].v);
        if (S.surface) {
            ws[i] = -ws[i];
            wd[i] = -wd[i];
        }
    }
    // east, north, up -> north, east, down
    us[0] = ws[1];
    us[1] = ws[0];
    us[2] = -ws[2];
    ud[0] = wd[1];
    ud[1] = wd[0];
    ud[2] = -wd[2];
}

}  // namespace ucy
