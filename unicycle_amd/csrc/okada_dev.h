// okada_dev.h -- Okada (1992) rectangular dislocation for one (receiver,
// source) pair per lane, evaluated for the strike-slip and dip-slip unit
// sources at once.
//
// Follows jdpmoore/unicycle fortran/3d/src/dc3d.f (DC3D, UA, UB, UC, DCCON0,
// DCCON2) and okada92.f90:109-207 (computeStressOkada92).  The reference
// evaluates one unit slip per call; both calls see identical geometry, so the
// geometry terms (one DCCON2 per corner, the I/J/K integrals of UB, the C/Z
// terms of UC) are shared here and only the slip-dependent combinations are
// formed twice.  Rounding points are the reference's: the DC3D interface is
// binary32 (okada92.f90:138-160, dc3d.f:168-179), XI = X-AL and D = DEPTH+-Z
// are binary32 sums (dc3d.f:53-54,60,113), the DC3D singularity EPS is the
// binary32 value of 1e-6 (dc3d.f:5-6,36), DCCON2 snaps in place with the
// binary64 EPS (dc3d.f:612-616).  Compile with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

namespace ucy {

// Per-source constants precomputed on the host with the C library's cos/sin,
// exactly as okada92.f90:146-160 and DCCON0 (dc3d.f:531-543) form them.
struct OkadaSource {
    double q1, q2;          // upper-left corner (north, east)
    double csst, ssst;      // cos/sin of strike
    double sd, cd;          // DCCON0 sin/cos of the binary32 dip in degrees
    double sdsd, cdcd, sdcd;
    float depth, al2, aw1;  // SNGL(q3), SNGL(L), -SNGL(W)
    float pad;
};

struct OkadaMedium {
    double alp1, alp2, alp3, alp4, alp5;  // DCCON0 (dc3d.f:525-529)
    double G, lambda;
};

struct DcGeo {
    double xi2, et2, q2, r, r2, r3, r5, y, d, tt, alx, ale, x11, y11, x32, y32;
    double ey, ez, fy, fz, gy, gz, hy, hz;
};

__device__ __forceinline__ void dc_geo(DcGeo &g, double &xi, double &et, double &q, double sd,
                                       double cd, int kxi, int ket)
{
    const double eps = 1.0e-6;
    if (fabs(xi) < eps) xi = 0.0;
    if (fabs(et) < eps) et = 0.0;
    if (fabs(q) < eps) q = 0.0;
    g.xi2 = xi * xi;
    g.et2 = et * et;
    g.q2 = q * q;
    g.r2 = g.xi2 + g.et2 + g.q2;
    g.r = sqrt(g.r2);
    if (g.r == 0.0) return;  // the reference keeps the previous corner's terms
    g.r3 = g.r * g.r2;
    g.r5 = g.r3 * g.r2;
    g.y = et * cd + q * sd;
    g.d = et * sd - q * cd;
    g.tt = (q == 0.0) ? 0.0 : atan(xi * et / (q * g.r));
    if (kxi == 1) {
        g.alx = -log(g.r - xi);
        g.x11 = 0.0;
        g.x32 = 0.0;
    } else {
        double rxi = g.r + xi;
        g.alx = log(rxi);
        g.x11 = 1.0 / (g.r * rxi);
        g.x32 = (g.r + rxi) * g.x11 * g.x11 / g.r;
    }
    if (ket == 1) {
        g.ale = -log(g.r - et);
        g.y11 = 0.0;
        g.y32 = 0.0;
    } else {
        double ret = g.r + et;
        g.ale = log(ret);
        g.y11 = 1.0 / (g.r * ret);
        g.y32 = (g.r + ret) * g.y11 * g.y11 / g.r;
    }
    g.ey = sd / g.r - g.y * q / g.r3;
    g.ez = cd / g.r + g.d * q / g.r3;
    g.fy = g.d / g.r3 + g.xi2 * g.y32 * sd;
    g.fz = g.y / g.r3 + g.xi2 * g.y32 * cd;
    g.gy = 2.0 * g.x11 * sd - g.y * q * g.x32;
    g.gz = 2.0 * g.x11 * cd + g.d * q * g.x32;
    g.hy = g.d * q * g.x32 + xi * q * g.y32 * sd;
    g.hz = g.y * q * g.x32 + xi * q * g.y32 * cd;
}

// UA for unit strike slip (s) and unit dip slip (d); dc3d.f:222-261.
// The reference scales each term by DISL/PI2 with DISL=1 (dc3d.f:242,261).
__device__ __forceinline__ void dc_ua2(const OkadaMedium &M, const OkadaSource &S, const DcGeo &g,
                                       double xi, double et, double q, double us[12], double ud[12])
{
    const double c = 1.0 / 6.283185307179586;
    const double a1 = M.alp1, a2 = M.alp2, sd = S.sd, cd = S.cd;
    const double xy = xi * g.y11, qx = q * g.x11, qy = q * g.y11;
    us[0] = c * (g.tt / 2.0 + a2 * xi * qy);
    us[1] = c * (a2 * q / g.r);
    us[2] = c * (a1 * g.ale - a2 * q * qy);
    us[3] = c * (-a1 * qy - a2 * g.xi2 * q * g.y32);
    us[4] = c * (-a2 * xi * q / g.r3);
    us[5] = c * (a1 * xy + a2 * xi * g.q2 * g.y32);
    us[6] = c * (a1 * xy * sd + a2 * xi * g.fy + g.d / 2.0 * g.x11);
    us[7] = c * (a2 * g.ey);
    us[8] = c * (a1 * (cd / g.r + qy * sd) - a2 * q * g.fy);
    us[9] = c * (a1 * xy * cd + a2 * xi * g.fz + g.y / 2.0 * g.x11);
    us[10] = c * (a2 * g.ez);
    us[11] = c * (-a1 * (sd / g.r - qy * cd) - a2 * q * g.fz);
    ud[0] = c * (a2 * q / g.r);
    ud[1] = c * (g.tt / 2.0 + a2 * et * qx);
    ud[2] = c * (a1 * g.alx - a2 * q * qx);
    ud[3] = c * (-a2 * xi * q / g.r3);
    ud[4] = c * (-qy / 2.0 - a2 * et * q / g.r3);
    ud[5] = c * (a1 / g.r + a2 * g.q2 / g.r3);
    ud[6] = c * (a2 * g.ey);
    ud[7] = c * (a1 * g.d * g.x11 + xy / 2.0 * sd + a2 * et * g.gy);
    ud[8] = c * (a1 * g.y * g.x11 - a2 * q * g.gy);
    ud[9] = c * (a2 * g.ez);
    ud[10] = c * (a1 * g.y * g.x11 + xy / 2.0 * cd + a2 * et * g.gz);
    ud[11] = c * (-a1 * g.d * g.x11 - a2 * q * g.gz);
}

// UB, dc3d.f:304-380
__device__ __forceinline__ void dc_ub2(const OkadaMedium &M, const OkadaSource &S, const DcGeo &g,
                                       double xi, double et, double q, double us[12], double ud[12])
{
    const double c = 1.0 / 6.283185307179586;
    const double sd = S.sd, cd = S.cd, a3 = M.alp3, sdcd = S.sdcd, cdcd = S.cdcd;
    const double r = g.r, y = g.y, d = g.d;
    const double rd = r + d;
    const double d11 = 1.0 / (r * rd);
    const double aj2 = xi * y / rd * d11;
    const double aj5 = -(d + y * y / rd) * d11;
    double ai3, ai4, ak1, ak3, aj3, aj6;
    if (cd != 0.0) {
        if (xi == 0.0) {
            ai4 = 0.0;
        } else {
            double x = sqrt(g.xi2 + g.q2);
            ai4 = 1.0 / cdcd * (xi / rd * sdcd
                                + 2.0 * atan((et * (x + q * cd) + x * (r + x) * sd)
                                             / (xi * (r + x) * cd)));
        }
        ai3 = (y * cd / rd - g.ale + sd * log(rd)) / cdcd;
        ak1 = xi * (d11 - g.y11 * sd) / cd;
        ak3 = (q * g.y11 - y * d11) / cd;
        aj3 = (ak1 - aj2 * sd) / cd;
        aj6 = (ak3 - aj5 * sd) / cd;
    } else {
        double rd2 = rd * rd;
        ai3 = (et / rd + y * q / rd2 - g.ale) / 2.0;
        ai4 = xi * y / rd2 / 2.0;
        ak1 = xi * q / rd * d11;
        ak3 = sd / rd * (g.xi2 * d11 - 1.0);
        aj3 = -xi / rd2 * (g.q2 * d11 - 1.0 / 2.0);
        aj6 = -y / rd2 * (g.xi2 * d11 - 1.0 / 2.0);
    }
    const double xy = xi * g.y11;
    const double ai1 = -xi / rd * cd - ai4 * sd;
    const double ai2 = log(rd) + ai3 * sd;
    const double ak2 = 1.0 / r + ak3 * sd;
    const double ak4 = xy * cd - ak1 * sd;
    const double aj1 = aj5 * cd - aj6 * sd;
    const double aj4 = -xy - aj2 * cd + aj3 * sd;
    const double qx = q * g.x11, qy = q * g.y11;
    us[0] = c * (-xi * qy - g.tt - a3 * ai1 * sd);
    us[1] = c * (-q / r + a3 * y / rd * sd);
    us[2] = c * (q * qy - a3 * ai2 * sd);
    us[3] = c * (g.xi2 * q * g.y32 - a3 * aj1 * sd);
    us[4] = c * (xi * q / g.r3 - a3 * aj2 * sd);
    us[5] = c * (-xi * g.q2 * g.y32 - a3 * aj3 * sd);
    us[6] = c * (-xi * g.fy - d * g.x11 + a3 * (xy + aj4) * sd);
    us[7] = c * (-g.ey + a3 * (1.0 / r + aj5) * sd);
    us[8] = c * (q * g.fy - a3 * (qy - aj6) * sd);
    us[9] = c * (-xi * g.fz - y * g.x11 + a3 * ak1 * sd);
    us[10] = c * (-g.ez + a3 * y * d11 * sd);
    us[11] = c * (q * g.fz + a3 * ak2 * sd);
    ud[0] = c * (-q / r + a3 * ai3 * sdcd);
    ud[1] = c * (-et * qx - g.tt - a3 * xi / rd * sdcd);
    ud[2] = c * (q * qx + a3 * ai4 * sdcd);
    ud[3] = c * (xi * q / g.r3 + a3 * aj4 * sdcd);
    ud[4] = c * (et * q / g.r3 + qy + a3 * aj5 * sdcd);
    ud[5] = c * (-g.q2 / g.r3 + a3 * aj6 * sdcd);
    ud[6] = c * (-g.ey + a3 * aj1 * sdcd);
    ud[7] = c * (-et * g.gy - xy * sd + a3 * aj2 * sdcd);
    ud[8] = c * (q * g.gy + a3 * aj3 * sdcd);
    ud[9] = c * (-g.ez - a3 * ak3 * sdcd);
    ud[10] = c * (-et * g.gz - xy * cd - a3 * xi * d11 * sdcd);
    ud[11] = c * (q * g.gz - a3 * ak4 * sdcd);
}

// UC, dc3d.f:422-483
__device__ __forceinline__ void dc_uc2(const OkadaMedium &M, const OkadaSource &S, const DcGeo &g,
                                       double xi, double et, double q, double z, double us[12],
                                       double ud[12])
{
    const double cc = 1.0 / 6.283185307179586;
    const double sd = S.sd, cd = S.cd, a4 = M.alp4, a5 = M.alp5, sdsd = S.sdsd, sdcd = S.sdcd;
    const double r = g.r, r2 = g.r2, r3 = g.r3, r5 = g.r5, y = g.y, d = g.d;
    const double xi2 = g.xi2, et2 = g.et2, q2 = g.q2, x11 = g.x11, y11 = g.y11;
    const double x32 = g.x32, y32 = g.y32;
    const double c = d + z;
    const double x53 = (8.0 * r2 + 9.0 * r * xi + 3.0 * xi2) * x11 * x11 * x11 / r2;
    const double y53 = (8.0 * r2 + 9.0 * r * et + 3.0 * et2) * y11 * y11 * y11 / r2;
    const double h = q * cd - z;
    const double z32 = sd / r3 - h * y32;
    const double z53 = 3.0 * sd / r5 - h * y53;
    const double y0 = y11 - xi2 * y32;
    const double z0 = z32 - xi2 * z53;
    const double ppy = cd / r3 + q * y32 * sd;
    const double ppz = sd / r3 - q * y32 * cd;
    const double qq = z * y32 + z32 + z0;
    const double qqy = 3.0 * c * d / r5 - qq * sd;
    const double qqz = 3.0 * c * y / r5 - qq * cd + q * y32;
    const double xy = xi * y11, qy = q * y11;
    const double qr = 3.0 * q / r5;
    const double cdr = (c + d) / r3;
    const double yy0 = y / r3 - y0 * cd;
    (void)y53;
    us[0] = cc * (a4 * xy * cd - a5 * xi * q * z32);
    us[1] = cc * (a4 * (cd / r + 2.0 * qy * sd) - a5 * c * q / r3);
    us[2] = cc * (a4 * qy * cd - a5 * (c * et / r3 - z * y11 + xi2 * z32));
    us[3] = cc * (a4 * y0 * cd - a5 * q * z0);
    us[4] = cc * (-a4 * xi * (cd / r3 + 2.0 * q * y32 * sd) + a5 * c * xi * qr);
    us[5] = cc * (-a4 * xi * q * y32 * cd + a5 * xi * (3.0 * c * et / r5 - qq));
    us[6] = cc * (-a4 * xi * ppy * cd - a5 * xi * qqy);
    us[7] = cc * (a4 * 2.0 * (d / r3 - y0 * sd) * sd - y / r3 * cd
                  - a5 * (cdr * sd - et / r3 - c * y * qr));
    us[8] = cc * (-a4 * q / r3 + yy0 * sd + a5 * (cdr * cd + c * d * qr - (y0 * cd + q * z0) * sd));
    us[9] = cc * (a4 * xi * ppz * cd - a5 * xi * qqz);
    us[10] = cc * (a4 * 2.0 * (y / r3 - y0 * cd) * sd + d / r3 * cd - a5 * (cdr * cd + c * d * qr));
    us[11] = cc * (yy0 * cd - a5 * (cdr * sd - c * y * qr - y0 * sdsd + q * z0 * cd));
    ud[0] = cc * (a4 * cd / r - qy * sd - a5 * c * q / r3);
    ud[1] = cc * (a4 * y * x11 - a5 * c * et * q * x32);
    ud[2] = cc * (-d * x11 - xy * sd - a5 * c * (x11 - q2 * x32));
    ud[3] = cc * (-a4 * xi / r3 * cd + a5 * c * xi * qr + xi * q * y32 * sd);
    ud[4] = cc * (-a4 * y / r3 + a5 * c * et * qr);
    ud[5] = cc * (d / r3 - y0 * sd + a5 * c / r3 * (1.0 - 3.0 * q2 / r2));
    ud[6] = cc * (-a4 * et / r3 + y0 * sdsd - a5 * (cdr * sd - c * y * qr));
    ud[7] = cc * (a4 * (x11 - y * y * x32) - a5 * c * ((d + 2.0 * q * cd) * x32 - y * et * q * x53));
    ud[8] = cc * (xi * ppy * sd + y * d * x32 + a5 * c * ((y + 2.0 * q * sd) * x32 - y * q2 * x53));
    ud[9] = cc * (-q / r3 + y0 * sdcd - a5 * (cdr * cd + c * d * qr));
    ud[10] = cc * (a4 * y * d * x32 - a5 * c * ((y - 2.0 * q * sd) * x32 + d * et * q * x53));
    ud[11] = cc * (-xi * ppz * sd + x11 - d * d * x32
                   - a5 * c * ((d - 2.0 * q * cd) * x32 - d * q2 * x53));
}

// Rotation of one corner's contribution into the DC3D frame and signed
// accumulation (dc3d.f:94-106 real source, 149-163 image source).
__device__ __forceinline__ void dc_accum_real(double U[12], const double a[12], double sd,
                                              double cd, bool minus)
{
#pragma unroll
    for (int i = 0; i < 12; i += 3) {
        double d0 = -a[i];
        double d1 = -a[i + 1] * cd + a[i + 2] * sd;
        double d2 = -a[i + 1] * sd - a[i + 2] * cd;
        if (i == 9) { d0 = -d0; d1 = -d1; d2 = -d2; }
        if (minus) { U[i] -= d0; U[i + 1] -= d1; U[i + 2] -= d2; }
        else       { U[i] += d0; U[i + 1] += d1; U[i + 2] += d2; }
    }
}

__device__ __forceinline__ void dc_accum_image(double U[12], const double a[12], const double b[12],
                                               const double cc[12], double Z, double sd, double cd,
                                               bool minus)
{
#pragma unroll
    for (int i = 0; i < 12; i += 3) {
        double d0 = a[i] + b[i] + Z * cc[i];
        double d1 = (a[i + 1] + b[i + 1] + Z * cc[i + 1]) * cd - (a[i + 2] + b[i + 2] + Z * cc[i + 2]) * sd;
        double d2 = (a[i + 1] + b[i + 1] - Z * cc[i + 1]) * sd + (a[i + 2] + b[i + 2] - Z * cc[i + 2]) * cd;
        if (i == 9) {
            d0 = d0 + cc[0];
            d1 = d1 + cc[1] * cd - cc[2] * sd;
            d2 = d2 - cc[1] * sd - cc[2] * cd;
        }
        if (minus) { U[i] -= d0; U[i + 1] -= d1; U[i + 2] -= d2; }
        else       { U[i] += d0; U[i + 1] += d1; U[i + 2] += d2; }
    }
}

// DC3D for unit strike slip and unit dip slip.  Returns 1 on the singular
// edge (all outputs zero), 0 otherwise.  Writes outputs u[OUT0 .. OUT0+N) of
// the reference's 12, rounded to binary32 as dc3d.f:168-179: the stress path
// takes the displacement gradient (OUT0 = 3, N = 9), the observation-point
// path the displacement (OUT0 = 0, N = 3).
template <int OUT0, int N>
__device__ __forceinline__ int dc3d_pair_out(const OkadaMedium &M, const OkadaSource &S, float X,
                                             float Y, float Z, float gs[N], float gd[N])
{
    const float epsf = 1.0e-6f;
    double Us[12], Ud[12];
#pragma unroll
    for (int i = 0; i < 12; i++) { Us[i] = 0.0; Ud[i] = 0.0; }
    const double sd = S.sd, cd = S.cd;
    double xi[2], et[2];
    xi[0] = (double)(float)(X - 0.0f);
    xi[1] = (double)(float)(X - S.al2);
    if (fabs(xi[0]) < (double)epsf) xi[0] = 0.0;
    if (fabs(xi[1]) < (double)epsf) xi[1] = 0.0;
    DcGeo g = {};
    double a[12], b[12], cu[12], ad[12], bd[12], cd_[12];
    for (int pass = 0; pass < 2; pass++) {
        const double D = pass == 0 ? (double)(float)(S.depth + Z) : (double)(float)(S.depth - Z);
        const double P = (double)Y * cd + D * sd;
        double Q = (double)Y * sd - D * cd;
        et[0] = P - (double)S.aw1;
        et[1] = P - 0.0;
        if (fabs(Q) < (double)epsf) Q = 0.0;
        if (fabs(et[0]) < (double)epsf) et[0] = 0.0;
        if (fabs(et[1]) < (double)epsf) et[1] = 0.0;
        if (Q == 0.0 && ((xi[0] * xi[1] <= 0.0 && et[0] * et[1] == 0.0) ||
                         (et[0] * et[1] <= 0.0 && xi[0] * xi[1] == 0.0))) {
#pragma unroll
            for (int i = 0; i < N; i++) { gs[i] = 0.0f; gd[i] = 0.0f; }
            return 1;
        }
        int kxi[2] = {0, 0}, ket[2] = {0, 0};
        const double r12 = sqrt(xi[0] * xi[0] + et[1] * et[1] + Q * Q);
        const double r21 = sqrt(xi[1] * xi[1] + et[0] * et[0] + Q * Q);
        const double r22 = sqrt(xi[1] * xi[1] + et[1] * et[1] + Q * Q);
        if (xi[0] < 0.0 && r21 + xi[1] < (double)epsf) kxi[0] = 1;
        if (xi[0] < 0.0 && r22 + xi[1] < (double)epsf) kxi[1] = 1;
        if (et[0] < 0.0 && r12 + et[1] < (double)epsf) ket[0] = 1;
        if (et[0] < 0.0 && r22 + et[1] < (double)epsf) ket[1] = 1;
        for (int k = 0; k < 2; k++) {
            for (int j = 0; j < 2; j++) {
                dc_geo(g, xi[j], et[k], Q, sd, cd, kxi[k], ket[j]);
                const bool minus = (j + k == 1);
                dc_ua2(M, S, g, xi[j], et[k], Q, a, ad);
                if (pass == 0) {
                    dc_accum_real(Us, a, sd, cd, minus);
                    dc_accum_real(Ud, ad, sd, cd, minus);
                } else {
                    dc_ub2(M, S, g, xi[j], et[k], Q, b, bd);
                    dc_uc2(M, S, g, xi[j], et[k], Q, (double)Z, cu, cd_);
                    dc_accum_image(Us, a, b, cu, (double)Z, sd, cd, minus);
                    dc_accum_image(Ud, ad, bd, cd_, (double)Z, sd, cd, minus);
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++) { gs[i] = (float)Us[i + OUT0]; gd[i] = (float)Ud[i + OUT0]; }
    return 0;
}

__device__ __forceinline__ int dc3d_pair(const OkadaMedium &M, const OkadaSource &S, float X,
                                         float Y, float Z, float gs[9], float gd[9])
{
    return dc3d_pair_out<3, 9>(M, S, X, Y, Z, gs, gd);
}

// computeStressOkada92 (okada92.f90:174-205) from the binary32 gradients.
// g = (uxx,uyx,uzx,uxy,uyy,uzy,uxz,uyz,uzz) in DC3D order.
__device__ __forceinline__ void okada_stress_from_grad(const OkadaMedium &M, double cs, double ss,
                                                       const float gr[9], double sig[6])
{
    const double uxx = gr[0], uyx = gr[1], uzx = gr[2], uxy = gr[3], uyy = gr[4], uzy = gr[5];
    const double uxz = gr[6], uyz = gr[7], uzz = gr[8];
    const double uxxr = (cs * uxx + ss * uyx) * cs - (-cs * uxy - ss * uyy) * ss;
    const double uxyr = (cs * uxx + ss * uyx) * ss + (-cs * uxy - ss * uyy) * cs;
    const double uxzr = -cs * uxz + ss * uyz;
    const double uyxr = (ss * uxx - cs * uyx) * cs - (-ss * uxy + cs * uyy) * ss;
    const double uyyr = (ss * uxx - cs * uyx) * ss + (-ss * uxy + cs * uyy) * cs;
    const double uyzr = -ss * uxz + cs * uyz;
    const double uzxr = -cs * uzx - ss * uzy;
    const double uzyr = -ss * uzx + cs * uzy;
    const double e11 = uxxr, e12 = (uxyr + uyxr) / 2.0, e13 = (uxzr + uzxr) / 2.0;
    const double e22 = uyyr, e23 = (uyzr + uzyr) / 2.0, e33 = uzz;
    const double ekk = e11 + e22 + e33;
    const double G = M.G, lam = M.lambda;
    sig[0] = lam * ekk + 2.0 * G * e11;
    sig[1] = 2.0 * G * e12;
    sig[2] = 2.0 * G * e13;
    sig[3] = lam * ekk + 2.0 * G * e22;
    sig[4] = 2.0 * G * e23;
    sig[5] = lam * ekk + 2.0 * G * e33;
}

}  // namespace ucy
