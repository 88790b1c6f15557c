/*
 * orc_okada.c -- oracle restatement of the Okada (1992) rectangle chain.
 *
 * TEST INFRASTRUCTURE ONLY (see ucy_oracle.h).  Build with -ffp-contract=off.
 *
 * Follows jdpmoore/unicycle fortran/3d/src/okada92.f90 and dc3d.f.  The
 * rounding points of the reference are reproduced:
 *   - computeStressOkada92 rounds x, y, z, depth, dip[deg], L, -W, slip and
 *     alpha to binary32 (okada92.f90:138-160);
 *   - DC3D forms XI = X-AL and D = DEPTH+/-Z in binary32 arithmetic
 *     (dc3d.f:53-54,60,113: both operands are REAL*4), compares against the
 *     binary32 EPS of dc3d.f:5-6,36, works in binary64 inside, and returns
 *     its 12 outputs rounded to binary32 (dc3d.f:168-179);
 *   - DCCON2 snaps XI, ET, Q in place with the binary64 EPS (dc3d.f:612-616),
 *     so a snap persists into later corners and into the image pass.
 * Okada's published expressions (Okada 1992, BSSA 82, Tables 6-8) are written
 * out below; the grouping of every product/quotient follows the reference's
 * left-to-right evaluation so that results agree to the last bit.
 */
#include <math.h>
#include <string.h>
#include "ucy_oracle.h"

/* dc3d.f:5-6,36: EPS is declared REAL*4 in DC3D */
static const float DC_EPS_F = 1.0e-6f;
/* dc3d.f:523,612: EPS is REAL*8 in DCCON0 / DCCON2 */
static const double DC_EPS_D = 1.0e-6;
static const double PI2 = 6.283185307179586;

/* medium and dip constants, DCCON0 (dc3d.f:509-545) */
typedef struct { double alp1, alp2, alp3, alp4, alp5, sd, cd, sdsd, cdcd, sdcd; } dc_med;

static void dc_med_init(dc_med *m, double alpha, double dipdeg)
{
    m->alp1 = (1.0 - alpha) / 2.0;
    m->alp2 = alpha / 2.0;
    m->alp3 = (1.0 - alpha) / alpha;
    m->alp4 = 1.0 - alpha;
    m->alp5 = alpha;
    double p18 = PI2 / 360.0;
    m->sd = sin(dipdeg * p18);
    m->cd = cos(dipdeg * p18);
    if (fabs(m->cd) < DC_EPS_D) {
        m->cd = 0.0;
        if (m->sd > 0.0) m->sd = 1.0;
        if (m->sd < 0.0) m->sd = -1.0;
    }
    m->sdsd = m->sd * m->sd;
    m->cdcd = m->cd * m->cd;
    m->sdcd = m->sd * m->cd;
}

/* station geometry for one corner, DCCON2 (dc3d.f:596-665) */
typedef struct {
    double xi2, et2, q2, r, r2, r3, r5, y, d, tt, alx, ale, x11, y11, x32, y32;
    double ey, ez, fy, fz, gy, gz, hy, hz;
} dc_geo;

static void dc_geo_init(dc_geo *g, double *xi, double *et, double *q, double sd, double cd,
                        int kxi, int ket)
{
    if (fabs(*xi) < DC_EPS_D) *xi = 0.0;
    if (fabs(*et) < DC_EPS_D) *et = 0.0;
    if (fabs(*q) < DC_EPS_D) *q = 0.0;
    double XI = *xi, ET = *et, Q = *q;
    g->xi2 = XI * XI;
    g->et2 = ET * ET;
    g->q2 = Q * Q;
    g->r2 = g->xi2 + g->et2 + g->q2;
    g->r = sqrt(g->r2);
    if (g->r == 0.0) return; /* reference leaves the previous corner's values */
    g->r3 = g->r * g->r2;
    g->r5 = g->r3 * g->r2;
    g->y = ET * cd + Q * sd;
    g->d = ET * sd - Q * cd;
    g->tt = (Q == 0.0) ? 0.0 : atan(XI * ET / (Q * g->r));
    if (kxi == 1) {
        g->alx = -log(g->r - XI);
        g->x11 = 0.0;
        g->x32 = 0.0;
    } else {
        double rxi = g->r + XI;
        g->alx = log(rxi);
        g->x11 = 1.0 / (g->r * rxi);
        g->x32 = (g->r + rxi) * g->x11 * g->x11 / g->r;
    }
    if (ket == 1) {
        g->ale = -log(g->r - ET);
        g->y11 = 0.0;
        g->y32 = 0.0;
    } else {
        double ret = g->r + ET;
        g->ale = log(ret);
        g->y11 = 1.0 / (g->r * ret);
        g->y32 = (g->r + ret) * g->y11 * g->y11 / g->r;
    }
    g->ey = sd / g->r - g->y * Q / g->r3;
    g->ez = cd / g->r + g->d * Q / g->r3;
    g->fy = g->d / g->r3 + g->xi2 * g->y32 * sd;
    g->fz = g->y / g->r3 + g->xi2 * g->y32 * cd;
    g->gy = 2.0 * g->x11 * sd - g->y * Q * g->x32;
    g->gz = 2.0 * g->x11 * cd + g->d * Q * g->x32;
    g->hy = g->d * Q * g->x32 + XI * Q * g->y32 * sd;
    g->hz = g->y * Q * g->x32 + XI * Q * g->y32 * cd;
}

static void add_scaled(double u[12], double disl, const double du[12])
{
    double c = disl / PI2;
    for (int i = 0; i < 12; i++) u[i] = u[i] + c * du[i];
}

/* part A, UA (dc3d.f:200-283) */
static void dc_ua(const dc_med *m, const dc_geo *g, double xi, double et, double q,
                  double d1, double d2, double d3, double u[12])
{
    double du[12];
    double xy = xi * g->y11, qx = q * g->x11, qy = q * g->y11;
    double sd = m->sd, cd = m->cd, a1 = m->alp1, a2 = m->alp2;
    double r = g->r, r3 = g->r3, y = g->y, d = g->d;
    for (int i = 0; i < 12; i++) u[i] = 0.0;
    if (d1 != 0.0) {
        du[0] = g->tt / 2.0 + a2 * xi * qy;
        du[1] = a2 * q / r;
        du[2] = a1 * g->ale - a2 * q * qy;
        du[3] = -a1 * qy - a2 * g->xi2 * q * g->y32;
        du[4] = -a2 * xi * q / r3;
        du[5] = a1 * xy + a2 * xi * g->q2 * g->y32;
        du[6] = a1 * xy * sd + a2 * xi * g->fy + d / 2.0 * g->x11;
        du[7] = a2 * g->ey;
        du[8] = a1 * (cd / r + qy * sd) - a2 * q * g->fy;
        du[9] = a1 * xy * cd + a2 * xi * g->fz + y / 2.0 * g->x11;
        du[10] = a2 * g->ez;
        du[11] = -a1 * (sd / r - qy * cd) - a2 * q * g->fz;
        add_scaled(u, d1, du);
    }
    if (d2 != 0.0) {
        du[0] = a2 * q / r;
        du[1] = g->tt / 2.0 + a2 * et * qx;
        du[2] = a1 * g->alx - a2 * q * qx;
        du[3] = -a2 * xi * q / r3;
        du[4] = -qy / 2.0 - a2 * et * q / r3;
        du[5] = a1 / r + a2 * g->q2 / r3;
        du[6] = a2 * g->ey;
        du[7] = a1 * d * g->x11 + xy / 2.0 * sd + a2 * et * g->gy;
        du[8] = a1 * y * g->x11 - a2 * q * g->gy;
        du[9] = a2 * g->ez;
        du[10] = a1 * y * g->x11 + xy / 2.0 * cd + a2 * et * g->gz;
        du[11] = -a1 * d * g->x11 - a2 * q * g->gz;
        add_scaled(u, d2, du);
    }
    if (d3 != 0.0) {
        du[0] = -a1 * g->ale - a2 * q * qy;
        du[1] = -a1 * g->alx - a2 * q * qx;
        du[2] = g->tt / 2.0 - a2 * (et * qx + xi * qy);
        du[3] = -a1 * xy + a2 * xi * g->q2 * g->y32;
        du[4] = -a1 / r + a2 * g->q2 / r3;
        du[5] = -a1 * qy - a2 * q * g->q2 * g->y32;
        du[6] = -a1 * (cd / r + qy * sd) - a2 * q * g->fy;
        du[7] = -a1 * y * g->x11 - a2 * q * g->gy;
        du[8] = a1 * (d * g->x11 + xy * sd) + a2 * q * g->hy;
        du[9] = a1 * (sd / r - qy * cd) - a2 * q * g->fz;
        du[10] = a1 * d * g->x11 - a2 * q * g->gz;
        du[11] = a1 * (y * g->x11 + xy * cd) + a2 * q * g->hz;
        add_scaled(u, d3, du);
    }
}

/* part B, UB (dc3d.f:284-401) */
static void dc_ub(const dc_med *m, const dc_geo *g, double xi, double et, double q,
                  double d1, double d2, double d3, double u[12])
{
    double du[12];
    double sd = m->sd, cd = m->cd, a3 = m->alp3;
    double sdsd = m->sdsd, cdcd = m->cdcd, sdcd = m->sdcd;
    double r = g->r, r3 = g->r3, y = g->y, d = g->d;
    double rd = r + d;
    double d11 = 1.0 / (r * rd);
    double aj2 = xi * y / rd * d11;
    double aj5 = -(d + y * y / rd) * d11;
    double ai1, ai2, ai3, ai4, ak1, ak2, ak3, ak4, aj1, aj3, aj4, aj6;
    if (cd != 0.0) {
        if (xi == 0.0) {
            ai4 = 0.0;
        } else {
            double x = sqrt(g->xi2 + g->q2);
            ai4 = 1.0 / cdcd * (xi / rd * sdcd
                                + 2.0 * atan((et * (x + q * cd) + x * (r + x) * sd)
                                             / (xi * (r + x) * cd)));
        }
        ai3 = (y * cd / rd - g->ale + sd * log(rd)) / cdcd;
        ak1 = xi * (d11 - g->y11 * sd) / cd;
        ak3 = (q * g->y11 - y * d11) / cd;
        aj3 = (ak1 - aj2 * sd) / cd;
        aj6 = (ak3 - aj5 * sd) / cd;
    } else {
        double rd2 = rd * rd;
        ai3 = (et / rd + y * q / rd2 - g->ale) / 2.0;
        ai4 = xi * y / rd2 / 2.0;
        ak1 = xi * q / rd * d11;
        ak3 = sd / rd * (g->xi2 * d11 - 1.0);
        aj3 = -xi / rd2 * (g->q2 * d11 - 1.0 / 2.0);
        aj6 = -y / rd2 * (g->xi2 * d11 - 1.0 / 2.0);
    }
    double xy = xi * g->y11;
    ai1 = -xi / rd * cd - ai4 * sd;
    ai2 = log(rd) + ai3 * sd;
    ak2 = 1.0 / r + ak3 * sd;
    ak4 = xy * cd - ak1 * sd;
    aj1 = aj5 * cd - aj6 * sd;
    aj4 = -xy - aj2 * cd + aj3 * sd;
    for (int i = 0; i < 12; i++) u[i] = 0.0;
    double qx = q * g->x11, qy = q * g->y11;
    if (d1 != 0.0) {
        du[0] = -xi * qy - g->tt - a3 * ai1 * sd;
        du[1] = -q / r + a3 * y / rd * sd;
        du[2] = q * qy - a3 * ai2 * sd;
        du[3] = g->xi2 * q * g->y32 - a3 * aj1 * sd;
        du[4] = xi * q / r3 - a3 * aj2 * sd;
        du[5] = -xi * g->q2 * g->y32 - a3 * aj3 * sd;
        du[6] = -xi * g->fy - d * g->x11 + a3 * (xy + aj4) * sd;
        du[7] = -g->ey + a3 * (1.0 / r + aj5) * sd;
        du[8] = q * g->fy - a3 * (qy - aj6) * sd;
        du[9] = -xi * g->fz - y * g->x11 + a3 * ak1 * sd;
        du[10] = -g->ez + a3 * y * d11 * sd;
        du[11] = q * g->fz + a3 * ak2 * sd;
        add_scaled(u, d1, du);
    }
    if (d2 != 0.0) {
        du[0] = -q / r + a3 * ai3 * sdcd;
        du[1] = -et * qx - g->tt - a3 * xi / rd * sdcd;
        du[2] = q * qx + a3 * ai4 * sdcd;
        du[3] = xi * q / r3 + a3 * aj4 * sdcd;
        du[4] = et * q / r3 + qy + a3 * aj5 * sdcd;
        du[5] = -g->q2 / r3 + a3 * aj6 * sdcd;
        du[6] = -g->ey + a3 * aj1 * sdcd;
        du[7] = -et * g->gy - xy * sd + a3 * aj2 * sdcd;
        du[8] = q * g->gy + a3 * aj3 * sdcd;
        du[9] = -g->ez - a3 * ak3 * sdcd;
        du[10] = -et * g->gz - xy * cd - a3 * xi * d11 * sdcd;
        du[11] = q * g->gz - a3 * ak4 * sdcd;
        add_scaled(u, d2, du);
    }
    if (d3 != 0.0) {
        du[0] = q * qy - a3 * ai3 * sdsd;
        du[1] = q * qx + a3 * xi / rd * sdsd;
        du[2] = et * qx + xi * qy - g->tt - a3 * ai4 * sdsd;
        du[3] = -xi * g->q2 * g->y32 - a3 * aj4 * sdsd;
        du[4] = -g->q2 / r3 - a3 * aj5 * sdsd;
        du[5] = q * g->q2 * g->y32 - a3 * aj6 * sdsd;
        du[6] = q * g->fy - a3 * aj1 * sdsd;
        du[7] = q * g->gy - a3 * aj2 * sdsd;
        du[8] = -q * g->hy - a3 * aj3 * sdsd;
        du[9] = q * g->fz + a3 * ak3 * sdsd;
        du[10] = q * g->gz + a3 * xi * d11 * sdsd;
        du[11] = -q * g->hz + a3 * ak4 * sdsd;
        add_scaled(u, d3, du);
    }
}

/* part C, UC (dc3d.f:402-508) */
static void dc_uc(const dc_med *m, const dc_geo *g, double xi, double et, double q,
                  double z, double d1, double d2, double d3, double u[12])
{
    double du[12];
    double sd = m->sd, cd = m->cd, a4 = m->alp4, a5 = m->alp5;
    double sdsd = m->sdsd, cdcd = m->cdcd, sdcd = m->sdcd;
    double r = g->r, r2 = g->r2, r3 = g->r3, r5 = g->r5, y = g->y, d = g->d;
    double xi2 = g->xi2, et2 = g->et2, q2 = g->q2, x11 = g->x11, y11 = g->y11;
    double x32 = g->x32, y32 = g->y32;
    double c = d + z;
    double x53 = (8.0 * r2 + 9.0 * r * xi + 3.0 * xi2) * x11 * x11 * x11 / r2;
    double y53 = (8.0 * r2 + 9.0 * r * et + 3.0 * et2) * y11 * y11 * y11 / r2;
    double h = q * cd - z;
    double z32 = sd / r3 - h * y32;
    double z53 = 3.0 * sd / r5 - h * y53;
    double y0 = y11 - xi2 * y32;
    double z0 = z32 - xi2 * z53;
    double ppy = cd / r3 + q * y32 * sd;
    double ppz = sd / r3 - q * y32 * cd;
    double qq = z * y32 + z32 + z0;
    double qqy = 3.0 * c * d / r5 - qq * sd;
    double qqz = 3.0 * c * y / r5 - qq * cd + q * y32;
    double xy = xi * y11, qx = q * x11, qy = q * y11;
    double qr = 3.0 * q / r5;
    double cqx = c * q * x53;
    double cdr = (c + d) / r3;
    double yy0 = y / r3 - y0 * cd;
    (void)qx; (void)cqx;
    for (int i = 0; i < 12; i++) u[i] = 0.0;
    if (d1 != 0.0) {
        du[0] = a4 * xy * cd - a5 * xi * q * z32;
        du[1] = a4 * (cd / r + 2.0 * qy * sd) - a5 * c * q / r3;
        du[2] = a4 * qy * cd - a5 * (c * et / r3 - z * y11 + xi2 * z32);
        du[3] = a4 * y0 * cd - a5 * q * z0;
        du[4] = -a4 * xi * (cd / r3 + 2.0 * q * y32 * sd) + a5 * c * xi * qr;
        du[5] = -a4 * xi * q * y32 * cd + a5 * xi * (3.0 * c * et / r5 - qq);
        du[6] = -a4 * xi * ppy * cd - a5 * xi * qqy;
        du[7] = a4 * 2.0 * (d / r3 - y0 * sd) * sd - y / r3 * cd
                - a5 * (cdr * sd - et / r3 - c * y * qr);
        du[8] = -a4 * q / r3 + yy0 * sd + a5 * (cdr * cd + c * d * qr - (y0 * cd + q * z0) * sd);
        du[9] = a4 * xi * ppz * cd - a5 * xi * qqz;
        du[10] = a4 * 2.0 * (y / r3 - y0 * cd) * sd + d / r3 * cd - a5 * (cdr * cd + c * d * qr);
        du[11] = yy0 * cd - a5 * (cdr * sd - c * y * qr - y0 * sdsd + q * z0 * cd);
        add_scaled(u, d1, du);
    }
    if (d2 != 0.0) {
        du[0] = a4 * cd / r - qy * sd - a5 * c * q / r3;
        du[1] = a4 * y * x11 - a5 * c * et * q * x32;
        du[2] = -d * x11 - xy * sd - a5 * c * (x11 - q2 * x32);
        du[3] = -a4 * xi / r3 * cd + a5 * c * xi * qr + xi * q * y32 * sd;
        du[4] = -a4 * y / r3 + a5 * c * et * qr;
        du[5] = d / r3 - y0 * sd + a5 * c / r3 * (1.0 - 3.0 * q2 / r2);
        du[6] = -a4 * et / r3 + y0 * sdsd - a5 * (cdr * sd - c * y * qr);
        du[7] = a4 * (x11 - y * y * x32) - a5 * c * ((d + 2.0 * q * cd) * x32 - y * et * q * x53);
        du[8] = xi * ppy * sd + y * d * x32 + a5 * c * ((y + 2.0 * q * sd) * x32 - y * q2 * x53);
        du[9] = -q / r3 + y0 * sdcd - a5 * (cdr * cd + c * d * qr);
        du[10] = a4 * y * d * x32 - a5 * c * ((y - 2.0 * q * sd) * x32 + d * et * q * x53);
        du[11] = -xi * ppz * sd + x11 - d * d * x32 - a5 * c * ((d - 2.0 * q * cd) * x32 - d * q2 * x53);
        add_scaled(u, d2, du);
    }
    if (d3 != 0.0) {
        du[0] = -a4 * (sd / r + qy * cd) - a5 * (z * y11 - q2 * z32);
        du[1] = a4 * 2.0 * xy * sd + d * x11 - a5 * c * (x11 - q2 * x32);
        du[2] = a4 * (y * x11 + xy * cd) + a5 * q * (c * et * x32 + xi * z32);
        du[3] = a4 * xi / r3 * sd + xi * q * y32 * cd + a5 * xi * (3.0 * c * et / r5 - 2.0 * z32 - z0);
        du[4] = a4 * 2.0 * y0 * sd - d / r3 + a5 * c / r3 * (1.0 - 3.0 * q2 / r2);
        du[5] = -a4 * yy0 - a5 * (c * et * qr - q * z0);
        du[6] = a4 * (q / r3 + y0 * sdcd) + a5 * (z / r3 * cd + c * d * qr - q * z0 * sd);
        du[7] = -a4 * 2.0 * xi * ppy * sd - y * d * x32
                + a5 * c * ((y + 2.0 * q * sd) * x32 - y * q2 * x53);
        du[8] = -a4 * (xi * ppy * cd - x11 + y * y * x32)
                + a5 * (c * ((d + 2.0 * q * cd) * x32 - y * et * q * x53) + xi * qqy);
        du[9] = -et / r3 + y0 * cdcd - a5 * (z / r3 * sd - c * y * qr - y0 * sdsd + q * z0 * cd);
        du[10] = a4 * 2.0 * xi * ppz * sd - x11 + d * d * x32
                 - a5 * c * ((d - 2.0 * q * cd) * x32 - d * q2 * x53);
        du[11] = a4 * (xi * ppz * cd + y * d * x32)
                 + a5 * (c * ((y - 2.0 * q * sd) * x32 + d * et * q * x53) + xi * qqz);
        add_scaled(u, d3, du);
    }
}

/* DC3D (dc3d.f:1-199).  Returns iret (0 normal, 1 singular). */
int orc_dc3d(float alpha, float X, float Y, float Z, float depth, float dip,
             float al1, float al2, float aw1, float aw2,
             float disl1, float disl2, float disl3, float uout[12])
{
    double U[12], du[12], dua[12], dub[12], duc[12];
    double xi[2], et[2];
    int kxi[2], ket[2];
    dc_med m;
    dc_geo g;
    memset(&g, 0, sizeof(g));
    for (int i = 0; i < 12; i++) U[i] = 0.0;
    dc_med_init(&m, (double)alpha, (double)dip);
    double sd = m.sd, cd = m.cd;
    double zz = (double)Z;
    double dd1 = disl1, dd2 = disl2, dd3 = disl3;
    const double epsf = (double)DC_EPS_F;
    xi[0] = (double)(float)(X - al1);
    xi[1] = (double)(float)(X - al2);
    if (fabs(xi[0]) < epsf) xi[0] = 0.0;
    if (fabs(xi[1]) < epsf) xi[1] = 0.0;

    for (int pass = 0; pass < 2; pass++) {
        /* pass 0: real source (dc3d.f:60-109); pass 1: image (dc3d.f:113-166) */
        double D = (pass == 0) ? (double)(float)(depth + Z) : (double)(float)(depth - Z);
        double P = (double)Y * cd + D * sd;
        double Q = (double)Y * sd - D * cd;
        et[0] = P - (double)aw1;
        et[1] = P - (double)aw2;
        if (fabs(Q) < epsf) Q = 0.0;
        if (fabs(et[0]) < epsf) et[0] = 0.0;
        if (fabs(et[1]) < epsf) et[1] = 0.0;
        if (Q == 0.0 && ((xi[0] * xi[1] <= 0.0 && et[0] * et[1] == 0.0) ||
                         (et[0] * et[1] <= 0.0 && xi[0] * xi[1] == 0.0))) {
            for (int i = 0; i < 12; i++) uout[i] = 0.0f;
            return 1;
        }
        kxi[0] = kxi[1] = ket[0] = ket[1] = 0;
        double r12 = sqrt(xi[0] * xi[0] + et[1] * et[1] + Q * Q);
        double r21 = sqrt(xi[1] * xi[1] + et[0] * et[0] + Q * Q);
        double r22 = sqrt(xi[1] * xi[1] + et[1] * et[1] + Q * Q);
        if (xi[0] < 0.0 && r21 + xi[1] < epsf) kxi[0] = 1;
        if (xi[0] < 0.0 && r22 + xi[1] < epsf) kxi[1] = 1;
        if (et[0] < 0.0 && r12 + et[1] < epsf) ket[0] = 1;
        if (et[0] < 0.0 && r22 + et[1] < epsf) ket[1] = 1;
        for (int k = 0; k < 2; k++) {
            for (int j = 0; j < 2; j++) {
                dc_geo_init(&g, &xi[j], &et[k], &Q, sd, cd, kxi[k], ket[j]);
                if (pass == 0) {
                    dc_ua(&m, &g, xi[j], et[k], Q, dd1, dd2, dd3, dua);
                    for (int i = 0; i < 12; i += 3) {
                        du[i] = -dua[i];
                        du[i + 1] = -dua[i + 1] * cd + dua[i + 2] * sd;
                        du[i + 2] = -dua[i + 1] * sd - dua[i + 2] * cd;
                        if (i == 9) {
                            du[i] = -du[i];
                            du[i + 1] = -du[i + 1];
                            du[i + 2] = -du[i + 2];
                        }
                    }
                } else {
                    dc_ua(&m, &g, xi[j], et[k], Q, dd1, dd2, dd3, dua);
                    dc_ub(&m, &g, xi[j], et[k], Q, dd1, dd2, dd3, dub);
                    dc_uc(&m, &g, xi[j], et[k], Q, zz, dd1, dd2, dd3, duc);
                    for (int i = 0; i < 12; i += 3) {
                        du[i] = dua[i] + dub[i] + (double)Z * duc[i];
                        du[i + 1] = (dua[i + 1] + dub[i + 1] + (double)Z * duc[i + 1]) * cd
                                    - (dua[i + 2] + dub[i + 2] + (double)Z * duc[i + 2]) * sd;
                        du[i + 2] = (dua[i + 1] + dub[i + 1] - (double)Z * duc[i + 1]) * sd
                                    + (dua[i + 2] + dub[i + 2] - (double)Z * duc[i + 2]) * cd;
                        if (i == 9) {
                            du[9] = du[9] + duc[0];
                            du[10] = du[10] + duc[1] * cd - duc[2] * sd;
                            du[11] = du[11] - duc[1] * sd - duc[2] * cd;
                        }
                    }
                }
                if (j + k != 1) { for (int i = 0; i < 12; i++) U[i] = U[i] + du[i]; }
                else            { for (int i = 0; i < 12; i++) U[i] = U[i] - du[i]; }
            }
        }
    }
    for (int i = 0; i < 12; i++) uout[i] = (float)U[i];
    return 0;
}

/* computeReferenceSystemOkada92 (okada92.f90:59-87) */
void orc_refsys_okada92(int ns, const double *x, const double *L, const double *W,
                        const double *strike, const double *dip,
                        double *sv, double *dv, double *nv, double *xc)
{
    for (int i = 0; i < ns; i++) {
        double cs = cos(strike[i]), ss = sin(strike[i]), cdp = cos(dip[i]), sdp = sin(dip[i]);
        sv[3 * i + 0] = cs;
        sv[3 * i + 1] = ss;
        sv[3 * i + 2] = 0.0;
        dv[3 * i + 0] = +ss * cdp;
        dv[3 * i + 1] = -cs * cdp;
        dv[3 * i + 2] = -sdp;
        nv[3 * i + 0] = -ss * sdp;
        nv[3 * i + 1] = +cs * sdp;
        nv[3 * i + 2] = -cdp;
        for (int k = 0; k < 3; k++)
            xc[3 * i + k] = x[3 * i + k] + L[i] / 2 * sv[3 * i + k] - W[i] / 2 * dv[3 * i + k];
    }
}

/* computeStressOkada92 (okada92.f90:109-207); sig = (s11,s12,s13,s22,s23,s33) */
void orc_stress_okada92(double x1, double x2, double x3, double q1, double q2, double q3,
                        double L, double W, double strike, double dipd,
                        double s, double d, double t, double G, double lambda,
                        double sig[6])
{
    const double RAD2DEG = 57.29577951308232;
    float alpha = (float)((lambda + G) / (lambda + 2.0 * G));
    float z = (float)(-x3);
    double csst = cos(strike), ssst = sin(strike);
    float x = (float)((x1 - q1) * csst + (x2 - q2) * ssst);
    float y = (float)((x1 - q1) * ssst - (x2 - q2) * csst);
    float depth = (float)q3;
    float dip = (float)(dipd * RAD2DEG);
    float al1 = 0.0f, al2 = (float)L, aw1 = -(float)W, aw2 = 0.0f;
    float disl1 = (float)s, disl2 = (float)d, disl3 = (float)t;
    float u[12];
    orc_dc3d(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, disl1, disl2, disl3, u);
    double uxx = u[3], uyx = u[4], uzx = u[5], uxy = u[6], uyy = u[7], uzy = u[8];
    double uxz = u[9], uyz = u[10], uzz = u[11];
    double uxxr = (csst * uxx + ssst * uyx) * csst - (-csst * uxy - ssst * uyy) * ssst;
    double uxyr = (csst * uxx + ssst * uyx) * ssst + (-csst * uxy - ssst * uyy) * csst;
    double uxzr = -csst * uxz + ssst * uyz;
    double uyxr = (ssst * uxx - csst * uyx) * csst - (-ssst * uxy + csst * uyy) * ssst;
    double uyyr = (ssst * uxx - csst * uyx) * ssst + (-ssst * uxy + csst * uyy) * csst;
    double uyzr = -ssst * uxz + csst * uyz;
    double uzxr = -csst * uzx - ssst * uzy;
    double uzyr = -ssst * uzx + csst * uzy;
    double e11 = uxxr, e12 = (uxyr + uyxr) / 2.0, e13 = (uxzr + uzxr) / 2.0;
    double e22 = uyyr, e23 = (uyzr + uzyr) / 2.0, e33 = uzz;
    double ekk = e11 + e22 + e33;
    sig[0] = lambda * ekk + 2.0 * G * e11;
    sig[1] = 2.0 * G * e12;
    sig[2] = 2.0 * G * e13;
    sig[3] = lambda * ekk + 2.0 * G * e22;
    sig[4] = 2.0 * G * e23;
    sig[5] = lambda * ekk + 2.0 * G * e33;
}

/* computeDisplacementOkada92 (okada92.f90:230-298): the same binary32
 * interface as the stress routine; DC3D's iret is honoured here (a receiver
 * on a fault edge gets zero displacement).  Returns iret. */
int orc_displacement_okada92(double x1, double x2, double x3, double q1, double q2, double q3,
                             double L, double W, double strike, double dipd,
                             double s, double d, double t, double G, double lambda,
                             double u[3])
{
    const double RAD2DEG = 57.29577951308232;
    float alpha = (float)((lambda + G) / (lambda + 2.0 * G));
    float z = (float)(-x3);
    double csst = cos(strike), ssst = sin(strike);
    float x = (float)((x1 - q1) * csst + (x2 - q2) * ssst);
    float y = (float)((x1 - q1) * ssst - (x2 - q2) * csst);
    float uu[12];
    int iret = orc_dc3d(alpha, x, y, z, (float)q3, (float)(dipd * RAD2DEG), 0.0f, (float)L,
                        -(float)W, 0.0f, (float)s, (float)d, (float)t, uu);
    if (iret == 0) {
        u[0] = (double)uu[0] * csst + (double)uu[1] * ssst;
        u[1] = (double)uu[0] * ssst - (double)uu[1] * csst;
        u[2] = -(double)uu[2];
    } else {
        u[0] = u[1] = u[2] = 0.0;
    }
    return iret;
}

/* traction (sigma.n) projected on (s,d,n); okada92.f90:390-398 */
static void project_traction(const double nv[3], const double sv[3], const double dv[3],
                             const double S[6], double *ts, double *td, double *tn)
{
    double s11 = S[0], s12 = S[1], s13 = S[2], s22 = S[3], s23 = S[4], s33 = S[5];
    double t1 = nv[0] * s11 + nv[1] * s12 + nv[2] * s13;
    double t2 = nv[0] * s12 + nv[1] * s22 + nv[2] * s23;
    double t3 = nv[0] * s13 + nv[1] * s23 + nv[2] * s33;
    *ts = t1 * sv[0] + t2 * sv[1] + t3 * sv[2];
    *td = t1 * dv[0] + t2 * dv[1] + t3 * dv[2];
    *tn = t1 * nv[0] + t2 * nv[1] + t3 * nv[2];
}

/* stress rotated into the receiver volume frame; okada92.f90:445-462 */
void orc_project_volume_frame(const double sv[3], const double dv[3], const double nv[3],
                              const double S[6], double out[6])
{
    double s11 = S[0], s12 = S[1], s13 = S[2], s22 = S[3], s23 = S[4], s33 = S[5];
    double a1 = sv[0] * s11 + sv[1] * s12 + sv[2] * s13;
    double a2 = sv[0] * s12 + sv[1] * s22 + sv[2] * s23;
    double a3 = sv[0] * s13 + sv[1] * s23 + sv[2] * s33;
    double b1 = nv[0] * s11 + nv[1] * s12 + nv[2] * s13;
    double b2 = nv[0] * s12 + nv[1] * s22 + nv[2] * s23;
    double b3 = nv[0] * s13 + nv[1] * s23 + nv[2] * s33;
    double c1 = -dv[0] * s11 - dv[1] * s12 - dv[2] * s13;
    double c2 = -dv[0] * s12 - dv[1] * s22 - dv[2] * s23;
    double c3 = -dv[0] * s13 - dv[1] * s23 - dv[2] * s33;
    out[0] = a1 * sv[0] + a2 * sv[1] + a3 * sv[2];
    out[1] = a1 * nv[0] + a2 * nv[1] + a3 * nv[2];
    out[2] = -a1 * dv[0] - a2 * dv[1] - a3 * dv[2];
    out[3] = b1 * nv[0] + b2 * nv[1] + b3 * nv[2];
    out[4] = -b1 * dv[0] - b2 * dv[1] - b3 * dv[2];
    out[5] = -c1 * dv[0] - c2 * dv[1] - c3 * dv[2];
}

/* computeTractionKernelsOkada92 (okada92.f90:368-402) */
void orc_traction_kernels_okada92(const double x[3], const double sv[3], const double dv[3],
                                  const double nv[3], int ns, const double *y,
                                  const double *L, const double *W, const double *strike,
                                  const double *dip, double s, double d, double t,
                                  double G, double lambda, double *ts, double *td, double *tn)
{
    for (int i = 0; i < ns; i++) {
        double S[6];
        orc_stress_okada92(x[0], x[1], x[2], y[3 * i], y[3 * i + 1], y[3 * i + 2], L[i], W[i],
                           strike[i], dip[i], s, d, t, G, lambda, S);
        project_traction(nv, sv, dv, S, &ts[i], &td[i], &tn[i]);
    }
}

/* computeStressKernelsOkada92 (okada92.f90:424-473) */
void orc_stress_kernels_okada92(const double x[3], const double sv[3], const double dv[3],
                                const double nv[3], int ns, const double *y,
                                const double *L, const double *W, const double *strike,
                                const double *dip, double s, double d, double t,
                                double G, double lambda, double *s11, double *s12,
                                double *s13, double *s22, double *s23, double *s33)
{
    for (int i = 0; i < ns; i++) {
        double S[6], R[6];
        orc_stress_okada92(x[0], x[1], x[2], y[3 * i], y[3 * i + 1], y[3 * i + 2], L[i], W[i],
                           strike[i], dip[i], s, d, t, G, lambda, S);
        orc_project_volume_frame(sv, dv, nv, S, R);
        s11[i] = R[0]; s12[i] = R[1]; s13[i] = R[2];
        s22[i] = R[3]; s23[i] = R[4]; s33[i] = R[5];
    }
}

/* exposed for the other oracle units */
void orc_project_traction(const double nv[3], const double sv[3], const double dv[3],
                          const double S[6], double *ts, double *td, double *tn)
{
    project_traction(nv, sv, dv, S, ts, td, tn);
}
