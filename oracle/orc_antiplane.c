/*
 * orc_antiplane.c -- oracle restatement of the 2-D antiplane rate-and-state
 * path of jdpmoore/unicycle fortran/2d/antiplane/src.
 *
 * TEST INFRASTRUCTURE ONLY (see ucy_oracle.h).  Build with -ffp-contract=off.
 *
 * Follows antiplane.f90:81-109 (patch frame), :131-156 (stress of a vertical
 * screw dislocation in a half space; the dip angle is ignored, as the
 * reference documents), :352-379 (traction kernels), greens.f90:93-248
 * (buildG, patches only: one traction row and one slip column per patch),
 * ratestate.f90:568-605 (initStateVector) and :619-724 (odefun).  The main
 * loop (ratestate.f90:254-276) is the 3-D one, run with orc_rkqs.
 *
 * The viscouscycles program of the same directory adds strain volumes
 * (antiplane shear zones): antiplane.f90:43-63 (boxcar), :213-255
 * (displacement), :401-433 and :479-635 (stress and traction kernels),
 * greens.f90:93-248 (buildG with volume rows and columns),
 * viscouscycles.f90:697-772 (initStateVector) and :786-957 (odefun).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ucy_oracle.h"

static const double AP_PI = 3.141592653589793;   /* antiplane.f90:35 */
static const double LG10 = 2.302585092994045684; /* log(10) */

/* antiplane.f90:81-109 */
void orc_refsys_antiplane(int ns, const double *x, const double *W, const double *dip,
                          double *sv, double *dv, double *nv, double *xc)
{
    for (int i = 0; i < ns; i++) {
        sv[3 * i] = 1.0;
        sv[3 * i + 1] = 0.0;
        sv[3 * i + 2] = 0.0;
        dv[3 * i] = +0.0;
        dv[3 * i + 1] = -cos(dip[i]);
        dv[3 * i + 2] = -sin(dip[i]);
        nv[3 * i] = -0.0;
        nv[3 * i + 1] = +sin(dip[i]);
        nv[3 * i + 2] = -cos(dip[i]);
        for (int k = 0; k < 3; k++) xc[3 * i + k] = x[3 * i + k] - W[i] / 2 * dv[3 * i + k];
    }
}

/* antiplane.f90:131-156: s12, s13 of unit slip on the vertical segment from
 * depth q3 to q3 + W at horizontal position q2, with its image */
void orc_stress_antiplane(double x2, double x3, double q2, double q3, double W, double dip,
                          double G, double *s12, double *s13)
{
    (void)dip;
    const double r2 = x2 - q2;
    const double a = x3 - q3, b = x3 + q3, c = x3 - q3 - W, d = x3 + q3 + W;
    *s12 = G * (-a / (r2 * r2 + a * a) + b / (r2 * r2 + b * b) + c / (r2 * r2 + c * c)
                - d / (r2 * r2 + d * d)) / 2.0 / AP_PI;
    *s13 = G * (r2 / (r2 * r2 + a * a) - r2 / (r2 * r2 + b * b) - r2 / (r2 * r2 + c * c)
                + r2 / (r2 * r2 + d * d)) / 2.0 / AP_PI;
}

/* antiplane.f90:178-191 (the displacement the stress derives from) */
double orc_displacement_antiplane(double x2, double x3, double q2, double q3, double W,
                                  double dip, double s)
{
    (void)dip;
    return s * (+atan((x3 - q3) / (x2 - q2)) - atan((x3 + q3) / (x2 - q2))
                - atan((x3 - q3 - W) / (x2 - q2)) + atan((x3 + q3 + W) / (x2 - q2))) / 2.0 / AP_PI;
}

/* antiplane.f90:352-379 */
void orc_traction_kernels_antiplane(const double x[3], const double sv[3], const double dv[3],
                                    const double nv[3], int ns, const double *y,
                                    const double *W, const double *dip, double G, double *ts)
{
    (void)dv;
    for (int i = 0; i < ns; i++) {
        double s12, s13;
        orc_stress_antiplane(x[1], x[2], y[3 * i + 1], y[3 * i + 2], W[i], dip[i], G, &s12, &s13);
        ts[i] = (nv[1] * s12 + nv[2] * s13) * sv[0] + (nv[0] * s12) * sv[1] + (nv[0] * s13) * sv[2];
    }
}

/* antiplane.f90:43-63 */
static double ap_heaviside(double x) { return 0.0 < x ? 1.0 : 0.0; }
static double ap_omega(double x) { return ap_heaviside(x + 0.5) - ap_heaviside(x - 0.5); }

/* antiplane.f90:479-515: stress of the primed eigenstrain of a volume of
 * thickness T (horizontal) and width W (depth) from q3 down, at horizontal
 * position q2, with its image; the eigenstrain is removed inside */
void orc_stress_sv_antiplane(double x2, double x3, double q2, double q3, double T, double W,
                             double dip, double e12p, double e13p, double G, double *s12,
                             double *s13)
{
    (void)dip;
    const double r2 = x2 - q2, hm = r2 - T / 2, hp = r2 + T / 2;
    const double a = x3 - q3, c = x3 - q3 - W, b = x3 + q3, d = x3 + q3 + W;
    const double in = ap_omega(r2 / T) * ap_omega((x3 - (2 * q3 + W) / 2.0) / W);
    *s12 = G / AP_PI * e12p * (atan(a / hp) - atan(a / hm) + atan(c / hm) - atan(c / hp)
                               - atan(d / hm) - atan(b / hp) + atan(b / hm) + atan(d / hp))
           + G / (2 * AP_PI) * e13p * (log(hm * hm + c * c) - log(hp * hp + c * c)
                                      + log(hm * hm + d * d) - log(hp * hp + d * d)
                                      - log(hm * hm + a * a) + log(hp * hp + a * a)
                                      - log(hm * hm + b * b) + log(hp * hp + b * b))
           - 2 * G * e12p * in;
    *s13 = G / (2 * AP_PI) * e12p * (log(hm * hm + c * c) - log(hp * hp + c * c)
                                    - log(hm * hm + d * d) + log(hp * hp + d * d)
                                    - log(hm * hm + a * a) + log(hp * hp + a * a)
                                    + log(hm * hm + b * b) - log(hp * hp + b * b))
           + G / AP_PI * e13p * (atan(hp / a) - atan(hm / a) - atan(hp / c) + atan(hm / c)
                                 + atan(hp / b) - atan(hm / b) - atan(hp / d) + atan(hm / d))
           - 2 * G * e13p * in;
}

/* antiplane.f90:213-255 */
double orc_displacement_sv_antiplane(double x2, double x3, double q2, double q3, double T,
                                     double W, double dip, double e12p, double e13p)
{
    (void)dip;
    const double r2 = x2 - q2, hm = r2 - T / 2, hp = r2 + T / 2;
    const double a = x3 - q3, c = x3 - q3 - W, b = x3 + q3, d = x3 + q3 + W;
    const double u12 = c * log(hm * hm + c * c) - c * log(hp * hp + c * c)
                       - a * log(hm * hm + a * a) + a * log(hp * hp + a * a)
                       + 2 * hm * (atan(c / hm) - atan(a / hm))
                       + 2 * hp * (atan(a / hp) - atan(c / hp))
                       + d * log(hp * hp + d * d) - d * log(hm * hm + d * d)
                       - b * log(hp * hp + b * b) + b * log(hm * hm + b * b)
                       + 2 * hp * (atan(d / hp) - atan(b / hp))
                       + 2 * hm * (atan(b / hm) - atan(d / hm));
    /* the multipliers x3 - W - q3 and x3 + W + q3 keep the reference's
     * operand order (antiplane.f90:245, 251) */
    const double cw = x3 - W - q3, dw = x3 + W + q3;
    const double u13 = hm * log(hm * hm + c * c) - hp * log(hp * hp + c * c)
                       - hm * log(hm * hm + a * a) + hp * log(hp * hp + a * a)
                       + 2 * cw * (atan(hm / c) - atan(hp / c))
                       + 2 * a * (atan(hp / a) - atan(hm / a))
                       + hm * log(hm * hm + d * d) - hp * log(hp * hp + d * d)
                       - hm * log(hm * hm + b * b) + hp * log(hp * hp + b * b)
                       + 2 * dw * (atan(hm / d) - atan(hp / d))
                       + 2 * b * (atan(hp / b) - atan(hm / b));
    return e12p / (2 * AP_PI) * u12 + e13p / (2 * AP_PI) * u13;
}

/* the projections of antiplane.f90:371-373 (traction on a patch) and
 * :418-425, :619-626 (stress in a volume's frame) */
static double ap_traction(const double sv[3], const double nv[3], double s12, double s13)
{
    return (nv[1] * s12 + nv[2] * s13) * sv[0] + (nv[0] * s12) * sv[1] + (nv[0] * s13) * sv[2];
}
static void ap_frame_stress(const double sv[3], const double dv[3], const double nv[3],
                            double s12, double s13, double *s12p, double *s13p)
{
    *s12p = (sv[1] * s12 + sv[2] * s13) * nv[0] + (sv[0] * s12) * nv[1] + (sv[0] * s13) * nv[2];
    *s13p = -(sv[1] * s12 + sv[2] * s13) * dv[0] - (sv[0] * s12) * dv[1] - (sv[0] * s13) * dv[2];
}

void orc_buildG_antiplane_volumes(int np, const double *px, const double *pW,
                                  const double *pdip, int nv, const double *vx,
                                  const double *vT, const double *vW, const double *vdip,
                                  double mu, double *K, int nthreads)
{
    const int ne = np + nv, n = np + 2 * nv;
    double *f = malloc(sizeof(double) * 12 * (size_t)(ne + 1));
    double *sv = f, *dv = f + 3 * (ne + 1), *nvv = f + 6 * (ne + 1), *xc = f + 9 * (ne + 1);
    orc_refsys_antiplane(np, px, pW, pdip, sv, dv, nvv, xc);
    orc_refsys_antiplane(nv, vx, vW, vdip, sv + 3 * np, dv + 3 * np, nvv + 3 * np, xc + 3 * np);
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int r = 0; r < ne; r++) {
        const double *x = xc + 3 * r, *s = sv + 3 * r, *d = dv + 3 * r, *nn = nvv + 3 * r;
        const int nrow = r < np ? 1 : 2;
        double *row0 = K + (size_t)(r < np ? r : np + 2 * (r - np)) * n;
        for (int c = 0; c < n; c++) {
            double s12, s13;
            if (c < np)  /* unit slip of patch c */
                orc_stress_antiplane(x[1], x[2], px[3 * c + 1], px[3 * c + 2], pW[c], pdip[c], mu,
                                     &s12, &s13);
            else {       /* unit e12' (j = 0) or e13' (j = 1) of volume (c - np) / 2 */
                const int q = (c - np) / 2, j = (c - np) % 2;
                orc_stress_sv_antiplane(x[1], x[2], vx[3 * q + 1], vx[3 * q + 2], vT[q], vW[q],
                                        vdip[q], j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, mu,
                                        &s12, &s13);
            }
            if (nrow == 1) {
                row0[c] = ap_traction(s, nn, s12, s13);
            } else {
                double a, b;
                ap_frame_stress(s, d, nn, s12, s13, &a, &b);
                row0[c] = a;
                row0[n + c] = b;
            }
        }
    }
    free(f);
}

/* buildG (greens.f90:93-248) for patches: K (ns x ns) row-major, row =
 * receiver (its traction), column = source (its slip) */
void orc_buildG_antiplane(int ns, const double *x, const double *W, const double *dip, double mu,
                          double *K, int nthreads)
{
    double *sv = malloc(sizeof(double) * 12 * (size_t)ns);
    double *dv = sv + 3 * ns, *nv = dv + 3 * ns, *xc = nv + 3 * ns;
    orc_refsys_antiplane(ns, x, W, dip, sv, dv, nv, xc);
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int r = 0; r < ns; r++)
        orc_traction_kernels_antiplane(xc + 3 * r, sv + 3 * r, dv + 3 * r, nv + 3 * r, ns, x, W,
                                       dip, mu, K + (size_t)r * ns);
    free(sv);
}

/* buildO (greens.f90:275-383) for patches: O (npts x ns) row-major, unit
 * slip of each segment at each point x(3,npts) (x2, x3 used) */
void orc_buildO_antiplane(int ns, const double *x, const double *W, const double *dip, int npts,
                          const double *pts, double *O)
{
    for (int k = 0; k < npts; k++)
        for (int s = 0; s < ns; s++)
            O[(size_t)k * ns + s] = orc_displacement_antiplane(pts[3 * k + 1], pts[3 * k + 2],
                                                               x[3 * s + 1], x[3 * s + 2], W[s],
                                                               dip[s], 1.0);
}

/* initStateVector (ratestate.f90:568-605): slip, traction, log10 theta,
 * log10 V = log10(0.98 Vpl) */
void orc_ap_init_state(const orc_ap_system *s, double *y)
{
    for (int i = 0; i < s->n; i++) {
        double *p = y + 4 * i;
        p[0] = 0.0;
        p[1] = s->sig[i] * (s->mu0[i] + (s->a[i] - s->b[i]) * log(s->Vpl[i] / s->Vo[i]))
               - s->damping[i] * s->Vpl[i];
        p[2] = log(s->L[i] / s->Vpl[i]) / LG10;
        p[3] = log(s->Vpl[i] * 0.98) / LG10;
    }
    /* volumes, viscouscycles.f90:740-762: the stress that flows at the
     * background rate, no strain */
    for (int i = 0; i < s->nvol; i++) {
        double *q = y + 4 * s->n + 4 * i;
        const double eps = sqrt(s->e12[i] * s->e12[i] + s->e13[i] * s->e13[i]);
        if (0.0 < eps && 0.0 < s->gammadot0[i]) {
            const double tau = s->Gm[i] * pow(eps / s->gammadot0[i] * exp(s->Q[i] / (s->R[i] * s->To[i])),
                                              1.0 / s->npower[i]);
            q[2] = tau * s->e12[i] / eps;
            q[3] = tau * s->e13[i] / eps;
        } else {
            q[2] = 0.0;
            q[3] = 0.0;
        }
        q[0] = 0.0;
        q[1] = 0.0;
    }
}

/* odefun (ratestate.f90:619-724); returns vMax.  vwork, twork: n each */
double orc_ap_odefun(const orc_ap_system *s, const double *y, double *dydt, double *vwork,
                     double *twork)
{
    const int n = s->n, nc = s->n + 2 * s->nvol;
    double vmax = 0.0;
    memset(dydt, 0, sizeof(double) * 4 * (size_t)(n + s->nvol));
    for (int i = 0; i < n; i++) {
        const double velocity = exp(y[4 * i + 3] * LG10);
        vmax = velocity > vmax ? velocity : vmax;
        dydt[4 * i] = velocity;
        vwork[i] = velocity - s->Vpl[i];
    }
    /* volumes: viscouscycles.f90:846-877 */
    for (int i = 0; i < s->nvol; i++) {
        const double *q = y + 4 * n + 4 * i;
        double *dq = dydt + 4 * n + 4 * i;
        const double s12 = q[2], s13 = q[3];
        const double sII = sqrt(s12 * s12 + s13 * s13);
        const double eII = s->gammadot0[i] * pow(sII / s->Gm[i], s->npower[i] - 1) / s->Gm[i]
                           * exp(-s->Q[i] / (s->R[i] * s->To[i]));
        const double e12 = eII * s12, e13 = eII * s13;
        dq[0] = e12;
        dq[1] = e13;
        vwork[n + 2 * i] = e12 - s->e12[i];
        vwork[n + 2 * i + 1] = e13 - s->e13[i];
    }
    orc_matvec(nc, nc, s->K, vwork, twork, 1);
    /* volumes: viscouscycles.f90:939-942, the stress rates */
    for (int i = 0; i < s->nvol; i++) {
        dydt[4 * n + 4 * i + 2] = twork[n + 2 * i];
        dydt[4 * n + 4 * i + 3] = twork[n + 2 * i + 1];
    }
    for (int i = 0; i < n; i++) {
        const double *p = y + 4 * i;
        double *dp = dydt + 4 * i;
        const double velocity = exp(p[3] * LG10);
        dp[2] = (exp(-p[2] * LG10) - velocity / s->L[i]) / LG10;
        const double dtau = twork[i];
        dp[3] = (dtau - s->b[i] * s->sig[i] * dp[2] * LG10)
                / (s->a[i] * s->sig[i] + s->damping[i] * velocity) / LG10;
        dp[1] = dtau - s->damping[i] * velocity * dp[3] * LG10;
    }
    return vmax;
}

typedef struct {
    const orc_ap_system *s;
    double *v, *t, vmax;
    long evals;
} ap_ctx;

static void ap_rhs(void *ctx, const double *y, double *dydt)
{
    ap_ctx *c = (ap_ctx *)ctx;
    c->vmax = orc_ap_odefun(c->s, y, dydt, c->v, c->t);
    c->evals++;
}

/* the main loop (ratestate.f90:254-276), as orc_integrate for the 3-D path */
long orc_ap_integrate(const orc_ap_system *s, double *y, double interval, double dt0,
                      long max_steps, double epsilon, double max_step, double *t_log,
                      double *dt_log, double *vmax_log, long log_cap, orc_run_stats *st)
{
    const int dim = 4 * (s->n + s->nvol), nc = s->n + 2 * s->nvol;
    double *work = malloc(sizeof(double) * (size_t)(11 * dim + 2 * nc));
    double *dydt = work, *yscal = work + dim, *rk = work + 2 * dim;
    ap_ctx c = {s, work + 11 * dim, work + 11 * dim + nc, 0.0, 0};
    double time = 0.0, dt_next = dt0;
    long steps = 0, attempts = 0;
    while (steps < max_steps) {
        ap_rhs(&c, y, dydt);
        const double vmax = c.vmax;
        const double dt_try = dt_next;
        for (int i = 0; i < dim; i++) yscal[i] = fabs(y[i]) + fabs(dt_try * dydt[i]) + 1e-30;
        double t0 = 0.0, hdid, hnext;
        int na = 0;
        if (orc_rkqs(dim, &t0, y, dydt, yscal, dt_try, &hdid, &hnext, epsilon, max_step, rk,
                     ap_rhs, &c, &na) != 0) {
            free(work);
            return -1;
        }
        attempts += na;
        time = time + hdid;
        dt_next = hnext;
        if (steps < log_cap) {
            t_log[steps] = time;
            dt_log[steps] = hdid;
            vmax_log[steps] = vmax;
        }
        steps++;
        if (interval <= time) break;
    }
    if (st) {
        st->time = time;
        st->dt_next = dt_next;
        st->steps = steps;
        st->attempts = attempts;
        st->rhs_evals = c.evals;
    }
    free(work);
    return steps;
}
