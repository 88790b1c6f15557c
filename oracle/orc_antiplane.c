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
}

/* odefun (ratestate.f90:619-724); returns vMax.  vwork, twork: n each */
double orc_ap_odefun(const orc_ap_system *s, const double *y, double *dydt, double *vwork,
                     double *twork)
{
    const int n = s->n;
    double vmax = 0.0;
    memset(dydt, 0, sizeof(double) * 4 * (size_t)n);
    for (int i = 0; i < n; i++) {
        const double velocity = exp(y[4 * i + 3] * LG10);
        vmax = velocity > vmax ? velocity : vmax;
        dydt[4 * i] = velocity;
        vwork[i] = velocity - s->Vpl[i];
    }
    orc_matvec(n, n, s->K, vwork, twork, 1);
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
    const int dim = 4 * s->n;
    double *work = malloc(sizeof(double) * (size_t)(11 * dim + 2 * s->n));
    double *dydt = work, *yscal = work + dim, *rk = work + 2 * dim;
    ap_ctx c = {s, work + 11 * dim, work + 11 * dim + s->n, 0.0, 0};
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
