/*
 * orc_ode.c -- oracle restatement of kernel assembly, the rate-and-state
 * right-hand side and the Cash-Karp RK45 stepper of jdpmoore/unicycle.
 *
 * TEST INFRASTRUCTURE ONLY (see ucy_oracle.h).  Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ucy_oracle.h"

void orc_project_traction(const double nv[3], const double sv[3], const double dv[3],
                          const double S[6], double *ts, double *td, double *tn);
void orc_project_volume_frame(const double sv[3], const double dv[3], const double nv[3],
                              const double S[6], double out[6]);

/* ratestate.f90:130 lg10=LOG(1.d1) */
static const double LG10 = 2.302585092994045684;

/* ------------------------------------------------------------------------ */
/* buildG (greens.f90:73-275).  Receivers in element order: rectangles,     */
/* triangles, volumes (ratestate.f90:838-867).  Source columns per         */
/* receiver row: 2 per rectangle, 2 per triangle, 6 per volume             */
/* (greens.f90:214-216, 266-271).  Rows: 3 traction components per patch   */
/* receiver (tractionRows), 6 stress components per volume (stressRows).   */
/* ------------------------------------------------------------------------ */
typedef struct {
    double *sv, *dv, *nv, *xc;
} frames;

static int build_frames(const orc_mesh *M, frames *fr)
{
    int ne = M->nrect + M->ntri + M->nvol;
    fr->sv = malloc(sizeof(double) * 3 * (ne + 1));
    fr->dv = malloc(sizeof(double) * 3 * (ne + 1));
    fr->nv = malloc(sizeof(double) * 3 * (ne + 1));
    fr->xc = malloc(sizeof(double) * 3 * (ne + 1));
    if (!fr->sv || !fr->dv || !fr->nv || !fr->xc) return -1;
    int o = 0;
    if (M->nrect)
        orc_refsys_okada92(M->nrect, M->rx, M->rL, M->rW, M->rstrike, M->rdip,
                           fr->sv, fr->dv, fr->nv, fr->xc);
    o += M->nrect;
    if (M->ntri)
        orc_refsys_nikkhoo15(M->nve, M->vert, M->ntri, M->i1, M->i2, M->i3,
                             fr->sv + 3 * o, fr->dv + 3 * o, fr->nv + 3 * o, fr->xc + 3 * o);
    o += M->ntri;
    if (M->nvol)
        orc_refsys_strainvolume(M->nvol, M->vx, M->vL, M->vW, M->vstrike,
                                fr->sv + 3 * o, fr->dv + 3 * o, fr->nv + 3 * o, fr->xc + 3 * o);
    return 0;
}

static void free_frames(frames *fr)
{
    free(fr->sv); free(fr->dv); free(fr->nv); free(fr->xc);
}

int orc_buildG(const orc_mesh *M, double *K, int nthreads)
{
    const int np = M->nrect + M->ntri, ne = np + M->nvol;
    const int n = 2 * np + 6 * M->nvol;
    const double eye[6][6] = {{1, 0, 0, 0, 0, 0}, {0, 1, 0, 0, 0, 0}, {0, 0, 1, 0, 0, 0},
                              {0, 0, 0, 1, 0, 0}, {0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 1}};
    frames fr;
    if (build_frames(M, &fr)) return -1;
    int rc = 0;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1) reduction(|:rc)
    for (int e = 0; e < ne; e++) {
        const double *xc = fr.xc + 3 * e, *sv = fr.sv + 3 * e, *dv = fr.dv + 3 * e,
                     *nv = fr.nv + 3 * e;
        int isvol = e >= np;
        int row0 = isvol ? 3 * np + 6 * (e - np) : 3 * e;
        int ncomp = isvol ? 6 : 3;
        int nsmax = M->nrect > M->ntri ? M->nrect : M->ntri;
        if (M->nvol > nsmax) nsmax = M->nvol;
        double *buf = malloc(sizeof(double) * 6 * (nsmax + 1));
        double *c[6];
        for (int k = 0; k < 6; k++) c[k] = buf + k * (nsmax + 1);
        for (int j = 0; j < 2 && M->nrect; j++) {
            if (!isvol)
                orc_traction_kernels_okada92(xc, sv, dv, nv, M->nrect, M->rx, M->rL, M->rW,
                                             M->rstrike, M->rdip, eye[j][0], eye[j][1],
                                             eye[j][2], M->mu, M->lambda, c[0], c[1], c[2]);
            else
                orc_stress_kernels_okada92(xc, sv, dv, nv, M->nrect, M->rx, M->rL, M->rW,
                                           M->rstrike, M->rdip, eye[j][0], eye[j][1],
                                           eye[j][2], M->mu, M->lambda, c[0], c[1], c[2],
                                           c[3], c[4], c[5]);
            for (int k = 0; k < ncomp; k++)
                for (int s = 0; s < M->nrect; s++)
                    K[(size_t)(row0 + k) * n + 2 * s + j] = c[k][s];
        }
        for (int j = 0; j < 2 && M->ntri; j++) {
            if (!isvol)
                orc_traction_kernels_nikkhoo15(xc, sv, dv, nv, M->ntri, M->i1, M->i2, M->i3,
                                               M->nve, M->vert, eye[j][0], eye[j][1],
                                               eye[j][2], M->mu, M->lambda, c[0], c[1], c[2]);
            else
                orc_stress_kernels_nikkhoo15(xc, sv, dv, nv, M->ntri, M->i1, M->i2, M->i3,
                                             M->nve, M->vert, eye[j][0], eye[j][1],
                                             eye[j][2], M->mu, M->lambda, c[0], c[1], c[2],
                                             c[3], c[4], c[5]);
            for (int k = 0; k < ncomp; k++)
                for (int s = 0; s < M->ntri; s++)
                    K[(size_t)(row0 + k) * n + 2 * M->nrect + 2 * s + j] = c[k][s];
        }
        for (int j = 0; j < 6 && M->nvol; j++) {
            int r;
            if (!isvol)
                r = orc_traction_kernels_strainvolume(xc, sv, dv, nv, M->nvol, M->vx, M->vL,
                                                      M->vT, M->vW, M->vstrike, eye[j],
                                                      M->mu, M->lambda, c[0], c[1], c[2]);
            else
                r = orc_stress_kernels_strainvolume(xc, sv, dv, nv, M->nvol, M->vx, M->vL,
                                                    M->vT, M->vW, M->vstrike, eye[j], M->mu,
                                                    M->lambda, c[0], c[1], c[2], c[3], c[4],
                                                    c[5]);
            if (r) rc |= 1;
            for (int k = 0; k < ncomp; k++)
                for (int s = 0; s < M->nvol; s++)
                    K[(size_t)(row0 + k) * n + 2 * np + 6 * s + j] = c[k][s];
        }
        free(buf);
    }
    free_frames(&fr);
    return rc ? -2 : 0;
}

/* buildO (greens.f90:302-462).  O, Of, Ol row-major (3 nobs x n): row 3k+c is
 * NED component c at observation point k (x[3k..3k+2]), columns as in buildG.
 * Of holds the fault columns and, of the volume columns, only the third
 * component (greens.f90:445 writes the volume's third component into Of);
 * Ol holds the volume columns' first two components.  Of, Ol may be NULL.
 * Returns 0, or -2 when a strain-volume parameter check fails. */
int orc_buildO(const orc_mesh *M, int nobs, const double *x, double *O, double *Of, double *Ol,
               int nthreads)
{
    const int np = M->nrect + M->ntri;
    const int n = 2 * np + 6 * M->nvol;
    const double nu = M->lambda / (M->lambda + M->mu) / 2.0;
    const double eye[6][6] = {{1, 0, 0, 0, 0, 0}, {0, 1, 0, 0, 0, 0}, {0, 0, 1, 0, 0, 0},
                              {0, 0, 0, 1, 0, 0}, {0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 1}};
    int rc = 0;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) reduction(|:rc)
    for (int k = 0; k < nobs; k++) {
        const double *p = x + 3 * k;
        double *row[3], *rf[3], *rl[3];
        for (int c = 0; c < 3; c++) {
            row[c] = O + (size_t)(3 * k + c) * n;
            rf[c] = Of ? Of + (size_t)(3 * k + c) * n : NULL;
            rl[c] = Ol ? Ol + (size_t)(3 * k + c) * n : NULL;
            for (int j = 0; j < n; j++) {
                if (rf[c]) rf[c][j] = 0.0;
                if (rl[c]) rl[c][j] = 0.0;
            }
        }
        double u[3];
        for (int s = 0; s < M->nrect; s++)
            for (int j = 0; j < 2; j++) {
                const double *q = M->rx + 3 * s;
                orc_displacement_okada92(p[0], p[1], p[2], q[0], q[1], q[2], M->rL[s], M->rW[s],
                                         M->rstrike[s], M->rdip[s], eye[j][0], eye[j][1],
                                         eye[j][2], M->mu, M->lambda, u);
                for (int c = 0; c < 3; c++) {
                    row[c][2 * s + j] = u[c];
                    if (rf[c]) rf[c][2 * s + j] = u[c];
                }
            }
        for (int s = 0; s < M->ntri; s++)
            for (int j = 0; j < 2; j++) {
                orc_displacement_nikkhoo15(p[0], p[1], p[2], M->vert + 3 * (M->i1[s] - 1),
                                           M->vert + 3 * (M->i2[s] - 1),
                                           M->vert + 3 * (M->i3[s] - 1), eye[j][0], eye[j][1],
                                           eye[j][2], M->mu, M->lambda, u);
                const int col = 2 * M->nrect + 2 * s + j;
                for (int c = 0; c < 3; c++) {
                    row[c][col] = u[c];
                    if (rf[c]) rf[c][col] = u[c];
                }
            }
        for (int s = 0; s < M->nvol; s++)
            for (int j = 0; j < 6; j++) {
                const double *q = M->vx + 3 * s;
                if (orc_displacement_strainvolume(p[0], p[1], p[2], q[0], q[1], q[2], M->vL[s],
                                                  M->vT[s], M->vW[s], M->vstrike[s], eye[j],
                                                  M->mu, nu, u))
                    rc |= 1;
                const int col = 2 * np + 6 * s + j;
                for (int c = 0; c < 3; c++) row[c][col] = u[c];
                if (rl[0]) rl[0][col] = u[0];
                if (rl[1]) rl[1][col] = u[1];
                if (rf[2]) rf[2][col] = u[2];
            }
    }
    return rc ? -2 : 0;
}

/* t = K v, rows summed left to right (DGEMV("T") on G(n,m), ratestate.f90:719) */
void orc_matvec(int m, int n, const double *K, const double *v, double *t, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < m; i++) {
        const double *row = K + (size_t)i * n;
        double s = 0.0;
        for (int j = 0; j < n; j++) s = s + row[j] * v[j];
        t[i] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* initStateVector (ratestate.f90:548-610)                                  */
/* ------------------------------------------------------------------------ */
void orc_init_state(const orc_system *S, double *y)
{
    int dim = 7 * S->n + 12 * S->nvol;
    for (int i = 0; i < dim; i++) y[i] = 0.0;
    for (int i = 0; i < S->n; i++) {
        double *p = y + 7 * i;
        double base = S->sig[i] * (S->mu0[i] + (S->a[i] - S->b[i]) * log(S->Vpl[i] / S->Vo[i]));
        /* ratestate.f90:589,604 vs viscouscycles.f90:750,765 */
        double tau = S->ic_variant ? base + S->damping[i] * S->Vpl[i]
                                   : base - S->damping[i] * S->Vpl[i];
        p[0] = 0.0;
        p[1] = 0.0;
        p[2] = tau * cos(S->rake[i]);
        p[3] = tau * sin(S->rake[i]);
        p[4] = 0.0;
        p[5] = log(S->L[i] / S->Vpl[i]) / LG10;
        p[6] = log(S->Vpl[i] * (S->ic_variant ? 0.98 : 0.99)) / LG10;
    }
    /* volumes, viscouscycles.f90:769-803 */
    for (int i = 0; i < S->nvol; i++) {
        double *q = y + 7 * S->n + 12 * i;
        const double *e = S->edot_bg + 6 * i;
        double eps = sqrt((e[0] * e[0] + 2 * (e[1] * e[1]) + 2 * (e[2] * e[2]) + e[3] * e[3]
                           + 2 * (e[4] * e[4]) + e[5] * e[5]) / 2.0);
        if (0.0 < eps) {
            double tau = S->Gm[i] * pow(eps / S->gammadot0[i] * exp(S->Qa[i] / (S->Rgas[i] * S->To[i])),
                                        1.0 / S->npower[i]);
            for (int k = 0; k < 6; k++) q[6 + k] = tau * e[k] / eps;
        } else {
            for (int k = 0; k < 6; k++) q[6 + k] = 0.0;
        }
        for (int k = 0; k < 6; k++) q[k] = 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* odefun (ratestate.f90:624-778; volume branch viscouscycles.f90:933-1080) */
/* ------------------------------------------------------------------------ */
double orc_odefun(const orc_system *S, const double *y, double *dydt, double *v, double *t)
{
    const int N = S->n, dim = 7 * N + 12 * S->nvol;
    double vMax = 0.0;
    for (int i = 0; i < dim; i++) dydt[i] = 0.0;
    /* step 1/3 - slip velocity (ratestate.f90:667-707) */
    for (int i = 0; i < N; i++) {
        const double *p = y + 7 * i;
        double ts = p[2], td = p[3];
        double rake = atan2(td, ts);
        double velocity = exp(p[6] * LG10);
        double vs = velocity * cos(rake), vd = velocity * sin(rake);
        vMax = velocity > vMax ? velocity : vMax;
        dydt[7 * i + 0] = vs;
        dydt[7 * i + 1] = vd;
        v[2 * i + 0] = vs - S->Vpl[i] * cos(S->rake[i]);
        v[2 * i + 1] = vd - S->Vpl[i] * sin(S->rake[i]);
    }
    /* volumes: deviatoric power-law strain rate (viscouscycles.f90:940-1000) */
    for (int i = 0; i < S->nvol; i++) {
        const double *q = y + 7 * N + 12 * i;
        double *dq = dydt + 7 * N + 12 * i;
        double s11 = q[6], s12 = q[7], s13 = q[8], s22 = q[9], s23 = q[10], s33 = q[11];
        double p = (s11 + s22 + s33) / 3.0;
        double s11p = s11 - p, s22p = s22 - p, s33p = s33 - p;
        double sII = sqrt((s11p * s11p + 2.0 * s12 * s12 + 2.0 * s13 * s13 + s22p * s22p
                           + 2.0 * s23 * s23 + s33p * s33p) / 2.0);
        double gd = S->gammadot0[i] * pow(sII, S->npower[i] - 1.0);
        double e[6] = {0.0 + gd * s11p, 0.0 + gd * s12, 0.0 + gd * s13,
                       0.0 + gd * s22p, 0.0 + gd * s23, 0.0 + gd * s33p};
        for (int k = 0; k < 6; k++) {
            dq[k] = e[k];
            v[2 * N + 6 * i + k] = e[k] - S->edot_bg[6 * i + k];
        }
    }
    /* step 2/3 - traction and stress rates */
    orc_matvec(S->m, S->ncol, S->K, v, t, S->nthreads);
    /* step 3/3 - remaining state rates (ratestate.f90:731-776) */
    for (int i = 0; i < N; i++) {
        const double *p = y + 7 * i;
        double *d = dydt + 7 * i;
        const double *tt = t + 3 * i;
        double rake = atan2(p[3], p[2]);
        double velocity = exp(p[6] * LG10);
        d[5] = (exp(-p[5] * LG10) - velocity / S->L[i]) / LG10;
        double dtau = tt[0] * cos(rake) + tt[1] * sin(rake);
        d[6] = (dtau - S->b[i] * S->sig[i] * d[5] * LG10)
               / (S->a[i] * S->sig[i] + S->damping[i] * velocity) / LG10;
        double correction = S->damping[i] * velocity * d[6] * LG10;
        d[2] = tt[0] - correction * cos(rake);
        d[3] = tt[1] - correction * sin(rake);
        d[4] = tt[2];
    }
    /* volumes: stress rate = kernel output (viscouscycles.f90:1072-1080) */
    for (int i = 0; i < S->nvol; i++) {
        double *dq = dydt + 7 * N + 12 * i;
        const double *tt = t + 3 * N + 6 * i;
        for (int k = 0; k < 6; k++) dq[6 + k] = tt[k];
    }
    return vMax;
}

/* ------------------------------------------------------------------------ */
/* RKCKm (ode45.f90:116-202).  BLAS-1 semantics: dscal y=a*y, daxpy y=y+a*x */
/* ------------------------------------------------------------------------ */
static const double A2 = 0.2, A3 = 0.3, A4 = 0.6, A5 = 1.0, A6 = 0.875;
static const double B21 = 0.2, B31 = 3.0 / 40.0, B32 = 9.0 / 40.0;
static const double B41 = 0.3, B42 = -0.9, B43 = 1.2;
static const double B51 = -11.0 / 54.0, B52 = 2.5, B53 = -70.0 / 27.0, B54 = 35.0 / 27.0;
static const double B61 = 1631.0 / 55296.0, B62 = 175.0 / 512.0, B63 = 575.0 / 13824.0,
                    B64 = 44275.0 / 110592.0, B65 = 253.0 / 4096.0;
static const double C1 = 37.0 / 378.0, C3 = 250.0 / 621.0, C4 = 125.0 / 594.0,
                    C6 = 512.0 / 1771.0;

void orc_rkck(int n, const double *yin, const double *dydt, double h, double *yout,
              double *yerr, double *w, orc_rhs_fn f, void *ctx)
{
    const double DC1 = C1 - 2825.0 / 27648.0, DC3 = C3 - 18575.0 / 48384.0,
                 DC4 = C4 - 13525.0 / 55296.0, DC5 = -277.0 / 14336.0, DC6 = C6 - 0.25;
    double *ak2 = w, *ak3 = w + n, *ak4 = w + 2 * n, *ak5 = w + 3 * n, *ak6 = w + 4 * n;
    double *yr = w + 5 * n, *yt = w + 6 * n;
    (void)A2; (void)A3; (void)A4; (void)A5; (void)A6;
    double hb21 = h * B21;
    for (int i = 0; i < n; i++) yt[i] = yin[i] + hb21 * dydt[i];
    f(ctx, yt, ak2);
    for (int i = 0; i < n; i++) {
        yr[i] = B31 * dydt[i];
        yr[i] = yr[i] + B32 * ak2[i];
        yt[i] = yin[i] + h * yr[i];
    }
    f(ctx, yt, ak3);
    for (int i = 0; i < n; i++) {
        yr[i] = B41 * dydt[i];
        yr[i] = yr[i] + B42 * ak2[i];
        yr[i] = yr[i] + B43 * ak3[i];
        yt[i] = yin[i] + h * yr[i];
    }
    f(ctx, yt, ak4);
    for (int i = 0; i < n; i++) {
        yr[i] = B51 * dydt[i];
        yr[i] = yr[i] + B52 * ak2[i];
        yr[i] = yr[i] + B53 * ak3[i];
        yr[i] = yr[i] + B54 * ak4[i];
        yt[i] = yin[i] + h * yr[i];
    }
    f(ctx, yt, ak5);
    for (int i = 0; i < n; i++) {
        yr[i] = B61 * dydt[i];
        yr[i] = yr[i] + B62 * ak2[i];
        yr[i] = yr[i] + B63 * ak3[i];
        yr[i] = yr[i] + B64 * ak4[i];
        yr[i] = yr[i] + B65 * ak5[i];
        yt[i] = yin[i] + h * yr[i];
    }
    f(ctx, yt, ak6);
    for (int i = 0; i < n; i++) {
        yr[i] = C1 * dydt[i];
        yr[i] = yr[i] + C3 * ak3[i];
        yr[i] = yr[i] + C4 * ak4[i];
        yr[i] = yr[i] + C6 * ak6[i];
        yout[i] = yin[i] + h * yr[i];
        double e = DC1 * dydt[i];
        e = e + DC3 * ak3[i];
        e = e + DC4 * ak4[i];
        e = e + DC5 * ak5[i];
        e = e + DC6 * ak6[i];
        yerr[i] = h * e;
    }
}

/* RKQSm (ode45.f90:51-108) */
int orc_rkqs(int n, double *t, double *y, double *dydt, const double *yscal,
             double htry, double *hdid, double *hnext, double epsilon, double max_step,
             double *w, orc_rhs_fn f, void *ctx, int *n_attempts)
{
    const double SAFETY = 0.7, PGROW = -0.2, PSHRNK = -0.15, ERRCON = 1.89e-4;
    double *ycand = w + 7 * n, *yerr = w + 8 * n;
    double h = htry;
    int att = 0;
    for (;;) {
        att++;
        orc_rkck(n, y, dydt, h, ycand, yerr, w, f, ctx);
        double errmax = 0.0;
        for (int i = 0; i < n; i++) {
            double e = fabs(yerr[i] / yscal[i]);
            if (e > errmax) errmax = e;
        }
        errmax = errmax / epsilon;
        if (errmax > 1.0) {
            h = h * SAFETY * pow(errmax, PSHRNK);
            double tnew = *t + h;
            if (tnew == *t) {
                if (n_attempts) *n_attempts = att;
                return -1;
            }
        } else {
            if (errmax > ERRCON) *hnext = SAFETY * h * pow(errmax, PGROW);
            else *hnext = 10.0 * h;
            if (*hnext > max_step) *hnext = max_step;
            *hdid = h;
            *t = *t + h;
            memcpy(y, ycand, sizeof(double) * n);
            break;
        }
    }
    if (n_attempts) *n_attempts = att;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* main loop (ratestate.f90:307-328)                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_system *S;
    double *v, *t;
    long evals;
    double vmax;
} rhs_ctx;

static void rhs_adapter(void *p, const double *y, double *dydt)
{
    rhs_ctx *c = (rhs_ctx *)p;
    c->evals++;
    c->vmax = orc_odefun(c->S, y, dydt, c->v, c->t);
}

long orc_integrate(const orc_system *S, double *y, double interval, double dt0,
                   long max_steps, double epsilon, double max_step,
                   double *t_log, double *dt_log, double *vmax_log, long log_cap,
                   orc_run_stats *st)
{
    const int dim = 7 * S->n + 12 * S->nvol;
    double *dydt = malloc(sizeof(double) * dim), *yscal = malloc(sizeof(double) * dim);
    double *w = malloc(sizeof(double) * 9 * dim);
    rhs_ctx c = {S, malloc(sizeof(double) * S->ncol), malloc(sizeof(double) * S->m), 0, 0.0};
    double time = 0.0, dt_next = dt0, dt_done = 0.0;
    long i, attempts = 0, rc = 0;
    for (i = 1; i <= max_steps; i++) {
        rhs_adapter(&c, y, dydt);
        double vmax = c.vmax;
        double dt_try = dt_next;
        for (int k = 0; k < dim; k++) yscal[k] = fabs(y[k]) + fabs(dt_try * dydt[k]) + 1.0e-30;
        double t0 = 0.0;
        int att = 0;
        if (orc_rkqs(dim, &t0, y, dydt, yscal, dt_try, &dt_done, &dt_next, epsilon, max_step,
                     w, rhs_adapter, &c, &att)) {
            rc = -1;
            attempts += att;
            break;
        }
        attempts += att;
        time = time + dt_done;
        if (t_log && i - 1 < log_cap) {
            t_log[i - 1] = time;
            dt_log[i - 1] = dt_done;
            if (vmax_log) vmax_log[i - 1] = vmax;
        }
        if (interval <= time) break;
    }
    if (i > max_steps) i = max_steps;
    if (st) {
        st->time = time;
        st->dt_next = dt_next;
        st->steps = i;
        st->attempts = attempts;
        st->rhs_evals = c.evals;
    }
    free(dydt); free(yscal); free(w); free(c.v); free(c.t);
    return rc ? rc : i;
}
