/* orc_planestrain.cpp -- CPU oracle of the 2-D plane-strain model of
 * jdpmoore/unicycle (fortran/2d/planestrain/src).  Test infrastructure only:
 * the tests, smoke() and the bench's CPU baseline load it; the product path
 * never does.
 *
 * Restates:
 *   computeStressStrainVolumePlaneStrain.f90 (stress of a dipping strain
 *     volume) as Hooke's law on the exact gradient of the displacement
 *     integrals of computeDisplacementStrainVolumePlaneStrain.f90, less the
 *     eigenstrain inside (:88-115 of the stress routine), with the parameter
 *     checks of :72-86;
 *   planestrain.f90:105-164 (line dislocation = zone of thickness W/1e4 with
 *     e23' = -s / (2 T)), :55-83 (frames), :266-291 (traction kernels);
 *   greens.f90:66-226 (buildG for patches: rows td, tn per receiver, one dip
 *     slip column per source);
 *   ratestate.f90:585-627 (initStateVector, with the tau0 option) and
 *     :637-753 (odefun); the main loop is ratestate.f90:266-295 with RKQSm
 *     of ode45.f90:51-108, the 3-D loop's arithmetic (orc_rkqs);
 *   with strain volumes, the viscouscycles program of the same directory:
 *     planestrain.f90:314-475 (volume-frame stress kernels, traction
 *     kernels of volumes), greens.f90:66-226 (all four blocks, unit strains
 *     of 1e-3), viscouscycles.f90:714-837 (initStateVector: the volume
 *     stress from the power law at the background rate, or s0) and
 *     :841-1065 (odefun: eII = gammadot0 sII^(n-1) exp(-Q/(R To)) with the
 *     2-D deviatoric stress, v = (e - e_bg) 1e3).
 *
 * The displacement integrals are shared with the device kernel
 * (unicycle_amd/csrc/ps_jint.h, templated on the number type); here they are
 * instantiated with a host dual number of the observation point and the C
 * library's functions.  The formulas are pinned against the reference's own
 * build (oracle/_ref) and against the mathematics, not against the device.
 */
#include <cmath>
#include <cstdlib>
#include <cstring>

extern "C" {
#include "ucy_oracle.h"
}

namespace {

struct PD {  // value and derivatives along x2, x3
    double v, g0, g1;
    PD() : v(0.0), g0(0.0), g1(0.0) {}
    PD(double c) : v(c), g0(0.0), g1(0.0) {}
    PD(double c, double a, double b) : v(c), g0(a), g1(b) {}
};
inline PD operator+(const PD &a, const PD &b) { return PD(a.v + b.v, a.g0 + b.g0, a.g1 + b.g1); }
inline PD operator-(const PD &a, const PD &b) { return PD(a.v - b.v, a.g0 - b.g0, a.g1 - b.g1); }
inline PD operator-(const PD &a) { return PD(-a.v, -a.g0, -a.g1); }
inline PD operator*(const PD &a, const PD &b)
{
    return PD(a.v * b.v, a.g0 * b.v + a.v * b.g0, a.g1 * b.v + a.v * b.g1);
}
inline PD operator/(const PD &a, const PD &b)
{
    const double q = a.v / b.v, r = 1.0 / b.v;
    return PD(q, (a.g0 - q * b.g0) * r, (a.g1 - q * b.g1) * r);
}
inline PD operator+(const PD &a, double b) { return PD(a.v + b, a.g0, a.g1); }
inline PD operator+(double b, const PD &a) { return PD(b + a.v, a.g0, a.g1); }
inline PD operator-(const PD &a, double b) { return PD(a.v - b, a.g0, a.g1); }
inline PD operator-(double b, const PD &a) { return PD(b - a.v, -a.g0, -a.g1); }
inline PD operator*(const PD &a, double b) { return PD(a.v * b, a.g0 * b, a.g1 * b); }
inline PD operator*(double b, const PD &a) { return PD(b * a.v, b * a.g0, b * a.g1); }
inline PD operator/(const PD &a, double b) { return PD(a.v / b, a.g0 / b, a.g1 / b); }
inline PD operator/(double b, const PD &a)
{
    const double q = b / a.v;
    return PD(q, -q * a.g0 / a.v, -q * a.g1 / a.v);
}
inline PD ps_sqrt(const PD &a)
{
    const double s = std::sqrt(a.v), h = 0.5 / s;
    return PD(s, h * a.g0, h * a.g1);
}
inline PD ps_log(const PD &a) { return PD(std::log(a.v), a.g0 / a.v, a.g1 / a.v); }
// atan2(u, 1) as the reference writes atan(u); d = du / (1 + u^2)
inline PD ps_atan2(const PD &u, double one)
{
    const double r = 1.0 / (one * one + u.v * u.v);
    return PD(std::atan2(u.v, one), one * u.g0 * r, one * u.g1 * r);
}
inline double ps_sqrt(double a) { return std::sqrt(a); }
inline double ps_log(double a) { return std::log(a); }
inline double ps_atan2(double u, double one) { return std::atan2(u, one); }

}  // namespace

#include "../unicycle_amd/csrc/ps_jint.h"

static psj::Geom ps_geom(double q3, double dip, double nu)
{
    psj::Geom g;
    g.q3 = q3;
    g.nu = nu;
    g.s1 = std::sin(dip);
    g.c1 = std::cos(dip);
    g.s2 = std::sin(2.0 * dip);
    g.c2 = std::cos(2.0 * dip);
    g.s3 = std::sin(3.0 * dip);
    g.c3 = std::cos(3.0 * dip);
    g.s4 = std::sin(4.0 * dip);
    g.c4 = std::cos(4.0 * dip);
    return g;
}

/* the parameter checks of computeStressStrainVolumePlaneStrain.f90:75-86 */
static int ps_check(double nu, double x3, double q3)
{
    if (-1.0 > nu || 0.5 < nu) return -2;
    if (0 > x3) return -3;
    if (0 > q3) return -4;
    return 0;
}

extern "C" int orc_stress_sv_planestrain(double r2, double x3, double q2, double q3, double T,
                                         double W, double dip, double e22p, double e23p,
                                         double e33p, double G, double lambda, double *s22,
                                         double *s23, double *s33)
{
    const double nu = lambda / 2.0 / (lambda + G);
    const int rc = ps_check(nu, x3, q3);
    if (rc) return rc;
    const psj::Geom g = ps_geom(q3, dip, nu);
    const psj::Strain e = psj::rotate(g, e22p, e23p, e33p, lambda, G);
    const double x2 = r2 - q2;
    psj::Point<PD> P{PD(x2, 1.0, 0.0), PD(x3, 0.0, 1.0)};
    PD u2, u3;
    psj::displacement(P, g, e, T, W, u2, u3);
    const double in = psj::omega((x2 * g.s1 - (x3 - q3) * g.c1) / T)
                      * psj::boxcar_s((x2 * g.c1 + (x3 - q3) * g.s1) / W);
    const double e22 = u2.g0 - e.e22 * in;
    const double e23 = (u2.g1 + u3.g0) / 2 - e.e23 * in;
    const double e33 = u3.g1 - e.e33 * in;
    *s22 = lambda * (e22 + e33) + 2 * G * e22;
    *s23 = 2 * G * e23;
    *s33 = lambda * (e22 + e33) + 2 * G * e33;
    return 0;
}

extern "C" int orc_displacement_sv_planestrain(double r2, double x3, double q2, double q3,
                                               double T, double W, double dip, double e22p,
                                               double e23p, double e33p, double G, double lambda,
                                               double *u2, double *u3)
{
    const double nu = lambda / 2.0 / (lambda + G);
    const int rc = ps_check(nu, x3, q3);
    if (rc) return rc;
    const psj::Geom g = ps_geom(q3, dip, nu);
    const psj::Strain e = psj::rotate(g, e22p, e23p, e33p, lambda, G);
    psj::Point<double> P{r2 - q2, x3};
    psj::displacement(P, g, e, T, W, *u2, *u3);
    return 0;
}

/* planestrain.f90:105-124 and :146-164: dip slip s on a line dislocation */
extern "C" int orc_stress_planestrain(double x2, double x3, double q2, double q3, double W,
                                      double dip, double s, double G, double lambda,
                                      double *s22, double *s23, double *s33)
{
    const double T = W / 1e4;
    const double eps23p = -s / (2.0 * T);
    return orc_stress_sv_planestrain(x2, x3, q2, q3, T, W, dip, 0.0, eps23p, 0.0, G, lambda,
                                     s22, s23, s33);
}

extern "C" int orc_displacement_planestrain(double x2, double x3, double q2, double q3,
                                            double W, double dip, double s, double G,
                                            double lambda, double *u2, double *u3)
{
    const double T = W / 1e4;
    const double eps23p = -s / (2.0 * T);
    return orc_displacement_sv_planestrain(x2, x3, q2, q3, T, W, dip, 0.0, eps23p, 0.0, G,
                                           lambda, u2, u3);
}

/* planestrain.f90:266-291 */
extern "C" int orc_traction_kernels_planestrain(const double x[3], const double sv[3],
                                               const double dv[3], const double nv[3], int ns,
                                               const double *y, const double *W,
                                               const double *dip, double G, double lambda,
                                               double *td, double *tn)
{
    (void)sv;
    for (int i = 0; i < ns; i++) {
        double s22, s23, s33;
        const int rc = orc_stress_planestrain(x[1], x[2], y[3 * i + 1], y[3 * i + 2], W[i],
                                              dip[i], 1.0, G, lambda, &s22, &s23, &s33);
        if (rc) return rc;
        td[i] = (nv[1] * s22 + nv[2] * s23) * dv[1] + (nv[1] * s23 + nv[2] * s33) * dv[2];
        tn[i] = (nv[1] * s22 + nv[2] * s23) * nv[1] + (nv[1] * s23 + nv[2] * s33) * nv[2];
    }
    return 0;
}

/* buildG (greens.f90:66-226) for patches: K (2 ns x ns) row-major; rows 2r
 * (dip traction) and 2r + 1 (normal traction) of receiver r, column s = dip
 * slip of source s.  The frames are those of the antiplane model
 * (planestrain.f90:55-83 is the same construction). */
extern "C" int orc_buildG_planestrain(int ns, const double *x, const double *W, const double *dip,
                                      double G, double lambda, double *K, int nthreads)
{
    double *f = (double *)std::malloc(sizeof(double) * 12 * (size_t)(ns + 1));
    double *sv = f, *dv = f + 3 * (ns + 1), *nv = f + 6 * (ns + 1), *xc = f + 9 * (ns + 1);
    orc_refsys_antiplane(ns, x, W, dip, sv, dv, nv, xc);
    int rc = 0;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1) reduction(min : rc)
    for (int r = 0; r < ns; r++) {
        const int e = orc_traction_kernels_planestrain(xc + 3 * r, sv + 3 * r, dv + 3 * r,
                                                       nv + 3 * r, ns, x, W, dip, G, lambda,
                                                       K + (size_t)(2 * r) * ns,
                                                       K + (size_t)(2 * r + 1) * ns);
        if (e < rc) rc = e;
    }
    std::free(f);
    return rc;
}

/* rotation into the receiver's frame of planestrain.f90:332-340 (and
 * :462-470): s22' = n.s.n, s23' = -n.s.d, s33' = d.s.d, in the reference's
 * operation order */
static void ps_volume_frame(const double nv[3], const double dv[3], double s22, double s23,
                            double s33, double *o22, double *o23, double *o33)
{
    *o22 = (nv[1] * s22 + nv[2] * s23) * nv[1] + (nv[1] * s23 + nv[2] * s33) * nv[2];
    *o23 = -(nv[1] * s22 + nv[2] * s23) * dv[1] - (nv[1] * s23 + nv[2] * s33) * dv[2];
    *o33 = -(-dv[1] * s22 - dv[2] * s23) * dv[1] - (-dv[1] * s23 - dv[2] * s33) * dv[2];
}

/* planestrain.f90:314-348 */
extern "C" int orc_stress_kernels_planestrain(const double x[3], const double sv[3],
                                              const double dv[3], const double nv[3], int ns,
                                              const double *y, const double *W,
                                              const double *dip, double G, double lambda,
                                              double *s22, double *s23, double *s33)
{
    (void)sv;
    for (int i = 0; i < ns; i++) {
        double a, b, c;
        const int rc = orc_stress_planestrain(x[1], x[2], y[3 * i + 1], y[3 * i + 2], W[i], dip[i],
                                              1.0, G, lambda, &a, &b, &c);
        if (rc) return rc;
        ps_volume_frame(nv, dv, a, b, c, s22 + i, s23 + i, s33 + i);
    }
    return 0;
}

/* planestrain.f90:375-410 */
extern "C" int orc_traction_kernels_sv_planestrain(const double x[3], const double sv[3],
                                                   const double dv[3], const double nv[3],
                                                   int ns, const double *y, const double *T,
                                                   const double *W, const double *dip,
                                                   double e22p, double e23p, double e33p,
                                                   double G, double lambda, double *td,
                                                   double *tn)
{
    (void)sv;
    for (int i = 0; i < ns; i++) {
        double s22, s23, s33;
        const int rc = orc_stress_sv_planestrain(x[1], x[2], y[3 * i + 1], y[3 * i + 2], T[i], W[i],
                                                 dip[i], e22p, e23p, e33p, G, lambda, &s22, &s23,
                                                 &s33);
        if (rc) return rc;
        td[i] = (nv[1] * s22 + nv[2] * s23) * dv[1] + (nv[1] * s23 + nv[2] * s33) * dv[2];
        tn[i] = (nv[1] * s22 + nv[2] * s23) * nv[1] + (nv[1] * s23 + nv[2] * s33) * nv[2];
    }
    return 0;
}

/* planestrain.f90:435-475 */
extern "C" int orc_stress_kernels_sv_planestrain(const double x[3], const double sv[3],
                                                 const double dv[3], const double nv[3], int ns,
                                                 const double *y, const double *T,
                                                 const double *W, const double *dip, double e22p,
                                                 double e23p, double e33p, double G,
                                                 double lambda, double *s22, double *s23,
                                                 double *s33)
{
    (void)sv;
    for (int i = 0; i < ns; i++) {
        double a, b, c;
        const int rc = orc_stress_sv_planestrain(x[1], x[2], y[3 * i + 1], y[3 * i + 2], T[i], W[i],
                                                 dip[i], e22p, e23p, e33p, G, lambda, &a, &b, &c);
        if (rc) return rc;
        ps_volume_frame(nv, dv, a, b, c, s22 + i, s23 + i, s33 + i);
    }
    return 0;
}

/* greens.f90:66-226 of fortran/2d/planestrain with strain volumes */
extern "C" int orc_buildG_planestrain_volumes(int np, const double *x, const double *W,
                                              const double *dip, int nv, const double *vx,
                                              const double *vT, const double *vW,
                                              const double *vdip, double G, double lambda,
                                              double *K, int nthreads)
{
    const int ne = np + nv;
    const size_t n = (size_t)np + 3 * (size_t)nv;
    double *f = (double *)std::malloc(sizeof(double) * 12 * (size_t)(ne + 1));
    double *sv = f, *dv = f + 3 * (ne + 1), *nvv = f + 6 * (ne + 1), *xc = f + 9 * (ne + 1);
    orc_refsys_antiplane(np, x, W, dip, sv, dv, nvv, xc);
    orc_refsys_antiplane(nv, vx, vW, vdip, sv + 3 * np, dv + 3 * np, nvv + 3 * np, xc + 3 * np);
    int rc = 0;
    (void)nthreads;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(min : rc)
    {
        double *w = (double *)std::malloc(sizeof(double) * 3 * (size_t)(nv + 1));
#pragma omp for schedule(dynamic, 4)
        for (int r = 0; r < ne; r++) {
            const double *X = xc + 3 * r, *S = sv + 3 * r, *D = dv + 3 * r, *N = nvv + 3 * r;
            int e = 0;
            if (r < np) {
                double *rd = K + (size_t)(2 * r) * n, *rn = rd + n;
                e = orc_traction_kernels_planestrain(X, S, D, N, np, x, W, dip, G, lambda, rd, rn);
                for (int j = 0; j < 3 && !e; j++) {
                    const double u[3] = {j == 0 ? 1e-3 : 0.0, j == 1 ? 1e-3 : 0.0,
                                         j == 2 ? 1e-3 : 0.0};
                    e = orc_traction_kernels_sv_planestrain(X, S, D, N, nv, vx, vT, vW, vdip, u[0],
                                                            u[1], u[2], G, lambda, w, w + nv);
                    for (int i = 0; i < nv && !e; i++) {
                        rd[np + 3 * i + j] = w[i];
                        rn[np + 3 * i + j] = w[nv + i];
                    }
                }
            } else {
                double *r0 = K + (size_t)(2 * np + 3 * (r - np)) * n, *r1 = r0 + n, *r2 = r1 + n;
                // the moduli in the order stressRows passes them (mu, lambda)
                // into the (lambda, G) of computeStressKernelsPlaneStrain
                e = orc_stress_kernels_planestrain(X, S, D, N, np, x, W, dip, lambda, G, r0, r1, r2);
                for (int j = 0; j < 3 && !e; j++) {
                    const double u[3] = {j == 0 ? 1e-3 : 0.0, j == 1 ? 1e-3 : 0.0,
                                         j == 2 ? 1e-3 : 0.0};
                    e = orc_stress_kernels_sv_planestrain(X, S, D, N, nv, vx, vT, vW, vdip, u[0],
                                                          u[1], u[2], G, lambda, w, w + nv,
                                                          w + 2 * (size_t)nv);
                    for (int i = 0; i < nv && !e; i++) {
                        r0[np + 3 * i + j] = w[i];
                        r1[np + 3 * i + j] = w[nv + i];
                        r2[np + 3 * i + j] = w[2 * nv + i];
                    }
                }
            }
            if (e < rc) rc = e;
        }
        std::free(w);
    }
    std::free(f);
    return rc;
}

/* ratestate.f90 (2-D plane strain):585-627 */
extern "C" void orc_ps_init_state(const orc_ps_system *s, double *y)
{
    const double lg10 = 2.302585092994045684;  /* LOG(1.d1) */
    for (int i = 0; i < s->n; i++) {
        double *p = y + 5 * i;
        p[0] = 0.0;
        if (!s->tau0 || 0 >= s->tau0[i])
            p[1] = s->sig[i] * (s->mu0[i] + (s->a[i] - s->b[i]) * std::log(s->Vpl[i] / s->Vo[i]))
                   - s->damping[i] * s->Vpl[i];
        else
            p[1] = s->tau0[i];
        p[2] = 0.0;
        p[3] = std::log(s->L[i] / s->Vpl[i]) / lg10;
        p[4] = std::log(s->Vpl[i] * 0.98) / lg10;
    }
    /* viscouscycles.f90 (2-D plane strain):777-821 */
    for (int i = 0; i < s->nv; i++) {
        double *q = y + 5 * (size_t)s->n + 6 * (size_t)i;
        const double *e = s->ebg + 3 * i;
        double s22, s23, s33;
        const double *z = s->s0 ? s->s0 + 3 * i : nullptr;
        if (!z || (0.0 == z[0] && 0.0 == z[1] && 0.0 == z[2])) {
            const double ekk = e[0] + e[2];
            const double e22p = e[0] - ekk / 2, e33p = e[2] - ekk / 2;
            const double eps = std::sqrt((e22p * e22p + 2.0 * (e[1] * e[1]) + e33p * e33p) / 2.0);
            if (0.0 < eps && 0.0 < s->gammadot0[i]) {
                const double tau = std::pow(eps / s->gammadot0[i] *
                                                std::exp(s->Q[i] / (s->R[i] * s->To[i])),
                                            1.0 / s->npower[i]);
                s22 = tau * e22p / eps;
                s23 = tau * e[1] / eps;
                s33 = tau * e33p / eps;
            } else {
                s22 = s23 = s33 = 0.0;
            }
        } else {
            s22 = z[0];
            s23 = z[1];
            s33 = z[2];
        }
        q[0] = q[1] = q[2] = 0.0;
        q[3] = s22;
        q[4] = s23;
        q[5] = s33;
    }
}

/* odefun (ratestate.f90:637-753): returns vMax */
extern "C" double orc_ps_odefun(const orc_ps_system *s, const double *y, double *dydt,
                                double *vwork, double *twork)
{
    const double lg10 = 2.302585092994045684;  /* LOG(1.d1) */
    const int n = s->n, nv = s->nv;
    double vmax = 0.0;
    std::memset(dydt, 0, sizeof(double) * (5 * (size_t)n + 6 * (size_t)nv));
    for (int i = 0; i < n; i++) {
        const double velocity = std::exp(y[5 * i + 4] * lg10);
        vmax = velocity > vmax ? velocity : vmax;
        dydt[5 * i] = velocity;
        vwork[i] = velocity - s->Vpl[i];
    }
    /* viscouscycles.f90 (2-D plane strain):930-967 */
    for (int i = 0; i < nv; i++) {
        const double *q = y + 5 * (size_t)n + 6 * (size_t)i;
        double *dq = dydt + 5 * (size_t)n + 6 * (size_t)i;
        const double p = (q[3] + q[5]) / 2.0;
        const double s22p = q[3] - p, s33p = q[5] - p, s23 = q[4];
        const double sII = std::sqrt((s22p * s22p + 2.0 * (s23 * s23) + s33p * s33p) / 2.0);
        const double eII = s->gammadot0[i] * std::pow(sII, s->npower[i] - 1.0)
                           * std::exp(-s->Q[i] / (s->R[i] * s->To[i]));
        const double e22 = eII * s22p, e23 = eII * s23, e33 = eII * s33p;
        dq[0] = e22;
        dq[1] = e23;
        dq[2] = e33;
        double *v = vwork + n + 3 * (size_t)i;
        v[0] = (e22 - s->ebg[3 * i]) * 1e3;
        v[1] = (e23 - s->ebg[3 * i + 1]) * 1e3;
        v[2] = (e33 - s->ebg[3 * i + 2]) * 1e3;
    }
    orc_matvec(2 * n + 3 * nv, n + 3 * nv, s->K, vwork, twork, 1);
    for (int i = 0; i < nv; i++) {  /* :1048-1053 */
        double *dq = dydt + 5 * (size_t)n + 6 * (size_t)i;
        const double *t = twork + 2 * (size_t)n + 3 * (size_t)i;
        dq[3] = t[0];
        dq[4] = t[1];
        dq[5] = t[2];
    }
    for (int i = 0; i < n; i++) {
        const double *p = y + 5 * i;
        double *dp = dydt + 5 * i;
        const double velocity = std::exp(p[4] * lg10);
        dp[3] = (std::exp(-p[3] * lg10) - velocity / s->L[i]) / lg10;
        const double dtau = twork[2 * i];
        const double sigma = s->sig[i];
        dp[4] = (dtau - s->b[i] * sigma * dp[3] * lg10) / (s->a[i] * sigma + s->damping[i] * velocity)
                / lg10;
        dp[1] = dtau - s->damping[i] * velocity * dp[4] * lg10;
        dp[2] = twork[2 * i + 1];
    }
    return vmax;
}

namespace {
struct ps_ctx {
    const orc_ps_system *s;
    double *v, *t, vmax;
    long evals;
};
void ps_rhs(void *ctx, const double *y, double *dydt)
{
    ps_ctx *c = (ps_ctx *)ctx;
    c->vmax = orc_ps_odefun(c->s, y, dydt, c->v, c->t);
    c->evals++;
}
}  // namespace

/* the main loop (ratestate.f90:266-295) */
extern "C" long orc_ps_integrate(const orc_ps_system *s, double *y, double interval, double dt0,
                                 long max_steps, double epsilon, double max_step, double *t_log,
                                 double *dt_log, double *vmax_log, long log_cap,
                                 orc_run_stats *st)
{
    const int dim = 5 * s->n + 6 * s->nv, nvv = s->n + 3 * s->nv;
    double *work = (double *)std::malloc(sizeof(double) * (size_t)(11 * dim + 3 * nvv));
    double *dydt = work, *yscal = work + dim, *rk = work + 2 * dim;
    ps_ctx c = {s, work + 11 * dim, work + 11 * dim + nvv, 0.0, 0};
    double time = 0.0, dt_next = dt0;
    long steps = 0, attempts = 0;
    while (steps < max_steps) {
        ps_rhs(&c, y, dydt);
        const double vmax = c.vmax;
        const double dt_try = dt_next;
        for (int i = 0; i < dim; i++) yscal[i] = std::fabs(y[i]) + std::fabs(dt_try * dydt[i]) + 1e-30;
        double t0 = 0.0, hdid, hnext;
        int na = 0;
        if (orc_rkqs(dim, &t0, y, dydt, yscal, dt_try, &hdid, &hnext, epsilon, max_step, rk,
                     ps_rhs, &c, &na) != 0) {
            std::free(work);
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
    std::free(work);
    return steps;
}
