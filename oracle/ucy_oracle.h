/*
 * ucy_oracle.h -- CPU oracle for the Unicycle quasi-dynamic hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against; it is never linked into, loaded by or called
 * from the product library (unicycle_amd/csrc).  Only tests/, the smoke()
 * entry point and the cpu_baseline leg of bench.py may load it.
 *
 * Every function restates one routine of jdpmoore/unicycle fortran/3d/src
 * (file:line cited at each definition in ucy_oracle.c) in plain C with the
 * reference's rounding points: the Okada chain rounds its inputs and outputs
 * to binary32 exactly where okada92.f90 / dc3d.f do, everything else is IEEE
 * binary64 with no contraction (build with -ffp-contract=off).
 *
 * Conventions follow the reference: arrays of 3-vectors are column-major
 * (3,ns) i.e. x[3*i+k]; the kernel K is the physical (m x n) matrix stored
 * row-major, which is the reference's Fortran G(n,m) column-major array.
 */
#ifndef UCY_ORACLE_H
#define UCY_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- rectangles (Okada 1992) ------------------------------------------ */
void orc_refsys_okada92(int ns, const double *x, const double *L, const double *W,
                        const double *strike, const double *dip,
                        double *sv, double *dv, double *nv, double *xc);
int orc_dc3d(float alpha, float x, float y, float z, float depth, float dip,
             float al1, float al2, float aw1, float aw2,
             float disl1, float disl2, float disl3, float u[12]);
void orc_stress_okada92(double x1, double x2, double x3, double q1, double q2, double q3,
                        double L, double W, double strike, double dipd,
                        double s, double d, double t, double G, double lambda,
                        double sig[6]);
/* computeDisplacementOkada92; returns DC3D's iret (u = 0 when nonzero) */
int orc_displacement_okada92(double x1, double x2, double x3, double q1, double q2, double q3,
                             double L, double W, double strike, double dipd,
                             double s, double d, double t, double G, double lambda,
                             double u[3]);
void orc_traction_kernels_okada92(const double x[3], const double sv[3], const double dv[3],
                                  const double nv[3], int ns, const double *y,
                                  const double *L, const double *W, const double *strike,
                                  const double *dip, double s, double d, double t,
                                  double G, double lambda, double *ts, double *td, double *tn);
void orc_stress_kernels_okada92(const double x[3], const double sv[3], const double dv[3],
                                const double nv[3], int ns, const double *y,
                                const double *L, const double *W, const double *strike,
                                const double *dip, double s, double d, double t,
                                double G, double lambda, double *s11, double *s12,
                                double *s13, double *s22, double *s23, double *s33);

/* ---- triangles (Nikkhoo & Walter 2015) --------------------------------- */
void orc_refsys_nikkhoo15(int nVe, const double *v, int ns, const int *i1, const int *i2,
                          const int *i3, double *sv, double *dv, double *nv, double *xc);
void orc_strain_in_halfspace(double stress[6], double strain[6], double X, double Y, double Z,
                             const double P1[3], const double P2[3], const double P3[3],
                             double Ss, double Ds, double Ts, double mu, double lambda);
/* dispinhalfspace_ (ENU) and computeDisplacementNikkhoo15 (NED) */
void orc_disp_in_halfspace(double u[3], double X, double Y, double Z, const double P1[3],
                           const double P2[3], const double P3[3], double Ss, double Ds,
                           double Ts, double nu);
void orc_displacement_nikkhoo15(double x1, double x2, double x3, const double P1[3],
                                const double P2[3], const double P3[3], double ss, double d,
                                double t, double G, double lambda, double u[3]);
void orc_traction_kernels_nikkhoo15(const double x[3], const double sv[3], const double dv[3],
                                    const double nv[3], int ns, const int *i1, const int *i2,
                                    const int *i3, int nVe, const double *vert,
                                    double s, double d, double t, double G, double lambda,
                                    double *ts, double *td, double *tn);
void orc_stress_kernels_nikkhoo15(const double x[3], const double sv[3], const double dv[3],
                                  const double nv[3], int ns, const int *i1, const int *i2,
                                  const int *i3, int nVe, const double *vert,
                                  double s, double d, double t, double G, double lambda,
                                  double *s11, double *s12, double *s13, double *s22,
                                  double *s23, double *s33);

/* ---- strain volumes (Barbot, Moore & Lambert 2017) --------------------- */
void orc_refsys_strainvolume(int ns, const double *x, const double *L, const double *W,
                             const double *strike, double *sv, double *dv, double *nv,
                             double *xc);
int orc_stress_strainvolume(double x1, double x2, double x3, double q1, double q2, double q3,
                            double L, double T, double W, double theta,
                            const double epsp[6], double G, double nu, double sig[6]);
int orc_displacement_strainvolume(double x1, double x2, double x3, double q1, double q2,
                                  double q3, double L, double T, double W, double theta,
                                  const double epsp[6], double G, double nu, double u[3]);
int orc_traction_kernels_strainvolume(const double x[3], const double sv[3], const double dv[3],
                                      const double nv[3], int ns, const double *y,
                                      const double *L, const double *T, const double *W,
                                      const double *strike, const double epsp[6],
                                      double G, double lambda,
                                      double *ts, double *td, double *tn);
int orc_stress_kernels_strainvolume(const double x[3], const double sv[3], const double dv[3],
                                    const double nv[3], int ns, const double *y,
                                    const double *L, const double *T, const double *W,
                                    const double *strike, const double epsp[6],
                                    double G, double lambda, double *s11, double *s12,
                                    double *s13, double *s22, double *s23, double *s33);

/* ---- assembly (greens.f90 buildG) -------------------------------------- */
typedef struct {
    double lambda, mu;
    int nrect;  const double *rx, *rL, *rW, *rstrike, *rdip;   /* rectangles */
    int ntri;   const int *i1, *i2, *i3; int nve; const double *vert; /* triangles */
    int nvol;   const double *vx, *vL, *vT, *vW, *vstrike;     /* volumes */
} orc_mesh;
/* K row-major (m x n); m = 3*(nrect+ntri)+6*nvol, n = 2*(nrect+ntri)+6*nvol.
 * nthreads > 1 splits receivers across OpenMP threads. Returns 0 or <0. */
int orc_buildG(const orc_mesh *mesh, double *K, int nthreads);
void orc_matvec(int m, int n, const double *K, const double *v, double *t, int nthreads);
/* buildO (greens.f90:302-462): O, Of, Ol row-major (3 nobs x n), x (3,nobs).
 * Of, Ol may be NULL.  Returns 0 or -2 (strain-volume parameter check). */
int orc_buildO(const orc_mesh *mesh, int nobs, const double *x, double *O, double *Of,
               double *Ol, int nthreads);

/* ---- rate-and-state RHS (ratestate.f90 odefun) ------------------------- */
typedef struct {
    int n;                 /* number of patches */
    const double *Vpl, *rake, *mu0, *sig, *a, *b, *L, *Vo, *damping;
    int nvol;              /* number of strain volumes (viscouscycles) */
    const double *gammadot0, *npower, *edot_bg; /* per volume: prefactor, exponent,
                                                   background strain rate (6 each) */
    const double *Gm, *Qa, *Rgas, *To;          /* per volume: initial-stress law */
    int ic_variant;        /* 0: ratestate.f90 IC, 1: viscouscycles.f90 IC */
    const double *K;       /* m x n row-major */
    int m, ncol;
    int nthreads;
} orc_system;
void orc_init_state(const orc_system *sys, double *y);
/* y, dydt length 7*n (+12*nvol); returns vMax */
double orc_odefun(const orc_system *sys, const double *y, double *dydt,
                  double *vwork /*ncol*/, double *twork /*m*/);

/* ---- Cash-Karp RK45 and step control (ode45.f90) ----------------------- */
typedef void (*orc_rhs_fn)(void *ctx, const double *y, double *dydt);
void orc_rkck(int n, const double *yin, const double *dydt, double h, double *yout,
              double *yerr, double *work /* 7*n */, orc_rhs_fn f, void *ctx);
/* returns 0 on success, -1 on step-size underflow */
int orc_rkqs(int n, double *t, double *y, double *dydt, const double *yscal,
             double htry, double *hdid, double *hnext, double epsilon, double max_step,
             double *work /* 9*n */, orc_rhs_fn f, void *ctx, int *n_attempts);

/* main loop of ratestate.f90:307-328.  Records accepted (time, dt) pairs.
 * Returns number of accepted steps or <0 on failure. */
typedef struct {
    double time, dt_next;
    long steps, attempts, rhs_evals;
} orc_run_stats;
long orc_integrate(const orc_system *sys, double *y, double interval, double dt0,
                   long max_steps, double epsilon, double max_step,
                   double *t_log, double *dt_log, double *vmax_log, long log_cap,
                   orc_run_stats *st);

/* ---- 2-D antiplane (fortran/2d/antiplane/src) ------------------------- */
void orc_refsys_antiplane(int ns, const double *x, const double *W, const double *dip,
                          double *sv, double *dv, double *nv, double *xc);
void orc_stress_antiplane(double x2, double x3, double q2, double q3, double W, double dip,
                          double G, double *s12, double *s13);
double orc_displacement_antiplane(double x2, double x3, double q2, double q3, double W,
                                  double dip, double s);
void orc_traction_kernels_antiplane(const double x[3], const double sv[3], const double dv[3],
                                    const double nv[3], int ns, const double *y,
                                    const double *W, const double *dip, double G, double *ts);
/* K (ns x ns) row-major: row = receiver traction, column = source slip */
void orc_buildG_antiplane(int ns, const double *x, const double *W, const double *dip, double mu,
                          double *K, int nthreads);
void orc_buildO_antiplane(int ns, const double *x, const double *W, const double *dip, int npts,
                          const double *pts, double *O);
/* strain volumes of the 2-D antiplane viscouscycles model
 * (fortran/2d/antiplane/src/antiplane.f90:213-255, 479-635): s12, s13 of the
 * primed eigenstrain (e12p, e13p) of a volume of thickness T and width W
 * whose top is at (q2, q3); the displacement it derives from */
void orc_stress_sv_antiplane(double x2, double x3, double q2, double q3, double T, double W,
                             double dip, double e12p, double e13p, double G, double *s12,
                             double *s13);
double orc_displacement_sv_antiplane(double x2, double x3, double q2, double q3, double T,
                                     double W, double dip, double e12p, double e13p);
/* buildG of patches and volumes (greens.f90:93-248 of fortran/2d/antiplane):
 * m = n = np + 2 nv; rows: one traction per patch, then (s12, s13) per volume
 * in its frame; columns: one slip per patch, then (e12, e13) per volume */
void orc_buildG_antiplane_volumes(int np, const double *px, const double *pW,
                                  const double *pdip, int nv, const double *vx,
                                  const double *vT, const double *vW, const double *vdip,
                                  double mu, double *K, int nthreads);
typedef struct {
    int n;                 /* patches; state 4 per patch: slip, traction, log10 theta, log10 V */
    const double *Vpl, *mu0, *sig, *a, *b, *L, *Vo, *damping;
    const double *K;       /* (n + 2 nvol) x (n + 2 nvol) row-major */
    /* volumes (viscouscycles.f90 of fortran/2d/antiplane); state 4 per volume:
     * e12, e13, s12, s13; background strain rate (e12, e13); power law
     * gammadot0 (sII/Gm)^(n-1)/Gm exp(-Q/(R To)) */
    int nvol;
    const double *gammadot0, *npower, *e12, *e13, *Gm, *Q, *R, *To;
} orc_ap_system;
/* ---- 2-D plane strain (fortran/2d/planestrain/src) --------------------- */
/* stress and displacement of a dipping strain volume (computeStressStrain-
 * VolumePlaneStrain.f90, computeDisplacementStrainVolumePlaneStrain.f90):
 * observation point (r2, x3), zone top (q2, q3), thickness T, width W, dip
 * (radians), strain e22', e23', e33' in the zone's frame.  0, or -2 / -3 / -4
 * for the reference's STOPs (Poisson's ratio, x3 < 0, q3 < 0). */
int orc_stress_sv_planestrain(double r2, double x3, double q2, double q3, double T, double W,
                              double dip, double e22p, double e23p, double e33p, double G,
                              double lambda, double *s22, double *s23, double *s33);
int orc_displacement_sv_planestrain(double r2, double x3, double q2, double q3, double T,
                                    double W, double dip, double e22p, double e23p,
                                    double e33p, double G, double lambda, double *u2,
                                    double *u3);
/* line dislocation with dip slip s (planestrain.f90:105-164) */
int orc_stress_planestrain(double x2, double x3, double q2, double q3, double W, double dip,
                           double s, double G, double lambda, double *s22, double *s23,
                           double *s33);
int orc_displacement_planestrain(double x2, double x3, double q2, double q3, double W,
                                 double dip, double s, double G, double lambda, double *u2,
                                 double *u3);
int orc_traction_kernels_planestrain(const double x[3], const double sv[3], const double dv[3],
                                     const double nv[3], int ns, const double *y,
                                     const double *W, const double *dip, double G,
                                     double lambda, double *td, double *tn);
/* K (2 ns x ns) row-major: rows 2r, 2r+1 = dip and normal traction of
 * receiver r, column s = dip slip of source s */
int orc_buildG_planestrain(int ns, const double *x, const double *W, const double *dip,
                           double G, double lambda, double *K, int nthreads);
/* the volume-frame stress kernels (planestrain.f90:314-348, 435-475;
 * s22' = n.s.n, s23' = -n.s.d, s33' = d.s.d) and the traction kernels of
 * strain volumes (:375-410), vectorised over ns sources as the reference's */
int orc_stress_kernels_planestrain(const double x[3], const double sv[3], const double dv[3],
                                   const double nv[3], int ns, const double *y,
                                   const double *W, const double *dip, double G,
                                   double lambda, double *s22, double *s23, double *s33);
int orc_traction_kernels_sv_planestrain(const double x[3], const double sv[3],
                                        const double dv[3], const double nv[3], int ns,
                                        const double *y, const double *T, const double *W,
                                        const double *dip, double e22p, double e23p,
                                        double e33p, double G, double lambda, double *td,
                                        double *tn);
int orc_stress_kernels_sv_planestrain(const double x[3], const double sv[3], const double dv[3],
                                      const double nv[3], int ns, const double *y,
                                      const double *T, const double *W, const double *dip,
                                      double e22p, double e23p, double e33p, double G,
                                      double lambda, double *s22, double *s23, double *s33);
/* buildG of the plane-strain viscouscycles program (greens.f90:66-226):
 * receivers np patches (rows td, tn) then nv volumes (rows s22', s23',
 * s33'); sources np dip slips then nv x 3 strains of 1e-3 (e22', e23',
 * e33'); K (2 np + 3 nv) x (np + 3 nv) row-major.  The volume rows of patch
 * sources take G and lambda swapped, as stressRows passes them (greens.f90:
 * 212-216 against the argument order of planestrain.f90:314). */
int orc_buildG_planestrain_volumes(int np, const double *x, const double *W, const double *dip,
                                   int nv, const double *vx, const double *vT, const double *vW,
                                   const double *vdip, double G, double lambda, double *K,
                                   int nthreads);
typedef struct {
    int n;   /* patches; state 5 per patch: dip slip, dip traction, normal
              * traction, log10 theta, log10 V */
    const double *Vpl, *tau0, *mu0, *sig, *a, *b, *L, *Vo, *damping;  /* tau0 may be NULL */
    const double *K;       /* (2n + 3nv) x (n + 3nv) row-major */
    /* strain volumes (viscouscycles.f90 of fortran/2d/planestrain): state 6
     * per volume after the patches: e22, e23, e33, s22, s23, s33; the
     * nonlinear Maxwell rheology gammadot0, n, Q, R, To; the background
     * strain rate e_bg (3 per volume: e22, e23, e33); s0 (3 per volume,
     * may be NULL) the initial stress, used where not all zero */
    int nv;
    const double *gammadot0, *npower, *Q, *R, *To, *ebg, *s0;
} orc_ps_system;
void orc_ps_init_state(const orc_ps_system *s, double *y);
double orc_ps_odefun(const orc_ps_system *s, const double *y, double *dydt, double *vwork,
                     double *twork);
long orc_ps_integrate(const orc_ps_system *s, double *y, double interval, double dt0,
                      long max_steps, double epsilon, double max_step, double *t_log,
                      double *dt_log, double *vmax_log, long log_cap, orc_run_stats *st);

void orc_ap_init_state(const orc_ap_system *s, double *y);
double orc_ap_odefun(const orc_ap_system *s, const double *y, double *dydt, double *vwork,
                     double *twork);
long orc_ap_integrate(const orc_ap_system *s, double *y, double interval, double dt0,
                      long max_steps, double epsilon, double max_step, double *t_log,
                      double *dt_log, double *vmax_log, long log_cap, orc_run_stats *st);

#ifdef __cplusplus
}
#endif
#endif
