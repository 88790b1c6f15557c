/*
 * ucy.h -- C ABI of the MI355X engine for the quasi-dynamic boundary-integral
 * hot path of jdpmoore/unicycle (fortran/3d).
 *
 * Plain pointers and sizes only.  All pointers are host pointers unless the
 * name ends in _dev (device pointers, stream-ordered on the context stream).
 * Every function returns 0 on success and a negative status on failure; the
 * message is available from ucy_last_error().  The library never aborts the
 * process (the reference STOPs, e.g. ode45.f90:86-90; the Fortran module
 * ucy_fortran (unicycle_amd/fortran/ucy_fortran.f90) maps a nonzero status
 * back to STOP).
 *
 * Layout conventions are the reference's:
 *   - arrays of 3-vectors are Fortran (3,ns) arrays, i.e. x[3*i+k];
 *   - the state vector is the reference's interleaved layout, 7 entries per
 *     patch [ss, sd, ts, td, tn, log10(theta), log10(V)] (macros.f90:57-64)
 *     followed by 12 per volume [e11..e33, s11..s33] (macros.f90:67-79);
 *   - the kernel K is the physical (m x n) matrix stored row-major, which is
 *     byte-for-byte the reference's column-major G(n,m) of greens.f90:116,
 *     m = 3*Npatch + 6*Nvol receiver rows, n = 2*Npatch + 6*Nvol source
 *     columns (greens.f90:109-111, 214-216, 266-271).
 * Threading: one host thread per context; calls are ordered on the context's
 * stream; a context is not reentrant.
 */
#ifndef UCY_H
#define UCY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UCY_OK 0
#define UCY_ERR_ARG -1
#define UCY_ERR_HIP -2
#define UCY_ERR_STATE -3
#define UCY_ERR_UNDERFLOW -4   /* ode45.f90:86-90 "stepsize underflow in RKQS" */
#define UCY_ERR_NONFINITE -5
#define UCY_ERR_GEOMETRY -6    /* strainvolume.f90:1059-1072 parameter checks */
#define UCY_ERR_LIMIT -7

#define UCY_PREC_F64 0         /* K stored in binary64 */
#define UCY_PREC_SPLIT_BF16 1  /* K stored as per-row-scaled bf16 hi+lo pair */
/* ensembles only: binary64 K for the step-start matvec t_0 = K v_0 and the
 * split-bf16 pair for the stage increments t_i = t_0 + K (v_i - v_0) */
#define UCY_PREC_MIXED 2
/* ensembles only: as UCY_PREC_SPLIT_BF16 with K stored in 3 bytes per entry,
 * the bf16 hi and an 8-bit code of the residual in steps of 1/128 of hi's
 * ulp (about 2^-16 relative), decoded to the bf16 lo operand in registers;
 * 25 % fewer stage-GEMM bytes */
#define UCY_PREC_BF16_I8 3

typedef struct ucy_ctx ucy_ctx;

/* ---- lifecycle -------------------------------------------------------- */
/* One context per GPU, matching one MPI rank of the reference. */
ucy_ctx *ucy_create(int device, int flags);
void ucy_destroy(ucy_ctx *ctx);
const char *ucy_last_error(const ucy_ctx *ctx);
const char *ucy_version(void);

/* ---- Green's-function routines with the reference argument lists ------- */
/* Evaluated on the GPU for one receiver against ns sources (parity hooks).
 * Replaces okada92.f90:368-377 computeTractionKernelsOkada92. */
int ucy_okada92_traction_kernels(ucy_ctx *ctx, const double x[3], const double sv[3],
                                 const double dv[3], const double nv[3], int ns,
                                 const double *y, const double *L, const double *W,
                                 const double *strike, const double *dip, double s, double d,
                                 double t, double G, double lambda, double *ts, double *td,
                                 double *tn);
/* Replaces okada92.f90:424-432 computeStressKernelsOkada92. */
int ucy_okada92_stress_kernels(ucy_ctx *ctx, const double x[3], const double sv[3],
                               const double dv[3], const double nv[3], int ns, const double *y,
                               const double *L, const double *W, const double *strike,
                               const double *dip, double s, double d, double t, double G,
                               double lambda, double *s11, double *s12, double *s13,
                               double *s22, double *s23, double *s33);
/* Replaces nikkhoo15.f90:187-199 computeTractionKernelsNikkhoo15 (1-based i1..i3). */
int ucy_nikkhoo15_traction_kernels(ucy_ctx *ctx, const double x[3], const double sv[3],
                                   const double dv[3], const double nv[3], int ns,
                                   const int *i1, const int *i2, const int *i3, int nVe,
                                   const double *vertices, double s, double d, double t,
                                   double G, double lambda, double *ts, double *td, double *tn);
/* Replaces nikkhoo15.f90:250-262 computeStressKernelsNikkhoo15. */
int ucy_nikkhoo15_stress_kernels(ucy_ctx *ctx, const double x[3], const double sv[3],
                                 const double dv[3], const double nv[3], int ns, const int *i1,
                                 const int *i2, const int *i3, int nVe, const double *vertices,
                                 double s, double d, double t, double G, double lambda,
                                 double *s11, double *s12, double *s13, double *s22,
                                 double *s23, double *s33);
/* straininhalfspace_ (StrainInHalfSpace.c:42-173), batched over n
 * (observation point, triangle) pairs: point i (X[i], Y[i], Z[i]) and
 * triangle i (P1, P2, P3 as n x 3, row i = one vertex), all east/north/up
 * with Z <= 0; TDCS slip (Ss, Ds, Ts) and moduli as the reference takes them
 * by value.  stress and strain are n x 6 in the reference's order (xx, xy,
 * xz, yy, yz, zz); NaN where the reference yields NaN (point on a side). */
int ucy_straininhalfspace(ucy_ctx *ctx, int n, double *stress, double *strain, const double *X,
                          const double *Y, const double *Z, const double *P1, const double *P2,
                          const double *P3, double Ss, double Ds, double Ts, double mu,
                          double lambda);
/* Replaces strainvolume.f90:4671-4684 computeTractionKernelsVerticalStrainVolume. */
int ucy_strainvolume_traction_kernels(ucy_ctx *ctx, const double x[3], const double sv[3],
                                      const double dv[3], const double nv[3], int ns,
                                      const double *y, const double *L, const double *T,
                                      const double *W, const double *strike, double e11p,
                                      double e12p, double e13p, double e22p, double e23p,
                                      double e33p, double G, double lambda, double *ts,
                                      double *td, double *tn);
/* Replaces strainvolume.f90:4743-4756 computeStressKernelsVerticalStrainVolume. */
int ucy_strainvolume_stress_kernels(ucy_ctx *ctx, const double x[3], const double sv[3],
                                    const double dv[3], const double nv[3], int ns,
                                    const double *y, const double *L, const double *T,
                                    const double *W, const double *strike, double e11p,
                                    double e12p, double e13p, double e22p, double e23p,
                                    double e33p, double G, double lambda, double *s11,
                                    double *s12, double *s13, double *s22, double *s23,
                                    double *s33);

/* Reference frames (host arithmetic, identical to the reference routines).
 * Replace okada92.f90:59 computeReferenceSystemOkada92,
 * nikkhoo15.f90:50 / stuart97.f90:44 computeReferenceSystemNikkhoo15/Stuart97,
 * strainvolume.f90:135 computeReferenceSystemVerticalStrainVolume. */
int ucy_reference_system_okada92(int ns, const double *x, const double *L, const double *W,
                                 const double *strike, const double *dip, double *sv,
                                 double *dv, double *nv, double *xc);
int ucy_reference_system_nikkhoo15(int nVe, const double *v, int ns, const int *i1,
                                   const int *i2, const int *i3, double *sv, double *dv,
                                   double *nv, double *xc);
int ucy_reference_system_strainvolume(int ns, const double *x, const double *L,
                                      const double *W, const double *strike, double *sv,
                                      double *dv, double *nv, double *xc);

/* ---- model definition (the reference's input-deck sections) ------------ */
/* ratestate.f90:1092 "elastic moduli" lambda, mu (strike/dip in radians below). */
int ucy_set_elastic(ucy_ctx *ctx, double lambda, double mu);
/* ratestate.f90:1142-1176 rectangular patches: upper-left corner x(3,ns), length,
 * width, strike, dip (radians). */
int ucy_set_rectangles(ucy_ctx *ctx, int ns, const double *x, const double *length,
                       const double *width, const double *strike, const double *dip);
/* ratestate.f90:1252-1290 triangles: 1-based vertex indices, vertices v(3,nVe). */
int ucy_set_triangles(ucy_ctx *ctx, int ns, const int *i1, const int *i2, const int *i3,
                      int nVe, const double *v);
/* viscouscycles.f90 strain volumes: upper-left corner x(3,ns), length, thickness,
 * width, strike (radians). */
int ucy_set_volumes(ucy_ctx *ctx, int ns, const double *x, const double *length,
                    const double *thickness, const double *width, const double *strike);
/* The 2-D antiplane model of fortran/2d/antiplane (ratestate.f90 there):
 * ns vertical segments (upper corner x(3,ns), north/east/down, x(2) and x(3)
 * used; width; dip in radians, which the Green's function ignores as the
 * reference does, antiplane.f90:126-127) in a half space of shear modulus mu.
 * Replaces any 3-D elements.  The state has 4 entries per patch (slip,
 * traction, log10 theta, log10 V); K is ns x ns (traction row of each
 * receiver, slip column of each source); ucy_set_friction takes the same
 * arrays (rake unused), ucy_build_kernel builds K (greens.f90:93-248), and
 * ucy_init_state (ratestate.f90:568-605), ucy_odefun (:619-724), ucy_rkck,
 * ucy_rkqs and ucy_integrate(_ex) run this model. */
int ucy_set_antiplane(ucy_ctx *ctx, double mu, int ns, const double *x, const double *width,
                      const double *dip);
/* Strain volumes of the 2-D antiplane viscouscycles model
 * (fortran/2d/antiplane/src/viscouscycles.f90 there; deck section
 * :1404-1474): nv volumes with upper centre x(3,nv) (x(2), x(3) used),
 * thickness (horizontal), width (depth extent) and dip (radians; the frame
 * uses it, the Green's functions ignore it, antiplane.f90:479-515).  Call
 * after ucy_set_antiplane, which clears them.  The state gains 4 entries per
 * volume (e12, e13, s12, s13); K becomes (ns + 2 nv) square: after the
 * segments' rows and columns, two stress rows (s12, s13 in the volume's
 * frame) and two strain columns (e12, e13) per volume (greens.f90:93-248).
 * ucy_set_rheology supplies the power law per volume with the background
 * strain rate in edot_bg[6 i + 1] (e12) and [6 i + 2] (e13):
 * eII = gammadot0 (sII/Gm)^(n-1) / Gm exp(-Q/(R To)) (viscouscycles.f90:
 * 846-877), the initial stress from the same law at the background rate
 * (:740-762).  Displacement kernels gain the volumes' (e12, e13) columns
 * (computeDisplacementStrainVolumeAntiplane, antiplane.f90:213-255). */
int ucy_set_antiplane_volumes(ucy_ctx *ctx, int nv, const double *x, const double *thickness,
                              const double *width, const double *dip);
/* The 2-D plane-strain model of fortran/2d/planestrain (ratestate.f90 there):
 * ns segments (upper corner x(3,ns), x(2) and x(3) used; width; dip in
 * radians) slipping along dip in a half space of Lame parameters lambda, mu.
 * Replaces any other elements.  A segment is a line dislocation evaluated as
 * the reference does, as a shear zone 1e-4 of its width thick
 * (planestrain.f90:105-164).  The state has 5 entries per patch (dip slip,
 * dip traction, normal traction, log10 theta, log10 V); K is 2 ns x ns (rows
 * 2r, 2r+1: dip and normal traction of receiver r; column s: dip slip of
 * source s, greens.f90:66-226).  ucy_set_friction takes the same arrays (rake
 * unused), ucy_build_kernel builds K, ucy_build_displacement gives two rows
 * per point (east, down; greens.f90:253-367), and ucy_init_state
 * (ratestate.f90:585-627), ucy_odefun (:637-753), ucy_rkck, ucy_rkqs and
 * ucy_integrate(_ex) run it.  The parameter checks of the stress routine
 * (Poisson's ratio in [-1, 0.5], depths >= 0) give UCY_ERR_GEOMETRY. */
int ucy_set_planestrain(ucy_ctx *ctx, double lambda, double mu, int ns, const double *x,
                        const double *width, const double *dip);
/* the plane-strain deck's tau0 per patch: a positive value replaces the
 * initial dip traction (ratestate.f90 of 2d/planestrain:604-609); reset by
 * ucy_set_planestrain */
int ucy_set_initial_traction(ucy_ctx *ctx, int np, const double *tau0);
/* Strain volumes of the plane-strain viscouscycles program
 * (fortran/2d/planestrain/src/viscouscycles.f90): nv volumes with top
 * x(3,nv) (x(2), x(3) used), thickness, width and dip (radians), after
 * ucy_set_planestrain (which clears them).  Depths below zero give
 * UCY_ERR_GEOMETRY and dimensions not positive UCY_ERR_ARG (:1559-1576).
 * The state gains 6 entries per volume after the patches (e22, e23, e33,
 * s22, s23, s33; macros.f90:73-79); K gains three stress rows per volume
 * (s22', s23', s33' in the volume's frame) and three strain columns per
 * volume, each a strain of 1e-3 (greens.f90:66-226), so that K is
 * (2 np + 3 nv) x (np + 3 nv).  As greens.f90:212-216 passes them, the
 * volume rows of patch sources take the moduli exchanged (G = lambda,
 * lambda = mu).  ucy_set_rheology supplies the nonlinear Maxwell law per
 * volume, with the background strain rate in edot_bg[6 i + 3] (e22),
 * [6 i + 4] (e23), [6 i + 5] (e33) and Gm unused: eII = gammadot0
 * sII^(n-1) exp(-Q/(R To)) on the 2-D deviatoric stress, v = (e - e_bg) 1e3
 * (viscouscycles.f90:929-967); the initial stress flows at the background
 * rate under the same law (:777-810).  Displacement kernels gain three
 * columns per volume, unit e22' and e23' strains and a zero column for e33'
 * (the reference's eye there has a zero third row, greens.f90:270-273,
 * 334-344). */
int ucy_set_planestrain_volumes(ucy_ctx *ctx, int nv, const double *x, const double *thickness,
                                const double *width, const double *dip);
/* the plane-strain volumes' initial stress s0(3,nv) (s22, s23, s33): a
 * volume whose three values are not all zero starts from them instead of
 * the background-rate stress (viscouscycles.f90 of 2d/planestrain:779-810);
 * reset by ucy_set_planestrain and ucy_set_planestrain_volumes */
int ucy_set_initial_volume_stress(ucy_ctx *ctx, int nv, const double *s0);
/* computeTractionKernelsPlaneStrain (planestrain.f90:266-291): dip and normal
 * traction on one receiver (x, sv, dv, nv) of unit dip slip on ns sources */
int ucy_planestrain_traction_kernels(ucy_ctx *ctx, const double x[3], const double sv[3],
                                     const double dv[3], const double nv[3], int ns,
                                     const double *y, const double *W, const double *dip,
                                     double G, double lambda, double *td, double *tn);
/* computeTractionKernelsAntiplane (antiplane.f90:352-379) */
int ucy_antiplane_traction_kernels(ucy_ctx *ctx, const double x[3], const double sv[3],
                                   const double dv[3], const double nv[3], int ns,
                                   const double *y, const double *W, const double *dip,
                                   double G, double *ts);
/* Friction per patch in element order (rectangles then triangles), as the
 * deck's PATCH_ELEMENT_STRUCT (types.f90:35-40); rake in radians. */
int ucy_set_friction(ucy_ctx *ctx, int np, const double *Vpl, const double *rake,
                     const double *mu0, const double *sig, const double *a, const double *b,
                     const double *L, const double *Vo, const double *damping);
/* Nonlinear Maxwell rheology per volume (viscouscycles.f90:963-995, 769-803):
 * prefactor gammadot0, power n, background strain rate e(6,nv), and the
 * initial-stress law parameters Gm, Q, R, To. */
int ucy_set_rheology(ucy_ctx *ctx, int nv, const double *gammadot0, const double *npower,
                     const double *edot_bg, const double *Gm, const double *Q, const double *R,
                     const double *To);

/* ---- assembly: buildG (greens.f90:73-275) -------------------------------- */
/* K is always built and kept in binary64.  UCY_PREC_SPLIT_BF16 and
 * UCY_PREC_MIXED also store K as packed split-bf16 pairs (4 bytes per entry,
 * power-of-two row scales; no counterpart in the reference), which
 * ucy_integrate then streams: t_i = t_0 + K~ (v_i - v_0) at every stage, with
 * t_0 = K~ v_0 (split) or K v_0 in binary64 (mixed) at the step start.  The
 * per-routine entry points (ucy_odefun, ucy_rkck, ucy_rkqs, ucy_matvec_f64)
 * always use the binary64 K. */
int ucy_build_kernel(ucy_ctx *ctx, int precision);
/* t = K~ v with the stored split-bf16 K (host vectors; reduced builds only) */
int ucy_matvec_reduced(ucy_ctx *ctx, const double *v, double *t);
/* wall seconds of the last ucy_build_kernel by source type: rectangle,
 * triangle and strain-volume columns (each ends with a stream synchronize;
 * no counterpart in the reference: for measuring assembly) */
int ucy_build_seconds(const ucy_ctx *ctx, double *rect, double *tri, double *vol);
int ucy_kernel_shape(const ucy_ctx *ctx, int64_t *m, int64_t *n);
/* copy K (m x n row-major, binary64) back to the host */
int ucy_get_kernel(ucy_ctx *ctx, double *K);
/* load an externally assembled K (m x n row-major) */
int ucy_set_kernel(ucy_ctx *ctx, const double *K, int64_t m, int64_t n);
/* device pointer to K (binary64, row-major) for zero-copy callers */
const double *ucy_kernel_dev(const ucy_ctx *ctx);

/* ---- observation points: buildO (greens.f90:302-462) and exportPoints ---- */
/* Displacement kernels at npts observation points x(3,npts) (north, east,
 * down), built on the device and kept there: O is the physical (3 npts x n)
 * matrix, row 3k+c = NED component c at point k, columns as K's (the
 * reference's O(n, 3 npts)).  Rectangles: computeDisplacementOkada92
 * (okada92.f90:230-298, zero at a DC3D-singular fault edge); triangles:
 * computeDisplacementNikkhoo15 (nikkhoo15.f90:324-367); volumes:
 * computeDisplacementVerticalStrainVolume (strainvolume.f90:215-268, its
 * STOPs as UCY_ERR_GEOMETRY).  npts = 0 drops the kernels.  The antiplane
 * model (ucy_set_antiplane) has one displacement component per point and one
 * column per segment: computeDisplacementAntiplane (fortran/2d/antiplane/src/
 * antiplane.f90:178-191, greens.f90:275-383); rows = npts; there d = O u with
 * u the segments' slip, df = d, dl = 0. */
int ucy_build_displacement(ucy_ctx *ctx, int npts, const double *x);
/* shape of O (rows = 3 npts, n); 0 x 0 when not built */
int ucy_displacement_shape(const ucy_ctx *ctx, int64_t *rows, int64_t *n);
/* copy O back to the host, and form buildO's Of and Ol from it (each may be
 * NULL): Of = the fault columns plus the volume columns' third component
 * (greens.f90:445), Ol = the volume columns' first two components */
int ucy_get_displacement_kernel(ucy_ctx *ctx, double *O, double *Of, double *Ol);
/* exportPoints (ratestate.f90:414-486, viscouscycles.f90:523-625): d = O u
 * with u the strike and dip slip of each patch and e11..e33 of each volume,
 * gathered from the state y; df = Of u and dl = Ol u (may be NULL).  Host
 * vectors; _dev takes device pointers. */
int ucy_displacement(ucy_ctx *ctx, const double *y, double *d, double *df, double *dl);
int ucy_displacement_dev(ucy_ctx *ctx, const double *y_dev, double *d_dev, double *df_dev,
                         double *dl_dev);

/* ---- RHS, matvec and the Cash-Karp stepper ------------------------------ */
/* DGEMV("T") of ratestate.f90:719-720: t(m) = K v(n). */
int ucy_matvec_f64(ucy_ctx *ctx, const double *v, double *t);
int ucy_matvec_f64_dev(ucy_ctx *ctx, const double *v_dev, double *t_dev);
/* initStateVector (ratestate.f90:548-610; variant 1 = viscouscycles.f90:710-820) */
int ucy_init_state(ucy_ctx *ctx, int variant, double *y);
/* odefun(n,time,y,dydt) of ratestate.f90:624-778 (+ viscouscycles.f90:933-1080);
 * vmax receives the maximum slip velocity (may be NULL). */
int ucy_odefun(ucy_ctx *ctx, int n, double time, const double *y, double *dydt, double *vmax);
/* No counterpart in the reference: odefun for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and arithmetic as ucy_odefun, on device
 * buffers: y_dev (n entries) in, dydt_dev (n) out, vmax_dev (1 entry, may be
 * NULL) out.  It allocates nothing: the slip-rate vector and the vMax
 * accumulator live in a static 1 MiB device array per GPU, so the kernel may
 * have at most 131040 columns (the n of ucy_kernel_shape); larger models get
 * UCY_ERR_LIMIT.  The kernel must be built and friction/rheology set for
 * every element (the set-up calls place them on the device).  stream is a
 * hipStream_t: NULL runs on the context's stream and waits for the result;
 * otherwise the call only enqueues on that stream.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (other contexts or
 * streams) race. */
int ucy_odefun_dev_static(ucy_ctx *ctx, int n, const double *y_dev, double *dydt_dev,
                          double *vmax_dev, void *stream);
/* RKCKm(n,t,yin,dydt,h,yout,yerr) of ode45.f90:116-202 */
int ucy_rkck(ucy_ctx *ctx, int n, double t, const double *yin, const double *dydt, double h,
             double *yout, double *yerr);
/* RKQSm(n,t,y,dydt,yscal,htry,hdid,hnext) of ode45.f90:51-108; epsilon and
 * maximumTimeStep as the module variables of ode45.f90:40-43. */
int ucy_rkqs(ucy_ctx *ctx, int n, double *t, double *y, double *dydt, const double *yscal,
             double htry, double *hdid, double *hnext, double epsilon, double max_step);

/* Main loop of ratestate.f90:307-328 run device-resident: one graph-captured
 * RK45 attempt per replay, state kept in HBM.  y is read at entry and written
 * at exit.  Optional logs receive, per accepted step, the time reached, dt_done
 * and the vMax of the step-start odefun (ratestate.f90:530-532). */
typedef struct {
    double time;        /* simulated time reached */
    double dt_next;     /* next step size */
    int64_t steps;      /* accepted steps */
    int64_t attempts;   /* RK45 attempts (accepted + rejected) */
    int64_t rhs_evals;  /* odefun evaluations */
    double vmax;        /* vMax of the last accepted step's step-start odefun */
    double wall_seconds;
} ucy_run_stats;
int ucy_integrate(ucy_ctx *ctx, double *y, double interval, double dt0, int64_t max_steps,
                  double epsilon, double max_step, double *t_log, double *dt_log,
                  double *vmax_log, int64_t log_cap, ucy_run_stats *stats);
/* The same main loop from time t0 (the run continues the reference's running
 * sum time = time + dt_done, so a run split into calls accumulates time as one
 * run does), with the step history the learned step-size proposal is trained
 * on: per accepted step also errmax/epsilon of the accepted attempt and the
 * number of rejected attempts before it (errmax and rejects may be NULL), and
 * optionally observed state entries.  Logged times are absolute.  No
 * counterpart in the reference. */
typedef struct {
    double *t, *dt, *vmax, *errmax, *rejects;
    int64_t cap;
    /* observed state entries, the per-iteration export of ratestate.f90:
     * 488-536 (needs the step log above): n_obs indices into the state
     * vector; per logged step their values at the step start (obs_y) and the
     * step-start derivatives (obs_dydt), [cap][n_obs] each.  n_obs = 0: off */
    int64_t n_obs;
    const int64_t *obs_index;
    double *obs_y, *obs_dydt;
    /* observation-point displacement at each logged step's start (needs the
     * step log and ucy_build_displacement): d, df, dl of ucy_displacement,
     * [cap][3 npts] each; any may be NULL (all NULL: off) */
    double *disp, *disp_f, *disp_l;
    /* eMax of each logged step's start: the largest strain-rate norm
     * sqrt(e:e/2) over the strain volumes (viscouscycles.f90:986; 0 without
     * volumes), [cap]; NULL: off */
    double *emax;
} ucy_step_log;
int ucy_integrate_ex(ucy_ctx *ctx, double *y, double t0, double interval, double dt0,
                     int64_t max_steps, double epsilon, double max_step,
                     const ucy_step_log *log, ucy_run_stats *stats);

/* Measurement hook (no counterpart in the reference): average duration in
 * seconds of the fused matvec + odefun post-pass kernel (the dominant kernel
 * of every RHS), timed with HIP events on the context stream over iters
 * launches at the state of the last ucy_odefun call (or the last stage of
 * ucy_integrate). */
int ucy_time_rhs_kernel(ucy_ctx *ctx, int iters, double *avg_seconds);

/* ---- batched ensembles (no counterpart in the reference) --------------- */
/* Independent fault instances in geometry groups: each group shares one
 * kernel K; every instance has its own friction parameters, state, time
 * and step size.  Stage GEMMs run on MFMA with K stored as split bf16. */
typedef struct ucy_ensemble ucy_ensemble;
ucy_ensemble *ucy_ensemble_create(ucy_ctx *ctx, int precision);
void ucy_ensemble_destroy(ucy_ensemble *ens);
/* add a group: rectangles geometry + n_inst instances of friction (each array
 * n_inst*np, instance-major) and initial states y0 (n_inst*7*np); returns the
 * group index or <0 */
int ucy_ensemble_add_group(ucy_ensemble *ens, int np, const double *x, const double *length,
                           const double *width, const double *strike, const double *dip,
                           double lambda, double mu, int n_inst, const double *Vpl,
                           const double *rake, const double *a, const double *b,
                           const double *sig, const double *L, const double *damping,
                           const double *y0);
/* assemble every group's K on the device, convert it to the ensemble
 * precision and tune the stage GEMM tiling on the device */
int ucy_ensemble_build(ucy_ensemble *ens);
/* back to the initial states given to ucy_ensemble_add_group, t = 0 */
int ucy_ensemble_reset(ucy_ensemble *ens, double dt0);
typedef struct {
    double inst_time_sum;   /* sum over instances of simulated time reached */
    double min_time, max_time;
    int64_t steps;          /* accepted steps, all instances */
    int64_t attempts;       /* RK45 attempts, all instances */
    int64_t replays;        /* graph replays (lock-step attempts) */
    int64_t n_inst, n_failed;
    double wall_seconds;
} ucy_ensemble_stats;
/* advance every instance from its current state until t >= t_end or its
 * attempt count (cumulative since build/reset) reaches max_attempts; a later
 * call with a larger t_end or max_attempts continues the same run, so a
 * caller can advance in chunks.  Failed instances (step-size underflow,
 * non-finite error norm) stay stopped.  stats: cumulative counters, wall
 * time of this call. */
int ucy_ensemble_run(ucy_ensemble *ens, double t_end, int64_t max_attempts, double epsilon,
                     double max_step, ucy_ensemble_stats *stats);
/* Measurement option (off by default; the reference's loop steps past the
 * interval, ratestate.f90:321-326): clamp each instance's last step so that
 * it lands exactly on t_end, so that runs in different precisions can be
 * compared at the same time. */
int ucy_ensemble_set_stop_at_end(ucy_ensemble *ens, int on);
/* Event recording: per instance, the start time of the first accepted step
 * whose step-start vMax exceeds vmax_threshold (default 1e-3 m/s; the vMax of
 * ratestate.f90:530-532).  With stop_at_event an instance stops there. */
int ucy_ensemble_set_events(ucy_ensemble *ens, double vmax_threshold, int stop_at_event);
/* first-event times of a group's instances (-1: no event yet) */
int ucy_ensemble_first_events(ucy_ensemble *ens, int group, double *t_event);
/* per-instance status: 0 running or finished, 1 step-size underflow
 * (ode45.f90:86-90), 2 non-finite error norm on four attempts in a row (an
 * attempt with a non-finite norm is rejected with a x0.1 shrink) */
int ucy_ensemble_status(ucy_ensemble *ens, int group, int *status);
/* per-step log of a group's accepted steps (the ensemble form of the step
 * logs of ucy_integrate_ex; no counterpart in the reference): cap steps per
 * instance from the last reset, cap = 0 turns it off.  get_step_log writes
 * [ninst][cap_out] arrays (any may be null; at most cap_out steps per
 * instance): time after the step, accepted h, step-start vMax,
 * errmax / epsilon, rejected attempts of the step; count[i] = steps written
 * for instance i */
int ucy_ensemble_set_step_log(ucy_ensemble *ens, int group, long long cap);
int ucy_ensemble_get_step_log(ucy_ensemble *ens, int group, long long cap_out, double *t,
                              double *dt, double *vmax, double *errmax, double *rejects,
                              long long *count);
/* y: n_inst x (7*np) in the reference layout; per-instance time, next step,
 * accepted steps and attempts (any output may be NULL) */
int ucy_ensemble_get_state(ucy_ensemble *ens, int group, double *y, double *time,
                           double *dt_next, int64_t *steps, int64_t *attempts);
/* one stage matvec of a group in the ensemble precision: V is n x n_inst
 * (row-major, column = instance), T is m x n_inst */
int ucy_ensemble_matvec(ucy_ensemble *ens, int group, const double *V, double *T);
/* the autotuner's choice for the stage GEMM: k-blocks (of 16) per split and
 * the bf16 kernel variant (0: 64 K rows per wave; 1: 64 rows with the next
 * k-block's loads issued ahead; 2: 128 rows, ahead; 3: 128 rows with the V
 * tiles staged in LDS per block in chunks of 4 k-blocks; 4: 64 rows, ahead,
 * V staged in chunks of 8 k-blocks; 5: 128 rows; 6: 64 rows, 7: 128 rows,
 * with every load a direct global->LDS load, K per wave and V per block,
 * 4 (6) or 3 (7) k-blocks deep; 8: 128 rows, K tiles per wave by direct
 * global->LDS loads and V in registers, 3 k-blocks deep, no barrier; 6-8
 * bf16+i8 only; 9: 128 rows in two register sets used in turn; 10: 128
 * rows per block of 4 waves, one K-family slice per wave, each K tile loaded
 * and decoded once per block into LDS), and its table of seconds
 * per launch (UCY_GEMM_TUNE_ENTRIES entries: k-blocks 32, 48, 64, 80, 96,
 * 104, 128, 160, 192, 256, 512 of variant 0, then of 1, ...; -1 = not
 * admissible); pointers may be NULL.  The tuner takes the candidate with the
 * least launch time plus the post-passes' read of its split partial
 * products charged at 4 TB/s. */
#define UCY_GEMM_TUNE_ENTRIES 121
/* override the autotuner: k-blocks per split and split-bf16 variant (as in
 * ucy_ensemble_tuning; binary64 ensembles take variant 0 only) */
int ucy_ensemble_set_tiling(ucy_ensemble *ens, int kb_per_split, int variant);
int ucy_ensemble_tuning(const ucy_ensemble *ens, int *kb_per_split, int *variant,
                        double *gemm_seconds);
/* HIP-event timing of the all-groups stage GEMM (seconds per launch) and its
 * algorithmic HBM bytes per launch */
int ucy_ensemble_time_gemm(ucy_ensemble *ens, int iters, double *avg_seconds, double *bytes);
/* HIP-event timing of the phases of an RK45 attempt, launched one by one
 * (not from the captured graph): phase_seconds[UCY_ENS_PHASES] receives the
 * summed seconds of each phase over `attempts` attempts, in execution order:
 * start, gemm 1, post 1, gemm 2, post 2, gemm 3, post 3, gemm 4, post 4,
 * gemm 5, post 5, gemm 6, finish, control.  Every instance advances by
 * `attempts` attempts (no end time), as run() would advance it. */
#define UCY_ENS_PHASES 14
int ucy_ensemble_time_phases(ucy_ensemble *ens, int attempts, double *phase_seconds);

/* ---- learned RK45 step-size proposal (no counterpart in the reference) --- */
/* An 8 -> 64 -> 64 -> 1 tanh perceptron (binary32, 4801 parameters) that
 * proposes log10 of the next step from per-instance step history; trained
 * on the plain integrator's step logs (ucy_integrate_ex) by hand-written
 * forward/backward kernels with Adam.  In an ensemble it replaces the growth
 * rule of RKQSm on acceptance; the embedded error estimate still accepts or
 * rejects, and a rejection shrinks with the classic rule (ode45.f90:82). */
#define UCY_MLP_INPUTS 8
#define UCY_MLP_PARAMS 4801
typedef struct ucy_mlp ucy_mlp;
ucy_mlp *ucy_mlp_create(ucy_ctx *ctx, uint64_t seed);   /* Glorot-uniform init */
void ucy_mlp_destroy(ucy_mlp *mlp);
int ucy_mlp_set_params(ucy_mlp *mlp, const float *theta);
int ucy_mlp_get_params(ucy_mlp *mlp, float *theta);
/* feature and target normalisation: x_n = (x - mean) / scale */
int ucy_mlp_set_normalization(ucy_mlp *mlp, const float *x_mean, const float *x_scale,
                              float y_mean, float y_scale);
/* raw features x (n x 8) -> raw outputs y (n) */
int ucy_mlp_forward(ucy_mlp *mlp, int64_t n, const float *x, float *y);
/* mean squared error of the normalised outputs over n samples and its
 * gradient with respect to the 4801 parameters */
int ucy_mlp_loss_grad(ucy_mlp *mlp, int64_t n, const float *x, const float *t, double *loss,
                      float *grad);
/* Adam over shuffled minibatches; loss_log (epochs entries, may be NULL) */
int ucy_mlp_train(ucy_mlp *mlp, int64_t n, const float *x, const float *t, int epochs,
                  int batch, float lr, uint64_t seed, double *loss_log);
/* use (a copy of) mlp for the step proposals of an ensemble; NULL restores
 * the classic controller */
int ucy_ensemble_set_mlp(ucy_ensemble *ens, const ucy_mlp *mlp);

#ifdef __cplusplus
}
#endif
#endif
