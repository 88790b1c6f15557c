// ucy_internal.h -- structures shared between the host-side C ABI and the
// HIP kernels of the plain (single-instance, binary64) path.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "okada_dev.h"

namespace ucy {

// Receiver description: centre, frame and first row of K it owns.
struct Receiver {
    double xc[3], sv[3], dv[3], nv[3];
};

// Per-patch friction constants for the RHS (types.f90:35-40).  vplc/vpls are
// Vpl*cos(rake) and Vpl*sin(rake) formed on the host with the C library, the
// same product the reference forms at ratestate.f90:701-702.
struct Friction {
    double a, b, sig, L, damping, vplc, vpls, pad;
};

// Per-volume rheology for the RHS (viscouscycles.f90:963-995).  The 2-D
// antiplane volumes (fortran/2d/antiplane/src/viscouscycles.f90:846-877) use
// ebg[1], ebg[2] (e12, e13), Gm and arr = exp(-Q/(R To)), formed on the host.
struct Rheology {
    double gammadot0, npower;
    double ebg[6];
    double Gm, arr;
};

// Device-side control block of the graph-captured integrator.
struct Control {
    double time;        // simulated time (ratestate.f90:176,321)
    double h;           // step size of the next attempt (dt_try, or shrunk h)
    double hnext;       // dt_next of the last accepted step
    double hdid;        // dt_done of the last accepted step
    double interval;    // end time (ratestate.f90:324)
    double epsilon;     // ode45.f90:40
    double max_step;    // ode45.f90:43
    unsigned long long errmax_bits;  // max |yerr/yscal| as ordered bits
    double vmax;        // max slip velocity of the current step-start odefun
    double last_vmax;   // same, for the last accepted step
    double emax;        // max strain-rate norm of the volumes, same odefun
                        // (viscouscycles.f90:986)
    int need_rhs;       // 1: next attempt starts a new step
    int accepted;       // 1: last attempt accepted, y <- ycand pending
    int done;
    int error;          // 0 ok, 1 underflow, 2 non-finite error norm
    long long steps, attempts, max_steps, rhs_evals;
    long long log_count, log_cap;
    int step_rejects;   // rejected attempts of the current step
    int obs_pending;    // set by the controller on acceptance, cleared at the next
                        // attempt's start: the observed entries are logged once per step
};

struct PlainDims {
    int np;      // patches
    int nv;      // volumes
    int dim;     // 7*np + 12*nv (4*(np + nv) for the antiplane model)
    int m, n;    // K rows, columns
    int ap;      // 1: the 2-D antiplane model (fortran/2d/antiplane: 4 state
                 // entries per element; per patch one slip rate and one
                 // traction row, per volume two strain rates e12, e13 and two
                 // stress rows s12, s13)
    int ps;      // 1: the 2-D plane-strain model (fortran/2d/planestrain: 5
                 // state entries per patch, one dip-slip rate, two traction
                 // rows: dip and normal; per volume 6 state entries e22, e23,
                 // e33, s22, s23, s33, three strain rates and three stress rows)
};

// ---- assembly launchers (k_assemble.hip) ----
// rows for receivers [r0, r0+nr) from all rectangle sources; kind 0 traction
// (3 rows per receiver), kind 1 volume-frame stress (6 rows per receiver).
void launch_okada_rows(hipStream_t st, const OkadaMedium &M, const OkadaSource *src, int ns,
                       const Receiver *rcv, int nr, int kind, double *K, int64_t ldk,
                       int64_t row0, int64_t col0, double *s_out, double *d_out, int out_stride);

// displacement rows 3k..3k+2 of O for points k in [0, npts) (x (3,npts)) from all
// rectangle sources, columns col0 + 2s + {0,1}
void launch_okada_disp(hipStream_t st, const OkadaMedium &M, const OkadaSource *src, int ns,
                       const double *pts, int npts, double *O, int64_t ldo, int64_t col0);

// antiplane buildG rows for receivers [0, nr) from ns segment sources
// src[3s..3s+2] = (q2, q3, W); out (one receiver) may replace K
void launch_ap_rows(hipStream_t st, const double *src, int ns, const Receiver *rcv, int nr,
                    double G, double *K, int64_t ldk, double *out);
// antiplane buildO: row k = point k (x(3,npts), x2 and x3 used), column s
void launch_ap_disp(hipStream_t st, const double *src, int ns, const double *pts, int npts,
                    double *O, int64_t ldo);
// plane-strain sources (fortran/2d/planestrain): top (q2, q3), width W and the
// sines and cosines of 1..4 times the dip, formed on the host
struct PsSource {
    double q2, q3, W;
    double sk[4], ck[4];
};
// plane-strain buildG rows 2r, 2r+1 (dip, normal traction) for receivers
// [0, nr) from ns sources; td, tn (one receiver) may replace K
void launch_ps_rows(hipStream_t st, const PsSource *src, int ns, const Receiver *rcv, int nr,
                    double G, double lam, double nu, double *K, int64_t ldk, double *td,
                    double *tn);
// plane-strain buildO rows 2k, 2k+1 (east, down) for points [0, npts)
void launch_ps_disp(hipStream_t st, const PsSource *src, int ns, const double *pts, int npts,
                    double G, double lam, double nu, double *O, int64_t ldo);
// the plane-strain buildG blocks that involve strain volumes (receivers: np
// patches then nv volumes in rcv; volume sources vsrc with thicknesses vT):
// patch rows <- volume columns, volume rows <- patch columns (G and lambda
// exchanged, as the reference passes them), volume rows <- volume columns
void launch_ps_blocks(hipStream_t st, const PsSource *psrc, int np, const PsSource *vsrc,
                      const double *vT, int nv, const Receiver *rcv, double G, double lam,
                      double *K, int64_t ldk);
// plane-strain buildO columns col0 + 3v + j of strain volumes
void launch_ps_volume_disp(hipStream_t st, const PsSource *vsrc, const double *vT, int nv,
                           const double *pts, int npts, double G, double lam, double *O,
                           int64_t ldo, int64_t col0);
// antiplane buildG blocks that involve strain volumes: every receiver (np
// patches, then nv volumes) against the volume sources vol[4q..4q+3] =
// (q2, q3, W, T), and the volume receivers against the segments seg
void launch_ap_volume_blocks(hipStream_t st, const double *seg, int np, const double *vol, int nv,
                             const Receiver *rcv, double G, double *K, int64_t ldk);
// antiplane buildO volume columns: O[k][2q + j], unit e12 (j = 0), e13 (j = 1)
void launch_ap_volume_disp(hipStream_t st, const double *vol, int nv, const double *pts, int npts,
                           double *O, int64_t ldo);

// ---- GEMV / RHS launchers (k_ode.hip) ----
void launch_matvec(hipStream_t st, const double *K, int m, int n, const double *v, double *t);
// odefun building blocks
void launch_prepass(hipStream_t st, const PlainDims &d, const Friction *fr, const Rheology *rh,
                    const double *y, double *dydt, double *v, double *vmax_out);
void launch_gemv_post(hipStream_t st, const PlainDims &d, const double *K, const Friction *fr,
                      const double *y, const double *v, double *dydt, double *t_out);
// integrator graph pieces
struct RkBuffers {
    double *y, *dydt, *yscal, *ytmp, *ak[7], *ycand, *yerr, *yr, *v, *t;
    double *v2;            // slip rates of the even stages (the attempt graph double-buffers v)
    // observed state entries logged per accepted step (null: off)
    const int64_t *obs_idx;
    int64_t n_obs;
    double *obs_y, *obs_dydt;
    // observation-point displacement logged per accepted step (disp null: off):
    // O (o_rows x o_n, row-major) times the slip / strain entries of y
    const double *O;
    int64_t o_rows, o_n;
    double *disp, *disp_f, *disp_l;
    // reduced-precision K (null in binary64): slip rates and tractions of the
    // step start; the stage matvecs then take v - v0 and add t0
    double *v0, *t0;
    double *tlog, *dtlog, *vlog;
    double *elog, *rlog;   // errmax/epsilon of the accepted attempt, rejections (optional)
    double *emlog;         // eMax of the step-start odefun (optional)
};
// the stored K of a context: binary64 always; packed split-bf16 pairs
// (hi << 16 | lo, row-major, row stride n) with power-of-two row scales when
// built in reduced precision
struct KOperand {
    const double *K;
    const uint32_t *Kp;
    const double *rs;
    int prec;   // UCY_PREC_F64, UCY_PREC_SPLIT_BF16 or UCY_PREC_MIXED
};
void launch_split_rowmajor(hipStream_t st, const double *K, int64_t m, int64_t n, double *rs,
                           uint32_t *Kp);
void launch_matvec_split(hipStream_t st, const uint32_t *Kp, const double *rs, int m, int n,
                         const double *v, double *t);
void launch_gemv_post_inc(hipStream_t st, const PlainDims &d, const KOperand &K, const Friction *fr,
                          const double *y, const double *v, double *dydt, double *t0);
// the attempt graph (k_ode.hip): start, rhs_stage 1..5, rhs_finish, controller
void launch_start(hipStream_t st, const PlainDims &d, const Friction *fr, const Rheology *rh,
                  const RkBuffers &b, Control *ctl);
void launch_rhs_stage(hipStream_t st, const PlainDims &d, const KOperand &K, const Friction *fr,
                      const Rheology *rh, const RkBuffers &b, Control *ctl, int stage);
void launch_rhs_finish(hipStream_t st, const PlainDims &d, const KOperand &K, const Friction *fr,
                       const RkBuffers &b, Control *ctl);
void launch_controller(hipStream_t st, const RkBuffers &b, Control *ctl);
void launch_obs_log(hipStream_t st, const RkBuffers &b, Control *ctl);
// observation-point displacement d = O u with u the slip (2 per patch) and
// strain (6 per volume) entries of y (ap: the slip, entry 0 of 4 per patch); df, dl (may be null): the parts of buildO's
// Of and Ol.  With ctl: written at log row log_count - 1 of the arrays, only in
// the attempt that accepted a step.
void launch_disp_gemv(hipStream_t st, const double *O, int64_t rows, int64_t n, int np, int model,
                      const double *y, double *d, double *df, double *dl, const Control *ctl);
// standalone Cash-Karp combination kernels for the ucy_rkck entry point
void launch_rk_combine(hipStream_t st, int n, int stage, double h, const double *yin,
                       double *const *ak, double *yr, double *ytmp, double *yout, double *yerr);

}  // namespace ucy
