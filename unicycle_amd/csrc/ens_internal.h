// ens_internal.h -- batched ensembles of independent fault instances.
//
// No counterpart in the reference: every instance runs the reference's
// per-instance algorithm (odefun ratestate.f90:624-778, RKQSm/RKCKm
// ode45.f90:51-202, main loop ratestate.f90:307-328) with its own friction,
// state, time and step size; instances of one geometry group share the
// kernel K, so the traction matvecs of all instances of a group become one
// GEMM  T(m x B) = K(m x n) V(n x B).
//
// Storage (per group, B = 64 instance columns, padded):
//   state vectors  SoA  X[(comp*N + e)*64 + inst]   (instance fastest)
//   K split bf16   pre-tiled in MFMA operand order, see k_ensemble.hip
//   row scales     exact powers of two, double
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "mlp_dev.h"

namespace ucy {

constexpr int kEnsCols = 64;     // instance columns per group (padded)
constexpr int kMaxGroups = 16;
// a geometry with more than 64 instances is stored as up to kMaxSlices
// consecutive groups ("slices") that share one K: the stage GEMM streams the
// family's K once for all its slices
constexpr int kMaxSlices = 4;
constexpr int kMaxRep = 16;      // replicas of the per-instance maxima

struct EnsFriction {             // per (element, instance), SoA arrays [e*64+inst]
    // a sig and b sig as products (the right-hand side uses a, b and sig
    // only there, multiplied first: same roundings, one read fewer)
    double *asig, *bsig, *L, *damping, *vplc, *vpls;
};

struct EnsGroup {
    int N;            // patches
    int ninst;        // live instances (<= 64)
    int m, n;         // K rows (3N), columns (2N)
    int mp, np;       // padded to 64 rows / 16 columns
    int nkb;          // np / 16
    int ksplit;       // split-K factor of the GEMM
    int kb_per_split; // k-blocks per split
    int mode;         // 0 split-bf16, 1 binary64, 2 mixed (see below), 3 bf16+i8
                      // (as 0 with the 3-byte K tiles)
    const uint16_t *Ktiles;  // split-bf16 tiles (modes 0, 2), bf16+i8 tile bytes (mode 3)
    const double *rscale;    // row scales (modes 0, 2)
    const double *K64;       // binary64 K tiles (modes 1, 2)
    uint16_t *Vtiles;        // V split-bf16 tiles (modes 0, 2)
    double *V64;             // V binary64 tiles (modes 1, 2)
    float *Tpart;            // [ksplit][mp][64] partial products (modes 0, 2)
    double *T64;             // [ksplit][mp][64] partial products (modes 1, 2)
    // modes 0 and 2 evaluate the stage matvecs as increments over the step
    // start:  t_i = t_0 + K (v_i - v_0),  t_0 = K v_0 (split bf16 in mode 0,
    // binary64 in mode 2), the increment always split bf16
    double *v0;              // [2N][64] slip-rate deviations at the step start
    double *t0;              // [mp][64] tractions at the step start
    int *any_rhs;            // some instance of the group starts a step
    // state (SoA, 7 components)
    double *y, *dydt, *ytmp, *ak[7], *ycand;   // the embedded error is only reduced
    EnsFriction fr;
    // per-instance control
    double *time, *h, *hnext, *hdid, *errmax, *vmax;
    // [kMaxRep][64] partial maxima of the step-start vMax (start kernel) and
    // of the attempt's error norm (finish kernel), one atomic per block into
    // replica blockIdx % kMaxRep; the control kernel folds them into vmax,
    // errmax (one address per instance would serialise thousands of atomics)
    double *vmax_rep, *errmax_rep;
    double *hstart;          // h of the step's first attempt: the dt_try of yscal
    double *t_event;         // time of the first step starting with vMax > event_vmax (-1: none)
    double *hist;            // [5][64] proposal history: log10 vMax, h, log10 errmax of the
                             // last accepted step (h = 0: none yet), rejections of this step,
                             // consecutive attempts with a non-finite error norm
    double *slog;            // optional per-step log [cap][5][64]: time after the step, h,
                             // step-start vMax, errmax / epsilon, rejections (null: off)
    long long slog_cap;
    int *need_rhs, *accepted, *done, *status;
    long long *steps, *attempts;
};

struct EnsTable {
    int ngroups;
    EnsGroup g[kMaxGroups];
    // K families: family f is groups fam_head[f] .. fam_head[f] + fam_slices[f] - 1,
    // which share the head's K tiles (and row scales)
    int nfam;
    int fam_head[kMaxGroups];
    int fam_slices[kMaxGroups];
    int blk_gemm[kMaxGroups + 1];   // prefix sums of split-bf16 GEMM blocks per family
    int blk_gemm64[kMaxGroups + 1]; // prefix sums of binary64 GEMM blocks per family
    int variant;                    // split-bf16 GEMM variant (rows per wave, pipelining)
    int blk_elem[kMaxGroups + 1];   // prefix sums of elementwise blocks
    int blk_inst[kMaxGroups + 1];   // prefix sums of per-instance blocks
    double t_end, epsilon, max_step;
    long long max_attempts;
    int stop_at_end;                // clamp the last step to land on t_end
    int stop_at_event;              // an instance stops at its first event
    double event_vmax;              // slip-velocity threshold of an event (m/s)
    const float *mlp;               // step-size proposal parameters (null: classic control)
    MlpNorm mlp_norm;
    int *all_done;                  // count of finished instances
    int total_inst;
};

// split-bf16 stage-GEMM variants of the autotuner: 0 = 64 rows per wave,
// 1 = 64 rows with the next k-block's loads issued ahead, 2 = 128 rows, ahead,
// 3 = 128 rows with V staged in LDS per block in chunks of 4 k-blocks,
// 4 = 64 rows, ahead, V staged in chunks of 8 k-blocks, 5 = 128 rows (200
// VGPRs; 128 rows ahead with V staged in chunks of 2 needs 264-272, one wave
// per SIMD, and measured 461 us against 265 us: not kept)
// 6 = 64 rows per wave, 7 = 128 rows: every load is a direct global->LDS
// load, each wave's K tiles in its own LDS ring and V in a ring shared by the
// block, 4 (6) or 3 (7) k-blocks deep (bf16+i8 only)
// 8 = 128 rows, wave-shaped as 0-5, K tiles by direct global->LDS loads
// into a ring private to the wave and V into a register ring, 3 k-blocks
// deep, no barrier (bf16+i8 only); 9 = 128 rows in two register sets used
// in turn, the loads of one k-block out while the other is multiplied;
// 10 = 128 rows, block-shaped for K families: one slice per wave, each K
// tile loaded and decoded once per block into LDS for the block's slices
constexpr int kGemmVariants = 11;
int ens_gemm_variant_rows(int variant);
int ens_gemm_variant_lds(int variant);
int ens_gemm_variant_dma(int variant);
int ens_gemm_variant_kshare(int variant);

// launchers (k_ensemble.hip)
void ens_launch_split_K(hipStream_t st, const double *K, int m, int n, int mp, int np, uint16_t *tiles,
                        double *rscale);
void ens_launch_split3_K(hipStream_t st, const double *K, int m, int n, int mp, int np,
                         uint8_t *tiles, double *rscale);
void ens_launch_tile_K64(hipStream_t st, const double *K, int m, int n, int mp, int np,
                         double *tiles);
// stage 1 = the step-start matvec (skipped when no instance of a group starts
// a step), 2..6 the stage matvecs, 0 a standalone matvec
void ens_launch_gemm(hipStream_t st, const EnsTable *dtab, const EnsTable &htab, int stage);
// one attempt = start (commit of the previous accepted step + step-start
// pre-pass), gemm 1, post_stage 1, gemm 2, ..., post_stage 5, gemm 6,
// post_finish, control; a run ends with one commit
void ens_launch_start(hipStream_t st, const EnsTable *dtab, const EnsTable &htab);
void ens_launch_post_stage(hipStream_t st, const EnsTable *dtab, const EnsTable &htab, int stage);
void ens_launch_post_finish(hipStream_t st, const EnsTable *dtab, const EnsTable &htab);
void ens_launch_control(hipStream_t st, const EnsTable *dtab, const EnsTable &htab);
void ens_launch_commit(hipStream_t st, const EnsTable *dtab, const EnsTable &htab);
void ens_launch_rearm(hipStream_t st, const EnsTable *dtab, const EnsTable &htab);
// standalone GEMM check: T(m x 64) = K V with the group's stage precision

}  // namespace ucy
