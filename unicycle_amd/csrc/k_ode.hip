// k_ode.hip -- rate-and-state right-hand side, the traction matvec and the
// Cash-Karp RK45 attempt, device-resident for graph capture.
//
// Reference: jdpmoore/unicycle fortran/3d/src
//   ratestate.f90:624-778   odefun (pre-pass, DGEMV, post-pass)
//   viscouscycles.f90:933-1080  volume branch of odefun
//   ode45.f90:51-202        RKQSm / RKCKm
//   ratestate.f90:307-328   main loop (yscal, time, exit test)
//
// Every kernel of the attempt reads the Control block first and returns when
// the run is done, so a captured attempt can be replayed any number of times.
#include <hip/hip_runtime.h>
#include "../../include/ucy.h"
#include "ucy_internal.h"

namespace ucy {

// ratestate.f90:130 lg10=LOG(1.d1)
__constant__ const double kLg10 = 2.302585092994045684;

// Cash-Karp tableau (ode45.f90:126-134)
#define B21 0.2
#define B31 (3.0 / 40.0)
#define B32 (9.0 / 40.0)
#define B41 0.3
#define B42 (-0.9)
#define B43 1.2
#define B51 (-11.0 / 54.0)
#define B52 2.5
#define B53 (-70.0 / 27.0)
#define B54 (35.0 / 27.0)
#define B61 (1631.0 / 55296.0)
#define B62 (175.0 / 512.0)
#define B63 (575.0 / 13824.0)
#define B64 (44275.0 / 110592.0)
#define B65 (253.0 / 4096.0)
#define C1 (37.0 / 378.0)
#define C3 (250.0 / 621.0)
#define C4 (125.0 / 594.0)
#define C6 (512.0 / 1771.0)
#define DC1 (C1 - 2825.0 / 27648.0)
#define DC3 (C3 - 18575.0 / 48384.0)
#define DC4 (C4 - 13525.0 / 55296.0)
#define DC5 (-277.0 / 14336.0)
#define DC6 (C6 - 0.25)

__device__ __forceinline__ unsigned long long dbits(double x)
{
    return (unsigned long long)__double_as_longlong(x);
}

// ---------------------------------------------------------------------------
// pre-pass for one patch (ratestate.f90:683-702) and one volume
// (viscouscycles.f90:940-995)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double patch_prepass(const double *p, const Friction &f, double *dp,
                                                double *vp)
{
    const double rake = atan2(p[3], p[2]);
    const double velocity = exp(p[6] * kLg10);
    const double vs = velocity * cos(rake), vd = velocity * sin(rake);
    dp[0] = vs;
    dp[1] = vd;
    vp[0] = vs - f.vplc;
    vp[1] = vd - f.vpls;
    return velocity;
}

// 2-D antiplane patch (fortran/2d/antiplane/src/ratestate.f90:660-677): slip
// rate V, v = V - Vpl (vplc holds Vpl)
__device__ __forceinline__ double ap_prepass(const double *p, const Friction &f, double *dp,
                                             double *vp)
{
    const double velocity = exp(p[3] * kLg10);
    dp[0] = velocity;
    vp[0] = velocity - f.vplc;
    return velocity;
}

// ratestate.f90 (2-D antiplane):698-722
__device__ __forceinline__ void ap_postpass(const double *p, const Friction &f, double t,
                                            double *dp)
{
    const double velocity = exp(p[3] * kLg10);
    dp[2] = (exp(-p[2] * kLg10) - velocity / f.L) / kLg10;
    const double dtau = t;
    dp[3] = (dtau - f.b * f.sig * dp[2] * kLg10) / (f.a * f.sig + f.damping * velocity) / kLg10;
    dp[1] = dtau - f.damping * velocity * dp[3] * kLg10;
}

// 2-D plane-strain patch (fortran/2d/planestrain/src/ratestate.f90:692-712):
// dip-slip rate V, v = V - Vpl (vplc holds Vpl)
__device__ __forceinline__ double ps_prepass(const double *p, const Friction &f, double *dp,
                                             double *vp)
{
    const double velocity = exp(p[4] * kLg10);
    dp[0] = velocity;
    vp[0] = velocity - f.vplc;
    return velocity;
}

// ratestate.f90 (2-D plane strain):723-750: the dip traction drives the
// friction law at the constant normal stress sig; the normal traction rate is
// integrated as is
__device__ __forceinline__ void ps_postpass(const double *p, const Friction &f, const double *t,
                                            double *dp)
{
    const double velocity = exp(p[4] * kLg10);
    dp[3] = (exp(-p[3] * kLg10) - velocity / f.L) / kLg10;
    const double dtau = t[0];
    const double sigma = f.sig;
    dp[4] = (dtau - f.b * sigma * dp[3] * kLg10) / (f.a * sigma + f.damping * velocity) / kLg10;
    dp[1] = dtau - f.damping * velocity * dp[4] * kLg10;
    dp[2] = t[1];
}

// 2-D antiplane volume (fortran/2d/antiplane/src/viscouscycles.f90:846-877):
// state [e12, e13, s12, s13]; strain rate eII s with
// eII = gammadot0 (sII/Gm)^(n-1) / Gm exp(-Q/(R To)), v = that less the
// background rate; returns the strain-rate norm (eMax of :869)
__device__ __forceinline__ double ap_volume_prepass(const double *q, const Rheology &r, double *dq,
                                                    double *vq)
{
    const double s12 = q[2], s13 = q[3];
    const double sII = sqrt(s12 * s12 + s13 * s13);
    const double eII = r.gammadot0 * pow(sII / r.Gm, r.npower - 1.0) / r.Gm * r.arr;
    const double e12 = eII * s12, e13 = eII * s13;
    dq[0] = e12;
    dq[1] = e13;
    vq[0] = e12 - r.ebg[1];
    vq[1] = e13 - r.ebg[2];
    return sqrt(e12 * e12 + e13 * e13);
}

// 2-D plane-strain volume (fortran/2d/planestrain/src/viscouscycles.f90:
// 929-967): state [e22, e23, e33, s22, s23, s33]; with the 2-D deviatoric
// stress s' (p = (s22 + s33) / 2), strain rate eII s' with
// eII = gammadot0 sII^(n-1) exp(-Q/(R To)); v = (that - the background
// rate) 1e3, as K's volume columns are strains of 1e-3; returns the
// strain-rate norm (eMax of :962)
__device__ __forceinline__ double ps_volume_prepass(const double *q, const Rheology &r, double *dq,
                                                    double *vq)
{
    const double s22 = q[3], s23 = q[4], s33 = q[5];
    const double p = (s22 + s33) / 2.0;
    const double s22p = s22 - p, s33p = s33 - p;
    const double sII = sqrt((s22p * s22p + 2.0 * (s23 * s23) + s33p * s33p) / 2.0);
    const double eII = r.gammadot0 * pow(sII, r.npower - 1.0) * r.arr;
    const double e22 = eII * s22p, e23 = eII * s23, e33 = eII * s33p;
    dq[0] = e22;
    dq[1] = e23;
    dq[2] = e33;
    vq[0] = (e22 - r.ebg[3]) * 1e3;
    vq[1] = (e23 - r.ebg[4]) * 1e3;
    vq[2] = (e33 - r.ebg[5]) * 1e3;
    return sqrt((e22 * e22 + 2.0 * (e23 * e23) + e33 * e33) / 2.0);
}

__device__ __forceinline__ void volume_prepass(const double *q, const Rheology &r, double *dq,
                                               double *vq)
{
    double s11 = q[6], s12 = q[7], s13 = q[8], s22 = q[9], s23 = q[10], s33 = q[11];
    const double p = (s11 + s22 + s33) / 3.0;
    s11 = s11 - p;
    s22 = s22 - p;
    s33 = s33 - p;
    const double sII = sqrt((s11 * s11 + 2.0 * (s12 * s12) + 2.0 * (s13 * s13) + s22 * s22
                             + 2.0 * (s23 * s23) + s33 * s33) / 2.0);
    const double eII = r.gammadot0 * pow(sII, r.npower - 1.0);
    const double e[6] = {0.0 + eII * s11, 0.0 + eII * s12, 0.0 + eII * s13,
                         0.0 + eII * s22, 0.0 + eII * s23, 0.0 + eII * s33};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        dq[k] = e[k];
        vq[k] = e[k] - r.ebg[k];
    }
}

// post-pass for one patch (ratestate.f90:747-771)
__device__ __forceinline__ void patch_postpass(const double *p, const Friction &f, const double *t,
                                               double *dp)
{
    const double rake = atan2(p[3], p[2]);
    const double velocity = exp(p[6] * kLg10);
    const double cr = cos(rake), sr = sin(rake);
    dp[5] = (exp(-p[5] * kLg10) - velocity / f.L) / kLg10;
    const double dtau = t[0] * cr + t[1] * sr;
    dp[6] = (dtau - f.b * f.sig * dp[5] * kLg10) / (f.a * f.sig + f.damping * velocity) / kLg10;
    const double correction = f.damping * velocity * dp[6] * kLg10;
    dp[2] = t[0] - correction * cr;
    dp[3] = t[1] - correction * sr;
    dp[4] = t[2];
}

__device__ __forceinline__ int elem_state_off(const PlainDims &d, int e)
{
    if (d.ps) return e < d.np ? 5 * e : 5 * d.np + 6 * (e - d.np);
    if (d.ap) return 4 * e;
    return e < d.np ? 7 * e : 7 * d.np + 12 * (e - d.np);
}
__device__ __forceinline__ int elem_v_off(const PlainDims &d, int e)
{
    if (d.ps) return e < d.np ? e : d.np + 3 * (e - d.np);
    if (d.ap) return e < d.np ? e : d.np + 2 * (e - d.np);
    return e < d.np ? 2 * e : 2 * d.np + 6 * (e - d.np);
}
__device__ __forceinline__ int elem_row_off(const PlainDims &d, int e)
{
    if (d.ps) return e < d.np ? 2 * e : 2 * d.np + 3 * (e - d.np);
    if (d.ap) return e < d.np ? e : d.np + 2 * (e - d.np);
    return e < d.np ? 3 * e : 3 * d.np + 6 * (e - d.np);
}
// state entries and slip-rate entries of element e
__device__ __forceinline__ int elem_nstate(const PlainDims &d, int e)
{
    return d.ps ? (e < d.np ? 5 : 6) : d.ap ? 4 : (e < d.np ? 7 : 12);
}
__device__ __forceinline__ int elem_nv(const PlainDims &d, int e)
{
    return d.ps ? (e < d.np ? 1 : 3) : d.ap ? (e < d.np ? 1 : 2) : (e < d.np ? 2 : 6);
}

// ---------------------------------------------------------------------------
// wave-level dot products of NR rows of K with v; one wavefront per element
// ---------------------------------------------------------------------------
// K is streamed once per RHS: non-temporal loads keep it from evicting v and
// the state from the caches
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <int NR>
__device__ __forceinline__ void wave_rows_dot(const double *__restrict__ K, int64_t row0, int n,
                                              const double *__restrict__ v, int lane, double acc[NR])
{
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = 0.0;
    const int n2 = n >> 1;
    const double2 *v2 = reinterpret_cast<const double2 *>(v);
    const f64x2 *kr[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) kr[r] = reinterpret_cast<const f64x2 *>(K + (row0 + r) * (int64_t)n);
    int j = lane;
    for (; j + 192 < n2; j += 256) {
        double2 vv[4];
        f64x2 kk[NR][4];
#pragma unroll
        for (int u = 0; u < 4; u++) vv[u] = v2[j + 64 * u];
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
            for (int u = 0; u < 4; u++) kk[r][u] = __builtin_nontemporal_load(kr[r] + j + 64 * u);
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                acc[r] = fma(kk[r][u].x, vv[u].x, acc[r]);
                acc[r] = fma(kk[r][u].y, vv[u].y, acc[r]);
            }
    }
    for (; j < n2; j += 64) {
        const double2 vv = v2[j];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const f64x2 kk = kr[r][j];
            acc[r] = fma(kk.x, vv.x, acc[r]);
            acc[r] = fma(kk.y, vv.y, acc[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        double a = acc[r];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        acc[r] = a;
    }
}

// the same with the packed split-bf16 K: entry = (hi + lo) * rs[row], each
// pair decoded exactly to binary64 and accumulated in binary64 (the GEMV is
// bound by HBM at 4 bytes per entry, so the wider arithmetic is free)
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double split_entry(uint32_t w)
{
    return (double)__uint_as_float(w & 0xFFFF0000u) + (double)__uint_as_float(w << 16);
}

template <int NR>
__device__ __forceinline__ void wave_rows_dot_split(const uint32_t *__restrict__ Kp,
                                                    const double *__restrict__ rs, int64_t row0,
                                                    int n, const double *__restrict__ v, int lane,
                                                    double acc[NR])
{
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = 0.0;
    const int n2 = n >> 1;
    const double2 *v2 = reinterpret_cast<const double2 *>(v);
    const u32x2 *kr[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) kr[r] = reinterpret_cast<const u32x2 *>(Kp + (row0 + r) * (int64_t)n);
    int j = lane;
    for (; j + 192 < n2; j += 256) {
        double2 vv[4];
        u32x2 kk[NR][4];
#pragma unroll
        for (int u = 0; u < 4; u++) vv[u] = v2[j + 64 * u];
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
            for (int u = 0; u < 4; u++) kk[r][u] = __builtin_nontemporal_load(kr[r] + j + 64 * u);
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
            for (int u = 0; u < 4; u++) {
                acc[r] = fma(split_entry(kk[r][u].x), vv[u].x, acc[r]);
                acc[r] = fma(split_entry(kk[r][u].y), vv[u].y, acc[r]);
            }
    }
    for (; j < n2; j += 64) {
        const double2 vv = v2[j];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const u32x2 kk = kr[r][j];
            acc[r] = fma(split_entry(kk.x), vv.x, acc[r]);
            acc[r] = fma(split_entry(kk.y), vv.y, acc[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        double a = acc[r];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        acc[r] = a * rs[row0 + r];
    }
}

__global__ void __launch_bounds__(256) matvec_split_kernel(const uint32_t *__restrict__ Kp,
                                                           const double *__restrict__ rs, int m,
                                                           int n, const double *__restrict__ v,
                                                           double *__restrict__ t)
{
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= m) return;
    double acc[1];
    wave_rows_dot_split<1>(Kp, rs, row, n, v, lane, acc);
    if (lane == 0) t[row] = acc[0];
}

void launch_matvec_split(hipStream_t st, const uint32_t *Kp, const double *rs, int m, int n,
                         const double *v, double *t)
{
    hipLaunchKernelGGL(matvec_split_kernel, dim3((m + 3) / 4), dim3(256), 0, st, Kp, rs, m, n, v, t);
}

// generic t = K v (m rows), one wavefront per row
__global__ void __launch_bounds__(256) matvec_kernel(const double *__restrict__ K, int m, int n,
                                                     const double *__restrict__ v,
                                                     double *__restrict__ t)
{
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= m) return;
    double acc[1];
    wave_rows_dot<1>(K, row, n, v, lane, acc);
    if (lane == 0) t[row] = acc[0];
}

void launch_matvec(hipStream_t st, const double *K, int m, int n, const double *v, double *t)
{
    hipLaunchKernelGGL(matvec_kernel, dim3((m + 3) / 4), dim3(256), 0, st, K, m, n, v, t);
}

// ---------------------------------------------------------------------------
// standalone odefun pieces (ucy_odefun)
// ---------------------------------------------------------------------------
__global__ void prepass_kernel(PlainDims d, const Friction *__restrict__ fr,
                               const Rheology *__restrict__ rh, const double *__restrict__ y,
                               double *__restrict__ dydt, double *__restrict__ v, double *vmax)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.np + d.nv) return;
    const int so = elem_state_off(d, e), vo = elem_v_off(d, e);
    if (d.ps && e >= d.np) {
        dydt[so + 3] = dydt[so + 4] = dydt[so + 5] = 0.0;
        ps_volume_prepass(y + so, rh[e - d.np], dydt + so, v + vo);
    } else if (d.ps) {
        for (int k = 1; k < 5; k++) dydt[so + k] = 0.0;
        const double vel = ps_prepass(y + so, fr[e], dydt + so, v + vo);
        if (vmax) atomicMax(reinterpret_cast<unsigned long long *>(vmax), dbits(vel));
    } else if (d.ap && e >= d.np) {
        dydt[so + 2] = dydt[so + 3] = 0.0;
        ap_volume_prepass(y + so, rh[e - d.np], dydt + so, v + vo);
    } else if (d.ap) {
        for (int k = 1; k < 4; k++) dydt[so + k] = 0.0;
        const double vel = ap_prepass(y + so, fr[e], dydt + so, v + vo);
        if (vmax) atomicMax(reinterpret_cast<unsigned long long *>(vmax), dbits(vel));
    } else if (e < d.np) {
        for (int k = 2; k < 7; k++) dydt[so + k] = 0.0;
        const double vel = patch_prepass(y + so, fr[e], dydt + so, v + vo);
        if (vmax) atomicMax(reinterpret_cast<unsigned long long *>(vmax), dbits(vel));
    } else {
        volume_prepass(y + so, rh[e - d.np], dydt + so, v + vo);
    }
}

void launch_prepass(hipStream_t st, const PlainDims &d, const Friction *fr, const Rheology *rh,
                    const double *y, double *dydt, double *v, double *vmax_out)
{
    const int ne = d.np + d.nv;
    hipLaunchKernelGGL(prepass_kernel, dim3((ne + 255) / 256), dim3(256), 0, st, d, fr, rh, y,
                       dydt, v, vmax_out);
}

// matvec modes of the RHS: binary64 K v; binary64 K v kept as t0 (step start
// of the mixed path); split K v kept as t0 (step start of the split path);
// t0 + split K dv (stages of both reduced paths, v holding v - v0)
enum { DOT_F64 = 0, DOT_F64_T0 = 1, DOT_SPLIT_T0 = 2, DOT_SPLIT_INC = 3 };

template <int NR, int MODE>
__device__ __forceinline__ void elem_rows(const KOperand &K, int64_t ro, int n,
                                          const double *__restrict__ v, double *t0, int lane,
                                          double acc[NR])
{
    if (MODE <= DOT_F64_T0) wave_rows_dot<NR>(K.K, ro, n, v, lane, acc);
    else wave_rows_dot_split<NR>(K.Kp, K.rs, ro, n, v, lane, acc);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (MODE == DOT_F64_T0 || MODE == DOT_SPLIT_T0) t0[ro + r] = acc[r];
            if (MODE == DOT_SPLIT_INC) acc[r] += t0[ro + r];
        }
    }
}

// K v for the rows of element e, then that element's post-pass.  One
// wavefront per element; 4 elements per 256-thread block.
template <int MODE>
__device__ __forceinline__ void elem_gemv_post(const PlainDims &d, const KOperand &K,
                                               const Friction *__restrict__ fr,
                                               const double *__restrict__ y,
                                               const double *__restrict__ v,
                                               double *__restrict__ dydt, double *t_out, int e,
                                               int lane, const double *yscal_h, double *yscal,
                                               double *t0)
{
    const int so = elem_state_off(d, e);
    const int64_t ro = elem_row_off(d, e);
    if (d.ps && e >= d.np) {
        double acc[3];
        elem_rows<3, MODE>(K, ro, d.n, v, t0, lane, acc);
        if (lane == 0) {
            for (int k = 0; k < 3; k++) {
                if (t_out) t_out[ro + k] = acc[k];
                dydt[so + 3 + k] = acc[k];  // viscouscycles.f90 (2-D plane strain):1048-1053
            }
            if (yscal) {
                const double h = *yscal_h;
                for (int k = 0; k < 6; k++)
                    yscal[so + k] = fabs(y[so + k]) + fabs(h * dydt[so + k]) + 1.0e-30;
            }
        }
    } else if (d.ps) {
        double acc[2];
        elem_rows<2, MODE>(K, ro, d.n, v, t0, lane, acc);
        if (lane == 0) {
            if (t_out) { t_out[ro] = acc[0]; t_out[ro + 1] = acc[1]; }
            ps_postpass(y + so, fr[e], acc, dydt + so);
            if (yscal) {
                const double h = *yscal_h;
                for (int k = 0; k < 5; k++)
                    yscal[so + k] = fabs(y[so + k]) + fabs(h * dydt[so + k]) + 1.0e-30;
            }
        }
    } else if (d.ap && e >= d.np) {
        double acc[2];
        elem_rows<2, MODE>(K, ro, d.n, v, t0, lane, acc);
        if (lane == 0) {
            if (t_out) { t_out[ro] = acc[0]; t_out[ro + 1] = acc[1]; }
            dydt[so + 2] = acc[0];  // viscouscycles.f90 (2-D antiplane):939-942
            dydt[so + 3] = acc[1];
            if (yscal) {
                const double h = *yscal_h;
                for (int k = 0; k < 4; k++)
                    yscal[so + k] = fabs(y[so + k]) + fabs(h * dydt[so + k]) + 1.0e-30;
            }
        }
    } else if (d.ap) {
        double acc[1];
        elem_rows<1, MODE>(K, ro, d.n, v, t0, lane, acc);
        if (lane == 0) {
            if (t_out) t_out[ro] = acc[0];
            ap_postpass(y + so, fr[e], acc[0], dydt + so);
            if (yscal) {
                const double h = *yscal_h;
                for (int k = 0; k < 4; k++)
                    yscal[so + k] = fabs(y[so + k]) + fabs(h * dydt[so + k]) + 1.0e-30;
            }
        }
    } else if (e < d.np) {
        double acc[3];
        elem_rows<3, MODE>(K, ro, d.n, v, t0, lane, acc);
        if (lane == 0) {
            if (t_out) { t_out[ro] = acc[0]; t_out[ro + 1] = acc[1]; t_out[ro + 2] = acc[2]; }
            patch_postpass(y + so, fr[e], acc, dydt + so);
            if (yscal) {
                const double h = *yscal_h;
                for (int k = 0; k < 7; k++)
                    yscal[so + k] = fabs(y[so + k]) + fabs(h * dydt[so + k]) + 1.0e-30;
            }
        }
    } else {
        double acc[6];
        elem_rows<6, MODE>(K, ro, d.n, v, t0, lane, acc);
        if (lane == 0) {
            for (int k = 0; k < 6; k++) {
                if (t_out) t_out[ro + k] = acc[k];
                dydt[so + 6 + k] = acc[k];  // viscouscycles.f90:1075-1080
            }
            if (yscal) {
                const double h = *yscal_h;
                for (int k = 0; k < 12; k++)
                    yscal[so + k] = fabs(y[so + k]) + fabs(h * dydt[so + k]) + 1.0e-30;
            }
        }
    }
}

__global__ void __launch_bounds__(256) gemv_post_kernel(PlainDims d, const double *__restrict__ K,
                                                        const Friction *__restrict__ fr,
                                                        const double *__restrict__ y,
                                                        const double *__restrict__ v,
                                                        double *__restrict__ dydt, double *t_out)
{
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= d.np + d.nv) return;
    const KOperand k{K, nullptr, nullptr, 0};
    elem_gemv_post<DOT_F64>(d, k, fr, y, v, dydt, t_out, e, lane, nullptr, nullptr, nullptr);
}

// a stage RHS of the reduced paths alone (t0 + split K dv, post-pass), for timing
__global__ void __launch_bounds__(256) gemv_post_inc_kernel(PlainDims d, KOperand K,
                                                            const Friction *__restrict__ fr,
                                                            const double *__restrict__ y,
                                                            const double *__restrict__ v,
                                                            double *__restrict__ dydt, double *t0)
{
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= d.np + d.nv) return;
    elem_gemv_post<DOT_SPLIT_INC>(d, K, fr, y, v, dydt, nullptr, e, lane, nullptr, nullptr, t0);
}

void launch_gemv_post_inc(hipStream_t st, const PlainDims &d, const KOperand &K, const Friction *fr,
                          const double *y, const double *v, double *dydt, double *t0)
{
    const int ne = d.np + d.nv;
    hipLaunchKernelGGL(gemv_post_inc_kernel, dim3((ne + 3) / 4), dim3(256), 0, st, d, K, fr, y, v,
                       dydt, t0);
}

void launch_gemv_post(hipStream_t st, const PlainDims &d, const double *K, const Friction *fr,
                      const double *y, const double *v, double *dydt, double *t_out)
{
    const int ne = d.np + d.nv;
    hipLaunchKernelGGL(gemv_post_kernel, dim3((ne + 3) / 4), dim3(256), 0, st, d, K, fr, y, v,
                       dydt, t_out);
}

// ---------------------------------------------------------------------------
// one RK45 attempt as 8 graph-capturable kernels:
//   start        commit of the previous accepted step (y = ycand,
//                ode45.f90:103) and the step-start pre-pass at y
//                (ratestate.f90:309, odefun :667-707), thread per element
//   rhs_stage 1  K v + post-pass at y (dydt, yscal of ratestate.f90:315),
//                then the input of stage 2 and its pre-pass, wave per element
//   rhs_stage s  K v + post-pass of stage s (2..5), then stage s+1's input
//   rhs_finish   K v + post-pass of stage 6, the 5th-order solution and the
//                embedded error (ode45.f90:185-200), max |yerr/yscal| (:75)
//   controller   RKQSm's decision and the main loop's time update
// The slip-rate vector is double-buffered (odd stages read v, even stages
// v2) because a wave writes the next stage's entries of its element while
// other waves still read the current ones.  A run ends with one commit.
// ---------------------------------------------------------------------------
// max of the bit patterns of a wavefront's non-negative values (NaN bits
// order above +inf), every lane active
__device__ __forceinline__ unsigned long long wave_max_bits(unsigned long long b)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long c = (unsigned long long)__shfl_xor((long long)b, o, 64);
        b = c > b ? c : b;
    }
    return b;
}

__global__ void start_kernel(PlainDims d, const Friction *__restrict__ fr,
                             const Rheology *__restrict__ rh, RkBuffers b, Control *ctl)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0) ctl->obs_pending = 0;
    const bool live = e < d.np + d.nv;
    // the step-start vMax and eMax, reduced over the wavefront and then
    // folded in with one atomic per wave (one per element would serialise
    // on the two addresses)
    double vel = 0.0, en = 0.0;
    if (live) {
        const int so = elem_state_off(d, e), vo = elem_v_off(d, e);
        if (ctl->accepted) {
            const int ns = elem_nstate(d, e);
            for (int k = 0; k < ns; k++) b.y[so + k] = b.ycand[so + k];
        }
        // ctl->vmax was cleared by the controller that accepted the previous step
        if (!ctl->done && ctl->need_rhs) {
            if (d.ps && e >= d.np) {
                // eMax of the step-start odefun (viscouscycles.f90 (2-D plane strain):962)
                en = ps_volume_prepass(b.y + so, rh[e - d.np], b.dydt + so, b.v + vo);
            } else if (d.ps) {
                vel = ps_prepass(b.y + so, fr[e], b.dydt + so, b.v + vo);
            } else if (d.ap && e >= d.np) {
                // eMax of the step-start odefun (viscouscycles.f90 (2-D antiplane):869)
                en = ap_volume_prepass(b.y + so, rh[e - d.np], b.dydt + so, b.v + vo);
            } else if (d.ap) {
                vel = ap_prepass(b.y + so, fr[e], b.dydt + so, b.v + vo);
            } else if (e < d.np) {
                // vMax of the step-start odefun (ratestate.f90:693), as exported every step
                vel = patch_prepass(b.y + so, fr[e], b.dydt + so, b.v + vo);
            } else {
                volume_prepass(b.y + so, rh[e - d.np], b.dydt + so, b.v + vo);
                // eMax of the step-start odefun (viscouscycles.f90:986), printed by export
                const double *r = b.dydt + so;
                en = sqrt((r[0] * r[0] + 2.0 * (r[1] * r[1]) + 2.0 * (r[2] * r[2]) +
                           r[3] * r[3] + 2.0 * (r[4] * r[4]) + r[5] * r[5]) / 2.0);
            }
            if (b.v0) {
                const int nvk = elem_nv(d, e);
                for (int k = 0; k < nvk; k++) b.v0[vo + k] = b.v[vo + k];
            }
        }
    }
    const unsigned long long bv = wave_max_bits(dbits(vel)), be = wave_max_bits(dbits(en));
    if ((threadIdx.x & 63) == 0) {
        if (bv) atomicMax(reinterpret_cast<unsigned long long *>(&ctl->vmax), bv);
        if (be) atomicMax(reinterpret_cast<unsigned long long *>(&ctl->emax), be);
    }
}

void launch_start(hipStream_t st, const PlainDims &d, const Friction *fr, const Rheology *rh,
                  const RkBuffers &b, Control *ctl)
{
    const int ne = d.np + d.nv;
    hipLaunchKernelGGL(start_kernel, dim3((ne + 255) / 256), dim3(256), 0, st, d, fr, rh, b, ctl);
}

// intermediate state of stage s (ode45.f90:137-180), BLAS-1 order kept:
// yrkck = B_s1*dydt; yrkck += B_sj*AKj; ytmp = yin + h*yrkck
__device__ __forceinline__ double stage_value(int stage, double h, double y0, double k1, double k2,
                                              double k3, double k4, double k5)
{
    double yr;
    switch (stage) {
    case 2:
        return y0 + (h * B21) * k1;
    case 3:
        yr = B31 * k1;
        yr = yr + B32 * k2;
        return y0 + h * yr;
    case 4:
        yr = B41 * k1;
        yr = yr + B42 * k2;
        yr = yr + B43 * k3;
        return y0 + h * yr;
    case 5:
        yr = B51 * k1;
        yr = yr + B52 * k2;
        yr = yr + B53 * k3;
        yr = yr + B54 * k4;
        return y0 + h * yr;
    default:
        yr = B61 * k1;
        yr = yr + B62 * k2;
        yr = yr + B63 * k3;
        yr = yr + B64 * k4;
        yr = yr + B65 * k5;
        return y0 + h * yr;
    }
}

// the input of stage `stage` (2..6) for element e and its pre-pass; the
// slip-rate entries go to vout (less v0 in the reduced paths)
__device__ __forceinline__ void stage_elem(const PlainDims &d, const Friction *__restrict__ fr,
                                           const Rheology *__restrict__ rh, const RkBuffers &b,
                                           double h, int e, int stage, double *vout)
{
    const int so = elem_state_off(d, e), vo = elem_v_off(d, e);
    const int ns = elem_nstate(d, e);
    for (int k = 0; k < ns; k++) {
        const int i = so + k;
        b.ytmp[i] = stage_value(stage, h, b.y[i], b.dydt[i], b.ak[2][i], b.ak[3][i], b.ak[4][i],
                                b.ak[5][i]);
    }
    double *out = b.ak[stage];
    if (d.ps && e >= d.np) ps_volume_prepass(b.ytmp + so, rh[e - d.np], out + so, vout + vo);
    else if (d.ps) ps_prepass(b.ytmp + so, fr[e], out + so, vout + vo);
    else if (d.ap && e >= d.np) ap_volume_prepass(b.ytmp + so, rh[e - d.np], out + so, vout + vo);
    else if (d.ap) ap_prepass(b.ytmp + so, fr[e], out + so, vout + vo);
    else if (e < d.np) patch_prepass(b.ytmp + so, fr[e], out + so, vout + vo);
    else volume_prepass(b.ytmp + so, rh[e - d.np], out + so, vout + vo);
    if (b.v0) {
        const int nvk = elem_nv(d, e);
        for (int k = 0; k < nvk; k++) vout[vo + k] -= b.v0[vo + k];
    }
}

template <int MODE>
__global__ void __launch_bounds__(256) rhs_stage_kernel(PlainDims d, KOperand K,
                                                        const Friction *__restrict__ fr,
                                                        const Rheology *__restrict__ rh,
                                                        RkBuffers b, Control *ctl, int stage)
{
    if (ctl->done) return;
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= d.np + d.nv) return;
    const double *vin = (stage & 1) ? b.v : b.v2;
    double *vout = (stage & 1) ? b.v2 : b.v;
    if (stage == 1) {
        // a retried attempt keeps the step-start derivative (the reference's
        // RKQSm loop reuses dydt)
        if (ctl->need_rhs)
            elem_gemv_post<MODE>(d, K, fr, b.y, vin, b.dydt, nullptr, e, lane, &ctl->h, b.yscal,
                                 b.t0);
        if (lane == 0 && e == 0) ctl->errmax_bits = 0ull;
    } else {
        elem_gemv_post<MODE>(d, K, fr, b.ytmp, vin, b.ak[stage], nullptr, e, lane, nullptr,
                             nullptr, b.t0);
    }
    if (lane == 0) stage_elem(d, fr, rh, b, ctl->h, e, stage + 1, vout);
}

template <int MODE>
__global__ void __launch_bounds__(256) rhs_finish_kernel(PlainDims d, KOperand K,
                                                         const Friction *__restrict__ fr,
                                                         RkBuffers b, Control *ctl)
{
    if (ctl->done) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int e = blockIdx.x * 4 + w;
    const bool live = e < d.np + d.nv;
    if (live)
        elem_gemv_post<MODE>(d, K, fr, b.ytmp, b.v2, b.ak[6], nullptr, e, lane, nullptr, nullptr,
                             b.t0);
    double emax = 0.0;
    if (live && lane == 0) {
        const double h = ctl->h;
        const int so = elem_state_off(d, e), ns = elem_nstate(d, e);
        for (int k = 0; k < ns; k++) {
            const int i = so + k;
            const double k1 = b.dydt[i], k3 = b.ak[3][i], k4 = b.ak[4][i], k5 = b.ak[5][i],
                         k6 = b.ak[6][i];
            double yr = C1 * k1;
            yr = yr + C3 * k3;
            yr = yr + C4 * k4;
            yr = yr + C6 * k6;
            b.ycand[i] = b.y[i] + h * yr;
            double er = DC1 * k1;
            er = er + DC3 * k3;
            er = er + DC4 * k4;
            er = er + DC5 * k5;
            er = er + DC6 * k6;
            er = h * er;
            const double r = fabs(er / b.yscal[i]);
            // non-finite ratios propagate as NaN bits, which order above +inf
            emax = (r > emax || r != r) ? r : emax;
        }
    }
    __shared__ unsigned long long red[4];
    if (lane == 0) red[w] = dbits(emax);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = red[0];
        for (int k = 1; k < 4; k++) m = red[k] > m ? red[k] : m;
        atomicMax(&ctl->errmax_bits, m);
    }
}

void launch_rhs_stage(hipStream_t st, const PlainDims &d, const KOperand &K, const Friction *fr,
                      const Rheology *rh, const RkBuffers &b, Control *ctl, int stage)
{
    const dim3 g((d.np + d.nv + 3) / 4), t(256);
    if (stage == 1 && K.prec == UCY_PREC_SPLIT_BF16)
        hipLaunchKernelGGL(rhs_stage_kernel<DOT_SPLIT_T0>, g, t, 0, st, d, K, fr, rh, b, ctl, stage);
    else if (stage == 1 && K.prec == UCY_PREC_MIXED)
        hipLaunchKernelGGL(rhs_stage_kernel<DOT_F64_T0>, g, t, 0, st, d, K, fr, rh, b, ctl, stage);
    else if (K.prec == UCY_PREC_F64)
        hipLaunchKernelGGL(rhs_stage_kernel<DOT_F64>, g, t, 0, st, d, K, fr, rh, b, ctl, stage);
    else
        hipLaunchKernelGGL(rhs_stage_kernel<DOT_SPLIT_INC>, g, t, 0, st, d, K, fr, rh, b, ctl, stage);
}

void launch_rhs_finish(hipStream_t st, const PlainDims &d, const KOperand &K, const Friction *fr,
                       const RkBuffers &b, Control *ctl)
{
    const dim3 g((d.np + d.nv + 3) / 4), t(256);
    if (K.prec == UCY_PREC_F64)
        hipLaunchKernelGGL(rhs_finish_kernel<DOT_F64>, g, t, 0, st, d, K, fr, b, ctl);
    else
        hipLaunchKernelGGL(rhs_finish_kernel<DOT_SPLIT_INC>, g, t, 0, st, d, K, fr, b, ctl);
}

// step control of RKQSm (ode45.f90:70-105) and the time update / exit test of
// the main loop (ratestate.f90:317-326).  RKQSm is called with t0 = 0
// (ratestate.f90:317), so its underflow test is h == 0.
__global__ void controller_kernel(RkBuffers b, Control *ctl)
{
    if (ctl->done) return;
    const double SAFETY = 0.7, PGROW = -0.2, PSHRNK = -0.15, ERRCON = 1.89e-4;
    ctl->attempts += 1;
    ctl->rhs_evals += ctl->need_rhs ? 6 : 5;
    const double raw = __longlong_as_double((long long)ctl->errmax_bits);
    if (!(raw == raw) || raw > 1.7976931348623157e308) {
        ctl->error = 2;
        ctl->done = 1;
        ctl->accepted = 0;
        return;
    }
    const double errmax = raw / ctl->epsilon;
    double h = ctl->h;
    if (errmax > 1.0) {
        h = h * SAFETY * pow(errmax, PSHRNK);
        ctl->h = h;
        ctl->need_rhs = 0;
        ctl->accepted = 0;
        ctl->step_rejects += 1;
        if (0.0 + h == 0.0) {
            ctl->error = 1;
            ctl->done = 1;
        }
        return;
    }
    double hnext = errmax > ERRCON ? SAFETY * h * pow(errmax, PGROW) : 10.0 * h;
    if (hnext > ctl->max_step) hnext = ctl->max_step;
    ctl->hdid = h;
    ctl->hnext = hnext;
    ctl->time = ctl->time + h;
    ctl->steps += 1;
    if (ctl->log_count < ctl->log_cap) {
        b.tlog[ctl->log_count] = ctl->time;
        b.dtlog[ctl->log_count] = h;
        b.vlog[ctl->log_count] = ctl->vmax;
        if (b.elog) {
            b.elog[ctl->log_count] = errmax;
            b.rlog[ctl->log_count] = (double)ctl->step_rejects;
        }
        if (b.emlog) b.emlog[ctl->log_count] = ctl->emax;
        ctl->log_count += 1;
    }
    ctl->step_rejects = 0;
    ctl->last_vmax = ctl->vmax;
    ctl->vmax = 0.0;
    ctl->emax = 0.0;
    ctl->accepted = 1;
    ctl->obs_pending = 1;
    ctl->need_rhs = 1;
    ctl->h = hnext;
    if (ctl->interval <= ctl->time || ctl->steps >= ctl->max_steps) ctl->done = 1;
}

// the step-start state and derivative of the observed entries after an
// accepted step, at the step's log row (the controller has advanced
// log_count; y is committed at the next attempt's start)
__global__ void obs_log_kernel(RkBuffers b, Control *ctl)
{
    // only in the attempt that accepted the step: replays issued after the
    // run ended keep `accepted` but have committed y already
    if (!ctl->obs_pending) return;
    const int64_t row = ctl->log_count - 1;
    if (row < 0 || row >= ctl->log_cap) return;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < b.n_obs;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = b.obs_idx[j];
        b.obs_y[row * b.n_obs + j] = b.y[i];
        b.obs_dydt[row * b.n_obs + j] = b.dydt[i];
    }
}

// exportPoints (ratestate.f90:414-486, viscouscycles.f90:523-625): d = O u,
// the reference's DGEMV("T") on O(n, 3 npts), with u gathered from the state
// (strike and dip slip of each patch, e11..e33 of each volume).  One block per
// row of O, four waves across its columns.  df and dl split the same products
// as buildO splits O into Of and Ol (greens.f90:388-445): fault columns and
// the volumes' third component go to df, the volumes' first two to dl.
__global__ void __launch_bounds__(256)
disp_gemv_kernel(const double *__restrict__ O, int64_t rows, int64_t n, int np, int model,
                 const double *__restrict__ y, double *d, double *df, double *dl,
                 const Control *ctl)
{
    int64_t off = 0;
    if (ctl) {
        if (!ctl->obs_pending) return;
        const int64_t row = ctl->log_count - 1;
        if (row < 0 || row >= ctl->log_cap) return;
        off = row * rows;
    }
    const int64_t r = blockIdx.x;
    // model 0: the 3-D model (3 rows per point), 1: antiplane (1 row),
    // 2: plane strain (2 rows: east, down; one dip slip per patch)
    const bool ap = model == 1, ps = model == 2;
    const bool vertical = model == 0 && (r % 3) == 2;
    const double *Or = O + r * n;
    // the 2-D antiplane model: one slip per patch (state entry 0 of 4), then
    // e12, e13 per volume (entries 0, 1 of 4); one displacement component per point
    const int64_t nf = (ap || ps) ? np : 2 * (int64_t)np, vbase = 7 * (int64_t)np;
    double a = 0.0, af = 0.0, al = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
        double u;
        if (ps) {
            // plane strain: one dip slip per patch (entry 0 of 5), then e22,
            // e23, e33 per volume (entries 0..2 of 6); viscouscycles.f90
            // (2-D plane strain):546-566
            u = j < np ? y[5 * j] : y[5 * (int64_t)np + 6 * ((j - np) / 3) + (j - np) % 3];
        } else if (ap) {
            u = j < np ? y[4 * j] : y[4 * np + 4 * ((j - np) >> 1) + ((j - np) & 1)];
        } else if (j < nf) {
            u = y[7 * (j >> 1) + (j & 1)];
        } else {
            const int64_t q = (j - nf) / 6;
            u = y[vbase + 12 * q + (j - nf - 6 * q)];
        }
        const double p = Or[j] * u;
        a += p;
        if (j < nf || vertical) af += p;
        else al += p;
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        af += __shfl_xor(af, o, 64);
        al += __shfl_xor(al, o, 64);
    }
    __shared__ double part[3][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = a;
        part[1][w] = af;
        part[2][w] = al;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < nw; k++) {
            s0 += part[0][k];
            s1 += part[1][k];
            s2 += part[2][k];
        }
        d[off + r] = s0;
        if (df) df[off + r] = s1;
        if (dl) dl[off + r] = s2;
    }
}

void launch_disp_gemv(hipStream_t st, const double *O, int64_t rows, int64_t n, int np, int model,
                      const double *y, double *d, double *df, double *dl, const Control *ctl)
{
    if (rows <= 0) return;
    hipLaunchKernelGGL(disp_gemv_kernel, dim3((unsigned)rows), dim3(256), 0, st, O, rows, n, np,
                       model, y, d, df, dl, ctl);
}

void launch_obs_log(hipStream_t st, const RkBuffers &b, Control *ctl)
{
    const int64_t blocks = b.n_obs < 256 * 64 ? (b.n_obs + 255) / 256 : 64;
    hipLaunchKernelGGL(obs_log_kernel, dim3((unsigned)blocks), dim3(256), 0, st, b, ctl);
}

// y <- ycand after an accepted attempt (ode45.f90:103)
__global__ void commit_kernel(int dim, RkBuffers b, Control *ctl)
{
    if (!ctl->accepted) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += gridDim.x * blockDim.x)
        b.y[i] = b.ycand[i];
}

void launch_controller(hipStream_t st, const RkBuffers &b, Control *ctl)
{
    hipLaunchKernelGGL(controller_kernel, dim3(1), dim3(1), 0, st, b, ctl);
}

void launch_commit(hipStream_t st, int dim, const RkBuffers &b, Control *ctl)
{
    int blocks = (dim + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(commit_kernel, dim3(blocks), dim3(256), 0, st, dim, b, ctl);
}

// ---------------------------------------------------------------------------
// standalone stage combination for ucy_rkck: stage 2..6 writes ytmp, stage 7
// writes yout and yerr.
// ---------------------------------------------------------------------------
__global__ void rk_combine_kernel(int n, int stage, double h, const double *__restrict__ yin,
                                  const double *k1, const double *k2, const double *k3,
                                  const double *k4, const double *k5, const double *k6,
                                  double *ytmp, double *yout, double *yerr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (stage <= 6) {
        ytmp[i] = stage_value(stage, h, yin[i], k1[i], stage > 2 ? k2[i] : 0.0,
                              stage > 3 ? k3[i] : 0.0, stage > 4 ? k4[i] : 0.0,
                              stage > 5 ? k5[i] : 0.0);
        return;
    }
    double yr = C1 * k1[i];
    yr = yr + C3 * k3[i];
    yr = yr + C4 * k4[i];
    yr = yr + C6 * k6[i];
    yout[i] = yin[i] + h * yr;
    double e = DC1 * k1[i];
    e = e + DC3 * k3[i];
    e = e + DC4 * k4[i];
    e = e + DC5 * k5[i];
    e = e + DC6 * k6[i];
    yerr[i] = h * e;
}

void launch_rk_combine(hipStream_t st, int n, int stage, double h, const double *yin,
                       double *const *ak, double *yr, double *ytmp, double *yout, double *yerr)
{
    (void)yr;
    hipLaunchKernelGGL(rk_combine_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, stage, h,
                       yin, ak[1], ak[2], ak[3], ak[4], ak[5], ak[6], ytmp, yout, yerr);
}

}  // namespace ucy
