// k_tri.hip -- K columns from triangular sources (Nikkhoo & Walter 2015),
// buildG's computeTractionKernelsNikkhoo15 / computeStressKernelsNikkhoo15
// rows (greens.f90:194-200, 246-252; nikkhoo15.f90:187-304).
//
// Host: per-source geometry (TD frames, vertex angles, side transforms) with
// the C library, as DispInHalfSpace.c / StrainInHalfSpace.c form it.
// Device: one (receiver, source) pair per lane, both unit slips per lane,
// each receiver row written as a contiguous run of 2 doubles per lane.
#include <hip/hip_runtime.h>
#include <cmath>
#include <string>
#include <vector>
#include "../../include/ucy.h"
#include "tri_dev.h"
#include "ucy_ctx.h"
#include "ucy_ext.h"

namespace ucy {

static const double kPiH = 3.14159265358979323846;

// TD frame, DispInHalfSpace.c:186-214 (image flip of the horizontal-TD strike
// only in the full-space routine)
static void host_td_frame(const double P1[3], const double P2[3], const double P3[3],
                          bool image_flip, double nv[3], double sv[3], double dv[3])
{
    const double a[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    const double b[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
    nv[0] = a[1] * b[2] - a[2] * b[1];
    nv[1] = a[2] * b[0] - a[0] * b[2];
    nv[2] = a[0] * b[1] - a[1] * b[0];
    double t = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    for (int k = 0; k < 3; k++) nv[k] = nv[k] / t;
    sv[0] = 0.0 * nv[2] - 1.0 * nv[1];
    sv[1] = 1.0 * nv[0] - 0.0 * nv[2];
    sv[2] = 0.0 * nv[1] - 0.0 * nv[0];
    t = sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
    if (t == 0.0) {
        sv[0] = 0.0 * nv[2];
        sv[1] = 1.0 * nv[2];
        sv[2] = 0.0 * nv[2];
        if (image_flip && P1[2] > 0.0)
            for (int k = 0; k < 3; k++) sv[k] = -sv[k];
    }
    t = sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
    for (int k = 0; k < 3; k++) sv[k] = sv[k] / t;
    dv[0] = nv[1] * sv[2] - nv[2] * sv[1];
    dv[1] = nv[2] * sv[0] - nv[0] * sv[2];
    dv[2] = nv[0] * sv[1] - nv[1] * sv[0];
    t = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    for (int k = 0; k < 3; k++) dv[k] = dv[k] / t;
}

static TdFull host_full(const double P1[3], const double P2[3], const double P3[3])
{
    TdFull T{};
    double nv[3], sv[3], dv[3];
    host_td_frame(P1, P2, P3, true, nv, sv, dv);
    for (int k = 0; k < 3; k++) {
        T.A[0][k] = nv[k];
        T.A[1][k] = sv[k];
        T.A[2][k] = dv[k];
        T.P2[k] = P2[k];
    }
    double p1[3], p3[3];
    const double p2[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++) {
        p1[i] = T.A[i][0] * (P1[0] - P2[0]) + T.A[i][1] * (P1[1] - P2[1]) + T.A[i][2] * (P1[2] - P2[2]);
        p3[i] = T.A[i][0] * (P3[0] - P2[0]) + T.A[i][1] * (P3[1] - P2[1]) + T.A[i][2] * (P3[2] - P2[2]);
    }
    double e12[3], e13[3], e23[3];
    const double n12 = sqrt((p2[0] - p1[0]) * (p2[0] - p1[0]) + (p2[1] - p1[1]) * (p2[1] - p1[1]) + (p2[2] - p1[2]) * (p2[2] - p1[2]));
    const double n13 = sqrt((p3[0] - p1[0]) * (p3[0] - p1[0]) + (p3[1] - p1[1]) * (p3[1] - p1[1]) + (p3[2] - p1[2]) * (p3[2] - p1[2]));
    const double n23 = sqrt((p3[0] - p2[0]) * (p3[0] - p2[0]) + (p3[1] - p2[1]) * (p3[1] - p2[1]) + (p3[2] - p2[2]) * (p3[2] - p2[2]));
    for (int i = 0; i < 3; i++) {
        e12[i] = (p2[i] - p1[i]) / n12;
        e13[i] = (p3[i] - p1[i]) / n13;
        e23[i] = (p3[i] - p2[i]) / n23;
    }
    const double ang[3] = {acos(e12[0] * e13[0] + e12[1] * e13[1] + e12[2] * e13[2]),
                           acos(-e12[0] * e23[0] - e12[1] * e23[1] - e12[2] * e23[2]),
                           acos(e23[0] * e13[0] + e23[1] * e13[1] + e23[2] * e13[2])};
    T.p1y = p1[1];
    T.p1z = p1[2];
    T.p3y = p3[1];
    T.p3z = p3[2];
    T.vy[0] = p1[1];
    T.vz[0] = p1[2];
    T.vy[1] = p2[1];
    T.vz[1] = p2[2];
    T.vy[2] = p3[1];
    T.vz[2] = p3[2];
    const double *c1[3] = {e13, e12, e23}, sg1[3] = {-1, 1, 1}, sg2[3] = {1, -1, -1};
    for (int v = 0; v < 3; v++) {
        T.s1[v][0] = sg1[v] * c1[v][1];
        T.s1[v][1] = sg1[v] * c1[v][2];
        T.s2[v][0] = sg2[v] * c1[v][1];
        T.s2[v][1] = sg2[v] * c1[v][2];
        T.ca[v] = cos(-kPiH + ang[v]);
        T.sa[v] = sin(-kPiH + ang[v]);
    }
    return T;
}

static TdSide host_side(const double PA[3], const double PB[3])
{
    TdSide T{};
    const double side[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
    const double len = sqrt(side[0] * side[0] + side[1] * side[1] + side[2] * side[2]);
    const double beta = acos(-side[2] / len);
    const double eps = 2.220446049250313e-16;
    T.beta = beta;
    if (fabs(beta) < eps || fabs(kPiH - beta) < eps) {
        T.skip = 1;
        return T;
    }
    double ey1[3] = {side[0], side[1], 0.0};
    const double n1 = sqrt(ey1[0] * ey1[0] + ey1[1] * ey1[1]);
    ey1[0] /= n1;
    ey1[1] /= n1;
    const double ey3[3] = {0.0, 0.0, -1.0};
    const double ey2[3] = {ey3[1] * ey1[2] - ey3[2] * ey1[1], ey3[2] * ey1[0] - ey3[0] * ey1[2],
                           ey3[0] * ey1[1] - ey3[1] * ey1[0]};
    for (int i = 0; i < 3; i++) {
        T.A[i][0] = ey1[i];
        T.A[i][1] = ey2[i];
        T.A[i][2] = ey3[i];
        T.PA[i] = PA[i];
    }
    for (int i = 0; i < 3; i++) T.sA[i] = T.A[i][0] * side[0] + T.A[i][1] * side[1] + T.A[i][2] * side[2];
    T.aA = -PA[2];
    T.aB = -PB[2];
    const double angs[2] = {-kPiH + beta, beta};
    for (int c = 0; c < 2; c++) {
        const double s = sin(angs[c]), co = cos(angs[c]);
        T.trig[c][0] = s;
        T.trig[c][1] = co;
        T.trig[c][2] = co / s;
        T.trig[c][3] = cos(angs[c] / 2.0) / sin(angs[c] / 2.0);
    }
    return T;
}

// source in NED vertices -> TriSource in the ENU frame of straininhalfspace_;
// the two directions are unit strike and unit dip slip, or, with gen set,
// the single slip (s, d, t) of nikkhoo15.f90:120-153 in direction 0
static TriSource host_tri(const double *P1, const double *P2, const double *P3,
                          bool gen = false, double s = 0.0, double d = 0.0, double t = 0.0)
{
    const double Q1[3] = {P1[1], P1[0], -P1[2]}, Q2[3] = {P2[1], P2[0], -P2[2]},
                 Q3[3] = {P3[1], P3[0], -P3[2]};
    const double M1[3] = {Q1[0], Q1[1], -Q1[2]}, M2[3] = {Q2[0], Q2[1], -Q2[2]},
                 M3[3] = {Q3[0], Q3[1], -Q3[2]};
    TriSource S{};
    S.full[0] = host_full(Q1, Q2, Q3);
    S.full[1] = host_full(M1, M2, M3);
    S.side[0] = host_side(Q1, Q2);
    S.side[1] = host_side(Q2, Q3);
    S.side[2] = host_side(Q3, Q1);
    double nv[3], sv[3], dv[3];
    host_td_frame(Q1, Q2, Q3, false, nv, sv, dv);
    // TDCS slips (tensile, strike, dip): strike (0,1,0), dip (0,0,-1); the
    // dip slip enters negated (nikkhoo15.f90:142)
    const double b0[3] = {gen ? t : 0.0, gen ? s : 1.0, gen ? -d : 0.0};
    const double b1[3] = {0.0, 0.0, gen ? 0.0 : -1.0};
    for (int i = 0; i < 3; i++) {
        S.btd[0][i] = b0[i];
        S.btd[1][i] = b1[i];
        S.bfs[0][i] = nv[i] * b0[0] + sv[i] * b0[1] + dv[i] * b0[2];
        S.bfs[1][i] = nv[i] * b1[0] + sv[i] * b1[1] + dv[i] * b1[2];
    }
    S.surface = (M1[2] == 0.0 && M2[2] == 0.0 && M3[2] == 0.0) ? 1 : 0;
    return S;
}

__device__ __forceinline__ void tri_proj_traction(const Receiver &R, const double S[6], double t[3])
{
    const double t1 = R.nv[0] * S[0] + R.nv[1] * S[1] + R.nv[2] * S[2];
    const double t2 = R.nv[0] * S[1] + R.nv[1] * S[3] + R.nv[2] * S[4];
    const double t3 = R.nv[0] * S[2] + R.nv[1] * S[4] + R.nv[2] * S[5];
    t[0] = t1 * R.sv[0] + t2 * R.sv[1] + t3 * R.sv[2];
    t[1] = t1 * R.dv[0] + t2 * R.dv[1] + t3 * R.dv[2];
    t[2] = t1 * R.nv[0] + t2 * R.nv[1] + t3 * R.nv[2];
}

__device__ __forceinline__ void tri_proj_volume(const Receiver &R, const double S[6], double o[6])
{
    const double *sv = R.sv, *dv = R.dv, *nv = R.nv;
    const double a1 = sv[0] * S[0] + sv[1] * S[1] + sv[2] * S[2];
    const double a2 = sv[0] * S[1] + sv[1] * S[3] + sv[2] * S[4];
    const double a3 = sv[0] * S[2] + sv[1] * S[4] + sv[2] * S[5];
    const double b1 = nv[0] * S[0] + nv[1] * S[1] + nv[2] * S[2];
    const double b2 = nv[0] * S[1] + nv[1] * S[3] + nv[2] * S[4];
    const double b3 = nv[0] * S[2] + nv[1] * S[4] + nv[2] * S[5];
    const double c1 = -dv[0] * S[0] - dv[1] * S[1] - dv[2] * S[2];
    const double c2 = -dv[0] * S[1] - dv[1] * S[3] - dv[2] * S[4];
    const double c3 = -dv[0] * S[2] - dv[1] * S[4] - dv[2] * S[5];
    o[0] = a1 * sv[0] + a2 * sv[1] + a3 * sv[2];
    o[1] = a1 * nv[0] + a2 * nv[1] + a3 * nv[2];
    o[2] = -a1 * dv[0] - a2 * dv[1] - a3 * dv[2];
    o[3] = b1 * nv[0] + b2 * nv[1] + b3 * nv[2];
    o[4] = -b1 * dv[0] - b2 * dv[1] - b3 * dv[2];
    o[5] = -c1 * dv[0] - c2 * dv[1] - c3 * dv[2];
}

template <int KIND, bool GEN>
__global__ void __launch_bounds__(128)
tri_rows_kernel(TriMedium M, const TriSource *__restrict__ src, int ns,
                const Receiver *__restrict__ rcv, double *__restrict__ K, int64_t ldk,
                int64_t row0, int64_t col0, double *s_out, double *d_out, int out_stride)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (s >= ns) return;
    const Receiver R = rcv[r];
    double Ss[6], Sd[6];
    tri_stress2<GEN>(M, src[s], R.xc[0], R.xc[1], R.xc[2], Ss, Sd);
    constexpr int NC = KIND == 0 ? 3 : 6;
    double os[NC], od[NC];
    if constexpr (KIND == 0) {
        tri_proj_traction(R, Ss, os);
        tri_proj_traction(R, Sd, od);
    } else {
        tri_proj_volume(R, Ss, os);
        tri_proj_volume(R, Sd, od);
    }
    if (K) {
        const int64_t base = (row0 + (int64_t)NC * r) * ldk + col0 + 2 * (int64_t)s;
#pragma unroll
        for (int c = 0; c < NC; c++)
            *reinterpret_cast<double2 *>(K + base + c * ldk) = make_double2(os[c], od[c]);
    }
    if (s_out) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            s_out[(int64_t)c * out_stride + s] = os[c];
            d_out[(int64_t)c * out_stride + s] = od[c];
        }
    }
}

// kind 0 traction rows, 1 volume-frame stress rows; gen: general slip
// vectors (per-pair hooks) instead of unit strike / dip
static void launch_tri_rows(hipStream_t st, const TriMedium &M, const TriSource *src, int ns,
                            const Receiver *rcv, int nr, int kind, double *K, int64_t ldk,
                            int64_t row0, int64_t col0, double *s_out, double *d_out, int stride,
                            bool gen = false)
{
    if (ns <= 0 || nr <= 0) return;
    dim3 grid((ns + 127) / 128, nr);
    if (kind == 0 && !gen)
        hipLaunchKernelGGL((tri_rows_kernel<0, false>), grid, dim3(128), 0, st, M, src, ns, rcv, K,
                           ldk, row0, col0, s_out, d_out, stride);
    else if (kind == 0)
        hipLaunchKernelGGL((tri_rows_kernel<0, true>), grid, dim3(128), 0, st, M, src, ns, rcv, K,
                           ldk, row0, col0, s_out, d_out, stride);
    else if (!gen)
        hipLaunchKernelGGL((tri_rows_kernel<1, false>), grid, dim3(128), 0, st, M, src, ns, rcv, K,
                           ldk, row0, col0, s_out, d_out, stride);
    else
        hipLaunchKernelGGL((tri_rows_kernel<1, true>), grid, dim3(128), 0, st, M, src, ns, rcv, K,
                           ldk, row0, col0, s_out, d_out, stride);
}

static TriMedium tri_medium(double G, double lambda)
{
    TriMedium M;
    M.mu = G;
    M.lambda = lambda;
    M.nu = lambda / (lambda + G) / 2.0;  // StrainInHalfSpace.c:193
    return M;
}

int ext_build_triangle_columns(hipStream_t st, double G, double lambda, int ntri, const int *i1,
                               const int *i2, const int *i3, int nve, const double *vert,
                               const Receiver *d_R, int np, int nvol, double *K, int64_t ldk,
                               int64_t col0, std::string &msg)
{
    (void)nve;
    std::vector<TriSource> S(ntri);
    for (int j = 0; j < ntri; j++)
        S[j] = host_tri(vert + 3 * (i1[j] - 1), vert + 3 * (i2[j] - 1), vert + 3 * (i3[j] - 1));
    TriSource *dS = nullptr;
    hipError_t e = hipMalloc((void **)&dS, sizeof(TriSource) * ntri);
    if (e != hipSuccess) {
        msg = std::string("hipMalloc: ") + hipGetErrorString(e);
        return UCY_ERR_HIP;
    }
    e = hipMemcpyAsync(dS, S.data(), sizeof(TriSource) * ntri, hipMemcpyHostToDevice, st);
    const TriMedium M = tri_medium(G, lambda);
    for (int r0 = 0; r0 < np && e == hipSuccess; r0 += 65535)
        launch_tri_rows(st, M, dS, ntri, d_R + r0, std::min(65535, np - r0), 0, K, ldk,
                        3 * (int64_t)r0, col0, nullptr, nullptr, 0);
    for (int r0 = 0; r0 < nvol && e == hipSuccess; r0 += 65535)
        launch_tri_rows(st, M, dS, ntri, d_R + np + r0, std::min(65535, nvol - r0), 1, K, ldk,
                        3 * (int64_t)np + 6 * (int64_t)r0, col0, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dS);
    if (e != hipSuccess) {
        msg = std::string("triangle assembly: ") + hipGetErrorString(e);
        return UCY_ERR_HIP;
    }
    return 0;
}

// displacement at observation points from triangle sources (buildO,
// greens.f90:302-462): one (point, triangle) pair per lane, both unit slips;
// rows 3k..3k+2 of O, columns col0 + 2s + {0,1}
__global__ void __launch_bounds__(128)
tri_disp_kernel(TriMedium M, const TriSource *__restrict__ src, int ns,
                const double *__restrict__ pts, double *__restrict__ O, int64_t ldo, int64_t col0)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (s >= ns) return;
    double us[3], ud[3];
    tri_disp2(M, src[s], pts[3 * k], pts[3 * k + 1], pts[3 * k + 2], us, ud);
    const int64_t base = 3 * (int64_t)k * ldo + col0 + 2 * (int64_t)s;
#pragma unroll
    for (int c = 0; c < 3; c++)
        *reinterpret_cast<double2 *>(O + base + c * ldo) = make_double2(us[c], ud[c]);
}

int ext_build_triangle_disp(hipStream_t st, double G, double lambda, int ntri, const int *i1,
                            const int *i2, const int *i3, const double *vert, const double *d_pts,
                            int npts, double *O, int64_t ldo, int64_t col0, std::string &msg)
{
    std::vector<TriSource> S(ntri);
    for (int j = 0; j < ntri; j++)
        S[j] = host_tri(vert + 3 * (i1[j] - 1), vert + 3 * (i2[j] - 1), vert + 3 * (i3[j] - 1));
    TriSource *dS = nullptr;
    hipError_t e = hipMalloc((void **)&dS, sizeof(TriSource) * ntri);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dS, S.data(), sizeof(TriSource) * ntri, hipMemcpyHostToDevice, st);
    // computeDisplacementNikkhoo15: nu from lambda (nikkhoo15.f90:339)
    const TriMedium M = tri_medium(G, lambda);
    for (int k0 = 0; k0 < npts && e == hipSuccess; k0 += 65535) {
        const int nk = std::min(65535, npts - k0);
        hipLaunchKernelGGL(tri_disp_kernel, dim3((ntri + 127) / 128, nk), dim3(128), 0, st, M, dS,
                           ntri, d_pts + 3 * (int64_t)k0, O + 3 * (int64_t)k0 * ldo, ldo, col0);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (dS) (void)hipFree(dS);
    if (e != hipSuccess) {
        msg = std::string("triangle displacement kernels: ") + hipGetErrorString(e);
        return UCY_ERR_HIP;
    }
    return 0;
}

// straininhalfspace_ (StrainInHalfSpace.c:42-173) over n (point, triangle)
// pairs, in the reference's east/north/up frame: main + image + harmonic
// parts, each as stress and strain, with the image's yz and zz components of
// both negated for a triangle with every vertex at the surface (:162-165; this
// makes the stress of such a triangle differ from Hooke's law applied to its
// strain, as in the reference).  Direction 0 of the source carries the slip.
__global__ void __launch_bounds__(128) strain_pairs_kernel(TriMedium M, const TriSource *__restrict__ S,
                                                           const double *__restrict__ X,
                                                           const double *__restrict__ Y,
                                                           const double *__restrict__ Z, int n,
                                                           double *__restrict__ stress,
                                                           double *__restrict__ strain)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TriSource &T = S[i];
    const G3 Xg = gmk(X[i], 1, 0, 0), Yg = gmk(Y[i], 0, 1, 0), Zg = gmk(Z[i], 0, 0, 1);
    G3 us[3], ud[3];
    double e[3][6], sg[3][6];
    td_full2(T.full[0], Xg, Yg, Zg, M.nu, T.btd[0], T.btd[1], us, ud);   // main
    sym_strain(us, e[0]);
    hooke6(e[0], M.mu, M.lambda, sg[0]);
    td_full2(T.full[1], Xg, Yg, Zg, M.nu, T.btd[0], T.btd[1], us, ud);   // image
    sym_strain(us, e[1]);
    hooke6(e[1], M.mu, M.lambda, sg[1]);
    if (T.surface) {
        e[1][4] = -e[1][4];
        e[1][5] = -e[1][5];
        sg[1][4] = -sg[1][4];
        sg[1][5] = -sg[1][5];
    }
    td_fsc2(T, Xg, Yg, Zg, M.nu, us, ud);                                 // harmonic
    sym_strain(us, e[2]);
    hooke6(e[2], M.mu, M.lambda, sg[2]);
    for (int k = 0; k < 6; k++) {
        stress[6 * (size_t)i + k] = sg[0][k] + sg[1][k] + sg[2][k];
        strain[6 * (size_t)i + k] = e[0][k] + e[1][k] + e[2][k];
    }
}

}  // namespace ucy

using namespace ucy;

extern "C" int ucy_straininhalfspace(ucy_ctx *c, int n, double *stress, double *strain,
                                     const double *X, const double *Y, const double *Z,
                                     const double *P1, const double *P2, const double *P3,
                                     double Ss, double Ds, double Ts, double mu, double lambda)
{
    if (!c || n < 0 || (n > 0 && (!stress || !strain || !X || !Y || !Z || !P1 || !P2 || !P3)))
        return UCY_ERR_ARG;
    if (n == 0) return UCY_OK;
    if (hipSetDevice(c->device) != hipSuccess) return UCY_ERR_HIP;
    // the reference takes ENU vertices; host_tri takes NED ones and converts
    // them back, and its general slip (s, d, t) enters as TDCS (t, s, -d)
    std::vector<TriSource> S(n);
    for (int i = 0; i < n; i++) {
        const double *p[3] = {P1 + 3 * (size_t)i, P2 + 3 * (size_t)i, P3 + 3 * (size_t)i};
        double q[3][3];
        for (int k = 0; k < 3; k++) {
            q[k][0] = p[k][1];
            q[k][1] = p[k][0];
            q[k][2] = -p[k][2];
        }
        S[i] = host_tri(q[0], q[1], q[2], true, Ss, -Ds, Ts);
    }
    TriSource *dS = nullptr;
    double *dbuf = nullptr;
    const size_t nn = (size_t)n;
    hipError_t e = hipMalloc((void **)&dS, sizeof(TriSource) * nn);
    if (e == hipSuccess) e = hipMalloc((void **)&dbuf, sizeof(double) * 15 * nn);
    if (e == hipSuccess) e = hipMemcpyAsync(dS, S.data(), sizeof(TriSource) * nn, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dbuf, X, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dbuf + nn, Y, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dbuf + 2 * nn, Z, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(strain_pairs_kernel, dim3((n + 127) / 128), dim3(128), 0, c->stream,
                           tri_medium(mu, lambda), dS, dbuf, dbuf + nn, dbuf + 2 * nn, n,
                           dbuf + 3 * nn, dbuf + 9 * nn);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(stress, dbuf + 3 * nn, sizeof(double) * 6 * nn, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(strain, dbuf + 9 * nn, sizeof(double) * 6 * nn, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dS);
    (void)hipFree(dbuf);
    if (e != hipSuccess) {
        c->err = std::string("straininhalfspace: ") + hipGetErrorString(e);
        return UCY_ERR_HIP;
    }
    return UCY_OK;
}

static int tri_hook(ucy_ctx *c, int kind, const double x[3], const double sv[3],
                    const double dv[3], const double nv[3], int ns, const int *i1, const int *i2,
                    const int *i3, int nVe, const double *vert, double s, double d, double t,
                    double G, double lambda, double **out)
{
    if (!c || ns < 0) return UCY_ERR_ARG;
    if (ns == 0) return UCY_OK;
    for (int j = 0; j < ns; j++)
        if (i1[j] < 1 || i1[j] > nVe || i2[j] < 1 || i2[j] > nVe || i3[j] < 1 || i3[j] > nVe) {
            c->err = "vertex index out of range";
            return UCY_ERR_ARG;
        }
    if (hipSetDevice(c->device) != hipSuccess) return UCY_ERR_HIP;
    const int NC = kind == 0 ? 3 : 6;
    std::vector<TriSource> S(ns);
    for (int j = 0; j < ns; j++)
        S[j] = host_tri(vert + 3 * (i1[j] - 1), vert + 3 * (i2[j] - 1), vert + 3 * (i3[j] - 1),
                        true, s, d, t);
    Receiver R;
    for (int k = 0; k < 3; k++) {
        R.xc[k] = x[k];
        R.sv[k] = sv[k];
        R.dv[k] = dv[k];
        R.nv[k] = nv[k];
    }
    TriSource *dS = nullptr;
    Receiver *dR = nullptr;
    double *dout = nullptr;
    hipError_t e = hipMalloc((void **)&dS, sizeof(TriSource) * ns);
    if (e == hipSuccess) e = hipMalloc((void **)&dR, sizeof(Receiver));
    if (e == hipSuccess) e = hipMalloc((void **)&dout, sizeof(double) * 2 * NC * (size_t)ns);
    if (e == hipSuccess) e = hipMemcpyAsync(dS, S.data(), sizeof(TriSource) * ns, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dR, &R, sizeof(Receiver), hipMemcpyHostToDevice, c->stream);
    std::vector<double> h(2 * NC * (size_t)ns);
    if (e == hipSuccess) {
        launch_tri_rows(c->stream, tri_medium(G, lambda), dS, ns, dR, 1, kind, nullptr, 0, 0, 0, dout,
                        dout + NC * (size_t)ns, ns, true);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dout, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dS);
    (void)hipFree(dR);
    (void)hipFree(dout);
    if (e != hipSuccess) {
        c->err = std::string("triangle kernels: ") + hipGetErrorString(e);
        return UCY_ERR_HIP;
    }
    for (int k = 0; k < NC; k++)
        for (int i = 0; i < ns; i++) out[k][i] = h[k * (size_t)ns + i];
    return UCY_OK;
}

extern "C" int ucy_nikkhoo15_traction_kernels(ucy_ctx *c, const double x[3], const double sv[3],
                                              const double dv[3], const double nv[3], int ns,
                                              const int *i1, const int *i2, const int *i3, int nVe,
                                              const double *vertices, double s, double d, double t,
                                              double G, double lambda, double *ts, double *td,
                                              double *tn)
{
    double *out[3] = {ts, td, tn};
    return tri_hook(c, 0, x, sv, dv, nv, ns, i1, i2, i3, nVe, vertices, s, d, t, G, lambda, out);
}

extern "C" int ucy_nikkhoo15_stress_kernels(ucy_ctx *c, const double x[3], const double sv[3],
                                            const double dv[3], const double nv[3], int ns,
                                            const int *i1, const int *i2, const int *i3, int nVe,
                                            const double *vertices, double s, double d, double t,
                                            double G, double lambda, double *s11, double *s12,
                                            double *s13, double *s22, double *s23, double *s33)
{
    double *out[6] = {s11, s12, s13, s22, s23, s33};
    return tri_hook(c, 1, x, sv, dv, nv, ns, i1, i2, i3, nVe, vertices, s, d, t, G, lambda, out);
}
