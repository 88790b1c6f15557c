// k_assemble.hip -- assembly of the traction/stress interaction kernel K
// (buildG, jdpmoore/unicycle fortran/3d/src/greens.f90:73-275) from
// rectangular (Okada 1992) sources.
//
// One (receiver, source) pair per lane, both unit slips per lane: a
// wavefront covers 64 consecutive sources of one receiver, so each of the
// receiver's rows is written as one contiguous 1 KiB run (2 doubles per lane,
// columns 2s and 2s+1 of greens.f90:214-216).
#include <hip/hip_runtime.h>
#include <algorithm>
#include "ucy_internal.h"

namespace ucy {

// traction (sigma.n) projected on (s,d,n), okada92.f90:390-398
__device__ __forceinline__ void proj_traction(const Receiver &R, const double S[6], double t[3])
{
    const double t1 = R.nv[0] * S[0] + R.nv[1] * S[1] + R.nv[2] * S[2];
    const double t2 = R.nv[0] * S[1] + R.nv[1] * S[3] + R.nv[2] * S[4];
    const double t3 = R.nv[0] * S[2] + R.nv[1] * S[4] + R.nv[2] * S[5];
    t[0] = t1 * R.sv[0] + t2 * R.sv[1] + t3 * R.sv[2];
    t[1] = t1 * R.dv[0] + t2 * R.dv[1] + t3 * R.dv[2];
    t[2] = t1 * R.nv[0] + t2 * R.nv[1] + t3 * R.nv[2];
}

// stress in the receiver volume frame, okada92.f90:445-462
__device__ __forceinline__ void proj_volume(const Receiver &R, const double S[6], double o[6])
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

template <int KIND>
__global__ void __launch_bounds__(256)
okada_rows_kernel(OkadaMedium M, const OkadaSource *__restrict__ src, int ns,
                  const Receiver *__restrict__ rcv, double *__restrict__ K, int64_t ldk,
                  int64_t row0, int64_t col0, double *s_out, double *d_out, int out_stride)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (s >= ns) return;
    const Receiver R = rcv[r];
    const OkadaSource S = src[s];
    // okada92.f90:144-153
    const double dx = R.xc[0] - S.q1, dy = R.xc[1] - S.q2;
    const float X = (float)(dx * S.csst + dy * S.ssst);
    const float Y = (float)(dx * S.ssst - dy * S.csst);
    const float Z = (float)(-R.xc[2]);
    float gs[9], gd[9];
    dc3d_pair(M, S, X, Y, Z, gs, gd);
    double sigs[6], sigd[6];
    okada_stress_from_grad(M, S.csst, S.ssst, gs, sigs);
    okada_stress_from_grad(M, S.csst, S.ssst, gd, sigd);
    constexpr int NC = KIND == 0 ? 3 : 6;
    double os[NC], od[NC];
    if constexpr (KIND == 0) {
        proj_traction(R, sigs, os);
        proj_traction(R, sigd, od);
    } else {
        proj_volume(R, sigs, os);
        proj_volume(R, sigd, od);
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

void launch_okada_rows(hipStream_t st, const OkadaMedium &M, const OkadaSource *src, int ns,
                       const Receiver *rcv, int nr, int kind, double *K, int64_t ldk,
                       int64_t row0, int64_t col0, double *s_out, double *d_out, int out_stride)
{
    if (ns <= 0 || nr <= 0) return;
    dim3 grid((ns + 255) / 256, nr);
    if (kind == 0)
        hipLaunchKernelGGL(okada_rows_kernel<0>, grid, dim3(256), 0, st, M, src, ns, rcv, K, ldk,
                           row0, col0, s_out, d_out, out_stride);
    else
        hipLaunchKernelGGL(okada_rows_kernel<1>, grid, dim3(256), 0, st, M, src, ns, rcv, K, ldk,
                           row0, col0, s_out, d_out, out_stride);
}

// Displacement at observation points from rectangle sources, buildO
// (greens.f90:302-462) with computeDisplacementOkada92 (okada92.f90:230-298):
// one (point, source) pair per lane, both unit slips; rows 3k..3k+2 of O
// (NED components at point k), columns col0 + 2s, col0 + 2s + 1.  A point on
// a fault edge (DC3D singular) gets zero displacement.
__global__ void __launch_bounds__(256)
okada_disp_kernel(OkadaMedium M, const OkadaSource *__restrict__ src, int ns,
                  const double *__restrict__ pts, double *__restrict__ O, int64_t ldo, int64_t col0)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (s >= ns) return;
    const OkadaSource S = src[s];
    // okada92.f90:262-268
    const double dx = pts[3 * k] - S.q1, dy = pts[3 * k + 1] - S.q2;
    const float X = (float)(dx * S.csst + dy * S.ssst);
    const float Y = (float)(dx * S.ssst - dy * S.csst);
    const float Z = (float)(-pts[3 * k + 2]);
    float us[3], ud[3];
    const int iret = dc3d_pair_out<0, 3>(M, S, X, Y, Z, us, ud);
    double os[3] = {0.0, 0.0, 0.0}, od[3] = {0.0, 0.0, 0.0};
    if (iret == 0) {  // okada92.f90:286-295
        os[0] = (double)us[0] * S.csst + (double)us[1] * S.ssst;
        os[1] = (double)us[0] * S.ssst - (double)us[1] * S.csst;
        os[2] = -(double)us[2];
        od[0] = (double)ud[0] * S.csst + (double)ud[1] * S.ssst;
        od[1] = (double)ud[0] * S.ssst - (double)ud[1] * S.csst;
        od[2] = -(double)ud[2];
    }
    const int64_t base = 3 * (int64_t)k * ldo + col0 + 2 * (int64_t)s;
#pragma unroll
    for (int c = 0; c < 3; c++)
        *reinterpret_cast<double2 *>(O + base + c * ldo) = make_double2(os[c], od[c]);
}

// buildG of the 2-D antiplane model (fortran/2d/antiplane/src/greens.f90:93-248):
// traction at receiver r from unit slip on source segment s,
// computeStressAntiplane (antiplane.f90:131-156; vertical segment from depth
// q3 to q3 + W at q2 with its image, the dip ignored as in the reference)
// projected by computeTractionKernelsAntiplane (antiplane.f90:371-373).  One
// pair per lane; row r is written contiguously.
__global__ void __launch_bounds__(256)
ap_rows_kernel(const double *__restrict__ src, int ns, const Receiver *__restrict__ rcv, double G,
               double *__restrict__ K, int64_t ldk, double *out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (s >= ns) return;
    const Receiver R = rcv[r];
    const double q2 = src[3 * s], q3 = src[3 * s + 1], W = src[3 * s + 2];
    const double pi = 3.141592653589793;  // antiplane.f90:35
    const double r2 = R.xc[1] - q2;
    const double a = R.xc[2] - q3, b = R.xc[2] + q3, c = R.xc[2] - q3 - W, d = R.xc[2] + q3 + W;
    const double s12 = G * (-a / (r2 * r2 + a * a) + b / (r2 * r2 + b * b) + c / (r2 * r2 + c * c)
                            - d / (r2 * r2 + d * d)) / 2.0 / pi;
    const double s13 = G * (r2 / (r2 * r2 + a * a) - r2 / (r2 * r2 + b * b)
                            - r2 / (r2 * r2 + c * c) + r2 / (r2 * r2 + d * d)) / 2.0 / pi;
    const double ts = (R.nv[1] * s12 + R.nv[2] * s13) * R.sv[0] + (R.nv[0] * s12) * R.sv[1]
                      + (R.nv[0] * s13) * R.sv[2];
    if (K) K[(int64_t)r * ldk + s] = ts;
    if (out) out[s] = ts;
}

// buildO of the 2-D antiplane model (fortran/2d/antiplane/src/greens.f90:275-383):
// displacement at point k (x2, x3) from unit slip on segment s,
// computeDisplacementAntiplane (antiplane.f90:178-191); row k, column s
__global__ void __launch_bounds__(256)
ap_disp_kernel(const double *__restrict__ src, int ns, const double *__restrict__ pts,
               double *__restrict__ O, int64_t ldo)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (s >= ns) return;
    const double x2 = pts[3 * k + 1], x3 = pts[3 * k + 2];
    const double q2 = src[3 * s], q3 = src[3 * s + 1], W = src[3 * s + 2];
    const double pi = 3.141592653589793;
    const double r = x2 - q2;
    O[(int64_t)k * ldo + s] = 1.0 * (+atan((x3 - q3) / r) - atan((x3 + q3) / r)
                                     - atan((x3 - q3 - W) / r) + atan((x3 + q3 + W) / r))
                              / 2.0 / pi;
}

void launch_ap_disp(hipStream_t st, const double *src, int ns, const double *pts, int npts,
                    double *O, int64_t ldo)
{
    for (int k0 = 0; k0 < npts && ns > 0; k0 += 65535) {
        const int nk = std::min(65535, npts - k0);
        hipLaunchKernelGGL(ap_disp_kernel, dim3((ns + 255) / 256, nk), dim3(256), 0, st, src, ns,
                           pts + 3 * (int64_t)k0, O + (int64_t)k0 * ldo, ldo);
    }
}

void launch_ap_rows(hipStream_t st, const double *src, int ns, const Receiver *rcv, int nr,
                    double G, double *K, int64_t ldk, double *out)
{
    for (int r0 = 0; r0 < nr && ns > 0; r0 += 65535) {
        const int nk = std::min(65535, nr - r0);
        hipLaunchKernelGGL(ap_rows_kernel, dim3((ns + 255) / 256, nk), dim3(256), 0, st, src, ns,
                           rcv + r0, G, K ? K + (int64_t)r0 * ldk : nullptr, ldk, out);
    }
}

void launch_okada_disp(hipStream_t st, const OkadaMedium &M, const OkadaSource *src, int ns,
                       const double *pts, int npts, double *O, int64_t ldo, int64_t col0)
{
    for (int k0 = 0; k0 < npts && ns > 0; k0 += 65535) {
        const int nk = std::min(65535, npts - k0);
        hipLaunchKernelGGL(okada_disp_kernel, dim3((ns + 255) / 256, nk), dim3(256), 0, st, M, src,
                           ns, pts + 3 * (int64_t)k0, O + 3 * (int64_t)k0 * ldo, ldo, col0);
    }
}

}  // namespace ucy
