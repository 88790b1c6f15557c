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
__device__ __forceinline__ void ap_segment_stress(double x2, double x3, const double *q, double G,
                                                  double &s12, double &s13)
{
    const double q2 = q[0], q3 = q[1], W = q[2];
    const double pi = 3.141592653589793;  // antiplane.f90:35
    const double r2 = x2 - q2;
    const double a = x3 - q3, b = x3 + q3, c = x3 - q3 - W, d = x3 + q3 + W;
    s12 = G * (-a / (r2 * r2 + a * a) + b / (r2 * r2 + b * b) + c / (r2 * r2 + c * c)
               - d / (r2 * r2 + d * d)) / 2.0 / pi;
    s13 = G * (r2 / (r2 * r2 + a * a) - r2 / (r2 * r2 + b * b)
               - r2 / (r2 * r2 + c * c) + r2 / (r2 * r2 + d * d)) / 2.0 / pi;
}

// the traction of computeTractionKernelsAntiplane (antiplane.f90:371-373) and
// the volume-frame stress of computeStressKernelsAntiplane (:418-425)
__device__ __forceinline__ double ap_traction(const Receiver &R, double s12, double s13)
{
    return (R.nv[1] * s12 + R.nv[2] * s13) * R.sv[0] + (R.nv[0] * s12) * R.sv[1]
           + (R.nv[0] * s13) * R.sv[2];
}
__device__ __forceinline__ void ap_frame_stress(const Receiver &R, double s12, double s13,
                                                double &s12p, double &s13p)
{
    s12p = (R.sv[1] * s12 + R.sv[2] * s13) * R.nv[0] + (R.sv[0] * s12) * R.nv[1]
           + (R.sv[0] * s13) * R.nv[2];
    s13p = -(R.sv[1] * s12 + R.sv[2] * s13) * R.dv[0] - (R.sv[0] * s12) * R.dv[1]
           - (R.sv[0] * s13) * R.dv[2];
}

__global__ void __launch_bounds__(256)
ap_rows_kernel(const double *__restrict__ src, int ns, const Receiver *__restrict__ rcv, double G,
               double *__restrict__ K, int64_t ldk, double *out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (s >= ns) return;
    const Receiver R = rcv[r];
    double s12, s13;
    ap_segment_stress(R.xc[1], R.xc[2], src + 3 * s, G, s12, s13);
    const double ts = ap_traction(R, s12, s13);
    if (K) K[(int64_t)r * ldk + s] = ts;
    if (out) out[s] = ts;
}

// boxcar of antiplane.f90:43-63
__device__ __forceinline__ double ap_omega(double x)
{
    return ((0.0 < x + 0.5) ? 1.0 : 0.0) - ((0.0 < x - 0.5) ? 1.0 : 0.0);
}

// computeStressStrainVolumeAntiplane (antiplane.f90:479-515) for the unit
// eigenstrains e12' (j = 0) and e13' (j = 1) of the volume q = (q2, q3, W, T)
// at once: the eight logarithms and sixteen arctangents serve both; the
// products by the zero component are kept, so non-finite terms propagate as
// they do in the reference's single-strain evaluation
__device__ __forceinline__ void ap_volume_stress(double x2, double x3, const double *q, double G,
                                                 double s12[2], double s13[2])
{
    const double pi = 3.141592653589793;
    const double q2 = q[0], q3 = q[1], W = q[2], T = q[3];
    const double r2 = x2 - q2, hm = r2 - T / 2, hp = r2 + T / 2;
    const double a = x3 - q3, c = x3 - q3 - W, b = x3 + q3, d = x3 + q3 + W;
    const double A = atan(a / hp) - atan(a / hm) + atan(c / hm) - atan(c / hp) - atan(d / hm)
                     - atan(b / hp) + atan(b / hm) + atan(d / hp);
    const double B = atan(hp / a) - atan(hm / a) - atan(hp / c) + atan(hm / c) + atan(hp / b)
                     - atan(hm / b) - atan(hp / d) + atan(hm / d);
    const double lmc = log(hm * hm + c * c), lpc = log(hp * hp + c * c);
    const double lmd = log(hm * hm + d * d), lpd = log(hp * hp + d * d);
    const double lma = log(hm * hm + a * a), lpa = log(hp * hp + a * a);
    const double lmb = log(hm * hm + b * b), lpb = log(hp * hp + b * b);
    const double L12 = lmc - lpc + lmd - lpd - lma + lpa - lmb + lpb;  // e13' term of s12
    const double L13 = lmc - lpc - lmd + lpd - lma + lpa + lmb - lpb;  // e12' term of s13
    const double o1 = ap_omega(r2 / T), o2 = ap_omega((x3 - (2 * q3 + W) / 2.0) / W);
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const double e12p = j == 0 ? 1.0 : 0.0, e13p = j == 1 ? 1.0 : 0.0;
        s12[j] = G / pi * e12p * A + G / (2 * pi) * e13p * L12 - 2 * G * e12p * o1 * o2;
        s13[j] = G / (2 * pi) * e12p * L13 + G / pi * e13p * B - 2 * G * e13p * o1 * o2;
    }
}

// buildG blocks of the antiplane model with strain volumes (greens.f90:93-248
// of fortran/2d/antiplane): receiver r of the elements (patches [0, np), then
// volumes) against source s; VSRC false: the patch segments (column s),
// true: the volumes (columns np + 2s + j for unit e12', e13').  A patch
// receiver fills its traction row r, a volume receiver its two frame-stress
// rows np + 2(r - np) + {0, 1}.  One pair per lane.
template <bool VSRC>
__global__ void __launch_bounds__(256)
ap_block_kernel(const double *__restrict__ src, int ns, const Receiver *__restrict__ rcv,
                int r_first, int np, double G, double *__restrict__ K, int64_t ldk)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = r_first + blockIdx.y;
    if (s >= ns) return;
    const Receiver R = rcv[r];
    double s12[2], s13[2];
    if (VSRC) ap_volume_stress(R.xc[1], R.xc[2], src + 4 * s, G, s12, s13);
    else ap_segment_stress(R.xc[1], R.xc[2], src + 3 * s, G, s12[0], s13[0]);
    const int nc = VSRC ? 2 : 1;
    const int64_t col = VSRC ? np + 2 * (int64_t)s : s;
#pragma unroll
    for (int j = 0; j < nc; j++) {
        if (r < np) {
            K[(int64_t)r * ldk + col + j] = ap_traction(R, s12[j], s13[j]);
        } else {
            double a, b;
            ap_frame_stress(R, s12[j], s13[j], a, b);
            const int64_t row = np + 2 * (int64_t)(r - np);
            K[row * ldk + col + j] = a;
            K[(row + 1) * ldk + col + j] = b;
        }
    }
}

// computeDisplacementStrainVolumeAntiplane (antiplane.f90:213-255) at point
// k for the unit e12 (column 2s) and e13 (column 2s + 1) of volume s
__global__ void __launch_bounds__(256)
ap_volume_disp_kernel(const double *__restrict__ src, int ns, const double *__restrict__ pts,
                      double *__restrict__ O, int64_t ldo)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (s >= ns) return;
    const double pi = 3.141592653589793;
    const double x2 = pts[3 * k + 1], x3 = pts[3 * k + 2];
    const double *q = src + 4 * s;
    const double q2 = q[0], q3 = q[1], W = q[2], T = q[3];
    const double r2 = x2 - q2, hm = r2 - T / 2, hp = r2 + T / 2;
    const double a = x3 - q3, c = x3 - q3 - W, b = x3 + q3, d = x3 + q3 + W;
    const double cw = x3 - W - q3, dw = x3 + W + q3;  // the reference's operand order
    const double lmc = log(hm * hm + c * c), lpc = log(hp * hp + c * c);
    const double lmd = log(hm * hm + d * d), lpd = log(hp * hp + d * d);
    const double lma = log(hm * hm + a * a), lpa = log(hp * hp + a * a);
    const double lmb = log(hm * hm + b * b), lpb = log(hp * hp + b * b);
    const double u12 = c * lmc - c * lpc - a * lma + a * lpa
                       + 2 * hm * (atan(c / hm) - atan(a / hm))
                       + 2 * hp * (atan(a / hp) - atan(c / hp))
                       + d * lpd - d * lmd - b * lpb + b * lmb
                       + 2 * hp * (atan(d / hp) - atan(b / hp))
                       + 2 * hm * (atan(b / hm) - atan(d / hm));
    const double u13 = hm * lmc - hp * lpc - hm * lma + hp * lpa
                       + 2 * cw * (atan(hm / c) - atan(hp / c))
                       + 2 * a * (atan(hp / a) - atan(hm / a))
                       + hm * lmd - hp * lpd - hm * lmb + hp * lpb
                       + 2 * dw * (atan(hm / d) - atan(hp / d))
                       + 2 * b * (atan(hp / b) - atan(hm / b));
    double *o = O + (int64_t)k * ldo + 2 * (int64_t)s;
    o[0] = 1.0 / (2 * pi) * u12 + 0.0 / (2 * pi) * u13;
    o[1] = 0.0 / (2 * pi) * u12 + 1.0 / (2 * pi) * u13;
}

void launch_ap_volume_blocks(hipStream_t st, const double *seg, int np, const double *vol, int nv,
                             const Receiver *rcv, double G, double *K, int64_t ldk)
{
    const int ne = np + nv;
    // every receiver against the volume sources
    for (int r0 = 0; r0 < ne && nv > 0; r0 += 65535) {
        const int nk = std::min(65535, ne - r0);
        hipLaunchKernelGGL(ap_block_kernel<true>, dim3((nv + 255) / 256, nk), dim3(256), 0, st,
                           vol, nv, rcv, r0, np, G, K, ldk);
    }
    // the volume receivers against the patch segments
    for (int r0 = np; r0 < ne && np > 0; r0 += 65535) {
        const int nk = std::min(65535, ne - r0);
        hipLaunchKernelGGL(ap_block_kernel<false>, dim3((np + 255) / 256, nk), dim3(256), 0, st,
                           seg, np, rcv, r0, np, G, K, ldk);
    }
}

void launch_ap_volume_disp(hipStream_t st, const double *vol, int nv, const double *pts, int npts,
                           double *O, int64_t ldo)
{
    for (int k0 = 0; k0 < npts && nv > 0; k0 += 65535) {
        const int nk = std::min(65535, npts - k0);
        hipLaunchKernelGGL(ap_volume_disp_kernel, dim3((nv + 255) / 256, nk), dim3(256), 0, st,
                           vol, nv, pts + 3 * (int64_t)k0, O + (int64_t)k0 * ldo, ldo);
    }
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
