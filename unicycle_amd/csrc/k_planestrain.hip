// k_planestrain.hip -- K and the displacement kernels of the 2-D plane-strain
// model (jdpmoore/unicycle fortran/2d/planestrain/src).
//
// A patch is a line dislocation with dip slip, evaluated as the reference
// does as a shear zone 1e-4 of its width thick carrying e23' = -s / (2 T)
// (planestrain.f90:105-164).  One lane per (receiver, source) pair: the zone's
// displacement integrals (ps_jint.h) differentiated in forward mode with
// respect to the observation point, strain less the eigenstrain inside,
// Hooke (computeStressStrainVolumePlaneStrain.f90:88-115), then the dip and
// normal tractions on the receiver (planestrain.f90:266-291).  Each source's
// sines and cosines of 1..4 times its dip are formed on the host with the C
// library, the values the reference's SIN(k dip) take.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "ucy_internal.h"

namespace ucy {

struct PD2 {  // value and derivatives along x2, x3
    double v, g0, g1;
    __host__ __device__ PD2() : v(0.0), g0(0.0), g1(0.0) {}
    __host__ __device__ PD2(double c) : v(c), g0(0.0), g1(0.0) {}
    __host__ __device__ PD2(double c, double a, double b) : v(c), g0(a), g1(b) {}
};
#define PD_FN __host__ __device__ __forceinline__
PD_FN PD2 operator+(const PD2 &a, const PD2 &b) { return PD2(a.v + b.v, a.g0 + b.g0, a.g1 + b.g1); }
PD_FN PD2 operator-(const PD2 &a, const PD2 &b) { return PD2(a.v - b.v, a.g0 - b.g0, a.g1 - b.g1); }
PD_FN PD2 operator-(const PD2 &a) { return PD2(-a.v, -a.g0, -a.g1); }
PD_FN PD2 operator*(const PD2 &a, const PD2 &b)
{
    return PD2(a.v * b.v, a.g0 * b.v + a.v * b.g0, a.g1 * b.v + a.v * b.g1);
}
PD_FN PD2 operator/(const PD2 &a, const PD2 &b)
{
    const double q = a.v / b.v, r = 1.0 / b.v;
    return PD2(q, (a.g0 - q * b.g0) * r, (a.g1 - q * b.g1) * r);
}
PD_FN PD2 operator+(const PD2 &a, double b) { return PD2(a.v + b, a.g0, a.g1); }
PD_FN PD2 operator+(double b, const PD2 &a) { return PD2(b + a.v, a.g0, a.g1); }
PD_FN PD2 operator-(const PD2 &a, double b) { return PD2(a.v - b, a.g0, a.g1); }
PD_FN PD2 operator-(double b, const PD2 &a) { return PD2(b - a.v, -a.g0, -a.g1); }
PD_FN PD2 operator*(const PD2 &a, double b) { return PD2(a.v * b, a.g0 * b, a.g1 * b); }
PD_FN PD2 operator*(double b, const PD2 &a) { return PD2(b * a.v, b * a.g0, b * a.g1); }
PD_FN PD2 operator/(const PD2 &a, double b) { return PD2(a.v / b, a.g0 / b, a.g1 / b); }
PD_FN PD2 operator/(double b, const PD2 &a)
{
    const double q = b / a.v;
    return PD2(q, -q * a.g0 / a.v, -q * a.g1 / a.v);
}
PD_FN PD2 ps_sqrt(const PD2 &a)
{
    const double s = sqrt(a.v), h = 0.5 / s;
    return PD2(s, h * a.g0, h * a.g1);
}
PD_FN PD2 ps_log(const PD2 &a) { return PD2(log(a.v), a.g0 / a.v, a.g1 / a.v); }
// atan2(u, 1), the reference's form of atan(u); d = du / (1 + u^2)
PD_FN PD2 ps_atan2(const PD2 &u, double one)
{
    const double r = 1.0 / (one * one + u.v * u.v);
    return PD2(atan2(u.v, one), one * u.g0 * r, one * u.g1 * r);
}
}  // namespace ucy

// the value-only forms, for the displacement kernels: declared where the
// integrals' unqualified calls find them
PD_FN double ps_sqrt(double a) { return sqrt(a); }
PD_FN double ps_log(double a) { return log(a); }
PD_FN double ps_atan2(double u, double one) { return atan2(u, one); }

#include "ps_jint.h"

namespace ucy {

__device__ __forceinline__ psj::Geom ps_geom(const PsSource &S, double nu)
{
    psj::Geom g;
    g.q3 = S.q3;
    g.nu = nu;
    g.s1 = S.sk[0];
    g.c1 = S.ck[0];
    g.s2 = S.sk[1];
    g.c2 = S.ck[1];
    g.s3 = S.sk[2];
    g.c3 = S.ck[2];
    g.s4 = S.sk[3];
    g.c4 = S.ck[3];
    return g;
}

// stress (s22, s23, s33) at (r2, x3) of the strain (e22', e23', e33') in
// the zone of source S with thickness T (computeStressStrainVolume-
// PlaneStrain.f90:88-115)
__device__ __forceinline__ void ps_zone_stress(double r2, double x3, const PsSource &S, double T,
                                               double e22p, double e23p, double e33p, double G,
                                               double lam, double nu, double &s22, double &s23,
                                               double &s33)
{
    const psj::Geom g = ps_geom(S, nu);
    const psj::Strain e = psj::rotate(g, e22p, e23p, e33p, lam, G);
    const double x2 = r2 - S.q2;
    const psj::Point<PD2> P{PD2(x2, 1.0, 0.0), PD2(x3, 0.0, 1.0)};
    PD2 u2, u3;
    psj::displacement(P, g, e, T, S.W, u2, u3);
    const double in = psj::omega((x2 * g.s1 - (x3 - S.q3) * g.c1) / T)
                      * psj::boxcar_s((x2 * g.c1 + (x3 - S.q3) * g.s1) / S.W);
    const double e22 = u2.g0 - e.e22 * in;
    const double e23 = (u2.g1 + u3.g0) / 2 - e.e23 * in;
    const double e33 = u3.g1 - e.e33 * in;
    s22 = lam * (e22 + e33) + 2 * G * e22;
    s23 = 2 * G * e23;
    s33 = lam * (e22 + e33) + 2 * G * e33;
}

// stress at (r2, x3) of a unit dip slip on source S: a zone W / 1e4 thick
// with e23' = -1 / (2 T) (planestrain.f90:105-124)
__device__ __forceinline__ void ps_fault_stress(double r2, double x3, const PsSource &S, double G,
                                                double lam, double nu, double &s22, double &s23,
                                                double &s33)
{
    const double T = S.W / 1e4;
    const double eps23p = -1.0 / (2.0 * T);
    ps_zone_stress(r2, x3, S, T, 0.0, eps23p, 0.0, G, lam, nu, s22, s23, s33);
}

// buildG of the plane-strain model (greens.f90:66-226 of fortran/2d/planestrain):
// rows 2r (dip traction) and 2r + 1 (normal traction) of receiver r, column
// s; td, tn (one receiver) may replace K
__global__ void __launch_bounds__(256)
ps_rows_kernel(const PsSource *__restrict__ src, int ns, const Receiver *__restrict__ rcv,
               double G, double lam, double nu, double *__restrict__ K, int64_t ldk, double *td,
               double *tn)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (s >= ns) return;
    const Receiver R = rcv[r];
    double s22, s23, s33;
    ps_fault_stress(R.xc[1], R.xc[2], src[s], G, lam, nu, s22, s23, s33);
    const double *nv = R.nv, *dv = R.dv;
    const double d = (nv[1] * s22 + nv[2] * s23) * dv[1] + (nv[1] * s23 + nv[2] * s33) * dv[2];
    const double n = (nv[1] * s22 + nv[2] * s23) * nv[1] + (nv[1] * s23 + nv[2] * s33) * nv[2];
    if (K) {
        K[(int64_t)(2 * r) * ldk + s] = d;
        K[(int64_t)(2 * r + 1) * ldk + s] = n;
    }
    if (td) td[s] = d;
    if (tn) tn[s] = n;
}

// buildO of the plane-strain model (greens.f90:253-367): displacement (u2, u3)
// at point k of a unit dip slip on source s (planestrain.f90:146-164), rows
// 2k, 2k + 1, column s
__global__ void __launch_bounds__(256)
ps_disp_kernel(const PsSource *__restrict__ src, int ns, const double *__restrict__ pts,
               double G, double lam, double nu, double *__restrict__ O, int64_t ldo)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (s >= ns) return;
    const PsSource S = src[s];
    const psj::Geom g = ps_geom(S, nu);
    const double T = S.W / 1e4;
    const double eps23p = -1.0 / (2.0 * T);
    const psj::Strain e = psj::rotate(g, 0.0, eps23p, 0.0, lam, G);
    const psj::Point<double> P{pts[3 * k + 1] - S.q2, pts[3 * k + 2]};
    double u2, u3;
    psj::displacement(P, g, e, T, S.W, u2, u3);
    O[(int64_t)(2 * k) * ldo + s] = u2;
    O[(int64_t)(2 * k + 1) * ldo + s] = u3;
}

// The blocks of buildG that involve strain volumes (greens.f90:66-226 of
// the plane-strain viscouscycles program), one lane per (receiver, source):
//   RV = false: a patch receiver, rows row0 + 2r (dip) and row0 + 2r + 1
//     (normal traction) (planestrain.f90:266-291, 375-410);
//   RV = true: a volume receiver, rows row0 + 3r + c, the stress in its
//     frame s22' = n.s.n, s23' = -n.s.d, s33' = d.s.d (:314-348, 435-475);
//   SV = false: a patch source, column col0 + s, unit dip slip;
//   SV = true: a volume source (thickness T[s]), columns col0 + 3s + j,
//     strains 1e-3 along e22', e23', e33' (the 1d-3 eye of :125-126, 192).
// The caller passes the moduli as the reference's kernel routine receives
// them.
template <bool RV, bool SV>
__global__ void __launch_bounds__(256)
ps_block_kernel(const PsSource *__restrict__ src, const double *__restrict__ T, int ns,
                const Receiver *__restrict__ rcv, double G, double lam, double nu,
                double *__restrict__ K, int64_t ldk, int64_t row0, int64_t col0)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (s >= ns) return;
    const Receiver R = rcv[r];
    const PsSource S = src[s];
    const double *nv = R.nv, *dv = R.dv;
    constexpr int NJ = SV ? 3 : 1;
#pragma unroll 1
    for (int j = 0; j < NJ; j++) {
        double s22, s23, s33;
        if (SV)
            ps_zone_stress(R.xc[1], R.xc[2], S, T[s], j == 0 ? 1e-3 : 0.0, j == 1 ? 1e-3 : 0.0,
                           j == 2 ? 1e-3 : 0.0, G, lam, nu, s22, s23, s33);
        else
            ps_fault_stress(R.xc[1], R.xc[2], S, G, lam, nu, s22, s23, s33);
        const int64_t col = col0 + (SV ? 3 * (int64_t)s + j : s);
        if (RV) {
            const double a = (nv[1] * s22 + nv[2] * s23) * nv[1] + (nv[1] * s23 + nv[2] * s33) * nv[2];
            const double b = -(nv[1] * s22 + nv[2] * s23) * dv[1] - (nv[1] * s23 + nv[2] * s33) * dv[2];
            const double c = -(-dv[1] * s22 - dv[2] * s23) * dv[1] - (-dv[1] * s23 - dv[2] * s33) * dv[2];
            double *Kr = K + (row0 + 3 * (int64_t)r) * ldk + col;
            Kr[0] = a;
            Kr[ldk] = b;
            Kr[2 * ldk] = c;
        } else {
            const double d = (nv[1] * s22 + nv[2] * s23) * dv[1] + (nv[1] * s23 + nv[2] * s33) * dv[2];
            const double n = (nv[1] * s22 + nv[2] * s23) * nv[1] + (nv[1] * s23 + nv[2] * s33) * nv[2];
            double *Kr = K + (row0 + 2 * (int64_t)r) * ldk + col;
            Kr[0] = d;
            Kr[ldk] = n;
        }
    }
}

// buildO columns of strain volumes (greens.f90:253-367): displacement (u2,
// u3) at point k of the unit strains of the reference's eye there, whose
// third row is zero: e22' = 1, e23' = 1, and nothing for e33' (its columns
// are zero); rows 2k, 2k + 1, columns col0 + 3v + j
__global__ void __launch_bounds__(256)
ps_volume_disp_kernel(const PsSource *__restrict__ src, const double *__restrict__ T, int ns,
                      const double *__restrict__ pts, double G, double lam, double nu,
                      double *__restrict__ O, int64_t ldo, int64_t col0)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (s >= ns) return;
    const PsSource S = src[s];
    const psj::Geom g = ps_geom(S, nu);
    const psj::Point<double> P{pts[3 * k + 1] - S.q2, pts[3 * k + 2]};
    double *o0 = O + (int64_t)(2 * k) * ldo + col0 + 3 * (int64_t)s, *o1 = o0 + ldo;
#pragma unroll 1
    for (int j = 0; j < 2; j++) {
        const psj::Strain e = psj::rotate(g, j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, 0.0, lam, G);
        double u2, u3;
        psj::displacement(P, g, e, T[s], S.W, u2, u3);
        o0[j] = u2;
        o1[j] = u3;
    }
    o0[2] = 0.0;
    o1[2] = 0.0;
}

void launch_ps_blocks(hipStream_t st, const PsSource *psrc, int np, const PsSource *vsrc,
                      const double *vT, int nv, const Receiver *rcv, double G, double lam,
                      double *K, int64_t ldk)
{
    const double nu = lam / 2.0 / (lam + G);
    const double nu_sw = G / 2.0 / (G + lam);   // lambda and G exchanged
    for (int r0 = 0; r0 < np && nv > 0; r0 += 65535) {   // patch receivers <- volumes
        const int nk = std::min(65535, np - r0);
        hipLaunchKernelGGL((ps_block_kernel<false, true>), dim3((nv + 255) / 256, nk), dim3(256), 0,
                           st, vsrc, vT, nv, rcv + r0, G, lam, nu, K, ldk, 2 * (int64_t)r0,
                           (int64_t)np);
    }
    for (int r0 = 0; r0 < nv && np > 0; r0 += 65535) {   // volume receivers <- patches
        const int nk = std::min(65535, nv - r0);
        hipLaunchKernelGGL((ps_block_kernel<true, false>), dim3((np + 255) / 256, nk), dim3(256), 0,
                           st, psrc, nullptr, np, rcv + np + r0, lam, G, nu_sw, K, ldk,
                           2 * (int64_t)np + 3 * (int64_t)r0, (int64_t)0);
    }
    for (int r0 = 0; r0 < nv; r0 += 65535) {             // volume receivers <- volumes
        const int nk = std::min(65535, nv - r0);
        hipLaunchKernelGGL((ps_block_kernel<true, true>), dim3((nv + 255) / 256, nk), dim3(256), 0,
                           st, vsrc, vT, nv, rcv + np + r0, G, lam, nu, K, ldk,
                           2 * (int64_t)np + 3 * (int64_t)r0, (int64_t)np);
    }
}

void launch_ps_volume_disp(hipStream_t st, const PsSource *vsrc, const double *vT, int nv,
                           const double *pts, int npts, double G, double lam, double *O,
                           int64_t ldo, int64_t col0)
{
    const double nu = lam / 2.0 / (lam + G);
    for (int k0 = 0; k0 < npts && nv > 0; k0 += 65535) {
        const int nk = std::min(65535, npts - k0);
        hipLaunchKernelGGL(ps_volume_disp_kernel, dim3((nv + 255) / 256, nk), dim3(256), 0, st, vsrc,
                           vT, nv, pts + 3 * (int64_t)k0, G, lam, nu,
                           O + (int64_t)(2 * k0) * ldo, ldo, col0);
    }
}

void launch_ps_rows(hipStream_t st, const PsSource *src, int ns, const Receiver *rcv, int nr,
                    double G, double lam, double nu, double *K, int64_t ldk, double *td,
                    double *tn)
{
    for (int r0 = 0; r0 < nr && ns > 0; r0 += 65535) {
        const int nk = std::min(65535, nr - r0);
        hipLaunchKernelGGL(ps_rows_kernel, dim3((ns + 255) / 256, nk), dim3(256), 0, st, src, ns,
                           rcv + r0, G, lam, nu, K ? K + (int64_t)(2 * r0) * ldk : nullptr, ldk,
                           td, tn);
    }
}

void launch_ps_disp(hipStream_t st, const PsSource *src, int ns, const double *pts, int npts,
                    double G, double lam, double nu, double *O, int64_t ldo)
{
    for (int k0 = 0; k0 < npts && ns > 0; k0 += 65535) {
        const int nk = std::min(65535, npts - k0);
        hipLaunchKernelGGL(ps_disp_kernel, dim3((ns + 255) / 256, nk), dim3(256), 0, st, src, ns,
                           pts + 3 * (int64_t)k0, G, lam, nu, O + (int64_t)(2 * k0) * ldo, ldo);
    }
}

}  // namespace ucy
