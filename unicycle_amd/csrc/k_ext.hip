// k_ext.hip -- vertical strain-volume (Barbot, Moore & Lambert 2017) source
// columns of K and the per-kernel hooks of strainvolume.f90:4671-4799.
//
// One lane per (receiver, source volume) pair computes the stress of all six
// unit eigenstrains at once: the 8-corner sum of the displacement
// J-integrals (sv_jint.h) differentiated in forward mode with respect to the
// observation point, then strain, rotation, eigenstrain removal inside the
// volume and Hooke (strainvolume.f90:1042-1138), then the traction projection
// for patch receivers or the volume-frame stress for volume receivers.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <vector>
#include "ucy_ext.h"
#include "ucy_ctx.h"
#include "../../include/ucy.h"

namespace ucy {

struct VD {  // value and derivatives along x1', x2', x3
    double v, g[3];
    __host__ __device__ VD() : v(0.0), g{0.0, 0.0, 0.0} {}
    __host__ __device__ VD(double c) : v(c), g{0.0, 0.0, 0.0} {}
    __host__ __device__ VD(double c, double a, double b, double e) : v(c), g{a, b, e} {}
    __host__ __device__ double d(int j) const { return g[j]; }
    __host__ __device__ double val() const { return v; }
};
#define VD_FN __host__ __device__ __forceinline__
VD_FN VD operator+(const VD &a, const VD &b) { return VD(a.v + b.v, a.g[0] + b.g[0], a.g[1] + b.g[1], a.g[2] + b.g[2]); }
VD_FN VD operator-(const VD &a, const VD &b) { return VD(a.v - b.v, a.g[0] - b.g[0], a.g[1] - b.g[1], a.g[2] - b.g[2]); }
VD_FN VD operator-(const VD &a) { return VD(-a.v, -a.g[0], -a.g[1], -a.g[2]); }
VD_FN VD operator*(const VD &a, const VD &b)
{
    return VD(a.v * b.v, a.g[0] * b.v + a.v * b.g[0], a.g[1] * b.v + a.v * b.g[1],
              a.g[2] * b.v + a.v * b.g[2]);
}
VD_FN VD operator/(const VD &a, const VD &b)
{
    const double q = a.v / b.v, r = 1.0 / b.v;
    return VD(q, (a.g[0] - q * b.g[0]) * r, (a.g[1] - q * b.g[1]) * r, (a.g[2] - q * b.g[2]) * r);
}
VD_FN VD operator+(const VD &a, double b) { return VD(a.v + b, a.g[0], a.g[1], a.g[2]); }
VD_FN VD operator+(double b, const VD &a) { return VD(a.v + b, a.g[0], a.g[1], a.g[2]); }
VD_FN VD operator-(const VD &a, double b) { return VD(a.v - b, a.g[0], a.g[1], a.g[2]); }
VD_FN VD operator-(double b, const VD &a) { return VD(b - a.v, -a.g[0], -a.g[1], -a.g[2]); }
VD_FN VD operator*(const VD &a, double b) { return VD(a.v * b, a.g[0] * b, a.g[1] * b, a.g[2] * b); }
VD_FN VD operator*(double b, const VD &a) { return VD(a.v * b, a.g[0] * b, a.g[1] * b, a.g[2] * b); }
VD_FN VD operator/(const VD &a, double b) { return VD(a.v / b, a.g[0] / b, a.g[1] / b, a.g[2] / b); }

VD_FN VD sv_sqrt(const VD &a)
{
    const double s = sqrt(a.v), h = 0.5 / s;
    return VD(s, h * a.g[0], h * a.g[1], h * a.g[2]);
}
// d atan2(y, x) = (x dy - y dx) / (x^2 + y^2)
VD_FN VD sv_atan2(const VD &y, const VD &x)
{
    const double r = 1.0 / (x.v * x.v + y.v * y.v);
    return VD(atan2(y.v, x.v), (x.v * y.g[0] - y.v * x.g[0]) * r, (x.v * y.g[1] - y.v * x.g[1]) * r,
              (x.v * y.g[2] - y.v * x.g[2]) * r);
}
// atan(y/x), sign(pi/2, y) at x = 0 (strainvolume.f90:91-100)
VD_FN VD sv_atan3(const VD &y, const VD &x)
{
    const double v = x.v == 0.0 ? copysign(1.5707963267948966, y.v) : atan(y.v / x.v);
    const double r = 1.0 / (x.v * x.v + y.v * y.v);
    return VD(v, (x.v * y.g[0] - y.v * x.g[0]) * r, (x.v * y.g[1] - y.v * x.g[1]) * r,
              (x.v * y.g[2] - y.v * x.g[2]) * r);
}
// x log y with 0 log y = 0 (strainvolume.f90:107-116), same rule for dx
VD_FN VD sv_xlogy(const VD &x, const VD &y)
{
    const double ly = log(y.v), iy = 1.0 / y.v;
    VD r;
    r.v = x.v == 0.0 ? 0.0 : x.v * ly;
    for (int j = 0; j < 3; j++) {
        const double t = x.g[j] == 0.0 ? 0.0 : x.g[j] * ly;
        r.g[j] = x.v == 0.0 ? t : t + x.v * y.g[j] * iy;
    }
    return r;
}
// 0.5 Re(log(X+1) - log(X-1)) (macros.f90:7)
VD_FN VD sv_acoth(const VD &X)
{
    const double v = 0.5 * (log(fabs(X.v + 1.0)) - log(fabs(X.v - 1.0)));
    const double d = 1.0 / (1.0 - X.v * X.v);
    return VD(v, d * X.g[0], d * X.g[1], d * X.g[2]);
}

// value only: the displacement kernels need no derivatives
struct VV {
    double v;
    __host__ __device__ VV() : v(0.0) {}
    __host__ __device__ VV(double c) : v(c) {}
    __host__ __device__ double d(int) const { return 0.0; }
    __host__ __device__ double val() const { return v; }
};
VD_FN VV operator+(const VV &a, const VV &b) { return VV(a.v + b.v); }
VD_FN VV operator-(const VV &a, const VV &b) { return VV(a.v - b.v); }
VD_FN VV operator-(const VV &a) { return VV(-a.v); }
VD_FN VV operator*(const VV &a, const VV &b) { return VV(a.v * b.v); }
VD_FN VV operator/(const VV &a, const VV &b) { return VV(a.v / b.v); }
VD_FN VV operator+(const VV &a, double b) { return VV(a.v + b); }
VD_FN VV operator+(double b, const VV &a) { return VV(a.v + b); }
VD_FN VV operator-(const VV &a, double b) { return VV(a.v - b); }
VD_FN VV operator-(double b, const VV &a) { return VV(b - a.v); }
VD_FN VV operator*(const VV &a, double b) { return VV(a.v * b); }
VD_FN VV operator*(double b, const VV &a) { return VV(a.v * b); }
VD_FN VV operator/(const VV &a, double b) { return VV(a.v / b); }
VD_FN VV sv_sqrt(const VV &a) { return VV(sqrt(a.v)); }
VD_FN VV sv_atan2(const VV &y, const VV &x) { return VV(atan2(y.v, x.v)); }
VD_FN VV sv_atan3(const VV &y, const VV &x)
{
    return VV(x.v == 0.0 ? copysign(1.5707963267948966, y.v) : atan(y.v / x.v));
}
VD_FN VV sv_xlogy(const VV &x, const VV &y) { return VV(x.v == 0.0 ? 0.0 : x.v * log(y.v)); }
VD_FN VV sv_acoth(const VV &X)
{
    return VV(0.5 * (log(fabs(X.v + 1.0)) - log(fabs(X.v - 1.0))));
}

}  // namespace ucy

#include "sv_jint.h"

namespace svj {
template <>
struct has_derivative<ucy::VV> {
    static constexpr bool value = false;
};
}  // namespace svj

namespace ucy {

struct SeedVD {
    __host__ __device__ VD operator()(double v, int j) const
    {
        VD r(v);
        r.g[j] = 1.0;
        return r;
    }
};

struct SvSource {
    double q[3], L, T, W, strike;
};

// KIND 0: traction rows (3) of patch receivers; 1: volume-frame stress rows (6)
template <int KIND>
__global__ void __launch_bounds__(128)
sv_rows_kernel(double G, double nu, const SvSource *__restrict__ src, int ns,
               const Receiver *__restrict__ rcv, double *__restrict__ K, int64_t ldk, int64_t row0,
               int64_t col0, double *out, int out_stride)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (s >= ns) return;
    const Receiver R = rcv[r];
    const SvSource S = src[s];
    double st[6][6];
    svj::stress6<VD, SeedVD, true>(R.xc[0], R.xc[1], R.xc[2], S.q[0], S.q[1], S.q[2], S.L, S.T,
                                   S.W, S.strike, G, nu, SeedVD(), st);
    constexpr int NC = KIND == 0 ? 3 : 6;
    double o[6][NC];
    const double *sv = R.sv, *dv = R.dv, *nv = R.nv;
    for (int e = 0; e < 6; e++) {
        const double *S6 = st[e];
        if constexpr (KIND == 0) {  // strainvolume.f90:4699-4707
            const double t1 = nv[0] * S6[0] + nv[1] * S6[1] + nv[2] * S6[2];
            const double t2 = nv[0] * S6[1] + nv[1] * S6[3] + nv[2] * S6[4];
            const double t3 = nv[0] * S6[2] + nv[1] * S6[4] + nv[2] * S6[5];
            o[e][0] = t1 * sv[0] + t2 * sv[1] + t3 * sv[2];
            o[e][1] = t1 * dv[0] + t2 * dv[1] + t3 * dv[2];
            o[e][2] = t1 * nv[0] + t2 * nv[1] + t3 * nv[2];
        } else {  // strainvolume.f90:4771-4793
            const double a1 = sv[0] * S6[0] + sv[1] * S6[1] + sv[2] * S6[2];
            const double a2 = sv[0] * S6[1] + sv[1] * S6[3] + sv[2] * S6[4];
            const double a3 = sv[0] * S6[2] + sv[1] * S6[4] + sv[2] * S6[5];
            const double b1 = nv[0] * S6[0] + nv[1] * S6[1] + nv[2] * S6[2];
            const double b2 = nv[0] * S6[1] + nv[1] * S6[3] + nv[2] * S6[4];
            const double b3 = nv[0] * S6[2] + nv[1] * S6[4] + nv[2] * S6[5];
            const double c1 = -dv[0] * S6[0] - dv[1] * S6[1] - dv[2] * S6[2];
            const double c2 = -dv[0] * S6[1] - dv[1] * S6[3] - dv[2] * S6[4];
            const double c3 = -dv[0] * S6[2] - dv[1] * S6[4] - dv[2] * S6[5];
            o[e][0] = a1 * sv[0] + a2 * sv[1] + a3 * sv[2];
            o[e][1] = a1 * nv[0] + a2 * nv[1] + a3 * nv[2];
            o[e][2] = -a1 * dv[0] - a2 * dv[1] - a3 * dv[2];
            o[e][3] = b1 * nv[0] + b2 * nv[1] + b3 * nv[2];
            o[e][4] = -b1 * dv[0] - b2 * dv[1] - b3 * dv[2];
            o[e][5] = -c1 * dv[0] - c2 * dv[1] - c3 * dv[2];
        }
    }
    if (K) {  // rows row0 + NC r + c, columns col0 + 6 s + e
        const int64_t base = (row0 + (int64_t)NC * r) * ldk + col0 + 6 * (int64_t)s;
        for (int c = 0; c < NC; c++) {
            double2 *p = reinterpret_cast<double2 *>(K + base + c * ldk);
            p[0] = make_double2(o[0][c], o[1][c]);
            p[1] = make_double2(o[2][c], o[3][c]);
            p[2] = make_double2(o[4][c], o[5][c]);
        }
    }
    if (out)  // hook layout: out[(e*NC + c)*stride + s]
        for (int e = 0; e < 6; e++)
            for (int c = 0; c < NC; c++) out[(int64_t)(e * NC + c) * out_stride + s] = o[e][c];
}

struct SeedVV {
    __host__ __device__ VV operator()(double v, int) const { return VV(v); }
};

// displacement at observation points from strain-volume sources (buildO,
// greens.f90:423-448 with computeDisplacementVerticalStrainVolume,
// strainvolume.f90:215-268): one (point, volume) pair per lane, the six unit
// eigenstrains from one 8-corner sum; rows 3k..3k+2, columns col0 + 6s + e
__global__ void __launch_bounds__(128)
sv_disp_kernel(double G, double nu, const SvSource *__restrict__ src, int ns,
               const double *__restrict__ pts, double *__restrict__ O, int64_t ldo, int64_t col0)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (s >= ns) return;
    const SvSource S = src[s];
    double u[6][3];
    svj::displacement6<VV>(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2], S.q[0], S.q[1], S.q[2], S.L,
                           S.T, S.W, S.strike, G, nu, SeedVV(), u);
    const int64_t base = 3 * (int64_t)k * ldo + col0 + 6 * (int64_t)s;
    for (int c = 0; c < 3; c++) {
        double2 *p = reinterpret_cast<double2 *>(O + base + c * ldo);
        p[0] = make_double2(u[0][c], u[1][c]);
        p[1] = make_double2(u[2][c], u[3][c]);
        p[2] = make_double2(u[4][c], u[5][c]);
    }
}

static void launch_sv_rows(hipStream_t st, double G, double nu, const SvSource *src, int ns,
                           const Receiver *rcv, int nr, int kind, double *K, int64_t ldk,
                           int64_t row0, int64_t col0, double *out, int stride)
{
    if (ns <= 0 || nr <= 0) return;
    const dim3 grid((ns + 127) / 128, nr);
    if (kind == 0)
        hipLaunchKernelGGL(sv_rows_kernel<0>, grid, dim3(128), 0, st, G, nu, src, ns, rcv, K, ldk,
                           row0, col0, out, stride);
    else
        hipLaunchKernelGGL(sv_rows_kernel<1>, grid, dim3(128), 0, st, G, nu, src, ns, rcv, K, ldk,
                           row0, col0, out, stride);
}

// strainvolume.f90:1059-1072 (STOP in the reference)
static bool sv_bad(double nu, double x3, double q3)
{
    return (-1.0 > nu) || (0.5 < nu) || (0.0 > x3) || (0.0 > q3);
}

int ext_build_volume_columns(hipStream_t st, double G, double lambda, int nvol, const double *x,
                             const double *L, const double *T, const double *W,
                             const double *strike, const Receiver *d_R, int np, double *K,
                             int64_t ldk, int64_t col0, std::string &msg)
{
    // computeTractionKernels/StressKernelsVerticalStrainVolume: nu from lambda
    const double nu = lambda / (lambda + G) / 2.0;
    std::vector<SvSource> S(nvol);
    for (int i = 0; i < nvol; i++) {
        S[i] = {{x[3 * i], x[3 * i + 1], x[3 * i + 2]}, L[i], T[i], W[i], strike[i]};
        if (sv_bad(nu, 0.0, S[i].q[2])) {
            msg = "strain volume: -1<=nu<=0.5 and source depth >= 0 required";
            return UCY_ERR_GEOMETRY;
        }
    }
    std::vector<Receiver> R(np + nvol);
    hipError_t e = hipMemcpy(R.data(), d_R, sizeof(Receiver) * (np + nvol), hipMemcpyDeviceToHost);
    for (int i = 0; e == hipSuccess && i < np + nvol; i++)
        if (R[i].xc[2] < 0.0) {
            msg = "strain volume: observation depth must be positive";
            return UCY_ERR_GEOMETRY;
        }
    SvSource *dS = nullptr;
    if (e == hipSuccess) e = hipMalloc((void **)&dS, sizeof(SvSource) * nvol);
    if (e == hipSuccess) e = hipMemcpyAsync(dS, S.data(), sizeof(SvSource) * nvol, hipMemcpyHostToDevice, st);
    for (int r0 = 0; r0 < np && e == hipSuccess; r0 += 65535)
        launch_sv_rows(st, G, nu, dS, nvol, d_R + r0, std::min(65535, np - r0), 0, K, ldk,
                       3 * (int64_t)r0, col0, nullptr, 0);
    for (int r0 = 0; r0 < nvol && e == hipSuccess; r0 += 65535)
        launch_sv_rows(st, G, nu, dS, nvol, d_R + np + r0, std::min(65535, nvol - r0), 1, K, ldk,
                       3 * (int64_t)np + 6 * (int64_t)r0, col0, nullptr, 0);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (dS) (void)hipFree(dS);
    if (e != hipSuccess) {
        msg = std::string("strain-volume assembly: ") + hipGetErrorString(e);
        return UCY_ERR_HIP;
    }
    return 0;
}

int ext_build_volume_disp(hipStream_t st, double G, double lambda, int nvol, const double *x,
                          const double *L, const double *T, const double *W, const double *strike,
                          const double *pts, const double *d_pts, int npts, double *O,
                          int64_t ldo, int64_t col0, std::string &msg)
{
    // buildO passes nu = lambda/(lambda+mu)/2 (greens.f90:337)
    const double nu = lambda / (lambda + G) / 2.0;
    std::vector<SvSource> S(nvol);
    for (int i = 0; i < nvol; i++) {
        S[i] = {{x[3 * i], x[3 * i + 1], x[3 * i + 2]}, L[i], T[i], W[i], strike[i]};
        if (sv_bad(nu, 0.0, S[i].q[2])) {
            msg = "strain volume: -1<=nu<=0.5 and source depth >= 0 required";
            return UCY_ERR_GEOMETRY;
        }
    }
    for (int k = 0; k < npts; k++)
        if (pts[3 * k + 2] < 0.0) {  // strainvolume.f90:236-239
            msg = "strain volume: observation depth must be positive";
            return UCY_ERR_GEOMETRY;
        }
    SvSource *dS = nullptr;
    hipError_t e = hipMalloc((void **)&dS, sizeof(SvSource) * nvol);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dS, S.data(), sizeof(SvSource) * nvol, hipMemcpyHostToDevice, st);
    for (int k0 = 0; k0 < npts && e == hipSuccess; k0 += 65535) {
        const int nk = std::min(65535, npts - k0);
        hipLaunchKernelGGL(sv_disp_kernel, dim3((nvol + 127) / 128, nk), dim3(128), 0, st, G, nu,
                           dS, nvol, d_pts + 3 * (int64_t)k0, O + 3 * (int64_t)k0 * ldo, ldo,
                           col0);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (dS) (void)hipFree(dS);
    if (e != hipSuccess) {
        msg = std::string("strain-volume displacement kernels: ") + hipGetErrorString(e);
        return UCY_ERR_HIP;
    }
    return 0;
}

}  // namespace ucy

using namespace ucy;

// one receiver, ns sources, eigenstrain (e11p..e33p): the six unit-eigenstrain
// columns from the device, combined on the host
static int sv_hook(ucy_ctx *c, int kind, const double x[3], const double sv[3], const double dv[3],
                   const double nv[3], int ns, const double *y, const double *L, const double *T,
                   const double *W, const double *strike, const double eps[6], double G,
                   double lambda, double **out)
{
    if (!c || ns < 0) return UCY_ERR_ARG;
    if (ns == 0) return UCY_OK;
    const double nu = lambda / (lambda + G) / 2.0;
    if (sv_bad(nu, x[2], 0.0)) {
        c->err = "strain volume: -1<=nu<=0.5 and observation depth >= 0 required";
        return UCY_ERR_GEOMETRY;
    }
    std::vector<SvSource> S(ns);
    for (int i = 0; i < ns; i++) {
        S[i] = {{y[3 * i], y[3 * i + 1], y[3 * i + 2]}, L[i], T[i], W[i], strike[i]};
        if (S[i].q[2] < 0.0) {
            c->err = "strain volume: source depth (q3) must be positive";
            return UCY_ERR_GEOMETRY;
        }
    }
    Receiver R;
    for (int k = 0; k < 3; k++) {
        R.xc[k] = x[k];
        R.sv[k] = sv[k];
        R.dv[k] = dv[k];
        R.nv[k] = nv[k];
    }
    if (hipSetDevice(c->device) != hipSuccess) return UCY_ERR_HIP;
    const int NC = kind == 0 ? 3 : 6;
    SvSource *dS = nullptr;
    Receiver *dR = nullptr;
    double *dout = nullptr;
    std::vector<double> h((size_t)6 * NC * ns);
    hipError_t e = hipMalloc((void **)&dS, sizeof(SvSource) * ns);
    if (e == hipSuccess) e = hipMalloc((void **)&dR, sizeof(Receiver));
    if (e == hipSuccess) e = hipMalloc((void **)&dout, sizeof(double) * h.size());
    if (e == hipSuccess) e = hipMemcpyAsync(dS, S.data(), sizeof(SvSource) * ns, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dR, &R, sizeof(Receiver), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_sv_rows(c->stream, G, nu, dS, ns, dR, 1, kind, nullptr, 0, 0, 0, dout, ns);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dout, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dS);
    (void)hipFree(dR);
    (void)hipFree(dout);
    if (e != hipSuccess) {
        c->err = std::string("strain-volume kernels: ") + hipGetErrorString(e);
        return UCY_ERR_HIP;
    }
    for (int k = 0; k < NC; k++)
        for (int i = 0; i < ns; i++) {
            double acc = 0.0;
            for (int j = 0; j < 6; j++)
                if (eps[j] != 0.0) acc += eps[j] * h[(size_t)(j * NC + k) * ns + i];
            out[k][i] = acc;
        }
    return UCY_OK;
}

extern "C" int ucy_strainvolume_traction_kernels(ucy_ctx *c, const double x[3], const double sv[3],
                                                 const double dv[3], const double nv[3], int ns,
                                                 const double *y, const double *L, const double *T,
                                                 const double *W, const double *strike,
                                                 double e11p, double e12p, double e13p,
                                                 double e22p, double e23p, double e33p, double G,
                                                 double lambda, double *ts, double *td, double *tn)
{
    const double eps[6] = {e11p, e12p, e13p, e22p, e23p, e33p};
    double *out[3] = {ts, td, tn};
    return sv_hook(c, 0, x, sv, dv, nv, ns, y, L, T, W, strike, eps, G, lambda, out);
}

extern "C" int ucy_strainvolume_stress_kernels(ucy_ctx *c, const double x[3], const double sv[3],
                                               const double dv[3], const double nv[3], int ns,
                                               const double *y, const double *L, const double *T,
                                               const double *W, const double *strike,
                                               double e11p, double e12p, double e13p, double e22p,
                                               double e23p, double e33p, double G, double lambda,
                                               double *s11, double *s12, double *s13, double *s22,
                                               double *s23, double *s33)
{
    const double eps[6] = {e11p, e12p, e13p, e22p, e23p, e33p};
    double *out[6] = {s11, s12, s13, s22, s23, s33};
    return sv_hook(c, 1, x, sv, dv, nv, ns, y, L, T, W, strike, eps, G, lambda, out);
}
