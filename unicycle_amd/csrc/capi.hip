// capi.hip -- host side of the C ABI declared in include/ucy.h.
//
// Context = one GPU (one MPI rank of the reference).  Geometry and friction
// arrive in the reference's layouts; per-element transcendental constants
// (frames, cos/sin of strike, DCCON0 dip terms, Vpl*cos(rake)) are formed
// here with the C library exactly as the reference forms them, so the device
// kernels only see the binary64 values the reference would.
//
// Build note: compiled with -fno-builtin-sin -fno-builtin-cos so the host
// compiler does not fuse sin/cos pairs into sincos, whose results differ in
// the last bit from separate sin and cos calls.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/ucy.h"
#include "ucy_internal.h"
#include "ucy_ext.h"
#include "ucy_ctx.h"

using namespace ucy;

namespace ucy {
void launch_commit(hipStream_t st, int dim, const RkBuffers &b, Control *ctl);
}

static int fail(ucy_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(c, UCY_ERR_HIP, "%s failed: %s (%s:%d)", #call,                   \
                        hipGetErrorString(e_), __FILE__, __LINE__);                       \
    } while (0)

static int ensure_buf(ucy_ctx *c, DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes) return 0;
    if (b.p) HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    HIPCHK(hipMalloc(&b.p, bytes));
    b.bytes = bytes;
    return 0;
}

// ---------------------------------------------------------------------------
// host-side reference arithmetic
// ---------------------------------------------------------------------------
static const double kRad2Deg = 57.29577951308232;  // okada92.f90:133

OkadaMedium make_medium(double G, double lambda)
{
    OkadaMedium M{};
    const float alpha = (float)((lambda + G) / (lambda + 2.0 * G));  // okada92.f90:138
    const double a = (double)alpha;                                 // dc3d.f:45
    M.alp1 = (1.0 - a) / 2.0;
    M.alp2 = a / 2.0;
    M.alp3 = (1.0 - a) / a;
    M.alp4 = 1.0 - a;
    M.alp5 = a;
    M.G = G;
    M.lambda = lambda;
    return M;
}

OkadaSource make_source(const double *x, double L, double W, double strike, double dipd)
{
    OkadaSource S{};
    S.q1 = x[0];
    S.q2 = x[1];
    S.csst = cos(strike);
    S.ssst = sin(strike);
    S.depth = (float)x[2];
    S.al2 = (float)L;
    S.aw1 = -(float)W;
    const double ddip = (double)(float)(dipd * kRad2Deg);  // okada92.f90:153, dc3d.f:46
    const double p18 = 6.283185307179586 / 360.0;          // dc3d.f:531-533
    double sd = sin(ddip * p18), cd = cos(ddip * p18);
    if (fabs(cd) < 1.0e-6) {
        cd = 0.0;
        if (sd > 0.0) sd = 1.0;
        if (sd < 0.0) sd = -1.0;
    }
    S.sd = sd;
    S.cd = cd;
    S.sdsd = sd * sd;
    S.cdcd = cd * cd;
    S.sdcd = sd * cd;
    return S;
}

// okada92.f90:59-87
void refsys_okada(int ns, const double *x, const double *L, const double *W,
                         const double *strike, const double *dip, double *sv, double *dv,
                         double *nv, double *xc)
{
    for (int i = 0; i < ns; i++) {
        const double cs = cos(strike[i]), ss = sin(strike[i]);
        const double cdp = cos(dip[i]), sdp = sin(dip[i]);
        sv[3 * i] = cs;
        sv[3 * i + 1] = ss;
        sv[3 * i + 2] = 0.0;
        dv[3 * i] = +ss * cdp;
        dv[3 * i + 1] = -cs * cdp;
        dv[3 * i + 2] = -sdp;
        nv[3 * i] = -ss * sdp;
        nv[3 * i + 1] = +cs * sdp;
        nv[3 * i + 2] = -cdp;
        for (int k = 0; k < 3; k++)
            xc[3 * i + k] = x[3 * i + k] + L[i] / 2 * sv[3 * i + k] - W[i] / 2 * dv[3 * i + k];
    }
}

// nikkhoo15.f90:50-101 (identical to stuart97.f90:44-95); 1-based indices
static void refsys_triangle(int nVe, const double *v, int ns, const int *i1, const int *i2,
                            const int *i3, double *sv, double *dv, double *nv, double *xc)
{
    (void)nVe;
    for (int j = 0; j < ns; j++) {
        const double *A = v + 3 * (i1[j] - 1), *B = v + 3 * (i2[j] - 1), *C = v + 3 * (i3[j] - 1);
        double n0 = (B[1] - A[1]) * (C[2] - A[2]) - (B[2] - A[2]) * (C[1] - A[1]);
        double n1 = (B[2] - A[2]) * (C[0] - A[0]) - (B[0] - A[0]) * (C[2] - A[2]);
        double n2 = (B[0] - A[0]) * (C[1] - A[1]) - (B[1] - A[1]) * (C[0] - A[0]);
        const double norm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        n0 = n0 / norm;
        n1 = n1 / norm;
        n2 = n2 / norm;
        if (0.0 < n2) {
            n0 = -n0;
            n1 = -n1;
            n2 = -n2;
        }
        double *N = nv + 3 * j, *S = sv + 3 * j, *D = dv + 3 * j;
        N[0] = n0;
        N[1] = n1;
        N[2] = n2;
        S[0] = cos(atan2(n0, n1));
        S[1] = -sin(atan2(n0, n1));
        S[2] = 0.0;
        D[0] = N[1] * S[2] - N[2] * S[1];
        D[1] = N[2] * S[0] - N[0] * S[2];
        D[2] = N[0] * S[1] - N[1] * S[0];
        for (int k = 0; k < 3; k++) xc[3 * j + k] = (A[k] + B[k] + C[k]) / 3.0;
    }
}

// strainvolume.f90:135-163
static void refsys_volume(int ns, const double *x, const double *L, const double *W,
                          const double *strike, double *sv, double *dv, double *nv, double *xc)
{
    for (int i = 0; i < ns; i++) {
        sv[3 * i] = cos(strike[i]);
        sv[3 * i + 1] = sin(strike[i]);
        sv[3 * i + 2] = 0.0;
        dv[3 * i] = 0.0;
        dv[3 * i + 1] = 0.0;
        dv[3 * i + 2] = -1.0;
        nv[3 * i] = -sin(strike[i]);
        nv[3 * i + 1] = +cos(strike[i]);
        nv[3 * i + 2] = -0.0;
        for (int k = 0; k < 3; k++)
            xc[3 * i + k] = x[3 * i + k] + L[i] / 2 * sv[3 * i + k] - W[i] / 2 * dv[3 * i + k];
    }
}

// ---------------------------------------------------------------------------
// lifecycle
// ---------------------------------------------------------------------------
extern "C" const char *ucy_version(void) { return "ucy-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// Static scratch of ucy_odefun_dev_static: a file-scope device array (each
// GPU holds its own copy until the process ends).  Host code never uses the
// array's name as an address: ucy_create resolves it once per device with
// hipGetSymbolAddress and every later use goes through that pointer.
// Layout (doubles): [0] the vMax accumulator, padded to 256 bytes, then the
// slip-rate vector v of up to kStaticMaxCols entries.
// ---------------------------------------------------------------------------
constexpr size_t kStaticBytes = size_t(1) << 20;                 // 1 MiB per device
constexpr int64_t kStaticLead = 32;                              // doubles before v
constexpr int64_t kStaticMaxCols = (int64_t)(kStaticBytes / sizeof(double)) - kStaticLead;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kStaticBytes / sizeof(double)];
static std::atomic<double *> g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// with `device` current: the device address of this GPU's scratch
static bool resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return false;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device].load()) return true;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return false;
    g_scratch[device].store(static_cast<double *>(p));
    return true;
}

extern "C" ucy_ctx *ucy_create(int device, int flags)
{
    (void)flags;
    ucy_ctx *c = new ucy_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return nullptr;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    if (!resolve_scratch(device)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return nullptr;
    }
    if (hipHostMalloc((void **)&c->h_ctl, sizeof(Control), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&c->d_ctl, sizeof(Control)) != hipSuccess) {
        delete c;
        return nullptr;
    }
    return c;
}

static void free_work(ucy_ctx *c)
{
    if (c->graph) hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    if (c->work) hipFree(c->work);
    c->work = nullptr;
    c->work_dim = 0;
}

extern "C" void ucy_destroy(ucy_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    while (!c->ensembles.empty()) ucy_ensemble_destroy(c->ensembles.back());
    free_work(c);
    if (c->K) hipFree(c->K);
    if (c->Kp) hipFree(c->Kp);
    if (c->Krs) hipFree(c->Krs);
    if (c->O) hipFree(c->O);
    if (c->d_fr) hipFree(c->d_fr);
    if (c->d_rh) hipFree(c->d_rh);
    if (c->d_ctl) hipFree(c->d_ctl);
    if (c->h_ctl) hipHostFree(c->h_ctl);
    if (c->hook.p) hipFree(c->hook.p);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *ucy_last_error(const ucy_ctx *c) { return c ? c->err.c_str() : "null context"; }

// ---------------------------------------------------------------------------
// reference frames
// ---------------------------------------------------------------------------
extern "C" int ucy_reference_system_okada92(int ns, const double *x, const double *L,
                                            const double *W, const double *strike,
                                            const double *dip, double *sv, double *dv,
                                            double *nv, double *xc)
{
    if (ns < 0) return UCY_ERR_ARG;
    refsys_okada(ns, x, L, W, strike, dip, sv, dv, nv, xc);
    return UCY_OK;
}

extern "C" int ucy_reference_system_nikkhoo15(int nVe, const double *v, int ns, const int *i1,
                                              const int *i2, const int *i3, double *sv,
                                              double *dv, double *nv, double *xc)
{
    if (ns < 0 || nVe < 0) return UCY_ERR_ARG;
    for (int j = 0; j < ns; j++)
        if (i1[j] < 1 || i1[j] > nVe || i2[j] < 1 || i2[j] > nVe || i3[j] < 1 || i3[j] > nVe)
            return UCY_ERR_ARG;
    refsys_triangle(nVe, v, ns, i1, i2, i3, sv, dv, nv, xc);
    return UCY_OK;
}

extern "C" int ucy_reference_system_strainvolume(int ns, const double *x, const double *L,
                                                 const double *W, const double *strike,
                                                 double *sv, double *dv, double *nv, double *xc)
{
    if (ns < 0) return UCY_ERR_ARG;
    refsys_volume(ns, x, L, W, strike, sv, dv, nv, xc);
    return UCY_OK;
}

// ---------------------------------------------------------------------------
// model definition
// ---------------------------------------------------------------------------
static int upload_params(ucy_ctx *c);

// set-up calls place the friction and rheology on the device as soon as the
// model is complete (every patch has friction, every volume a rheology), so
// that entry points which may not allocate find them there
static int upload_if_complete(ucy_ctx *c)
{
    const int np = c->nrect + c->ntri;
    if (np + c->nvol == 0) return UCY_OK;
    if (np && (!c->have_friction || (int)c->Vpl.size() != np)) return UCY_OK;
    if (c->nvol && (!c->have_rheology || (int)c->gd0.size() != c->nvol)) return UCY_OK;
    HIPCHK(hipSetDevice(c->device));
    return upload_params(c);
}

extern "C" int ucy_set_elastic(ucy_ctx *c, double lambda, double mu)
{
    if (!c) return UCY_ERR_ARG;
    if (!(mu > 0)) return fail(c, UCY_ERR_ARG, "shear modulus must be positive");
    const double nu = lambda / 2.0 / (lambda + mu);  // ratestate.f90:1102
    if (nu < -1.0 || nu > 0.5) return fail(c, UCY_ERR_ARG, "Poisson's ratio out of [-1,0.5]");
    c->lambda = lambda;
    c->mu = mu;
    c->have_elastic = true;
    return UCY_OK;
}

extern "C" int ucy_set_rectangles(ucy_ctx *c, int ns, const double *x, const double *length,
                                  const double *width, const double *strike, const double *dip)
{
    if (!c || ns < 0) return UCY_ERR_ARG;
    for (int i = 0; i < ns; i++)
        if (!(std::max(length[i], width[i]) > 0.0))  // ratestate.f90:1184
            return fail(c, UCY_ERR_ARG, "patch %d: length and width must be positive", i + 1);
    c->model = 0;
    c->nrect = ns;
    c->rx.assign(x, x + 3 * (size_t)ns);
    c->rL.assign(length, length + ns);
    c->rW.assign(width, width + ns);
    c->rstrike.assign(strike, strike + ns);
    c->rdip.assign(dip, dip + ns);
    c->params_dirty = true;
    return upload_if_complete(c);
}

// the 2-D antiplane model (fortran/2d/antiplane): ns segments with upper
// corner x(3,ns) (x(2), x(3) used), width and dip (radians); replaces any
// 3-D elements
extern "C" int ucy_set_antiplane(ucy_ctx *c, double mu, int ns, const double *x,
                                 const double *width, const double *dip)
{
    if (!c || ns < 0) return UCY_ERR_ARG;
    if (!(mu > 0.0)) return fail(c, UCY_ERR_ARG, "shear modulus must be positive");
    for (int i = 0; i < ns; i++)
        if (!(width[i] > 0.0))  // ratestate.f90 (2-D antiplane):1090-1094
            return fail(c, UCY_ERR_ARG, "patch %d: width must be positive", i + 1);
    c->model = 1;
    c->mu = mu;
    c->lambda = 0.0;
    c->have_elastic = true;
    c->nrect = ns;
    c->rx.assign(x, x + 3 * (size_t)ns);
    c->rW.assign(width, width + ns);
    c->rL.assign(width, width + ns);
    c->rstrike.assign(ns, 0.0);
    c->rdip.assign(dip, dip + ns);
    c->ntri = c->nvol = c->nve = 0;
    c->params_dirty = true;
    return upload_if_complete(c);
}

// antiplane.f90:81-109
static void refsys_antiplane(int ns, const double *x, const double *W, const double *dip,
                             std::vector<Receiver> &R)
{
    R.resize(ns);
    for (int i = 0; i < ns; i++) {
        Receiver &r = R[i];
        r.sv[0] = 1.0;
        r.sv[1] = 0.0;
        r.sv[2] = 0.0;
        r.dv[0] = +0.0;
        r.dv[1] = -cos(dip[i]);
        r.dv[2] = -sin(dip[i]);
        r.nv[0] = -0.0;
        r.nv[1] = +sin(dip[i]);
        r.nv[2] = -cos(dip[i]);
        for (int k = 0; k < 3; k++) r.xc[k] = x[3 * i + k] - W[i] / 2 * r.dv[k];
    }
}

extern "C" int ucy_set_triangles(ucy_ctx *c, int ns, const int *i1, const int *i2, const int *i3,
                                 int nVe, const double *v)
{
    if (!c || ns < 0 || nVe < 0) return UCY_ERR_ARG;
    for (int j = 0; j < ns; j++)
        if (i1[j] < 1 || i1[j] > nVe || i2[j] < 1 || i2[j] > nVe || i3[j] < 1 || i3[j] > nVe)
            return fail(c, UCY_ERR_ARG, "triangle %d: vertex index out of range", j + 1);
    if (c->model) c->nrect = 0;
    c->model = 0;
    c->ntri = ns;
    c->nve = nVe;
    c->i1.assign(i1, i1 + ns);
    c->i2.assign(i2, i2 + ns);
    c->i3.assign(i3, i3 + ns);
    c->vert.assign(v, v + 3 * (size_t)nVe);
    c->params_dirty = true;
    return upload_if_complete(c);
}

extern "C" int ucy_set_volumes(ucy_ctx *c, int ns, const double *x, const double *length,
                               const double *thickness, const double *width, const double *strike)
{
    if (!c || ns < 0) return UCY_ERR_ARG;
    for (int i = 0; i < ns; i++)
        if (x[3 * i + 2] < 0)  // strainvolume.f90:1069-1072
            return fail(c, UCY_ERR_GEOMETRY, "volume %d: source depth (q3) must be positive", i + 1);
    if (c->model) c->nrect = 0;
    c->model = 0;
    c->nvol = ns;
    c->vx.assign(x, x + 3 * (size_t)ns);
    c->vL.assign(length, length + ns);
    c->vT.assign(thickness, thickness + ns);
    c->vW.assign(width, width + ns);
    c->vstrike.assign(strike, strike + ns);
    c->params_dirty = true;
    return upload_if_complete(c);
}

// strain volumes of the 2-D antiplane model (fortran/2d/antiplane/src/
// viscouscycles.f90:1404-1474): upper centre x(3,nv) (x(2), x(3) used),
// thickness (horizontal), width (depth extent) and dip (radians, used by the
// frame only); after ucy_set_antiplane, which clears them
extern "C" int ucy_set_antiplane_volumes(ucy_ctx *c, int nv, const double *x,
                                         const double *thickness, const double *width,
                                         const double *dip)
{
    if (!c || nv < 0 || (nv && (!x || !thickness || !width || !dip))) return UCY_ERR_ARG;
    if (c->model != 1) return fail(c, UCY_ERR_STATE, "no antiplane model: call ucy_set_antiplane first");
    for (int i = 0; i < nv; i++) {
        if (x[3 * i + 2] < 0)  // viscouscycles.f90 (2-D antiplane):1455-1460
            return fail(c, UCY_ERR_GEOMETRY, "volume %d: depth must be positive", i + 1);
        if (!(thickness[i] > 0.0) || !(width[i] > 0.0))  // :1469-1473
            return fail(c, UCY_ERR_ARG, "volume %d: dimensions must be positive", i + 1);
    }
    c->nvol = nv;
    c->vx.assign(x, x + 3 * (size_t)nv);
    c->vT.assign(thickness, thickness + nv);
    c->vW.assign(width, width + nv);
    c->vdip.assign(dip, dip + nv);
    c->vL.assign(nv, 0.0);
    c->vstrike.assign(nv, 0.0);
    c->params_dirty = true;
    return upload_if_complete(c);
}

// the 2-D plane-strain model (fortran/2d/planestrain): ns segments with
// upper corner x(3,ns) (x(2), x(3) used), width and dip (radians); replaces
// any other elements
extern "C" int ucy_set_planestrain(ucy_ctx *c, double lambda, double mu, int ns, const double *x,
                                   const double *width, const double *dip)
{
    if (!c || ns < 0 || (ns && (!x || !width || !dip))) return UCY_ERR_ARG;
    if (0 > mu) return fail(c, UCY_ERR_ARG, "shear modulus must be positive");  // ratestate.f90:1046-1051
    for (int i = 0; i < ns; i++)
        if (!(width[i] > 0.0))  // ratestate.f90 (2-D plane strain):1110-1114
            return fail(c, UCY_ERR_ARG, "patch %d: width must be positive", i + 1);
    c->model = 2;
    c->mu = mu;
    c->lambda = lambda;
    c->have_elastic = true;
    c->nrect = ns;
    c->rx.assign(x, x + 3 * (size_t)ns);
    c->rW.assign(width, width + ns);
    c->rL.assign(width, width + ns);
    c->rstrike.assign(ns, 0.0);
    c->rdip.assign(dip, dip + ns);
    c->ntri = c->nvol = c->nve = 0;
    c->tau0.clear();
    c->s0.clear();
    c->params_dirty = true;
    return upload_if_complete(c);
}

// strain volumes of the 2-D plane-strain model (fortran/2d/planestrain/src/
// viscouscycles.f90:1535-1578): top x(3,nv) (x(2), x(3) used), thickness,
// width and dip (radians); after ucy_set_planestrain, which clears them
extern "C" int ucy_set_planestrain_volumes(ucy_ctx *c, int nv, const double *x,
                                           const double *thickness, const double *width,
                                           const double *dip)
{
    if (!c || nv < 0 || (nv && (!x || !thickness || !width || !dip))) return UCY_ERR_ARG;
    if (c->model != 2)
        return fail(c, UCY_ERR_STATE, "no plane-strain model: call ucy_set_planestrain first");
    for (int i = 0; i < nv; i++) {
        if (0 > x[3 * i + 2])  // :1559-1564
            return fail(c, UCY_ERR_GEOMETRY, "volume %d: depth must be positive", i + 1);
        if (!(std::min(width[i], thickness[i]) > 0.0))  // :1572-1576
            return fail(c, UCY_ERR_ARG, "volume %d: strain volume dimension must be positive",
                        i + 1);
    }
    c->nvol = nv;
    c->vx.assign(x, x + 3 * (size_t)nv);
    c->vT.assign(thickness, thickness + nv);
    c->vW.assign(width, width + nv);
    c->vdip.assign(dip, dip + nv);
    c->vL.assign(nv, 0.0);
    c->vstrike.assign(nv, 0.0);
    c->s0.clear();
    c->params_dirty = true;
    return upload_if_complete(c);
}

extern "C" int ucy_set_initial_volume_stress(ucy_ctx *c, int nv, const double *s0)
{
    if (!c || nv < 0 || (nv && !s0)) return UCY_ERR_ARG;
    if (c->model != 2) return fail(c, UCY_ERR_STATE, "initial volume stress: plane-strain model only");
    c->s0.assign(s0, s0 + 3 * (size_t)nv);
    return UCY_OK;
}

extern "C" int ucy_set_initial_traction(ucy_ctx *c, int np, const double *tau0)
{
    if (!c || np < 0 || (np && !tau0)) return UCY_ERR_ARG;
    c->tau0.assign(tau0, tau0 + np);
    return UCY_OK;
}

// plane-strain sources: top, width, sin / cos of 1..4 times the dip
static std::vector<PsSource> ps_sources(int ns, const double *x, const double *W,
                                        const double *dip)
{
    std::vector<PsSource> S(ns);
    for (int i = 0; i < ns; i++) {
        S[i].q2 = x[3 * i + 1];
        S[i].q3 = x[3 * i + 2];
        S[i].W = W[i];
        for (int k = 0; k < 4; k++) {
            S[i].sk[k] = sin((k + 1.0) * dip[i]);
            S[i].ck[k] = cos((k + 1.0) * dip[i]);
        }
    }
    return S;
}

// the parameter checks of computeStressStrainVolumePlaneStrain.f90:75-86
static int ps_check(ucy_ctx *c, double nu, int nr, const double *x3, int xstride, int ns,
                    const double *q3, int qstride)
{
    if (-1.0 > nu || 0.5 < nu)
        return fail(c, UCY_ERR_GEOMETRY, "Poisson's ratio %g outside [-1, 0.5]", nu);
    for (int i = 0; i < nr; i++)
        if (0 > x3[(size_t)xstride * i])
            return fail(c, UCY_ERR_GEOMETRY, "observation depth (x3) must be positive");
    for (int i = 0; i < ns; i++)
        if (0 > q3[(size_t)qstride * i])
            return fail(c, UCY_ERR_GEOMETRY, "source depth (q3) must be positive");
    return 0;
}

extern "C" int ucy_set_friction(ucy_ctx *c, int np, const double *Vpl, const double *rake,
                                const double *mu0, const double *sig, const double *a,
                                const double *b, const double *L, const double *Vo,
                                const double *damping)
{
    if (!c || np < 0) return UCY_ERR_ARG;
    c->Vpl.assign(Vpl, Vpl + np);
    c->rake.assign(rake, rake + np);
    c->mu0.assign(mu0, mu0 + np);
    c->sig.assign(sig, sig + np);
    c->a.assign(a, a + np);
    c->b.assign(b, b + np);
    c->Lc.assign(L, L + np);
    c->Vo.assign(Vo, Vo + np);
    c->damping.assign(damping, damping + np);
    c->have_friction = true;
    c->params_dirty = true;
    return upload_if_complete(c);
}

extern "C" int ucy_set_rheology(ucy_ctx *c, int nv, const double *gammadot0,
                                const double *npower, const double *edot_bg, const double *Gm,
                                const double *Q, const double *R, const double *To)
{
    if (!c || nv < 0) return UCY_ERR_ARG;
    c->gd0.assign(gammadot0, gammadot0 + nv);
    c->npow.assign(npower, npower + nv);
    c->ebg.assign(edot_bg, edot_bg + 6 * (size_t)nv);
    c->Gm.assign(Gm, Gm + nv);
    c->Qa.assign(Q, Q + nv);
    c->Rg.assign(R, R + nv);
    c->To.assign(To, To + nv);
    c->have_rheology = true;
    c->params_dirty = true;
    return upload_if_complete(c);
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
static void element_frames(const ucy_ctx *c, std::vector<Receiver> &R)
{
    const int ne = c->nrect + c->ntri + c->nvol;
    std::vector<double> sv(3 * (size_t)ne + 3), dv(3 * (size_t)ne + 3), nv(3 * (size_t)ne + 3),
        xc(3 * (size_t)ne + 3);
    if (c->nrect)
        refsys_okada(c->nrect, c->rx.data(), c->rL.data(), c->rW.data(), c->rstrike.data(),
                     c->rdip.data(), sv.data(), dv.data(), nv.data(), xc.data());
    int o = c->nrect;
    if (c->ntri)
        refsys_triangle(c->nve, c->vert.data(), c->ntri, c->i1.data(), c->i2.data(), c->i3.data(),
                        sv.data() + 3 * o, dv.data() + 3 * o, nv.data() + 3 * o, xc.data() + 3 * o);
    o += c->ntri;
    if (c->nvol)
        refsys_volume(c->nvol, c->vx.data(), c->vL.data(), c->vW.data(), c->vstrike.data(),
                      sv.data() + 3 * o, dv.data() + 3 * o, nv.data() + 3 * o, xc.data() + 3 * o);
    R.resize(ne);
    for (int e = 0; e < ne; e++)
        for (int k = 0; k < 3; k++) {
            R[e].xc[k] = xc[3 * e + k];
            R[e].sv[k] = sv[3 * e + k];
            R[e].dv[k] = dv[3 * e + k];
            R[e].nv[k] = nv[3 * e + k];
        }
}

extern "C" int ucy_build_kernel(ucy_ctx *c, int precision)
{
    if (!c) return UCY_ERR_ARG;
    if (precision != UCY_PREC_F64 && precision != UCY_PREC_SPLIT_BF16 &&
        precision != UCY_PREC_MIXED)
        return fail(c, UCY_ERR_ARG, "unknown precision %d", precision);
    if (!c->have_elastic) return fail(c, UCY_ERR_STATE, "elastic moduli not set");
    HIPCHK(hipSetDevice(c->device));
    const int np = c->nrect + c->ntri, ne = np + c->nvol;
    if (ne == 0) return fail(c, UCY_ERR_STATE, "no elements");
    const bool ap = c->model == 1, ps = c->model == 2;
    const int64_t nap = np + 2 * (int64_t)c->nvol;  // antiplane rows and columns
    const int64_t nps = np + 3 * (int64_t)c->nvol;  // plane-strain columns
    const int64_t m = ap ? nap
                    : ps ? 2 * (int64_t)np + 3 * (int64_t)c->nvol
                         : 3 * (int64_t)np + 6 * (int64_t)c->nvol;
    const int64_t n = ap ? nap + (nap & 1)
                    : ps ? nps + (nps & 1) : 2 * (int64_t)np + 6 * (int64_t)c->nvol;
    if (ps) {
        // receivers (patches, then volumes) and sources as the kernels see
        // them; the volume rows of patch sources take G and lambda exchanged
        std::vector<Receiver> R0, RV0;
        refsys_antiplane(np, c->rx.data(), c->rW.data(), c->rdip.data(), R0);
        refsys_antiplane(c->nvol, c->vx.data(), c->vW.data(), c->vdip.data(), RV0);
        R0.insert(R0.end(), RV0.begin(), RV0.end());
        const int stride = (int)(sizeof(Receiver) / sizeof(double));
        const double nu = c->lambda / 2.0 / (c->lambda + c->mu);
        int rc = ps_check(c, nu, np + c->nvol, R0.empty() ? nullptr : &R0[0].xc[2], stride, np,
                          c->rx.data() + 2, 3);
        if (!rc) rc = ps_check(c, nu, 0, nullptr, 1, c->nvol, c->vx.data() + 2, 3);
        if (!rc && c->nvol && np)
            rc = ps_check(c, c->mu / 2.0 / (c->mu + c->lambda), 0, nullptr, 1, 0, nullptr, 1);
        if (rc) return rc;
    }
    free_work(c);
    if (c->K) HIPCHK(hipFree(c->K));
    c->K = nullptr;
    if (c->Kp) HIPCHK(hipFree(c->Kp));
    c->Kp = nullptr;
    if (c->Krs) HIPCHK(hipFree(c->Krs));
    c->Krs = nullptr;
    c->kprec = UCY_PREC_F64;
    if (n % 2) return fail(c, UCY_ERR_STATE, "odd column count");
    HIPCHK(hipMalloc((void **)&c->K, (size_t)(m * n) * sizeof(double)));
    c->m = m;
    c->n = n;
    c->nl = ap ? nap : ps ? nps : n;
    std::vector<Receiver> R;
    if (ps) {  // buildG of fortran/2d/planestrain (greens.f90:66-226)
        // receivers: the patches, then the volumes (the same frames,
        // planestrain.f90:55-83 for both, viscouscycles.f90:1187-1213);
        // sources: the patches, then the volumes with their thicknesses
        const int nv = c->nvol, ne = np + nv;
        std::vector<Receiver> RV;
        refsys_antiplane(np, c->rx.data(), c->rW.data(), c->rdip.data(), R);
        refsys_antiplane(nv, c->vx.data(), c->vW.data(), c->vdip.data(), RV);
        R.insert(R.end(), RV.begin(), RV.end());
        std::vector<PsSource> S = ps_sources(np, c->rx.data(), c->rW.data(), c->rdip.data());
        const std::vector<PsSource> SV = ps_sources(nv, c->vx.data(), c->vW.data(), c->vdip.data());
        S.insert(S.end(), SV.begin(), SV.end());
        Receiver *d_R = nullptr;
        PsSource *d_S = nullptr;
        double *d_T = nullptr;
        HIPCHK(hipMalloc((void **)&d_R, sizeof(Receiver) * ne));
        HIPCHK(hipMalloc((void **)&d_S, sizeof(PsSource) * ne));
        HIPCHK(hipMalloc((void **)&d_T, sizeof(double) * (nv + 1)));
        HIPCHK(hipMemcpyAsync(d_R, R.data(), sizeof(Receiver) * ne, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(d_S, S.data(), sizeof(PsSource) * ne, hipMemcpyHostToDevice,
                              c->stream));
        if (nv)
            HIPCHK(hipMemcpyAsync(d_T, c->vT.data(), sizeof(double) * nv, hipMemcpyHostToDevice,
                                  c->stream));
        auto t0 = std::chrono::steady_clock::now();
        if (n != nps) HIPCHK(hipMemsetAsync(c->K, 0, (size_t)(m * n) * sizeof(double), c->stream));
        const double nu = c->lambda / 2.0 / (c->lambda + c->mu);
        if (np)
            launch_ps_rows(c->stream, d_S, np, d_R, np, c->mu, c->lambda, nu, c->K, n, nullptr,
                           nullptr);
        if (nv)
            launch_ps_blocks(c->stream, d_S, np, d_S + np, d_T, nv, d_R, c->mu, c->lambda, c->K, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        c->build_seconds[0] =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        c->build_seconds[1] = c->build_seconds[2] = 0.0;
        HIPCHK(hipFree(d_R));
        HIPCHK(hipFree(d_S));
        HIPCHK(hipFree(d_T));
        if (precision != UCY_PREC_F64) {
            HIPCHK(hipMalloc((void **)&c->Kp, (size_t)(m * n) * sizeof(uint32_t)));
            HIPCHK(hipMalloc((void **)&c->Krs, (size_t)m * sizeof(double)));
            launch_split_rowmajor(c->stream, c->K, m, n, c->Krs, c->Kp);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(c->stream));
            c->kprec = precision;
        }
        return UCY_OK;
    }
    if (ap) {  // buildG of fortran/2d/antiplane (greens.f90:93-248)
        // receivers: the patches, then the volumes (frames of
        // computeReferenceSystemAntiplane for both, viscouscycles.f90:1085-1107);
        // sources: (q2, q3, W) per patch, then (q2, q3, W, T) per volume
        const int nv = c->nvol, ne = np + nv;
        std::vector<Receiver> RV;
        refsys_antiplane(np, c->rx.data(), c->rW.data(), c->rdip.data(), R);
        refsys_antiplane(nv, c->vx.data(), c->vW.data(), c->vdip.data(), RV);
        R.insert(R.end(), RV.begin(), RV.end());
        std::vector<double> S(3 * (size_t)np + 4 * (size_t)nv + 1);
        for (int i = 0; i < np; i++) {
            S[3 * i] = c->rx[3 * i + 1];
            S[3 * i + 1] = c->rx[3 * i + 2];
            S[3 * i + 2] = c->rW[i];
        }
        double *SV = S.data() + 3 * (size_t)np;
        for (int i = 0; i < nv; i++) {
            SV[4 * i] = c->vx[3 * i + 1];
            SV[4 * i + 1] = c->vx[3 * i + 2];
            SV[4 * i + 2] = c->vW[i];
            SV[4 * i + 3] = c->vT[i];
        }
        Receiver *d_R = nullptr;
        double *d_S = nullptr;
        HIPCHK(hipMalloc((void **)&d_R, sizeof(Receiver) * ne));
        HIPCHK(hipMalloc((void **)&d_S, sizeof(double) * S.size()));
        HIPCHK(hipMemcpyAsync(d_R, R.data(), sizeof(Receiver) * ne, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(d_S, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice,
                              c->stream));
        auto t0 = std::chrono::steady_clock::now();
        if (n != nap) HIPCHK(hipMemsetAsync(c->K, 0, (size_t)(m * n) * sizeof(double), c->stream));
        if (np) launch_ap_rows(c->stream, d_S, np, d_R, np, c->mu, c->K, n, nullptr);
        if (nv) launch_ap_volume_blocks(c->stream, d_S, np, d_S + 3 * (size_t)np, nv, d_R, c->mu,
                                        c->K, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        c->build_seconds[0] =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        c->build_seconds[1] = c->build_seconds[2] = 0.0;
        HIPCHK(hipFree(d_R));
        HIPCHK(hipFree(d_S));
        if (precision != UCY_PREC_F64) {
            HIPCHK(hipMalloc((void **)&c->Kp, (size_t)(m * n) * sizeof(uint32_t)));
            HIPCHK(hipMalloc((void **)&c->Krs, (size_t)m * sizeof(double)));
            launch_split_rowmajor(c->stream, c->K, m, n, c->Krs, c->Kp);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(c->stream));
            c->kprec = precision;
        }
        return UCY_OK;
    }
    element_frames(c, R);
    Receiver *d_R = nullptr;
    HIPCHK(hipMalloc((void **)&d_R, sizeof(Receiver) * ne));
    HIPCHK(hipMemcpyAsync(d_R, R.data(), sizeof(Receiver) * ne, hipMemcpyHostToDevice, c->stream));
    int rc = 0;
    using clk = std::chrono::steady_clock;
    auto since = [](clk::time_point t0) {
        return std::chrono::duration<double>(clk::now() - t0).count();
    };
    c->build_seconds[0] = c->build_seconds[1] = c->build_seconds[2] = 0.0;
    HIPCHK(hipStreamSynchronize(c->stream));
    auto t0 = clk::now();
    // rectangle sources -> all receivers
    if (c->nrect) {
        std::vector<OkadaSource> S(c->nrect);
        for (int i = 0; i < c->nrect; i++)
            S[i] = make_source(&c->rx[3 * i], c->rL[i], c->rW[i], c->rstrike[i], c->rdip[i]);
        OkadaSource *d_S = nullptr;
        HIPCHK(hipMalloc((void **)&d_S, sizeof(OkadaSource) * c->nrect));
        HIPCHK(hipMemcpyAsync(d_S, S.data(), sizeof(OkadaSource) * c->nrect,
                              hipMemcpyHostToDevice, c->stream));
        const OkadaMedium M = make_medium(c->mu, c->lambda);
        for (int r0 = 0; r0 < np; r0 += 65535) {
            const int nr = std::min(65535, np - r0);
            launch_okada_rows(c->stream, M, d_S, c->nrect, d_R + r0, nr, 0, c->K, n, 3 * (int64_t)r0,
                              0, nullptr, nullptr, 0);
        }
        for (int r0 = 0; r0 < c->nvol; r0 += 65535) {
            const int nr = std::min(65535, c->nvol - r0);
            launch_okada_rows(c->stream, M, d_S, c->nrect, d_R + np + r0, nr, 1, c->K, n,
                              3 * (int64_t)np + 6 * (int64_t)r0, 0, nullptr, nullptr, 0);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(d_S));
    }
    c->build_seconds[0] = since(t0);
    t0 = clk::now();
    if (c->ntri) rc = ext_build_triangle_columns(c->stream, c->mu, c->lambda, c->ntri, c->i1.data(),
                                                  c->i2.data(), c->i3.data(), c->nve,
                                                  c->vert.data(), d_R, np, c->nvol, c->K, n,
                                                  2 * (int64_t)c->nrect, c->err);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->build_seconds[1] = since(t0);
    t0 = clk::now();
    if (!rc && c->nvol)
        rc = ext_build_volume_columns(c->stream, c->mu, c->lambda, c->nvol, c->vx.data(),
                                      c->vL.data(), c->vT.data(), c->vW.data(),
                                      c->vstrike.data(), d_R, np, c->K, n, 2 * (int64_t)np, c->err);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->build_seconds[2] = since(t0);
    HIPCHK(hipFree(d_R));
    if (!rc && precision != UCY_PREC_F64) {
        // the split pairs of K (4 bytes per entry) beside the binary64 K
        HIPCHK(hipMalloc((void **)&c->Kp, (size_t)(m * n) * sizeof(uint32_t)));
        HIPCHK(hipMalloc((void **)&c->Krs, (size_t)m * sizeof(double)));
        launch_split_rowmajor(c->stream, c->K, m, n, c->Krs, c->Kp);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        c->kprec = precision;
    }
    return rc;
}

extern "C" int ucy_build_seconds(const ucy_ctx *c, double *rect, double *tri, double *vol)
{
    if (!c) return UCY_ERR_ARG;
    if (rect) *rect = c->build_seconds[0];
    if (tri) *tri = c->build_seconds[1];
    if (vol) *vol = c->build_seconds[2];
    return UCY_OK;
}

extern "C" int ucy_kernel_shape(const ucy_ctx *c, int64_t *m, int64_t *n)
{
    if (!c) return UCY_ERR_ARG;
    if (m) *m = c->m;
    if (n) *n = c->nl;
    return UCY_OK;
}

extern "C" const double *ucy_kernel_dev(const ucy_ctx *c) { return c ? c->K : nullptr; }

extern "C" int ucy_get_kernel(ucy_ctx *c, double *K)
{
    if (!c || !K) return UCY_ERR_ARG;
    if (!c->K) return fail(c, UCY_ERR_STATE, "kernel not built");
    HIPCHK(hipMemcpy2DAsync(K, sizeof(double) * c->nl, c->K, sizeof(double) * c->n,
                            sizeof(double) * c->nl, c->m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return UCY_OK;
}

extern "C" int ucy_set_kernel(ucy_ctx *c, const double *K, int64_t m, int64_t n)
{
    if (!c || !K || m <= 0 || n <= 0) return UCY_ERR_ARG;
    const int np = c->nrect + c->ntri;
    const bool shape_ok = c->model == 2
        ? (m == 2 * (int64_t)np + 3 * (int64_t)c->nvol && n == np + 3 * (int64_t)c->nvol)
        : c->model == 1
        ? (m == np + 2 * (int64_t)c->nvol && n == m)
        : (m == 3 * (int64_t)np + 6 * (int64_t)c->nvol && n == 2 * (int64_t)np + 6 * (int64_t)c->nvol);
    if (!shape_ok) return fail(c, UCY_ERR_ARG, "kernel shape does not match the mesh");
    HIPCHK(hipSetDevice(c->device));
    free_work(c);
    if (c->K) HIPCHK(hipFree(c->K));
    c->K = nullptr;
    if (c->Kp) HIPCHK(hipFree(c->Kp));
    c->Kp = nullptr;
    if (c->Krs) HIPCHK(hipFree(c->Krs));
    c->Krs = nullptr;
    c->kprec = UCY_PREC_F64;
    const int64_t ld = c->model ? n + (n & 1) : n;
    if (ld % 2) return fail(c, UCY_ERR_STATE, "odd column count");
    HIPCHK(hipMalloc((void **)&c->K, (size_t)(m * ld) * sizeof(double)));
    if (ld != n) HIPCHK(hipMemsetAsync(c->K, 0, (size_t)(m * ld) * sizeof(double), c->stream));
    HIPCHK(hipMemcpy2DAsync(c->K, sizeof(double) * ld, K, sizeof(double) * n, sizeof(double) * n,
                            m, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->m = m;
    c->n = ld;
    c->nl = n;
    return UCY_OK;
}

// computeTractionKernelsAntiplane (fortran/2d/antiplane/src/antiplane.f90:352-379):
// one receiver (x, sv, dv, nv), ns segment sources y(3,ns), W, dip
extern "C" int ucy_antiplane_traction_kernels(ucy_ctx *c, const double x[3], const double sv[3],
                                              const double dv[3], const double nv[3], int ns,
                                              const double *y, const double *W,
                                              const double *dip, double G, double *ts)
{
    if (!c || ns < 0) return UCY_ERR_ARG;
    (void)dip;  // the dip of a source does not enter (antiplane.f90:126-127)
    if (ns == 0) return UCY_OK;
    HIPCHK(hipSetDevice(c->device));
    Receiver R;
    for (int k = 0; k < 3; k++) {
        R.xc[k] = x[k];
        R.sv[k] = sv[k];
        R.dv[k] = dv[k];
        R.nv[k] = nv[k];
    }
    std::vector<double> S(3 * (size_t)ns);
    for (int i = 0; i < ns; i++) {
        S[3 * i] = y[3 * i + 1];
        S[3 * i + 1] = y[3 * i + 2];
        S[3 * i + 2] = W[i];
    }
    char *buf = nullptr;
    const size_t bytes = sizeof(Receiver) + sizeof(double) * 4 * (size_t)ns;
    HIPCHK(hipMalloc((void **)&buf, bytes));
    Receiver *d_R = (Receiver *)buf;
    double *d_S = (double *)(buf + sizeof(Receiver)), *d_o = d_S + 3 * ns;
    HIPCHK(hipMemcpyAsync(d_R, &R, sizeof(Receiver), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_S, S.data(), sizeof(double) * 3 * ns, hipMemcpyHostToDevice,
                          c->stream));
    launch_ap_rows(c->stream, d_S, ns, d_R, 1, G, nullptr, 0, d_o);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ts, d_o, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(buf));
    return UCY_OK;
}

// computeTractionKernelsPlaneStrain (fortran/2d/planestrain/src/
// planestrain.f90:266-291): one receiver, ns sources y(3,ns), W, dip
extern "C" int ucy_planestrain_traction_kernels(ucy_ctx *c, const double x[3], const double sv[3],
                                                const double dv[3], const double nv[3], int ns,
                                                const double *y, const double *W,
                                                const double *dip, double G, double lambda,
                                                double *td, double *tn)
{
    if (!c || ns < 0) return UCY_ERR_ARG;
    if (ns == 0) return UCY_OK;
    const double nu = lambda / 2.0 / (lambda + G);
    const int rc = ps_check(c, nu, 1, x + 2, 1, ns, y + 2, 3);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    Receiver R;
    for (int k = 0; k < 3; k++) {
        R.xc[k] = x[k];
        R.sv[k] = sv[k];
        R.dv[k] = dv[k];
        R.nv[k] = nv[k];
    }
    const std::vector<PsSource> S = ps_sources(ns, y, W, dip);
    char *buf = nullptr;
    const size_t bytes = sizeof(Receiver) + sizeof(PsSource) * (size_t)ns + sizeof(double) * 2 * (size_t)ns;
    HIPCHK(hipMalloc((void **)&buf, bytes));
    Receiver *d_R = (Receiver *)buf;
    PsSource *d_S = (PsSource *)(buf + sizeof(Receiver));
    double *d_o = (double *)(buf + sizeof(Receiver) + sizeof(PsSource) * (size_t)ns);
    HIPCHK(hipMemcpyAsync(d_R, &R, sizeof(Receiver), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_S, S.data(), sizeof(PsSource) * ns, hipMemcpyHostToDevice, c->stream));
    launch_ps_rows(c->stream, d_S, ns, d_R, 1, G, lambda, nu, nullptr, 0, d_o, d_o + ns);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(td, d_o, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(tn, d_o + ns, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(buf));
    return UCY_OK;
}

// ---------------------------------------------------------------------------
// observation points: buildO (greens.f90:302-462) and exportPoints
// ---------------------------------------------------------------------------
extern "C" int ucy_build_displacement(ucy_ctx *c, int npts, const double *x)
{
    if (!c || npts < 0 || (npts > 0 && !x)) return UCY_ERR_ARG;
    HIPCHK(hipSetDevice(c->device));
    if (c->graph) {  // a captured attempt may log through the old O
        hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    if (c->O) HIPCHK(hipFree(c->O));
    c->O = nullptr;
    c->o_rows = c->o_n = 0;
    if (npts == 0) return UCY_OK;
    if (!c->have_elastic) return fail(c, UCY_ERR_STATE, "elastic moduli not set");
    if (c->model == 2) {  // fortran/2d/planestrain/src/greens.f90:253-367
        // columns: one dip slip per patch, then e22, e23, e33 per volume
        const int ns = c->nrect, nv = c->nvol, ne = ns + nv;
        const int64_t on = ns + 3 * (int64_t)nv;
        if (ne == 0) return fail(c, UCY_ERR_STATE, "no elements");
        const double nu = c->lambda / 2.0 / (c->lambda + c->mu);
        int rc = ps_check(c, nu, npts, x + 2, 3, ns, c->rx.data() + 2, 3);
        if (!rc) rc = ps_check(c, nu, 0, nullptr, 1, nv, c->vx.data() + 2, 3);
        if (rc) return rc;
        std::vector<PsSource> S = ps_sources(ns, c->rx.data(), c->rW.data(), c->rdip.data());
        const std::vector<PsSource> SV = ps_sources(nv, c->vx.data(), c->vW.data(), c->vdip.data());
        S.insert(S.end(), SV.begin(), SV.end());
        char *buf = nullptr;
        double *O = nullptr;
        const size_t sb = sizeof(PsSource) * ne, tb = sizeof(double) * (nv + 1);
        HIPCHK(hipMalloc((void **)&buf, sb + tb + sizeof(double) * 3 * (size_t)npts));
        HIPCHK(hipMalloc((void **)&O, sizeof(double) * 2 * (size_t)npts * on));
        PsSource *d_S = (PsSource *)buf;
        double *d_T = (double *)(buf + sb);
        double *pts = (double *)(buf + sb + tb);
        HIPCHK(hipMemcpyAsync(d_S, S.data(), sb, hipMemcpyHostToDevice, c->stream));
        if (nv)
            HIPCHK(hipMemcpyAsync(d_T, c->vT.data(), sizeof(double) * nv, hipMemcpyHostToDevice,
                                  c->stream));
        HIPCHK(hipMemcpyAsync(pts, x, sizeof(double) * 3 * npts, hipMemcpyHostToDevice,
                              c->stream));
        launch_ps_disp(c->stream, d_S, ns, pts, npts, c->mu, c->lambda, nu, O, on);
        launch_ps_volume_disp(c->stream, d_S + ns, d_T, nv, pts, npts, c->mu, c->lambda, O, on, ns);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(buf));
        c->O = O;
        c->o_rows = 2 * (int64_t)npts;
        c->o_n = on;
        return UCY_OK;
    }
    if (c->model == 1) {  // fortran/2d/antiplane/src/greens.f90:275-383
        // columns: one slip per segment, then e12, e13 per volume
        const int ns = c->nrect, nv = c->nvol;
        const int64_t on = ns + 2 * (int64_t)nv;
        if (on == 0) return fail(c, UCY_ERR_STATE, "no elements");
        std::vector<double> S(3 * (size_t)ns + 4 * (size_t)nv + 1);
        for (int i = 0; i < ns; i++) {
            S[3 * i] = c->rx[3 * i + 1];
            S[3 * i + 1] = c->rx[3 * i + 2];
            S[3 * i + 2] = c->rW[i];
        }
        for (int i = 0; i < nv; i++) {
            double *q = S.data() + 3 * (size_t)ns + 4 * i;
            q[0] = c->vx[3 * i + 1];
            q[1] = c->vx[3 * i + 2];
            q[2] = c->vW[i];
            q[3] = c->vT[i];
        }
        double *buf = nullptr, *O = nullptr;
        HIPCHK(hipMalloc((void **)&buf, sizeof(double) * (S.size() + 3 * (size_t)npts)));
        HIPCHK(hipMalloc((void **)&O, sizeof(double) * (size_t)npts * on));
        double *pts = buf + S.size();
        HIPCHK(hipMemcpyAsync(buf, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(pts, x, sizeof(double) * 3 * npts, hipMemcpyHostToDevice,
                              c->stream));
        if (ns) launch_ap_disp(c->stream, buf, ns, pts, npts, O, on);
        if (nv) launch_ap_volume_disp(c->stream, buf + 3 * (size_t)ns, nv, pts, npts, O + ns, on);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(buf));
        c->O = O;
        c->o_rows = npts;
        c->o_n = on;
        return UCY_OK;
    }
    const int np = c->nrect + c->ntri;
    if (np + c->nvol == 0) return fail(c, UCY_ERR_STATE, "no elements");
    const int64_t rows = 3 * (int64_t)npts, n = 2 * (int64_t)np + 6 * (int64_t)c->nvol;
    double *d_pts = nullptr, *O = nullptr;
    HIPCHK(hipMalloc((void **)&d_pts, sizeof(double) * 3 * (size_t)npts));
    HIPCHK(hipMalloc((void **)&O, sizeof(double) * (size_t)(rows * n)));
    HIPCHK(hipMemcpyAsync(d_pts, x, sizeof(double) * 3 * (size_t)npts, hipMemcpyHostToDevice,
                          c->stream));
    int rc = 0;
    if (c->nrect) {
        std::vector<OkadaSource> S(c->nrect);
        for (int i = 0; i < c->nrect; i++)
            S[i] = make_source(&c->rx[3 * i], c->rL[i], c->rW[i], c->rstrike[i], c->rdip[i]);
        OkadaSource *d_S = nullptr;
        HIPCHK(hipMalloc((void **)&d_S, sizeof(OkadaSource) * c->nrect));
        HIPCHK(hipMemcpyAsync(d_S, S.data(), sizeof(OkadaSource) * c->nrect,
                              hipMemcpyHostToDevice, c->stream));
        launch_okada_disp(c->stream, make_medium(c->mu, c->lambda), d_S, c->nrect, d_pts, npts, O,
                          n, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(d_S));
    }
    if (c->ntri)
        rc = ext_build_triangle_disp(c->stream, c->mu, c->lambda, c->ntri, c->i1.data(),
                                     c->i2.data(), c->i3.data(), c->vert.data(), d_pts, npts, O, n,
                                     2 * (int64_t)c->nrect, c->err);
    if (!rc && c->nvol)
        rc = ext_build_volume_disp(c->stream, c->mu, c->lambda, c->nvol, c->vx.data(),
                                   c->vL.data(), c->vT.data(), c->vW.data(), c->vstrike.data(), x,
                                   d_pts, npts, O, n, 2 * (int64_t)np, c->err);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(d_pts));
    if (rc) {
        HIPCHK(hipFree(O));
        return rc;
    }
    c->O = O;
    c->o_rows = rows;
    c->o_n = n;
    return UCY_OK;
}

extern "C" int ucy_displacement_shape(const ucy_ctx *c, int64_t *rows, int64_t *n)
{
    if (!c) return UCY_ERR_ARG;
    if (rows) *rows = c->O ? c->o_rows : 0;
    if (n) *n = c->O ? c->o_n : 0;
    return UCY_OK;
}

// O built, and for the current mesh
static int check_disp(ucy_ctx *c)
{
    if (!c->O) return fail(c, UCY_ERR_STATE, "displacement kernels not built");
    const int64_t want = c->model == 2 ? (int64_t)c->nrect + 3 * (int64_t)c->nvol
                       : c->model == 1 ? (int64_t)c->nrect + 2 * (int64_t)c->nvol
                                       : 2 * (int64_t)(c->nrect + c->ntri) + 6 * (int64_t)c->nvol;
    if (c->o_n != want)
        return fail(c, UCY_ERR_STATE, "displacement kernels built for another mesh");
    return 0;
}

extern "C" int ucy_get_displacement_kernel(ucy_ctx *c, double *O, double *Of, double *Ol)
{
    if (!c) return UCY_ERR_ARG;
    int rc = check_disp(c);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    const size_t cnt = (size_t)(c->o_rows * c->o_n);
    std::vector<double> tmp;
    double *h = O;
    if (!h) {
        tmp.resize(cnt);
        h = tmp.data();
    }
    HIPCHK(hipMemcpyAsync(h, c->O, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const bool two_d = c->model != 0;
    const int64_t nf = two_d ? c->nrect : 2 * (int64_t)(c->nrect + c->ntri);
    for (int64_t r = 0; r < c->o_rows; r++) {
        const bool vertical = !two_d && r % 3 == 2;
        for (int64_t j = 0; j < c->o_n; j++) {
            const double v = h[r * c->o_n + j];
            const bool f = j < nf || vertical;
            if (Of) Of[r * c->o_n + j] = f ? v : 0.0;
            if (Ol) Ol[r * c->o_n + j] = f ? 0.0 : v;
        }
    }
    return UCY_OK;
}

extern "C" int ucy_displacement_dev(ucy_ctx *c, const double *y, double *d, double *df,
                                    double *dl)
{
    if (!c || !y || !d) return UCY_ERR_ARG;
    int rc = check_disp(c);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    launch_disp_gemv(c->stream, c->O, c->o_rows, c->o_n, c->nrect + c->ntri, c->model, y, d,
                     df, dl, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return UCY_OK;
}

extern "C" int ucy_displacement(ucy_ctx *c, const double *y, double *d, double *df, double *dl)
{
    if (!c || !y || !d) return UCY_ERR_ARG;
    int rc = check_disp(c);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    const int64_t dim = c->model == 2 ? 5 * (int64_t)c->nrect + 6 * (int64_t)c->nvol
                      : c->model == 1 ? 4 * (int64_t)(c->nrect + c->nvol)
                                      : 7 * (int64_t)(c->nrect + c->ntri) + 12 * (int64_t)c->nvol;
    const int64_t R = c->o_rows;
    double *buf = nullptr;
    HIPCHK(hipMalloc((void **)&buf, sizeof(double) * (size_t)(dim + 3 * R)));
    double *dy = buf, *dd = buf + dim, *ddf = dd + R, *ddl = ddf + R;
    HIPCHK(hipMemcpyAsync(dy, y, sizeof(double) * dim, hipMemcpyHostToDevice, c->stream));
    launch_disp_gemv(c->stream, c->O, R, c->o_n, c->nrect + c->ntri, c->model, dy, dd,
                     df ? ddf : nullptr, dl ? ddl : nullptr, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(d, dd, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
    if (df) HIPCHK(hipMemcpyAsync(df, ddf, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
    if (dl) HIPCHK(hipMemcpyAsync(dl, ddl, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(buf));
    return UCY_OK;
}

// ---------------------------------------------------------------------------
// matvec / RHS / stepper
// ---------------------------------------------------------------------------
static int upload_params(ucy_ctx *c)
{
    if (!c->params_dirty && c->d_fr) return 0;
    const int np = c->nrect + c->ntri;
    if (np && (!c->have_friction || (int)c->Vpl.size() != np))
        return fail(c, UCY_ERR_STATE, "friction properties not set for all %d patches", np);
    if (c->nvol && (!c->have_rheology || (int)c->gd0.size() != c->nvol))
        return fail(c, UCY_ERR_STATE, "rheology not set for all %d volumes", c->nvol);
    std::vector<Friction> F(np + 1);
    for (int i = 0; i < np; i++) {
        F[i].a = c->a[i];
        F[i].b = c->b[i];
        F[i].sig = c->sig[i];
        F[i].L = c->Lc[i];
        F[i].damping = c->damping[i];
        // 2-D: v = V - Vpl (ratestate.f90 of 2d/antiplane:676, 2d/planestrain:710);
        // 3-D: ratestate.f90:701-702
        F[i].vplc = c->model ? c->Vpl[i] : c->Vpl[i] * cos(c->rake[i]);
        F[i].vpls = c->model ? 0.0 : c->Vpl[i] * sin(c->rake[i]);
    }
    std::vector<Rheology> H(c->nvol + 1);
    for (int i = 0; i < c->nvol; i++) {
        H[i].gammadot0 = c->gd0[i];
        H[i].npower = c->npow[i];
        for (int k = 0; k < 6; k++) H[i].ebg[k] = c->ebg[6 * i + k];
        H[i].Gm = c->Gm[i];
        H[i].arr = exp(-c->Qa[i] / (c->Rg[i] * c->To[i]));
    }
    // the device arrays are reused while they are large enough; a captured
    // attempt holds their addresses, so a reallocation drops it
    if ((c->fr_cap < np + 1 || c->rh_cap < c->nvol + 1) && c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    if (c->fr_cap < np + 1) {
        if (c->d_fr) HIPCHK(hipFree(c->d_fr));
        c->d_fr = nullptr;
        c->fr_cap = 0;
        HIPCHK(hipMalloc((void **)&c->d_fr, sizeof(Friction) * (np + 1)));
        c->fr_cap = np + 1;
    }
    if (c->rh_cap < c->nvol + 1) {
        if (c->d_rh) HIPCHK(hipFree(c->d_rh));
        c->d_rh = nullptr;
        c->rh_cap = 0;
        HIPCHK(hipMalloc((void **)&c->d_rh, sizeof(Rheology) * (c->nvol + 1)));
        c->rh_cap = c->nvol + 1;
    }
    HIPCHK(hipMemcpyAsync(c->d_fr, F.data(), sizeof(Friction) * (np + 1), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(c->d_rh, H.data(), sizeof(Rheology) * (c->nvol + 1),
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->params_dirty = false;
    return 0;
}

static PlainDims dims(const ucy_ctx *c)
{
    PlainDims d;
    d.np = c->nrect + c->ntri;
    d.nv = c->nvol;
    d.ap = c->model == 1;
    d.ps = c->model == 2;
    d.dim = d.ps ? 5 * d.np + 6 * d.nv : d.ap ? 4 * (d.np + d.nv) : 7 * d.np + 12 * d.nv;
    d.m = (int)c->m;
    d.n = (int)c->n;
    return d;
}

// work arrays: y dydt yscal ytmp ak2..ak6 ycand yerr v t (+ logs separately)
static int ensure_work(ucy_ctx *c)
{
    const PlainDims d = dims(c);
    if (c->work && c->work_dim == (size_t)d.dim) return 0;
    free_work(c);
    const size_t nvec = 11;
    const size_t bytes =
        sizeof(double) * (nvec * (size_t)d.dim + 3 * (size_t)d.n + 2 * (size_t)d.m + 64);
    HIPCHK(hipMalloc((void **)&c->work, bytes));
    HIPCHK(hipMemsetAsync(c->work, 0, bytes, c->stream));
    double *p = c->work;
    RkBuffers &b = c->rb;
    b.y = p; p += d.dim;
    b.dydt = p; p += d.dim;
    b.yscal = p; p += d.dim;
    b.ytmp = p; p += d.dim;
    b.ak[0] = b.ak[1] = b.dydt;
    for (int k = 2; k <= 6; k++) { b.ak[k] = p; p += d.dim; }
    b.ycand = p; p += d.dim;
    b.yerr = p; p += d.dim;
    b.yr = nullptr;
    b.v = p; p += d.n;
    b.t = p; p += d.m;
    b.v0 = p; p += d.n;
    b.t0 = p; p += d.m;
    b.v2 = p; p += d.n;
    if (c->kprec == UCY_PREC_F64) b.v0 = b.t0 = nullptr;
    b.tlog = b.dtlog = b.vlog = b.elog = b.rlog = b.emlog = nullptr;
    c->work_dim = d.dim;
    return 0;
}

static int check_ready(ucy_ctx *c, int n)
{
    if (!c) return UCY_ERR_ARG;
    if (!c->K) return fail(c, UCY_ERR_STATE, "kernel not built");
    const PlainDims d = dims(c);
    if (n != d.dim) return fail(c, UCY_ERR_ARG, "state length %d, expected %d", n, d.dim);
    HIPCHK(hipSetDevice(c->device));
    int rc = upload_params(c);
    if (rc) return rc;
    return ensure_work(c);
}

extern "C" int ucy_matvec_reduced(ucy_ctx *c, const double *v, double *t)
{
    if (!c || !v || !t) return UCY_ERR_ARG;
    if (!c->Kp) return fail(c, UCY_ERR_STATE, "kernel not built in reduced precision");
    HIPCHK(hipSetDevice(c->device));
    double *dv = nullptr, *dt = nullptr;
    HIPCHK(hipMalloc((void **)&dv, sizeof(double) * c->n));
    HIPCHK(hipMalloc((void **)&dt, sizeof(double) * c->m));
    if (c->nl != c->n)
        HIPCHK(hipMemsetAsync(dv + c->nl, 0, sizeof(double) * (c->n - c->nl), c->stream));
    HIPCHK(hipMemcpyAsync(dv, v, sizeof(double) * c->nl, hipMemcpyHostToDevice, c->stream));
    launch_matvec_split(c->stream, c->Kp, c->Krs, (int)c->m, (int)c->n, dv, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(t, dt, sizeof(double) * c->m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(dv));
    HIPCHK(hipFree(dt));
    return UCY_OK;
}

extern "C" int ucy_matvec_f64_dev(ucy_ctx *c, const double *v_dev, double *t_dev)
{
    if (!c || !v_dev || !t_dev) return UCY_ERR_ARG;
    if (!c->K) return fail(c, UCY_ERR_STATE, "kernel not built");
    if (c->n % 2) return fail(c, UCY_ERR_STATE, "odd column count");
    HIPCHK(hipSetDevice(c->device));
    if (c->nl != c->n) {  // the caller's v has nl entries: pad it with zeros
        if (ensure_buf(c, c->hook, sizeof(double) * (size_t)c->n)) return UCY_ERR_HIP;
        double *pv = (double *)c->hook.p;
        if (pv != v_dev) {
            HIPCHK(hipMemsetAsync(pv + c->nl, 0, sizeof(double) * (c->n - c->nl), c->stream));
            HIPCHK(hipMemcpyAsync(pv, v_dev, sizeof(double) * c->nl, hipMemcpyDeviceToDevice,
                                  c->stream));
        }
        v_dev = pv;
    }
    launch_matvec(c->stream, c->K, (int)c->m, (int)c->n, v_dev, t_dev);
    HIPCHK(hipGetLastError());
    return UCY_OK;
}

extern "C" int ucy_matvec_f64(ucy_ctx *c, const double *v, double *t)
{
    if (!c || !v || !t) return UCY_ERR_ARG;
    if (!c->K) return fail(c, UCY_ERR_STATE, "kernel not built");
    HIPCHK(hipSetDevice(c->device));
    if (ensure_buf(c, c->hook, sizeof(double) * (size_t)(c->m + c->n))) return UCY_ERR_HIP;
    double *dv = (double *)c->hook.p, *dt = dv + c->n;
    if (c->nl != c->n)
        HIPCHK(hipMemsetAsync(dv + c->nl, 0, sizeof(double) * (c->n - c->nl), c->stream));
    HIPCHK(hipMemcpyAsync(dv, v, sizeof(double) * c->nl, hipMemcpyHostToDevice, c->stream));
    int rc = ucy_matvec_f64_dev(c, dv, dt);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(t, dt, sizeof(double) * c->m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return UCY_OK;
}

// initStateVector, ratestate.f90:548-610 (variant 0) / viscouscycles.f90:710-820 (1)
extern "C" int ucy_init_state(ucy_ctx *c, int variant, double *y)
{
    if (!c || !y) return UCY_ERR_ARG;
    const int np = c->nrect + c->ntri;
    if (np && (int)c->Vpl.size() != np) return fail(c, UCY_ERR_STATE, "friction not set");
    if (c->nvol && (int)c->gd0.size() != c->nvol) return fail(c, UCY_ERR_STATE, "rheology not set");
    const double lg10 = 2.302585092994045684;
    if (c->model == 2) {  // ratestate.f90 (2-D plane strain):585-627
        for (int i = 0; i < np; i++) {
            double *p = y + 5 * i;
            const bool t0 = i < (int)c->tau0.size() && 0 < c->tau0[i];
            p[0] = 0.0;
            p[1] = t0 ? c->tau0[i]
                      : c->sig[i] * (c->mu0[i] + (c->a[i] - c->b[i]) * log(c->Vpl[i] / c->Vo[i]))
                            - c->damping[i] * c->Vpl[i];
            p[2] = 0.0;
            p[3] = log(c->Lc[i] / c->Vpl[i]) / lg10;
            p[4] = log(c->Vpl[i] * 0.98) / lg10;
        }
        // volumes, viscouscycles.f90 (2-D plane strain):777-821: s0 where
        // not all zero, else the stress that flows at the background rate
        // (e22, e23, e33 in ebg[3..5]) under the power law with Arrhenius
        for (int i = 0; i < c->nvol; i++) {
            double *q = y + 5 * (size_t)np + 6 * (size_t)i;
            const double e22 = c->ebg[6 * i + 3], e23 = c->ebg[6 * i + 4], e33 = c->ebg[6 * i + 5];
            const bool given = 3 * (size_t)(i + 1) <= c->s0.size() &&
                               !(0.0 == c->s0[3 * i] && 0.0 == c->s0[3 * i + 1] &&
                                 0.0 == c->s0[3 * i + 2]);
            double s22 = 0.0, s23 = 0.0, s33 = 0.0;
            if (given) {
                s22 = c->s0[3 * i];
                s23 = c->s0[3 * i + 1];
                s33 = c->s0[3 * i + 2];
            } else {
                const double ekk = e22 + e33;
                const double e22p = e22 - ekk / 2, e33p = e33 - ekk / 2;
                const double eps = sqrt((e22p * e22p + 2.0 * (e23 * e23) + e33p * e33p) / 2.0);
                if (0.0 < eps && 0.0 < c->gd0[i]) {
                    const double tau =
                        pow(eps / c->gd0[i] * exp(c->Qa[i] / (c->Rg[i] * c->To[i])), 1.0 / c->npow[i]);
                    s22 = tau * e22p / eps;
                    s23 = tau * e23 / eps;
                    s33 = tau * e33p / eps;
                }
            }
            q[0] = q[1] = q[2] = 0.0;
            q[3] = s22;
            q[4] = s23;
            q[5] = s33;
        }
        return UCY_OK;
    }
    if (c->model == 1) {  // ratestate.f90 (2-D antiplane):568-605
        for (int i = 0; i < np; i++) {
            double *p = y + 4 * i;
            p[0] = 0.0;
            p[1] = c->sig[i] * (c->mu0[i] + (c->a[i] - c->b[i]) * log(c->Vpl[i] / c->Vo[i]))
                   - c->damping[i] * c->Vpl[i];
            p[2] = log(c->Lc[i] / c->Vpl[i]) / lg10;
            p[3] = log(c->Vpl[i] * 0.98) / lg10;
        }
        // volumes, viscouscycles.f90 (2-D antiplane):740-762: the stress
        // that flows at the background rate (e12, e13 in ebg[1], ebg[2])
        for (int i = 0; i < c->nvol; i++) {
            double *q = y + 4 * (np + i);
            const double e12 = c->ebg[6 * i + 1], e13 = c->ebg[6 * i + 2];
            const double eps = sqrt(e12 * e12 + e13 * e13);
            if (0.0 < eps && 0.0 < c->gd0[i]) {
                const double tau = c->Gm[i] * pow(eps / c->gd0[i] * exp(c->Qa[i] / (c->Rg[i] * c->To[i])),
                                                  1.0 / c->npow[i]);
                q[2] = tau * e12 / eps;
                q[3] = tau * e13 / eps;
            } else {
                q[2] = q[3] = 0.0;
            }
            q[0] = q[1] = 0.0;
        }
        return UCY_OK;
    }
    for (int i = 0; i < np; i++) {
        double *p = y + 7 * i;
        const double base = c->sig[i] * (c->mu0[i] + (c->a[i] - c->b[i]) * log(c->Vpl[i] / c->Vo[i]));
        const double tau = variant ? base + c->damping[i] * c->Vpl[i] : base - c->damping[i] * c->Vpl[i];
        p[0] = 0.0;
        p[1] = 0.0;
        p[2] = tau * cos(c->rake[i]);
        p[3] = tau * sin(c->rake[i]);
        p[4] = 0.0;
        p[5] = log(c->Lc[i] / c->Vpl[i]) / lg10;
        p[6] = log(c->Vpl[i] * (variant ? 0.98 : 0.99)) / lg10;
    }
    for (int i = 0; i < c->nvol; i++) {
        double *q = y + 7 * np + 12 * i;
        const double *e = &c->ebg[6 * i];
        const double eps = sqrt((e[0] * e[0] + 2 * (e[1] * e[1]) + 2 * (e[2] * e[2]) + e[3] * e[3]
                                 + 2 * (e[4] * e[4]) + e[5] * e[5]) / 2.0);
        if (0.0 < eps) {
            const double tau = c->Gm[i] * pow(eps / c->gd0[i] * exp(c->Qa[i] / (c->Rg[i] * c->To[i])),
                                              1.0 / c->npow[i]);
            for (int k = 0; k < 6; k++) q[6 + k] = tau * e[k] / eps;
        } else {
            for (int k = 0; k < 6; k++) q[6 + k] = 0.0;
        }
        for (int k = 0; k < 6; k++) q[k] = 0.0;
    }
    return UCY_OK;
}

extern "C" int ucy_odefun(ucy_ctx *c, int n, double time, const double *y, double *dydt,
                          double *vmax)
{
    (void)time;  // odefun does not depend on time (ratestate.f90:624-778)
    int rc = check_ready(c, n);
    if (rc) return rc;
    const PlainDims d = dims(c);
    RkBuffers &b = c->rb;
    double *d_vmax = b.t + d.m;  // spare slot after t
    HIPCHK(hipMemcpyAsync(b.ytmp, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(d_vmax, 0, sizeof(double), c->stream));
    launch_prepass(c->stream, d, c->d_fr, c->d_rh, b.ytmp, b.ak[2], b.v, d_vmax);
    launch_gemv_post(c->stream, d, c->K, c->d_fr, b.ytmp, b.v, b.ak[2], b.t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dydt, b.ak[2], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (vmax) HIPCHK(hipMemcpyAsync(vmax, d_vmax, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return UCY_OK;
}

// clears the vMax accumulator and the padding column of v (a kernel, not
// hipMemsetAsync, so that the call can be captured into a graph)
__global__ void static_prep_kernel(double *vmax, double *vpad, int npad)
{
    if (threadIdx.x == 0) *vmax = 0.0;
    if ((int)threadIdx.x < npad) vpad[threadIdx.x] = 0.0;
}

extern "C" int ucy_odefun_dev_static(ucy_ctx *c, int n, const double *y_dev, double *dydt_dev,
                                     double *vmax_dev, void *stream)
{
    if (!c || !y_dev || !dydt_dev) return UCY_ERR_ARG;
    if (!c->K) return fail(c, UCY_ERR_STATE, "kernel not built");
    const PlainDims d = dims(c);
    if (n != d.dim) return fail(c, UCY_ERR_ARG, "state length %d, expected %d", n, d.dim);
    if (c->params_dirty || !c->d_fr || !c->d_rh)
        return fail(c, UCY_ERR_STATE,
                    "friction/rheology not on the device: set them for every element first");
    if (c->n > kStaticMaxCols)
        return fail(c, UCY_ERR_LIMIT, "%lld kernel columns exceed the static scratch (%lld)",
                    (long long)c->n, (long long)kStaticMaxCols);
    double *base = g_scratch[c->device].load();
    if (!base) return fail(c, UCY_ERR_STATE, "no static scratch on device %d", c->device);
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    double *vm = vmax_dev ? vmax_dev : base;
    double *v = base + kStaticLead;
    const int npad = (int)(c->n - c->nl);  // 0, or the antiplane model's padding column
    hipLaunchKernelGGL(static_prep_kernel, dim3(1), dim3(64), 0, st, vm, v + c->nl, npad);
    launch_prepass(st, d, c->d_fr, c->d_rh, y_dev, dydt_dev, v, vm);
    launch_gemv_post(st, d, c->K, c->d_fr, y_dev, v, dydt_dev, nullptr);
    HIPCHK(hipGetLastError());
    if (!stream) HIPCHK(hipStreamSynchronize(st));
    return UCY_OK;
}

// one Cash-Karp attempt on device buffers (y, dydt already resident)
static void rkck_device(ucy_ctx *c, const PlainDims &d, double h)
{
    RkBuffers &b = c->rb;
    for (int s = 2; s <= 6; s++) {
        launch_rk_combine(c->stream, d.dim, s, h, b.y, b.ak, nullptr, b.ytmp, nullptr, nullptr);
        launch_prepass(c->stream, d, c->d_fr, c->d_rh, b.ytmp, b.ak[s], b.v, nullptr);
        launch_gemv_post(c->stream, d, c->K, c->d_fr, b.ytmp, b.v, b.ak[s], nullptr);
    }
    launch_rk_combine(c->stream, d.dim, 7, h, b.y, b.ak, nullptr, nullptr, b.ycand, b.yerr);
}

extern "C" int ucy_rkck(ucy_ctx *c, int n, double t, const double *yin, const double *dydt,
                        double h, double *yout, double *yerr)
{
    (void)t;
    int rc = check_ready(c, n);
    if (rc) return rc;
    const PlainDims d = dims(c);
    RkBuffers &b = c->rb;
    HIPCHK(hipMemcpyAsync(b.y, yin, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.dydt, dydt, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    rkck_device(c, d, h);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(yout, b.ycand, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(yerr, b.yerr, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return UCY_OK;
}

// RKQSm (ode45.f90:51-108) with the Cash-Karp stages on the device and the
// scalar step control on the host (the integrator below keeps it on device).
extern "C" int ucy_rkqs(ucy_ctx *c, int n, double *t, double *y, double *dydt,
                        const double *yscal, double htry, double *hdid, double *hnext,
                        double epsilon, double max_step)
{
    int rc = check_ready(c, n);
    if (rc) return rc;
    const PlainDims d = dims(c);
    RkBuffers &b = c->rb;
    HIPCHK(hipMemcpyAsync(b.y, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.dydt, dydt, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    std::vector<double> yerr(n), ycand(n);
    const double SAFETY = 0.7, PGROW = -0.2, PSHRNK = -0.15, ERRCON = 1.89e-4;
    double h = htry;
    for (;;) {
        rkck_device(c, d, h);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(yerr.data(), b.yerr, sizeof(double) * n, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double errmax = 0.0;
        for (int i = 0; i < n; i++) {
            const double e = fabs(yerr[i] / yscal[i]);
            if (e > errmax) errmax = e;
        }
        errmax = errmax / epsilon;
        if (errmax > 1.0) {
            h = h * SAFETY * pow(errmax, PSHRNK);
            if (*t + h == *t) return fail(c, UCY_ERR_UNDERFLOW, "stepsize underflow in RKQS");
        } else {
            *hnext = errmax > ERRCON ? SAFETY * h * pow(errmax, PGROW) : 10.0 * h;
            if (*hnext > max_step) *hnext = max_step;
            *hdid = h;
            *t = *t + h;
            HIPCHK(hipMemcpyAsync(y, b.ycand, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            return UCY_OK;
        }
    }
}

// ---------------------------------------------------------------------------
// device-resident main loop
// ---------------------------------------------------------------------------
static int capture_attempt(ucy_ctx *c)
{
    if (c->graph) return 0;
    const PlainDims d = dims(c);
    RkBuffers &b = c->rb;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const KOperand K{c->K, c->Kp, c->Krs, c->kprec};
    launch_start(c->stream, d, c->d_fr, c->d_rh, b, c->d_ctl);
    for (int s = 1; s <= 5; s++) launch_rhs_stage(c->stream, d, K, c->d_fr, c->d_rh, b, c->d_ctl, s);
    launch_rhs_finish(c->stream, d, K, c->d_fr, b, c->d_ctl);
    launch_controller(c->stream, b, c->d_ctl);
    if (b.n_obs > 0) launch_obs_log(c->stream, b, c->d_ctl);
    if (b.disp)
        launch_disp_gemv(c->stream, b.O, b.o_rows, b.o_n, d.np, c->model, b.y, b.disp, b.disp_f,
                         b.disp_l, c->d_ctl);
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return fail(c, UCY_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int ucy_integrate(ucy_ctx *c, double *y, double interval, double dt0,
                             int64_t max_steps, double epsilon, double max_step, double *t_log,
                             double *dt_log, double *vmax_log, int64_t log_cap,
                             ucy_run_stats *stats)
{
    ucy_step_log log = {t_log, dt_log, vmax_log, nullptr, nullptr, log_cap, 0, nullptr, nullptr,
                        nullptr, nullptr, nullptr, nullptr, nullptr};
    return ucy_integrate_ex(c, y, 0.0, interval, dt0, max_steps, epsilon, max_step, &log, stats);
}

extern "C" int ucy_integrate_ex(ucy_ctx *c, double *y, double t0, double interval, double dt0,
                                int64_t max_steps, double epsilon, double max_step,
                                const ucy_step_log *log, ucy_run_stats *stats)
{
    double *t_log = log ? log->t : nullptr, *dt_log = log ? log->dt : nullptr;
    double *vmax_log = log ? log->vmax : nullptr;
    double *err_log = log ? log->errmax : nullptr, *rej_log = log ? log->rejects : nullptr;
    const int64_t log_cap = log ? log->cap : 0;
    const int64_t n_obs = log ? log->n_obs : 0;
    if (!c || !y || n_obs < 0) return UCY_ERR_ARG;
    const PlainDims d0 = c->K ? dims(c) : PlainDims{};
    int rc = check_ready(c, d0.dim);
    if (rc) return rc;
    if (!(interval > 0) || !(dt0 > 0) || max_steps <= 0 || !(epsilon > 0) || !(t0 == t0))
        return fail(c, UCY_ERR_ARG, "invalid integration parameters");
    const PlainDims d = dims(c);
    RkBuffers &b = c->rb;
    // device log buffers (reallocated per call; logs are parity aids)
    double *d_log = nullptr;
    const int64_t cap = (t_log && dt_log && vmax_log) ? log_cap : 0;
    const bool ext = cap > 0 && err_log && rej_log;
    double *em_log = log ? log->emax : nullptr;
    const bool with_em = cap > 0 && em_log;
    if (cap > 0)
        HIPCHK(hipMalloc((void **)&d_log, sizeof(double) * (6 * cap)));
    if (c->graph && (b.tlog != d_log)) {
        hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    b.tlog = d_log;
    b.dtlog = d_log ? d_log + cap : nullptr;
    b.vlog = d_log ? d_log + 2 * cap : nullptr;
    b.elog = ext ? d_log + 3 * cap : nullptr;
    b.rlog = ext ? d_log + 4 * cap : nullptr;
    b.emlog = with_em ? d_log + 5 * cap : nullptr;
    // observed entries (needs the step log for the row index)
    int64_t *d_oidx = nullptr;
    double *d_obs = nullptr;
    b.obs_idx = nullptr;
    b.n_obs = 0;
    b.obs_y = b.obs_dydt = nullptr;
    if (n_obs > 0) {
        if (cap <= 0 || !log->obs_index || !log->obs_y || !log->obs_dydt) {
            if (d_log) hipFree(d_log);
            return fail(c, UCY_ERR_ARG, "observed entries need the step log and their buffers");
        }
        for (int64_t j = 0; j < n_obs; j++)
            if (log->obs_index[j] < 0 || log->obs_index[j] >= d.dim) {
                if (d_log) hipFree(d_log);
                return fail(c, UCY_ERR_ARG, "observed index %lld outside the state",
                            (long long)log->obs_index[j]);
            }
        HIPCHK(hipMalloc((void **)&d_oidx, sizeof(int64_t) * n_obs));
        HIPCHK(hipMalloc((void **)&d_obs, sizeof(double) * 2 * n_obs * cap));
        HIPCHK(hipMemcpyAsync(d_oidx, log->obs_index, sizeof(int64_t) * n_obs,
                              hipMemcpyHostToDevice, c->stream));
        b.obs_idx = d_oidx;
        b.n_obs = n_obs;
        b.obs_y = d_obs;
        b.obs_dydt = d_obs + n_obs * cap;
    }
    // observation-point displacement (needs the step log and O)
    const bool want_disp = log && (log->disp || log->disp_f || log->disp_l);
    double *d_disp = nullptr;
    b.O = nullptr;
    b.o_rows = b.o_n = 0;
    b.disp = b.disp_f = b.disp_l = nullptr;
    if (want_disp) {
        rc = cap > 0 ? check_disp(c)
                     : fail(c, UCY_ERR_ARG, "displacement logging needs the step log");
        if (rc) {
            if (d_log) hipFree(d_log);
            if (d_oidx) {
                hipFree(d_oidx);
                hipFree(d_obs);
            }
            b.tlog = b.dtlog = b.vlog = b.elog = b.rlog = b.emlog = nullptr;
            b.obs_idx = nullptr;
            b.n_obs = 0;
            b.obs_y = b.obs_dydt = nullptr;
            return rc;
        }
        const int64_t R = c->o_rows;
        HIPCHK(hipMalloc((void **)&d_disp, sizeof(double) * 3 * R * cap));
        b.O = c->O;
        b.o_rows = R;
        b.o_n = c->o_n;
        b.disp = d_disp;
        b.disp_f = log->disp_f ? d_disp + R * cap : nullptr;
        b.disp_l = log->disp_l ? d_disp + 2 * R * cap : nullptr;
    }
    Control &h = *c->h_ctl;
    memset(&h, 0, sizeof(Control));
    h.time = t0;
    h.h = dt0;
    h.hnext = dt0;
    h.interval = interval;
    h.epsilon = epsilon;
    h.max_step = max_step;
    h.need_rhs = 1;
    h.max_steps = max_steps;
    h.log_cap = cap;
    HIPCHK(hipMemcpyAsync(c->d_ctl, &h, sizeof(Control), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.y, y, sizeof(double) * d.dim, hipMemcpyHostToDevice, c->stream));
    rc = capture_attempt(c);
    if (rc) {
        if (d_log) hipFree(d_log);
        if (d_oidx) {
            hipFree(d_oidx);
            hipFree(d_obs);
        }
        if (d_disp) hipFree(d_disp);
        b.tlog = b.dtlog = b.vlog = b.elog = b.rlog = b.emlog = nullptr;
        b.obs_idx = nullptr;
        b.n_obs = 0;
        b.obs_y = b.obs_dydt = nullptr;
        b.O = nullptr;
        b.disp = b.disp_f = b.disp_l = nullptr;
        return rc;
    }
    auto w0 = std::chrono::steady_clock::now();
    int batch = 16;
    for (;;) {
        for (int i = 0; i < batch; i++) HIPCHK(hipGraphLaunch(c->graph, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_ctl, c->d_ctl, sizeof(Control), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (h.done) break;
        if (batch < 256) batch *= 2;
    }
    // the last accepted step (later attempts commit at their start)
    launch_commit(c->stream, d.dim, b, c->d_ctl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    auto t1 = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(y, b.y, sizeof(double) * d.dim, hipMemcpyDeviceToHost, c->stream));
    if (cap > 0 && h.log_count > 0) {
        HIPCHK(hipMemcpyAsync(t_log, b.tlog, sizeof(double) * h.log_count, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipMemcpyAsync(dt_log, b.dtlog, sizeof(double) * h.log_count,
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(vmax_log, b.vlog, sizeof(double) * h.log_count,
                              hipMemcpyDeviceToHost, c->stream));
        if (ext) {
            HIPCHK(hipMemcpyAsync(err_log, b.elog, sizeof(double) * h.log_count,
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(rej_log, b.rlog, sizeof(double) * h.log_count,
                                  hipMemcpyDeviceToHost, c->stream));
        }
        if (with_em)
            HIPCHK(hipMemcpyAsync(em_log, b.emlog, sizeof(double) * h.log_count,
                                  hipMemcpyDeviceToHost, c->stream));
        if (n_obs > 0) {
            HIPCHK(hipMemcpyAsync(log->obs_y, b.obs_y, sizeof(double) * n_obs * h.log_count,
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(log->obs_dydt, b.obs_dydt, sizeof(double) * n_obs * h.log_count,
                                  hipMemcpyDeviceToHost, c->stream));
        }
        const size_t dbytes = sizeof(double) * (size_t)(b.o_rows * h.log_count);
        if (want_disp && log->disp)
            HIPCHK(hipMemcpyAsync(log->disp, b.disp, dbytes, hipMemcpyDeviceToHost, c->stream));
        if (want_disp && log->disp_f)
            HIPCHK(hipMemcpyAsync(log->disp_f, b.disp_f, dbytes, hipMemcpyDeviceToHost, c->stream));
        if (want_disp && log->disp_l)
            HIPCHK(hipMemcpyAsync(log->disp_l, b.disp_l, dbytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (d_log) {
        // the graph refers to the log buffers; drop it with them
        hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
        HIPCHK(hipFree(d_log));
        b.tlog = b.dtlog = b.vlog = b.elog = b.rlog = b.emlog = nullptr;
    }
    if (d_oidx) {
        HIPCHK(hipFree(d_oidx));
        HIPCHK(hipFree(d_obs));
        b.obs_idx = nullptr;
        b.n_obs = 0;
        b.obs_y = b.obs_dydt = nullptr;
    }
    if (d_disp) {
        HIPCHK(hipFree(d_disp));
        b.O = nullptr;
        b.disp = b.disp_f = b.disp_l = nullptr;
    }
    if (stats) {
        stats->time = h.time;
        stats->dt_next = h.h;
        stats->steps = h.steps;
        stats->attempts = h.attempts;
        stats->rhs_evals = h.rhs_evals;
        stats->vmax = h.last_vmax;
        stats->wall_seconds = std::chrono::duration<double>(t1 - w0).count();
    }
    if (h.error == 1) return fail(c, UCY_ERR_UNDERFLOW, "stepsize underflow in RKQS");
    if (h.error == 2) return fail(c, UCY_ERR_NONFINITE, "non-finite error norm");
    return UCY_OK;
}

// ---------------------------------------------------------------------------
// per-kernel parity hooks (one receiver, ns sources)
// ---------------------------------------------------------------------------
static int okada_hook(ucy_ctx *c, int kind, const double x[3], const double sv[3],
                      const double dv[3], const double nv[3], int ns, const double *y,
                      const double *L, const double *W, const double *strike, const double *dip,
                      double s, double dd, double t, double G, double lambda, double **out)
{
    if (!c || ns < 0) return UCY_ERR_ARG;
    if (ns == 0) return UCY_OK;
    if (t != 0.0) return fail(c, UCY_ERR_ARG, "tensile opening is not part of the slip kernels");
    HIPCHK(hipSetDevice(c->device));
    const int NC = kind == 0 ? 3 : 6;
    std::vector<OkadaSource> S(ns);
    for (int i = 0; i < ns; i++) S[i] = make_source(y + 3 * i, L[i], W[i], strike[i], dip[i]);
    Receiver R;
    for (int k = 0; k < 3; k++) {
        R.xc[k] = x[k];
        R.sv[k] = sv[k];
        R.dv[k] = dv[k];
        R.nv[k] = nv[k];
    }
    const size_t bytes = sizeof(OkadaSource) * ns + sizeof(Receiver) + sizeof(double) * 2 * NC * (size_t)ns;
    if (ensure_buf(c, c->hook, bytes + 256)) return UCY_ERR_HIP;
    char *base = (char *)c->hook.p;
    OkadaSource *d_S = (OkadaSource *)base;
    Receiver *d_R = (Receiver *)(base + sizeof(OkadaSource) * ns);
    double *d_s = (double *)(base + ((sizeof(OkadaSource) * ns + sizeof(Receiver) + 15) & ~(size_t)15));
    double *d_d = d_s + NC * (size_t)ns;
    HIPCHK(hipMemcpyAsync(d_S, S.data(), sizeof(OkadaSource) * ns, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d_R, &R, sizeof(Receiver), hipMemcpyHostToDevice, c->stream));
    launch_okada_rows(c->stream, make_medium(G, lambda), d_S, ns, d_R, 1, kind, nullptr, 0, 0, 0,
                      d_s, d_d, ns);
    HIPCHK(hipGetLastError());
    std::vector<double> hs(NC * (size_t)ns), hd(NC * (size_t)ns);
    HIPCHK(hipMemcpyAsync(hs.data(), d_s, sizeof(double) * NC * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hd.data(), d_d, sizeof(double) * NC * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // unit slips are exact; other (s, d) combine the two unit responses
    for (int k = 0; k < NC; k++)
        for (int i = 0; i < ns; i++) {
            const double vs = hs[k * (size_t)ns + i], vd = hd[k * (size_t)ns + i];
            double r;
            if (s == 1.0 && dd == 0.0) r = vs;
            else if (s == 0.0 && dd == 1.0) r = vd;
            else r = s * vs + dd * vd;
            out[k][i] = r;
        }
    return UCY_OK;
}

extern "C" int ucy_okada92_traction_kernels(ucy_ctx *c, const double x[3], const double sv[3],
                                            const double dv[3], const double nv[3], int ns,
                                            const double *y, const double *L, const double *W,
                                            const double *strike, const double *dip, double s,
                                            double d, double t, double G, double lambda,
                                            double *ts, double *td, double *tn)
{
    double *out[3] = {ts, td, tn};
    return okada_hook(c, 0, x, sv, dv, nv, ns, y, L, W, strike, dip, s, d, t, G, lambda, out);
}

extern "C" int ucy_okada92_stress_kernels(ucy_ctx *c, const double x[3], const double sv[3],
                                          const double dv[3], const double nv[3], int ns,
                                          const double *y, const double *L, const double *W,
                                          const double *strike, const double *dip, double s,
                                          double d, double t, double G, double lambda,
                                          double *s11, double *s12, double *s13, double *s22,
                                          double *s23, double *s33)
{
    double *out[6] = {s11, s12, s13, s22, s23, s33};
    return okada_hook(c, 1, x, sv, dv, nv, ns, y, L, W, strike, dip, s, d, t, G, lambda, out);
}

// ---------------------------------------------------------------------------
// measurement hook: HIP-event timing of the fused matvec + post-pass kernel
// ---------------------------------------------------------------------------
extern "C" int ucy_time_rhs_kernel(ucy_ctx *c, int iters, double *avg_seconds)
{
    if (!c || iters <= 0 || !avg_seconds) return UCY_ERR_ARG;
    const PlainDims d0 = c->K ? dims(c) : PlainDims{};
    int rc = check_ready(c, d0.dim);
    if (rc) return rc;
    const PlainDims d = dims(c);
    RkBuffers &b = c->rb;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // at the state of the last ucy_odefun call (or the last stage of ucy_integrate);
    // in reduced precision a stage RHS: t = t0 + split(K) v over the b.t buffer
    launch_prepass(c->stream, d, c->d_fr, c->d_rh, b.ytmp, b.ak[2], b.v, nullptr);
    const KOperand K{c->K, c->Kp, c->Krs, c->kprec};
    auto one = [&]() {
        if (c->kprec == UCY_PREC_F64)
            launch_gemv_post(c->stream, d, c->K, c->d_fr, b.ytmp, b.v, b.ak[2], nullptr);
        else
            launch_gemv_post_inc(c->stream, d, K, c->d_fr, b.ytmp, b.v, b.ak[2], b.t);
    };
    one();  // warm
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; i++) one();
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *avg_seconds = 1e-3 * ms / iters;
    return UCY_OK;
}
