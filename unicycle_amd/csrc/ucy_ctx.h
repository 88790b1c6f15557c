// ucy_ctx.h -- the context behind the opaque ucy_ctx handle of include/ucy.h.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "ucy_internal.h"

using ucy::Control;
using ucy::Friction;
using ucy::Rheology;
using ucy::RkBuffers;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct ucy_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // elastic moduli
    double lambda = 0, mu = 0;
    bool have_elastic = false;
    // 0: the 3-D model; 1: the 2-D antiplane model; 2: the 2-D plane-strain
    // model (the 2-D segments use the rectangle arrays: x, width, dip)
    int model = 0;
    // plane strain: initial dip traction per patch where positive (the deck's
    // tau0, fortran/2d/planestrain/src/ratestate.f90:604-609); empty: none
    std::vector<double> tau0;
    // initial stress s22, s23, s33 of each plane-strain volume (used where
    // not all zero; fortran/2d/planestrain/src/viscouscycles.f90:779-810);
    // empty: none
    std::vector<double> s0;
    // rectangles
    int nrect = 0;
    std::vector<double> rx, rL, rW, rstrike, rdip;
    // triangles
    int ntri = 0, nve = 0;
    std::vector<int> i1, i2, i3;
    std::vector<double> vert;
    // volumes (the 2-D models': upper centre x(3,nv), thickness vT, width
    // vW, dip vdip in radians)
    int nvol = 0;
    std::vector<double> vx, vL, vT, vW, vstrike, vdip;
    // friction / rheology (element order)
    std::vector<double> Vpl, rake, mu0, sig, a, b, Lc, Vo, damping;
    std::vector<double> gd0, npow, ebg, Gm, Qa, Rg, To;
    bool have_friction = false, have_rheology = false;
    // device kernel: binary64 always; packed split-bf16 pairs and row scales
    // when built in reduced precision (kprec)
    double *K = nullptr;
    int64_t m = 0, n = 0;
    // columns of the model (n is that, padded to even for the pairwise row
    // reads of the GEMVs; the antiplane model with an odd segment count pads
    // one zero column)
    int64_t nl = 0;
    uint32_t *Kp = nullptr;
    double *Krs = nullptr;
    int kprec = 0;
    // displacement kernels at observation points (buildO): o_rows x o_n
    // row-major on the device, built for a mesh with o_n columns
    double *O = nullptr;
    int64_t o_rows = 0, o_n = 0;
    // wall seconds of the last ucy_build_kernel by source type (rectangle,
    // triangle, volume columns; each phase ends with a stream synchronize)
    double build_seconds[3] = {0, 0, 0};
    // device per-element data
    Friction *d_fr = nullptr;
    Rheology *d_rh = nullptr;
    int fr_cap = 0, rh_cap = 0;   // entries the two device arrays hold
    bool params_dirty = true;
    // work buffers
    double *work = nullptr;
    size_t work_dim = 0;
    RkBuffers rb{};
    Control *d_ctl = nullptr;
    Control *h_ctl = nullptr;  // pinned
    // captured attempt
    hipGraphExec_t graph = nullptr;
    // scratch for the per-kernel hooks
    DevBuf hook;
    // ensembles created on this context (destroyed with it)
    std::vector<struct ucy_ensemble *> ensembles;
};


// host-side reference arithmetic shared by the single-instance and ensemble
// paths (capi.hip)
ucy::OkadaMedium make_medium(double G, double lambda);
ucy::OkadaSource make_source(const double *x, double L, double W, double strike, double dipd);
void refsys_okada(int ns, const double *x, const double *L, const double *W, const double *strike,
                  const double *dip, double *sv, double *dv, double *nv, double *xc);
