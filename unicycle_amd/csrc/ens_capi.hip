// ens_capi.hip -- host side of the batched-ensemble C ABI (include/ucy.h).
//
// A group = one geometry (Okada rectangles) shared by up to 64 instances.
// build(): K assembled on the device by the same Okada kernel as buildG,
// converted to the ensemble precision, tiling tuned by timing the stage GEMM
// on the device.  run(): one graph-captured lock-step RK45 attempt for all
// instances of all groups per replay.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/ucy.h"
#include "ens_internal.h"
#include "ucy_ctx.h"

using namespace ucy;

struct HostGroup {
    int N = 0, ninst = 0;
    int share = -1;   // a slice of the K family headed by group `share` (-1: a head)
    double lambda = 0, mu = 0;
    std::vector<double> x, L, W, strike, dip;
    // per instance-major [inst][e]
    std::vector<double> Vpl, rake, a, b, sig, Lc, damping, y0;
};

struct ucy_ensemble {
    ucy_ctx *ctx = nullptr;
    int mode = 0;  // 0 split-bf16, 1 binary64, 2 mixed, 3 bf16+i8
    std::vector<HostGroup> hg;
    EnsTable tab{};
    EnsTable *d_tab = nullptr;
    std::vector<void *> allocs;
    hipGraphExec_t graph = nullptr;
    bool built = false;
    int kb_per_split = 0;
    int variant = 0;  // split-bf16 GEMM variant chosen by the autotuner
    int stop_at_end = 0;
    int stop_at_event = 0;
    double event_vmax = 1e-3;
    float *d_mlp = nullptr;  // copy of the proposal network's parameters
    std::vector<double *> slogs;  // per-group step logs (null: off)
    MlpNorm mlp_norm{};
    bool use_mlp = false;
    std::vector<double> tune_seconds;
    double gemm_bytes = 0;
    int *h_done = nullptr;   // pinned host copy of the finished-instance count
};

static int efail(ucy_ensemble *E, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (E && E->ctx) E->ctx->err = buf;
    return code;
}

#define EHIP(call)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return efail(E, UCY_ERR_HIP, "%s failed: %s (%s:%d)", #call,                   \
                         hipGetErrorString(e_), __FILE__, __LINE__);                       \
    } while (0)

template <typename T>
static int ealloc(ucy_ensemble *E, T **p, size_t count)
{
    void *q = nullptr;
    EHIP(hipMalloc(&q, sizeof(T) * (count ? count : 1)));
    EHIP(hipMemsetAsync(q, 0, sizeof(T) * (count ? count : 1), E->ctx->stream));
    E->allocs.push_back(q);
    *p = (T *)q;
    return 0;
}

extern "C" ucy_ensemble *ucy_ensemble_create(ucy_ctx *ctx, int precision)
{
    if (!ctx) return nullptr;
    if (precision != UCY_PREC_SPLIT_BF16 && precision != UCY_PREC_F64 &&
        precision != UCY_PREC_MIXED && precision != UCY_PREC_BF16_I8)
        return nullptr;
    ucy_ensemble *E = new ucy_ensemble();
    E->ctx = ctx;
    E->mode = precision == UCY_PREC_SPLIT_BF16 ? 0
              : precision == UCY_PREC_F64      ? 1
              : precision == UCY_PREC_MIXED    ? 2
                                               : 3;
    ctx->ensembles.push_back(E);
    return E;
}

static void release(ucy_ensemble *E)
{
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    E->graph = nullptr;
    for (void *p : E->allocs) (void)hipFree(p);
    E->allocs.clear();
    for (double *p : E->slogs)
        if (p) (void)hipFree(p);
    E->slogs.clear();
    if (E->h_done) (void)hipHostFree(E->h_done);
    E->h_done = nullptr;
    E->d_tab = nullptr;
    E->built = false;
}

extern "C" void ucy_ensemble_destroy(ucy_ensemble *E)
{
    if (!E) return;
    (void)hipSetDevice(E->ctx->device);
    (void)hipStreamSynchronize(E->ctx->stream);
    release(E);
    if (E->d_mlp) (void)hipFree(E->d_mlp);
    auto &reg = E->ctx->ensembles;
    reg.erase(std::remove(reg.begin(), reg.end(), E), reg.end());
    delete E;
}

extern "C" int ucy_ensemble_add_group(ucy_ensemble *E, int np, const double *x, const double *length,
                                      const double *width, const double *strike, const double *dip,
                                      double lambda, double mu, int n_inst, const double *Vpl,
                                      const double *rake, const double *a, const double *b,
                                      const double *sig, const double *L, const double *damping,
                                      const double *y0)
{
    if (!E || np <= 0 || n_inst <= 0) return UCY_ERR_ARG;
    if (n_inst > kEnsCols * kMaxSlices)
        return efail(E, UCY_ERR_LIMIT, "at most %d instances per geometry", kEnsCols * kMaxSlices);
    // more than 64 instances: consecutive groups of 64 (the last one partly
    // filled) that share the first one's K
    const int slices = (n_inst + kEnsCols - 1) / kEnsCols;
    if ((int)E->hg.size() + slices > kMaxGroups)
        return efail(E, UCY_ERR_LIMIT, "at most %d groups of 64 instances", kMaxGroups);
    if (!(mu > 0)) return efail(E, UCY_ERR_ARG, "shear modulus must be positive");
    const int head = (int)E->hg.size();
    for (int sl = 0; sl < slices; sl++) {
        HostGroup G;
        G.N = np;
        G.ninst = std::min(kEnsCols, n_inst - kEnsCols * sl);
        G.share = sl ? head : -1;
        G.lambda = lambda;
        G.mu = mu;
        if (sl == 0) {   // the geometry is only read for the head's K
            G.x.assign(x, x + 3 * (size_t)np);
            G.L.assign(length, length + np);
            G.W.assign(width, width + np);
            G.strike.assign(strike, strike + np);
            G.dip.assign(dip, dip + np);
        }
        const size_t o = (size_t)np * kEnsCols * sl, ne = (size_t)np * G.ninst;
        G.Vpl.assign(Vpl + o, Vpl + o + ne);
        G.rake.assign(rake + o, rake + o + ne);
        G.a.assign(a + o, a + o + ne);
        G.b.assign(b + o, b + o + ne);
        G.sig.assign(sig + o, sig + o + ne);
        G.Lc.assign(L + o, L + o + ne);
        G.damping.assign(damping + o, damping + o + ne);
        G.y0.assign(y0 + 7 * o, y0 + 7 * (o + ne));
        E->hg.push_back(std::move(G));
    }
    release(E);
    return head;
}

// upload y0 (reference layout per instance) into the SoA state and reset control
static int load_initial(ucy_ensemble *E, double dt0)
{
    ucy_ctx *c = E->ctx;
    for (size_t gi = 0; gi < E->hg.size(); gi++) {
        const HostGroup &H = E->hg[gi];
        EnsGroup &G = E->tab.g[gi];
        const int N = H.N;
        std::vector<double> Y((size_t)7 * N * kEnsCols, 0.0);
        for (int i = 0; i < H.ninst; i++)
            for (int e = 0; e < N; e++)
                for (int k = 0; k < 7; k++)
                    Y[((size_t)k * N + e) * kEnsCols + i] = H.y0[((size_t)i * N + e) * 7 + k];
        EHIP(hipMemcpyAsync(G.y, Y.data(), sizeof(double) * Y.size(), hipMemcpyHostToDevice, c->stream));
        std::vector<double> hv(kEnsCols, dt0), zero(kEnsCols, 0.0);
        std::vector<int> one(kEnsCols, 1), izero(kEnsCols, 0);
        std::vector<long long> lzero(kEnsCols, 0);
        EHIP(hipMemcpyAsync(G.h, hv.data(), sizeof(double) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.hnext, hv.data(), sizeof(double) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.time, zero.data(), sizeof(double) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.vmax, zero.data(), sizeof(double) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        std::vector<double> none(kEnsCols, -1.0);
        EHIP(hipMemcpyAsync(G.t_event, none.data(), sizeof(double) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        std::vector<double> hz(5 * kEnsCols, 0.0);
        EHIP(hipMemcpyAsync(G.hist, hz.data(), sizeof(double) * hz.size(), hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.errmax, zero.data(), sizeof(double) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        std::vector<double> zrep(kMaxRep * kEnsCols, 0.0);
        EHIP(hipMemcpyAsync(G.vmax_rep, zrep.data(), sizeof(double) * zrep.size(), hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.errmax_rep, zrep.data(), sizeof(double) * zrep.size(), hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.need_rhs, one.data(), sizeof(int) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.accepted, izero.data(), sizeof(int) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.done, izero.data(), sizeof(int) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.status, izero.data(), sizeof(int) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.steps, lzero.data(), sizeof(long long) * kEnsCols, hipMemcpyHostToDevice, c->stream));
        EHIP(hipMemcpyAsync(G.attempts, lzero.data(), sizeof(long long) * kEnsCols, hipMemcpyHostToDevice, c->stream));
    }
    int z = 0;
    EHIP(hipMemcpyAsync(E->tab.all_done, &z, sizeof(int), hipMemcpyHostToDevice, c->stream));
    EHIP(hipStreamSynchronize(c->stream));
    return 0;
}

static void set_split(EnsTable &T, int kbps, int variant)
{
    T.variant = variant;
    const int rb = ens_gemm_variant_rows(variant) / 32;
    for (int gi = 0; gi < T.ngroups; gi++) {
        EnsGroup &G = T.g[gi];
        int k = std::max(1, std::min(kbps, G.nkb));
        G.ksplit = (G.nkb + k - 1) / k;
        G.kb_per_split = (G.nkb + G.ksplit - 1) / G.ksplit;
    }
    // GEMM blocks per K family: work items (row item, k-split, slice), 4
    // waves per block
    int acc = 0, acc64 = 0;
    for (int f = 0; f < T.nfam; f++) {
        const EnsGroup &G = T.g[T.fam_head[f]];
        const int S = T.fam_slices[f];
        T.blk_gemm[f] = acc;
        T.blk_gemm64[f] = acc64;
        const int nri = (G.mp / 32 + rb - 1) / rb;
        // LDS variants (one slice only): one block per (4 consecutive row
        // items, k-split); K-sharing variant: one block per (row item,
        // k-split), its waves the slices
        acc += ens_gemm_variant_lds(variant)      ? ((nri + 3) / 4) * G.ksplit
               : ens_gemm_variant_kshare(variant) ? nri * G.ksplit
                                                  : (nri * G.ksplit * S + 3) / 4;
        acc64 += ((G.mp / 64) * G.ksplit * S + 3) / 4;
    }
    T.blk_gemm[T.nfam] = acc;
    T.blk_gemm64[T.nfam] = acc64;
}

static bool has_shared_families(const EnsTable &T)
{
    for (int f = 0; f < T.nfam; f++)
        if (T.fam_slices[f] > 1) return true;
    return false;
}

static int upload_table(ucy_ensemble *E)
{
    EHIP(hipMemcpyAsync(E->d_tab, &E->tab, sizeof(EnsTable), hipMemcpyHostToDevice, E->ctx->stream));
    EHIP(hipStreamSynchronize(E->ctx->stream));
    return 0;
}

static const int kMaxSplit = 16;
// the direct global->LDS variants (6, 7) of the bf16+i8 stage GEMM
static const bool kDmaVariantsOn = true;

extern "C" int ucy_ensemble_build(ucy_ensemble *E)
{
    if (!E) return UCY_ERR_ARG;
    if (E->hg.empty()) return efail(E, UCY_ERR_STATE, "no groups");
    ucy_ctx *c = E->ctx;
    EHIP(hipSetDevice(c->device));
    release(E);
    EnsTable &T = E->tab;
    memset(&T, 0, sizeof(T));
    T.ngroups = (int)E->hg.size();
    int eacc = 0, iacc = 0, total = 0;
    double bytes = 0;
    for (int gi = 0; gi < T.ngroups; gi++) {
        const HostGroup &H = E->hg[gi];
        EnsGroup &G = T.g[gi];
        G.N = H.N;
        G.ninst = H.ninst;
        G.m = 3 * H.N;
        G.n = 2 * H.N;
        G.mp = (G.m + 63) / 64 * 64;
        G.np = (G.n + 15) / 16 * 16;
        G.nkb = G.np / 16;
        G.mode = E->mode;
        total += H.ninst;
        // ---- K on the device (greens.f90:73-275 for rectangle receivers),
        // once per K family: its other slices take the head's tiles
        const bool head = H.share < 0;
        double *K64 = nullptr;
        if (head) {
            EHIP(hipMalloc((void **)&K64, sizeof(double) * (size_t)G.m * G.n));
            E->allocs.push_back(K64);   // freed below, or by release() on an error
        }
        if (head) {
            std::vector<double> sv(3 * (size_t)H.N), dv(3 * (size_t)H.N), nv(3 * (size_t)H.N),
                xc(3 * (size_t)H.N);
            refsys_okada(H.N, H.x.data(), H.L.data(), H.W.data(), H.strike.data(), H.dip.data(),
                         sv.data(), dv.data(), nv.data(), xc.data());
            std::vector<Receiver> R(H.N);
            std::vector<OkadaSource> S(H.N);
            for (int e = 0; e < H.N; e++) {
                for (int k = 0; k < 3; k++) {
                    R[e].xc[k] = xc[3 * e + k];
                    R[e].sv[k] = sv[3 * e + k];
                    R[e].dv[k] = dv[3 * e + k];
                    R[e].nv[k] = nv[3 * e + k];
                }
                S[e] = make_source(&H.x[3 * e], H.L[e], H.W[e], H.strike[e], H.dip[e]);
            }
            Receiver *dR = nullptr;
            OkadaSource *dS = nullptr;
            EHIP(hipMalloc((void **)&dR, sizeof(Receiver) * H.N));
            EHIP(hipMalloc((void **)&dS, sizeof(OkadaSource) * H.N));
            EHIP(hipMemcpyAsync(dR, R.data(), sizeof(Receiver) * H.N, hipMemcpyHostToDevice, c->stream));
            EHIP(hipMemcpyAsync(dS, S.data(), sizeof(OkadaSource) * H.N, hipMemcpyHostToDevice, c->stream));
            const OkadaMedium M = make_medium(H.mu, H.lambda);
            for (int r0 = 0; r0 < H.N; r0 += 65535)
                launch_okada_rows(c->stream, M, dS, H.N, dR + r0, std::min(65535, H.N - r0), 0, K64,
                                  G.n, 3 * (int64_t)r0, 0, nullptr, nullptr, 0);
            EHIP(hipGetLastError());
            EHIP(hipStreamSynchronize(c->stream));
            EHIP(hipFree(dR));
            EHIP(hipFree(dS));
        }
        // stage-GEMM bytes per launch: split bf16 (4 bytes per entry) for
        // modes 0 and 2, bf16+i8 (3 bytes) for mode 3
        if (G.mode != 1) {
            if (head) {
                const double kbytes = G.mode == 3 ? 3.0 : 4.0;
                uint16_t *tiles = nullptr;
                double *rs = nullptr;
                if (ealloc(E, &tiles, (size_t)G.mp * G.np * (G.mode == 3 ? 3 : 4) / 2))
                    return UCY_ERR_HIP;
                if (ealloc(E, &rs, G.mp)) return UCY_ERR_HIP;
                if (G.mode == 3)
                    ens_launch_split3_K(c->stream, K64, G.m, G.n, G.mp, G.np, (uint8_t *)tiles, rs);
                else
                    ens_launch_split_K(c->stream, K64, G.m, G.n, G.mp, G.np, tiles, rs);
                EHIP(hipGetLastError());
                G.Ktiles = tiles;
                G.rscale = rs;
                bytes += kbytes * G.mp * G.np + 8.0 * G.mp;
            } else {
                G.Ktiles = T.g[H.share].Ktiles;
                G.rscale = T.g[H.share].rscale;
            }
            if (ealloc(E, &G.Vtiles, (size_t)G.np * kEnsCols * 2)) return UCY_ERR_HIP;
            if (ealloc(E, &G.Tpart, (size_t)kMaxSplit * G.mp * kEnsCols)) return UCY_ERR_HIP;
            if (ealloc(E, &G.v0, (size_t)G.np * kEnsCols)) return UCY_ERR_HIP;
            if (ealloc(E, &G.t0, (size_t)G.mp * kEnsCols)) return UCY_ERR_HIP;
            bytes += 4.0 * G.np * kEnsCols;
        }
        if (G.mode == 1 || G.mode == 2) {
            if (head) {
                double *tiles = nullptr;
                if (ealloc(E, &tiles, (size_t)G.mp * G.np)) return UCY_ERR_HIP;
                ens_launch_tile_K64(c->stream, K64, G.m, G.n, G.mp, G.np, tiles);
                EHIP(hipGetLastError());
                G.K64 = tiles;
                if (G.mode == 1) bytes += 8.0 * G.mp * G.np;
            } else {
                G.K64 = T.g[H.share].K64;
            }
            if (ealloc(E, &G.V64, (size_t)G.np * kEnsCols)) return UCY_ERR_HIP;
            if (ealloc(E, &G.T64, (size_t)kMaxSplit * G.mp * kEnsCols)) return UCY_ERR_HIP;
            if (G.mode == 1) bytes += 8.0 * G.np * kEnsCols;
        }
        EHIP(hipStreamSynchronize(c->stream));
        if (K64) {
            E->allocs.erase(std::find(E->allocs.begin(), E->allocs.end(), (void *)K64));
            EHIP(hipFree(K64));
        }
        if (ealloc(E, &G.any_rhs, 1)) return UCY_ERR_HIP;
        // ---- state and control
        const size_t S7 = (size_t)7 * G.N * kEnsCols;
        double **vecs[] = {&G.y, &G.dydt, &G.ytmp, &G.ak[2], &G.ak[3], &G.ak[4],
                           &G.ak[5], &G.ak[6], &G.ycand};
        for (double **v : vecs)
            if (ealloc(E, v, S7)) return UCY_ERR_HIP;
        G.ak[0] = G.ak[1] = G.dydt;
        // friction SoA [e][inst]
        const size_t F = (size_t)G.N * kEnsCols;
        double **fv[] = {&G.fr.asig, &G.fr.bsig, &G.fr.L, &G.fr.damping, &G.fr.vplc, &G.fr.vpls};
        for (double **v : fv)
            if (ealloc(E, v, F)) return UCY_ERR_HIP;
        {
            std::vector<double> fas(F, 0.0), fbs(F, 0.0), fl(F, 1.0), fd(F, 0.0), fc(F, 0.0),
                fsn(F, 0.0);
            for (int i = 0; i < H.ninst; i++)
                for (int e = 0; e < H.N; e++) {
                    const size_t src = (size_t)i * H.N + e, dst = (size_t)e * kEnsCols + i;
                    fas[dst] = H.a[src] * H.sig[src];
                    fbs[dst] = H.b[src] * H.sig[src];
                    fl[dst] = H.Lc[src];
                    fd[dst] = H.damping[src];
                    fc[dst] = H.Vpl[src] * cos(H.rake[src]);
                    fsn[dst] = H.Vpl[src] * sin(H.rake[src]);
                }
            const std::vector<double> *hs[] = {&fas, &fbs, &fl, &fd, &fc, &fsn};
            double *ds[] = {G.fr.asig, G.fr.bsig, G.fr.L, G.fr.damping, G.fr.vplc, G.fr.vpls};
            for (int k = 0; k < 6; k++)
                EHIP(hipMemcpyAsync(ds[k], hs[k]->data(), sizeof(double) * F, hipMemcpyHostToDevice, c->stream));
            EHIP(hipStreamSynchronize(c->stream));
        }
        double **cv[] = {&G.time, &G.h, &G.hnext, &G.hdid, &G.errmax, &G.vmax, &G.t_event,
                         &G.hstart};
        for (double **v : cv)
            if (ealloc(E, v, kEnsCols)) return UCY_ERR_HIP;
        int **ci[] = {&G.need_rhs, &G.accepted, &G.done, &G.status};
        for (int **v : ci)
            if (ealloc(E, v, kEnsCols)) return UCY_ERR_HIP;
        if (ealloc(E, &G.hist, 5 * kEnsCols)) return UCY_ERR_HIP;
        if (ealloc(E, &G.vmax_rep, kMaxRep * kEnsCols)) return UCY_ERR_HIP;
        if (ealloc(E, &G.errmax_rep, kMaxRep * kEnsCols)) return UCY_ERR_HIP;
        G.slog = nullptr;
        G.slog_cap = 0;
        if (ealloc(E, &G.steps, kEnsCols)) return UCY_ERR_HIP;
        if (ealloc(E, &G.attempts, kEnsCols)) return UCY_ERR_HIP;
        T.blk_elem[gi] = eacc;
        eacc += (G.N + 3) / 4;
        T.blk_inst[gi] = iacc;
        iacc += 1;
    }
    T.blk_elem[T.ngroups] = eacc;
    T.blk_inst[T.ngroups] = iacc;
    // K families: a head and the slices that follow it
    T.nfam = 0;
    for (int gi = 0; gi < T.ngroups; gi++) {
        if (E->hg[gi].share < 0) {
            T.fam_head[T.nfam] = gi;
            T.fam_slices[T.nfam] = 1;
            T.nfam++;
        } else {
            T.fam_slices[T.nfam - 1]++;
        }
    }
    T.total_inst = total;
    T.t_end = 0;
    T.epsilon = 1e-6;
    T.max_step = 1.7976931348623157e308;
    T.max_attempts = 1;
    if (ealloc(E, &T.all_done, 1)) return UCY_ERR_HIP;
    void *dt = nullptr;
    EHIP(hipMalloc(&dt, sizeof(EnsTable)));
    E->allocs.push_back(dt);
    E->d_tab = (EnsTable *)dt;
    EHIP(hipHostMalloc((void **)&E->h_done, sizeof(int), hipHostMallocDefault));
    E->gemm_bytes = bytes;
    // ---- tune the split-K tiling on the device: time the all-groups GEMM
    E->tune_seconds.clear();
    int best = 0, best_v = 0;
    double best_t = 1e30;
    // split depths (k-blocks of 16 per split); the table below is
    // UCY_GEMM_TUNE_ENTRIES = kGemmVariants x kDepths entries
    // (48, 80 and 104 give the odd split counts: 5 splits of a 512-k-block
    // K at 104, where 4 and 6 leave a second, half-empty round of blocks)
    constexpr int kDepths = 11;
    const int cands[kDepths] = {32, 48, 64, 80, 96, 104, 128, 160, 192, 256, 512};
    static_assert(kGemmVariants * kDepths == UCY_GEMM_TUNE_ENTRIES, "tuning table size");
    {
        hipEvent_t e0, e1;
        EHIP(hipEventCreate(&e0));
        EHIP(hipEventCreate(&e1));
        // (split-bf16 variant, k-blocks per split); the binary64 GEMM has one
        // variant
        for (int ci = 0; ci < kGemmVariants * kDepths; ci++) {
            const int kb = cands[ci % kDepths], v = ci / kDepths;
            // binary64 has one variant; the LDS variants stage one V per
            // block and so serve families of one slice only
            if ((v > 0 && E->mode == 1) || (ens_gemm_variant_lds(v) && has_shared_families(T)) ||
                (ens_gemm_variant_dma(v) && (E->mode != 3 || !kDmaVariantsOn))) {
                E->tune_seconds.push_back(-1.0);
                continue;
            }
            set_split(T, kb, v);
            bool ok = true;
            for (int gi = 0; gi < T.ngroups; gi++)
                if (T.g[gi].ksplit > kMaxSplit) ok = false;
            if (!ok) {
                E->tune_seconds.push_back(-1.0);
                continue;
            }
            if (upload_table(E)) return UCY_ERR_HIP;
            // 2 warm launches, 15 timed (single launches vary by ~10 %); a
            // candidate whose launch fails is not admissible
            for (int r = 0; r < 2; r++) ens_launch_gemm(c->stream, E->d_tab, T, 2);
            if (hipGetLastError() != hipSuccess) {
                EHIP(hipStreamSynchronize(c->stream));
                E->tune_seconds.push_back(-1.0);
                continue;
            }
            EHIP(hipEventRecord(e0, c->stream));
            for (int r = 0; r < 15; r++) ens_launch_gemm(c->stream, E->d_tab, T, 2);
            EHIP(hipGetLastError());
            EHIP(hipEventRecord(e1, c->stream));
            EHIP(hipEventSynchronize(e1));
            float ms = 0;
            EHIP(hipEventElapsedTime(&ms, e0, e1));
            const double t = 1e-3 * ms / 15;
            E->tune_seconds.push_back(t);
            // the choice also charges the post-pass's read of the partial
            // products, one per split, at 4 TB/s (what the attempt measures:
            // a deeper split's extra partials cost the post kernels about
            // that), so that a split that only moves time from the GEMM into
            // the post-pass is not taken
            double pbytes = 0.0;
            for (int gi = 0; gi < T.ngroups; gi++)
                pbytes += (double)T.g[gi].ksplit * T.g[gi].mp * kEnsCols * (E->mode == 1 ? 8 : 4);
            const double teff = t + pbytes / 4.0e12;
            if (teff < best_t) {
                best_t = teff;
                best = kb;
                best_v = v;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (!best) return efail(E, UCY_ERR_STATE, "no admissible GEMM tiling");
    }
    E->kb_per_split = best;
    E->variant = best_v;
    set_split(T, best, best_v);
    if (upload_table(E)) return UCY_ERR_HIP;
    E->built = true;
    return load_initial(E, 1.0e-3);
}

extern "C" int ucy_ensemble_reset(ucy_ensemble *E, double dt0)
{
    if (!E) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    if (!(dt0 > 0)) return efail(E, UCY_ERR_ARG, "dt0 must be positive");
    EHIP(hipSetDevice(E->ctx->device));
    return load_initial(E, dt0);
}

static int capture(ucy_ensemble *E)
{
    if (E->graph) return 0;
    ucy_ctx *c = E->ctx;
    const EnsTable &T = E->tab;
    hipGraph_t g = nullptr;
    EHIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    ens_launch_start(c->stream, E->d_tab, T);
    for (int s = 1; s <= 5; s++) {
        ens_launch_gemm(c->stream, E->d_tab, T, s);
        ens_launch_post_stage(c->stream, E->d_tab, T, s);
    }
    ens_launch_gemm(c->stream, E->d_tab, T, 6);
    ens_launch_post_finish(c->stream, E->d_tab, T);
    ens_launch_control(c->stream, E->d_tab, T);
    EHIP(hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(&E->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return efail(E, UCY_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int ucy_ensemble_run(ucy_ensemble *E, double t_end, int64_t max_attempts,
                                double epsilon, double max_step, ucy_ensemble_stats *stats)
{
    if (!E) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    if (!(t_end > 0) || max_attempts <= 0 || !(epsilon > 0))
        return efail(E, UCY_ERR_ARG, "invalid run parameters");
    ucy_ctx *c = E->ctx;
    EHIP(hipSetDevice(c->device));
    E->tab.t_end = t_end;
    E->tab.epsilon = epsilon;
    E->tab.max_step = max_step;
    E->tab.max_attempts = max_attempts;
    E->tab.stop_at_end = E->stop_at_end;
    E->tab.stop_at_event = E->stop_at_event;
    E->tab.event_vmax = E->event_vmax;
    E->tab.mlp = E->use_mlp ? E->d_mlp : nullptr;
    E->tab.mlp_norm = E->mlp_norm;
    if (upload_table(E)) return UCY_ERR_HIP;
    if (capture(E)) return UCY_ERR_HIP;
    // continue from the current state: re-arm every instance short of t_end
    EHIP(hipMemsetAsync(E->tab.all_done, 0, sizeof(int), c->stream));
    ens_launch_rearm(c->stream, E->d_tab, E->tab);
    EHIP(hipGetLastError());
    int *h_done = E->h_done;   // allocated with the ensemble, freed by release()
    EHIP(hipMemcpyAsync(h_done, E->tab.all_done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    EHIP(hipStreamSynchronize(c->stream));
    auto t0 = std::chrono::steady_clock::now();
    int batch = 8;
    int64_t replays = 0;
    for (; *h_done < E->tab.total_inst;) {
        for (int i = 0; i < batch; i++) EHIP(hipGraphLaunch(E->graph, c->stream));
        replays += batch;
        EHIP(hipMemcpyAsync(h_done, E->tab.all_done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        EHIP(hipStreamSynchronize(c->stream));
        if (*h_done >= E->tab.total_inst) break;
        if (batch < 64) batch *= 2;
    }
    // the last attempt's accepted steps (later attempts commit at their start)
    ens_launch_commit(c->stream, E->d_tab, E->tab);
    EHIP(hipStreamSynchronize(c->stream));
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->min_time = 1e300;
        for (int gi = 0; gi < E->tab.ngroups; gi++) {
            const EnsGroup &G = E->tab.g[gi];
            std::vector<double> tm(kEnsCols);
            std::vector<long long> st(kEnsCols), at(kEnsCols);
            std::vector<int> status(kEnsCols);
            EHIP(hipMemcpy(tm.data(), G.time, sizeof(double) * kEnsCols, hipMemcpyDeviceToHost));
            EHIP(hipMemcpy(st.data(), G.steps, sizeof(long long) * kEnsCols, hipMemcpyDeviceToHost));
            EHIP(hipMemcpy(at.data(), G.attempts, sizeof(long long) * kEnsCols, hipMemcpyDeviceToHost));
            EHIP(hipMemcpy(status.data(), G.status, sizeof(int) * kEnsCols, hipMemcpyDeviceToHost));
            for (int i = 0; i < G.ninst; i++) {
                stats->inst_time_sum += tm[i];
                stats->min_time = std::min(stats->min_time, tm[i]);
                stats->max_time = std::max(stats->max_time, tm[i]);
                stats->steps += st[i];
                stats->attempts += at[i];
                stats->n_failed += status[i] != 0;
            }
            stats->n_inst += G.ninst;
        }
        stats->replays = replays;
        stats->wall_seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return UCY_OK;
}

extern "C" int ucy_ensemble_set_stop_at_end(ucy_ensemble *E, int on)
{
    if (!E) return UCY_ERR_ARG;
    E->stop_at_end = on ? 1 : 0;
    return UCY_OK;
}

namespace ucy {
const float *mlp_device_params(const ucy_mlp *M);
MlpNorm mlp_norm(const ucy_mlp *M);
}

extern "C" int ucy_ensemble_set_mlp(ucy_ensemble *E, const ucy_mlp *M)
{
    if (!E) return UCY_ERR_ARG;
    EHIP(hipSetDevice(E->ctx->device));
    if (!M) {
        E->use_mlp = false;
        return UCY_OK;
    }
    if (!E->d_mlp) EHIP(hipMalloc((void **)&E->d_mlp, sizeof(float) * kMlpParams));
    EHIP(hipMemcpy(E->d_mlp, mlp_device_params(M), sizeof(float) * kMlpParams, hipMemcpyDeviceToDevice));
    E->mlp_norm = mlp_norm(M);
    E->use_mlp = true;
    return UCY_OK;
}

extern "C" int ucy_ensemble_set_events(ucy_ensemble *E, double vmax_threshold, int stop_at_event)
{
    if (!E || !(vmax_threshold > 0)) return UCY_ERR_ARG;
    E->event_vmax = vmax_threshold;
    E->stop_at_event = stop_at_event ? 1 : 0;
    return UCY_OK;
}

extern "C" int ucy_ensemble_first_events(ucy_ensemble *E, int group, double *t_event)
{
    if (!E || group < 0 || group >= (int)E->hg.size() || !t_event) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    EHIP(hipSetDevice(E->ctx->device));
    EHIP(hipStreamSynchronize(E->ctx->stream));
    std::vector<double> tmp(kEnsCols);
    EHIP(hipMemcpy(tmp.data(), E->tab.g[group].t_event, sizeof(double) * kEnsCols, hipMemcpyDeviceToHost));
    std::copy(tmp.begin(), tmp.begin() + E->hg[group].ninst, t_event);
    return UCY_OK;
}

extern "C" int ucy_ensemble_set_step_log(ucy_ensemble *E, int group, long long cap)
{
    if (!E || group < 0 || group >= (int)E->hg.size() || cap < 0) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    EHIP(hipSetDevice(E->ctx->device));
    EHIP(hipStreamSynchronize(E->ctx->stream));
    if (E->slogs.size() < E->hg.size()) E->slogs.resize(E->hg.size(), nullptr);
    EnsGroup &G = E->tab.g[group];
    if (E->slogs[group]) EHIP(hipFree(E->slogs[group]));
    E->slogs[group] = nullptr;
    G.slog = nullptr;
    G.slog_cap = 0;
    if (cap > 0) {
        double *p = nullptr;
        const size_t bytes = sizeof(double) * 5 * kEnsCols * (size_t)cap;
        if (hipMalloc((void **)&p, bytes) != hipSuccess)
            return efail(E, UCY_ERR_HIP, "step log of %lld steps does not fit", cap);
        E->slogs[group] = p;
        G.slog = p;
        G.slog_cap = cap;
    }
    return UCY_OK;
}

extern "C" int ucy_ensemble_get_step_log(ucy_ensemble *E, int group, long long cap_out,
                                         double *t, double *dt, double *vmax, double *errmax,
                                         double *rejects, long long *count)
{
    if (!E || group < 0 || group >= (int)E->hg.size() || !count || cap_out < 0)
        return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    const EnsGroup &G = E->tab.g[group];
    if (!G.slog) return efail(E, UCY_ERR_STATE, "no step log on group %d", group);
    EHIP(hipSetDevice(E->ctx->device));
    EHIP(hipStreamSynchronize(E->ctx->stream));
    const int ni = E->hg[group].ninst;
    std::vector<long long> steps(kEnsCols);
    EHIP(hipMemcpy(steps.data(), G.steps, sizeof(long long) * kEnsCols, hipMemcpyDeviceToHost));
    const size_t cap = (size_t)G.slog_cap;
    std::vector<double> buf(5 * kEnsCols * cap);
    EHIP(hipMemcpy(buf.data(), G.slog, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    double *out[5] = {t, dt, vmax, errmax, rejects};
    for (int i = 0; i < ni; i++) {
        count[i] = std::min<long long>(std::min<long long>(steps[i], (long long)cap), cap_out);
        for (int f = 0; f < 5; f++) {
            if (!out[f]) continue;
            for (size_t k = 0; k < (size_t)count[i]; k++)
                out[f][(size_t)i * cap_out + k] = buf[(k * 5 + f) * kEnsCols + i];
        }
    }
    return UCY_OK;
}

extern "C" int ucy_ensemble_status(ucy_ensemble *E, int group, int *status)
{
    if (!E || group < 0 || group >= (int)E->hg.size() || !status) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    EHIP(hipSetDevice(E->ctx->device));
    EHIP(hipStreamSynchronize(E->ctx->stream));
    std::vector<int> tmp(kEnsCols);
    EHIP(hipMemcpy(tmp.data(), E->tab.g[group].status, sizeof(int) * kEnsCols, hipMemcpyDeviceToHost));
    std::copy(tmp.begin(), tmp.begin() + E->hg[group].ninst, status);
    return UCY_OK;
}

extern "C" int ucy_ensemble_get_state(ucy_ensemble *E, int group, double *y, double *time,
                                      double *dt_next, int64_t *steps, int64_t *attempts)
{
    if (!E || group < 0 || group >= (int)E->hg.size()) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    EHIP(hipSetDevice(E->ctx->device));
    EHIP(hipStreamSynchronize(E->ctx->stream));
    const EnsGroup &G = E->tab.g[group];
    const HostGroup &H = E->hg[group];
    if (y) {
        std::vector<double> Y((size_t)7 * G.N * kEnsCols);
        EHIP(hipMemcpy(Y.data(), G.y, sizeof(double) * Y.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < H.ninst; i++)
            for (int e = 0; e < H.N; e++)
                for (int k = 0; k < 7; k++)
                    y[((size_t)i * H.N + e) * 7 + k] = Y[((size_t)k * G.N + e) * kEnsCols + i];
    }
    std::vector<double> tmp(kEnsCols);
    std::vector<long long> lt(kEnsCols);
    if (time) {
        EHIP(hipMemcpy(tmp.data(), G.time, sizeof(double) * kEnsCols, hipMemcpyDeviceToHost));
        std::copy(tmp.begin(), tmp.begin() + H.ninst, time);
    }
    if (dt_next) {
        EHIP(hipMemcpy(tmp.data(), G.h, sizeof(double) * kEnsCols, hipMemcpyDeviceToHost));
        std::copy(tmp.begin(), tmp.begin() + H.ninst, dt_next);
    }
    if (steps) {
        EHIP(hipMemcpy(lt.data(), G.steps, sizeof(long long) * kEnsCols, hipMemcpyDeviceToHost));
        for (int i = 0; i < H.ninst; i++) steps[i] = lt[i];
    }
    if (attempts) {
        EHIP(hipMemcpy(lt.data(), G.attempts, sizeof(long long) * kEnsCols, hipMemcpyDeviceToHost));
        for (int i = 0; i < H.ninst; i++) attempts[i] = lt[i];
    }
    return UCY_OK;
}

extern "C" int ucy_ensemble_set_tiling(ucy_ensemble *E, int kb_per_split, int variant)
{
    if (!E || kb_per_split <= 0 || variant < 0 || variant >= kGemmVariants) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    if (E->mode == 1 && variant != 0)
        return efail(E, UCY_ERR_ARG, "the binary64 GEMM has one variant");
    if (ens_gemm_variant_lds(variant) && has_shared_families(E->tab))
        return efail(E, UCY_ERR_ARG, "variant %d stages one V per block: not with shared K",
                     variant);
    if (ens_gemm_variant_dma(variant) && E->mode != 3)
        return efail(E, UCY_ERR_ARG, "variant %d streams bf16+i8 K tiles only", variant);
    if (ens_gemm_variant_dma(variant) && !kDmaVariantsOn)
        return efail(E, UCY_ERR_ARG, "variant %d is switched off", variant);
    EnsTable T = E->tab;
    set_split(T, kb_per_split, variant);
    for (int gi = 0; gi < T.ngroups; gi++)
        if (T.g[gi].ksplit > kMaxSplit)
            return efail(E, UCY_ERR_LIMIT, "%d k-blocks per split needs more than %d splits",
                         kb_per_split, kMaxSplit);
    EHIP(hipSetDevice(E->ctx->device));
    EHIP(hipStreamSynchronize(E->ctx->stream));
    E->tab = T;
    E->kb_per_split = kb_per_split;
    E->variant = variant;
    if (E->graph) (void)hipGraphExecDestroy(E->graph);   // grid sizes change
    E->graph = nullptr;
    return upload_table(E);
}

extern "C" int ucy_ensemble_tuning(const ucy_ensemble *E, int *kb_per_split, int *variant,
                                   double *gemm_seconds)
{
    if (!E) return UCY_ERR_ARG;
    if (kb_per_split) *kb_per_split = E->kb_per_split;
    if (variant) *variant = E->variant;
    if (gemm_seconds)
        for (size_t i = 0; i < UCY_GEMM_TUNE_ENTRIES; i++)
            gemm_seconds[i] = i < E->tune_seconds.size() ? E->tune_seconds[i] : -1.0;
    return UCY_OK;
}

extern "C" int ucy_ensemble_time_gemm(ucy_ensemble *E, int iters, double *avg_seconds, double *bytes)
{
    if (!E || iters <= 0) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    ucy_ctx *c = E->ctx;
    EHIP(hipSetDevice(c->device));
    hipEvent_t e0, e1;
    EHIP(hipEventCreate(&e0));
    EHIP(hipEventCreate(&e1));
    ens_launch_gemm(c->stream, E->d_tab, E->tab, 2);
    EHIP(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; i++) ens_launch_gemm(c->stream, E->d_tab, E->tab, 2);
    EHIP(hipEventRecord(e1, c->stream));
    EHIP(hipEventSynchronize(e1));
    float ms = 0;
    EHIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (avg_seconds) *avg_seconds = 1e-3 * ms / iters;
    if (bytes) *bytes = E->gemm_bytes;
    return UCY_OK;
}

extern "C" int ucy_ensemble_time_phases(ucy_ensemble *E, int attempts, double *phase_seconds)
{
    if (!E || attempts <= 0 || !phase_seconds) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    ucy_ctx *c = E->ctx;
    EHIP(hipSetDevice(c->device));
    // an unbounded run: no end time, no attempt cap, no event stop
    E->tab.t_end = 1e300;
    E->tab.max_attempts = (long long)1 << 62;
    E->tab.stop_at_end = 0;
    E->tab.stop_at_event = 0;
    E->tab.event_vmax = E->event_vmax;
    E->tab.mlp = E->use_mlp ? E->d_mlp : nullptr;
    E->tab.mlp_norm = E->mlp_norm;
    if (upload_table(E)) return UCY_ERR_HIP;
    EHIP(hipMemsetAsync(E->tab.all_done, 0, sizeof(int), c->stream));
    ens_launch_rearm(c->stream, E->d_tab, E->tab);
    EHIP(hipGetLastError());
    hipEvent_t ev[UCY_ENS_PHASES + 1];
    for (int i = 0; i <= UCY_ENS_PHASES; i++) {
        hipError_t e = hipEventCreate(&ev[i]);
        if (e != hipSuccess) {
            for (int j = 0; j < i; j++) (void)hipEventDestroy(ev[j]);
            return efail(E, UCY_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
        }
    }
    for (int i = 0; i < UCY_ENS_PHASES; i++) phase_seconds[i] = 0.0;
    int rc = UCY_OK;
    for (int a = 0; a < attempts && rc == UCY_OK; a++) {
        int p = 0;
        (void)hipEventRecord(ev[p++], c->stream);
        ens_launch_start(c->stream, E->d_tab, E->tab);
        (void)hipEventRecord(ev[p++], c->stream);
        for (int s = 1; s <= 5; s++) {
            ens_launch_gemm(c->stream, E->d_tab, E->tab, s);
            (void)hipEventRecord(ev[p++], c->stream);
            ens_launch_post_stage(c->stream, E->d_tab, E->tab, s);
            (void)hipEventRecord(ev[p++], c->stream);
        }
        ens_launch_gemm(c->stream, E->d_tab, E->tab, 6);
        (void)hipEventRecord(ev[p++], c->stream);
        ens_launch_post_finish(c->stream, E->d_tab, E->tab);
        (void)hipEventRecord(ev[p++], c->stream);
        ens_launch_control(c->stream, E->d_tab, E->tab);
        (void)hipEventRecord(ev[p++], c->stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipEventSynchronize(ev[UCY_ENS_PHASES]);
        for (int i = 0; i < UCY_ENS_PHASES && e == hipSuccess; i++) {
            float ms = 0;
            e = hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            phase_seconds[i] += 1e-3 * ms;
        }
        if (e != hipSuccess) rc = efail(E, UCY_ERR_HIP, "phase timing: %s", hipGetErrorString(e));
    }
    if (rc == UCY_OK) {
        ens_launch_commit(c->stream, E->d_tab, E->tab);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = efail(E, UCY_ERR_HIP, "commit: %s", hipGetErrorString(e));
    }
    for (int i = 0; i <= UCY_ENS_PHASES; i++) (void)hipEventDestroy(ev[i]);
    return rc;
}

namespace ucy {
void ens_launch_matvec_group(hipStream_t st, const EnsTable *dtab, const EnsTable &htab, int g,
                             const double *V, double *T);
}

extern "C" int ucy_ensemble_matvec(ucy_ensemble *E, int group, const double *V, double *T)
{
    if (!E || !V || !T || group < 0 || group >= (int)E->hg.size()) return UCY_ERR_ARG;
    if (!E->built) return efail(E, UCY_ERR_STATE, "ensemble not built");
    ucy_ctx *c = E->ctx;
    EHIP(hipSetDevice(c->device));
    const EnsGroup &G = E->tab.g[group];
    double *dV = nullptr, *dT = nullptr;
    EHIP(hipMalloc((void **)&dV, sizeof(double) * (size_t)G.n * G.ninst));
    EHIP(hipMalloc((void **)&dT, sizeof(double) * (size_t)G.m * G.ninst));
    EHIP(hipMemcpyAsync(dV, V, sizeof(double) * (size_t)G.n * G.ninst, hipMemcpyHostToDevice, c->stream));
    ens_launch_matvec_group(c->stream, E->d_tab, E->tab, group, dV, dT);
    EHIP(hipGetLastError());
    EHIP(hipMemcpyAsync(T, dT, sizeof(double) * (size_t)G.m * G.ninst, hipMemcpyDeviceToHost, c->stream));
    EHIP(hipStreamSynchronize(c->stream));
    EHIP(hipFree(dV));
    EHIP(hipFree(dT));
    return UCY_OK;
}
