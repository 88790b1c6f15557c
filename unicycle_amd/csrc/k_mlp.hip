// k_mlp.hip -- training and inference of the step-size proposal network
// (mlp_dev.h) on the GPU: per-sample forward/backward in one wavefront per
// 64 samples, per-block gradient partials reduced deterministically, Adam.
// No counterpart in the reference.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>
#include "mlp_dev.h"
#include "ucy_ctx.h"
#include "../../include/ucy.h"

namespace ucy {

constexpr int kMlpBlock = 64;          // samples per block (one wavefront)
constexpr int kMlpStride = kMlpH + 1;  // LDS row stride (bank spread)

// Mean squared error of the normalised target over the batch idx[0..nb) and
// its gradient: each block writes its partial sums (4801 gradient entries,
// then the loss) to partial[blockIdx.x].
__global__ void __launch_bounds__(kMlpBlock)
mlp_grad_kernel(const float *__restrict__ P, MlpNorm nm, const float *__restrict__ X,
                const float *__restrict__ T, const int *__restrict__ idx, int nb, float inv_n,
                float *__restrict__ partial)
{
    __shared__ float sX[kMlpBlock][kMlpIn];
    __shared__ float sH1[kMlpBlock][kMlpStride];
    __shared__ float sH2[kMlpBlock][kMlpStride];
    __shared__ float sD1[kMlpBlock][kMlpStride];
    __shared__ float sD2[kMlpBlock][kMlpStride];
    __shared__ float sDy[kMlpBlock];
    __shared__ float sLoss[kMlpBlock];
    const int s = threadIdx.x;
    const int g = blockIdx.x * kMlpBlock + s;
    const bool live = g < nb;
    // ---- forward (normalised input and target)
    float x[kMlpIn];
    float tn = 0.0f;
    if (live) {
        const int r = idx ? idx[g] : g;
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) x[i] = (X[(size_t)r * kMlpIn + i] - nm.xmean[i]) / nm.xscale[i];
        tn = (T[r] - nm.ymean) / nm.yscale;
    } else {
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) x[i] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kMlpIn; i++) sX[s][i] = x[i];
    for (int j = 0; j < kMlpH; j++) {
        float a = P[kMlpOffB1 + j];
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) a = fmaf(P[j * kMlpIn + i], x[i], a);
        sH1[s][j] = tanhf(a);
    }
    float y = P[kMlpOffB3];
    for (int k = 0; k < kMlpH; k++) {
        float a = P[kMlpOffB2 + k];
        for (int j = 0; j < kMlpH; j++) a = fmaf(P[kMlpOffW2 + k * kMlpH + j], sH1[s][j], a);
        const float h2 = tanhf(a);
        sH2[s][k] = h2;
        y = fmaf(P[kMlpOffW3 + k], h2, y);
    }
    // ---- backward of (y - t)^2 / n
    const float e = live ? y - tn : 0.0f;
    const float dy = 2.0f * e * inv_n;
    sDy[s] = dy;
    sLoss[s] = e * e * inv_n;
    for (int k = 0; k < kMlpH; k++) {
        const float h2 = sH2[s][k];
        sD2[s][k] = dy * P[kMlpOffW3 + k] * (1.0f - h2 * h2);
    }
    for (int j = 0; j < kMlpH; j++) {
        float a = 0.0f;
        for (int k = 0; k < kMlpH; k++) a = fmaf(P[kMlpOffW2 + k * kMlpH + j], sD2[s][k], a);
        const float h1 = sH1[s][j];
        sD1[s][j] = a * (1.0f - h1 * h1);
    }
    __syncthreads();
    // ---- block sums of the per-sample gradient terms, entry by entry
    float *out = partial + (size_t)blockIdx.x * (kMlpParams + 1);
    for (int p = s; p < kMlpParams; p += kMlpBlock) {
        float acc = 0.0f;
        if (p < kMlpOffB1) {  // W1[j][i]
            const int j = p / kMlpIn, i = p % kMlpIn;
            for (int q = 0; q < kMlpBlock; q++) acc = fmaf(sD1[q][j], sX[q][i], acc);
        } else if (p < kMlpOffW2) {  // b1[j]
            const int j = p - kMlpOffB1;
            for (int q = 0; q < kMlpBlock; q++) acc += sD1[q][j];
        } else if (p < kMlpOffB2) {  // W2[k][j]
            const int k = (p - kMlpOffW2) / kMlpH, j = (p - kMlpOffW2) % kMlpH;
            for (int q = 0; q < kMlpBlock; q++) acc = fmaf(sD2[q][k], sH1[q][j], acc);
        } else if (p < kMlpOffW3) {  // b2[k]
            const int k = p - kMlpOffB2;
            for (int q = 0; q < kMlpBlock; q++) acc += sD2[q][k];
        } else if (p < kMlpOffB3) {  // w3[k]
            const int k = p - kMlpOffW3;
            for (int q = 0; q < kMlpBlock; q++) acc = fmaf(sDy[q], sH2[q][k], acc);
        } else {  // b3
            for (int q = 0; q < kMlpBlock; q++) acc += sDy[q];
        }
        out[p] = acc;
    }
    if (s == 0) {
        float l = 0.0f;
        for (int q = 0; q < kMlpBlock; q++) l += sLoss[q];
        out[kMlpParams] = l;
    }
}

// grad[p] = sum over blocks of partial[b][p] (p = kMlpParams: the loss)
__global__ void mlp_reduce_kernel(const float *__restrict__ partial, int nblk, float *__restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float acc = 0.0f;
    for (int b = 0; b < nblk; b++) acc += partial[(size_t)b * (kMlpParams + 1) + p];
    grad[p] = acc;
}

// Adam (Kingma & Ba 2015) with bias correction; step t >= 1
__global__ void mlp_adam_kernel(float *__restrict__ P, const float *__restrict__ grad,
                                float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// raw features -> raw proposal (log10 step)
__global__ void mlp_forward_kernel(const float *__restrict__ P, MlpNorm nm,
                                   const float *__restrict__ X, int64_t n, float *__restrict__ Y)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    float x[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; i++) x[i] = (X[g * kMlpIn + i] - nm.xmean[i]) / nm.xscale[i];
    Y[g] = mlp_eval(P, x) * nm.yscale + nm.ymean;
}

}  // namespace ucy

using namespace ucy;

struct ucy_mlp {
    ucy_ctx *ctx = nullptr;
    std::vector<float> theta;
    MlpNorm norm{};
    float *d_theta = nullptr;
};

#define MCHK(call)                                                                   \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            M->ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);         \
            return UCY_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

extern "C" ucy_mlp *ucy_mlp_create(ucy_ctx *ctx, uint64_t seed)
{
    if (!ctx) return nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return nullptr;
    ucy_mlp *M = new ucy_mlp();
    M->ctx = ctx;
    M->theta.assign(kMlpParams, 0.0f);
    // Glorot-uniform weights, zero biases
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int count) {
        const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> U(-a, a);
        for (int i = 0; i < count; i++) M->theta[off + i] = U(rng);
    };
    fill(0, kMlpIn, kMlpH, kMlpH * kMlpIn);
    fill(kMlpOffW2, kMlpH, kMlpH, kMlpH * kMlpH);
    fill(kMlpOffW3, kMlpH, 1, kMlpH);
    for (int i = 0; i < kMlpIn; i++) {
        M->norm.xmean[i] = 0.0f;
        M->norm.xscale[i] = 1.0f;
    }
    M->norm.ymean = 0.0f;
    M->norm.yscale = 1.0f;
    if (hipMalloc((void **)&M->d_theta, sizeof(float) * kMlpParams) != hipSuccess ||
        hipMemcpy(M->d_theta, M->theta.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess) {
        if (M->d_theta) (void)hipFree(M->d_theta);
        delete M;
        return nullptr;
    }
    return M;
}

extern "C" void ucy_mlp_destroy(ucy_mlp *M)
{
    if (!M) return;
    (void)hipSetDevice(M->ctx->device);
    if (M->d_theta) (void)hipFree(M->d_theta);
    delete M;
}

extern "C" int ucy_mlp_set_params(ucy_mlp *M, const float *theta)
{
    if (!M || !theta) return UCY_ERR_ARG;
    MCHK(hipSetDevice(M->ctx->device));
    M->theta.assign(theta, theta + kMlpParams);
    MCHK(hipMemcpy(M->d_theta, theta, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    return UCY_OK;
}

extern "C" int ucy_mlp_get_params(ucy_mlp *M, float *theta)
{
    if (!M || !theta) return UCY_ERR_ARG;
    MCHK(hipSetDevice(M->ctx->device));
    MCHK(hipMemcpy(M->theta.data(), M->d_theta, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    std::copy(M->theta.begin(), M->theta.end(), theta);
    return UCY_OK;
}

extern "C" int ucy_mlp_set_normalization(ucy_mlp *M, const float *x_mean, const float *x_scale,
                                         float y_mean, float y_scale)
{
    if (!M || !x_mean || !x_scale || !(y_scale > 0)) return UCY_ERR_ARG;
    for (int i = 0; i < kMlpIn; i++) {
        if (!(x_scale[i] > 0)) return UCY_ERR_ARG;
        M->norm.xmean[i] = x_mean[i];
        M->norm.xscale[i] = x_scale[i];
    }
    M->norm.ymean = y_mean;
    M->norm.yscale = y_scale;
    return UCY_OK;
}

namespace {
struct MlpBuf {
    void *p = nullptr;
    ~MlpBuf() { if (p) (void)hipFree(p); }
};
}  // namespace

extern "C" int ucy_mlp_forward(ucy_mlp *M, int64_t n, const float *x, float *y)
{
    if (!M || n < 0 || (n && (!x || !y))) return UCY_ERR_ARG;
    if (n == 0) return UCY_OK;
    MCHK(hipSetDevice(M->ctx->device));
    MlpBuf dx, dy;
    MCHK(hipMalloc(&dx.p, sizeof(float) * kMlpIn * n));
    MCHK(hipMalloc(&dy.p, sizeof(float) * n));
    MCHK(hipMemcpyAsync(dx.p, x, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, M->ctx->stream));
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       M->ctx->stream, M->d_theta, M->norm, (const float *)dx.p, n, (float *)dy.p);
    MCHK(hipGetLastError());
    MCHK(hipMemcpyAsync(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, M->ctx->stream));
    MCHK(hipStreamSynchronize(M->ctx->stream));
    return UCY_OK;
}

// full-batch loss and gradient (the kernels of one training step)
extern "C" int ucy_mlp_loss_grad(ucy_mlp *M, int64_t n, const float *x, const float *t,
                                 double *loss, float *grad)
{
    if (!M || n <= 0 || n > (int64_t)1 << 30 || !x || !t) return UCY_ERR_ARG;
    MCHK(hipSetDevice(M->ctx->device));
    hipStream_t st = M->ctx->stream;
    const int nblk = (int)((n + kMlpBlock - 1) / kMlpBlock);
    MlpBuf dx, dt, dp, dg;
    MCHK(hipMalloc(&dx.p, sizeof(float) * kMlpIn * n));
    MCHK(hipMalloc(&dt.p, sizeof(float) * n));
    MCHK(hipMalloc(&dp.p, sizeof(float) * (kMlpParams + 1) * nblk));
    MCHK(hipMalloc(&dg.p, sizeof(float) * (kMlpParams + 1)));
    MCHK(hipMemcpyAsync(dx.p, x, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, st));
    MCHK(hipMemcpyAsync(dt.p, t, sizeof(float) * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(nblk), dim3(kMlpBlock), 0, st, M->d_theta, M->norm,
                       (const float *)dx.p, (const float *)dt.p, (const int *)nullptr, (int)n,
                       1.0f / (float)n, (float *)dp.p);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 256) / 256), dim3(256), 0, st,
                       (const float *)dp.p, nblk, (float *)dg.p);
    MCHK(hipGetLastError());
    std::vector<float> h(kMlpParams + 1);
    MCHK(hipMemcpyAsync(h.data(), dg.p, sizeof(float) * (kMlpParams + 1), hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    if (grad) std::copy(h.begin(), h.begin() + kMlpParams, grad);
    if (loss) *loss = h[kMlpParams];
    return UCY_OK;
}

// Adam over shuffled minibatches; loss_log[e] = mean batch loss of epoch e
extern "C" int ucy_mlp_train(ucy_mlp *M, int64_t n, const float *x, const float *t, int epochs,
                             int batch, float lr, uint64_t seed, double *loss_log)
{
    if (!M || n <= 0 || n > (int64_t)1 << 30 || !x || !t || epochs <= 0 || batch <= 0 || !(lr > 0))
        return UCY_ERR_ARG;
    MCHK(hipSetDevice(M->ctx->device));
    hipStream_t st = M->ctx->stream;
    const int nb_max = (int)std::min<int64_t>(batch, n);
    const int nblk_max = (nb_max + kMlpBlock - 1) / kMlpBlock;
    MlpBuf dx, dt, didx, dp, dg, dm, dv;
    const int nbatches = (int)((n + nb_max - 1) / nb_max);
    MCHK(hipMalloc(&dx.p, sizeof(float) * kMlpIn * n));
    MCHK(hipMalloc(&dt.p, sizeof(float) * n));
    MCHK(hipMalloc(&didx.p, sizeof(int) * n));
    MCHK(hipMalloc(&dp.p, sizeof(float) * (kMlpParams + 1) * nblk_max));
    MCHK(hipMalloc(&dg.p, sizeof(float) * (kMlpParams + 1) * (size_t)nbatches));
    MCHK(hipMalloc(&dm.p, sizeof(float) * kMlpParams));
    MCHK(hipMalloc(&dv.p, sizeof(float) * kMlpParams));
    MCHK(hipMemcpyAsync(dx.p, x, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, st));
    MCHK(hipMemcpyAsync(dt.p, t, sizeof(float) * n, hipMemcpyHostToDevice, st));
    std::vector<float> zero(kMlpParams, 0.0f);
    MCHK(hipMemcpyAsync(dm.p, zero.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st));
    MCHK(hipMemcpyAsync(dv.p, zero.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st));
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int64_t step = 0;
    std::vector<float> bl(kMlpParams + 1);
    for (int e = 0; e < epochs; e++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        MCHK(hipMemcpyAsync(didx.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        for (int b = 0; b < nbatches; b++) {
            const int64_t off = (int64_t)b * nb_max;
            const int nb = (int)std::min<int64_t>(nb_max, n - off);
            const int nblk = (nb + kMlpBlock - 1) / kMlpBlock;
            float *g = (float *)dg.p + (size_t)b * (kMlpParams + 1);
            hipLaunchKernelGGL(mlp_grad_kernel, dim3(nblk), dim3(kMlpBlock), 0, st, M->d_theta,
                               M->norm, (const float *)dx.p, (const float *)dt.p,
                               (const int *)didx.p + off, nb, 1.0f / (float)nb, (float *)dp.p);
            hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 256) / 256), dim3(256), 0, st,
                               (const float *)dp.p, nblk, g);
            step++;
            const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
            hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, st,
                               M->d_theta, (const float *)g, (float *)dm.p, (float *)dv.p, lr, b1,
                               b2, eps, c1, c2);
        }
        MCHK(hipGetLastError());
        if (loss_log) {
            double acc = 0.0;
            for (int b = 0; b < nbatches; b++) {
                float l = 0.0f;
                MCHK(hipMemcpyAsync(&l, (float *)dg.p + (size_t)b * (kMlpParams + 1) + kMlpParams,
                                    sizeof(float), hipMemcpyDeviceToHost, st));
                MCHK(hipStreamSynchronize(st));
                acc += l;
            }
            loss_log[e] = acc / nbatches;
        }
    }
    MCHK(hipStreamSynchronize(st));
    MCHK(hipMemcpy(M->theta.data(), M->d_theta, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    return UCY_OK;
}

// accessors for the ensemble (ens_capi.hip)
namespace ucy {
const float *mlp_device_params(const ucy_mlp *M) { return M ? M->d_theta : nullptr; }
MlpNorm mlp_norm(const ucy_mlp *M) { return M->norm; }
}  // namespace ucy
