// mlp_dev.h -- learned RK45 step-size proposal (no counterpart in the
// reference): an 8 -> 64 -> 64 -> 1 perceptron with tanh activations and
// binary32 parameters, evaluated per instance by the ensemble controller and
// trained by the kernels of k_mlp.hip.
//
// Parameter layout (4801 floats):
//   W1[64][8]  (hidden unit j, input i)      offset    0
//   b1[64]                                    offset  512
//   W2[64][64] (unit k of layer 2, unit j)    offset  576
//   b2[64]                                    offset 4672
//   w3[64]                                    offset 4736
//   b3                                        offset 4800
//
// Features at the end of accepted step k, for the proposal of step k+1 (the
// same definitions are computed from the plain integrator's step history by
// unicycle_amd/stepnet.py):
//   f0 log10 vMax_k (vMax of the step-start odefun, floor 1e-30)
//   f1 f0 - log10 vMax_{k-1}                     (0 at the first step)
//   f2 log10 h_k
//   f3 log10 (errmax_k / epsilon)                (floor 1e-10)
//   f4 log10 h_{k-1}                             (f2 at the first step)
//   f5 rejected attempts of step k
//   f6 log10 of the classic proposal (ode45.f90:92-99)
//   f7 f3 of step k-1                            (f3 at the first step)
// The output is the normalised log10 of the proposed step; the controller
// keeps it within a factor 4 of the classic proposal and below 10 h_k.
#pragma once
#include <hip/hip_runtime.h>

namespace ucy {

constexpr int kMlpIn = 8;
constexpr int kMlpH = 64;
constexpr int kMlpOffB1 = kMlpH * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpH;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpH * kMlpH;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpH;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpH;
constexpr int kMlpParams = kMlpOffB3 + 1;  // 4801

struct MlpNorm {
    float xmean[kMlpIn], xscale[kMlpIn];
    float ymean, yscale;
};

// network output for a normalised input (tanh of binary32)
__device__ __forceinline__ float mlp_eval(const float *__restrict__ P, const float x[kMlpIn])
{
    float h1[kMlpH];
#pragma unroll 8
    for (int j = 0; j < kMlpH; j++) {
        float a = P[kMlpOffB1 + j];
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) a = fmaf(P[j * kMlpIn + i], x[i], a);
        h1[j] = tanhf(a);
    }
    float y = P[kMlpOffB3];
    for (int k = 0; k < kMlpH; k++) {
        float a = P[kMlpOffB2 + k];
        const float *w = P + kMlpOffW2 + k * kMlpH;
#pragma unroll 16
        for (int j = 0; j < kMlpH; j++) a = fmaf(w[j], h1[j], a);
        y = fmaf(P[kMlpOffW3 + k], tanhf(a), y);
    }
    return y;
}

// features of the proposal (see the header comment), binary64 in
__device__ __forceinline__ void mlp_features(double vmax, double vmax_prev_log, double h,
                                             double h_prev, double errmax, double lerr_prev,
                                             double rejects, double h_classic, bool first,
                                             float f[kMlpIn])
{
    const double lv = log10(fmax(vmax, 1e-30));
    const double le = log10(fmax(errmax, 1e-10));
    f[0] = (float)lv;
    f[1] = first ? 0.0f : (float)(lv - vmax_prev_log);
    f[2] = (float)log10(h);
    f[3] = (float)le;
    f[4] = first ? f[2] : (float)log10(h_prev);
    f[5] = (float)rejects;
    f[6] = (float)log10(h_classic);
    f[7] = first ? f[3] : (float)lerr_prev;
}

}  // namespace ucy
