// k_ensemble.hip -- batched ensembles: split-bf16 K on MFMA, binary64 state,
// per-instance Cash-Karp RK45 step control.
//
// Per-instance arithmetic is the reference's (ratestate.f90:624-778 odefun,
// ode45.f90:51-202 RKQSm/RKCKm, ratestate.f90:307-328 main loop); only the
// traction matvec changes: for the 64 instance columns of a group,
//     T = K V   with  K = diag(s) (Khi + Klo),  V = Vhi + Vlo  (bf16 pairs)
//     T ~= diag(s) (Khi Vhi + Khi Vlo + Klo Vhi)   accumulated in fp32
// on v_mfma_f32_32x32x16_bf16.  s are exact powers of two (no rounding).
//
// Tiled operand layouts (one 1 KiB tile = one wavefront's 16-byte-per-lane
// load, lane l holds 8 consecutive k of row/column l%32, k-half l/32):
//   K tile (rb, kb, part)   rows 32*rb.., k 16*kb..,  part 0 hi / 1 lo
//       offset  (((rb*nkb + kb)*2 + part)*64 + lane)*8 + j
//   V tile (kb, cb, part)   k 16*kb.., columns 32*cb..
//       offset  (((kb*2 + cb)*2 + part)*64 + lane)*8 + j
// The bf16+i8 K (3 bytes per entry) keeps the hi tile and replaces the lo
// tile by 8-bit codes relative to hi's exponent, decoded to the same bf16 lo
// operand in registers (split3_kernel, decode_lo8):
//   K tile (rb, kb)         1536 bytes at ((rb*nkb + kb)*1536):
//       hi  bytes 16*lane + 2*j,   lo code  byte 1024 + 8*lane + j
#include <hip/hip_runtime.h>
#include <type_traits>
#include "ens_internal.h"

namespace ucy {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__constant__ const double kELg10 = 2.302585092994045684;

// ---------------------------------------------------------------------------
// binary64 -> bf16 with round-to-nearest-even, directly from the binary64
// value (no intermediate binary32 rounding).  Returns the bf16 bits and the
// rounded value as a double.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t bf16_rne(double x, double &xr)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned long long exp = (u >> 52) & 0x7ff;
    if (exp == 0 || exp == 0x7ff) {  // zero / subnormal-binary64 / inf / nan
        xr = exp == 0 ? 0.0 : x;
        const float f = (float)xr;
        return (uint16_t)(__float_as_uint(f) >> 16);
    }
    // 45 dropped bits: add half their range less one, plus the kept lsb
    // (ties to even), and truncate
    const unsigned long long lsb = (u >> 45) & 1ull;
    u += 0x00000FFFFFFFFFFFull + lsb;
    u &= ~0x00001FFFFFFFFFFFull;
    xr = __longlong_as_double((long long)u);
    const float f = (float)xr;  // exact: 8 significant bits, binary32 range
    return (uint16_t)(__float_as_uint(f) >> 16);
}

__device__ __forceinline__ void split2(double x, uint16_t &hi, uint16_t &lo)
{
    double h, l;
    hi = bf16_rne(x, h);
    lo = bf16_rne(x - h, l);
}

__device__ __forceinline__ size_t ktile_off(int r, int c, int nkb, int part)
{
    const int rb = r >> 5, kb = c >> 4, w = c & 15;
    const int lane = (r & 31) + 32 * (w >> 3), j = w & 7;
    return ((((size_t)rb * nkb + kb) * 2 + part) * 64 + lane) * 8 + j;
}

__device__ __forceinline__ size_t vtile_off(int c, int col, int part)
{
    const int kb = c >> 4, w = c & 15, cb = col >> 5;
    const int lane = (col & 31) + 32 * (w >> 3), j = w & 7;
    return ((((size_t)kb * 2 + cb) * 2 + part) * 64 + lane) * 8 + j;
}

// binary64 tiles for v_mfma_f64_16x16x4f64: one 512-byte tile = 16 rows (or
// columns) x 4 k, lane l holds row/column l%16, k l/16
//   K tile (rt, kt)  offset (rt*nkt + kt)*64 + lane
//   V tile (kt, ct)  offset (kt*4 + ct)*64 + lane
__device__ __forceinline__ size_t k64_off(int r, int c, int nkt)
{
    return ((size_t)(r >> 4) * nkt + (c >> 2)) * 64 + (r & 15) + 16 * (c & 3);
}
__device__ __forceinline__ size_t v64_off(int c, int col)
{
    return ((size_t)(c >> 2) * 4 + (col >> 4)) * 64 + (col & 15) + 16 * (c & 3);
}

__global__ void tile64_kernel(const double *__restrict__ K, int m, int n, int mp, int np,
                              double *__restrict__ tiles)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)mp * np) return;
    const int r = (int)(i / np), c = (int)(i % np);
    tiles[k64_off(r, c, np >> 2)] = (r < m && c < n) ? K[(size_t)r * n + c] : 0.0;
}

void ens_launch_tile_K64(hipStream_t st, const double *K, int m, int n, int mp, int np,
                         double *tiles)
{
    const long long tot = (long long)mp * np;
    hipLaunchKernelGGL(tile64_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, K, m,
                       n, mp, np, tiles);
}

// ---------------------------------------------------------------------------
// K -> split bf16 tiles with exact power-of-two row scales
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) rowscale_kernel(const double *__restrict__ K, int m, int n,
                                                       int mp, double *rscale)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= mp) return;
    double mx = 0.0;
    if (r < m)
        for (int c = lane; c < n; c += 64) mx = fmax(mx, fabs(K[(size_t)r * n + c]));
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if (lane == 0) {
        double s = 1.0;
        if (mx > 0.0) {
            int e;
            frexp(mx, &e);          // mx = f * 2^e, f in [0.5, 1)
            s = ldexp(1.0, e);      // mx / s in [0.5, 1)
        }
        rscale[r] = s;
    }
}

__global__ void split_kernel(const double *__restrict__ K, int m, int n, int mp, int np,
                             const double *__restrict__ rscale, uint16_t *__restrict__ tiles)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)mp * np) return;
    const int r = (int)(i / np), c = (int)(i % np);
    const double x = (r < m && c < n) ? K[(size_t)r * n + c] / rscale[r] : 0.0;
    uint16_t hi, lo;
    split2(x, hi, lo);
    const int nkb = np >> 4;
    tiles[ktile_off(r, c, nkb, 0)] = hi;
    tiles[ktile_off(r, c, nkb, 1)] = lo;
}

void ens_launch_split_K(hipStream_t st, const double *K, int m, int n, int mp, int np,
                        uint16_t *tiles, double *rscale)
{
    hipLaunchKernelGGL(rowscale_kernel, dim3((mp + 3) / 4), dim3(256), 0, st, K, m, n, mp, rscale);
    const long long tot = (long long)mp * np;
    hipLaunchKernelGGL(split_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, K, m, n,
                       mp, np, rscale, tiles);
}

// K -> bf16+i8 tiles with the same exact row scales.  For x = K / s, hi =
// bf16(x) and the residual r = x - hi (exact in binary64) is coded in steps
// of 2^(E-142), E the biased exponent of hi, i.e. 1/128 of hi's ulp:
// q = rint(r / 2^(E-142)) lies in [-128, 128] (|r| <= ulp/2), and the byte
// stored is -q (the decoder's scale carries a minus sign).  Where q would
// be -128, hi moves one ulp down and q is taken again (then in [64, 128]),
// so every entry keeps |x - (hi + q 2^(E-142))| <= 2^(E-143), 2^-16
// (1 + 2^-8) relative.  Entries with E < 32 (|x| < 2^-95, |x| <= 1 by the
// row scale) are flushed to zero: their products lie far below the fp32
// accumulation's resolution.
__global__ void split3_kernel(const double *__restrict__ K, int m, int n, int mp, int np,
                              const double *__restrict__ rscale, uint8_t *__restrict__ tiles)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)mp * np) return;
    const int r = (int)(i / np), c = (int)(i % np);
    const double x = (r < m && c < n) ? K[(size_t)r * n + c] / rscale[r] : 0.0;
    double h;
    uint16_t hi = bf16_rne(x, h);
    int E = (hi >> 7) & 0xFF;
    int q = 0;
    if (E < 32) {
        hi = 0;
    } else {
        double t = rint(ldexp(x - h, 142 - E));
        if (t < -127.0) {
            hi = bf16_rne(h - ldexp(1.0, E - 134), h);   // exact: the next bf16 down
            E = (hi >> 7) & 0xFF;
            t = rint(ldexp(x - h, 142 - E));
        }
        q = -(int)fmin(fmax(t, -127.0), 128.0);
    }
    const int nkb = np >> 4;
    const int rb = r >> 5, kb = c >> 4, w = c & 15;
    const int lane = (r & 31) + 32 * (w >> 3), j = w & 7;
    uint8_t *tile = tiles + ((size_t)rb * nkb + kb) * 1536;
    reinterpret_cast<uint16_t *>(tile)[lane * 8 + j] = hi;
    tile[1024 + lane * 8 + j] = (uint8_t)(int8_t)q;
}

void ens_launch_split3_K(hipStream_t st, const double *K, int m, int n, int mp, int np,
                         uint8_t *tiles, double *rscale)
{
    hipLaunchKernelGGL(rowscale_kernel, dim3((mp + 3) / 4), dim3(256), 0, st, K, m, n, mp, rscale);
    const long long tot = (long long)mp * np;
    hipLaunchKernelGGL(split3_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, K, m,
                       n, mp, np, rscale, tiles);
}

// bf16 lo operand of 8 bf16+i8 entries from their hi words (2 bf16 each, low
// half first) and lo bytes c_j = -q_j: lo_j = c_j (-2^(E_j - 142)).  Both
// halves' scales come from one mask and one add: per 16-bit half, E << 7
// plus 0x7880 is (E - 15) << 7 with the sign bit set for E >= 15, i.e. the
// top half of the binary32 -2^(E-142) (no carry between the halves: E <=
// 127 keeps each half below 0xB801); for E = 0 (a zero hi, code 0) it is
// +2^114, finite.  The packer leaves no 0 < E < 32.  c_j times the scale has
// at most 8 significant bits, so its top 16 bits are the exact bf16.
// (Scaling hi by 2^15 instead would save the adjust but keeps a second copy
// of hi live: past 256 VGPRs, half the waves.)
__device__ __forceinline__ bf16x8 decode_lo8(const u32x4 &h, const u32x2 &q)
{
    u32x4 out;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const unsigned M = (h[i] & 0x7F807F80u) + 0x78807880u;
        const unsigned Q = q[i >> 1] >> (16 * (i & 1));
        const float s0 = __uint_as_float(M << 16);
        const float s1 = __uint_as_float(M & 0xFFFF0000u);
        const float l0 = (float)(int)(signed char)(Q & 0xFFu) * s0;
        const float l1 = (float)(int)(signed char)((Q >> 8) & 0xFFu) * s1;
        // bytes 2, 3 of l0 then bytes 2, 3 of l1
        out[i] = __builtin_amdgcn_perm(__float_as_uint(l1), __float_as_uint(l0), 0x07060302u);
    }
    return __builtin_bit_cast(bf16x8, out);
}

// bytes of one K tile (rb, kb): split bf16 hi + lo, or bf16+i8
template <bool I8>
constexpr size_t kTileBytes = I8 ? 1536 : 2048;

// single-instance K -> packed split-bf16 pairs, row-major with the row
// stride of K: word = hi << 16 | lo of K / rs[row]
__global__ void __launch_bounds__(256) split_rowmajor_kernel(const double *__restrict__ K,
                                                             int64_t m, int64_t n,
                                                             const double *__restrict__ rs,
                                                             uint32_t *__restrict__ Kp)
{
    const int64_t tot = m * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint16_t hi, lo;
        split2(K[i] / rs[i / n], hi, lo);
        Kp[i] = ((uint32_t)hi << 16) | lo;
    }
}

void launch_split_rowmajor(hipStream_t st, const double *K, int64_t m, int64_t n, double *rs,
                           uint32_t *Kp)
{
    hipLaunchKernelGGL(rowscale_kernel, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, st, K,
                       (int)m, (int)n, (int)m, rs);
    hipLaunchKernelGGL(split_rowmajor_kernel, dim3(256 * 64), dim3(256), 0, st, K, m, n, rs, Kp);
}

// ---------------------------------------------------------------------------
// group lookup from a block index and a prefix table
// ---------------------------------------------------------------------------
__device__ __forceinline__ int group_of(const int *prefix, int ng, int blk)
{
    int g = 0;
    while (g + 1 < ng && blk >= prefix[g + 1]) g++;
    return g;
}

// ---------------------------------------------------------------------------
// split-bf16 GEMM: one wavefront = RB 32-row blocks x 64 columns over one
// k-split.  RB = 4 halves the V-tile reads per K byte against RB = 2 at
// twice the accumulators; the autotuner picks.
// ---------------------------------------------------------------------------
// K loads: non-temporal when the wave is the only reader of its K tiles
// (NT); plain when the slices of a K family read the same tiles in the
// waves of one block, so that the first read serves the others from cache
// Loads through the global address space: a load through a generic (flat)
// pointer may reach LDS, its counter then retires out of order, and the
// compiler can only wait for all outstanding loads at once, which drains
// every load issued ahead.
template <typename T>
using gptr = __attribute__((address_space(1))) const T *;

template <typename T>
__device__ __forceinline__ gptr<T> as_global(const T *p)
{
    return (gptr<T>)p;
}

template <bool NT, typename T>
__device__ __forceinline__ T kload(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(as_global(p));
    else return *as_global(p);
}

template <int RB, bool PIPE, bool I8, bool NT>
__device__ __forceinline__ void gemm_tile_bf16(const EnsGroup &G, int rb0, int sp, int lane)
{
    const int kb0 = sp * G.kb_per_split;
    int kb1 = kb0 + G.kb_per_split;
    if (kb1 > G.nkb) kb1 = G.nkb;
    floatx16 acc[RB][2];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[a][b][i] = 0.0f;
    // tile (rb, kb): hi then lo, 64 lanes x 16 bytes each.  Pipelining is
    // by hand: loops holding MFMA builtins are not runtime-unrolled by the
    // compiler.
    const char *A = reinterpret_cast<const char *>(G.Ktiles) + (size_t)rb0 * G.nkb * kTileBytes<I8>;
    const size_t rstride = (size_t)G.nkb * kTileBytes<I8>;
    const u32x4 *Bv = reinterpret_cast<const u32x4 *>(G.Vtiles) + lane;
    // per 32-row block: hi, and the bf16 lo (split) or the lo codes (I8)
    u32x4 ca[RB][2], cb[4], na[RB][2], nb[4];
    u32x2 cq[RB], nq[RB];
    auto load = [&](int kb, u32x4 (&ra)[RB][2], u32x2 (&rq)[RB], u32x4 (&rbv)[4]) {
#pragma unroll
        for (int a = 0; a < RB; a++) {
            const char *t = A + a * rstride + (size_t)kb * kTileBytes<I8>;
            ra[a][0] = kload<NT>(reinterpret_cast<const u32x4 *>(t) + lane);
            if constexpr (I8)
                rq[a] = kload<NT>(reinterpret_cast<const u32x2 *>(t + 1024) + lane);
            else
                ra[a][1] = kload<NT>(reinterpret_cast<const u32x4 *>(t) + 64 + lane);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) rbv[q] = *as_global(Bv + (size_t)kb * 256 + 64 * q);
    };
    if (kb0 < kb1) load(kb0, ca, cq, cb);
    for (int kb = kb0; kb < kb1; kb++) {
        // PIPE: the next k-block's loads go out before this one's MFMAs
        // (more bytes in flight per wave, more registers: fewer waves)
        if (!PIPE && kb > kb0) load(kb, ca, cq, cb);
        if (PIPE && kb + 1 < kb1) load(kb + 1, na, nq, nb);
        const bf16x8 B[2][2] = {{__builtin_bit_cast(bf16x8, cb[0]), __builtin_bit_cast(bf16x8, cb[1])},
                                {__builtin_bit_cast(bf16x8, cb[2]), __builtin_bit_cast(bf16x8, cb[3])}};
#pragma unroll
        for (int a = 0; a < RB; a++) {
            const bf16x8 Ah = __builtin_bit_cast(bf16x8, ca[a][0]);
            bf16x8 Al;
            if constexpr (I8)
                Al = decode_lo8(ca[a][0], cq[a]);
            else
                Al = __builtin_bit_cast(bf16x8, ca[a][1]);
#pragma unroll
            for (int b = 0; b < 2; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][1], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, B[b][0], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][0], acc[a][b], 0, 0, 0);
            }
        }
        if (PIPE) {
#pragma unroll
            for (int a = 0; a < RB; a++) {
                ca[a][0] = na[a][0];
                if constexpr (I8)
                    cq[a] = nq[a];
                else
                    ca[a][1] = na[a][1];
            }
#pragma unroll
            for (int q = 0; q < 4; q++) cb[q] = nb[q];
        }
    }
    // 32x32 fp32 accumulator layout: value i of lane l is row 8*(i/4) + 4*(l/32) + i%4,
    // column l%32 of the tile
    float *T = G.Tpart + (size_t)sp * G.mp * kEnsCols;
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int row = 32 * (rb0 + a) + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const int col = 32 * b + (lane & 31);
                T[(size_t)row * kEnsCols + col] = acc[a][b][i];
            }
}

// The same tile with two register sets used in turn (variant 9): the loop
// takes k-blocks in pairs, loads the odd one into set 1 before multiplying
// set 0 and the next even one into set 0 before multiplying set 1, so every
// wait covers only the older set (a copy from a "next" set to a "current"
// one would wait for the loads just issued).  Loads are unconditional: a
// k-block past the range loads the last one again and multiplies a zero V.
template <int RB, bool I8, bool NT>
__device__ __forceinline__ void gemm_tile_bf16_2set(const EnsGroup &G, int rb0, int sp, int lane)
{
    const int kb0 = sp * G.kb_per_split;
    const int kb1 = min(kb0 + G.kb_per_split, G.nkb);
    floatx16 acc[RB][2];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[a][b][i] = 0.0f;
    const char *A = reinterpret_cast<const char *>(G.Ktiles) + (size_t)rb0 * G.nkb * kTileBytes<I8>;
    const size_t rstride = (size_t)G.nkb * kTileBytes<I8>;
    const u32x4 *Bv = reinterpret_cast<const u32x4 *>(G.Vtiles) + lane;
    u32x4 ca[2][RB][2], cb[2][4];
    u32x2 cq[2][RB];
    auto load = [&](int kb, auto set_c) {
        constexpr int s = decltype(set_c)::value;
        kb = min(kb, kb1 - 1);
#pragma unroll
        for (int a = 0; a < RB; a++) {
            const char *t = A + a * rstride + (size_t)kb * kTileBytes<I8>;
            ca[s][a][0] = kload<NT>(reinterpret_cast<const u32x4 *>(t) + lane);
            if constexpr (I8)
                cq[s][a] = kload<NT>(reinterpret_cast<const u32x2 *>(t + 1024) + lane);
            else
                ca[s][a][1] = kload<NT>(reinterpret_cast<const u32x4 *>(t) + 64 + lane);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) cb[s][q] = *as_global(Bv + (size_t)kb * 256 + 64 * q);
    };
    auto mult = [&](bool live, auto set_c) {
        constexpr int s = decltype(set_c)::value;
        const u32x4 zero = {0u, 0u, 0u, 0u};
        u32x4 v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = live ? cb[s][q] : zero;
        const bf16x8 B[2][2] = {{__builtin_bit_cast(bf16x8, v[0]), __builtin_bit_cast(bf16x8, v[1])},
                                {__builtin_bit_cast(bf16x8, v[2]), __builtin_bit_cast(bf16x8, v[3])}};
#pragma unroll
        for (int a = 0; a < RB; a++) {
            const bf16x8 Ah = __builtin_bit_cast(bf16x8, ca[s][a][0]);
            bf16x8 Al;
            if constexpr (I8)
                Al = decode_lo8(ca[s][a][0], cq[s][a]);
            else
                Al = __builtin_bit_cast(bf16x8, ca[s][a][1]);
#pragma unroll
            for (int b = 0; b < 2; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][1], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, B[b][0], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][0], acc[a][b], 0, 0, 0);
            }
        }
    };
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    // the scheduler keeps the four phases in order (it would otherwise
    // hoist both sets' loads and interleave the sets' products, so that
    // every iteration waits for all its loads)
    if (kb0 < kb1) {
        load(kb0, S0{});
        for (int kb = kb0; kb < kb1; kb += 2) {
            load(kb + 1, S1{});
            __builtin_amdgcn_sched_barrier(0);
            mult(true, S0{});
            __builtin_amdgcn_sched_barrier(0);
            load(kb + 2, S0{});
            __builtin_amdgcn_sched_barrier(0);
            mult(kb + 1 < kb1, S1{});
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    float *T = G.Tpart + (size_t)sp * G.mp * kEnsCols;
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int row = 32 * (rb0 + a) + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const int col = 32 * b + (lane & 31);
                T[(size_t)row * kEnsCols + col] = acc[a][b][i];
            }
}

// row items of RB 32-row blocks: mp is a multiple of 64, so the last item of
// RB = 4 may hold two blocks
__host__ __device__ __forceinline__ int bf16_row_items(int mp, int rb)
{
    return ((mp >> 5) + rb - 1) / rb;
}

// ---------------------------------------------------------------------------
// Variants 6 and 7: every load of the k-loop is a direct global->LDS load
// (global_load_lds; the 16- and 4-byte forms land lane l's bytes at 16 l and
// 4 l, which keeps a tile's layout; the 12-byte form lands lane l at 16 l as
// well, so it would not).  The 4 waves of a block take 4 consecutive row items of
// one k-split (as the V-staging variants); each wave streams its own K tiles
// into a private LDS ring and one quarter of the shared V tile into a
// block-wide ring, D k-blocks deep.  Per k-block: wait for this wave's loads
// of the k-block (the counter retires in issue order), a barrier without a
// memory fence (every quarter of V has landed, and every wave has finished
// reading the slot about to be refilled), issue the loads D - 1 k-blocks
// ahead, read the slot from LDS, multiply.  The LDS reads are one asm block
// that also waits for its own results: the compiler does not order a plain
// LDS read after the DMA that fills it, and would otherwise wait for every
// outstanding load (the whole ring) before it.
// ---------------------------------------------------------------------------

// f(std::integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time
template <int N, int I = 0, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

template <int N>
__device__ __forceinline__ void wait_vm()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// One ring slot: RB bf16+i8 K tiles (hi 16 bytes and codes 8 bytes per lane)
// from the wave's ring at kh = kring + 16 lane, kq = kring + 1024 + 8 lane,
// and the 4 x 16 bytes of V per lane from vv = vring + 16 lane; KO and VO
// are the slot's byte offsets
template <int RB, int KO, int VO>
__device__ __forceinline__ void lds_read_slot(unsigned kh, unsigned kq, unsigned vv,
                                              u32x4 (&h)[RB], u32x2 (&q)[RB], u32x4 (&v)[4])
{
    static_assert(KO + (RB - 1) * 1536 + 1024 < 65536 && VO + 3072 < 65536, "ds offsets");
    if constexpr (RB == 4)
        asm volatile(
            "ds_read_b128 %0, %12 offset:%15\n"
            "ds_read_b128 %1, %12 offset:%16\n"
            "ds_read_b128 %2, %12 offset:%17\n"
            "ds_read_b128 %3, %12 offset:%18\n"
            "ds_read_b64 %4, %13 offset:%15\n"
            "ds_read_b64 %5, %13 offset:%16\n"
            "ds_read_b64 %6, %13 offset:%17\n"
            "ds_read_b64 %7, %13 offset:%18\n"
            "ds_read_b128 %8, %14 offset:%19\n"
            "ds_read_b128 %9, %14 offset:%20\n"
            "ds_read_b128 %10, %14 offset:%21\n"
            "ds_read_b128 %11, %14 offset:%22\n"
            "s_waitcnt lgkmcnt(0)"
            : "=v"(h[0]), "=v"(h[1]), "=v"(h[2]), "=v"(h[3]), "=v"(q[0]), "=v"(q[1]), "=v"(q[2]),
              "=v"(q[3]), "=v"(v[0]), "=v"(v[1]), "=v"(v[2]), "=v"(v[3])
            : "v"(kh), "v"(kq), "v"(vv), "n"(KO), "n"(KO + 1536), "n"(KO + 3072),
              "n"(KO + 4608), "n"(VO), "n"(VO + 1024), "n"(VO + 2048), "n"(VO + 3072)
            : "memory");
    else
        asm volatile(
            "ds_read_b128 %0, %8 offset:%11\n"
            "ds_read_b128 %1, %8 offset:%12\n"
            "ds_read_b64 %2, %9 offset:%11\n"
            "ds_read_b64 %3, %9 offset:%12\n"
            "ds_read_b128 %4, %10 offset:%13\n"
            "ds_read_b128 %5, %10 offset:%14\n"
            "ds_read_b128 %6, %10 offset:%15\n"
            "ds_read_b128 %7, %10 offset:%16\n"
            "s_waitcnt lgkmcnt(0)"
            : "=v"(h[0]), "=v"(h[1]), "=v"(q[0]), "=v"(q[1]), "=v"(v[0]), "=v"(v[1]), "=v"(v[2]),
              "=v"(v[3])
            : "v"(kh), "v"(kq), "v"(vv), "n"(KO), "n"(KO + 1536), "n"(VO), "n"(VO + 1024),
              "n"(VO + 2048), "n"(VO + 3072)
            : "memory");
}

template <int RB, int D>
__device__ __forceinline__ void gemm_block_bf16_dma(const EnsGroup &G, int ri, int sp, int wave,
                                                    int lane, char *kring, char *vring)
{
    constexpr int TB = 1536;
    constexpr int PER = 3 * RB + 1;   // loads per k-block and wave: 3 per K tile, 1 V quarter
    static_assert(D >= 2 && (D - 2) * PER < 64, "ring depth");
    const int nrb = G.mp >> 5;
    const int nri = bf16_row_items(G.mp, RB);
    const bool valid = ri < nri;
    const int rb0 = (valid ? ri : 0) * RB;
    const int nvr = valid ? min(RB, nrb - rb0) : 0;   // RB, or 2 on the tail item
    const int kb0 = sp * G.kb_per_split;
    const int nk = min(kb0 + G.kb_per_split, G.nkb) - kb0;
    floatx16 acc[RB][2];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[a][b][i] = 0.0f;
    const char *Kt = reinterpret_cast<const char *>(G.Ktiles);
    const char *Vt = reinterpret_cast<const char *>(G.Vtiles);
    typedef __attribute__((address_space(3))) void lds_void;
    // the loads of k-block kb0 + j (clamped to the last one; waves past the
    // last row item and rows past the last block load valid tiles they do
    // not use, so every wave issues PER loads per k-block) into ring slot `slot`
    auto issue = [&](int j, int slot) {
        const int kb = kb0 + min(j, nk - 1);
#pragma unroll
        for (int a = 0; a < RB; a++) {
            const char *t = Kt + ((size_t)min(rb0 + a, nrb - 1) * G.nkb + kb) * TB;
            char *l = kring + (slot * RB + a) * TB;
            // the tile's 1024 bytes of hi (64 lanes x 16), then its 512 bytes
            // of codes as two loads of 64 lanes x 4.  All lanes take part in
            // every DMA: the LDS base is a wave-uniform register (M0), and a
            // load under a lane mask has been merged by the compiler with a
            // neighbouring one of another base, the mask's lanes then
            // writing at the wrong base
            __builtin_amdgcn_global_load_lds((const void *)(t + 16 * lane), (lds_void *)l, 16, 0,
                                             0);
            __builtin_amdgcn_global_load_lds((const void *)(t + 1024 + 4 * lane),
                                             (lds_void *)(l + 1024), 4, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(t + 1280 + 4 * lane),
                                             (lds_void *)(l + 1280), 4, 0, 0);
        }
        __builtin_amdgcn_global_load_lds((const void *)(Vt + (size_t)kb * 4096 + wave * 1024 + lane * 16),
                                         (lds_void *)(vring + slot * 4096 + wave * 1024), 16, 0, 0);
    };
    const unsigned kl = (unsigned)(uintptr_t)(__attribute__((address_space(3))) char *)kring;
    const unsigned vl = (unsigned)(uintptr_t)(__attribute__((address_space(3))) char *)vring;
    const unsigned kh = kl + 16 * lane, kq = kl + 1024 + 8 * lane, vv = vl + 16 * lane;
    auto step = [&](int j, auto slot_c) {
        constexpr int slot = decltype(slot_c)::value;
        wait_vm<(D - 2) * PER>();          // this wave's loads of k-block j
        __builtin_amdgcn_s_barrier();       // all quarters of V j; slot j - 1 read by all
        issue(j + D - 1, (slot + D - 1) % D);
        if (!valid) return;
        u32x4 hh[RB], vv4[4];
        u32x2 qq[RB];
        lds_read_slot<RB, slot * RB * TB, slot * 4096>(kh, kq, vv, hh, qq, vv4);
        // past the range V counts as 0: the products are exact zeros
        const bool live = j < nk;
        const u32x4 zero = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int q = 0; q < 4; q++) vv4[q] = live ? vv4[q] : zero;
        const bf16x8 B[2][2] = {{__builtin_bit_cast(bf16x8, vv4[0]), __builtin_bit_cast(bf16x8, vv4[1])},
                                {__builtin_bit_cast(bf16x8, vv4[2]), __builtin_bit_cast(bf16x8, vv4[3])}};
#pragma unroll
        for (int a = 0; a < RB; a++) {
            if (a >= nvr) continue;
            const bf16x8 Ah = __builtin_bit_cast(bf16x8, hh[a]);
            const bf16x8 Al = decode_lo8(hh[a], qq[a]);
#pragma unroll
            for (int b = 0; b < 2; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][1], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, B[b][0], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][0], acc[a][b], 0, 0, 0);
            }
        }
    };
    if (nk > 0) {
#pragma unroll
        for (int s = 0; s < D - 1; s++) issue(s, s);
        for (int j = 0; j < nk; j += D)
            static_for<D>([&](auto s_c) { step(j + decltype(s_c)::value, s_c); });
        wait_vm<0>();   // no LDS write may land after the block has left
    }
    if (!valid) return;
    float *T = G.Tpart + (size_t)sp * G.mp * kEnsCols;
#pragma unroll
    for (int a = 0; a < RB; a++) {
        if (a >= nvr) continue;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int row = 32 * (rb0 + a) + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const int col = 32 * b + (lane & 31);
                T[(size_t)row * kEnsCols + col] = acc[a][b][i];
            }
    }
}

// The K tiles of one ring slot (no V): RB x (hi 16 bytes, codes 8 bytes) per
// lane, waiting for its own reads as lds_read_slot does
template <int RB, int KO>
__device__ __forceinline__ void lds_read_k(unsigned kh, unsigned kq, u32x4 (&h)[RB], u32x2 (&q)[RB])
{
    static_assert(KO + (RB - 1) * 1536 + 1024 < 65536, "ds offsets");
    if constexpr (RB == 4)
        asm volatile(
            "ds_read_b128 %0, %8 offset:%10\n"
            "ds_read_b128 %1, %8 offset:%11\n"
            "ds_read_b128 %2, %8 offset:%12\n"
            "ds_read_b128 %3, %8 offset:%13\n"
            "ds_read_b64 %4, %9 offset:%10\n"
            "ds_read_b64 %5, %9 offset:%11\n"
            "ds_read_b64 %6, %9 offset:%12\n"
            "ds_read_b64 %7, %9 offset:%13\n"
            "s_waitcnt lgkmcnt(0)"
            : "=v"(h[0]), "=v"(h[1]), "=v"(h[2]), "=v"(h[3]), "=v"(q[0]), "=v"(q[1]), "=v"(q[2]),
              "=v"(q[3])
            : "v"(kh), "v"(kq), "n"(KO), "n"(KO + 1536), "n"(KO + 3072), "n"(KO + 4608)
            : "memory");
    else
        asm volatile(
            "ds_read_b128 %0, %4 offset:%6\n"
            "ds_read_b128 %1, %4 offset:%7\n"
            "ds_read_b64 %2, %5 offset:%6\n"
            "ds_read_b64 %3, %5 offset:%7\n"
            "s_waitcnt lgkmcnt(0)"
            : "=v"(h[0]), "=v"(h[1]), "=v"(q[0]), "=v"(q[1])
            : "v"(kh), "v"(kq), "n"(KO), "n"(KO + 1536)
            : "memory");
}

// Variants 8 and 9: the wave-shaped work items of variants 0-5 (one row
// item, k-split and slice per wave, no barrier), with the wave's K tiles
// streamed D k-blocks ahead by direct global->LDS loads into a ring private
// to the wave, and the V tile of the same k-blocks loaded into a register
// ring of the same depth.  K in flight per wave is (D - 1) k-blocks without
// their registers.  Per k-block: wait for this wave's loads of the k-block
// (retired in issue order: the loads of the D - 2 later k-blocks may still
// be out), issue the k-block D - 1 ahead into the slot read at the previous
// step (its reads waited for themselves), read the slot, multiply.  NT: the
// K loads are non-temporal (one slice per family; slices of a shared K find
// its tiles in L2).
template <int RB, int D, bool NT>
__device__ __forceinline__ void gemm_wave_bf16_dmav(const EnsGroup &G, int item, int lane,
                                                    char *kring)
{
    constexpr int TB = 1536;
    constexpr int PER = 3 * RB + 4;   // loads per k-block: 3 per K tile, 4 of V
    static_assert(D >= 2 && (D - 2) * PER < 64, "ring depth");
    const int nrb = G.mp >> 5;
    const int nri = bf16_row_items(G.mp, RB);
    const int rb0 = (item % nri) * RB, sp = item / nri;
    const int nvr = min(RB, nrb - rb0);   // RB, or 2 on the tail item
    const int kb0 = sp * G.kb_per_split;
    const int nk = min(kb0 + G.kb_per_split, G.nkb) - kb0;
    floatx16 acc[RB][2];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[a][b][i] = 0.0f;
    const char *Kt = reinterpret_cast<const char *>(G.Ktiles);
    const u32x4 *Bv = reinterpret_cast<const u32x4 *>(G.Vtiles) + lane;
    typedef __attribute__((address_space(3))) void lds_void;
    u32x4 vr[D][4];
    // the loads of k-block kb0 + j (clamped to the last one; rows past the
    // last block load a valid tile they do not use, so every k-block is PER
    // loads) into slot `slot`
    auto issue = [&](int j, auto slot_c) {
        constexpr int slot = decltype(slot_c)::value;
        const int kb = kb0 + min(j, nk - 1);
#pragma unroll
        for (int a = 0; a < RB; a++) {
            const char *t = Kt + ((size_t)min(rb0 + a, nrb - 1) * G.nkb + kb) * TB;
            char *l = kring + (slot * RB + a) * TB;
            // full-wave loads only (the LDS base is wave-uniform); cache
            // policy 2 = nt
            __builtin_amdgcn_global_load_lds((const void *)(t + 16 * lane), (lds_void *)l, 16, 0,
                                             NT ? 2 : 0);
            __builtin_amdgcn_global_load_lds((const void *)(t + 1024 + 4 * lane),
                                             (lds_void *)(l + 1024), 4, 0, NT ? 2 : 0);
            __builtin_amdgcn_global_load_lds((const void *)(t + 1280 + 4 * lane),
                                             (lds_void *)(l + 1280), 4, 0, NT ? 2 : 0);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) vr[slot][q] = *as_global(Bv + (size_t)kb * 256 + 64 * q);
    };
    const unsigned kl = (unsigned)(uintptr_t)(__attribute__((address_space(3))) char *)kring;
    const unsigned kh = kl + 16 * lane, kq = kl + 1024 + 8 * lane;
    auto step = [&](int j, auto slot_c) {
        constexpr int slot = decltype(slot_c)::value;
        wait_vm<(D - 2) * PER>();          // this wave's loads of k-block j
        issue(j + D - 1, std::integral_constant<int, (slot + D - 1) % D>{});
        u32x4 hh[RB];
        u32x2 qq[RB];
        lds_read_k<RB, slot * RB * TB>(kh, kq, hh, qq);
        // past the range V counts as 0: the products are exact zeros
        const bool live = j < nk;
        const u32x4 zero = {0u, 0u, 0u, 0u};
        u32x4 v4[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v4[q] = live ? vr[slot][q] : zero;
        const bf16x8 B[2][2] = {{__builtin_bit_cast(bf16x8, v4[0]), __builtin_bit_cast(bf16x8, v4[1])},
                                {__builtin_bit_cast(bf16x8, v4[2]), __builtin_bit_cast(bf16x8, v4[3])}};
#pragma unroll
        for (int a = 0; a < RB; a++) {
            if (a >= nvr) continue;
            const bf16x8 Ah = __builtin_bit_cast(bf16x8, hh[a]);
            const bf16x8 Al = decode_lo8(hh[a], qq[a]);
#pragma unroll
            for (int b = 0; b < 2; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][1], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, B[b][0], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][0], acc[a][b], 0, 0, 0);
            }
        }
    };
    if (nk > 0) {
        static_for<D - 1>([&](auto s_c) { issue(decltype(s_c)::value, s_c); });
        for (int j = 0; j < nk; j += D)
            static_for<D>([&](auto s_c) { step(j + decltype(s_c)::value, s_c); });
        wait_vm<0>();   // no LDS write may land after the wave has left
    }
    float *T = G.Tpart + (size_t)sp * G.mp * kEnsCols;
#pragma unroll
    for (int a = 0; a < RB; a++) {
        if (a >= nvr) continue;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int row = 32 * (rb0 + a) + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const int col = 32 * b + (lane & 31);
                T[(size_t)row * kEnsCols + col] = acc[a][b][i];
            }
    }
}


// Variant 10, for K families of several slices (more than 64 instances of
// one geometry): the 4 waves of a block take one row item (4 32-row blocks)
// and one k-split, wave w multiplying slice w.  Per k-block, wave w loads
// 32-row K tile w, decodes it (the lo codes of bf16+i8) and writes its hi
// and lo operands to LDS; after a barrier every wave multiplies all four
// tiles with its own slice's V.  The K tile leaves memory and is decoded once
// per block instead of once per slice: with 4 slices the GEMM carries 4x the
// products per K byte and is bound by the matrix cores, whose issue the
// other variants share with 4x the decode work.
// Pipeline: LDS holds two k-blocks (2 x 4 tiles x hi, lo x 1 KiB); the raw
// tiles of the two k-blocks after them are in flight in two register sets,
// and V one k-block pair ahead in two more, so that every wait covers loads
// issued a k-block pair earlier; one barrier per k-block.  A wave without a
// slice (S < 4) or whose slice skips the matvec still loads, decodes and
// meets every barrier.
template <bool I8>
__device__ __forceinline__ void gemm_block_bf16_kshare(const EnsTable *__restrict__ tab, int head,
                                                       int S, int ri, int sp, int wave, int lane,
                                                       int stage, u32x4 *lds)
{
    constexpr int RB = 4;
    const EnsGroup &H = tab->g[head];
    const int nrb = H.mp >> 5;
    const int rb0 = ri * RB;
    const int nvr = min(RB, nrb - rb0);   // 4, or 2 on the tail item
    const int kb0 = sp * H.kb_per_split;
    const int kb1 = min(kb0 + H.kb_per_split, H.nkb);
    // this wave's slice, and whether its product is stored.  Every wave
    // multiplies (a wave without a slice reads slice 0's V): the products
    // stay free of branches, which would move the accumulators between
    // register files at every join.  A block none of whose slices needs the
    // matvec returns before the first barrier.
    const bool mine = wave < S;
    const EnsGroup &G = tab->g[head + (mine ? wave : 0)];
    const bool active = mine && !(stage == 1 && !*G.any_rhs);
    if (stage == 1) {
        bool any = false;
        for (int w = 0; w < S; w++) any = any || *tab->g[head + w].any_rhs;
        if (!any) return;   // block-uniform
    }
    const bool dec = wave < nvr;   // this wave's K tile exists (else zeros)
    floatx16 acc[RB][2];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[a][b][i] = 0.0f;
    const char *A = reinterpret_cast<const char *>(H.Ktiles) +
                    (size_t)(rb0 + (dec ? wave : 0)) * H.nkb * kTileBytes<I8>;
    const u32x4 *Bv = reinterpret_cast<const u32x4 *>(G.Vtiles) + lane;
    u32x4 rh[2], rl[2], bv[2][4];
    u32x2 rq[2];
    auto loadk = [&](int kb, auto set_c) {
        constexpr int s = decltype(set_c)::value;
        const char *t = A + (size_t)min(kb, kb1 - 1) * kTileBytes<I8>;
        rh[s] = kload<true>(reinterpret_cast<const u32x4 *>(t) + lane);
        if constexpr (I8)
            rq[s] = kload<true>(reinterpret_cast<const u32x2 *>(t + 1024) + lane);
        else
            rl[s] = kload<true>(reinterpret_cast<const u32x4 *>(t) + 64 + lane);
    };
    auto loadv = [&](int kb, auto set_c) {
        constexpr int s = decltype(set_c)::value;
        const u32x4 *v = Bv + (size_t)min(kb, kb1 - 1) * 256;
#pragma unroll
        for (int q = 0; q < 4; q++) bv[s][q] = *as_global(v + 64 * q);
    };
    // LDS: [buffer][tile][hi, lo][lane]
    auto put = [&](int buf, auto set_c) {
        constexpr int s = decltype(set_c)::value;
        u32x4 *d = lds + ((buf * RB + wave) * 2) * 64 + lane;
        d[0] = rh[s];
        if constexpr (I8)
            d[64] = __builtin_bit_cast(u32x4, decode_lo8(rh[s], rq[s]));
        else
            d[64] = rl[s];
    };
    // a k-block past the range multiplies a zero V
    auto mult = [&](int buf, bool live, auto set_c) {
        constexpr int s = decltype(set_c)::value;
        const u32x4 zero = {0u, 0u, 0u, 0u};
        u32x4 v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = live ? bv[s][q] : zero;
        const bf16x8 B[2][2] = {{__builtin_bit_cast(bf16x8, v[0]), __builtin_bit_cast(bf16x8, v[1])},
                                {__builtin_bit_cast(bf16x8, v[2]), __builtin_bit_cast(bf16x8, v[3])}};
#pragma unroll
        for (int a = 0; a < RB; a++) {
            const u32x4 *t = lds + ((buf * RB + a) * 2) * 64 + lane;
            const bf16x8 Ah = __builtin_bit_cast(bf16x8, t[0]);
            const bf16x8 Al = __builtin_bit_cast(bf16x8, t[64]);
#pragma unroll
            for (int b = 0; b < 2; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][1], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, B[b][0], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][0], acc[a][b], 0, 0, 0);
            }
        }
    };
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    if (kb0 < kb1) {
        // a wave without a K tile (tail item) leaves zeros in its LDS slots
        if (!dec) {
            const u32x4 zero = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int i = 0; i < 4; i++) lds[(((i >> 1) * RB + wave) * 2 + (i & 1)) * 64 + lane] = zero;
        }
        // LDS buffers 0, 1 <- kb0, kb0 + 1; then, in the order the loop
        // issues them (so that the waits the compiler derives at the loop
        // head hold for every iteration): V set 0 <- kb0, raw set 0 <-
        // kb0 + 2, V set 1 <- kb0 + 1, raw set 1 <- kb0 + 3 (past the
        // range: the last k-block again)
        loadk(kb0, S0{});
        loadk(kb0 + 1, S1{});
        if (dec) {
            put(0, S0{});
            put(1, S1{});
        }
        __builtin_amdgcn_sched_barrier(0);
        loadv(kb0, S0{});
        __builtin_amdgcn_sched_barrier(0);
        loadk(kb0 + 2, S0{});
        __builtin_amdgcn_sched_barrier(0);
        loadv(kb0 + 1, S1{});
        __builtin_amdgcn_sched_barrier(0);
        loadk(kb0 + 3, S1{});
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        for (int kb = kb0; kb < kb1; kb += 2) {
            // buffer 0 holds kb, buffer 1 kb + 1, raw set 0 is kb + 2 and
            // set 1 kb + 3, V set 0 is kb and set 1 kb + 1
            __builtin_amdgcn_sched_barrier(0);
            mult(0, true, S0{});
            __builtin_amdgcn_sched_barrier(0);
            loadv(kb + 2, S0{});
            __syncthreads();
            if (dec) put(0, S0{});
            loadk(kb + 4, S0{});
            __builtin_amdgcn_sched_barrier(0);
            mult(1, kb + 1 < kb1, S1{});
            __builtin_amdgcn_sched_barrier(0);
            loadv(kb + 3, S1{});
            __syncthreads();
            if (dec) put(1, S1{});
            loadk(kb + 5, S1{});
        }
    }
    if (!active) return;
    float *T = G.Tpart + (size_t)sp * G.mp * kEnsCols;
#pragma unroll
    for (int a = 0; a < RB; a++) {
        if (a >= nvr) break;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int row = 32 * (rb0 + a) + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const int col = 32 * b + (lane & 31);
                T[(size_t)row * kEnsCols + col] = acc[a][b][i];
            }
    }
}

template <int RB, bool PIPE, bool TWO, bool I8, bool NT>
__device__ __forceinline__ void gemm_wave_bf16(const EnsGroup &G, int item, int lane)
{
    const int nrb = G.mp >> 5;
    const int nri = bf16_row_items(G.mp, RB);
    const int rb0 = (item % nri) * RB, sp = item / nri;
    if constexpr (TWO) {
        if (RB == 2 || rb0 + RB <= nrb) gemm_tile_bf16_2set<RB, I8, NT>(G, rb0, sp, lane);
        else gemm_tile_bf16_2set<2, I8, NT>(G, rb0, sp, lane);
        return;
    }
    if (RB == 2 || rb0 + RB <= nrb) gemm_tile_bf16<RB, PIPE, I8, NT>(G, rb0, sp, lane);
    else gemm_tile_bf16<2, PIPE, I8, NT>(G, rb0, sp, lane);
}

// split-bf16 GEMM variants of the autotuner: K rows per wave, pipelined
// loads, V staged in LDS (chunk of k-blocks, 0 = V read by every wave)
struct Bf16Variant {
    int rb;
    bool pipe;
    int lds_chunk;
    int dma;        // > 0: K tiles through the wave's LDS ring this many k-blocks deep
                    // (bf16+i8 only)
    bool vreg;      // with dma: V in a register ring of the same depth, wave-shaped
                    // (gemm_wave_bf16_dmav); else V in a ring shared by the block
                    // (gemm_block_bf16_dma)
    bool two_sets;  // two register sets in turn (gemm_tile_bf16_2set)
    bool kshare;    // block-shaped, one slice per wave, K tiles decoded once
                    // per block into LDS (gemm_block_bf16_kshare)
};
constexpr Bf16Variant kBf16Variants[kGemmVariants] = {
    {2, false, 0, 0, false, false, false}, {2, true, 0, 0, false, false, false},
    {4, true, 0, 0, false, false, false},  {4, false, 4, 0, false, false, false},
    {2, true, 8, 0, false, false, false},  {4, false, 0, 0, false, false, false},
    {2, false, 0, 4, false, false, false}, {4, false, 0, 3, false, false, false},
    {4, false, 0, 3, true, false, false},  {4, false, 0, 0, false, true, false},
    {4, false, 0, 0, false, false, true}};

// LDS variants: the 4 waves of a block take 4 consecutive row items of one
// k-split; the block stages the split's V tiles through LDS in chunks of CK
// k-blocks (4 KiB each), double-buffered, one barrier per chunk, so V is read
// from L2 once per block instead of once per wave.  Waves past the last row
// item only help with the staging.  The last row item of a group may hold
// two 32-row blocks (mp is a multiple of 64).
template <int RB, bool PIPE, int CK, bool I8>
__device__ __forceinline__ void gemm_block_bf16_lds(const EnsGroup &G, int ri, int sp, int wave,
                                                    int lane, u32x4 *lds)
{
    const int nrb = G.mp >> 5;
    const int nri = bf16_row_items(G.mp, RB);
    const bool valid = ri < nri;
    const int rb0 = ri * RB;
    const int nvr = valid ? min(RB, nrb - rb0) : 0;   // RB, or 2 on the tail item
    const int kb0 = sp * G.kb_per_split;
    const int kb1 = min(kb0 + G.kb_per_split, G.nkb);
    const int tid = wave * 64 + lane;
    floatx16 acc[RB][2];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[a][b][i] = 0.0f;
    const char *A = reinterpret_cast<const char *>(G.Ktiles) + (size_t)(valid ? rb0 : 0) * G.nkb * kTileBytes<I8>;
    const size_t rstride = (size_t)G.nkb * kTileBytes<I8>;
    const u32x4 *Vg = reinterpret_cast<const u32x4 *>(G.Vtiles);
    const u32x4 zero = {0u, 0u, 0u, 0u};
    // chunk 0 of V into buffer 0
#pragma unroll
    for (int u = 0; u < CK; u++) {
        const int i = tid + 256 * u, kb = kb0 + (i >> 8);
        lds[i] = kb < kb1 ? *as_global(Vg + (size_t)kb * 256 + (i & 255)) : zero;
    }
    __syncthreads();
    u32x4 ca[RB][2], na[RB][2], stg[CK];
    u32x2 cq[RB], nq[RB];
    auto loadK = [&](int kb, u32x4 (&ra)[RB][2], u32x2 (&rq)[RB]) {
#pragma unroll
        for (int a = 0; a < RB; a++)
            if (a < nvr) {
                const char *t = A + a * rstride + (size_t)kb * kTileBytes<I8>;
                ra[a][0] = kload<true>(reinterpret_cast<const u32x4 *>(t) + lane);
                if constexpr (I8)
                    rq[a] = kload<true>(reinterpret_cast<const u32x2 *>(t + 1024) + lane);
                else
                    ra[a][1] = kload<true>(reinterpret_cast<const u32x4 *>(t) + 64 + lane);
            }
    };
    if (valid) loadK(kb0, ca, cq);
    int buf = 0;
    for (int kb = kb0; kb < kb1; kb++) {
        const int k = (kb - kb0) % CK;
        const int next = kb - k + CK;          // first k-block of the next chunk
        if (k == 0 && next < kb1) {
#pragma unroll
            for (int u = 0; u < CK; u++) {
                const int i = tid + 256 * u, kbn = next + (i >> 8);
                stg[u] = kbn < kb1 ? *as_global(Vg + (size_t)kbn * 256 + (i & 255)) : zero;
            }
        }
        if (PIPE && valid && kb + 1 < kb1) loadK(kb + 1, na, nq);
        if (!PIPE && valid && kb > kb0) loadK(kb, ca, cq);
        if (valid) {
            const u32x4 *Bl = lds + buf * (CK * 256) + k * 256 + lane;
            const bf16x8 B[2][2] = {{__builtin_bit_cast(bf16x8, Bl[0]), __builtin_bit_cast(bf16x8, Bl[64])},
                                    {__builtin_bit_cast(bf16x8, Bl[128]), __builtin_bit_cast(bf16x8, Bl[192])}};
#pragma unroll
            for (int a = 0; a < RB; a++) {
                if (a >= nvr) continue;
                const bf16x8 Ah = __builtin_bit_cast(bf16x8, ca[a][0]);
                bf16x8 Al;
                if constexpr (I8)
                    Al = decode_lo8(ca[a][0], cq[a]);
                else
                    Al = __builtin_bit_cast(bf16x8, ca[a][1]);
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][1], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, B[b][0], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, B[b][0], acc[a][b], 0, 0, 0);
                }
            }
            if (PIPE) {
#pragma unroll
                for (int a = 0; a < RB; a++) {
                    ca[a][0] = na[a][0];
                    if constexpr (I8)
                        cq[a] = nq[a];
                    else
                        ca[a][1] = na[a][1];
                }
            }
        }
        // end of a chunk with another to come: every wave has read this
        // chunk's buffer; the other buffer (read two chunks ago, before the
        // last barrier) takes the next chunk
        if (k == CK - 1 && next < kb1) {
#pragma unroll
            for (int u = 0; u < CK; u++) lds[(buf ^ 1) * (CK * 256) + tid + 256 * u] = stg[u];
            __syncthreads();
            buf ^= 1;
        }
    }
    if (!valid) return;
    float *T = G.Tpart + (size_t)sp * G.mp * kEnsCols;
#pragma unroll
    for (int a = 0; a < RB; a++) {
        if (a >= nvr) continue;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int row = 32 * (rb0 + a) + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                const int col = 32 * b + (lane & 31);
                T[(size_t)row * kEnsCols + col] = acc[a][b][i];
            }
    }
}

// binary64 batched matvec on v_mfma_f64_16x16x4f64: one wavefront = 64 rows
// x 64 columns over one k-split (4 x 4 tiles of 16 x 16)
typedef double doublex4 __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ void gemm_wave_f64(const EnsGroup &G, int item, int lane)
{
    const int nrq = G.mp >> 6;
    const int rq = item % nrq, sp = item / nrq;
    const int nkt = G.np >> 2;
    const int kt0 = sp * G.kb_per_split * 4;
    int kt1 = kt0 + G.kb_per_split * 4;
    if (kt1 > nkt) kt1 = nkt;
    doublex4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[a][c] = (doublex4){0.0, 0.0, 0.0, 0.0};
    const double *A = G.K64 + (size_t)(4 * rq) * nkt * 64 + lane;
    const double *B = G.V64 + lane;
    for (int kt = kt0; kt < kt1; kt++) {
        double av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; a++) av[a] = kload<NT>(A + ((size_t)a * nkt + kt) * 64);
#pragma unroll
        for (int c = 0; c < 4; c++) bv[c] = *as_global(B + ((size_t)kt * 4 + c) * 64);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 4; c++)
                acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
    }
    // 16x16 f64 accumulator layout (unlike the f32 16x16 forms, rows are
    // interleaved): value i of lane l is row 4*i + l/16, column l%16
    double *T = G.T64 + (size_t)sp * G.mp * kEnsCols;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = 64 * rq + 16 * a + 4 * i + (lane >> 4);
                const int col = 16 * c + (lane & 15);
                T[(size_t)row * kEnsCols + col] = acc[a][c][i];
            }
}

// binary64 path: mode 1 always, mode 2 for the step-start matvec
__host__ __device__ __forceinline__ bool gemm_f64_path(int mode, int stage)
{
    return mode == 1 || (mode == 2 && stage == 1);
}

// stage 1 = the step-start matvec (skipped when every instance of the group
// retries: t_0 stands), 2..6 the stage matvecs, 0 a standalone matvec
template <int V, bool I8>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(const EnsTable *__restrict__ tab, int stage)
{
    constexpr int RB = kBf16Variants[V].rb;
    constexpr int CK = kBf16Variants[V].lds_chunk;
    const int blk = blockIdx.x;
    const int f = group_of(tab->blk_gemm, tab->nfam, blk);
    const int head = tab->fam_head[f], S = tab->fam_slices[f];
    if constexpr (kBf16Variants[V].kshare) {
        // one block per (row item, k-split), one slice per wave
        __shared__ u32x4 lds[2 * 4 * 2 * 64];
        const EnsGroup &H = tab->g[head];
        const int nri = bf16_row_items(H.mp, 4);
        const int local = blk - tab->blk_gemm[f];
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        gemm_block_bf16_kshare<I8>(tab, head, S, local % nri, local / nri, wave, threadIdx.x & 63,
                                   stage, lds);
    } else if constexpr (CK > 0) {
        // one slice per family (the tuner admits the LDS variants only then)
        const EnsGroup &G = tab->g[head];
        if (stage == 1 && !*G.any_rhs) return;   // block-uniform
        __shared__ u32x4 lds[2 * (CK > 0 ? CK : 1) * 256];
        const int nrg = (bf16_row_items(G.mp, RB) + 3) / 4;
        const int local = blk - tab->blk_gemm[f];
        const int wave = threadIdx.x >> 6;
        gemm_block_bf16_lds<RB, kBf16Variants[V].pipe, CK, I8>(G, 4 * (local % nrg) + wave,
                                                               local / nrg, wave, threadIdx.x & 63,
                                                               lds);
    } else if constexpr (kBf16Variants[V].dma > 0 && kBf16Variants[V].vreg) {
        // wave-shaped as below, each wave with its own K ring; bf16+i8 K
        // only (the tuner admits it only then)
        if constexpr (I8) {
            constexpr int RB = kBf16Variants[V].rb, DD = kBf16Variants[V].dma;
            __shared__ __attribute__((aligned(16))) char kring[4 * DD * RB * 1536];
            // wave-uniform, stated so the tile addresses stay scalar
            const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            const int item = (blk - tab->blk_gemm[f]) * 4 + wave;
            const EnsGroup &H = tab->g[head];
            if (item >= bf16_row_items(H.mp, RB) * H.ksplit * S) return;
            const EnsGroup &G = tab->g[head + item % S];
            if (stage == 1 && !*G.any_rhs) return;
            char *ring = kring + wave * (DD * RB * 1536);
            if (S == 1) gemm_wave_bf16_dmav<RB, DD, true>(G, item, threadIdx.x & 63, ring);
            else gemm_wave_bf16_dmav<RB, DD, false>(G, item / S, threadIdx.x & 63, ring);
        }
    } else if constexpr (kBf16Variants[V].dma > 0) {
        // one slice per family (block-shaped, as the V-staging variants);
        // bf16+i8 K only (the tuner admits it only then)
        if constexpr (I8) {
            constexpr int DD = kBf16Variants[V].dma;
            const EnsGroup &G = tab->g[head];
            if (stage == 1 && !*G.any_rhs) return;   // block-uniform
            __shared__ __attribute__((aligned(16))) char kring[4 * DD * RB * 1536];
            __shared__ __attribute__((aligned(16))) char vring[DD * 4096];
            const int nrg = (bf16_row_items(G.mp, RB) + 3) / 4;
            const int local = blk - tab->blk_gemm[f];
            const int wave = threadIdx.x >> 6;
            gemm_block_bf16_dma<RB, DD>(G, 4 * (local % nrg) + wave, local / nrg, wave,
                                        threadIdx.x & 63, kring + wave * (DD * RB * 1536),
                                        vring);
        }
    } else {
        // work items (row item, k-split, slice), slice fastest: the waves of
        // a block take the slices of one row item and k-split, so the
        // family's K tiles leave HBM once for all of them
        // wave-uniform, stated so that the group's fields come by scalar
        // loads and the tile addresses stay scalar
        const int item = (blk - tab->blk_gemm[f]) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const EnsGroup &H = tab->g[head];
        if (item >= bf16_row_items(H.mp, RB) * H.ksplit * S) return;
        const EnsGroup &G = tab->g[head + item % S];
        if (stage == 1 && !*G.any_rhs) return;
        constexpr bool P = kBf16Variants[V].pipe, TW = kBf16Variants[V].two_sets;
        if (S == 1) gemm_wave_bf16<RB, P, TW, I8, true>(G, item, threadIdx.x & 63);
        else gemm_wave_bf16<RB, P, TW, I8, false>(G, item / S, threadIdx.x & 63);
    }
}

int ens_gemm_variant_dma(int v)
{
    return kBf16Variants[v].dma > 0;
}

// block-shaped variants: V shared by the 4 waves of a block (one k-split
// per block, one slice per K family)
int ens_gemm_variant_lds(int v)
{
    return kBf16Variants[v].lds_chunk > 0 || (kBf16Variants[v].dma > 0 && !kBf16Variants[v].vreg);
}

int ens_gemm_variant_kshare(int v)
{
    return kBf16Variants[v].kshare;
}

int ens_gemm_variant_rows(int v)
{
    return 32 * kBf16Variants[v].rb;
}

__global__ void __launch_bounds__(256) gemm_f64_kernel(const EnsTable *__restrict__ tab, int stage)
{
    const int blk = blockIdx.x;
    const int f = group_of(tab->blk_gemm64, tab->nfam, blk);
    const int head = tab->fam_head[f], S = tab->fam_slices[f];
    const EnsGroup &H = tab->g[head];
    const int item = (blk - tab->blk_gemm64[f]) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= (H.mp >> 6) * H.ksplit * S) return;
    const EnsGroup &G = tab->g[head + item % S];
    if (stage == 1 && !*G.any_rhs) return;
    if (S == 1) gemm_wave_f64<true>(G, item, threadIdx.x & 63);
    else gemm_wave_f64<false>(G, item / S, threadIdx.x & 63);
}

template <bool I8>
static void launch_gemm_bf16(hipStream_t st, const EnsTable *dtab, const EnsTable &htab, int stage)
{
    const dim3 grid(htab.blk_gemm[htab.nfam]);
    switch (htab.variant) {
    case 0: hipLaunchKernelGGL((gemm_bf16_kernel<0, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 1: hipLaunchKernelGGL((gemm_bf16_kernel<1, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 2: hipLaunchKernelGGL((gemm_bf16_kernel<2, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 3: hipLaunchKernelGGL((gemm_bf16_kernel<3, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 4: hipLaunchKernelGGL((gemm_bf16_kernel<4, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 5: hipLaunchKernelGGL((gemm_bf16_kernel<5, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 6: hipLaunchKernelGGL((gemm_bf16_kernel<6, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 7: hipLaunchKernelGGL((gemm_bf16_kernel<7, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 8: hipLaunchKernelGGL((gemm_bf16_kernel<8, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    case 9: hipLaunchKernelGGL((gemm_bf16_kernel<9, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    default: hipLaunchKernelGGL((gemm_bf16_kernel<10, I8>), grid, dim3(256), 0, st, dtab, stage); break;
    }
}

void ens_launch_gemm(hipStream_t st, const EnsTable *dtab, const EnsTable &htab, int stage)
{
    if (gemm_f64_path(htab.g[0].mode, stage))
        hipLaunchKernelGGL(gemm_f64_kernel, dim3(htab.blk_gemm64[htab.nfam]), dim3(256), 0, st,
                           dtab, stage);
    else if (htab.g[0].mode == 3)
        launch_gemm_bf16<true>(st, dtab, htab, stage);
    else
        launch_gemm_bf16<false>(st, dtab, htab, stage);
}

// ---------------------------------------------------------------------------
// per-(element, instance) pieces of odefun
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t sidx(const EnsGroup &G, int comp, int e, int inst)
{
    return ((size_t)comp * G.N + e) * kEnsCols + inst;
}

__device__ __forceinline__ void put_split(const EnsGroup &G, int c, int inst, double v)
{
    uint16_t hi, lo;
    split2(v, hi, lo);
    G.Vtiles[vtile_off(c, inst, 0)] = hi;
    G.Vtiles[vtile_off(c, inst, 1)] = lo;
}

// pre-pass (ratestate.f90:683-702) of element e of instance inst at state X;
// writes the slip rates to out[0..1] and the GEMM operand entries: binary64
// v (mode 1); at the step start (base) v_0, kept, as split bf16 (modes 0, 3)
// or binary64 (mode 2); at a stage the split-bf16 increment v - v_0
// x2, x3, x6: the state's traction components 2, 3 and log10 V, passed in
// registers by callers that have just formed them
__device__ __forceinline__ double ens_prepass(const EnsGroup &G, double x2, double x3, double x6,
                                              double *out, int e, int inst, bool base)
{
    const size_t f = (size_t)e * kEnsCols + inst;
    const double rake = atan2(x3, x2);
    const double velocity = exp(x6 * kELg10);
    const double vs = velocity * cos(rake), vd = velocity * sin(rake);
    out[sidx(G, 0, e, inst)] = vs;
    out[sidx(G, 1, e, inst)] = vd;
    const double v0 = vs - G.fr.vplc[f], v1 = vd - G.fr.vpls[f];
    double *b0 = G.v0 + (size_t)(2 * e) * kEnsCols + inst, *b1 = b0 + kEnsCols;
    if (G.mode == 1 || (G.mode == 2 && base)) {
        G.V64[v64_off(2 * e, inst)] = v0;
        G.V64[v64_off(2 * e + 1, inst)] = v1;
        if (G.mode == 2) {
            *b0 = v0;
            *b1 = v1;
        }
    } else if (base) {
        *b0 = v0;
        *b1 = v1;
        put_split(G, 2 * e, inst, v0);
        put_split(G, 2 * e + 1, inst, v1);
    } else {
        put_split(G, 2 * e, inst, v0 - *b0);
        put_split(G, 2 * e + 1, inst, v1 - *b1);
    }
    return velocity;
}

__device__ __forceinline__ double sum_bf16(const EnsGroup &G, int r, int inst)
{
    double acc = 0.0;
    for (int k = 0; k < G.ksplit; k++) acc += (double)G.Tpart[((size_t)k * G.mp + r) * kEnsCols + inst];
    return acc * G.rscale[r];
}

__device__ __forceinline__ double sum_f64(const EnsGroup &G, int r, int inst)
{
    double acc = 0.0;
    for (int k = 0; k < G.ksplit; k++) acc += G.T64[((size_t)k * G.mp + r) * kEnsCols + inst];
    return acc;
}

// traction rates of element e at the current stage (stage 1: step start,
// whose value is kept as t_0 in modes 0 and 2)
__device__ __forceinline__ void ens_traction(const EnsGroup &G, int e, int inst, int stage,
                                             double t[3])
{
    for (int c = 0; c < 3; c++) {
        const int r = 3 * e + c;
        double *t0 = G.t0 + (size_t)r * kEnsCols + inst;
        if (G.mode == 1) {
            t[c] = sum_f64(G, r, inst);
        } else if (stage == 1) {
            t[c] = G.mode == 2 ? sum_f64(G, r, inst) : sum_bf16(G, r, inst);
            *t0 = t[c];
        } else {
            t[c] = *t0 + sum_bf16(G, r, inst);
        }
    }
}

// elementwise launch geometry: blocks of 256 threads = 4 elements x 64 instances
__device__ __forceinline__ bool elem_coords(const EnsTable *tab, int &g, int &e, int &inst)
{
    g = group_of(tab->blk_elem, tab->ngroups, blockIdx.x);
    const EnsGroup &G = tab->g[g];
    inst = threadIdx.x & 63;
    e = (blockIdx.x - tab->blk_elem[g]) * 4 + (threadIdx.x >> 6);
    return e < G.N && inst < G.ninst;
}

// attempt start: commit of the previous attempt's accepted step (y = ycand,
// ode45.f90:100-101), then the odefun pre-pass at y for instances that start
// a new step.  The commit is idempotent (a finished instance keeps its
// accepted flag and its ycand), and the run ends with one commit launch.
// returns whether it committed; then yv holds the committed state
__device__ __forceinline__ bool commit_elem(const EnsGroup &G, int e, int inst, double yv[7])
{
    if (!G.accepted[inst]) return false;
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const size_t i = sidx(G, k, e, inst);
        yv[k] = G.ycand[i];
        G.y[i] = yv[k];
    }
    return true;
}

__global__ void __launch_bounds__(256) ens_start_kernel(const EnsTable *__restrict__ tab)
{
    int g, e, inst;
    const bool live = elem_coords(tab, g, e, inst);
    const EnsGroup &G = tab->g[g];
    double vel = 0.0;
    bool need = false;
    if (live) {
        double yv[7];
        const bool fresh = commit_elem(G, e, inst, yv);
        need = !G.done[inst] && G.need_rhs[inst];
        if (need) {
            if (!fresh) {
                yv[2] = G.y[sidx(G, 2, e, inst)];
                yv[3] = G.y[sidx(G, 3, e, inst)];
                yv[6] = G.y[sidx(G, 6, e, inst)];
            }
            vel = ens_prepass(G, yv[2], yv[3], yv[6], G.dydt, e, inst, true);
        }
    }
    // vMax per instance: the 4 elements of this block share instance
    // columns, so reduce across the waves first, then one atomic per block
    // and instance into the block's replica (slip rates are positive, so
    // their bit patterns order as they do)
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = (unsigned long long)__double_as_longlong(vel);
    if (__syncthreads_or(need) && threadIdx.x == 0) *G.any_rhs = 1;
    if (threadIdx.x < 64) {
        unsigned long long b = red[threadIdx.x];
        for (int w = 1; w < 4; w++) b = red[threadIdx.x + 64 * w] > b ? red[threadIdx.x + 64 * w] : b;
        if (inst < G.ninst && b)
            atomicMax(reinterpret_cast<unsigned long long *>(
                          &G.vmax_rep[(blockIdx.x % kMaxRep) * kEnsCols + inst]), b);
    }
}

// post-pass (ratestate.f90:747-771) writing stage derivative `stage`
// (1 = step start, which also records the step's dt_try for yscal)
// keep: the derivative's components 2-6 also go to keep[2..6], for the same
// thread's next pass; store = false leaves memory alone (stage 6, whose
// only reader is the same thread's finish_elem)
__device__ __forceinline__ void post_elem(const EnsGroup &G, int e, int inst, int stage,
                                          double *keep = nullptr, bool store = true)
{
    const double *X = stage == 1 ? G.y : G.ytmp;
    double *out = G.ak[stage];
    const size_t f = (size_t)e * kEnsCols + inst;
    double t[3];
    ens_traction(G, e, inst, stage, t);
    const double rake = atan2(X[sidx(G, 3, e, inst)], X[sidx(G, 2, e, inst)]);
    const double velocity = exp(X[sidx(G, 6, e, inst)] * kELg10);
    const double cr = cos(rake), sr = sin(rake);
    const double d5 = (exp(-X[sidx(G, 5, e, inst)] * kELg10) - velocity / G.fr.L[f]) / kELg10;
    const double dtau = t[0] * cr + t[1] * sr;
    const double d6 = (dtau - G.fr.bsig[f] * d5 * kELg10)
                      / (G.fr.asig[f] + G.fr.damping[f] * velocity) / kELg10;
    const double correction = G.fr.damping[f] * velocity * d6 * kELg10;
    const double d2 = t[0] - correction * cr, d3 = t[1] - correction * sr;
    if (keep) {
        keep[2] = d2;
        keep[3] = d3;
        keep[4] = t[2];
        keep[5] = d5;
        keep[6] = d6;
    }
    if (store) {
        out[sidx(G, 2, e, inst)] = d2;
        out[sidx(G, 3, e, inst)] = d3;
        out[sidx(G, 4, e, inst)] = t[2];
        out[sidx(G, 5, e, inst)] = d5;
        out[sidx(G, 6, e, inst)] = d6;
    }
    if (stage == 1 && e == 0) G.hstart[inst] = G.h[inst];
}

__device__ __forceinline__ double ens_stage_value(int stage, double h, double y0, double k1,
                                                  double k2, double k3, double k4, double k5)
{
    double yr;
    switch (stage) {
    case 2:
        return y0 + (h * 0.2) * k1;
    case 3:
        yr = (3.0 / 40.0) * k1;
        yr = yr + (9.0 / 40.0) * k2;
        return y0 + h * yr;
    case 4:
        yr = 0.3 * k1;
        yr = yr + (-0.9) * k2;
        yr = yr + 1.2 * k3;
        return y0 + h * yr;
    case 5:
        yr = (-11.0 / 54.0) * k1;
        yr = yr + 2.5 * k2;
        yr = yr + (-70.0 / 27.0) * k3;
        yr = yr + (35.0 / 27.0) * k4;
        return y0 + h * yr;
    default:
        yr = (1631.0 / 55296.0) * k1;
        yr = yr + (175.0 / 512.0) * k2;
        yr = yr + (575.0 / 13824.0) * k3;
        yr = yr + (44275.0 / 110592.0) * k4;
        yr = yr + (253.0 / 4096.0) * k5;
        return y0 + h * yr;
    }
}

// stage input `stage` (2..6) of element e: y + h sum_j a_ij k_j
// (ode45.f90:133-161) and its odefun pre-pass.  Only the components the
// right-hand side reads are formed: traction 2, 3 (rake), log10 theta 5 and
// log10 V 6.  Slip (0, 1) and normal traction (4) of a stage input feed
// nothing (the derivatives do not depend on them; the fifth-order solution
// and the error are built from the stored derivatives), and skipping them
// saves 3/7 of this loop's traffic.
// klast: components of k_(stage-1), the derivative the same thread's
// post-pass has just formed (read from registers, not back from memory)
__device__ __forceinline__ void stage_elem(const EnsGroup &G, int e, int inst, int stage,
                                           const double *klast)
{
    if (stage == 2 && e == 0)
        for (int r = 0; r < kMaxRep; r++) G.errmax_rep[r * kEnsCols + inst] = 0.0;
    const double h = G.h[inst];
    constexpr int kRhsComps[4] = {2, 3, 5, 6};
    double yt[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const size_t i = sidx(G, kRhsComps[c], e, inst);
        const double kl = klast[kRhsComps[c]];
        const double k1 = stage == 2 ? kl : G.dydt[i];
        const double k2 = stage > 2 ? (stage == 3 ? kl : G.ak[2][i]) : 0.0;
        const double k3 = stage > 3 ? (stage == 4 ? kl : G.ak[3][i]) : 0.0;
        const double k4 = stage > 4 ? (stage == 5 ? kl : G.ak[4][i]) : 0.0;
        const double k5 = stage > 5 ? kl : 0.0;
        yt[c] = ens_stage_value(stage, h, G.y[i], k1, k2, k3, k4, k5);
        G.ytmp[i] = yt[c];
    }
    ens_prepass(G, yt[0], yt[1], yt[3], G.ak[stage], e, inst, false);
}

// after the GEMM of stage s (1..5): the post-pass of s, then the input and
// pre-pass of stage s+1 (one launch per stage instead of two).  k_s goes to
// the stage input from registers; the kernel is held to 64 VGPRs (8 waves
// per SIMD), which its loads need more than the compiler's deeper hoisting
// (a form of it left at 78 VGPRs ran 8-10 % slower)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
ens_post_stage_kernel(const EnsTable *__restrict__ tab, int stage)
{
    int g, e, inst;
    if (!elem_coords(tab, g, e, inst)) return;
    const EnsGroup &G = tab->g[g];
    if (G.done[inst]) return;
    // a retried attempt keeps the step-start derivative (ratestate.f90:313)
    double kv[7];
    if (stage != 1 || G.need_rhs[inst]) {
        post_elem(G, e, inst, stage, kv);
    } else {
        constexpr int kRhsComps[4] = {2, 3, 5, 6};
#pragma unroll
        for (int c = 0; c < 4; c++) kv[kRhsComps[c]] = G.dydt[sidx(G, kRhsComps[c], e, inst)];
    }
    stage_elem(G, e, inst, stage + 1, kv);
}

// fifth-order solution and embedded error of element e (ode45.f90:162-172);
// returns max_k |yerr/yscal| (NaN propagates).  yscal (ratestate.f90:315)
// is formed here from the y and dydt this pass reads anyway, with the dt_try
// of the step's first attempt (retries keep it, as the reference's RKQSm
// keeps its yscal argument): the same arithmetic as a stored yscal, without
// its 7-component write and read.
// k6: stage 6's derivative components 2-6 from the post-pass (its slip
// rates, components 0 and 1, come from the stage-6 pre-pass in memory)
__device__ __forceinline__ double finish_elem(const EnsGroup &G, int e, int inst, const double *k6v)
{
    double emax = 0.0;
    const double h = G.h[inst];
    const double hs = G.hstart[inst];
    const double DC1 = 37.0 / 378.0 - 2825.0 / 27648.0, DC3 = 250.0 / 621.0 - 18575.0 / 48384.0,
                 DC4 = 125.0 / 594.0 - 13525.0 / 55296.0, DC5 = -277.0 / 14336.0,
                 DC6 = 512.0 / 1771.0 - 0.25;
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const size_t i = sidx(G, k, e, inst);
        const double k1 = G.dydt[i], k3 = G.ak[3][i], k4 = G.ak[4][i], k5 = G.ak[5][i],
                     k6 = k >= 2 ? k6v[k] : G.ak[6][i];
        double yr = (37.0 / 378.0) * k1;
        yr = yr + (250.0 / 621.0) * k3;
        yr = yr + (125.0 / 594.0) * k4;
        yr = yr + (512.0 / 1771.0) * k6;
        G.ycand[i] = G.y[i] + h * yr;
        double er = DC1 * k1;
        er = er + DC3 * k3;
        er = er + DC4 * k4;
        er = er + DC5 * k5;
        er = er + DC6 * k6;
        er = h * er;
        const double yscal = fabs(G.y[i]) + fabs(hs * k1) + 1.0e-30;
        const double r = fabs(er / yscal);
        emax = (r > emax || r != r) ? r : emax;
    }
    return emax;
}

// after the GEMM of stage 6: its post-pass, the fifth-order solution, the
// embedded error and the per-instance max |yerr/yscal|.  k_6's components
// 2-6 stay in registers (nothing else reads them) and are not stored; held
// to 64 VGPRs as the post-stage kernel (96 unheld: 115 against 105 µs)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
ens_post_finish_kernel(const EnsTable *__restrict__ tab)
{
    int g, e, inst;
    const bool live = elem_coords(tab, g, e, inst);
    const EnsGroup &G = tab->g[g];
    // all GEMMs of this attempt have run: clear the step-start flag
    if (blockIdx.x == tab->blk_elem[g] && threadIdx.x == 0) *G.any_rhs = 0;
    double emax = 0.0;
    if (live && !G.done[inst]) {
        double k6[7];
        post_elem(G, e, inst, 6, k6, false);
        emax = finish_elem(G, e, inst, k6);
    }
    // the 4 elements of this block share instance columns: reduce across waves
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = (unsigned long long)__double_as_longlong(emax);
    __syncthreads();
    if (threadIdx.x < 64) {
        unsigned long long b = red[threadIdx.x];
        for (int w = 1; w < 4; w++) b = red[threadIdx.x + 64 * w] > b ? red[threadIdx.x + 64 * w] : b;
        if (inst < G.ninst && b)
            atomicMax(reinterpret_cast<unsigned long long *>(
                          &G.errmax_rep[(blockIdx.x % kMaxRep) * kEnsCols + inst]), b);
    }
}

// per-instance RKQSm step control + main-loop time update (ode45.f90:70-105,
// ratestate.f90:317-326).  One block of 256 threads per group: wave 0 runs
// the rule, one lane per instance.  With a step-size network, the four waves
// evaluate it together for the lanes that ask: parameters staged in LDS,
// layer 1 and the tanh of layer 2 split by hidden unit (16 per wave), the
// output summed by wave 0 over the 64 units in order.  Every value is formed
// in the order of mlp_eval, so the proposal is the one a single thread
// computes; the split takes the network's ~5,000 dependent FMAs and global
// loads per lane off the attempt's critical path.
__global__ void __launch_bounds__(256) ens_control_kernel(EnsTable *tab)
{
    // parameters, inputs [input][instance], hidden layers [instance][unit]
    // (rows padded by 4 floats: a lane's 16-byte reads of its own row fall
    // in distinct banks across 16 lanes)
    __shared__ __attribute__((aligned(16))) float sP[kMlpParams + 3];
    __shared__ float sX[kMlpIn][64];
    __shared__ __attribute__((aligned(16))) float sH[64][kMlpH + 4];
    __shared__ __attribute__((aligned(16))) float sT[64][kMlpH + 4];
    const int blk = blockIdx.x;
    const int g = group_of(tab->blk_inst, tab->ngroups, blk);
    const EnsGroup &G = tab->g[g];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int inst = (blk - tab->blk_inst[g]) * 64 + lane;
    const double SAFETY = 0.7, PGROW = -0.2, PSHRNK = -0.15, ERRCON = 1.89e-4;
    double *hist = G.hist;
    bool capcheck = false, finish = false, want = false;
    // the network's parameters into LDS, all of a thread's loads issued at
    // once and before the rule, whose latency they overlap (a strided copy
    // loop would wait for each load in turn); the barrier before the
    // network makes them visible
    if (tab->mlp) {
        constexpr int kPer = (kMlpParams + 255) / 256;
        float w[kPer];
#pragma unroll
        for (int r = 0; r < kPer; r++) {
            const int i = threadIdx.x + 256 * r;
            w[r] = i < kMlpParams ? tab->mlp[i] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < kPer; r++) {
            const int i = threadIdx.x + 256 * r;
            if (i < kMlpParams) sP[i] = w[r];
        }
    }
    double h = 0.0, hnext = 0.0, errmax = 0.0;
    // ---- the rule (wave 0); an accepted step waits for the proposal
    if (wave == 0 && inst < G.ninst && !G.done[inst]) {
        G.attempts[inst] += 1;
        // fold the replicas of the step-start vMax and the error norm (bit
        // patterns: the maximum of non-negative values, NaN above +inf)
        unsigned long long bv = 0ull, be = 0ull;
#pragma unroll
        for (int r = 0; r < kMaxRep; r++) {
            const unsigned long long a = (unsigned long long)__double_as_longlong(
                                             G.vmax_rep[r * kEnsCols + inst]),
                                     c = (unsigned long long)__double_as_longlong(
                                             G.errmax_rep[r * kEnsCols + inst]);
            bv = a > bv ? a : bv;
            be = c > be ? c : be;
        }
        G.vmax[inst] = __longlong_as_double((long long)bv);
        G.errmax[inst] = __longlong_as_double((long long)be);
        const double raw = G.errmax[inst];
        G.accepted[inst] = 0;
        h = G.h[inst];
        if (!(raw == raw) || raw > 1.7976931348623157e308) {
            // an attempt whose error norm is not finite (an overflow inside
            // the stages) is rejected with the largest shrink of the rule,
            // x0.1 (ode45.f90:81 would accept it: NaN > 1 is false); the
            // fourth such attempt in a row fails the instance
            const double nf = hist[4 * kEnsCols + inst] + 1.0;
            hist[4 * kEnsCols + inst] = nf;
            hist[3 * kEnsCols + inst] += 1.0;
            h *= 0.1;
            G.h[inst] = h;
            G.need_rhs[inst] = 0;
            if (nf >= 4.0 || 0.0 + h == 0.0) {
                G.status[inst] = 2;
                G.done[inst] = 1;
                atomicAdd(tab->all_done, 1);
            }
        } else {
            capcheck = true;
            hist[4 * kEnsCols + inst] = 0.0;
            errmax = raw / tab->epsilon;
            if (errmax > 1.0) {
                h = h * SAFETY * pow(errmax, PSHRNK);
                G.h[inst] = h;
                G.need_rhs[inst] = 0;
                hist[3 * kEnsCols + inst] += 1.0;
                if (0.0 + h == 0.0) {
                    G.status[inst] = 1;
                    G.done[inst] = 1;
                    atomicAdd(tab->all_done, 1);
                }
            } else {
                finish = true;
                hnext = errmax > ERRCON ? SAFETY * h * pow(errmax, PGROW) : 10.0 * h;
                if (hnext > tab->max_step) hnext = tab->max_step;
                // learned proposal of the next step where the error estimate
                // is informative (below ERRCON the classic rule already takes
                // the largest growth, x10); the embedded error estimate still
                // judges the proposal and a rejection falls back to the
                // classic shrink
                want = tab->mlp && errmax > ERRCON;
                if (want) {
                    const bool first = hist[1 * kEnsCols + inst] == 0.0;
                    float f[kMlpIn];
                    mlp_features(G.vmax[inst], hist[inst], h, hist[1 * kEnsCols + inst], errmax,
                                 hist[2 * kEnsCols + inst], hist[3 * kEnsCols + inst], hnext, first,
                                 f);
                    const MlpNorm &nm = tab->mlp_norm;
#pragma unroll
                    for (int i = 0; i < kMlpIn; i++) sX[i][lane] = (f[i] - nm.xmean[i]) / nm.xscale[i];
                }
            }
        }
    }
    if (wave == 0 && !want) {
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) sX[i][lane] = 0.0f;
    }
    // ---- the network (all waves, every thread reaches the barriers)
    // (each thread holds its instance's inputs and hidden row in registers
    // and reads the weights, the same for every lane, 4 at a time; every
    // sum runs in mlp_eval's order)
    if (tab->mlp && __syncthreads_or(want)) {
        float x[kMlpIn];
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) x[i] = sX[i][lane];
        for (int jj = 0; jj < 16; jj++) {
            const int j = 16 * wave + jj;
            float a = sP[kMlpOffB1 + j];
#pragma unroll
            for (int i = 0; i < kMlpIn; i++) a = fmaf(sP[j * kMlpIn + i], x[i], a);
            sH[lane][j] = tanhf(a);
        }
        __syncthreads();
        float hv[kMlpH];
#pragma unroll
        for (int j4 = 0; j4 < kMlpH / 4; j4++) {
            const float4 t = reinterpret_cast<const float4 *>(&sH[lane][0])[j4];
            hv[4 * j4] = t.x;
            hv[4 * j4 + 1] = t.y;
            hv[4 * j4 + 2] = t.z;
            hv[4 * j4 + 3] = t.w;
        }
        for (int kk = 0; kk < 16; kk++) {
            const int k = 16 * wave + kk;
            float a = sP[kMlpOffB2 + k];
            const float4 *w4 = reinterpret_cast<const float4 *>(sP + kMlpOffW2 + k * kMlpH);
#pragma unroll
            for (int j4 = 0; j4 < kMlpH / 4; j4++) {
                const float4 w = w4[j4];
                a = fmaf(w.x, hv[4 * j4], a);
                a = fmaf(w.y, hv[4 * j4 + 1], a);
                a = fmaf(w.z, hv[4 * j4 + 2], a);
                a = fmaf(w.w, hv[4 * j4 + 3], a);
            }
            sT[lane][k] = tanhf(a);
        }
        __syncthreads();
    }
    // ---- the accepted step (wave 0)
    if (finish) {
        if (want) {
            float y = sP[kMlpOffB3];
            const float4 *t4 = reinterpret_cast<const float4 *>(&sT[lane][0]);
            const float4 *w4 = reinterpret_cast<const float4 *>(sP + kMlpOffW3);
#pragma unroll
            for (int k4 = 0; k4 < kMlpH / 4; k4++) {
                const float4 t = t4[k4], w = w4[k4];
                y = fmaf(w.x, t.x, y);
                y = fmaf(w.y, t.y, y);
                y = fmaf(w.z, t.z, y);
                y = fmaf(w.w, t.w, y);
            }
            const MlpNorm &nm = tab->mlp_norm;
            const double lp = (double)(y * nm.yscale + nm.ymean);
            // within a factor 4 of the classic proposal and the classic
            // growth limit: the network corrects the rule, it does not
            // extrapolate past it
            double hp = pow(10.0, lp);
            hp = fmin(fmax(hp, 0.25 * hnext), fmin(4.0 * hnext, 10.0 * h));
            hnext = (hp == hp) ? fmin(hp, tab->max_step) : hnext;
        }
        if (G.slog && G.steps[inst] < G.slog_cap) {
            double *r = G.slog + (size_t)G.steps[inst] * 5 * kEnsCols + inst;
            r[0] = G.time[inst] + h;
            r[kEnsCols] = h;
            r[2 * kEnsCols] = G.vmax[inst];
            r[3 * kEnsCols] = errmax;
            r[4 * kEnsCols] = hist[3 * kEnsCols + inst];
        }
        hist[inst] = log10(fmax(G.vmax[inst], 1e-30));
        hist[1 * kEnsCols + inst] = h;
        hist[2 * kEnsCols + inst] = log10(fmax(errmax, 1e-10));
        hist[3 * kEnsCols + inst] = 0.0;
        G.hdid[inst] = h;
        // the step that was clamped to the end lands on t_end exactly
        const double tnew = G.time[inst] + h;
        const bool last = tab->stop_at_end && tnew >= tab->t_end * (1.0 - 1e-15);
        G.time[inst] = last ? tab->t_end : tnew;
        if (tab->stop_at_end && !last && G.time[inst] + hnext > tab->t_end)
            hnext = tab->t_end - G.time[inst];
        G.hnext[inst] = hnext;
        G.steps[inst] += 1;
        G.accepted[inst] = 1;
        G.need_rhs[inst] = 1;
        G.h[inst] = hnext;
        // vMax of this step's start (ratestate.f90:530-532 logs it per step)
        const bool event = G.t_event[inst] < 0.0 && G.vmax[inst] > tab->event_vmax;
        if (event) G.t_event[inst] = G.time[inst] - h;
        G.vmax[inst] = 0.0;
        for (int r = 0; r < kMaxRep; r++) G.vmax_rep[r * kEnsCols + inst] = 0.0;
        if (tab->t_end <= G.time[inst] || (event && tab->stop_at_event)) {
            G.done[inst] = 1;
            atomicAdd(tab->all_done, 1);
        }
    }
    if (capcheck && !G.done[inst] && G.attempts[inst] >= tab->max_attempts) {
        G.done[inst] = 1;
        atomicAdd(tab->all_done, 1);
    }
}

// start of a run call: an instance that has not failed, is short of the
// call's end time and still has attempts left steps again; the finished
// count is rebuilt from the flags (all_done is zeroed by the caller)
__global__ void ens_rearm_kernel(EnsTable *tab)
{
    const int blk = blockIdx.x;
    const int g = group_of(tab->blk_inst, tab->ngroups, blk);
    const EnsGroup &G = tab->g[g];
    const int inst = (blk - tab->blk_inst[g]) * 64 + threadIdx.x;
    if (inst >= G.ninst) return;
    const int done = !(G.status[inst] == 0 && G.time[inst] < tab->t_end &&
                       G.attempts[inst] < tab->max_attempts &&
                       !(tab->stop_at_event && G.t_event[inst] >= 0.0));
    G.done[inst] = done;
    if (done) atomicAdd(tab->all_done, 1);
}

void ens_launch_rearm(hipStream_t st, const EnsTable *dtab, const EnsTable &h)
{
    hipLaunchKernelGGL(ens_rearm_kernel, dim3(h.blk_inst[h.ngroups]), dim3(64), 0, st,
                       const_cast<EnsTable *>(dtab));
}

__global__ void __launch_bounds__(256) ens_commit_kernel(const EnsTable *__restrict__ tab)
{
    int g, e, inst;
    if (!elem_coords(tab, g, e, inst)) return;
    double yv[7];
    commit_elem(tab->g[g], e, inst, yv);
}

void ens_launch_start(hipStream_t st, const EnsTable *dtab, const EnsTable &h)
{
    hipLaunchKernelGGL(ens_start_kernel, dim3(h.blk_elem[h.ngroups]), dim3(256), 0, st, dtab);
}
void ens_launch_post_stage(hipStream_t st, const EnsTable *dtab, const EnsTable &h, int stage)
{
    hipLaunchKernelGGL(ens_post_stage_kernel, dim3(h.blk_elem[h.ngroups]), dim3(256), 0, st, dtab,
                       stage);
}
void ens_launch_post_finish(hipStream_t st, const EnsTable *dtab, const EnsTable &h)
{
    hipLaunchKernelGGL(ens_post_finish_kernel, dim3(h.blk_elem[h.ngroups]), dim3(256), 0, st, dtab);
}
void ens_launch_control(hipStream_t st, const EnsTable *dtab, const EnsTable &h)
{
    hipLaunchKernelGGL(ens_control_kernel, dim3(h.blk_inst[h.ngroups]), dim3(256), 0, st,
                       const_cast<EnsTable *>(dtab));
}
void ens_launch_commit(hipStream_t st, const EnsTable *dtab, const EnsTable &h)
{
    hipLaunchKernelGGL(ens_commit_kernel, dim3(h.blk_elem[h.ngroups]), dim3(256), 0, st, dtab);
}

}  // namespace ucy

namespace ucy {
// V (n x ninst, column = instance, row-major [c][inst]) -> the group's V storage
__global__ void ens_fill_v_kernel(const EnsTable *__restrict__ tab, int g, const double *__restrict__ V)
{
    const EnsGroup &G = tab->g[g];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.n * kEnsCols) return;
    const int c = i / kEnsCols, inst = i % kEnsCols;
    const double v = inst < G.ninst ? V[(size_t)c * G.ninst + inst] : 0.0;
    if (G.mode == 1) G.V64[v64_off(c, inst)] = v;
    else put_split(G, c, inst, v);
}

// T (m x ninst) from the partial products
__global__ void ens_reduce_t_kernel(const EnsTable *__restrict__ tab, int g, double *__restrict__ T)
{
    const EnsGroup &G = tab->g[g];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.m * G.ninst) return;
    const int r = i / G.ninst, inst = i % G.ninst;
    T[i] = G.mode == 1 ? sum_f64(G, r, inst) : sum_bf16(G, r, inst);
}

void ens_launch_matvec_group(hipStream_t st, const EnsTable *dtab, const EnsTable &htab, int g,
                             const double *V, double *T)
{
    const EnsGroup &G = htab.g[g];
    hipLaunchKernelGGL(ens_fill_v_kernel, dim3((G.n * kEnsCols + 255) / 256), dim3(256), 0, st, dtab, g, V);
    // the all-groups stage GEMM with the current tiling (the one the runs
    // use); the other groups' partial products are scratch
    ens_launch_gemm(st, dtab, htab, 0);
    hipLaunchKernelGGL(ens_reduce_t_kernel, dim3((G.m * G.ninst + 255) / 256), dim3(256), 0, st, dtab, g, T);
}
}  // namespace ucy
