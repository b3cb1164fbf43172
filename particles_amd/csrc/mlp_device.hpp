// Step-size network, device side: a 16 -> 64 -> 64 -> 1 MLP with GELU
// activations whose two hidden layers run on gfx950 BF16 matrix cores
// (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32 accumulation), one 64-sample
// batch per wavefront.  The integrator runs it fused (one proposal per
// particle of the wave); the training kernels (mlp.hip) run the same forward
// and its backward.
//
// MFMA operand layouts (32x32 output tile, K = 16 per instruction), lane l:
//   A (32 x 16): row l % 32, k = 8 (l / 32) .. 8 (l / 32) + 7
//   B (16 x 32): col l % 32, k = 8 (l / 32) .. 8 (l / 32) + 7
//   C (32 x 32): col l % 32, element v = 0..15 at row 8 (v / 4) + 4 (l / 32) + v % 4
// Operands are staged in LDS with K contiguous, so each lane's operand is one
// 16-byte read; a product C = A Bt^T takes A as [rows][K] and Bt as [cols][K].
// Rows are padded by 8 elements (16 bytes) against LDS bank conflicts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptc {

constexpr int kMlpIn = 16;
constexpr int kMlpHidden = 64;
// FP32 master parameters, packed: w1[16][64], b1[64], w2[64][64], b2[64],
// w3[64], b3, mu[16], sig[16]  (weights stored [in][out])
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpHidden;
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;
constexpr int kMlpB3 = kMlpW3 + kMlpHidden;
constexpr int kMlpMu = kMlpB3 + 1;
constexpr int kMlpSig = kMlpMu + kMlpIn;
constexpr int kMlpParams = kMlpSig + kMlpIn;  // 5345 floats
constexpr int kMlpTrainable = kMlpMu;         // 5313; mu and sig are data statistics

// a training record: 16 features, log(h_accepted / h_prev), log(error ratio
// of the accepted step), rejections before it, order q of the error estimate
constexpr int kRecord = 20;

constexpr int kLdIn = kMlpIn + 8;       // padded row of a [.][16] operand
constexpr int kLdH = kMlpHidden + 8;    // padded row of a [.][64] operand

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// BF16 weights in the layouts the forward products read, FP32 vectors
struct MlpInfer {
    __bf16 w1t[kMlpHidden][kLdIn];  // [out][in]
    __bf16 w2t[kMlpHidden][kLdH];   // [out][in]
    float b1[kMlpHidden], b2[kMlpHidden], w3[kMlpHidden], b3;
    float mu[kMlpIn], rsig[kMlpIn];  // inputs are (f - mu) * rsig
};

// ... plus W2 as [in][out] for the backward product dA1 = dZ2 W2^T
struct MlpPacked {
    MlpInfer f;
    __bf16 w2[kMlpHidden][kLdH];
};

// GELU, tanh form (Hendrycks & Gimpel 2016):
//   gelu(z) = z (1 + tanh(u)) / 2,  u = sqrt(2/pi) (z + 0.044715 z^3),
// evaluated through the identity (1 + tanh(u)) / 2 = 1 / (1 + e^{-2u}) with
// the hardware base-2 exponential and reciprocal and explicit FMAs.  Clean
// limits: e^{-2u} = inf gives 0 * z, e^{-2u} = 0 gives z.
constexpr float kGeluC1 = -2.0f * 0.7978845608028654f * 1.4426950408889634f;  // -2 c log2(e)
constexpr float kGeluC3 = kGeluC1 * 0.044715f;

// sigma(2u) = 1 / (1 + e^{-2u})
__device__ __forceinline__ float gelu_sigma(float z)
{
    const float w = z * __builtin_fmaf(kGeluC3, z * z, kGeluC1);  // -2u log2(e)
    return __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(w) + 1.0f);
}

__device__ __forceinline__ float gelu(float z) { return z * gelu_sigma(z); }

// d gelu / dz = sigma + z sigma (1 - sigma) 2u'(z), 2u' = 2c (1 + 3 * 0.044715 z^2)
__device__ __forceinline__ float gelu_grad(float z)
{
    const float sg = gelu_sigma(z);
    const float du = 2.0f * 0.7978845608028654f * __builtin_fmaf(3.0f * 0.044715f, z * z, 1.0f);
    return __builtin_fmaf(z * sg * (1.0f - sg), du, sg);
}

// Orders the wave's LDS accesses: everything before completes before anything
// after (lanes exchange data through LDS; one wave, no workgroup barrier).
__device__ __forceinline__ void lds_sync_wave()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acc[n] += A[32 m + ., 0:K] Bt[32 n + ., 0:K]^T, n = 0, 1 (a 32 x 64 strip of
// a 64 x 64 output); A and Bt in LDS with K contiguous, K a multiple of 16
__device__ __forceinline__ void wave_gemm_strip(const __bf16 *A, int lda, int m, const __bf16 *Bt,
                                                int ldb, int K, f32x16 acc[2])
{
    const int lane = (int)__lane_id();
    const int r = lane & 31, kh = (lane >> 5) * 8;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 a = *(const bf16x8 *)(A + (32 * m + r) * lda + k0 + kh);
        const bf16x8 b0 = *(const bf16x8 *)(Bt + r * ldb + k0 + kh);
        const bf16x8 b1 = *(const bf16x8 *)(Bt + (32 + r) * ldb + k0 + kh);
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc[1], 0, 0, 0);
    }
}

// row and column of element v of output tile (m, n) held by this lane
__device__ __forceinline__ int tile_row(int m, int v)
{
    return 32 * m + 8 * (v >> 2) + 4 * ((int)__lane_id() >> 5) + (v & 3);
}
__device__ __forceinline__ int tile_col(int n) { return 32 * n + ((int)__lane_id() & 31); }

// Per-wave LDS of the fused forward: the inputs, then the layer-1
// activations, [sample][unit] (a product has read all of its operand before
// the next layer's values are written).
struct MlpWaveBuf {
    __bf16 h[64][kLdH];
};

// four consecutive values of one row as four BF16 in one 8-byte store
__device__ __forceinline__ void store_bf16x4(__bf16 *dst, float a, float b, float c, float d)
{
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    bf16x4 v;
    v[0] = __bf16(a); v[1] = __bf16(b); v[2] = __bf16(c); v[3] = __bf16(d);
    *(bf16x4 *)dst = v;
}

// Forward pass for the wave's 64 samples: lane l supplies the raw features
// of sample l and receives its output.  Every lane of the wave must call it,
// with the wave converged.  The layers are computed transposed, Z^T = W^T X^T
// (units as rows, samples as columns), so each lane holds four consecutive
// units of one sample: activations leave in 8-byte stores, and the output
// layer reduces in registers.
__device__ __forceinline__ float mlp_forward_wave(const MlpInfer &W, MlpWaveBuf &B, const float *feat)
{
    const int lane = (int)__lane_id();
    {
        bf16x8 x0, x1;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            x0[i] = __bf16((feat[i] - W.mu[i]) * W.rsig[i]);
            x1[i] = __bf16((feat[8 + i] - W.mu[8 + i]) * W.rsig[8 + i]);
        }
        *(bf16x8 *)&B.h[lane][0] = x0;
        *(bf16x8 *)&B.h[lane][8] = x1;
    }
    lds_sync_wave();

    // the accumulators start at the bias of their unit (row)
    const int hi4 = 4 * (lane >> 5);
    f32x16 acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; m++) {
#pragma unroll
        for (int v = 0; v < 16; v++)
            acc[m][0][v] = acc[m][1][v] = W.b1[32 * m + 8 * (v >> 2) + hi4 + (v & 3)];
        wave_gemm_strip(&W.w1t[0][0], kLdIn, m, &B.h[0][0], kLdH, kMlpIn, acc[m]);
    }
    lds_sync_wave();
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++) {
            const int s = tile_col(n);
#pragma unroll
            for (int q = 0; q < 4; q++)
                store_bf16x4(&B.h[s][32 * m + 8 * q + hi4], gelu(acc[m][n][4 * q + 0]),
                             gelu(acc[m][n][4 * q + 1]), gelu(acc[m][n][4 * q + 2]),
                             gelu(acc[m][n][4 * q + 3]));
        }
    lds_sync_wave();

#pragma unroll
    for (int m = 0; m < 2; m++) {
#pragma unroll
        for (int v = 0; v < 16; v++)
            acc[m][0][v] = acc[m][1][v] = W.b2[32 * m + 8 * (v >> 2) + hi4 + (v & 3)];
        wave_gemm_strip(&W.w2t[0][0], kLdH, m, &B.h[0][0], kLdH, kMlpHidden, acc[m]);
    }
    lds_sync_wave();  // the buffer may be reused right after

    // y(s) = b3 + sum_o bf16(gelu(z2(o, s))) w3(o): this lane's 32 units of
    // samples s = l % 32 and 32 + l % 32, then the other half from lane l ^ 32
    float y[2] = {0.0f, 0.0f};
#pragma unroll
    for (int n = 0; n < 2; n++)
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int o = 32 * m + 8 * (v >> 2) + hi4 + (v & 3);
                y[n] = __builtin_fmaf(float(__bf16(gelu(acc[m][n][v]))), W.w3[o], y[n]);
            }
    y[0] += __shfl_xor(y[0], 32);
    y[1] += __shfl_xor(y[1], 32);
    return W.b3 + (lane < 32 ? y[0] : y[1]);
}

// Forward pass for the groups of G lanes of a wave (G = 6: ten groups,
// lanes 60..63 in none; G = 3: 21 groups, lane 63 in none; lanes G g ..
// G g + G - 1 hold the same features, those of sample g): the samples fill
// one 32-sample column tile, so the layers take half the matrix products of
// mlp_forward_wave and a quarter of its accumulators, with the same
// operations per sample (identical outputs).  Every lane of group g
// receives sample g's output.  Every lane of the wave must call it, with the
// wave converged.  Buffer rows 32..63 hold the inputs of samples 0..31
// (samples past the last group zero), rows 0..31 the layer-1 activations.
template <int G = 6>
__device__ __forceinline__ float mlp_forward_groups(const MlpInfer &W, MlpWaveBuf &B,
                                                    const float *feat)
{
    constexpr int kGroups = 64 / G;
    static_assert(kGroups <= 32, "one 32-sample tile");
    const int lane = (int)__lane_id();
    const int g = lane / G;
    if (lane < G * kGroups && lane == G * g) {
        bf16x8 x0, x1;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            x0[i] = __bf16((feat[i] - W.mu[i]) * W.rsig[i]);
            x1[i] = __bf16((feat[8 + i] - W.mu[8 + i]) * W.rsig[8 + i]);
        }
        *(bf16x8 *)&B.h[32 + g][0] = x0;
        *(bf16x8 *)&B.h[32 + g][8] = x1;
    }
    if (lane >= 32 + kGroups) {
        bf16x8 z;
#pragma unroll
        for (int i = 0; i < 8; i++) z[i] = __bf16(0.0f);
        *(bf16x8 *)&B.h[lane][0] = z;
        *(bf16x8 *)&B.h[lane][8] = z;
    }
    lds_sync_wave();

    const int r = lane & 31, kh = (lane >> 5) * 8, hi4 = 4 * (lane >> 5);
    const bf16x8 xb = *(const bf16x8 *)&B.h[32 + r][kh];
#pragma unroll
    for (int m = 0; m < 2; m++) {
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; v++) acc[v] = W.b1[32 * m + 8 * (v >> 2) + hi4 + (v & 3)];
        const bf16x8 a = *(const bf16x8 *)(&W.w1t[0][0] + (32 * m + r) * kLdIn + kh);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xb, acc, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; q++)
            store_bf16x4(&B.h[r][32 * m + 8 * q + hi4], gelu(acc[4 * q + 0]),
                         gelu(acc[4 * q + 1]), gelu(acc[4 * q + 2]), gelu(acc[4 * q + 3]));
    }
    lds_sync_wave();

    float y = 0.0f;
#pragma unroll
    for (int m = 0; m < 2; m++) {
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; v++) acc[v] = W.b2[32 * m + 8 * (v >> 2) + hi4 + (v & 3)];
#pragma unroll
        for (int k0 = 0; k0 < kMlpHidden; k0 += 16) {
            const bf16x8 a = *(const bf16x8 *)(&W.w2t[0][0] + (32 * m + r) * kLdH + k0 + kh);
            const bf16x8 b = *(const bf16x8 *)&B.h[r][k0 + kh];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int o = 32 * m + 8 * (v >> 2) + hi4 + (v & 3);
            y = __builtin_fmaf(float(__bf16(gelu(acc[v]))), W.w3[o], y);
        }
    }
    lds_sync_wave();  // the buffer may be reused right after
    y += __shfl_xor(y, 32);
    return __shfl(W.b3 + y, g);
}

}  // namespace ptc
