// Step-size network kernels for gfx950: packing of the FP32 master
// parameters into the BF16 MFMA layouts, batched inference, the fused
// forward + backward pass that accumulates the mean-squared-error gradient on
// BF16 matrix cores, the deterministic gradient reduction with an Adam
// update, feature statistics, and a probe of the wave-level product.
//
// The network (16 -> 64 -> 64 -> 1, GELU, log step-factor output) is defined
// in mlp_device.hpp.  Gradients are those of the network as evaluated: BF16
// rounding of inputs, weights and activations is treated as the identity in
// the backward pass, FP32 everywhere else.
#include "internal.hpp"
#include "mlp_device.hpp"

namespace ptc {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    // splitmix64 finaliser (Steele, Lea & Flood 2014)
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// training target of a record: log(h_acc / h_prev) + alpha * log(1 / err) / q,
// the accepted step extrapolated to an error ratio of one with the local
// error model err ~ h^q (alpha = 1) or taken as is (alpha = 0), clamped
__device__ __forceinline__ float record_target(const float *r, float alpha, float lo, float hi)
{
    float y = r[16] - alpha * r[17] / r[19];
    y = y < lo ? lo : y;
    return y > hi ? hi : y;
}

// ---------------------------------------------------------------------------
// parameters -> packed BF16 layouts
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pack_one(int p, float w, MlpPacked *__restrict__ P)
{
    if (p < kMlpB1) {
        const int f = p / kMlpHidden, o = p % kMlpHidden;
        P->f.w1t[o][f] = __bf16(w);
    } else if (p < kMlpW2) {
        P->f.b1[p - kMlpB1] = w;
    } else if (p < kMlpB2) {
        const int i = (p - kMlpW2) / kMlpHidden, o = (p - kMlpW2) % kMlpHidden;
        P->f.w2t[o][i] = __bf16(w);
        P->w2[i][o] = __bf16(w);
    } else if (p < kMlpW3) {
        P->f.b2[p - kMlpB2] = w;
    } else if (p < kMlpB3) {
        P->f.w3[p - kMlpW3] = w;
    } else if (p == kMlpB3) {
        P->f.b3 = w;
    } else if (p < kMlpSig) {
        P->f.mu[p - kMlpMu] = w;
    } else if (p < kMlpParams) {
        P->f.rsig[p - kMlpSig] = 1.0f / w;
    }
}

__global__ void mlp_pack_kernel(const float *__restrict__ params, MlpPacked *__restrict__ P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) pack_one(p, params[p], P);
}

// ---------------------------------------------------------------------------
// batched inference: 4 waves per block, 64 samples per wave and pass
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
mlp_forward_kernel(const MlpPacked *__restrict__ Wg, int64_t n, const float *__restrict__ feat,
                   int ld, float *__restrict__ out)
{
    __shared__ MlpInfer W;
    __shared__ MlpWaveBuf buf[4];
    for (int i = threadIdx.x; i < (int)(sizeof(MlpInfer) / 4); i += blockDim.x)
        ((uint32_t *)&W)[i] = ((const uint32_t *)&Wg->f)[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t chunks = (n + 63) / 64;
    for (int64_t c = (int64_t)blockIdx.x * 4 + wave; c < chunks; c += (int64_t)gridDim.x * 4) {
        const int64_t j = c * 64 + lane;
        float f[kMlpIn];
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) f[i] = j < n ? feat[j * ld + i] : W.mu[i];
        const float y = mlp_forward_wave(W, buf[wave], f);
        if (j < n) out[j] = y;
    }
}

// the same through mlp_forward_groups: ten samples per wave and pass
__global__ void __launch_bounds__(256)
mlp_forward_groups_kernel(const MlpPacked *__restrict__ Wg, int64_t n,
                          const float *__restrict__ feat, int ld, float *__restrict__ out)
{
    __shared__ MlpInfer W;
    __shared__ MlpWaveBuf buf[4];
    for (int i = threadIdx.x; i < (int)(sizeof(MlpInfer) / 4); i += blockDim.x)
        ((uint32_t *)&W)[i] = ((const uint32_t *)&Wg->f)[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane / 6;
    const int64_t chunks = (n + 9) / 10;
    for (int64_t c = (int64_t)blockIdx.x * 4 + wave; c < chunks; c += (int64_t)gridDim.x * 4) {
        const int64_t j = c * 10 + g;
        const bool own = lane < 60 && j < n;
        float f[kMlpIn];
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) f[i] = own ? feat[j * ld + i] : W.mu[i];
        const float y = mlp_forward_groups(W, buf[wave], f);
        if (own && lane == 6 * g) out[j] = y;
    }
}

// ---------------------------------------------------------------------------
// forward + backward, one wave per block
// ---------------------------------------------------------------------------
struct MlpTrainBuf {
    __bf16 x[64][kLdIn];   // normalised inputs [sample][feature]
    __bf16 xT[32][kLdH];   // [feature][sample]; rows 16..31 stay zero
    __bf16 a1T[64][kLdH];  // layer-1 activations [unit][sample]
    __bf16 hA[64][kLdH];   // a1 [s][u], then dZ2 [s][o], then dZ1^T [o][s]
    __bf16 hB[64][kLdH];   // a2 [s][u], then dZ2^T [o][s]
    float dy[64];
};

// Per-block gradient partials of the loss mean((y - target)^2) over `batch`
// samples: sample j of the batch is record j (seed == 0) or a record drawn
// uniformly with replacement from n (seed != 0, stream `step`).  partials has
// one row of kMlpTrainable + 1 floats per block (the last is the block's sum
// of squared residuals).
__global__ void __launch_bounds__(64)
mlp_grad_kernel(const MlpPacked *__restrict__ Wg, const float *__restrict__ rec, int64_t n,
                int64_t batch, uint64_t seed, uint64_t step, float alpha, float lo, float hi,
                float *__restrict__ partials)
{
    __shared__ MlpPacked W;
    __shared__ MlpTrainBuf B;
    const int lane = threadIdx.x;
    for (int i = lane; i < (int)(sizeof(MlpPacked) / 4); i += 64)
        ((uint32_t *)&W)[i] = ((const uint32_t *)Wg)[i];
    for (int i = lane; i < 16 * kLdH; i += 64) B.xT[16 + i / kLdH][i % kLdH] = __bf16(0.0f);
    __syncthreads();

    f32x16 gW2[2][2], gW1[2];
#pragma unroll
    for (int n2 = 0; n2 < 2; n2++) {
        gW1[n2] = f32x16{};
        gW2[0][n2] = f32x16{};
        gW2[1][n2] = f32x16{};
    }
    float gb1[2] = {0.0f, 0.0f}, gb2[2] = {0.0f, 0.0f}, gw3[2] = {0.0f, 0.0f};
    float gb3 = 0.0f, loss = 0.0f;
    const float inv_batch = 1.0f / (float)batch;
    const int64_t chunks = (batch + 63) / 64;

    for (int64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const int64_t j = c * 64 + lane;
        const bool valid = j < batch;
        float target = 0.0f;
        {
            float xin[kMlpIn];
            if (valid) {
                const int64_t r = seed ? (int64_t)(mix64(seed ^ mix64(step * 0x9e3779b97f4a7c15ull +
                                                                       (uint64_t)j)) %
                                                   (uint64_t)n)
                                       : j;
                const float *q = rec + r * kRecord;
#pragma unroll
                for (int i = 0; i < kMlpIn; i++) xin[i] = (q[i] - W.f.mu[i]) * W.f.rsig[i];
                target = record_target(q, alpha, lo, hi);
            } else {
#pragma unroll
                for (int i = 0; i < kMlpIn; i++) xin[i] = 0.0f;
            }
#pragma unroll
            for (int i = 0; i < kMlpIn; i++) {
                const __bf16 b = __bf16(xin[i]);
                B.x[lane][i] = b;
                B.xT[i][lane] = b;
            }
        }
        lds_sync_wave();

        // layer 1: z1 = x W1 + b1 (kept in registers), a1 = gelu(z1)
        f32x16 z1[2][2], z2[2][2];
#pragma unroll
        for (int m = 0; m < 2; m++) {
            z1[m][0] = f32x16{};
            z1[m][1] = f32x16{};
            wave_gemm_strip(&B.x[0][0], kLdIn, m, &W.f.w1t[0][0], kLdIn, kMlpIn, z1[m]);
        }
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int n2 = 0; n2 < 2; n2++) {
                const int u = tile_col(n2);
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int s = tile_row(m, v);
                    z1[m][n2][v] += W.f.b1[u];
                    const __bf16 a = __bf16(gelu(z1[m][n2][v]));
                    B.hA[s][u] = a;
                    B.a1T[u][s] = a;
                }
            }
        lds_sync_wave();

        // layer 2
#pragma unroll
        for (int m = 0; m < 2; m++) {
            z2[m][0] = f32x16{};
            z2[m][1] = f32x16{};
            wave_gemm_strip(&B.hA[0][0], kLdH, m, &W.f.w2t[0][0], kLdH, kMlpHidden, z2[m]);
        }
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int n2 = 0; n2 < 2; n2++) {
                const int u = tile_col(n2);
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    z2[m][n2][v] += W.f.b2[u];
                    B.hB[tile_row(m, v)][u] = __bf16(gelu(z2[m][n2][v]));
                }
            }
        lds_sync_wave();

        // output and residual
        float y = W.f.b3;
#pragma unroll 8
        for (int u = 0; u < kMlpHidden; u++) y += float(B.hB[lane][u]) * W.f.w3[u];
        const float res = valid ? y - target : 0.0f;
        loss += res * res;
        const float dy = 2.0f * res * inv_batch;
        gb3 += dy;
        B.dy[lane] = dy;
        lds_sync_wave();

        // dW3, dZ2 = dy w3 gelu'(z2), dB2; dZ2 -> hA [s][o] and hB [o][s]
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int n2 = 0; n2 < 2; n2++) {
                const int u = tile_col(n2);
                const float w3 = W.f.w3[u];
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int s = tile_row(m, v);
                    const float z = z2[m][n2][v];
                    const float d = B.dy[s];
                    gw3[n2] += d * float(__bf16(gelu(z)));
                    const float dz = d * w3 * gelu_grad(z);
                    gb2[n2] += dz;
                    const __bf16 b = __bf16(dz);
                    B.hA[s][u] = b;
                    B.hB[u][s] = b;
                }
            }
        lds_sync_wave();

        // dA1 = dZ2 W2^T (rows s, cols i, the layout of z1); dW2 += a1^T dZ2
        f32x16 da1[2][2];
#pragma unroll
        for (int m = 0; m < 2; m++) {
            da1[m][0] = f32x16{};
            da1[m][1] = f32x16{};
            wave_gemm_strip(&B.hA[0][0], kLdH, m, &W.w2[0][0], kLdH, kMlpHidden, da1[m]);
            wave_gemm_strip(&B.a1T[0][0], kLdH, m, &B.hB[0][0], kLdH, 64, gW2[m]);
        }
        lds_sync_wave();

        // dZ1 = dA1 gelu'(z1), dB1; dZ1^T -> hA [o][s]
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int n2 = 0; n2 < 2; n2++) {
                const int u = tile_col(n2);
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const float dz = da1[m][n2][v] * gelu_grad(z1[m][n2][v]);
                    gb1[n2] += dz;
                    B.hA[u][tile_row(m, v)] = __bf16(dz);
                }
            }
        lds_sync_wave();

        // dW1 += x^T dZ1 (rows = features, 16 of the 32 used)
        wave_gemm_strip(&B.xT[0][0], kLdH, 0, &B.hA[0][0], kLdH, 64, gW1);
        lds_sync_wave();
    }

    // column sums: the two lane halves hold the two halves of the rows
#pragma unroll
    for (int n2 = 0; n2 < 2; n2++) {
        gb1[n2] += __shfl_xor(gb1[n2], 32);
        gb2[n2] += __shfl_xor(gb2[n2], 32);
        gw3[n2] += __shfl_xor(gw3[n2], 32);
    }
    for (int off = 32; off > 0; off >>= 1) {
        gb3 += __shfl_xor(gb3, off);
        loss += __shfl_xor(loss, off);
    }

    float *__restrict__ P = partials + (size_t)blockIdx.x * (kMlpTrainable + 1);
#pragma unroll
    for (int n2 = 0; n2 < 2; n2++) {
        const int o = tile_col(n2);
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int f = tile_row(0, v);
            if (f < kMlpIn) P[kMlpW1 + f * kMlpHidden + o] = gW1[n2][v];
        }
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int v = 0; v < 16; v++) P[kMlpW2 + tile_row(m, v) * kMlpHidden + o] = gW2[m][n2][v];
        if (lane < 32) {
            P[kMlpB1 + o] = gb1[n2];
            P[kMlpB2 + o] = gb2[n2];
            P[kMlpW3 + o] = gw3[n2];
        }
    }
    if (lane == 0) {
        P[kMlpB3] = gb3;
        P[kMlpTrainable] = loss;
    }
}

// Sums the block partials in block order (deterministic).  grad_out, when
// given, receives the gradient and the summed squared residual; with
// apply != 0 an Adam step (Kingma & Ba 2015) updates the master parameters,
// its moments and the packed copies.
__global__ void mlp_reduce_kernel(const float *__restrict__ partials, int nblocks,
                                  float *__restrict__ grad_out, int apply, float lr, float bc1,
                                  float bc2, float *__restrict__ params, float *__restrict__ m1,
                                  float *__restrict__ m2, MlpPacked *__restrict__ P,
                                  float *__restrict__ loss_out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpTrainable) return;
    float g = 0.0f;
    for (int b = 0; b < nblocks; b++) g += partials[(size_t)b * (kMlpTrainable + 1) + p];
    if (grad_out) grad_out[p] = g;
    if (p == kMlpTrainable) {
        if (loss_out) *loss_out = g;
        return;
    }
    if (!apply) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float a = b1 * m1[p] + (1.0f - b1) * g;
    const float v = b2 * m2[p] + (1.0f - b2) * g * g;
    m1[p] = a;
    m2[p] = v;
    const float w = params[p] - lr * (a / bc1) / (sqrtf(v / bc2) + eps);
    params[p] = w;
    pack_one(p, w, P);
}

// per-block sums and sums of squares (FP64) of the 16 features of n records
__global__ void __launch_bounds__(256)
mlp_stats_kernel(const float *__restrict__ rec, int64_t n, double *__restrict__ partials)
{
    __shared__ double red[256];
    double s[kMlpIn], q[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; i++) s[i] = q[i] = 0.0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < kMlpIn; i++) {
            const double x = rec[j * kRecord + i];
            s[i] += x;
            q[i] += x * x;
        }
    }
#pragma unroll
    for (int i = 0; i < 2 * kMlpIn; i++) {
        red[threadIdx.x] = i < kMlpIn ? s[i] : q[i - kMlpIn];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(size_t)blockIdx.x * 2 * kMlpIn + i] = red[0];
        __syncthreads();
    }
}

// C[64][64] = A[64][K] Bt[64][K]^T through the wave product (inputs rounded
// to BF16), for checking the operand layouts
__global__ void __launch_bounds__(64)
mlp_gemm_probe_kernel(const float *__restrict__ A, const float *__restrict__ Bt, float *__restrict__ C)
{
    __shared__ __bf16 a[64][kLdH], b[64][kLdH];
    const int lane = threadIdx.x;
    for (int i = 0; i < 64; i++) {
        a[i][lane] = __bf16(A[i * 64 + lane]);
        b[i][lane] = __bf16(Bt[i * 64 + lane]);
    }
    __syncthreads();
    f32x16 acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; m++) {
        acc[m][0] = f32x16{};
        acc[m][1] = f32x16{};
        wave_gemm_strip(&a[0][0], kLdH, m, &b[0][0], kLdH, 64, acc[m]);
    }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n2 = 0; n2 < 2; n2++)
#pragma unroll
            for (int v = 0; v < 16; v++) C[tile_row(m, v) * 64 + tile_col(n2)] = acc[m][n2][v];
}

}  // namespace

hipError_t mlp_pack(const float *params, MlpPacked *packed, hipStream_t stream)
{
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, params,
                       packed);
    return hipGetLastError();
}

hipError_t mlp_forward(const MlpPacked *packed, int64_t n, const float *feat, int ld, float *out,
                       hipStream_t stream, int groups)
{
    if (n <= 0) return hipSuccess;
    const int64_t per_block = groups ? 40 : 256;
    int64_t blocks = (n + per_block - 1) / per_block;
    if (blocks > 2048) blocks = 2048;
    if (groups)
        hipLaunchKernelGGL(mlp_forward_groups_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                           packed, n, feat, ld, out);
    else
        hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                           packed, n, feat, ld, out);
    return hipGetLastError();
}

int mlp_grad_blocks(int64_t batch)
{
    const int64_t chunks = (batch + 63) / 64;
    return (int)(chunks < kMlpMaxGradBlocks ? chunks : kMlpMaxGradBlocks);
}

hipError_t mlp_gradient(const MlpPacked *packed, const float *rec, int64_t n, int64_t batch,
                        uint64_t seed, uint64_t step, float alpha, float lo, float hi,
                        float *partials, hipStream_t stream)
{
    const int blocks = mlp_grad_blocks(batch);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64), 0, stream, packed, rec, n, batch,
                       seed, step, alpha, lo, hi, partials);
    return hipGetLastError();
}

hipError_t mlp_reduce(const float *partials, int nblocks, float *grad_out, int apply, float lr,
                      float bc1, float bc2, float *params, float *m1, float *m2, MlpPacked *packed,
                      float *loss_out, hipStream_t stream)
{
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpTrainable + 1 + 255) / 256), dim3(256), 0,
                       stream, partials, nblocks, grad_out, apply, lr, bc1, bc2, params, m1, m2,
                       packed, loss_out);
    return hipGetLastError();
}

hipError_t mlp_stats(const float *rec, int64_t n, double *partials, int blocks, hipStream_t stream)
{
    hipLaunchKernelGGL(mlp_stats_kernel, dim3(blocks), dim3(256), 0, stream, rec, n, partials);
    return hipGetLastError();
}

hipError_t mlp_gemm_probe(const float *a, const float *bt, float *c, hipStream_t stream)
{
    hipLaunchKernelGGL(mlp_gemm_probe_kernel, dim3(1), dim3(64), 0, stream, a, bt, c);
    return hipGetLastError();
}

}  // namespace ptc
