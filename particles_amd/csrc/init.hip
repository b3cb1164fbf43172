// Initial particle states generated on the device, and the moments the
// reference prints for them (sources/particles.F90:291-312 generates
// positions uniform in the seeding region and momenta normal(0, vth) on the
// host with the intrinsic random_number, random.F90:104-220, and prints the
// mean, standard deviation, skewness and kurtosis of each component,
// random.F90:231-353).
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC 2011), counter = (particle index,
// draw block), key = seed -- every particle's draws depend only on its index
// and the seed.  Eight doubles per particle from four blocks, 53 bits each.
#include "internal.hpp"

#include "particles_amd.h"

namespace ptc {
namespace {

__device__ __forceinline__ void philox_round(uint32_t c[4], const uint32_t k[2])
{
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint64_t seed)
{
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int r = 0; r < 10; r++) {
        if (r) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        philox_round(c, k);
    }
}

// 53-bit uniform in [0, 1) from two 32-bit words
__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * (1.0 / 9007199254740992.0);
}

constexpr double kTwoPi = 6.283185307179586476925286766559;

__global__ void init_kernel(DeviceParticles S, int m, int kind, uint64_t seed, double lo0,
                            double lo1, double lo2, double hi0, double hi1, double hi2, double a,
                            double b)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S.np;
         i += (int64_t)gridDim.x * blockDim.x) {
        double d[8];
#pragma unroll
        for (int blk = 0; blk < 4; blk++) {
            uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)blk, 0u};
            philox4x32_10(c, seed);
            d[2 * blk] = u53(c[0], c[1]);
            d[2 * blk + 1] = u53(c[2], c[3]);
        }
        double s[6];
        s[0] = lo0 + (hi0 - lo0) * d[0];
        s[1] = lo1 + (hi1 - lo1) * d[1];
        s[2] = lo2 + (hi2 - lo2) * d[2];
        if (kind == PTC_INIT_MAXWELL) {
            // Box-Muller: (d3, d4) -> two normals, (d5, d6) -> one
            const double r0 = sqrt(-2.0 * log(1.0 - d[3])), r1 = sqrt(-2.0 * log(1.0 - d[5]));
            s[3] = a * r0 * cos(kTwoPi * d[4]);
            s[4] = a * r0 * sin(kTwoPi * d[4]);
            s[5] = a * r1 * cos(kTwoPi * d[6]);
        } else {
            // gamma fixed (a) or log-uniform in [a, b]; isotropic direction
            const double g = kind == PTC_INIT_GAMMA_LOG ? exp(log(a) + (log(b) - log(a)) * d[3]) : a;
            const double um = sqrt(g * g - 1.0 > 0.0 ? g * g - 1.0 : 0.0);
            const double ct = 2.0 * d[4] - 1.0;
            const double st = sqrt(1.0 - ct * ct > 0.0 ? 1.0 - ct * ct : 0.0);
            const double ph = kTwoPi * d[5];
            s[3] = um * st * cos(ph);
            s[4] = um * st * sin(ph);
            s[5] = um * ct;
        }
        double *o = S.out + (size_t)8 * m * i;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            S.x[c][i] = s[c];
            o[c] = s[c];
        }
        o[6] = 0.0;
        o[7] = 0.0;
    }
}

// per-block sums of x^1..x^4 (about `shift`) of the six components of
// state(1:6,1,:): partials[block][c][p]
__global__ void __launch_bounds__(256)
moments_kernel(const DeviceParticles S, int m, const double *__restrict__ shift,
               double *__restrict__ partials)
{
    __shared__ double red[256];
    double acc[24];
#pragma unroll
    for (int j = 0; j < 24; j++) acc[j] = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S.np;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double *o = S.out + (size_t)8 * m * i;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const double x = o[c] - shift[c], x2 = x * x;
            acc[4 * c] += x;
            acc[4 * c + 1] += x2;
            acc[4 * c + 2] += x2 * x;
            acc[4 * c + 3] += x2 * x2;
        }
    }
#pragma unroll
    for (int j = 0; j < 24; j++) {
        red[threadIdx.x] = acc[j];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(size_t)blockIdx.x * 24 + j] = red[0];
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_init(const DeviceParticles &S, int m, int kind, uint64_t seed, const double lo[3],
                       const double hi[3], double a, double b, hipStream_t stream)
{
    int64_t blocks = (S.np + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(init_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, S, m, kind, seed,
                       lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], a, b);
    return hipGetLastError();
}

hipError_t launch_moments(const DeviceParticles &S, int m, const double *shift, double *partials,
                          int blocks, hipStream_t stream)
{
    hipLaunchKernelGGL(moments_kernel, dim3(blocks), dim3(256), 0, stream, S, m, shift, partials);
    return hipGetLastError();
}

}  // namespace ptc
