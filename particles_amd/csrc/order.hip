// Work-queue orders.  Particles are independent, so the order in which the
// persistent kernel hands them to lanes changes no result; it changes how
// the work is spread.  Finished particles sort last in every order.
//   * spatial: Morton order of the current cell, so the lanes of a wave (and
//     the waves of a CU) start on neighbouring cells and their gathers share
//     cache lines, L2 and translation entries;
//   * cost: most RHS evaluations in the previous segment first (longest
//     processing time first), so the last particles handed out are the cheap
//     ones and lanes finish together.
#include "internal.hpp"
#include "particles_amd.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace ptc {
namespace {

// spreads the low 21 bits of v to every third bit
__device__ __forceinline__ uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ void morton_keys_kernel(const DeviceParticles S, const StepParams P,
                                   uint64_t *__restrict__ keys, int32_t *__restrict__ idx)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S.np;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t key = ~0ull;
        if (S.status[i] == PTC_STATUS_RUNNING) {
            uint64_t c[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const double x = (S.x[a][i] - P.xl[a]) / P.xs[a];
                const double f = P.dmd[a] * (x - floor(x));
                int64_t l = (int64_t)f;
                l = l < 0 ? 0 : (l >= P.dm[a] ? P.dm[a] - 1 : l);  // non-finite -> clamped
                c[a] = (uint64_t)l;
            }
            key = spread3(c[0]) | spread3(c[1]) << 1 | spread3(c[2]) << 2;
        }
        keys[i] = key;
        idx[i] = (int32_t)i;
    }
}

// key = ~(RHS evaluations since the last call), ascending = most work first
// (with no history every key is equal and the stable sort keeps index
// order; a lowest-gamma-first guess for that first launch measured slower)
__global__ void cost_keys_kernel(const DeviceParticles S, int64_t *__restrict__ prev,
                                 uint64_t *__restrict__ keys, int32_t *__restrict__ idx)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < S.np;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t now = S.nrhs[i];
        int64_t d = now - prev[i];
        d = d < 0 ? 0 : (d > 0x7fffffffll ? 0x7fffffffll : d);
        keys[i] = S.status[i] == PTC_STATUS_RUNNING ? (uint64_t)(0x7fffffffll - d) : ~0ull;
        idx[i] = (int32_t)i;
        prev[i] = now;
    }
}

}  // namespace

size_t order_scratch_bytes(int64_t np)
{
    size_t temp = 0;
    (void)rocprim::radix_sort_pairs(nullptr, temp, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                    (const int32_t *)nullptr, (int32_t *)nullptr, (size_t)np, 0,
                                    64);
    // keys in/out, indices in, then the sort's temporary storage
    return 2 * sizeof(uint64_t) * (size_t)np + sizeof(int32_t) * (size_t)np + temp + 256;
}

hipError_t launch_cost_order(const DeviceParticles &S, int64_t *prev, int32_t *order, void *scratch,
                             size_t scratch_bytes, hipStream_t stream)
{
    const int64_t n = S.np;
    if (n <= 0) return hipSuccess;
    uint64_t *k_in = (uint64_t *)scratch;
    uint64_t *k_out = k_in + n;
    int32_t *i_in = (int32_t *)(k_out + n);
    void *temp = (void *)(((uintptr_t)(i_in + n) + 255) & ~(uintptr_t)255);
    size_t temp_bytes = scratch_bytes - ((char *)temp - (char *)scratch);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(cost_keys_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, S, prev,
                       k_in, i_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // 31-bit costs, plus the top bit set for finished particles
    return rocprim::radix_sort_pairs(temp, temp_bytes, k_in, k_out, i_in, order, (size_t)n, 0, 33,
                                     stream);
}

hipError_t launch_spatial_order(const DeviceParticles &S, const StepParams &P, int32_t *order,
                                void *scratch, size_t scratch_bytes, hipStream_t stream)
{
    const int64_t n = S.np;
    if (n <= 0) return hipSuccess;
    uint64_t *k_in = (uint64_t *)scratch;
    uint64_t *k_out = k_in + n;
    int32_t *i_in = (int32_t *)(k_out + n);
    void *temp = (void *)(((uintptr_t)(i_in + n) + 255) & ~(uintptr_t)255);
    size_t temp_bytes = scratch_bytes - ((char *)temp - (char *)scratch);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(morton_keys_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, S, P,
                       k_in, i_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // 3 bits per level of the largest dimension (+1 so that ~0 sorts last)
    int bits = 1;
    const int dmax = P.dm[0] > P.dm[1] ? (P.dm[0] > P.dm[2] ? P.dm[0] : P.dm[2])
                                       : (P.dm[1] > P.dm[2] ? P.dm[1] : P.dm[2]);
    while ((1 << bits) < dmax) bits++;
    int end_bit = 3 * bits + 1;
    if (end_bit > 64) end_bit = 64;
    return rocprim::radix_sort_pairs(temp, temp_bytes, k_in, k_out, i_in, order, (size_t)n, 0,
                                     end_bit, stream);
}

}  // namespace ptc
