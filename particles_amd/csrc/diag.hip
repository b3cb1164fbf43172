// Line-gather diagnostic: the rate at which the device serves random,
// aligned 128-byte lines from a buffer far larger than its caches.  That is
// the access pattern of the integrator's cell loads at 1024^3 (a cell's
// corner rows are four such lines in FP64 pair rows, two in FP32 quad
// rows, at addresses no two particles share for long), so this rate, not
// a streaming read, is the ceiling a gather-bound launch can reach.
#include "internal.hpp"

#include <cstdint>

namespace ptc {
namespace {

constexpr int kLineLanes = kDiagLineLanes;  // 8 lanes x 16 bytes = one 128-byte line
constexpr int kInFlight = kDiagInFlight;    // independent line loads per lane before use

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Each group of eight lanes reads one line per load, at a line drawn
// uniformly from [0, nlines) by a hash of (group, iteration); the loads of
// kInFlight iterations are issued before any is used.  The xor of the data
// keeps the loads (it is written only when it hits an unlikely value).
__global__ void __launch_bounds__(256) line_gather_kernel(const uint4 *__restrict__ buf,
                                                          uint64_t nlines, uint64_t seed,
                                                          int rounds, uint32_t *sink)
{
    const uint32_t slot = threadIdx.x % kLineLanes;
    const uint64_t group = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLineLanes;
    const uint64_t key = mix64(seed ^ group);
    uint32_t acc = 0;
    for (int r = 0; r < rounds; r++) {
        uint4 v[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; u++) {
            const uint64_t h = mix64(key + (uint64_t)(r * kInFlight + u));
            const uint64_t line = __umul64hi(h, nlines);  // < nlines
            v[u] = buf[line * kLineLanes + slot];
        }
#pragma unroll
        for (int u = 0; u < kInFlight; u++) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    if (acc == 0x9e3779b9u) sink[0] = acc;
}

}  // namespace

// the launcher declared in internal.hpp
hipError_t launch_line_gather(const void *buf, uint64_t nlines, uint64_t seed, int rounds,
                              uint32_t *sink, unsigned blocks, hipStream_t stream)
{
    hipLaunchKernelGGL(line_gather_kernel, dim3(blocks), dim3(256), 0, stream,
                       (const uint4 *)buf, nlines, seed, rounds, sink);
    return hipGetLastError();
}

}  // namespace ptc
