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

// Cell loads as the group and trio kernels make them: ten groups of six
// lanes per wave (lanes 60-63 idle), each group one cell load at a time,
// each lane 16-byte loads.  PAT 0: four random aligned 128-byte lines per
// cell (pair rows), lane c reading bytes 16 c of each; PAT 1: one 384-byte
// span at a random 192-byte-aligned offset (column rows: planes k and k + 1
// of one column), lane c reading bytes 32 c .. 32 c + 31 of each 192-byte
// half.  nunits: lines (PAT 0) or 192-byte rows (PAT 1) of the buffer; the
// last row is never the first of a span, so every read stays inside it.
template <int PAT>
__global__ void __launch_bounds__(256) cell_gather_kernel(const uint4 *__restrict__ buf,
                                                          uint64_t nunits, uint64_t seed,
                                                          int cells, uint32_t *sink)
{
    const int lane = threadIdx.x & 63;
    const int g = lane / 6, c = lane % 6;
    if (g >= 10) return;
    const uint64_t grp = ((uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * 10 + g;
    const uint64_t key = mix64(seed ^ (grp * 0x9e3779b97f4a7c15ull));
    uint32_t acc = 0;
    for (int it = 0; it < cells; it++) {
        const uint64_t h0 = mix64(key + (uint64_t)it * 4);
        uint4 v[4];
        if (PAT == 0) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint64_t line = __umul64hi(mix64(h0 + j), nunits);  // < nunits
                v[j] = buf[line * 8 + c];
            }
        } else {
            const uint64_t r = __umul64hi(h0, nunits - 1);  // < nunits - 1
            const uint64_t b0 = r * 12 + 2 * c, b1 = (r + 1) * 12 + 2 * c;
            v[0] = buf[b0];
            v[1] = buf[b0 + 1];
            v[2] = buf[b1];
            v[3] = buf[b1 + 1];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) acc ^= v[j].x ^ v[j].w;
    }
    if (acc == 0x9e3779b9u) sink[0] = acc;
}

}  // namespace

// the launchers declared in internal.hpp
hipError_t launch_cell_gather(const void *buf, uint64_t nunits, int pattern, uint64_t seed,
                              int cells, uint32_t *sink, unsigned blocks, hipStream_t stream)
{
    if (pattern == 0)
        hipLaunchKernelGGL(cell_gather_kernel<0>, dim3(blocks), dim3(256), 0, stream,
                           (const uint4 *)buf, nunits, seed, cells, sink);
    else
        hipLaunchKernelGGL(cell_gather_kernel<1>, dim3(blocks), dim3(256), 0, stream,
                           (const uint4 *)buf, nunits, seed, cells, sink);
    return hipGetLastError();
}

hipError_t launch_line_gather(const void *buf, uint64_t nlines, uint64_t seed, int rounds,
                              uint32_t *sink, unsigned blocks, hipStream_t stream)
{
    hipLaunchKernelGGL(line_gather_kernel, dim3(blocks), dim3(256), 0, stream,
                       (const uint4 *)buf, nlines, seed, rounds, sink);
    return hipGetLastError();
}

}  // namespace ptc
