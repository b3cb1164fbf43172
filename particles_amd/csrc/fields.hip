// Device-side synthesis of the benchmark fields on large grids (no
// counterpart in the reference, which reads its fields from FITS files,
// sources/fields.F90:51-78): sums of real Fourier modes with integer wave
// vectors on the unit periodic box, the reconnection-like sheet profile and
// its E = -V x B + eta curl B.  The values are the same functions the host
// generator (particles_amd/workloads.py) evaluates with an inverse FFT; here
// every node sums its modes directly, so grids of 1024^3 are made in seconds
// without a host round trip.
//
// Output: six SOA planes (U or E x,y,z, then B x,y,z) of dm1*dm2*dm3 FP64
// values, x fastest -- the layout ptc_fields_create uploads.
#include "internal.hpp"

namespace ptc {
namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

// out[c][node] += sum_m amp_m pol_m[c] cos(2 pi (n_m . x) + phase_m),
// x = (i/dm0, j/dm1, k/dm2).  The phase is reduced to (-1, 1) turns per
// axis in integer arithmetic before the FP64 cosine.
__global__ void __launch_bounds__(256)
mode_sum_kernel(int dm0, int dm1, int dm2, int nm, const int32_t *__restrict__ n,
                const double *__restrict__ amp, const double *__restrict__ pol,
                const double *__restrict__ phase, double *__restrict__ out)
{
    const int64_t nodes = (int64_t)dm0 * dm1 * dm2;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nodes;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(g % dm0);
        const int j = (int)((g / dm0) % dm1);
        const int k = (int)(g / ((int64_t)dm0 * dm1));
        const double r0 = 1.0 / dm0, r1 = 1.0 / dm1, r2 = 1.0 / dm2;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int m = 0; m < nm; m++) {
            // turns along each axis: (n * i mod dm) / dm, the product exact
            const int64_t ax = ((int64_t)n[3 * m] * i) % dm0;
            const int64_t ay = ((int64_t)n[3 * m + 1] * j) % dm1;
            const int64_t az = ((int64_t)n[3 * m + 2] * k) % dm2;
            const double turns = (double)ax * r0 + (double)ay * r1 + (double)az * r2;
            const double c = amp[m] * cos(kTwoPi * turns + phase[m]);
            s0 += c * pol[3 * m];
            s1 += c * pol[3 * m + 1];
            s2 += c * pol[3 * m + 2];
        }
        out[g] += s0;
        out[nodes + g] += s1;
        out[2 * nodes + g] += s2;
    }
}

__global__ void fill_kernel(double *__restrict__ p, int64_t n, double v, int add)
{
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x)
        p[g] = add ? p[g] + v : v;
}

// periodic double current sheet tanh((y-1/4)/d) - tanh((y-3/4)/d) - 1
__device__ __forceinline__ double sheet(double y, double d)
{
    return tanh((y - 0.25) / d) - tanh((y - 0.75) / d) - 1.0;
}

// B += sheet(y) x + guide z
__global__ void sheet_kernel(int dm0, int dm1, int dm2, double delta, double guide,
                             double *__restrict__ b)
{
    const int64_t nodes = (int64_t)dm0 * dm1 * dm2;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nodes;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)((g / dm0) % dm1);
        b[g] += sheet((double)j / dm1, delta);
        b[2 * nodes + g] += guide;
    }
}

// E = -V x B + eta curl B with V = (0, -vin sheet(y), 0) and curl by
// periodic central differences (spacing 1/dm per axis)
__global__ void reconnection_e_kernel(int dm0, int dm1, int dm2, double delta, double vin,
                                      double eta, const double *__restrict__ b,
                                      double *__restrict__ e)
{
    const int64_t nodes = (int64_t)dm0 * dm1 * dm2;
    const int64_t sy = dm0, sz = (int64_t)dm0 * dm1;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nodes;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(g % dm0);
        const int j = (int)((g / dm0) % dm1);
        const int k = (int)(g / sz);
        const int64_t xp = g - i + (i + 1 == dm0 ? 0 : i + 1), xm = g - i + (i == 0 ? dm0 - 1 : i - 1);
        const int64_t yp = g + ((j + 1 == dm1 ? 0 : j + 1) - j) * sy;
        const int64_t ym = g + ((j == 0 ? dm1 - 1 : j - 1) - j) * sy;
        const int64_t zp = g + ((k + 1 == dm2 ? 0 : k + 1) - k) * sz;
        const int64_t zm = g + ((k == 0 ? dm2 - 1 : k - 1) - k) * sz;
        const double *bx = b, *by = b + nodes, *bz = b + 2 * nodes;
        const double hx = 0.5 * dm0, hy = 0.5 * dm1, hz = 0.5 * dm2;  // 1 / (2 h)
        const double jx = (bz[yp] - bz[ym]) * hy - (by[zp] - by[zm]) * hz;
        const double jy = (bx[zp] - bx[zm]) * hz - (bz[xp] - bz[xm]) * hx;
        const double jz = (by[xp] - by[xm]) * hx - (bx[yp] - bx[ym]) * hy;
        const double vy = -vin * sheet((double)j / dm1, delta);
        e[g] = -(vy * bz[g]) + eta * jx;
        e[nodes + g] = eta * jy;
        e[2 * nodes + g] = vy * bx[g] + eta * jz;
    }
}

int grid_for(int64_t n)
{
    int64_t b = (n + 255) / 256;
    return (int)(b < 65536 ? (b < 1 ? 1 : b) : 65536);
}

}  // namespace

hipError_t synth_fill(double *p, int64_t n, double v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, n, v, 0);
    return hipGetLastError();
}

hipError_t synth_add(double *p, int64_t n, double v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, n, v, 1);
    return hipGetLastError();
}

hipError_t synth_modes(const int dm[3], int nm, const int32_t *n, const double *amp,
                       const double *pol, const double *phase, double *out, hipStream_t s)
{
    if (nm <= 0) return hipSuccess;
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    hipLaunchKernelGGL(mode_sum_kernel, dim3(grid_for(nodes)), dim3(256), 0, s, dm[0], dm[1],
                       dm[2], nm, n, amp, pol, phase, out);
    return hipGetLastError();
}

hipError_t synth_sheet(const int dm[3], double delta, double guide, double *b, hipStream_t s)
{
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    hipLaunchKernelGGL(sheet_kernel, dim3(grid_for(nodes)), dim3(256), 0, s, dm[0], dm[1], dm[2],
                       delta, guide, b);
    return hipGetLastError();
}

hipError_t synth_reconnection_e(const int dm[3], double delta, double vin, double eta,
                                const double *b, double *e, hipStream_t s)
{
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    hipLaunchKernelGGL(reconnection_e_kernel, dim3(grid_for(nodes)), dim3(256), 0, s, dm[0],
                       dm[1], dm[2], delta, vin, eta, b, e);
    return hipGetLastError();
}

}  // namespace ptc
